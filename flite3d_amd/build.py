"""Build recipe for the native libraries of flite3d_amd (in-tree).

* ``libflite3d_prepro.so`` -- host C++ grid preprocessor (g++).
* ``libflite3d_amd.so``    -- HIP kernels for gfx950 plus the C-ABI (hipcc).

Both land next to this file so that they travel with the source tree.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")

PREPRO_SO = os.path.join(HERE, "libflite3d_prepro.so")
DEVICE_SO = os.path.join(HERE, "libflite3d_amd.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# packed-FP32 instructions are disabled for all device code on this target
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd):
    print("[build]", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd[:3]))
    return r


def build_prepro():
    src = os.path.join(CSRC, "prepro.cpp")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I", INCLUDE,
          src, "-o", PREPRO_SO])
    return PREPRO_SO


def device_sources():
    return [os.path.join(CSRC, f) for f in sorted(os.listdir(CSRC))
            if f.endswith(".hip")]


def build_device():
    srcs = device_sources()
    _run([HIPCC] + HIP_FLAGS + ["-I", INCLUDE, "-I", CSRC] + srcs +
         ["-o", DEVICE_SO])
    return DEVICE_SO


def build_all():
    build_prepro()
    build_device()


if __name__ == "__main__":
    build_all()
