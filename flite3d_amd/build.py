"""Build recipe for the native libraries of flite3d_amd (in-tree).

* ``libflite3d_prepro.so`` -- host C++ grid preprocessor (g++).
* ``libflite3d_amd.so``    -- HIP kernels for gfx950 plus the C-ABI (hipcc).
* ``fortran/_build/mgnsg3d_gpu`` -- the Fortran host driver over the C-ABI
  (flang), with its ISO_C_BINDING modules.

All land under this directory so that they travel with the source tree.
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


FORTRAN = os.path.join(HERE, "fortran")
FORTRAN_BUILD = os.path.join(FORTRAN, "_build")
FORTRAN_EXE = os.path.join(FORTRAN_BUILD, "mgnsg3d_gpu")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
# module order: each file uses only the ones before it
FORTRAN_SOURCES = ["flite_gpu.f90", "flite_input.f90", "flite_gridio.f90", "mgnsg3d_gpu.f90"]


def build_fortran():
    """Compile the Fortran host (modules + driver) and link it against
    libflite3d_amd.so, found at run time next to the package ($ORIGIN)."""
    os.makedirs(FORTRAN_BUILD, exist_ok=True)
    objs = []
    for f in FORTRAN_SOURCES:
        o = os.path.join(FORTRAN_BUILD, f.replace(".f90", ".o"))
        _run([FLANG, "-O2", "-c", "-module-dir", FORTRAN_BUILD, "-I", FORTRAN_BUILD,
              os.path.join(FORTRAN, f), "-o", o])
        objs.append(o)
    _run([FLANG, "-O2"] + objs + ["-o", FORTRAN_EXE, "-L", HERE, "-lflite3d_amd",
                                  "-Wl,-rpath,$ORIGIN/../.."])
    return FORTRAN_EXE


def build_all():
    build_prepro()
    build_device()
    build_fortran()


if __name__ == "__main__":
    build_all()
