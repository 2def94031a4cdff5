"""Build recipe for the CPU oracle (test infrastructure, see oracle.h).

Compiles oracle/oracle.c with gcc into oracle/_build/liboracle.so.  The
reference solver itself is Fortran + MPI and is not built here (see the
header of oracle.c for how the oracle is pinned instead).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "liboracle.so")


def build():
    os.makedirs(OUT, exist_ok=True)
    cmd = ["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-fno-fast-math",
           "-ffp-contract=off", "-Wall", "-fopenmp", os.path.join(HERE, "oracle.c"),
           "-lm", "-o", LIB]
    print("[build]", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("oracle build failed")
    return LIB


if __name__ == "__main__":
    build()
