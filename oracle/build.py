"""Build recipe of the C restatement (oracle/c/vv1d_oracle.c) -> oracle/libvv1d_oracle.so.

Checker / CPU-baseline infrastructure only.  -ffp-contract=off keeps the
bracket arithmetic unfused (the reference's order); -march=x86-64-v3 keeps the
library runnable on any AVX2 host.

The reference itself (Cython + petsc4py + PETSc + FFTW + MPI) cannot be built
in this image, so there is no oracle/_ref build.
"""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "c", "vv1d_oracle.c")
OUT = os.path.join(HERE, "libvv1d_oracle.so")


def build(verbose=False):
    tmp = OUT + ".tmp"
    cmd = ["gcc", "-O3", "-std=gnu11", "-march=x86-64-v3", "-mtune=generic", "-ffp-contract=off", "-fopenmp",
           "-fPIC", "-shared", "-o", tmp, SRC, "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
