"""Build recipe for the CPU FP64 oracle (test infrastructure only).

Compiles oracle/oracle.cpp and oracle/oracle_vtk.cpp into oracle/liboracle.so with g++ -O3
-ffp-contract=off (no FMA contraction, so the operation order written in the
restatement is the one executed) and -fopenmp (FFT lines and pointwise loops
across threads; no reduction depends on the thread count).  The reference itself cannot be built here:
it needs FFTW3-MPI, Eigen3, fftw_arr and beadrodpmer, none of which exist in
this image, so there is no oracle/_ref build.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


def build(verbose=False):
    srcs = [os.path.join(HERE, f) for f in ("oracle.cpp", "oracle_vtk.cpp")]
    tmp = LIB + ".tmp"
    cmd = ["g++", "-O3", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
           "-fPIC", "-shared", "-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
