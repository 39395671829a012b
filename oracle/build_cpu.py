"""Builds oracle/ns2d_cpu (the compiled CPU restatement of the RK4CN step,
measurement and test infrastructure) with gcc + OpenMP against the FFTW3
interface of the image's Intel MKL (libmkl_rt)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "ns2d_cpu.c")
EXE = os.path.join(HERE, "ns2d_cpu")
MKL_LIB = os.environ.get("NS2D_MKL_LIB", "/opt/conda/lib")


def build_cpu(verbose: bool = False) -> str:
    cmd = [os.environ.get("CC", "gcc"), "-O3", "-std=c99", "-march=x86-64-v3", "-ffp-contract=off", "-fopenmp",
           "-W", "-Wall", SRC, "-o", EXE, "-L" + MKL_LIB, "-lmkl_rt", "-Wl,-rpath," + MKL_LIB, "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return EXE


if __name__ == "__main__":
    print(build_cpu(verbose="-v" in sys.argv))
