"""Builds, in-tree: libns2d_gpu.so (HIP kernels + C ABI) for gfx950 with hipcc;
libns2d_host.so (command line, initial conditions, HDF5 output; plain C) and
the ns2d_solver executable (the reference-compatible driver) with gcc; the
probes/ measurement programs with hipcc."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libns2d_gpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
SOURCES = ["ns2d_kernels.hip", "ns2d_capi.cpp"]
DRIVER = os.path.join(HERE, "driver")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
HOST_LIB = os.path.join(HERE, "libns2d_host.so")
SOLVER = os.path.join(HERE, "ns2d_solver")
CC = os.environ.get("CC", "gcc")
# serial HDF5 of the image (1.10.6); the reference links -lhdf5 -lhdf5_hl
HDF5_PREFIX = os.environ.get("NS2D_HDF5_PREFIX", "/opt/conda")
CFLAGS = ["-O3", "-std=gnu99", "-W", "-Wall", "-fPIC"]


def build(verbose: bool = False) -> str:
    objs = []
    for src in SOURCES:
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC] + FLAGS + (["-x", "hip"] if src.endswith(".cpp") else []) + \
              ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIB] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    for o in objs:
        os.remove(o)
    build_driver(verbose)
    build_probes(verbose)
    return LIB


def build_probes(verbose: bool = False) -> None:
    """Stand-alone measurement programs under probes/ (not part of the library)."""
    src = os.path.join(HERE, "probes", "column_bandwidth.hip")
    cmd = [HIPCC] + FLAGS + [src, "-o", os.path.splitext(src)[0]]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_driver(verbose: bool = False) -> str:
    inc = ["-I" + INCLUDE, "-I" + DRIVER, "-I" + os.path.join(HDF5_PREFIX, "include")]
    h5 = ["-L" + os.path.join(HDF5_PREFIX, "lib"), "-lhdf5_hl", "-lhdf5",
          "-Wl,-rpath," + os.path.join(HDF5_PREFIX, "lib")]
    srcs = [os.path.join(DRIVER, f) for f in ("ns2d_cli.c", "ns2d_ics.c", "ns2d_h5.c")]
    # HDF5's directory is a run path of libns2d_host.so only: that directory
    # also holds an older libstdc++ that the HIP runtime must not pick up
    cmds = [[CC] + CFLAGS + inc + ["-shared", "-o", HOST_LIB] + srcs + h5 + ["-lm"],
            [CC] + CFLAGS + inc + [os.path.join(DRIVER, "ns2d_main.c"), "-o", SOLVER, "-L" + HERE,
                                   "-lns2d_host", "-lns2d_gpu", "-Wl,-rpath,$ORIGIN", "-lm"]]
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    return SOLVER


if __name__ == "__main__":
    if "--driver" in sys.argv:
        print(build_driver(verbose="-v" in sys.argv))
    else:
        print(build(verbose="-v" in sys.argv))
