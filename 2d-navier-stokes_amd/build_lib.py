"""Builds, in-tree: libns2d_gpu.so (HIP kernels + C ABI) for gfx950 with hipcc;
libns2d_host.so (command line, initial conditions, HDF5 output; plain C) and
the ns2d_solver executable (the reference-compatible driver) with gcc.

The kernel shapes are instantiated by ns2d_pass_tu.hip, compiled once per
transform-length range (PASS_TUS, which must match NS2D_PASS_TUS in
ns2d_kernels.hip); all translation units compile in parallel."""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libns2d_gpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"] + os.environ.get("NS2D_HIPCC_FLAGS", "").split()
SOURCES = ["ns2d_kernels.hip", "ns2d_adj.hip", "ns2d_noalloc.hip", "ns2d_capi.cpp"]
PASS_TUS = [(8, 128), (256, 256), (512, 512), (1024, 1024), (2048, 2048), (4096, 4096)]
DRIVER = os.path.join(HERE, "driver")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
HOST_LIB = os.path.join(HERE, "libns2d_host.so")
SOLVER = os.path.join(HERE, "ns2d_solver")
CC = os.environ.get("CC", "gcc")
# serial HDF5 of the image (1.10.6); the reference links -lhdf5 -lhdf5_hl
HDF5_PREFIX = os.environ.get("NS2D_HDF5_PREFIX", "/opt/conda")
CFLAGS = ["-O3", "-std=gnu99", "-W", "-Wall", "-fPIC"]


def _compile_jobs():
    jobs = []
    for src in SOURCES:
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        jobs.append([HIPCC] + FLAGS + (["-x", "hip"] if src.endswith(".cpp") else []) +
                    ["-c", os.path.join(CSRC, src), "-o", obj])
    for lo, hi in PASS_TUS:
        obj = os.path.join(CSRC, f"ns2d_pass_tu_{lo}_{hi}.o")
        jobs.append([HIPCC] + FLAGS + [f"-DNS2D_TU_LO={lo}", f"-DNS2D_TU_HI={hi}", "-c",
                                       os.path.join(CSRC, "ns2d_pass_tu.hip"), "-o", obj])
    return jobs


def build(verbose: bool = False, jobs: int = 0) -> str:
    cmds = _compile_jobs()
    # the largest units first, so that they do not finish last
    cmds.sort(key=lambda c: ("ns2d_pass_tu" not in c[-1], c[-1]))
    objs = [c[-1] for c in cmds]
    nj = jobs or max(1, min(len(cmds), os.cpu_count() or 1, 16))

    def run(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)

    with ThreadPoolExecutor(max_workers=nj) as ex:
        for f in [ex.submit(run, c) for c in cmds]:
            f.result()
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIB] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    for o in objs:
        os.remove(o)
    build_driver(verbose)
    return LIB


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
