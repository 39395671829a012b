"""Builds tests/cpuemu/_build/libns2d_emu.so: the library's own sources
(2d-navier-stokes_amd/csrc, unchanged apart from empty inline-asm statements,
which have no host meaning) compiled for the host with g++ and
AddressSanitizer against the CPU stand-in for HIP in tests/cpuemu/hip/.

TEST INFRASTRUCTURE: every GPU thread of a workgroup runs as a host thread
(barriers are real barriers), device memory is host malloc with ASan
redzones, and graph capture records and replays the launches, so an
out-of-bounds global or LDS access in any kernel shape, or in the host code's
buffer sizes, is reported by ASan on the CPU."""
import os
import re
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "2d-navier-stokes_amd", "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libns2d_emu.so")
PASS_TUS = [(8, 128), (256, 256), (512, 512), (1024, 1024), (2048, 2048), (4096, 4096)]


def asan_runtime() -> str:
    return subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True,
                          check=True).stdout.strip()


def build(verbose: bool = False) -> str:
    src = os.path.join(OUT, "src", "csrc")
    inc = os.path.join(OUT, "include")
    obj = os.path.join(OUT, "obj")
    shutil.rmtree(OUT, ignore_errors=True)
    for d in (src, inc, obj):
        os.makedirs(d)
    shutil.copy(os.path.join(ROOT, "include", "ns2d_gpu.h"), inc)
    for f in os.listdir(CSRC):
        text = open(os.path.join(CSRC, f)).read()
        # register-laundering and memory-clobber asm statements are empty
        text = re.sub(r'asm volatile\(""[^;]*\);', ";", text)
        open(os.path.join(src, f), "w").write(text)
    cxx = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address", "-fPIC",
           "-I" + HERE, "-I" + src, "-w"] + os.environ.get("NS2D_EMU_FLAGS", "").split()
    jobs = [cxx + ["-c", os.path.join(HERE, "emu_rt.cpp"), "-o", os.path.join(obj, "emu_rt.o")]]
    for s in ("ns2d_kernels.hip", "ns2d_adj.hip", "ns2d_noalloc.hip", "ns2d_capi.cpp"):
        if not os.path.exists(os.path.join(src, s)):
            continue
        jobs.append(cxx + ["-x", "c++", "-c", os.path.join(src, s), "-o", os.path.join(obj, s + ".o")])
    for lo, hi in PASS_TUS:
        jobs.append(cxx + ["-x", "c++", f"-DNS2D_TU_LO={lo}", f"-DNS2D_TU_HI={hi}", "-c",
                           os.path.join(src, "ns2d_pass_tu.hip"), "-o", os.path.join(obj, f"pass_{lo}.o")])

    def run(c):
        if verbose:
            print(" ".join(c), flush=True)
        subprocess.run(c, check=True)

    with ThreadPoolExecutor(max_workers=max(1, min(len(jobs), os.cpu_count() or 1))) as ex:
        for f in [ex.submit(run, c) for c in jobs]:
            f.result()
    objs = sorted(os.path.join(obj, o) for o in os.listdir(obj))
    run(["g++", "-shared", "-fsanitize=address", "-o", LIB] + objs + ["-lpthread"])
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
