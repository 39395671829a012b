"""The library under AddressSanitizer on the CPU (tests/cpuemu): its own
kernel and host sources compiled for the host against a CPU stand-in for HIP
(every GPU thread a host thread, device memory host malloc with redzones), so
that an out-of-bounds global or LDS access is reported on the CPU.

Regression: the column pass's forward epilogue applied the dealias factor on
every slot of a column group, also on slots past the last column, whose ky
can lie past the row's end (ky >= ld); with the per-mode Hou-Li filter table
that read past the table's end (at 128 x 8 a 16-column work group reads
ky = 5..15 of the last rows).  On the GPU such a read faults only when it
crosses into an unmapped page, so the fault came and went.  The sweep runs
every compiled shape of the 8- and 16-point passes the same way."""
import os
import shutil
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++ with AddressSanitizer")

CASE = r'''
import sys
sys.path[:0] = [{pkg!r}, {oracle!r}]
import numpy as np
import ns2d
import ns2d_oracle as O
Nx, Ny = 128, 8
p = O.Params(Nx, Ny, nu=1e-3, hyper_flag=0, hyper_pow=1.0, filter="houli")
rng = np.random.default_rng(Nx + 3 * Ny)
w0 = O.ic_random_spectrum(p, seed=Nx + Ny, k0=3.0)
a = np.abs(w0).max()
w0[:, -1] = 0.01 * a * (rng.standard_normal(Nx) + 1j * rng.standard_normal(Nx))
w0[Nx - 1, 0] += 0.2 * a
w0[0, 0] = 0.0
st = O.State(w0, p)
O.step(st, 1e-4, 3)
for var in range(len(ns2d.kpass_variants(Nx))):
    cfg = ns2d.Config(Nx=Nx, Ny=Ny, nu=1e-3, hyper_flag=0, hyper_pow=1.0, filter="houli")
    with ns2d.Solver(cfg) as s:
        s.set_kpass_variant(var)
        s.set_state(w0)
        s.spectral_step(1e-4, 3)
        got = s.get_state()
        s.nonlinear_rhs(w0)
    assert np.linalg.norm(got - st.w_hat) <= 1e-12 * np.linalg.norm(st.w_hat), var
print("CASE OK", flush=True)
'''


@pytest.fixture(scope="module")
def emu_env():
    sys.path.insert(0, os.path.join(HERE, "cpuemu"))
    import build_emu
    lib = build_emu.build()
    env = dict(os.environ, LD_PRELOAD=build_emu.asan_runtime(), NS2D_LIB=lib, NS2D_NO_TORCH="1",
               NS2D_ENS_CACHE="", ASAN_OPTIONS="detect_leaks=0")
    return env


def _run(args, env, timeout):
    r = subprocess.run([sys.executable] + args, env=env, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0 and "AddressSanitizer" not in r.stderr, (r.stdout[-2000:], r.stderr[-6000:])
    return r.stdout


def test_hou_li_table_reads_stay_in_bounds(emu_env, tmp_path):
    script = tmp_path / "case.py"
    script.write_text(CASE.format(pkg=os.path.join(ROOT, "2d-navier-stokes_amd"), oracle=os.path.join(ROOT, "oracle")))
    assert "CASE OK" in _run([str(script)], emu_env, 300)


def test_small_length_shape_sweep(emu_env):
    out = _run([os.path.join(HERE, "cpuemu", "shape_sweep.py"), "8", "16"], emu_env, 600)
    assert "SWEEP OK" in out
