"""Runs the ASan-emulated library (tests/cpuemu/build_emu.py) through every
compiled column- and row-pass shape, on small grids whose column groups reach
past the last column and past the row's end (ky >= ld), with the per-mode
Hou-Li filter table, a non-Hermitian ky = 0 column and data in the Nyquist
column: RK4CN steps (direct and graph-replayed), CN-AB4 steps, the RHS, the
measures with flux terms (and the allocation-free variant), r2c, c2r, the
CFL rate,
adjoint, and an ensemble whose tuner launches every batched shape.  Any
out-of-bounds access aborts the process with an AddressSanitizer report.

Run with NS2D_LIB pointing at the emulated library and the ASan runtime
preloaded (tests/test_cpu_emulation.py does this)."""
import ctypes
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "2d-navier-stokes_amd"))

import numpy as np  # noqa: E402

import ns2d  # noqa: E402
from ns2d import ics  # noqa: E402

LENGTHS = [8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096]


def field(nx, ny, seed):
    w = ics.random_phase(nx, ny, seed=seed, k0=3.0, e0=0.5, filt="houli")
    rng = np.random.default_rng(seed)
    a = np.abs(w).max()
    w[:, -1] = 0.01 * a * (rng.standard_normal(nx) + 1j * rng.standard_normal(nx))
    w[nx - 1, 0] += 0.2 * a
    w[0, 0] = 0.0
    return w


def cfg(nx, ny, **kw):
    base = dict(Nx=nx, Ny=ny, nu=1e-3, hyper_flag=0, hyper_pow=1.0, filter="houli")
    base.update(kw)
    return ns2d.Config(**base)


def exercise(s, w, x):
    s.set_state(w)
    s.spectral_step(1e-4, 3)
    s.get_state()
    s.nonlinear_rhs(w)
    s.system_measures(with_flux=True)
    s.r2c(x)
    s.c2r(w)
    s.max_vort()
    s.cfl_rate()


def main():
    only = [int(a) for a in sys.argv[1:]] or LENGTHS
    lib = ns2d.load_library()
    for n in only:
        for nx, ny, pas in ((n, 8, "k"), (n, 32, "k"), (16, n, "r")):
            w = field(nx, ny, n + ny)
            x = np.random.default_rng(n).standard_normal((nx, ny))
            count = len(ns2d.kpass_variants(nx)) if pas == "k" else len(ns2d.rpass_variants(ny))
            for v in range(count):
                with ns2d.Solver(cfg(nx, ny)) as s:
                    (s.set_kpass_variant if pas == "k" else s.set_rpass_variant)(v)
                    exercise(s, w, x)
            # CN-AB4 (start-up, then graph-replayed periods of four steps)
            with ns2d.Solver(cfg(nx, ny, integrator="AB4", forcing="KOLM", force_k=1, force_scale=0.5)) as s:
                s.set_state(w)
                s.spectral_step(1e-4, 9)
                if hasattr(lib, "ns2d_measures_noalloc"):
                    m = ns2d._Measures()
                    sp = np.zeros(s.n_spect)
                    fl = np.zeros(6 * s.n_spect)
                    assert lib.ns2d_measures_noalloc(s._h, 1, ctypes.byref(m), ns2d._dptr(sp), ns2d._dptr(sp),
                                                     ns2d._dptr(fl)) == 0
            # the adjoint (checkpointed, graph-replayed segments)
            with ns2d.Solver(cfg(nx, ny, quirks=0, check_every=0)) as s:
                s.vjp(w, 1e-4, 5, w, ckpt_every=2)
            print(f"{nx}x{ny}: {count} {'column' if pas == 'k' else 'row'}-pass shapes", flush=True)
        # ensemble: two sizes of this length; its tuner launches every batched shape
        with ns2d.Ensemble([cfg(n, 8), cfg(n, 8), cfg(16, n)]) as e:
            for i, m in enumerate(e.members):
                m.set_state(field(m.Nx, m.Ny, i))
            e.step([1e-4, 2e-4, 1e-4], 3)
        print(f"ensemble at {n}", flush=True)
    print("SWEEP OK", flush=True)


if __name__ == "__main__":
    main()
