"""The compiled CPU restatement used as the CPU baseline (oracle/ns2d_cpu.c)
against the numpy oracle: same RK4CN trajectory (reference quirks on) and the
same per-step max|w| check, on square and rectangular grids with both
filters the baseline supports."""
import json
import os
import subprocess

import numpy as np
import pytest

import ns2d_oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "ns2d_cpu")


@pytest.fixture(scope="module")
def exe():
    if not os.path.exists(EXE):
        import build_cpu
        build_cpu.build_cpu()
    return EXE


@pytest.mark.parametrize("nx,ny,filt", [(64, 64, "houli_ref"), (32, 64, "23"), (64, 32, "houli_ref")])
def test_compiled_baseline_matches_oracle(exe, tmp_path, nx, ny, filt):
    from ns2d import ics
    nu, dt, n = 1e-3, 1e-3, 7
    p = O.Params(nx, ny, nu=nu, hyper_flag=0, hyper_pow=1.0, filter=filt)
    w0 = ics.random_phase(nx, ny, seed=nx + ny, k0=4.0, e0=0.5, filt=filt)
    w0[nx - 3, 0] += 0.2 * np.abs(w0).max()      # non-Hermitian ky=0 entry: the conjugacy write
    icp, outp = tmp_path / "ic.bin", tmp_path / "out.bin"
    w0.astype(np.complex128).tofile(icp)
    env = dict(os.environ, OMP_NUM_THREADS="2")
    r = subprocess.run([exe, str(nx), str(ny), repr(nu), "1.0", filt, repr(dt), str(n), "1", "1", str(icp),
                        str(outp)], capture_output=True, text=True, env=env, check=True)
    res = json.loads(r.stdout.strip().splitlines()[-1])
    got = np.fromfile(outp, dtype=np.complex128).reshape(nx, ny // 2 + 1)
    st = O.State(w0, p)
    O.step(st, dt, n)
    assert np.linalg.norm(got - st.w_hat) / np.linalg.norm(st.w_hat) < 1e-13
    assert abs(res["last_max_vort"] - O.max_vort(st.w_hat, p)) <= 1e-13 * res["last_max_vort"]
    assert res["threads"] == 2 and res["steps"] == n and res["median_steps_per_s"] > 0
