"""GPU parity at the benchmark configurations' own sizes, through the C ABI,
against the oracle (numpy/scipy restatement of the reference, run with up to
16 FFT threads).

Tolerances (FP64), as the parity criterion states them: relative L2 of w_hat
<= 1e-10 and energy / enstrophy / palinstrophy <= 1e-12 relative at every
save; spectra <= 1e-12 relative L2.

* decaying turbulence, 1024^2 (Hou-Li in the reference's integer-division
  form, RK4CN, nu = 5e-5, dt = 5e-4, seeded k0 = 8, E = 0.5): 100 steps with
  the measures every 20 steps (SpectralSolve's save cadence, solver.c:177-193;
  ComputeSystemMeasurables, sys_msr.c:32-373);
* the same workload at 4096^2, 3 RK4CN steps (the 4096 column-pass default);
* forced Kolmogorov flow, 4096^2 (KOLM k_f = 4, F = 1, drag alpha = 0.01 with
  flag 1 and power 0, nu = 2e-5, 2/3 filter, low-amplitude IC on |k| <= 16):
  3 RK4CN start-up steps and 3 CN-AB4 steps (AB4Step, solver.c:261-350;
  force.c:535-538);
* ensemble with two members each at 256^2, 512^2 and 1024^2 (k0 ~ U[4, 12],
  2/3 filter, CFL-0.5 steps): bit for bit equal to plain solves with the
  autotuner off, within 1e-13 with it on, and the 256^2 members against the
  oracle;
* adjoint at 512^2 over 2000 steps (nu = 1e-4, dt = 1e-3, checkpoints every
  100 steps, J = Z(T), IC projected to E = 0.5 and P = 5000): central finite
  differences of the GPU map along 16 seeded directions (Richardson-
  extrapolated) to 1e-6 relative, and torch autograd through the FP64 torch
  restatement over a short horizon to 1e-11.
"""
import os

import numpy as np
import pytest

import ns2d_oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(autouse=True)
def _oracle_threads():
    O.set_fft_workers(max(1, min(16, os.cpu_count() or 1)))
    yield
    O.set_fft_workers(1)


def solver(p: O.Params, check=1):
    import ns2d
    cfg = ns2d.Config(Nx=p.Nx, Ny=p.Ny, nu=p.nu, hyper_flag=p.hyper_flag, hyper_pow=p.hyper_pow,
                      alpha=p.alpha, drag_flag=p.drag_flag, drag_pow=p.drag_pow, filter=p.filter,
                      forcing=p.forcing, force_k=p.force_k, force_scale=p.force_scale,
                      integrator=p.integrator, quirks=p.quirks, check_every=check)
    return ns2d.Solver(cfg)


def check_save(tag, s, st, p):
    """w_hat and the measures of the GPU state against the oracle's."""
    got = s.get_state()
    m = s.system_measures()
    mo = O.measures(st.w_hat, p, hyper_flag=st.hyper_flag)
    e = rel(got, st.w_hat)
    q = {k: abs(m[k] - mo[k]) / abs(mo[k]) for k in ("energy", "enstrophy", "palinstrophy")}
    print(f"{tag}: w_hat rel L2 {e:.2e}, E/Z/P rel {q['energy']:.1e} {q['enstrophy']:.1e} "
          f"{q['palinstrophy']:.1e}", flush=True)
    assert e <= 1e-10, tag
    for k, v in q.items():
        assert v <= 1e-12, (tag, k)
    assert rel(m["enrg_spect"], mo["enrg_spect"]) <= 1e-12, tag
    assert rel(m["enst_spect"], mo["enst_spect"]) <= 1e-12, tag


# ---------------------------------------------------------------------------
# decaying turbulence, RK4CN, 1024^2 and 4096^2
# ---------------------------------------------------------------------------
@pytest.mark.timeout(900)
def test_decaying_turbulence_1024_series():
    from ns2d import ics
    N, nu, dt, steps, save = 1024, 5e-5, 5e-4, 100, 20
    p = O.Params(N, N, nu=nu, hyper_flag=0, hyper_pow=1.0, filter="houli_ref")
    w0 = ics.random_phase(N, N, seed=1, k0=8.0, e0=0.5, filt="houli_ref")
    st = O.State(w0, p)
    with solver(p) as s:
        s.set_state(w0)
        check_save("1024^2 RK4CN step 0", s, st, p)
        for k in range(steps // save):
            s.spectral_step(dt, save)
            O.step(st, dt, save)
            check_save(f"1024^2 RK4CN step {(k + 1) * save}", s, st, p)
        info = s.info()
        assert info["iters"] == steps and info["first_blowup_iter"] == -1
        assert abs(info["last_max_vort"] - O.max_vort(st.w_hat, p)) <= 1e-12 * info["last_max_vort"]


@pytest.mark.timeout(900)
def test_decaying_turbulence_4096_rk4cn():
    from ns2d import ics
    N, nu, dt = 4096, 5e-5, 5e-4
    p = O.Params(N, N, nu=nu, hyper_flag=0, hyper_pow=1.0, filter="houli_ref")
    w0 = ics.random_phase(N, N, seed=1, k0=8.0, e0=0.5, filt="houli_ref")
    st = O.State(w0, p)
    with solver(p) as s:
        s.set_state(w0)
        for k in range(3):
            s.spectral_step(dt, 1)
            O.step(st, dt, 1)
            print(f"4096^2 RK4CN step {k + 1} done", flush=True)
        check_save("4096^2 RK4CN step 3", s, st, p)


# ---------------------------------------------------------------------------
# forced Kolmogorov flow, CN-AB4, 4096^2
# ---------------------------------------------------------------------------
@pytest.mark.timeout(900)
def test_kolmogorov_4096_ab4():
    from ns2d import ics
    N = 4096
    p = O.Params(N, N, nu=2e-5, hyper_flag=0, hyper_pow=1.0, alpha=0.01, drag_flag=1, drag_pow=0.0,
                 filter="23", forcing="KOLM", force_k=4, force_scale=1.0, integrator="AB4")
    w0 = ics.random_phase(N, N, seed=3, e0=1e-6, filt="23", spectrum="flat", kmax=16.0)
    dt = 2.5e-4
    st = O.State(w0, p)
    with solver(p) as s:
        s.set_state(w0)
        for k in range(6):
            s.spectral_step(dt, 1)
            O.step(st, dt, 1)
            kind = "RK4CN start-up" if k < 3 else "CN-AB4"
            if k in (2, 5):
                check_save(f"4096^2 Kolmogorov step {k + 1} ({kind})", s, st, p)
            else:
                print(f"4096^2 Kolmogorov step {k + 1} ({kind}) done", flush=True)
        m = s.system_measures()
        mo = O.measures(st.w_hat, p, hyper_flag=st.hyper_flag)
        assert abs(m["forcing"] - mo["forcing"]) <= 1e-12 * abs(mo["forcing"])


# ---------------------------------------------------------------------------
# ensemble: two members each at 256^2, 512^2, 1024^2
# ---------------------------------------------------------------------------
def _ensemble_members(seed=2024):
    import ns2d
    from ns2d import ics
    rng = np.random.default_rng(seed)
    sizes = [256, 512, 1024, 256, 512, 1024]
    nu = 5e-5
    params, cfgs, ws, dts = [], [], [], []
    for i, n in enumerate(sizes):
        k0 = float(rng.uniform(4.0, 12.0))
        p = O.Params(n, n, nu=nu, hyper_flag=0, hyper_pow=1.0, filter="23")
        w0 = ics.random_phase(n, n, seed=seed + i, k0=k0, e0=0.5, filt="23")
        params.append(p)
        cfgs.append(ns2d.Config(Nx=n, Ny=n, nu=nu, hyper_flag=0, hyper_pow=1.0, alpha=0.0, drag_flag=0,
                                drag_pow=0.0, filter="23", integrator="RK4CN", check_every=1))
        ws.append(w0)
        dts.append(ics.cfl_timestep(w0, n, n, nu, 0.5))
    return params, cfgs, ws, dts


@pytest.mark.timeout(600)
@pytest.mark.parametrize("tuned", [False, True])
def test_ensemble_config_sizes(monkeypatch, tuned):
    import ns2d
    params, cfgs, ws, dts = _ensemble_members()
    nsteps = 10
    monkeypatch.setenv("NS2D_ENS_TUNE", "1" if tuned else "0")
    with ns2d.Ensemble(cfgs) as e:
        for m, w in zip(e.members, ws):
            m.set_state(w)
        e.step(dts, nsteps)
        got = [m.get_state() for m in e.members]
        shapes = [e.variants(i) for i in range(len(e))]
    print(f"ensemble (tuned={tuned}) shapes per member: {shapes}", flush=True)
    for i, (p, c) in enumerate(zip(params, cfgs)):
        with ns2d.Solver(c) as s:
            s.set_state(ws[i])
            s.spectral_step(dts[i], nsteps)
            plain = s.get_state()
        if tuned:
            assert rel(got[i], plain) <= 1e-13, i
        else:
            assert np.array_equal(got[i], plain), i
        if p.Nx == 256:
            st = O.State(ws[i], p)
            O.step(st, dts[i], nsteps)
            assert rel(got[i], st.w_hat) <= 1e-10, i


# ---------------------------------------------------------------------------
# adjoint: dZ(T)/dw0 at 512^2 over 2000 steps
# ---------------------------------------------------------------------------
def _adjoint_params():
    return O.Params(512, 512, nu=1e-4, hyper_flag=0, hyper_pow=1.0, filter="23", quirks=0)


def _enstrophy_grad(p, w):
    c = np.full((1, p.Ny // 2 + 1), 2.0)
    c[0, 0] = c[0, -1] = 1.0
    g = 2.0 * c * w * (4 * np.pi ** 2) * (0.5 / float(p.Nx * p.Ny) ** 2)
    g[0, 0] = 0.0
    return g


@pytest.mark.timeout(900)
def test_adjoint_512_2000_steps_finite_differences():
    from ns2d import ics
    p = _adjoint_params()
    w0 = ics.random_phase_ep(512, 512, seed=5, k0=6.0, e0=0.5, p0=5000.0)
    dt, n, ckpt = 1e-3, 2000, 100
    with solver(p, check=0) as s:
        def Z(w):
            s.set_state(w)
            s.spectral_step(dt, n)
            wT = s.get_state()
            return O.measures(wT, p)["enstrophy"], wT
        Z0, wT = Z(w0)
        g = s.vjp(w0, dt, n, _enstrophy_grad(p, wT), ckpt_every=ckpt)
        rng = np.random.default_rng(17)
        amp = np.abs(w0)
        for k in range(16):
            d = amp * np.exp(2j * np.pi * rng.random(w0.shape))
            h = 1e-4 * np.linalg.norm(w0) / np.linalg.norm(d)
            D = lambda e: (Z(w0 + e * d)[0] - Z(w0 - e * d)[0]) / (2 * e)
            fd = (4.0 * D(h / 2) - D(h)) / 3.0
            an = float(np.sum(g.real * d.real + g.imag * d.imag))
            print(f"512^2 x 2000 direction {k}: fd {fd:.12e} adjoint {an:.12e} rel {abs(fd - an) / abs(an):.1e}",
                  flush=True)
            assert abs(fd - an) <= 1e-6 * abs(an), (k, fd, an)


@pytest.mark.timeout(600)
def test_adjoint_512_matches_torch_autograd():
    import torch
    from ns2d import ics
    from ns2d_torch_ref import TorchRK4CN
    p = _adjoint_params()
    w0 = ics.random_phase_ep(512, 512, seed=5, k0=6.0, e0=0.5, p0=5000.0)
    dt, n = 1e-3, 3
    T = TorchRK4CN(p)
    w = torch.tensor(w0, requires_grad=True)
    x = w
    for _ in range(n):
        x = T.step(x, dt)
    T.enstrophy(x).backward()
    g_ref = w.grad.numpy()
    with solver(p, check=0) as s:
        g = s.vjp(w0, dt, n, _enstrophy_grad(p, x.detach().numpy()), ckpt_every=2)
    e = rel(g, g_ref)
    print(f"512^2 adjoint of {n} steps vs torch autograd: rel L2 {e:.2e}", flush=True)
    assert e <= 1e-11
