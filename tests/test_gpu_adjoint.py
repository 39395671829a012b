"""GPU: the hand-written discrete adjoint of the RK4CN map against torch
autograd through a pure-torch FP64 restatement (CPU) and against central
finite differences of the GPU forward map."""
import numpy as np
import pytest

import ns2d_oracle as O

pytestmark = pytest.mark.gpu


def make(p):
    import ns2d
    cfg = ns2d.Config(Nx=p.Nx, Ny=p.Ny, nu=p.nu, hyper_flag=p.hyper_flag, hyper_pow=p.hyper_pow,
                      alpha=p.alpha, drag_flag=p.drag_flag, drag_pow=p.drag_pow, filter=p.filter,
                      quirks=p.quirks, check_every=0)
    return ns2d.Solver(cfg)


def torch_grad(p, w0, dt, n):
    import torch
    from ns2d_torch_ref import TorchRK4CN
    T = TorchRK4CN(p)
    w = torch.tensor(w0, requires_grad=True)
    x = w
    for _ in range(n):
        x = T.step(x, dt)
    J = T.enstrophy(x)
    J.backward()
    return float(J.detach()), w.grad.numpy(), x.detach().numpy()


def enstrophy_grad(p, w):
    kx, ky = O.wavenumbers(p.Nx, p.Ny)
    c = np.full((1, p.Ny // 2 + 1), 2.0)
    c[0, 0] = c[0, -1] = 1.0
    g = 2.0 * c * w * (4 * np.pi ** 2) * (0.5 / float(p.Nx * p.Ny) ** 2)
    g[0, 0] = 0.0
    return g


# Ny = 256, 512, 1024 run the single-row adjoint row pass (their default)
@pytest.mark.parametrize("Nx,Ny,n,ckpt", [(32, 32, 5, 0), (64, 32, 6, 4), (64, 64, 7, 3), (128, 128, 4, 0),
                                          (256, 256, 3, 2), (64, 512, 2, 0), (32, 1024, 2, 0)])
def test_vjp_matches_torch_autograd(Nx, Ny, n, ckpt):
    print(f"vjp {Nx}x{Ny} n={n} ckpt={ckpt}", flush=True)
    p = O.Params(Nx, Ny, nu=2e-3, hyper_flag=0, hyper_pow=1.0, filter="23", quirks=0)
    w0 = O.ic_random_spectrum(p, seed=Nx + n, k0=4.0)
    w0[Nx - 2, 0] += 0.2 * np.abs(w0).max()   # non-Hermitian ky=0 entry
    dt = 2e-3
    J, g_ref, wT = torch_grad(p, w0, dt, n)
    with make(p) as s:
        g = s.vjp(w0, dt, n, enstrophy_grad(p, wT), ckpt_every=ckpt)
    assert np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref) < 1e-11


def test_vjp_finite_differences():
    """dZ(T)/dw0 along 16 seeded directions against central differences of
    the GPU map (Richardson-extrapolated in the step, so the O(eps^2) term
    drops out): agreement to 1e-6 relative."""
    p = O.Params(128, 128, nu=1e-3, hyper_flag=0, hyper_pow=1.0, quirks=0)
    w0 = O.ic_random_spectrum(p, seed=5, k0=6.0)
    dt, n = 2e-3, 20
    with make(p) as s:
        def Z(w):
            s.set_state(w)
            s.spectral_step(dt, n)
            wT = s.get_state()
            return O.measures(wT, p)["enstrophy"], wT
        _, wT = Z(w0)
        g = s.vjp(w0, dt, n, enstrophy_grad(p, wT), ckpt_every=7)
        rng = np.random.default_rng(11)
        amp = np.abs(w0)
        for k in range(16):
            # directions with the initial condition's spectral support and
            # amplitudes, random phases (non-Hermitian on the ky=0 column too)
            d = amp * np.exp(2j * np.pi * rng.random(w0.shape))
            h = 1e-3 * np.linalg.norm(w0) / np.linalg.norm(d)
            D = lambda e: (Z(w0 + e * d)[0] - Z(w0 - e * d)[0]) / (2 * e)
            fd = (4.0 * D(h / 2) - D(h)) / 3.0
            an = float(np.sum(g.real * d.real + g.imag * d.imag))
            print(f"direction {k}: fd {fd:.12e} adjoint {an:.12e}", flush=True)
            assert abs(fd - an) <= 1e-6 * abs(an), (k, fd, an)


@pytest.mark.parametrize("device", ["cpu", "cuda"])
def test_autograd_function_enstrophy(device):
    import torch

    from ns2d import autograd as nag
    if device == "cuda" and not torch.cuda.is_available():
        pytest.skip("torch sees no GPU in this process")
    p = O.Params(64, 64, nu=2e-3, hyper_flag=0, hyper_pow=1.0, quirks=0)
    w0 = O.ic_random_spectrum(p, seed=8, k0=4.0)
    dt, n = 2e-3, 6
    J_ref, g_ref, _ = torch_grad(p, w0, dt, n)
    with make(p) as s:
        w = torch.tensor(w0, device=device, requires_grad=True)
        wT = nag.rk4cn(w, s, dt, n, 0)
        J = nag.enstrophy(wT, p.Nx, p.Ny)
        J.backward()
        g = w.grad.detach().cpu().numpy()
    assert abs(float(J.detach()) - J_ref) < 1e-12 * abs(J_ref)
    assert np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref) < 1e-11


def test_vjp_after_forward_steps_on_the_same_handle():
    """A handle that has stepped another state before keeps no stale
    transforms into the adjoint's forward sweep."""
    p = O.Params(64, 64, nu=1e-3, hyper_flag=0, hyper_pow=1.0, quirks=0)
    w0 = O.ic_random_spectrum(p, seed=9, k0=6.0)
    dt, n = 2e-3, 9
    J, g_ref, wT = torch_grad(p, w0, dt, n)
    with make(p) as s:
        s.set_state(O.ic_random_spectrum(p, seed=10, k0=3.0))
        s.spectral_step(dt, 5)
        g = s.vjp(w0, dt, n, enstrophy_grad(p, wT), ckpt_every=4)
    assert np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref) < 1e-11


def test_vjp_graph_replayed_segments(monkeypatch):
    """Eight checkpoint segments: the middle ones run as replayed hipGraphs
    (one per buffer parity); equal to the direct launches bit for bit and to
    torch autograd."""
    p = O.Params(64, 64, nu=1e-3, hyper_flag=0, hyper_pow=1.0, quirks=0)
    w0 = O.ic_random_spectrum(p, seed=12, k0=5.0)
    dt, n = 2e-3, 30
    J, g_ref, wT = torch_grad(p, w0, dt, n)
    lam = enstrophy_grad(p, wT)
    with make(p) as s:
        g1 = s.vjp(w0, dt, n, lam, ckpt_every=4)
        g1b = s.vjp(w0, dt, n, lam, ckpt_every=4)      # cached graphs
        g3 = s.vjp(w0, dt, n, lam, ckpt_every=3)       # odd segment length
    monkeypatch.setenv("NS2D_GRAPH", "0")
    with make(p) as s:
        g2 = s.vjp(w0, dt, n, lam, ckpt_every=4)
    np.testing.assert_array_equal(g1, g2)
    np.testing.assert_array_equal(g1, g1b)
    assert np.linalg.norm(g1 - g_ref) / np.linalg.norm(g_ref) < 1e-11
    assert np.linalg.norm(g3 - g_ref) / np.linalg.norm(g_ref) < 1e-11


@pytest.mark.parametrize("ckpt", [4, 3, 7])
def test_vjp_overlapped_recompute(monkeypatch, ckpt):
    """Two tapes (the recompute of segment s-1 on a second stream beside the
    adjoint sweep of segment s) against one tape in sequential order, with and
    without graph replay: bit for bit, and against torch autograd."""
    p = O.Params(64, 64, nu=1e-3, hyper_flag=0, hyper_pow=1.0, quirks=0)
    w0 = O.ic_random_spectrum(p, seed=13, k0=5.0)
    dt, n = 2e-3, 29
    J, g_ref, wT = torch_grad(p, w0, dt, n)
    lam = enstrophy_grad(p, wT)
    got = {}
    for ov in ("1", "0"):
        for gr in ("1", "0"):
            monkeypatch.setenv("NS2D_ADJ_OVERLAP", ov)
            monkeypatch.setenv("NS2D_GRAPH", gr)
            with make(p) as s:
                got[ov, gr] = s.vjp(w0, dt, n, lam, ckpt_every=ckpt)
                # a second call reuses tapes, streams and cached graphs
                np.testing.assert_array_equal(s.vjp(w0, dt, n, lam, ckpt_every=ckpt), got[ov, gr])
    for k, g in got.items():
        np.testing.assert_array_equal(g, got["0", "0"], err_msg=str(k))
    assert np.linalg.norm(got["1", "1"] - g_ref) / np.linalg.norm(g_ref) < 1e-11


def test_vjp_rejects_reference_conjugacy_quirk():
    import ns2d
    p = O.Params(32, 32, nu=1e-3)
    with make(p) as s:
        with pytest.raises(ns2d.NS2DError):
            s.vjp(np.zeros((32, 17), complex), 1e-3, 2, np.zeros((32, 17), complex))
