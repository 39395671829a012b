"""Pins the oracle (oracle/ns2d_oracle.py) against mathematics the reference
fixes, since the reference itself cannot be built here and ships no golden
vectors.  Each test names the oracle function it pins and the plausible
mistake it would catch.
"""
import math

import numpy as np
import pytest

import ns2d_oracle as O


def dft2_c2r_bruteforce(X, Nx, Ny):
    """FFTW c2r semantics from the definition: complex inverse DFT along x of
    every ky column, then along y the real field sum with only the real parts
    of the ky=0 and ky=Ny/2 entries."""
    g = np.zeros((Nx, Ny // 2 + 1), dtype=complex)
    for x in range(Nx):
        for kxi in range(Nx):
            g[x] += X[kxi] * np.exp(2j * np.pi * kxi * x / Nx)
    out = np.zeros((Nx, Ny))
    for x in range(Nx):
        for y in range(Ny):
            s = g[x, 0].real + g[x, Ny // 2].real * (-1) ** y
            for k in range(1, Ny // 2):
                s += 2.0 * (g[x, k] * np.exp(2j * np.pi * k * y / Ny)).real
            out[x, y] = s
    return out


def test_c2r_matches_definition_including_projection():
    """c2r: x first, Re of the ky=0/Nyquist entries (a y-first or projection-free
    transform fails on the non-Hermitian input)."""
    rng = np.random.default_rng(0)
    for Nx, Ny in ((8, 8), (4, 8), (8, 4)):
        X = rng.standard_normal((Nx, Ny // 2 + 1)) + 1j * rng.standard_normal((Nx, Ny // 2 + 1))
        assert np.abs(O.c2r(X, Nx, Ny) - dft2_c2r_bruteforce(X, Nx, Ny)).max() < 1e-12


def test_r2c_matches_definition():
    rng = np.random.default_rng(1)
    Nx, Ny = 8, 8
    w = rng.standard_normal((Nx, Ny))
    ref = np.zeros((Nx, Ny // 2 + 1), dtype=complex)
    for kx in range(Nx):
        for ky in range(Ny // 2 + 1):
            for x in range(Nx):
                for y in range(Ny):
                    ref[kx, ky] += w[x, y] * np.exp(-2j * np.pi * (kx * x / Nx + ky * y / Ny))
    assert np.abs(O.r2c(w) - ref).max() < 1e-12


def test_wavenumbers_layout():
    kx, ky = O.wavenumbers(8, 6)
    assert kx[:, 0].tolist() == [0, 1, 2, 3, 4, -3, -2, -1]
    assert ky[0].tolist() == [0, 1, 2, 3]


@pytest.mark.parametrize("N", [16, 32])
def test_rhs_analytic_jacobian(N):
    """w = a cos x + b cos 2y: psi = a cos x + b/4 cos 2y, u = psi_y, v = -psi_x,
    u.grad w = -(3/2) a b sin x sin 2y.  Pins the + sign (reference Q1), the
    1/|k|^2, kx<->ky placement and the (1/(NxNy))^2 normalisation."""
    a, b = 0.7, -1.3
    p = O.Params(N, N, nu=0.0, filter="none")
    x, y = O.collocation(N, N)
    w_hat = O.r2c(a * np.cos(x) + b * np.cos(2 * y))
    K = O.nonlinear_rhs(w_hat, p)
    ref = O.r2c(-1.5 * a * b * np.sin(x) * np.sin(2 * y))
    assert np.abs(K - ref).max() < 1e-12 * N * N


def test_rhs_taylor_green_vanishes():
    """Taylor-Green is a steady Euler solution: u.grad w = 0."""
    p = O.Params(64, 64)
    K = O.nonlinear_rhs(O.ic_taylor_green(p), p)
    assert np.abs(K).max() < 1e-9


@pytest.mark.parametrize("N", [32, 64])
def test_rhs_conserves_energy_and_enstrophy(N):
    """2/3-masked Galerkin Jacobian: sum c Re(conj(w)K)/k^2 = sum c Re(conj(w)K) = 0
    for a field inside the 2/3 disk (dropping the mask or the 1/k^2 breaks it)."""
    p = O.Params(N, N, nu=0.0)
    w = O.ic_random_spectrum(p, seed=4, k0=N / 8)
    K = O.nonlinear_rhs(w, p)
    kx, ky = O.wavenumbers(N, N)
    k2 = (kx * kx + ky * ky).astype(float)
    c = np.full((1, N // 2 + 1), 2.0)
    c[0, 0] = c[0, -1] = 1.0
    dz = np.sum(c * (np.conj(w) * K).real)
    de = np.sum(np.where(k2 > 0, c * (np.conj(w) * K).real / np.where(k2 > 0, k2, 1), 0))
    scale = np.sum(c * np.abs(w) * np.abs(K))
    assert abs(dz) < 1e-12 * scale
    assert abs(de) < 1e-12 * scale


def test_dealias_23_bruteforce():
    for N in (8, 12, 16, 32):
        p = O.Params(N, N)
        f = O.dealias_factor(p)
        kx, ky = O.wavenumbers(N, N)
        kc = int(N / 3.0)
        for i in range(N):
            for j in range(N // 2 + 1):
                want = 0.0 if kx[i, 0] ** 2 + ky[0, j] ** 2 >= kc * kc else 1.0
                assert f[i, j] == want


def test_dealias_kolmogorov_kills_kx0():
    p = O.Params(16, 16, forcing="KOLM", force_k=2)
    f = O.dealias_factor(p)
    assert np.all(f[0, :] == 0.0)
    q = O.Params(16, 16, forcing="KOLM", force_k=2, quirks=O.QUIRKS_REF & ~O.QUIRK_KOLM_KX0_KILL)
    assert O.dealias_factor(q)[0, 1] == 1.0


def test_houli_integer_division_form():
    """Reference Hou-Li (solver.c:1088): integer k/(N/2) -> 1 everywhere except the
    Nyquist row and column (exp(-36)) and their corner (0)."""
    N = 16
    f = O.dealias_factor(O.Params(N, N, filter="houli_ref"))
    want = np.ones((N, N // 2 + 1))
    want[N // 2, :] = math.exp(-36.0)
    want[:, N // 2] = math.exp(-36.0)
    want[N // 2, N // 2] = 0.0
    assert np.array_equal(f, want)
    g = O.dealias_factor(O.Params(N, N, filter="houli"))
    assert g[0, 0] == 1.0 and abs(g[N // 2, 0] - math.exp(-36.0)) < 1e-30 and 0.9 < g[3, 3] <= 1.0


def test_force_conjugacy_indices():
    for N in (4, 8):
        rng = np.random.default_rng(N)
        arg = rng.standard_normal((N, 3)) + 1j * rng.standard_normal((N, 3))
        state = np.zeros_like(arg)
        O.force_conjugacy(arg, state)
        for i in range(N):
            kx = i if i <= N // 2 else i - N
            if kx < 0:
                assert state[i, 0] == np.conj(arg[-kx, 0])
            else:
                assert state[i, 0] == 0
        assert np.all(state[:, 1:] == 0)


def test_taylor_green_decay_closed_form():
    """RK4CN with nu k^2 (-v nu -v 0 -v 1.0) on Taylor-Green, whose nonlinear term
    vanishes: the scheme's closed form is w_n = w_0 ((2-D)/(2+D))^n with D = dt nu |k|^2
    (a missing 2 in the CN factor, a wrong viscous power or a dropped dt fails),
    and w(t) = 2 cos x cos y exp(-2 nu t) up to the CN truncation D^3/12 per step."""
    dt, nu, n = 1e-3, 1e-2, 200
    p = O.Params(32, 32, nu=nu, hyper_flag=0, hyper_pow=1.0)
    st = O.State(O.ic_taylor_green(p), p)
    O.step(st, dt, n)
    x, y = O.collocation(32, 32)
    D = dt * nu * 2.0
    w_cn = 2 * np.cos(x) * np.cos(y) * ((2 - D) / (2 + D)) ** n
    got = O.real_vorticity(st.w_hat, p)
    assert np.abs(got - w_cn).max() < 1e-14
    exact = 2 * np.cos(x) * np.cos(y) * np.exp(-2 * nu * n * dt)
    assert np.abs(got - exact).max() < 2 * n * D ** 3 / 12 + 1e-14
    # Q3: with the drag flag off the reference applies nu*pow(k^2, hyper_pow)
    q = O.Params(32, 32, nu=nu, hyper_flag=0, hyper_pow=2.0)
    st2 = O.State(O.ic_taylor_green(q), q)
    O.step(st2, dt, n)
    D2 = dt * nu * 4.0
    w_cn2 = 2 * np.cos(x) * np.cos(y) * ((2 - D2) / (2 + D2)) ** n
    assert np.abs(O.real_vorticity(st2.w_hat, q) - w_cn2).max() < 1e-14
    assert st2.hyper_flag == 1


def _order_ratio(integ, quirks):
    N = 32
    p = O.Params(N, N, nu=0.0, integrator=integ, quirks=quirks)
    w0 = O.ic_random_spectrum(p, seed=9, k0=3.0, energy=0.5)
    T = 0.4

    def run(dt):
        st = O.State(w0, p)
        O.step(st, dt, int(round(T / dt)))
        return st.w_hat

    ref = run(T / 640)
    return np.linalg.norm(run(T / 40) - ref) / np.linalg.norm(run(T / 80) - ref)


@pytest.mark.parametrize("integ", ["RK4CN", "AB4"])
def test_fourth_order_in_time(integ):
    """Without the reference quirks and with nu = 0 (isolating the explicit
    part) halving dt divides the error by ~16: a wrong Butcher weight, stage
    time or AB4 coefficient drops the order."""
    r = _order_ratio(integ, 0)
    assert 13.0 < r < 19.0, r


def test_euler_build_is_rk4cn_without_viscosity():
    """The reference built with -D__EULER (inviscid) takes the explicit update
    w + (dt/6)(K1 + 2K2 + 2K3 + K4) (solver.c:827-830) after the same four
    stages; the RK4CN update with nu = alpha = 0 (D = 0: factors 1 and dt)
    is that step up to the association of the sums, so Euler runs are RK4CN
    runs with zero viscosity and drag."""
    N = 32
    p = O.Params(N, N, nu=0.0, hyper_flag=0, hyper_pow=1.0, filter="23")
    w0 = O.ic_random_spectrum(p, seed=11, k0=4.0, energy=0.5)
    dt = 1e-3
    st = O.State(w0, p)
    O.rk4cn_step(st, dt)
    # the explicit update from the same stages (quirks as in rk4cn_step)
    st2 = O.State(w0, p)
    w = st2.w_hat
    K1 = O.rhs_with_side_effect(w, w, p)
    K2 = O.rhs_with_side_effect(w + (dt * 0.5) * K1, w, p)
    K3 = O.rhs_with_side_effect(w + (dt * 0.5) * K2, w, p)
    K4 = O.rhs_with_side_effect(w + dt * K3, w, p)
    euler = w + (dt / 6.0) * (K1 + 2.0 * K2 + 2.0 * K3 + K4)
    euler[0, 0] = 0.0
    err = np.linalg.norm(st.w_hat - euler) / np.linalg.norm(euler)
    assert err < 1e-15, err
    assert np.linalg.norm(euler - w0) > 1e-6 * np.linalg.norm(w0)   # the step does something


@pytest.mark.parametrize("integ,lo,hi", [("RK4CN", 1.7, 2.4), ("AB4", 3.4, 4.6)])
def test_reference_conjugacy_write_lowers_order(integ, lo, hi):
    """With the reference's ForceConjugacy write into the global w_hat (Q2,
    solver.c:1031/2101), the ky=0, kx<0 entries pick up conj(stage input) in the
    middle of every step: RK4CN becomes first order (ratio ~2) and AB4 second
    order (ratio ~4).  An oracle that skipped or misplaced the write would
    show ~16 here."""
    r = _order_ratio(integ, O.QUIRKS_REF)
    assert lo < r < hi, r


def test_measures_parseval():
    """E, Z, P from the spectral sums equal the real-space integrals
    (1/2) int |u|^2, (1/2) int w^2, (1/2) int |grad w|^2 over [0, 2pi)^2, and the
    spectra sum to the totals."""
    N = 32
    p = O.Params(N, N, nu=1e-3, quirks=O.QUIRKS_REF & ~O.QUIRK_ENST_DISS)
    w = O.ic_random_spectrum(p, seed=3, k0=4.0)
    m = O.measures(w, p)
    kx, ky = O.wavenumbers(N, N)
    k2 = kx * kx + ky * ky
    inv = np.where(k2 > 0, 1.0 / np.where(k2 > 0, k2, 1), 0.0)
    nrm = 1.0 / (N * N)
    u = O.c2r(1j * ky * inv * w, N, N) * nrm
    v = O.c2r(-1j * kx * inv * w, N, N) * nrm
    wr = O.c2r(w, N, N) * nrm
    wx = O.c2r(1j * kx * w, N, N) * nrm
    wy = O.c2r(1j * ky * w, N, N) * nrm
    area = 4 * math.pi ** 2
    assert abs(m["energy"] - 0.5 * area * np.mean(u * u + v * v)) < 1e-12 * m["energy"]
    assert abs(m["enstrophy"] - 0.5 * area * np.mean(wr * wr)) < 1e-12 * m["enstrophy"]
    assert abs(m["palinstrophy"] - 0.5 * area * np.mean(wx * wx + wy * wy)) < 1e-12 * m["palinstrophy"]
    assert abs(m["enrg_spect"].sum() - m["energy"]) < 1e-12 * m["energy"]
    assert abs(m["enst_spect"].sum() - m["enstrophy"]) < 1e-12 * m["enstrophy"]
    # without Q7 the enstrophy dissipation is 2 nu P for nu k^2 viscosity
    assert abs(m["enst_diss"] - 2 * p.nu * m["palinstrophy"]) < 1e-12 * m["enst_diss"]
    assert abs(m["enrg_diss"] - 2 * p.nu * m["enstrophy"]) < 1e-12 * m["enrg_diss"]


def test_n_spect():
    assert O.n_spect(128, 128) == 92
    assert O.n_spect(1024, 1024) == 725
    assert O.n_spect(4096, 4096) == 2897


def test_ab4_startup_stores_k1():
    p = O.Params(16, 16, nu=1e-3, integrator="AB4")
    w0 = O.ic_random_spectrum(p, seed=2, k0=3.0)
    st = O.State(w0, p)
    K1 = O.nonlinear_rhs(w0, p)
    O.step(st, 1e-3, 1)
    assert np.allclose(st.ab_hist[0], K1, rtol=0, atol=1e-14 * np.abs(K1).max())


def test_kolmogorov_forcing_value():
    p = O.Params(16, 16, forcing="KOLM", force_k=4, force_scale=1.0)
    assert O.forcing_value(p) == -2.0
    q = O.Params(16, 16, forcing="KOLM", force_k=4, force_scale=1.0,
                 quirks=O.QUIRKS_REF & ~O.QUIRK_KOLM_UNSCALED)
    assert O.forcing_value(q) == -2.0 * 256
    w = np.zeros((16, 9), dtype=complex)
    K = O.nonlinear_rhs(w, p)
    assert K[0, 4] == -2.0 and np.count_nonzero(K) == 1


# ---------------------------------------------------------------------------
# flux diagnostics (sys_msr.c:125-133, 195-305, 363-372) and the CFL step
# (solver.c:2381-2450), pinned by what the dynamics fixes
# ---------------------------------------------------------------------------
def _subset_sums(w, p, lim=64.0):
    """Z and E restricted to the flux subset 0 <= |k|^2 < lim^2 (data_types.h:174-175),
    with the half-spectrum weights and normalisation of the totals."""
    kx, ky = O.wavenumbers(p.Nx, p.Ny)
    k2 = (kx * kx + ky * ky).astype(np.float64)
    c = np.full((1, p.Ny // 2 + 1), 2.0)
    c[0, 0] = c[0, -1] = 1.0
    sb = (k2 > 0) & (k2 < lim * lim)
    a2 = np.abs(w) ** 2
    C = 4.0 * np.pi ** 2 * 0.5 / float(p.Nx * p.Ny) ** 2
    return (float(np.sum(np.where(sb, c * a2, 0.0))) * C,
            float(np.sum(np.where(sb, c * a2 / np.where(k2 > 0, k2, 1.0), 0.0))) * C)


def _richardson_rate(f, w0, p, h):
    """d f(w(t))/dt at t = 0 along the inviscid RK4CN trajectory: two forward
    differences (steps h and h/2), Richardson-combined (error O(h^2))."""
    def D(dt):
        st = O.State(w0, p)
        O.step(st, dt, 1)
        return (f(st.w_hat) - f(w0)) / dt
    return 2.0 * D(h / 2) - D(h)


def test_flux_subset_and_spectra_are_the_inviscid_rates_of_change():
    """At 256^2 the 2/3 disk (|k| < 85) extends past the flux subset's
    UPR_SBST_LIM = 64, so energy and enstrophy flow across its edge.  With
    nu = 0 and no forcing, dw/dt = K = NonlinearRHSBatch(w), so the subset
    fluxes are d/dt of the subset enstrophy/energy and the flux spectra are
    d/dt of the enstrophy/energy spectra; the time derivatives come from the
    oracle's own time stepping."""
    p = O.Params(256, 256, nu=0.0, hyper_flag=0, hyper_pow=1.0, filter="23", quirks=0)
    w0 = O.ic_random_spectrum(p, seed=21, k0=30.0)
    K = O.nonlinear_rhs(w0, p)
    m = O.measures(w0, p, rhs=K)
    h = 2e-6
    z_rate = _richardson_rate(lambda w: _subset_sums(w, p)[0], w0, p, h)
    e_rate = _richardson_rate(lambda w: _subset_sums(w, p)[1], w0, p, h)
    assert abs(m["enst_flux_sbst"]) > 1e-3 * abs(z_rate) and abs(z_rate) > 0
    assert abs(z_rate - m["enst_flux_sbst"]) <= 1e-6 * abs(z_rate)
    assert abs(e_rate - m["enrg_flux_sbst"]) <= 1e-6 * abs(e_rate)
    zs = _richardson_rate(lambda w: O.measures(w, p)["enst_spect"], w0, p, h)
    es = _richardson_rate(lambda w: O.measures(w, p)["enrg_spect"], w0, p, h)
    assert np.linalg.norm(zs - m["enst_flux_spect"]) <= 1e-6 * np.linalg.norm(zs)
    assert np.linalg.norm(es - m["enrg_flux_spect"]) <= 1e-6 * np.linalg.norm(es)
    # without viscosity the d/dt series are the unnormalised flux sums
    C = 4.0 * np.pi ** 2 * 0.5 / float(p.Nx * p.Ny) ** 2
    assert abs(m["d_enst_dt_sbst"] * C - m["enst_flux_sbst"]) <= 1e-12 * abs(m["enst_flux_sbst"])
    assert abs(m["d_enrg_dt_sbst"] * C - m["enrg_flux_sbst"]) <= 1e-12 * abs(m["enrg_flux_sbst"])


def test_flux_vanishes_on_the_full_disk_and_dissipation_subsets():
    """At 64^2 the 2/3 disk (|k| < 21) lies inside the subset, so the subset
    fluxes are the total fluxes, zero by the Galerkin conservation of energy
    and enstrophy; every flux spectrum sums to zero.  The dissipation subsets
    are then 2 nu P and 2 nu Z (enst_diss with Q7 off, enrg_diss), and the
    d/dt series are flux + dissipation in the reference's normalisation."""
    p = O.Params(64, 64, nu=3e-3, hyper_flag=0, hyper_pow=1.0, filter="23", quirks=0)
    w0 = O.ic_random_spectrum(p, seed=4, k0=8.0)
    K = O.nonlinear_rhs(w0, p)
    m = O.measures(w0, p, rhs=K)
    kx, ky = O.wavenumbers(p.Nx, p.Ny)
    k2 = (kx * kx + ky * ky).astype(np.float64)
    C = 4.0 * np.pi ** 2 * 0.5 / float(p.Nx * p.Ny) ** 2
    scale_z = 2.0 * C * float(np.sum(np.abs(w0) * np.abs(K)))
    scale_e = 2.0 * C * float(np.sum(np.abs(w0) * np.abs(K) / np.where(k2 > 0, k2, 1.0)))
    assert abs(m["enst_flux_sbst"]) <= 1e-13 * scale_z
    assert abs(m["enrg_flux_sbst"]) <= 1e-13 * scale_e
    assert abs(np.sum(m["enst_flux_spect"])) <= 1e-13 * scale_z
    assert abs(np.sum(m["enrg_flux_spect"])) <= 1e-13 * scale_e
    assert np.max(np.abs(m["enst_flux_spect"])) > 1e-6 * scale_z
    np.testing.assert_allclose(m["enst_diss_sbst"], 2.0 * p.nu * m["palinstrophy"], rtol=1e-13)
    np.testing.assert_allclose(m["enrg_diss_sbst"], 2.0 * p.nu * m["enstrophy"], rtol=1e-13)
    np.testing.assert_allclose(m["enst_diss_sbst"], m["enst_diss"], rtol=1e-13)
    np.testing.assert_allclose(m["enrg_diss_sbst"], m["enrg_diss"], rtol=1e-13)
    np.testing.assert_allclose(C * m["d_enst_dt_sbst"], m["enst_flux_sbst"] + 0.5 * m["enst_diss_sbst"],
                               rtol=1e-12, atol=1e-12 * scale_z)
    np.testing.assert_allclose(C * m["d_enrg_dt_sbst"], m["enrg_flux_sbst"] + 0.5 * m["enrg_diss_sbst"],
                               rtol=1e-12, atol=1e-12 * scale_e)
    np.testing.assert_allclose(m["d_enst_dt_spect"], m["enst_flux_spect"] + m["enst_diss_spect"], rtol=1e-13,
                               atol=1e-13 * scale_z)
    np.testing.assert_allclose(m["enst_diss_spect"].sum(), 0.5 * m["enst_diss"], rtol=1e-13)


@pytest.mark.parametrize("N,nu", [(64, 0.0), (64, 1e-3), (128, 5e-4)])
def test_cfl_timestep_closed_form(N, nu):
    """Taylor-Green w = 2 cos x cos y: psi = cos x cos y, u = -cos x sin y,
    v = sin x cos y, so max(|u| + |v|) = max |sin(x + y)| or |sin(x - y)| = 1,
    reached on the grid (x = y = pi/4); with dx = dy = 2 pi/N the CFL step is
    cfl / (N/2 + pi^2 nu N^2 / (2 pi^2))."""
    p = O.Params(N, N, nu=nu)
    x = np.arange(N) * 2 * np.pi / N
    w = 2.0 * np.cos(x)[:, None] * np.cos(x)[None, :]
    dt = O.cfl_timestep(O.r2c(w), p, 0.9)
    dx = 2 * np.pi / N
    expect = 0.9 / (np.pi * (1.0 / dx) + np.pi ** 2 * nu * 2.0 / dx ** 2)
    assert abs(dt - expect) <= 1e-14 * expect
    # velocity of a single Fourier mode: w = cos(3 y) -> u = -sin(3y)/3, v = 0
    w = np.cos(3 * x)[None, :] * np.ones((N, 1))
    dt = O.cfl_timestep(O.r2c(w), p, 0.5)
    expect = 0.5 / (np.pi * (1.0 / 3.0) / dx + np.pi ** 2 * nu * 2.0 / dx ** 2)
    assert abs(dt - expect) <= 1e-13 * expect
