"""GPU parity: the HIP path, called through the C ABI, against the oracle.

Tolerances (FP64): single RHS evaluations and transforms agree to 1e-12
relative L2; trajectories to 1e-10 relative L2 of w_hat (the bar for the
Taylor-Green 1000-step run); energy / enstrophy series to 1e-12 relative.
"""
import numpy as np
import pytest

import ns2d_oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def solver(p: O.Params, check=1):
    import ns2d
    cfg = ns2d.Config(Nx=p.Nx, Ny=p.Ny, nu=p.nu, hyper_flag=p.hyper_flag, hyper_pow=p.hyper_pow,
                      alpha=p.alpha, drag_flag=p.drag_flag, drag_pow=p.drag_pow, filter=p.filter,
                      forcing=p.forcing, force_k=p.force_k, force_scale=p.force_scale,
                      integrator=p.integrator, quirks=p.quirks, check_every=check)
    return ns2d.Solver(cfg)


SIZES = [(8, 8), (16, 16), (32, 64), (64, 32), (128, 128), (256, 256), (512, 512), (1024, 1024)]


@pytest.mark.parametrize("Nx,Ny", SIZES)
def test_r2c_matches_numpy(Nx, Ny):
    print(f"r2c {Nx}x{Ny}", flush=True)
    p = O.Params(Nx, Ny)
    rng = np.random.default_rng(Nx * 7 + Ny)
    w = rng.standard_normal((Nx, Ny))
    with solver(p) as s:
        got = s.r2c(w)
    assert rel(got, O.r2c(w)) < 1e-13


@pytest.mark.parametrize("Nx,Ny", SIZES)
@pytest.mark.parametrize("filt", ["23", "houli_ref"])
def test_rhs_matches_oracle(Nx, Ny, filt):
    print(f"rhs {Nx}x{Ny} {filt}", flush=True)
    p = O.Params(Nx, Ny, nu=1e-3, filter=filt)
    w = O.ic_random_spectrum(p, seed=Nx + 3 * Ny, k0=max(2.0, Nx / 16))
    # make the ky=0 column non-Hermitian to exercise the c2r projection
    w[Nx - 1, 0] += 0.3 * (1 + 1j) * np.abs(w).max()
    K_ref = O.nonlinear_rhs(w, p)
    st = w.copy()
    O.force_conjugacy(w, st)
    with solver(p) as s:
        s.set_state(w)
        K = s.nonlinear_rhs(w)
        st_gpu = s.get_state()
    assert rel(K, K_ref) < 1e-12
    assert rel(st_gpu, st) < 1e-15


@pytest.mark.parametrize("Nx,Ny", [(16, 16), (64, 64), (128, 128), (256, 128)])
def test_rhs_kolmogorov_forcing(Nx, Ny):
    print(f"rhs kolm {Nx}x{Ny}", flush=True)
    p = O.Params(Nx, Ny, nu=1e-3, forcing="KOLM", force_k=4, force_scale=1.0)
    w = O.ic_random_spectrum(p, seed=5)
    with solver(p) as s:
        K = s.nonlinear_rhs(w)
    assert rel(K, O.nonlinear_rhs(w, p)) < 1e-12


@pytest.mark.parametrize("Nx,Ny", [(16, 16), (64, 64), (128, 128), (512, 256)])
def test_max_and_real_vorticity(Nx, Ny):
    print(f"maxvort {Nx}x{Ny}", flush=True)
    p = O.Params(Nx, Ny)
    w = O.ic_random_spectrum(p, seed=11)
    w[3, 0] += 0.1 * np.abs(w).max()
    with solver(p) as s:
        s.set_state(w)
        m = s.max_vort()
        wr = s.real_vorticity()
    ref = O.real_vorticity(w, p)
    assert rel(wr, ref) < 1e-13
    assert abs(m - O.max_vort(w, p)) <= 1e-13 * abs(m)


@pytest.mark.parametrize("integ", ["RK4CN", "AB4"])
@pytest.mark.parametrize("Nx,Ny,filt", [(32, 32, "23"), (64, 64, "houli_ref"), (128, 64, "23")])
def test_steps_match_oracle(Nx, Ny, filt, integ):
    print(f"steps {Nx}x{Ny} {filt} {integ}", flush=True)
    p = O.Params(Nx, Ny, nu=2e-3, hyper_flag=0, hyper_pow=1.0, filter=filt, integrator=integ)
    w0 = O.ic_random_spectrum(p, seed=21, k0=4.0)
    st = O.State(w0, p)
    O.step(st, 1e-3, 12)
    with solver(p) as s:
        s.set_state(w0)
        s.spectral_step(1e-3, 5)
        s.spectral_step(1e-3, 7)
        got = s.get_state()
        info = s.info()
    assert info["iters"] == 12
    assert info["hyper_flag"] == st.hyper_flag
    assert rel(got, st.w_hat) < 1e-12


def test_taylor_green_1000_steps():
    """Config 1: TG 128^2, nu=1e-3 (-v 1e-3 -v 0 -v 1.0), 2/3, RK4CN, dt=1e-3, 1000 steps."""
    print("tg 128 x1000", flush=True)
    p = O.Params(128, 128, nu=1e-3, hyper_flag=0, hyper_pow=1.0)
    w0 = O.ic_taylor_green(p)
    st = O.State(w0, p)
    O.step(st, 1e-3, 1000)
    with solver(p) as s:
        s.set_state(w0)
        s.spectral_step(1e-3, 1000)
        got = s.get_state()
        wr = s.real_vorticity()
    assert rel(got, st.w_hat) < 1e-10
    x, y = O.collocation(128, 128)
    exact = 2.0 * np.cos(x) * np.cos(y) * np.exp(-2.0 * 1e-3 * 1.0)
    assert np.abs(wr - exact).max() < 1e-12


@pytest.mark.parametrize("with_flux", [False, True])
@pytest.mark.parametrize("Nx,Ny,forcing", [(64, 64, "NONE"), (128, 128, "KOLM"), (256, 128, "NONE")])
def test_measures_match_oracle(Nx, Ny, forcing, with_flux):
    print(f"measures {Nx}x{Ny} {forcing} flux={with_flux}", flush=True)
    p = O.Params(Nx, Ny, nu=1e-3, hyper_flag=1, hyper_pow=2.0, alpha=0.1, drag_flag=1,
                 drag_pow=-2.0, forcing=forcing, force_k=3, force_scale=1.0)
    w = O.ic_random_spectrum(p, seed=2)
    w[Nx - 2, 0] += 0.05 * np.abs(w).max()     # non-Hermitian ky=0 entry (measures see it)
    wref = w.copy()
    rhs = None
    if with_flux:
        rhs = O.nonlinear_rhs(wref, p)
        O.force_conjugacy(wref, wref)
        if forcing == "KOLM":
            rhs[0, p.force_k] -= O.forcing_value(p)
    ref = O.measures(wref, p, rhs=rhs)
    with solver(p) as s:
        s.set_state(w)
        got = s.system_measures(with_flux=with_flux)
    for k in ("energy", "enstrophy", "palinstrophy", "enrg_diss", "enst_diss", "forcing"):
        assert abs(got[k] - ref[k]) <= 1e-12 * max(abs(ref[k]), 1e-300), k
    assert got["divergence"] < 1e-20
    assert rel(got["enrg_spect"], ref["enrg_spect"]) < 1e-12
    assert rel(got["enst_spect"], ref["enst_spect"]) < 1e-12
    if with_flux:
        for k in ("enrg_flux_sbst", "enrg_diss_sbst", "enst_flux_sbst", "enst_diss_sbst",
                  "d_enrg_dt_sbst", "d_enst_dt_sbst"):
            assert abs(got[k] - ref[k]) <= 1e-10 * max(abs(ref[k]), 1e-30) + 1e-14, k
        for k in ("d_enst_dt_spect", "enst_diss_spect", "enst_flux_spect",
                  "d_enrg_dt_spect", "enrg_diss_spect", "enrg_flux_spect"):
            assert rel(got[k], ref[k]) < 1e-10, k


def _variant_cases():
    import ns2d
    try:
        ns2d.load_library()
    except OSError:
        return []
    cases = []
    for nx, ny in [(128, 32), (256, 32), (512, 32), (1024, 32), (2048, 16), (4096, 16)]:
        for i, _ in enumerate(ns2d.kpass_variants(nx)):
            cases.append((nx, ny, i, 0))
    for ny, nx in [(32, 16), (64, 16), (128, 16), (256, 16), (512, 16), (1024, 16), (2048, 16), (4096, 16)]:
        for j, _ in enumerate(ns2d.rpass_variants(ny)):
            cases.append((nx, ny, 0, j))
    return cases


def _pair_cases():
    import ns2d
    try:
        ns2d.load_library()
    except OSError:
        return []
    cases = []
    for nx, ny in [(64, 32), (256, 64), (512, 128), (1024, 32), (2048, 16), (4096, 8), (128, 8)]:
        vs = ns2d.kpass_variants(nx)
        packed = [i for i, v in enumerate(vs) if v["packed_pair"]]
        unpacked = [i for i, v in enumerate(vs) if not v["packed_pair"]]
        cases.append((nx, ny, packed[0], unpacked[0]))
        # the shape that keeps the pair's second column in registers
        regs = [i for i, v in enumerate(vs) if v["pair_in_registers"]]
        if regs:
            cases.append((nx, ny, regs[0], unpacked[0]))
    return cases


@pytest.mark.parametrize("Nx,Ny,packed,unpacked", _pair_cases())
@pytest.mark.parametrize("filt", ["none", "houli"])
def test_packed_pair_column_pass(Nx, Ny, packed, unpacked, filt):
    """Column-pass shapes that carry ky = 0 and ky = Ny/2 as one packed
    complex line, against the oracle and an unpacked shape.  With no filter
    (or the smooth Hou-Li one) the Nyquist column holds data, and the ky = 0
    column is made non-Hermitian.  Packing shares rounding between the two
    columns (each column's error scales with the larger of the two), so the
    Nyquist column here is 1 % of the largest mode and the packed/unpacked
    agreement bound is 1e-12."""
    print(f"packed pair {Nx}x{Ny} {filt}", flush=True)
    p = O.Params(Nx, Ny, nu=1e-3, hyper_flag=0, hyper_pow=1.0, filter=filt)
    rng = np.random.default_rng(Nx + 3 * Ny)
    w0 = O.ic_random_spectrum(p, seed=Nx + Ny, k0=3.0)
    scale = np.abs(w0).max()
    w0[:, -1] = 0.01 * scale * (rng.standard_normal(Nx) + 1j * rng.standard_normal(Nx))
    w0[Nx - 1, 0] += 0.2 * scale
    w0[0, 0] = 0.0
    st = O.State(w0, p)
    O.step(st, 1e-4, 3)
    x = rng.standard_normal((Nx, Ny))
    out = {}
    for var in (packed, unpacked):
        with solver(p) as s:
            s.set_kpass_variant(var)
            s.set_state(w0)
            s.spectral_step(1e-4, 3)
            out[var] = (s.get_state(), s.nonlinear_rhs(w0), s.r2c(x), s.c2r(w0))
    for var in (packed, unpacked):
        got, K, xf, wr = out[var]
        assert rel(got, st.w_hat) < 1e-12
        assert rel(K, O.nonlinear_rhs(w0, p)) < 1e-12
        assert rel(xf, O.r2c(x)) < 1e-13
        assert rel(wr, O.c2r(w0, Nx, Ny)) < 1e-13
    for a, b in zip(out[packed], out[unpacked]):
        assert rel(a, b) < 1e-12


@pytest.mark.parametrize("Nx,Ny,var,rvar", _variant_cases())
@pytest.mark.parametrize("integ", ["RK4CN", "AB4"])
def test_pass_variants_match_oracle(Nx, Ny, var, rvar, integ):
    """Every compiled column-pass and row-pass shape gives the oracle's trajectory."""
    print(f"variant {Nx}x{Ny} k#{var} r#{rvar} {integ}", flush=True)
    p = O.Params(Nx, Ny, nu=2e-3, hyper_flag=1, hyper_pow=2.0, alpha=0.05, drag_flag=1, drag_pow=0.0,
                 filter="23", integrator=integ, forcing="KOLM", force_k=2, force_scale=0.5)
    w0 = O.ic_random_spectrum(p, seed=Nx + var, k0=3.0)
    w0[Nx - 1, 0] += 0.2 * np.abs(w0).max()
    st = O.State(w0, p)
    O.step(st, 1e-3, 6)
    with solver(p) as s:
        s.set_kpass_variant(var)
        s.set_rpass_variant(rvar)
        s.set_state(w0)
        s.spectral_step(1e-3, 6)
        got = s.get_state()
        mv = s.max_vort()
        cr = s.cfl_rate()
        K = s.nonlinear_rhs(w0)
        wr = s.c2r(w0)
    assert rel(got, st.w_hat) < 1e-12
    assert rel(K, O.nonlinear_rhs(w0, p)) < 1e-12
    # GetTimestep's convective rate, reduced in the row pass (RP_CFL): with
    # nu = 0 and cfl = 1 the oracle's CFL step is 1 / rate
    cr_ref = 1.0 / O.cfl_timestep(st.w_hat, O.Params(Nx, Ny, nu=0.0), 1.0)
    assert abs(cr - cr_ref) <= 1e-12 * cr_ref
    # the row pass's vorticity transform (max|w| check, c2r)
    assert abs(mv - O.max_vort(st.w_hat, p)) <= 1e-12 * abs(mv)
    assert rel(wr, O.c2r(w0, Nx, Ny)) < 1e-13


def test_flux_terms_are_inviscid_rates_of_change_on_gpu():
    """The GPU's flux subsets and flux spectra (ns2d_measures with_flux) at
    256^2, nu = 0: d/dt of the subset enstrophy/energy and of the spectra
    along the GPU's own RK4CN trajectory (Richardson-combined forward
    differences), to 1e-6 relative."""
    p = O.Params(256, 256, nu=0.0, hyper_flag=0, hyper_pow=1.0, filter="23", quirks=0)
    w0 = O.ic_random_spectrum(p, seed=21, k0=30.0)
    h = 2e-6
    kx, ky = O.wavenumbers(256, 256)
    k2 = (kx * kx + ky * ky).astype(np.float64)
    c = np.full((1, 129), 2.0)
    c[0, 0] = c[0, -1] = 1.0
    sb = (k2 > 0) & (k2 < 64.0 ** 2)
    C = 4.0 * np.pi ** 2 * 0.5 / float(256 * 256) ** 2

    def sums(w, m):
        a2 = np.abs(w) ** 2
        return np.array([np.sum(np.where(sb, c * a2, 0.0)) * C,
                         np.sum(np.where(sb, c * a2 / np.where(k2 > 0, k2, 1.0), 0.0)) * C]), m

    with solver(p, check=0) as s:
        s.set_state(w0)
        m0 = s.system_measures(with_flux=True)
        base = sums(w0, m0)

        def rate(dt):
            s.set_state(w0)
            s.spectral_step(dt, 1)
            w = s.get_state()
            m = s.system_measures()
            x = sums(w, m)
            return ((x[0] - base[0]) / dt, (m["enst_spect"] - m0["enst_spect"]) / dt,
                    (m["enrg_spect"] - m0["enrg_spect"]) / dt)

        a, b = rate(h / 2), rate(h)
        sub = 2 * a[0] - b[0]
        zs = 2 * a[1] - b[1]
        es = 2 * a[2] - b[2]
    assert abs(sub[0] - m0["enst_flux_sbst"]) <= 1e-6 * abs(sub[0])
    assert abs(sub[1] - m0["enrg_flux_sbst"]) <= 1e-6 * abs(sub[1])
    assert np.linalg.norm(zs - m0["enst_flux_spect"]) <= 1e-6 * np.linalg.norm(zs)
    assert np.linalg.norm(es - m0["enrg_flux_spect"]) <= 1e-6 * np.linalg.norm(es)
