"""CPU tests of the driver's host side (libns2d_host.so): glibc rand()
restatement, command line, output naming, the reference's initial
conditions, dealiasing, mean flows and the HDF5 layout."""
import ctypes
import math
import os

import numpy as np
import pytest

import ns2d_ics_ref as R
import ns2d_oracle as O
from h5lite import File, write_double_dataset

host = pytest.importorskip("ns2d.host")

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def r2c(a):
    return np.fft.rfft2(a)


def c2r_for(Nx, Ny):
    return lambda a: np.fft.irfft2(a, s=(Nx, Ny), norm="forward")


def test_glibc_rand_restatement_matches_fixture():
    """The TYPE_3 restatement against rand() outputs recorded from glibc."""
    rows = [tuple(map(int, l.split())) for l in open(os.path.join(GOLDEN, "glibc_rand_123456789.txt"))
            if not l.startswith("#")]
    g = R.GlibcRand(R.SEED)
    vals = [g.rand() for _ in range(rows[-1][0] + 1)]
    assert all(vals[i] == v for i, v in rows)
    assert len(rows) == 2048 + 20


def test_glibc_rand_restatement_matches_libc_in_process():
    libc = ctypes.CDLL(None)
    libc.srand(12345)
    g = R.GlibcRand(12345)
    assert [libc.rand() for _ in range(5000)] == [g.rand() for _ in range(5000)]
    g0, g1 = R.GlibcRand(0), R.GlibcRand(1)   # glibc maps seed 0 to 1
    assert [g0.rand() for _ in range(10)] == [g1.rand() for _ in range(10)]


def test_cli_defaults_are_the_reference_defaults():
    d = host.defaults().as_dict()
    # utils.c:48-88
    assert d["N"] == (64, 64) and d["dt"] == 1e-4 and d["T"] == 1.0 and d["t0"] == 0.0
    assert d["u0"] == "TG_VORT" and d["forcing"] == "NONE" and d["save_every"] == 100
    assert d["nu"] == 1e-4 and d["hyper_visc_pow"] == 2.0 and d["ekmn_drag_pow"] == -2.0
    assert d["cfl_const"] == 0.9 and d["cfl_cond_flag"] == 1 and d["trans_iters_frac"] == 0.2
    assert d["output_dir"] == "./Data/Tmp/" and d["output_tag"] == "No-Tag"
    assert d["integrator"] == "RK4CN" and d["filter"] == "23" and d["quirks"] == 0x1F


def test_cli_parses_the_run_script_command():
    """The argument list run_solver.py:232-245 builds, with its repeated flags."""
    args = ("-o /tmp/out/ -n 64 -n 32 -s 0.00000 -e 1.00000 -T 0 -c 1 -c 0.500000 -h 0.001000 -h 0 "
            "-v 0.0010000000 -v 1 -v 2.0 -d 0.010000 -d 1 -d -2.0 -i DECAY_TURB_BB -t Run-A "
            "-f KOLM -f 4 -f 1.0 -p 50").split()
    d = host.parse_cli(args).as_dict()
    assert d["output_dir"] == "/tmp/out/" and d["N"] == (64, 32)
    assert d["T"] == 1.0 and d["trans_iters_flag"] == 0
    assert d["cfl_cond_flag"] == 1 and d["cfl_const"] == 0.5
    assert d["dt"] == 0.001 and d["adapt_step_flag"] == 0
    assert d["nu"] == 0.001 and d["hyper_visc_flag"] == 1 and d["hyper_visc_pow"] == 2.0
    assert d["ekmn_alpha"] == 0.01 and d["ekmn_drag_flag"] == 1 and d["ekmn_drag_pow"] == -2.0
    assert d["u0"] == "DECAY_TURB_BB" and d["output_tag"] == "Run-A"
    assert d["forcing"] == "KOLM" and d["force_k"] == 4 and d["force_scale_var"] == 1.0
    assert d["save_every"] == 50 and d["file_only"] == 0
    # a second -o selects file-only output; unknown initial conditions become NONE
    d = host.parse_cli(["-o", "/tmp/x/", "-o", "1", "-i", "NOT_AN_IC", "-I", "AB4", "-D", "HOU_LI"]).as_dict()
    assert d["file_only"] == 1 and d["u0"] == "NONE" and d["integrator"] == "AB4" and d["filter"] == "HOU_LI"


@pytest.mark.parametrize("args,text", [
    (["-n", "63"], "must be a multiple of 2"),
    (["-n", "2"], "must be greater than 2"),
    (["-s", "2.0", "-e", "1.0"], "must be greater than the initial time"),
    (["-h", "0"], "must be strictly positive"),
    (["-h", "0.1", "-h", "3"], "Incorrect Adaptive stepping flag"),
    (["-v", "-1"], "viscosity"),
    (["-v", "1", "-v", "2"], "Incorrect Hyperviscosity flag"),
    (["-v", "1", "-v", "1", "-v", "0"], "hyperviscosity power"),
    (["-d", "0.1", "-d", "1", "-d", "1.0"], "Hypodiffusivity power"),
    (["-c", "2"], "Incorrect CFL condition flag"),
    (["-T", "1", "-T", "1.5"], "Incorrect transient iterations fraction"),
    (["-z", "/nonexistent/file.ini"], "cannot be found"),
    (["-q"], "Incorrect command line flag"),
])
def test_cli_rejects_what_the_reference_rejects(args, text):
    with pytest.raises(host.HostError, match=text):
        host.parse_cli(args)


def test_output_paths_follow_the_reference_naming():
    c = host.parse_cli("-o /tmp/o/ -n 64 -n 32 -e 1.0 -h 0.001 -v 0.001 -v 0 -v 2.0 -i TG_VORT -t TG-Test".split())
    p = host.output_paths(c)
    # hdf5_funcs.c:572-586 format, this configuration
    folder = ("SIM_DATA_NAVIER_RK4CN_FULL_N[64,32]_T[0.0,0.001,1.000]_NU[0.001,0,2.0]_DRAG[0,0,-2.0]_CFL[0.90]_"
              "FORC[NONE,0,1]_u0[TG_VORT]_TAG[TG-Test]/")
    assert p["dir"] == "/tmp/o/" + folder
    assert p["main_file"] == "/tmp/o/" + folder + "Main_HDF_Data.h5"
    assert p["spectra_file"] == "/tmp/o/" + folder + "Spectra_HDF_Data.h5"
    assert p["sys_msr_file"] == "/tmp/o/" + folder + "SystemMeasures_HDF_Data.h5"
    c = host.parse_cli("-o /tmp/o/ -o 1 -n 64 -n 64 -I AB4".split())
    p = host.output_paths(c)
    tail = "_SIM[NAV-AB4-FULL]_N[64,64]_T[0.0,0.0001,1.000]_NU[0.0001,0,2.0]_DRAG[0,0,-2.0]_CFL[0.90]_FORC[NONE,0,1]_u0[TG_VORT].h5"
    assert p["main_file"] == "/tmp/o/Main_HDF_Data" + tail
    assert p["spectra_file"] == "/tmp/o/Spectra_HDF_Data" + tail
    assert p["sys_msr_file"] == p["main_file"]


SPECTRAL_ICS = ["DECAY_TURB_BB", "DECAY_TURB_NB", "DECAY_TURB_EXP", "GAUSS_DECAY_TURB", "EXTRM_ENS", "RING",
                "UNIF", "RANDOM", "TESTING", "NONE", "ZERO", "ZERO_KOLM"]
REAL_ICS = ["TG_VORT", "TG_VEL", "DOUBLE_SHEAR_LAYER", "GAUSS_BLOB", "DECAY_TURB_ALT"]


def _cfg(u0, Nx, Ny, filt="23", forcing="NONE", ranks=None):
    args = ["-n", str(Nx), "-n", str(Ny), "-i", u0, "-D", filt]
    if forcing != "NONE":
        args += ["-f", forcing, "-f", "2", "-f", "1.0"]
    if ranks is not None:
        args += ["-R", str(ranks)]
    return host.parse_cli(args)


@pytest.mark.parametrize("u0", SPECTRAL_ICS + REAL_ICS)
@pytest.mark.parametrize("Nx,Ny,filt,forcing", [(32, 32, "23", "NONE"), (64, 32, "HOU_LI", "NONE"),
                                                (32, 64, "23", "KOLM")])
def test_initial_conditions_match_restatement(u0, Nx, Ny, filt, forcing):
    got = host.initial_condition(_cfg(u0, Nx, Ny, filt, forcing), r2c=r2c, c2r=c2r_for(Nx, Ny))
    ref = R.initial_condition(u0 if u0 != "NONE" else "NONE", Nx, Ny, filt, forcing, r2c=r2c, c2r=c2r_for(Nx, Ny))
    assert got.shape == (Nx, Ny // 2 + 1)
    np.testing.assert_array_equal(got, ref)


RANDOM_ICS = ["DECAY_TURB_BB", "DECAY_TURB_NB", "DECAY_TURB_EXP", "DECAY_TURB_ALT", "GAUSS_DECAY_TURB",
              "EXTRM_ENS", "RING", "UNIF", "RANDOM", "NONE"]
# (Nx, Ny, P): equal slabs, a short last slab, an empty last rank, one row per rank
RANK_CASES = [(32, 32, 2), (32, 32, 4), (30, 16, 3), (32, 16, 6), (16, 16, 5), (24, 16, 7), (8, 16, 16)]


@pytest.mark.parametrize("u0", RANDOM_ICS)
@pytest.mark.parametrize("Nx,Ny,P", RANK_CASES)
def test_rank_count_initial_conditions_match_restatement(u0, Nx, Ny, P):
    """-R P against the restatement's P-rank form (a generator per slab kept
    across the loops, per-slab sums reduced by recursive doubling)."""
    got = host.initial_condition(_cfg(u0, Nx, Ny, ranks=P), r2c=r2c, c2r=c2r_for(Nx, Ny))
    ref = R.initial_condition(u0, Nx, Ny, r2c=r2c, c2r=c2r_for(Nx, Ny), ranks=P)
    np.testing.assert_array_equal(got, ref)


def test_rank_count_one_is_the_default_field():
    for u0 in ("DECAY_TURB_BB", "RING", "NONE"):
        a = host.initial_condition(_cfg(u0, 32, 32), r2c=r2c)
        b = host.initial_condition(_cfg(u0, 32, 32, ranks=1), r2c=r2c)
        np.testing.assert_array_equal(a, b)
    assert host.defaults().as_dict()["ic_ranks"] == 1
    assert host.parse_cli(["-R", "8"]).as_dict()["ic_ranks"] == 8
    for bad in ("0", "-2", "x", "3.5", "65537"):
        with pytest.raises(host.HostError, match="rank count"):
            host.parse_cli(["-R", bad])


def _libc_uniforms(n):
    """the first n values of (double) rand() / RAND_MAX after srand(123456789),
    from the process's own libc"""
    libc = ctypes.CDLL(None)
    libc.srand(R.SEED)
    return np.array([libc.rand() for _ in range(n)], dtype=np.float64) / float(R.RAND_MAX)


def _hou_li_kept(Nx, Ny):
    """modes the Hou-Li filter with integer division multiplies by exactly 1
    (|kx| < Nx/2, ky < Ny/2) and the ky = 0 conjugacy fix-up leaves alone"""
    return [(i, j) for i in range(Nx) for j in range(1, Ny // 2)
            if abs(i if i <= Nx // 2 else i - Nx) < Nx // 2]


@pytest.mark.parametrize("P", [1, 2, 3, 5])
def test_rank_slabs_restart_the_generator(P):
    """NONE draws an amplitude and a phase per mode (solver.c:1984-1995); on P
    ranks each slab of ceil(Nx/P) kx rows draws the seed's sequence from its
    start (every rank calls srand, solver.c:1121).  Checked against libc's own
    rand(), not the restatement."""
    Nx, Ny = 24, 16
    Nyf, rows = Ny // 2 + 1, -(-Nx // P)
    u = _libc_uniforms(2 * rows * Nyf)
    w = host.initial_condition(_cfg("NONE", Nx, Ny, "HOU_LI", ranks=P), r2c=r2c)
    for i, j in _hou_li_kept(Nx, Ny):
        n = (i % rows) * Nyf + j               # the mode's place in its rank's walk
        want = u[2 * n] * complex(math.cos(2.0 * math.pi * u[2 * n + 1]), math.sin(2.0 * math.pi * u[2 * n + 1]))
        assert abs(w[i, j] - want) <= 1e-15, (P, i, j)


@pytest.mark.parametrize("P", [1, 2, 4])
def test_rank_ring_phases_continue_after_the_slab_discards(P):
    """RING draws one discarded value per mode of its slab (solver.c:1690),
    then one phase per ring mode: rank r's k-th ring mode takes draw
    rows*Nyf + k of the seed's sequence (slabs of equal length here)."""
    Nx, Ny = 32, 32
    Nyf, rows = Ny // 2 + 1, Nx // P
    w = host.initial_condition(_cfg("RING", Nx, Ny, "HOU_LI", ranks=P), r2c=r2c)
    u = _libc_uniforms(rows * Nyf + rows * Nyf)
    for r in range(P):
        k = 0
        for i in range(r * rows, (r + 1) * rows):
            k0 = i if i <= Nx // 2 else i - Nx
            for j in range(Nyf):
                ak = math.sqrt(k0 * k0 + j * j)
                if 3.0 <= ak <= 10.0 and (j != 0 or k0 > 0):
                    ph = 2.0 * math.pi * u[rows * Nyf + k]
                    z = w[i, j] / abs(w[i, j])
                    assert abs(z - complex(math.cos(ph), math.sin(ph))) <= 1e-13, (P, i, j)
                    k += 1
        assert k > 0


def test_rank_sums_reduce_by_recursive_doubling():
    """Sums where the association decides the last bit: recursive doubling
    adds pairs first ((a0+a1)+(a2+a3)), with ranks beyond the largest power of
    two folded into pairs beforehand; a left-to-right sum gives 1.0."""
    e = 2.0 ** -53
    red = lambda P, v: R._Ranks(P, P).allreduce(v)
    assert red(4, [1.0, e, e, e]) == 1.0 + 2 * e
    assert sum([1.0, e, e, e]) == 1.0
    assert red(3, [1.0, e, e]) == 1.0                       # (a0 + a1) + a2
    assert red(6, [1.0, e, e, 0.0, e, e]) == 1.0 + 2 * e    # ((a0+a1) + (a2+a3)) + (a4+a5)
    assert red(5, [1.0, e, e, e, e]) == 1.0 + 2 * e         # ((a0+a1) + a2) + (a3+a4)
    assert red(5, [e, e, 1.0, e, e]) == 1.0 + 4 * e         # ((a0+a1) + a2) + (a3+a4): every add exact
    assert red(3, [e, 1.0, e]) == 1.0                       # (a0+a1) + a2 rounds twice
    assert red(1, [0.25]) == 0.25
    # and runs whose normalisation the association does change (few cases
    # do: a slab's sum has few terms): the C library matches the restatement
    # and not a restatement that reduced left to right
    for u0, Nx, Ny, P in (("DECAY_TURB_BB", 32, 16, 6), ("DECAY_TURB_BB", 24, 16, 7), ("RING", 32, 16, 6)):
        ref = R.initial_condition(u0, Nx, Ny, ranks=P)
        got = host.initial_condition(_cfg(u0, Nx, Ny, ranks=P), r2c=r2c)
        np.testing.assert_array_equal(got, ref)
        orig = R._Ranks.allreduce
        try:
            R._Ranks.allreduce = lambda self, part: sum(part)
            planted = R.initial_condition(u0, Nx, Ny, ranks=P)
        finally:
            R._Ranks.allreduce = orig
        assert not np.array_equal(got, planted), (u0, Nx, Ny, P)


def test_initial_conditions_properties():
    """What the mathematics fixes, independent of the restatement."""
    Nx = Ny = 64
    # TG_VORT = 2 cos x cos y: four modes of magnitude Nx*Ny/2
    w = host.initial_condition(_cfg("TG_VORT", Nx, Ny), r2c=r2c)
    nz = np.argwhere(np.abs(w) > 1e-9)
    assert sorted(map(tuple, nz)) == [(1, 1), (Nx - 1, 1)]
    np.testing.assert_allclose(np.abs(w[1, 1]), Nx * Ny / 2, rtol=1e-14)
    # TG_VEL: u = cos x sin y, v = -sin x cos y has vorticity
    # dv/dx - du/dy = -2 cos x cos y (the reference's TESTING solution for
    # TG_VEL, -2 sin x sin y at solver.c:2579, does not match its own IC)
    w = host.initial_condition(_cfg("TG_VEL", Nx, Ny), r2c=r2c)
    wr = O.c2r(w, Nx, Ny) / (Nx * Ny)
    x, y = np.meshgrid(np.arange(Nx) * 2 * np.pi / Nx, np.arange(Ny) * 2 * np.pi / Ny, indexing="ij")
    np.testing.assert_allclose(wr, -2 * np.cos(x) * np.cos(y), atol=1e-13)
    # decaying turbulence spectra carry the reference's energy E0 (reference formula)
    p = O.Params(Nx, Ny)
    for u0, e0 in (("DECAY_TURB_BB", 0.5), ("DECAY_TURB_NB", 0.5), ("DECAY_TURB_EXP", 343.0 / 96.0)):
        w = host.initial_condition(_cfg(u0, Nx, Ny), r2c=r2c)
        kx, ky = O.wavenumbers(Nx, Ny)
        k2 = (kx * kx + ky * ky).astype(float)
        c = np.full(Ny // 2 + 1, 2.0); c[0] = c[-1] = 1.0
        keep = (kx == 0) | (ky != 0)
        raw = np.sum(np.where(k2 > 0, c * np.abs(w) ** 2 / np.where(k2 > 0, k2, 1), 0))
        # ky=0 rows with kx != 0 are cleared after normalisation (solver.c:1344)
        assert np.all(w[~keep] == 0)
        assert raw * 4 * np.pi ** 2 * 0.5 / (Nx * Ny) ** 2 <= e0 * (1 + 1e-12)
    # UNIF: unit-modulus modes times the normalisation on the annulus 2.5 < |k| < 9.5
    w = host.initial_condition(_cfg("UNIF", Nx, Ny), r2c=r2c)
    kx, ky = O.wavenumbers(Nx, Ny)
    ka = np.sqrt(kx * kx + ky * ky)
    band = (ka > 2.5) & (ka < 9.5)
    c = np.full(Ny // 2 + 1, 2.0); c[0] = c[-1] = 1.0
    norm = math.sqrt(1.0 / np.sum(np.broadcast_to(c, band.shape)[band]))
    np.testing.assert_allclose(np.abs(w[band]), norm, rtol=1e-14)
    assert np.all(w[~band] == 0)
    # the conjugacy fix-up makes the ky=0 column Hermitian
    w = host.initial_condition(_cfg("RANDOM", Nx, Ny), r2c=r2c)
    for i in range(1, Nx // 2):
        assert w[Nx - i, 0] == np.conj(w[i, 0])


def test_max_palin_reads_the_stream_function_file(tmp_path):
    Nx, Ny = 32, 32
    rng = np.random.default_rng(3)
    psi = rng.standard_normal((Nx, Ny))
    d = tmp_path / "MaxPalinstrophy"
    d.mkdir()
    write_double_dataset(str(d / f"maxdpdt_P10000_N{Nx}_IG0_psi_f.h5"), "psi", psi)
    c = _cfg("MAX_PALIN", Nx, Ny)
    c.input_dir = (str(tmp_path) + "/").encode()
    got = host.initial_condition(c, r2c=r2c)
    ref = R.initial_condition("MAX_PALIN", Nx, Ny, psi=psi)
    np.testing.assert_array_equal(got, ref)
    c.input_dir = b"/nonexistent/"
    with pytest.raises(host.HostError, match="Unable to open input file"):
        host.initial_condition(c, r2c=r2c)


def test_real_space_ics_need_transforms():
    with pytest.raises(host.HostError, match="needs transforms"):
        host.initial_condition(_cfg("TG_VORT", 16, 16))


@pytest.mark.parametrize("filt,forcing,oracle_filter", [("23", "NONE", "23"), ("HOU_LI", "NONE", "houli_ref"),
                                                        ("23", "KOLM", "23")])
def test_host_dealias_matches_oracle(filt, forcing, oracle_filter):
    Nx, Ny = 64, 48
    rng = np.random.default_rng(0)
    f = rng.standard_normal((Nx, Ny // 2 + 1)) + 1j * rng.standard_normal((Nx, Ny // 2 + 1))
    c = _cfg("ZERO", Nx, Ny, filt, forcing)
    p = O.Params(Nx, Ny, filter=oracle_filter, forcing=forcing if forcing == "KOLM" else "NONE", force_k=2)
    np.testing.assert_array_equal(host.dealias(c, f), f * O.dealias_factor(p))


def test_mean_flows_match_restatement():
    Nx, Ny = 32, 64
    w = host.initial_condition(_cfg("DECAY_TURB_BB", Nx, Ny), r2c=r2c)
    my, mx = host.mean_flows(_cfg("ZERO", Nx, Ny), w)
    ry, rx = R.mean_flows(w, Nx, Ny)
    np.testing.assert_allclose(my, ry, rtol=1e-14)
    np.testing.assert_allclose(mx, rx, rtol=1e-14)


def test_hdf5_layout_round_trip(tmp_path):
    Nx, Ny = 16, 8
    c = host.parse_cli(["-o", str(tmp_path) + "/", "-n", str(Nx), "-n", str(Ny), "-t", "h5"])
    paths = host.output_paths(c, mkdirs=True)
    ns = O.n_spect(Nx, Ny)
    rng = np.random.default_rng(1)
    saves = []
    with host.H5Writer(c, paths, ns) as f:
        for k in range(3):
            s = dict(w=rng.standard_normal((Nx, Ny)),
                     w_hat=rng.standard_normal((Nx, Ny // 2 + 1)) + 1j * rng.standard_normal((Nx, Ny // 2 + 1)),
                     enst_spect=rng.random(ns), enrg_spect=rng.random(ns), flux_spect=rng.random(6 * ns))
            f.save(k, t=0.1 * k, dt=0.1, **s)
            saves.append(s)
        series = {k: rng.random(3) for k in host.SERIES_KEYS}
        series["mean_flow_y"] = rng.random(Nx)
        series["mean_flow_x"] = rng.random(Ny // 2 + 1)
        f.final(series)
    with File(paths["main_file"]) as h:
        assert h.keys() == ["Iter_00000", "Iter_00001", "Iter_00002"]
        for k, s in enumerate(saves):
            g = f"/Iter_{k:05d}"
            assert sorted(h.keys(g)) == ["w", "w_hat"]
            assert h.attr(g, "TimeValue") == 0.1 * k and h.attr(g, "TimeStep") == 0.1
            np.testing.assert_array_equal(h.read(g + "/w"), s["w"])
            np.testing.assert_array_equal(h.read(g + "/w_hat"), s["w_hat"])
            assert h.members(g + "/w_hat") == [("r", 0), ("i", 8)]
    names = ["TimeDerivativeEnstrophySpectrum", "EnstrophyDissipationSpectrum", "EnstrophyFluxSpectrum",
             "TimeDerivativeEnergySpectrum", "EnergyDissipationSpectrum", "EnergyFluxSpectrum"]
    with File(paths["spectra_file"]) as h:
        assert h.keys() == ["Iter_00000", "Iter_00001", "Iter_00002"]
        for k, s in enumerate(saves):
            g = f"/Iter_{k:05d}/"
            np.testing.assert_array_equal(h.read(g + "EnstrophySpectrum"), s["enst_spect"])
            np.testing.assert_array_equal(h.read(g + "EnergySpectrum"), s["enrg_spect"])
            for q, nm in enumerate(names):
                np.testing.assert_array_equal(h.read(g + nm), s["flux_spect"][q * ns:(q + 1) * ns])
    with File(paths["sys_msr_file"]) as h:
        np.testing.assert_array_equal(h.read("kx"), np.r_[0:Nx // 2 + 1, -Nx // 2 + 1:0])
        np.testing.assert_array_equal(h.read("ky"), np.arange(Ny // 2 + 1))
        np.testing.assert_array_equal(h.read("x"), np.arange(Nx) * 2.0 * np.pi / Nx)
        np.testing.assert_array_equal(h.read("y"), np.arange(Ny) * 2.0 * np.pi / Ny)
        ref_names = {"time": "Time", "tot_energy": "TotalEnergy", "tot_enstr": "TotalEnstrophy",
                     "tot_palin": "TotalPalinstrophy", "tot_forc": "TotalForcing", "tot_div": "TotalDivergence",
                     "enrg_diss": "EnergyDissipation", "enst_diss": "EnstrophyDissipation",
                     "d_enst_dt_sbst": "TimeDerivativeEnstrophySubset", "enst_flux_sbst": "EnstrophyFluxSubset",
                     "enst_diss_sbst": "EnstrophyDissSubset", "d_enrg_dt_sbst": "TimeDerivativeEnergySubset",
                     "enrg_flux_sbst": "EnergyFluxSubset", "enrg_diss_sbst": "EnergyDissSubset",
                     "mean_flow_y": "MeanFlow_y", "mean_flow_x": "MeanFlow_x"}
        for k, nm in ref_names.items():
            np.testing.assert_array_equal(h.read(nm), series[k])
