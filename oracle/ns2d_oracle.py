"""CPU restatement (numpy) of the reference solver's hot path.

TEST INFRASTRUCTURE ONLY.  This module is the checker against which the HIP
path is compared; only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it.  The product path never
calls it.

It restates, function by function, the reference's
(EndCar808/2D-Navier-Stokes) vorticity right-hand side, dealiasing, conjugacy
fix-up, RK4-Crank-Nicolson and Adams-Bashforth-4 integrators, Kolmogorov
forcing, max-vorticity check and system measures.  Every function names the
reference lines it follows (paths relative to the reference root).

Conventions (reference ``Solver/src/solver.c:2022-2068``, FFTW plans at
``:2873-2901``):

* ``w_hat`` is complex128 ``[Nx][Ny//2+1]``, x slow, in *unnormalised* FFTW
  units: ``w_hat = r2c(w)``.
* ``kx = 0..Nx/2, -Nx/2+1..-1`` (rows), ``ky = 0..Ny/2`` (columns).
* ``c2r`` is FFTW's unnormalised 2-D complex-to-real transform: a complex
  inverse transform along x, then a 1-D c2r along y that ignores the
  imaginary parts of the ky=0 and ky=Ny/2 entries.  numpy's ``irfft2`` with
  ``norm="forward"`` has exactly these semantics (checked against the DFT
  definition in ``tests/test_oracle.py``).

Reference quirks reproduced when the matching ``QUIRK_*`` bit is set (all set
by default, ``QUIRKS_REF``):

* Q1 the RHS is ``+u.grad(w)`` (``solver.c:981``) -- always reproduced.
* Q2 ``ForceConjugacy(arg)`` writes ``conj(arg[|kx|,0])`` into the *global*
  state's ky=0, kx<0 entries (``solver.c:1031, 2089-2101``).
* Q3 ``HYPER_VISC_FLAG = HYPER_VISC`` is an assignment in every RK4/RK4CN
  viscous branch (``solver.c:675,714,753,792,843``): with the drag flag off
  the viscous factor is always ``nu*pow(k^2, hyper_pow)`` and the flag stays 1.
* Q4 the Hou-Li filter divides integers (``solver.c:1088``): filter
  ``"houli_ref"``.
* Q6 Kolmogorov forcing ``-0.5*F*k_f`` is added to the N^2-scaled RHS
  unscaled (``force.c:537``, ``solver.c:1022-1027``) and the 2/3 filter zeroes
  every kx=0 mode when forcing is KOLM (``solver.c:1072-1078``).
* Q7 ``enst_diss`` drops ``pre_fac`` off the ky=0/Nyquist edge
  (``sys_msr.c:225``).

Parity pinning: every function here is checked in ``tests/test_oracle.py``
against the DFT definition, closed forms (Taylor-Green), analytic Jacobians,
conservation invariants, brute-force enumeration or Parseval -- see that
file's docstrings for which check pins which function.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

try:  # multi-threaded pocketfft for the timed CPU baseline; numpy otherwise
    import scipy.fft as _sfft
except Exception:  # pragma: no cover
    _sfft = None

# ---------------------------------------------------------------------------
# quirk bits (mirrors include/ns2d_gpu.h NS2D_QUIRK_*)
# ---------------------------------------------------------------------------
QUIRK_CONJ_GLOBAL = 1 << 0   # Q2
QUIRK_VISC_ASSIGN = 1 << 1   # Q3
QUIRK_KOLM_UNSCALED = 1 << 2  # Q6 forcing amplitude not N^2-scaled
QUIRK_KOLM_KX0_KILL = 1 << 3  # Q6 2/3 filter zeroes kx=0 under KOLM
QUIRK_ENST_DISS = 1 << 4     # Q7
QUIRKS_REF = (QUIRK_CONJ_GLOBAL | QUIRK_VISC_ASSIGN | QUIRK_KOLM_UNSCALED
              | QUIRK_KOLM_KX0_KILL | QUIRK_ENST_DISS)

# Butcher / AB4 constants, solver.c:29-47
RK4_A21, RK4_A32, RK4_A43 = 0.5, 0.5, 1.0
RK4_B1, RK4_B2, RK4_B3, RK4_B4 = 1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0
AB4_1, AB4_2, AB4_3, AB4_4 = 55.0 / 24.0, -59.0 / 24.0, 37.0 / 24.0, -3.0 / 8.0
MAX_VORT_LIM = 1e100          # data_types.h:173
MIN_STEP_SIZE = 1e-10         # data_types.h:171
KAPPA = 1.0                   # data_types.h:140

_FFT_WORKERS = 1

# element-wise libm functions: the reference calls the C library's exp/pow
# (glibc, correctly rounded here); numpy's vectorised exp can differ by 1 ulp
_libm_exp = np.vectorize(math.exp, otypes=[np.float64])
_libm_pow = np.vectorize(math.pow, otypes=[np.float64])


def set_fft_workers(n: int) -> None:
    """Threads used by the oracle's FFTs (only the CPU baseline sets >1)."""
    global _FFT_WORKERS
    _FFT_WORKERS = max(1, int(n))


@dataclasses.dataclass
class Params:
    """Solver parameters; names follow ``system_vars_struct`` (data_types.h:190-304)
    and the CLI of ``utils.c:34-461`` (-n Nx -n Ny, -v nu -v hyper_flag
    -v hyper_pow, -d alpha -d drag_flag -d drag_pow, -f F -f k_f)."""
    Nx: int
    Ny: int
    nu: float = 1e-3
    hyper_flag: int = 0
    hyper_pow: float = 1.0
    alpha: float = 0.0
    drag_flag: int = 0
    drag_pow: float = 0.0
    filter: str = "23"            # "23" | "houli_ref" | "houli" | "none"
    forcing: str = "NONE"         # "NONE" | "KOLM"
    force_k: int = 0
    force_scale: float = 0.0
    integrator: str = "RK4CN"     # "RK4CN" | "AB4" | "RK4"
    quirks: int = QUIRKS_REF

    @property
    def Nyf(self) -> int:
        return self.Ny // 2 + 1


def wavenumbers(Nx: int, Ny: int):
    """kx (Nx,1) and ky (1,Nyf) int64 arrays, solver.c:2048-2067."""
    kx = np.array([i if i <= Nx // 2 else i - Nx for i in range(Nx)], dtype=np.int64)
    ky = np.arange(Ny // 2 + 1, dtype=np.int64)
    return kx[:, None], ky[None, :]


def collocation(Nx: int, Ny: int):
    """x (Nx,1), y (1,Ny) collocation points, solver.c:2040-2066."""
    x = np.arange(Nx, dtype=np.float64) * 2.0 * math.pi / Nx
    y = np.arange(Ny, dtype=np.float64) * 2.0 * math.pi / Ny
    return x[:, None], y[None, :]


# ---------------------------------------------------------------------------
# FFTW-convention transforms (plans solver.c:2883-2895)
# ---------------------------------------------------------------------------
def c2r(X: np.ndarray, Nx: int, Ny: int) -> np.ndarray:
    """Unnormalised 2-D c2r (fftw_mpi_execute_dft_c2r, solver.c:932,964,2221)."""
    if _sfft is not None:
        return _sfft.irfft2(X, s=(Nx, Ny), norm="forward", workers=_FFT_WORKERS)
    return np.fft.irfft2(X, s=(Nx, Ny), norm="forward")


def r2c(x: np.ndarray) -> np.ndarray:
    """Unnormalised 2-D r2c (fftw_mpi_execute_dft_r2c, solver.c:994,1178)."""
    if _sfft is not None:
        return _sfft.rfft2(x, workers=_FFT_WORKERS)
    return np.fft.rfft2(x)


# ---------------------------------------------------------------------------
# dealiasing, conjugacy, forcing
# ---------------------------------------------------------------------------
def dealias_factor(p: Params) -> np.ndarray:
    """Per-mode multiplier applied by ApplyDealiasing, solver.c:1039-1099.

    "23": 0 where kx^2+ky^2 >= (int)(Nx/3)^2 (both axes use Nx, :1048,:1065);
          with KOLM forcing (Q6) also 0 on every kx=0 mode (:1072-1078).
    "houli_ref": exp(-36*(sqrt((kx/(Nx/2))^2+(ky/(Ny/2))^2))^36) with C integer
          (truncating) division (Q4, :1088).
    "houli": the same filter with real division (physics form).
    """
    kx, ky = wavenumbers(p.Nx, p.Ny)
    if p.filter == "23":
        k2 = (kx * kx + ky * ky).astype(np.float64)
        kmax_sqr = float(int(p.Nx / 3.0)) ** 2
        f = np.where(k2 >= kmax_sqr, 0.0, 1.0)
        if p.forcing == "KOLM" and (p.quirks & QUIRK_KOLM_KX0_KILL):
            f = np.where((kx == 0) & (k2 < kmax_sqr), 0.0, f)
        return f
    if p.filter == "houli_ref":
        hx, hy = p.Nx // 2, p.Ny // 2
        qx = (np.sign(kx) * (np.abs(kx) // hx)).astype(np.float64)  # C truncation
        qy = (np.sign(ky) * (np.abs(ky) // hy)).astype(np.float64)
        return _libm_exp(-36.0 * _libm_pow(np.sqrt(qx ** 2 + qy ** 2), 36.0))
    if p.filter == "houli":
        qx = kx / (p.Nx / 2.0)
        qy = ky / (p.Ny / 2.0)
        return _libm_exp(-36.0 * _libm_pow(np.sqrt(qx ** 2 + qy ** 2), 36.0))
    if p.filter == "none":
        return np.ones((p.Nx, p.Nyf))
    raise ValueError(p.filter)


def apply_dealiasing(arr: np.ndarray, p: Params) -> np.ndarray:
    """ApplyDealiasing(array, 1, N), solver.c:1039-1099 (returns a new array)."""
    return arr * dealias_factor(p)


def forcing_value(p: Params) -> complex:
    """ComputeForcing KOLM value, force.c:535-538: -0.5*F*k_f (Q6: not N^2 scaled).
    Without Q6 the physical amplitude is scaled to FFTW units (x Nx*Ny)."""
    if p.forcing != "KOLM":
        return 0.0
    f = -0.5 * p.force_scale * float(p.force_k)
    if not (p.quirks & QUIRK_KOLM_UNSCALED):
        f *= float(p.Nx * p.Ny)
    return f


def force_conjugacy(arg: np.ndarray, w_hat_global: np.ndarray) -> None:
    """ForceConjugacy, solver.c:2074-2107: gather arg's ky=0 column, then write
    conj(arg[|kx|,0]) into the *global* state's kx<0 rows (Q2)."""
    Nx = arg.shape[0]
    col = arg[:, 0].copy()
    for i in range(Nx // 2 + 1, Nx):      # rows with kx<0
        w_hat_global[i, 0] = np.conj(col[Nx - i])


# ---------------------------------------------------------------------------
# the vorticity right-hand side
# ---------------------------------------------------------------------------
def nonlinear_rhs(w_hat: np.ndarray, p: Params, with_force: bool = True) -> np.ndarray:
    """NonlinearRHSBatch, solver.c:884-1032 (without the ForceConjugacy side
    effect, which callers apply with ``force_conjugacy``).

    u_hat = i*ky/|k|^2*w_hat, v_hat = -i*kx/|k|^2*w_hat (:914-915);
    d_x w = i*kx*w_hat, d_y w = i*ky*w_hat (:946-947); the (0,0) mode of each is 0;
    n = u*w_x + v*w_y in real space (:981, Q1 sign); K = r2c(n)*(1/(Nx*Ny))^2
    (:1005) with K(0,0)=0 (:1008); then ApplyDealiasing (:1019) and the forcing
    added on (kx=0, ky=k_f) (:1022-1027, force.c:93).
    """
    Nx, Ny = p.Nx, p.Ny
    kx, ky = wavenumbers(Nx, Ny)
    k2 = kx * kx + ky * ky
    nz = k2 != 0
    inv = np.where(nz, 1.0 / np.where(nz, k2, 1).astype(np.float64), 0.0)
    fkx = kx.astype(np.float64)
    fky = ky.astype(np.float64)
    u_hat = np.where(nz, 1j * fky * inv * w_hat, 0.0)
    v_hat = np.where(nz, -1.0 * 1j * fkx * inv * w_hat, 0.0)
    u = c2r(u_hat, Nx, Ny)
    v = c2r(v_hat, Nx, Ny)
    wx_hat = np.where(nz, 1j * fkx * w_hat, 0.0)
    wy_hat = np.where(nz, 1j * fky * w_hat, 0.0)
    wx = c2r(wx_hat, Nx, Ny)
    wy = c2r(wy_hat, Nx, Ny)
    nl = 1.0 * (u * wx + v * wy)
    out = r2c(nl)
    norm_fac = 1.0 / (Nx * Ny)
    out = np.where(nz, out * (norm_fac * norm_fac), 0.0)
    out = apply_dealiasing(out, p)
    if with_force and p.forcing == "KOLM":
        out[0, p.force_k] += forcing_value(p)
    return out


def rhs_with_side_effect(arg: np.ndarray, state: np.ndarray, p: Params) -> np.ndarray:
    """NonlinearRHSBatch(arg, ...) including its ForceConjugacy(arg) write into
    ``state`` (the global w_hat) when Q2 is on, solver.c:1031."""
    K = nonlinear_rhs(arg, p)
    if p.quirks & QUIRK_CONJ_GLOBAL:
        force_conjugacy(arg, state)
    return K


# ---------------------------------------------------------------------------
# viscous / drag factor
# ---------------------------------------------------------------------------
def _pw(k2: np.ndarray, e: float) -> np.ndarray:
    # the (0,0) mode is excluded by every caller (its update is forced to 0)
    return _libm_pow(np.where(k2 > 0, k2, 1.0), e)


def visc_coeff(p: Params, hyper_flag: int, branch: str = "rk") -> np.ndarray:
    """The D_fac/dt coefficient of the CN update.

    branch "rk": RK4Step/RK4CN branches solver.c:835-850 (Q3 assignment makes the
    third test always true once the first two fail, i.e. whenever drag_flag!=1).
    branch "ab": AB4Step branches solver.c:315-330 (plain comparisons).
    """
    kx, ky = wavenumbers(p.Nx, p.Ny)
    k2 = (kx * kx + ky * ky).astype(np.float64)
    hyp = hyper_flag == 1
    drag = p.drag_flag == 1
    if hyp and drag:
        return p.nu * _pw(k2, p.hyper_pow) + p.alpha * _pw(k2, p.drag_pow)
    if (not hyp) and drag:
        return p.nu * k2 + p.alpha * _pw(k2, p.drag_pow)
    if branch == "rk" and (p.quirks & QUIRK_VISC_ASSIGN):
        return p.nu * _pw(k2, p.hyper_pow)          # (flag = 1) && !drag
    if hyp and not drag:
        return p.nu * _pw(k2, p.hyper_pow)
    return p.nu * k2


class State:
    """Integration state: w_hat plus what the reference keeps in Int_data /
    sys_vars between steps (AB4 history, the possibly-mutated hyper flag)."""

    def __init__(self, w_hat: np.ndarray, p: Params):
        self.p = p
        self.w_hat = np.array(w_hat, dtype=np.complex128, copy=True)
        self.iters = 0                  # completed steps
        self.hyper_flag = p.hyper_flag  # mutated by Q3
        self.ab_hist = [None, None, None]  # AB_tmp_nonlin[0..2]


def _rk_visc_side_effect(st: State) -> None:
    p = st.p
    if (p.quirks & QUIRK_VISC_ASSIGN) and p.drag_flag != 1:
        st.hyper_flag = 1


def rk4cn_step(st: State, dt: float) -> np.ndarray:
    """RK4Step with __RK4CN, solver.c:625-873.  Returns K1 (RK1), which AB4's
    pre-steps store (solver.c:283)."""
    p = st.p
    w = st.w_hat
    K1 = rhs_with_side_effect(w, w, p)
    tmp = w + (dt * RK4_A21) * K1                                   # :690
    K2 = rhs_with_side_effect(tmp, w, p)
    tmp = w + (dt * RK4_A32) * K2                                   # :729
    K3 = rhs_with_side_effect(tmp, w, p)
    tmp = w + (dt * RK4_A43) * K3                                   # :768
    K4 = rhs_with_side_effect(tmp, w, p)
    D = dt * visc_coeff(p, st.hyper_flag, "rk")                     # :833-850
    _rk_visc_side_effect(st)
    ssum = RK4_B1 * K1 + RK4_B2 * K2 + RK4_B3 * K3 + RK4_B4 * K4
    new = w * ((2.0 - D) / (2.0 + D)) + (2.0 * dt / (2.0 + D)) * ssum  # :853
    new[0, 0] = 0.0                                                  # :860
    st.w_hat = new
    st.iters += 1
    return K1


def rk4_step(st: State, dt: float) -> None:
    """Classic explicit RK4 with the linear term inside every stage (the
    physics form of the reference's __RK4 integrator; the reference's own
    __RK4/__NAVIER branch adds +dt*nu*k^2*w to each stage, solver.c:664-684,
    and is not reproduced)."""
    p = st.p
    lam = visc_coeff(p, p.hyper_flag, "ab")
    w = st.w_hat

    def F(s):
        return rhs_with_side_effect(s, w, p) - lam * s

    R1 = F(w)
    R2 = F(w + (dt * RK4_A21) * R1)
    R3 = F(w + (dt * RK4_A32) * R2)
    R4 = F(w + (dt * RK4_A43) * R3)
    new = w + (dt / 6.0) * (R1 + 2.0 * R2 + 2.0 * R3 + R4)
    new[0, 0] = 0.0
    st.w_hat = new
    st.iters += 1


def ab4_step(st: State, dt: float) -> None:
    """AB4Step, solver.c:261-350, with real RK4CN start-up steps (Q5: the
    shipped header makes AB4 unreachable; the start-up uses RK4Step's CN form)."""
    p = st.p
    iters = st.iters + 1
    if iters <= 3:
        K1 = rk4cn_step(st, dt)
        st.ab_hist[iters - 1] = K1.copy()                           # :283
        return
    w = st.w_hat
    N = rhs_with_side_effect(w, w, p)                                # :296
    D = dt * visc_coeff(p, st.hyper_flag, "ab")                     # :313-330
    H = st.ab_hist
    ssum = AB4_1 * N + AB4_2 * H[2] + AB4_3 * H[1] + AB4_4 * H[0]
    new = w * ((2.0 - D) / (2.0 + D)) + (2.0 * dt / (2.0 + D)) * ssum  # :333
    new[0, 0] = 0.0                                                  # :337
    st.w_hat = new
    st.ab_hist = [H[1], H[2], N]                                     # :346-348
    st.iters += 1


def step(st: State, dt: float, nsteps: int = 1) -> None:
    for _ in range(nsteps):
        if st.p.integrator == "RK4CN":
            rk4cn_step(st, dt)
        elif st.p.integrator == "AB4":
            ab4_step(st, dt)
        elif st.p.integrator == "RK4":
            rk4_step(st, dt)
        else:
            raise ValueError(st.p.integrator)


# ---------------------------------------------------------------------------
# diagnostics
# ---------------------------------------------------------------------------
def real_vorticity(w_hat: np.ndarray, p: Params) -> np.ndarray:
    """GetMaxData("VORT") transform: w = c2r(w_hat)/(Nx*Ny), solver.c:2219-2231."""
    return c2r(w_hat, p.Nx, p.Ny) * (1.0 / (p.Nx * p.Ny))


def max_vort(w_hat: np.ndarray, p: Params) -> float:
    """GetMaxData("VORT") max|w|, solver.c:2237-2270."""
    return float(np.max(np.abs(real_vorticity(w_hat, p))))


def cfl_timestep(w_hat: np.ndarray, p: Params, cfl: float) -> float:
    """GetTimestep's CFL branch, solver.c:2381-2450: u_hat = i ky/|k|^2 w_hat,
    v_hat = -i kx/|k|^2 w_hat (:2392-2401), u, v = c2r(.)/(Nx Ny) (:2409-2420),
    dt = cfl / (pi (max|u|/dx + max|v|/dy) + pi^2 nu (1/dx^2 + 1/dy^2))
    (:2425-2447), max taken over the sum per point as the reference's loop
    does.  The reference's u_hat index at :2394 offsets the rows (its u is
    wrong); this uses every mode's own velocity, as the driver does."""
    Nx, Ny = p.Nx, p.Ny
    kx, ky = wavenumbers(Nx, Ny)
    k2 = (kx * kx + ky * ky).astype(np.float64)
    k2s = np.where(k2 > 0, k2, 1.0)
    u_hat = np.where(k2 > 0, 1j * ky * w_hat / k2s, 0.0)
    v_hat = np.where(k2 > 0, -1j * kx * w_hat / k2s, 0.0)
    u = c2r(u_hat, Nx, Ny) * (1.0 / (Nx * Ny))
    v = c2r(v_hat, Nx, Ny) * (1.0 / (Nx * Ny))
    dx, dy = 2.0 * math.pi / Nx, 2.0 * math.pi / Ny
    conv = float(np.max(math.pi * (np.abs(u) / dx + np.abs(v) / dy)))
    diff = math.pi ** 2 * (p.nu / dx ** 2 + p.nu / dy ** 2)
    return cfl / (conv + diff)


def n_spect(Nx: int, Ny: int) -> int:
    """sys_msr.c:394."""
    return int(round(math.sqrt((Nx / 2.0) * (Nx / 2.0) + (Ny / 2.0) * (Ny / 2.0)) + 1))


def measures(w_hat: np.ndarray, p: Params, hyper_flag: int | None = None,
             rhs: np.ndarray | None = None) -> dict:
    """ComputeSystemMeasurables, sys_msr.c:32-373 (E, Z, P, divergence, energy
    and enstrophy dissipation, forcing input, energy/enstrophy spectra; and,
    when ``rhs`` = NonlinearRHSBatch(w_hat) - forcing is given, the flux
    subsets and flux spectra of :200-305).

    Weights c=1 on the ky=0 and ky=Ny/2 columns and 2 elsewhere; C=4*pi^2*0.5/(NxNy)^2.
    """
    if hyper_flag is None:
        hyper_flag = p.hyper_flag
    Nx, Ny = p.Nx, p.Ny
    Nyf = Ny // 2 + 1
    kx, ky = wavenumbers(Nx, Ny)
    k2i = kx * kx + ky * ky
    k2 = k2i.astype(np.float64)
    nz = k2i != 0
    k2s = np.where(nz, k2, 1.0)
    norm_fac = 0.5 / float(Nx * Ny) ** 2
    const_fac = 4.0 * math.pi ** 2
    c = np.full((1, Nyf), 2.0)
    c[0, 0] = 1.0
    c[0, Nyf - 1] = 1.0
    # pre_fac, sys_msr.c:160-175 (plain comparisons on the current flag)
    hyp = hyper_flag == 1
    drag = p.drag_flag == 1
    if hyp and drag:
        pre = p.nu * _libm_pow(k2, p.hyper_pow) + p.alpha * _libm_pow(k2s, p.drag_pow)
    elif (not hyp) and drag:
        pre = p.nu * k2 + p.alpha * _libm_pow(k2s, p.drag_pow)
    elif hyp and not drag:
        pre = p.nu * _libm_pow(k2, p.hyper_pow)
    else:
        pre = p.nu * k2
    a2 = (w_hat.real ** 2 + w_hat.imag ** 2)   # cabs(w * conj(w)), sys_msr.c:191
    a2 = np.where(nz, a2, 0.0)
    u_z = 1j * ky * w_hat / k2s
    v_z = -1j * kx * w_hat / k2s
    div = 1j * (kx * u_z + ky * v_z)
    uv2 = np.abs(u_z * np.conj(u_z) + v_z * np.conj(v_z))
    cw = np.where(nz, c, 0.0)
    out = {}
    out["energy"] = float(np.sum(cw * a2 / k2s)) * const_fac * norm_fac
    out["enstrophy"] = float(np.sum(cw * a2)) * const_fac * norm_fac
    out["palinstrophy"] = float(np.sum(cw * k2 * a2)) * const_fac * norm_fac
    out["divergence"] = float(np.sum(cw * np.abs(div * np.conj(div)))) * const_fac * norm_fac
    out["enrg_diss"] = float(np.sum(cw * pre * uv2)) * 2.0 * const_fac * norm_fac
    if p.quirks & QUIRK_ENST_DISS:
        # sys_msr.c:197 (edge, with pre_fac) and :225 (interior, pre_fac missing)
        ew = np.where(c == 1.0, pre, 1.0)
    else:
        ew = pre
    out["enst_diss"] = float(np.sum(cw * ew * a2)) * 2.0 * const_fac * norm_fac
    if p.forcing == "KOLM":
        out["forcing"] = abs(forcing_value(p) * w_hat[0, p.force_k]) * const_fac * norm_fac
    else:
        out["forcing"] = 0.0
    ns = n_spect(Nx, Ny)
    bins = np.rint(np.sqrt(k2)).astype(np.int64)   # round(): no half-way ties occur
    ek = np.zeros(ns)
    zk = np.zeros(ns)
    np.add.at(ek, bins[nz], (cw * const_fac * norm_fac * a2 / k2s)[nz])
    np.add.at(zk, bins[nz], (cw * const_fac * norm_fac * a2)[nz])
    out["enrg_spect"] = ek
    out["enst_spect"] = zk
    if rhs is not None:
        deriv = (w_hat * np.conj(rhs) + np.conj(w_hat) * rhs).real
        diss = pre * a2
        lo, hi = 0.0, 64.0 ** 2           # LWR/UPR_SBST_LIM, data_types.h:174-175
        sb = nz & (k2 >= lo) & (k2 < hi)
        wgt = np.where(sb, c, 0.0)
        out["d_enrg_dt_sbst"] = float(np.sum(wgt * (deriv + diss) / k2s))
        out["enrg_diss_sbst"] = float(np.sum(wgt * diss / k2s)) * 2.0 * const_fac * norm_fac
        out["enrg_flux_sbst"] = float(np.sum(wgt * deriv / k2s)) * const_fac * norm_fac
        out["d_enst_dt_sbst"] = float(np.sum(wgt * (deriv + diss)))
        out["enst_diss_sbst"] = float(np.sum(wgt * diss)) * 2.0 * const_fac * norm_fac
        out["enst_flux_sbst"] = float(np.sum(wgt * deriv)) * const_fac * norm_fac
        td = deriv * const_fac * norm_fac
        ts = diss * const_fac * norm_fac
        for name, val in (("d_enst_dt_spect", td + ts), ("enst_diss_spect", ts),
                          ("enst_flux_spect", td), ("d_enrg_dt_spect", (td + ts) / k2s),
                          ("enrg_diss_spect", ts / k2s), ("enrg_flux_spect", td / k2s)):
            s = np.zeros(ns)
            np.add.at(s, bins[nz], (cw * val)[nz])
            out[name] = s
    return out


# ---------------------------------------------------------------------------
# initial conditions used by the parity configurations
# ---------------------------------------------------------------------------
def ic_taylor_green(p: Params) -> np.ndarray:
    """TG_VORT, solver.c:1161-1179: w = 2*kappa*cos(kappa x)cos(kappa y), r2c;
    then ApplyDealiasing (:1994) and ForceConjugacy (:1997)."""
    x, y = collocation(p.Nx, p.Ny)
    w = 2.0 * KAPPA * np.cos(KAPPA * x) * np.cos(KAPPA * y)
    w_hat = r2c(w)
    w_hat = apply_dealiasing(w_hat, p)
    force_conjugacy(w_hat, w_hat)
    return w_hat


def hermitian_columns(w_hat: np.ndarray) -> None:
    """Make the ky=0 and ky=Ny/2 columns Hermitian in kx (w(-kx)=conj(w(kx)),
    real at kx=0 and kx=Nx/2), as a real field's r2c output is."""
    Nx = w_hat.shape[0]
    for j in (0, w_hat.shape[1] - 1):
        for i in range(Nx // 2 + 1, Nx):
            w_hat[i, j] = np.conj(w_hat[Nx - i, j])
        w_hat[0, j] = w_hat[0, j].real
        w_hat[Nx // 2, j] = w_hat[Nx // 2, j].real


def ic_random_spectrum(p: Params, seed: int, k0: float = 8.0, energy: float = 0.5,
                       kmin: float = 0.0, kmax: float | None = None,
                       spectrum: str = "k4exp") -> np.ndarray:
    """Seeded random-phase vorticity with energy spectrum E(k) ~ k^4 exp(-(k/k0)^2)
    ("k4exp") or flat on kmin<|k|<=kmax ("flat"), normalised so that the
    reference's energy formula (sys_msr.c:191) gives ``energy``.  Not a reference
    IC (nearest: DECAY_TURB_EXP, solver.c:1275-1277); used to build the parity
    workloads.  Hermitian on the ky=0 column, (0,0)=0, dealiased."""
    Nx, Ny = p.Nx, p.Ny
    rng = np.random.default_rng(seed)
    kx, ky = wavenumbers(Nx, Ny)
    k2 = (kx * kx + ky * ky).astype(np.float64)
    kk = np.sqrt(k2)
    if spectrum == "k4exp":
        spec = kk ** 4 * np.exp(-(kk / k0) ** 2)
    elif spectrum == "flat":
        spec = np.where((kk > kmin) & (kk <= (kmax if kmax is not None else np.inf)), 1.0, 0.0)
    else:
        raise ValueError(spectrum)
    # |w_hat|^2 per mode ~ k^2 * E(k)/k (annulus area)
    amp = np.sqrt(np.where(k2 > 0, spec * kk, 0.0))
    phase = rng.uniform(0.0, 2.0 * np.pi, size=(Nx, Ny // 2 + 1))
    w_hat = amp * np.exp(1j * phase)
    w_hat[0, 0] = 0.0
    w_hat = apply_dealiasing(w_hat, p)
    hermitian_columns(w_hat)
    e = measures(w_hat, p)["energy"]
    if e > 0:
        w_hat *= math.sqrt(energy / e)
    return w_hat


def ic_kolmogorov_lowamp(p: Params, seed: int, kmax: float = 16.0, amp: float = 1e-3) -> np.ndarray:
    """Low-amplitude random vorticity with uniform phases on 0<|k|<=kmax, scaled
    so that max |w_hat|/(Nx*Ny) = amp (physical units).  Used for the
    Kolmogorov-flow workload."""
    w_hat = ic_random_spectrum(p, seed, spectrum="flat", kmin=0.0, kmax=kmax, energy=1.0)
    m = np.max(np.abs(w_hat)) / (p.Nx * p.Ny)
    if m > 0:
        w_hat *= amp / m
    return w_hat
