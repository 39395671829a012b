"""Host-side initial conditions (reference InitialConditions, Solver/src/solver.c:1107-2014).

These run once before the time loop.  Spectral ICs are built directly in
w_hat; real-space ICs go through the GPU's unnormalised r2c (``Solver.r2c``).
Every IC ends, as in the reference, with ApplyDealiasing (solver.c:1994) and
the ky=0 conjugacy fix-up (solver.c:1997).
"""
from __future__ import annotations

import math

import numpy as np


def wavenumbers(Nx: int, Ny: int):
    """kx (Nx,1), ky (1,Ny/2+1) integer wavenumbers, solver.c:2048-2067."""
    kx = np.where(np.arange(Nx) <= Nx // 2, np.arange(Nx), np.arange(Nx) - Nx).astype(np.int64)
    return kx[:, None], np.arange(Ny // 2 + 1, dtype=np.int64)[None, :]


def dealias_factor(Nx: int, Ny: int, filt: str = "23", kolm: bool = False) -> np.ndarray:
    """ApplyDealiasing multiplier (solver.c:1039-1099) for the IC."""
    kx, ky = wavenumbers(Nx, Ny)
    k2 = (kx * kx + ky * ky).astype(np.float64)
    if filt == "23":
        kmax_sqr = float(int(Nx / 3.0)) ** 2
        f = np.where(k2 >= kmax_sqr, 0.0, 1.0)
        if kolm:
            f = np.where(kx == 0, 0.0, f)
        return f
    # the C library's exp/pow, as the reference evaluates the filter
    libm_exp = np.vectorize(math.exp, otypes=[np.float64])
    libm_pow = np.vectorize(math.pow, otypes=[np.float64])
    if filt == "houli_ref":
        qx = (np.sign(kx) * (np.abs(kx) // (Nx // 2))).astype(np.float64)
        qy = (ky // (Ny // 2)).astype(np.float64)
        return libm_exp(-36.0 * libm_pow(np.sqrt(qx ** 2 + qy ** 2), 36.0))
    if filt == "houli":
        r = np.sqrt((kx / (Nx / 2.0)) ** 2 + (ky / (Ny / 2.0)) ** 2)
        return libm_exp(-36.0 * libm_pow(r, 36.0))
    return np.ones((Nx, Ny // 2 + 1))


def force_conjugacy(w_hat: np.ndarray) -> None:
    """ForceConjugacy on the IC (solver.c:2074-2107): w[kx<0,0] = conj(w[|kx|,0])."""
    Nx = w_hat.shape[0]
    w_hat[Nx // 2 + 1:, 0] = np.conj(w_hat[Nx // 2 - 1:0:-1, 0])


def energy(w_hat: np.ndarray, Nx: int, Ny: int) -> float:
    """Total energy as ComputeSystemMeasurables forms it (sys_msr.c:191,220,362)."""
    kx, ky = wavenumbers(Nx, Ny)
    k2 = (kx * kx + ky * ky).astype(np.float64)
    c = np.full((1, Ny // 2 + 1), 2.0)
    c[0, 0] = c[0, -1] = 1.0
    a2 = w_hat.real ** 2 + w_hat.imag ** 2   # cabs(w * conj(w))
    s = np.sum(np.where(k2 > 0, c * a2 / np.where(k2 > 0, k2, 1.0), 0.0))
    const_fac = 4.0 * math.pi ** 2
    norm_fac = 0.5 / float(Nx * Ny) ** 2
    return float(s) * const_fac * norm_fac


def palinstrophy(w_hat: np.ndarray, Nx: int, Ny: int) -> float:
    """Total palinstrophy as ComputeSystemMeasurables forms it (sys_msr.c:195,223,362)."""
    kx, ky = wavenumbers(Nx, Ny)
    k2 = (kx * kx + ky * ky).astype(np.float64)
    c = np.full((1, Ny // 2 + 1), 2.0)
    c[0, 0] = c[0, -1] = 1.0
    a2 = w_hat.real ** 2 + w_hat.imag ** 2
    s = np.sum(c * k2 * a2)
    return float(s) * 4.0 * math.pi ** 2 * (0.5 / float(Nx * Ny) ** 2)


def random_phase_ep(Nx: int, Ny: int, seed: int, k0: float = 6.0, e0: float = 0.5, p0: float = 5000.0,
                    filt: str = "23") -> np.ndarray:
    """random_phase's k^4 exp(-(k/k0)^2) field projected onto fixed energy e0
    and palinstrophy p0 (the initial conditions of extreme-enstrophy searches):
    w_hat -> a (|k|/k0)^b w_hat with b found by bisection so that P/E = p0/e0
    (the ratio grows monotonically with b), then a so that E = e0."""
    base = random_phase(Nx, Ny, seed, k0=k0, e0=e0, filt=filt)
    kx, ky = wavenumbers(Nx, Ny)
    q = np.sqrt((kx * kx + ky * ky).astype(np.float64)) / k0
    q[0, 0] = 1.0

    def ratio(b):
        w = base * q ** b
        return palinstrophy(w, Nx, Ny) / energy(w, Nx, Ny)

    target = p0 / e0
    lo, hi = -4.0, 4.0
    if not (ratio(lo) < target < ratio(hi)):
        raise ValueError(f"palinstrophy {p0} not reachable at energy {e0} for k0={k0}")
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if ratio(mid) < target:
            lo = mid
        else:
            hi = mid
    w = base * q ** (0.5 * (lo + hi))
    return w * math.sqrt(e0 / energy(w, Nx, Ny))


def random_phase(Nx: int, Ny: int, seed: int, k0: float = 8.0, e0: float = 0.5, filt: str = "23",
                 spectrum: str = "k4exp", kmin: float = 0.0, kmax: float | None = None) -> np.ndarray:
    """Seeded random-phase vorticity: |w_hat|^2 = |k| E(|k|) with E(k) ~ k^4 exp(-(k/k0)^2)
    ("k4exp") or E = 1 on kmin < |k| <= kmax ("flat"); uniform phases from
    numpy's PCG64 stream of ``seed``; Hermitian on the ky=0 and ky=Ny/2
    columns; dealiased; scaled to energy e0 (reference energy formula).  The
    nearest reference IC is DECAY_TURB_EXP (solver.c:1275-1292)."""
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
    amp = np.sqrt(np.where(k2 > 0, spec * kk, 0.0))
    phase = rng.uniform(0.0, 2.0 * np.pi, size=(Nx, Ny // 2 + 1))
    w_hat = amp * np.exp(1j * phase)
    w_hat[0, 0] = 0.0
    w_hat = w_hat * dealias_factor(Nx, Ny, filt)
    for j in (0, Ny // 2):
        w_hat[Nx // 2 + 1:, j] = np.conj(w_hat[Nx // 2 - 1:0:-1, j])
        w_hat[0, j] = w_hat[0, j].real
        w_hat[Nx // 2, j] = w_hat[Nx // 2, j].real
    e = energy(w_hat, Nx, Ny)
    if e > 0:
        w_hat *= math.sqrt(e0 / e)
    return w_hat


def cfl_timestep(w_hat: np.ndarray, Nx: int, Ny: int, nu: float, cfl: float = 0.5) -> float:
    """GetTimestep's CFL formula (solver.c:2381-2450) on the host, with every
    mode's velocity (the reference's u_hat indexing at solver.c:2394 offsets
    the rows): cfl / (pi (max|u|/dx + max|v|/dy) + pi^2 nu (1/dx^2 + 1/dy^2)),
    u = c2r(i ky w_hat/|k|^2)/(Nx Ny), v = c2r(-i kx w_hat/|k|^2)/(Nx Ny)."""
    kx, ky = wavenumbers(Nx, Ny)
    k2 = (kx * kx + ky * ky).astype(np.float64)
    inv = np.where(k2 > 0, 1.0 / np.where(k2 > 0, k2, 1.0), 0.0)
    u = np.fft.irfft2(1j * ky * inv * w_hat, s=(Nx, Ny), norm="forward") / (Nx * Ny)
    v = np.fft.irfft2(-1j * kx * inv * w_hat, s=(Nx, Ny), norm="forward") / (Nx * Ny)
    dx, dy = 2 * math.pi / Nx, 2 * math.pi / Ny
    conv = float(np.max(math.pi * (np.abs(u) / dx + np.abs(v) / dy)))
    return cfl / (conv + math.pi ** 2 * nu * (1 / dx ** 2 + 1 / dy ** 2))


def taylor_green(solver, kappa: float = 1.0, filt: str = "23") -> np.ndarray:
    """TG_VORT (solver.c:1161-1179): w = 2 kappa cos(kappa x) cos(kappa y), r2c on
    the GPU, then dealias and conjugacy fix-up."""
    Nx, Ny = solver.Nx, solver.Ny
    x = (np.arange(Nx) * 2.0 * math.pi / Nx)[:, None]
    y = (np.arange(Ny) * 2.0 * math.pi / Ny)[None, :]
    w = 2.0 * kappa * np.cos(kappa * x) * np.cos(kappa * y)
    w_hat = solver.r2c(w) * dealias_factor(Nx, Ny, filt)
    force_conjugacy(w_hat)
    return w_hat
