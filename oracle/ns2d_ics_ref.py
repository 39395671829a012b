"""CPU restatement of the reference's initial conditions and of glibc rand().

TEST INFRASTRUCTURE ONLY (imported by tests/ alone): the checker for the
driver's host library (libns2d_host.so), never part of the product path.

* ``GlibcRand`` restates glibc's ``srand``/``rand`` (random_r.c, the default
  TYPE_3 additive-feedback generator, degree 31, separation 3): the state is
  seeded with the Park-Miller recurrence ``r[i] = 16807 r[i-1] mod (2^31-1)``
  (Schrage's decomposition), ``r[31..33] = r[0..2]``, then
  ``r[i] = r[i-31] + r[i-3] mod 2^32`` and output ``r[i] >> 1``, discarding
  the first 310 outputs.  glibc 2.35 is the version of this image; the
  generator has been unchanged since glibc 2.0.  Pinned against outputs of the
  real ``rand()`` (tests/golden/glibc_rand_123456789.txt, made by
  tests/golden/gen_glibc_rand.c).
* ``initial_condition`` restates ``InitialConditions``
  (Solver/src/solver.c:1107-2014) for one process holding the whole grid, or
  as ``ranks`` MPI processes form it (``_Ranks``: one generator per slab of
  kx rows, the normalisation sums reduced as ``MPI_Allreduce`` does), then ``ApplyDealiasing`` (:1039-1099) and ``ForceConjugacy`` (:2074-2107),
  as :1996-2000 do.  The scalar libm calls (``math.*``) and the operation
  order of each formula follow the C source so that spectral initial
  conditions agree bit for bit with a C build; the real-space ones go through
  the transforms passed in.

Conventions as ``ns2d_oracle``: ``w_hat`` complex128 ``[Nx][Ny//2+1]``,
FFTW units, ``kx = 0..Nx/2, -Nx/2+1..-1``.
"""
from __future__ import annotations

import ctypes
import ctypes.util
import math

import numpy as np

RAND_MAX = 2147483647
SEED = 123456789            # solver.c:1121

# data_types.h:141-165
KAPPA = 1.0
SIGMA = 15.0 / math.pi
DELTA = 0.005
BETA = 1.0
S = 6.0
DT_K0 = 6.0
DT_E0 = 0.5
DT2_K0 = 30
DT2_E0 = 0.5
DTEXP_K0 = 8.0
DTEXP_E0 = 343.0 / 96.0
RING_MIN_K = 3.0
RING_MAX_K = 10.0
EXTRM_ENS_MIN_K = 1.5
EXTRM_ENS_POW = 1.5
UNIF_MIN_K = 2.5
UNIF_MAX_K = 9.5
RAND_ENST0 = 1.0


class GlibcRand:
    """glibc srand(seed); rand() restated (TYPE_3, see module docstring)."""

    def __init__(self, seed: int):
        seed &= 0xFFFFFFFF
        if seed == 0:
            seed = 1
        r = [0] * 34
        r[0] = seed - (1 << 32) if seed >= (1 << 31) else seed
        for i in range(1, 31):
            # Schrage: 16807 * r mod (2^31 - 1) without overflow; C's division
            # truncates toward zero
            a = r[i - 1]
            hi = abs(a) // 127773 if a >= 0 else -(abs(a) // 127773)
            lo = a - hi * 127773
            word = 16807 * lo - 2836 * hi
            if word < 0:
                word += 2147483647
            r[i] = word
        for i in range(31, 34):
            r[i] = r[i - 31]
        self._r = [x & 0xFFFFFFFF for x in r]
        for _ in range(310):
            self._next()

    def _next(self) -> int:
        r = self._r
        v = (r[-31] + r[-3]) & 0xFFFFFFFF
        r.append(v)
        if len(r) > 64:
            del r[:-34]
        return v >> 1

    def rand(self) -> int:
        return self._next()

    def uniform(self) -> float:
        """(double) rand() / (double) RAND_MAX"""
        return float(self._next()) / float(RAND_MAX)


class _Ranks:
    """The MPI ranks of a P-process reference run as InitialConditions sees
    them: the kx rows in FFTW's default block distribution (solver.c:2615,
    ``FFTW_MPI_DEFAULT_BLOCK``: ceil(Nx/P) rows per rank, the last ranks
    short or empty), one generator per rank, each seeded by the rank's own
    ``srand(123456789)`` (solver.c:1121) and kept across the IC's loops, and
    one partial sum per rank for each ``MPI_Allreduce``."""

    def __init__(self, Nx: int, P: int):
        if P < 1:
            raise ValueError("rank count must be at least 1")
        self.P = P
        self.rows = -(-Nx // P)
        self._gens = {}

    def rng(self, i: int) -> GlibcRand:
        """the generator of the rank that owns kx row i"""
        r = i // self.rows
        if r not in self._gens:
            self._gens[r] = GlibcRand(SEED)
        return self._gens[r]

    def rank(self, i: int) -> int:
        return i // self.rows

    def sums(self) -> list:
        return [0.0] * self.P

    def allreduce(self, part: list) -> float:
        """MPI_Allreduce(MPI_SUM) of one double per rank by recursive
        doubling, message by message as MPICH's
        MPIR_Allreduce_intra_recursive_doubling and Open MPI's
        ompi_coll_base_allreduce_intra_recursivedoubling exchange them: the
        first 2*(P - P2) ranks fold pairwise into the odd rank (P2 the
        largest power of two <= P), then P2 ranks swap and add with the rank
        whose new index differs in bit 0, 1, 2, ...  Returns rank 0's
        result (every rank ends with the same bits)."""
        P = self.P
        buf = list(part)
        pof2 = 1 << (P.bit_length() - 1)
        rem = P - pof2
        real = {}
        for r in range(P):
            if r < 2 * rem:
                if r % 2 == 0:
                    continue          # sends to r + 1 and takes no further part
                buf[r] = buf[r - 1] + buf[r]
                real[r // 2] = r
            else:
                real[r - rem] = r
        mask = 1
        while mask < pof2:
            buf_next = dict((r, buf[real[nr ^ mask]] + buf[r]) for nr, r in real.items())
            for r, v in buf_next.items():
                buf[r] = v
            mask <<= 1
        out = [buf[r] for r in real.values()]
        assert all(v == out[0] for v in out), "ranks disagree after the reduction"
        return buf[real[0]]


def _kx(i: int, Nx: int) -> int:
    return i if i <= Nx // 2 else i - Nx


def _x(i: int, N: int) -> float:
    return float(i) * 2.0 * math.pi / float(N)


_libm = ctypes.CDLL(ctypes.util.find_library("m"))
_libm.sincos.argtypes = [ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
_libm.sincos.restype = None


def _expi(theta: float) -> complex:
    """cexp(i theta) of glibc: sincos(theta) times exp(0) = 1 (glibc's cexp
    and gcc's expansion of it both call sincos, which can differ from
    separate sin and cos calls in the last bit)."""
    s, c = ctypes.c_double(), ctypes.c_double()
    _libm.sincos(theta, ctypes.byref(s), ctypes.byref(c))
    return complex(c.value, s.value)


def _scale(s: float, z: complex) -> complex:
    """real * complex in C99: each part multiplied once."""
    return complex(s * z.real, s * z.imag)


def _abs2(z: complex) -> float:
    """cabs(z * conj(z))"""
    return z.real * z.real + z.imag * z.imag


def dealias(f: np.ndarray, Nx: int, Ny: int, filt: str, forcing: str) -> None:
    """ApplyDealiasing, solver.c:1039-1099 (in place, one field)."""
    Nyf = Ny // 2 + 1
    kmax_sqr = math.pow(int(Nx / 3.0), 2.0)
    for i in range(Nx):
        k0 = _kx(i, Nx)
        for j in range(Nyf):
            if filt == "HOU_LI":
                q0 = int(k0 / (Nx // 2))              # C integer division
                q1 = int(j / (Ny // 2))
                g = math.exp(-36.0 * math.pow(math.sqrt(math.pow(float(q0), 2.0) + math.pow(float(q1), 2.0)), 36.0))
                f[i, j] = _scale(g, complex(f[i, j]))
            else:
                k_sqr = float(k0) * k0 + j * j
                if k_sqr >= kmax_sqr or (forcing == "KOLM" and k0 == 0):
                    f[i, j] = 0.0


def force_conjugacy(w: np.ndarray, Nx: int) -> None:
    """ForceConjugacy(w_hat, N), solver.c:2074-2107."""
    for i in range(Nx):
        k0 = _kx(i, Nx)
        if k0 < 0:
            w[i, 0] = np.conj(w[-k0, 0])


def _decay_turb(Nx: int, Ny: int, u0: str, R: _Ranks) -> np.ndarray:
    """solver.c:1246-1361"""
    Nyf = Ny // 2 + 1
    w = np.zeros((Nx, Nyf), complex)
    for i in range(Nx):
        k0 = _kx(i, Nx)
        for j in range(Nyf):
            if k0 == 0 and j == 0:
                continue
            sk = math.sqrt(float(k0 * k0 + j * j))
            if u0 == "DECAY_TURB_BB":
                spec = sk / (1.0 + math.pow(sk, 4.0) / DT_K0)
            elif u0 == "DECAY_TURB_NB":
                spec = math.pow(sk, 6.0) / math.pow((1.0 + sk / (2.0 * DT2_K0)), 18.0)
            else:
                spec = math.pow(sk, 7.0) / math.pow(DTEXP_K0, 8.0) * math.exp(-3.5 * math.pow(sk / DTEXP_K0, 2.0))
            u1 = R.rng(i).uniform()
            if u0 == "DECAY_TURB_EXP":
                amp = math.sqrt(sk * spec / math.pi)
            else:
                amp = math.sqrt(spec * (sk / (2.0 * math.pi)))
            w[i, j] = _scale(amp, _expi(2.0 * math.pi * u1))
    part = R.sums()
    for i in range(Nx):
        k0 = _kx(i, Nx)
        for j in range(Nyf):
            if k0 != 0 or j != 0:
                inv = 1.0 / float(k0 * k0 + j * j)
                a2 = _abs2(complex(w[i, j]))
                part[R.rank(i)] += inv * a2 if (j == 0 or j == Nyf - 1) else 2.0 * inv * a2
    # solver.c:1331-1334: each rank scales its sum, then the sums are reduced
    enrg = R.allreduce([e * ((0.5 / math.pow(Nx * Ny, 2.0)) * 4.0 * math.pow(math.pi, 2.0)) for e in part])
    e0 = {"DECAY_TURB_BB": DT_E0, "DECAY_TURB_NB": DT2_E0}.get(u0, DTEXP_E0)
    s = math.sqrt(e0 / enrg)
    for i in range(Nx):
        k0 = _kx(i, Nx)
        for j in range(Nyf):
            w[i, j] = _scale(s, complex(w[i, j])) if (k0 == 0 or j != 0) else 0.0
    return w


def _decay_turb_alt(Nx: int, Ny: int, R: _Ranks, r2c, c2r) -> np.ndarray:
    """solver.c:1362-1522"""
    Nyf = Ny // 2 + 1
    psi_hat = np.zeros((Nx, Nyf), complex)
    for i in range(Nx):
        k0 = _kx(i, Nx)
        for j in range(Nyf):
            u1 = R.rng(i).uniform()
            u2 = R.rng(i).uniform()
            # cos and sin of one argument: gcc -O2 and above emits one sincos call
            e = _expi(2.0 * math.pi * u2)
            z1 = math.sqrt(-2.0 * math.log(u1)) * e.real
            z2 = math.sqrt(-2.0 * math.log(u1)) * e.imag
            if k0 == 0 and j == 0:
                continue
            ks = math.sqrt(float(k0 * k0 + j * j))
            spec = 1.0 / (ks * (1.0 + math.pow(ks / 6.0, 4.0)))
            psi_hat[i, j] = _scale(math.sqrt(spec * math.pow(ks, 2.0) / (math.pow(ks, 3.0) * 2.0 * math.pi)),
                                   complex(z1, z2))
    psi = np.array(c2r(psi_hat), dtype=np.float64)
    part = R.sums()
    for i in range(Nx):
        for j in range(Ny):
            psi[i, j] /= float(Nx * Ny)
            part[R.rank(i)] += psi[i, j]
    mean = R.allreduce(part)
    psi -= mean / float(Nx * Ny)
    psi_hat = np.array(r2c(psi), dtype=complex)
    part = R.sums()
    for i in range(Nx):
        k0 = _kx(i, Nx)
        for j in range(Nyf):
            k2 = float(k0 * k0 + j * j)
            a2 = _abs2(complex(psi_hat[i, j]))
            part[R.rank(i)] += k2 * a2 if (j == 0 or j == Nyf - 1) else 2.0 * k2 * a2
    # solver.c:1487-1490
    enrg = R.allreduce([e * (4.0 * math.pow(math.pi, 2.0) * (0.5 / math.pow(Nx * Ny, 2.0))) for e in part])
    s = math.sqrt(0.5 / enrg)
    w = np.zeros((Nx, Nyf), complex)
    for i in range(Nx):
        k0 = _kx(i, Nx)
        for j in range(Nyf):
            p = _scale(s, complex(psi_hat[i, j]))
            if k0 != 0 or j != 0:
                w[i, j] = _scale(-float(k0 * k0 + j * j), p)
    return w


def _gauss_decay_turb(Nx: int, Ny: int, R: _Ranks) -> np.ndarray:
    """solver.c:1523-1606"""
    Nyf = Ny // 2 + 1
    psi = np.zeros((Nx, Nyf), complex)
    for i in range(Nx):
        k0 = _kx(i, Nx)
        for j in range(Nyf):
            if k0 == 0 and j == 0:
                continue
            ks = math.sqrt(float(k0 * k0 + j * j))
            spec = math.pow(ks, 6.0) / math.pow((1.0 + ks / 60.0), 18.0)
            u1 = R.rng(i).uniform()
            psi[i, j] = _scale(math.sqrt(spec * (1.0 / ks)), _expi(2.0 * math.pi * u1))
    part = R.sums()
    for i in range(Nx):
        k0 = _kx(i, Nx)
        for j in range(Nyf):
            if k0 != 0 or j != 0:
                k2 = float(k0 * k0 + j * j)
                a2 = _abs2(complex(psi[i, j]))
                part[R.rank(i)] += k2 * a2 if (j == 0 or j == Nyf - 1) else 2.0 * k2 * a2
    # solver.c:1578-1581
    enrg = R.allreduce([e * (4.0 * math.pow(math.pi, 2.0) * (0.5 / math.pow(Nx * Ny, 2.0))) for e in part])
    se, sh = math.sqrt(enrg), math.sqrt(0.5)
    w = np.zeros((Nx, Nyf), complex)
    for i in range(Nx):
        k0 = _kx(i, Nx)
        for j in range(Nyf):
            z = complex(psi[i, j])
            z = complex(z.real / se, z.imag / se)
            z = _scale(sh, z)
            if k0 != 0 or j != 0:
                w[i, j] = _scale(float(k0 * k0 + j * j), z)
    return w


def _extrm_ens(Nx: int, Ny: int, R: _Ranks) -> np.ndarray:
    """solver.c:1607-1670"""
    Nyf = Ny // 2 + 1
    P, K = EXTRM_ENS_POW, EXTRM_ENS_MIN_K
    part = R.sums()
    for i in range(Nx):
        k0 = _kx(i, Nx)
        for j in range(Nyf):
            k2 = float(k0 * k0 + j * j)
            edge = j == 0 or j == Nyf - 1
            if k2 > 0 and math.sqrt(k2) > K:
                p = math.pow(math.sqrt(k2), -P)
                part[R.rank(i)] += p if edge else 2.0 * p
            elif math.sqrt(k2) <= K:
                p = math.pow(K, -P)
                e = math.exp((math.sqrt(k2) - K) / 3.5)
                part[R.rank(i)] += p * e if edge else 2.0 * p * e
    norm = math.sqrt(K / R.allreduce(part))
    w = np.zeros((Nx, Nyf), complex)
    for i in range(Nx):
        k0 = _kx(i, Nx)
        for j in range(Nyf):
            k2 = float(k0 * k0 + j * j)
            if k2 > 0 and math.sqrt(k2) > K:
                u1 = R.rng(i).uniform()
                w[i, j] = _scale(math.pow(math.sqrt(k2), -P / 2.0) * norm, _expi(2.0 * math.pi * u1))
            elif math.sqrt(k2) <= K:
                u1 = R.rng(i).uniform()
                a = math.pow(K, -P / 2.0) * math.exp((math.sqrt(k2) - K) / 3.5) * norm
                w[i, j] = _scale(a, _expi(2.0 * math.pi * u1))
    return w


def _ring(Nx: int, Ny: int, R: _Ranks) -> np.ndarray:
    """solver.c:1671-1718"""
    Nyf = Ny // 2 + 1
    mid = (RING_MIN_K + RING_MAX_K) / 2.0
    inside = lambda ak, k0, j: (RING_MIN_K <= ak <= RING_MAX_K) and (j != 0 or k0 > 0)
    part = R.sums()
    for i in range(Nx):
        k0 = _kx(i, Nx)
        for j in range(Nyf):
            ak = math.sqrt(float(k0 * k0 + j * j))
            R.rng(i).uniform()
            if inside(ak, k0, j):
                part[R.rank(i)] += math.exp(-math.pow(ak - mid, 2.0))
    norm = R.allreduce(part)
    w = np.zeros((Nx, Nyf), complex)
    for i in range(Nx):
        k0 = _kx(i, Nx)
        for j in range(Nyf):
            ak = math.sqrt(float(k0 * k0 + j * j))
            if inside(ak, k0, j):
                u1 = R.rng(i).uniform()
                a = math.sqrt(72.0 * math.exp(-math.pow(ak - mid, 2.0)) / (math.pi * ak * norm))
                w[i, j] = _scale(a, _expi(2.0 * math.pi * u1))
    return w


def _unif(Nx: int, Ny: int, R: _Ranks) -> np.ndarray:
    """solver.c:1719-1765"""
    Nyf = Ny // 2 + 1
    part = R.sums()
    for i in range(Nx):
        k0 = _kx(i, Nx)
        for j in range(Nyf):
            ka = math.sqrt(float(k0 * k0 + j * j))
            if UNIF_MIN_K < ka < UNIF_MAX_K:
                part[R.rank(i)] += 1.0 if (j == 0 or j == Nyf - 1) else 2.0
    norm = math.sqrt(1.0 / R.allreduce(part))
    w = np.zeros((Nx, Nyf), complex)
    for i in range(Nx):
        k0 = _kx(i, Nx)
        for j in range(Nyf):
            ka = math.sqrt(float(k0 * k0 + j * j))
            if UNIF_MIN_K < ka < UNIF_MAX_K:
                u1 = R.rng(i).uniform()
                w[i, j] = _scale(norm, _expi(2.0 * math.pi * u1))
    return w


def _random(Nx: int, Ny: int, R: _Ranks) -> np.ndarray:
    """solver.c:1900-1951 (its enstrophy sum starts at zero here)"""
    Nyf = Ny // 2 + 1
    w = np.zeros((Nx, Nyf), complex)
    part = R.sums()
    for i in range(Nx):
        k0 = _kx(i, Nx)
        for j in range(Nyf):
            sk = math.sqrt(float(k0 * k0 + j * j))
            if UNIF_MIN_K < sk < UNIF_MAX_K:
                ek = 1.0 / sk
                zk = math.pow(sk, 2.0) * ek
                r1 = R.rng(i).uniform()
                w[i, j] = _scale(math.sqrt(zk / ek), _expi(r1 * 2.0 * math.pi))
            z = complex(w[i, j])
            part[R.rank(i)] += math.pow(math.hypot(z.real, z.imag), 2.0)
    enst = R.allreduce(part) * 2.0
    s = math.sqrt(RAND_ENST0 / enst)
    for i in range(Nx):
        k0 = _kx(i, Nx)
        for j in range(Nyf):
            w[i, j] = _scale(s, complex(w[i, j])) if (k0 != 0 or j != 0) else 0.0
    return w


def _testing(Nx: int, Ny: int) -> np.ndarray:
    """solver.c:1952-1983"""
    Nyf = Ny // 2 + 1
    w = np.zeros((Nx, Nyf), complex)
    for i in range(Nx):
        k0 = _kx(i, Nx)
        for j in range(Nyf):
            if k0 == 0 and j == 0:
                continue
            inv = 1.0 / float(k0 * k0 + j * j)
            th = -math.pi / 4.0 if (j == 0 and k0 < 0) else math.pi / 4.0
            w[i, j] = _scale(inv, _expi(th))
    return w


def _none(Nx: int, Ny: int, R: _Ranks) -> np.ndarray:
    """solver.c:1984-1995: amplitude drawn first, then the phase."""
    Nyf = Ny // 2 + 1
    w = np.zeros((Nx, Nyf), complex)
    for i in range(Nx):
        for j in range(Nyf):
            amp = R.rng(i).uniform()
            ph = R.rng(i).uniform()
            w[i, j] = _scale(amp, _expi(ph * 2.0 * math.pi))
    return w


def _zero(Nx: int, Ny: int, kolm: bool) -> np.ndarray:
    """solver.c:1214-1245"""
    Nyf = Ny // 2 + 1
    w = np.zeros((Nx, Nyf), complex)
    if kolm:
        for i in range(Nx):
            k0 = _kx(i, Nx)
            for j in range(Nyf):
                if j == 1 and k0 == 1:
                    w[i, j] = -0.5
                elif j == 2 and k0 in (3, 4):
                    w[i, j] = -4.0 + 1.0j
                elif j == 3 and k0 in (4, 5):
                    w[i, j] = -2.0 + 3.0j
                elif j == 4 and k0 in (5, 6):
                    w[i, j] = 8.0 - 6.0j
    return w


def _real_space(Nx: int, Ny: int, u0: str, r2c) -> np.ndarray:
    """TG_VORT (solver.c:1161-1179), DOUBLE_SHEAR_LAYER (:1180-1213) and
    GAUSS_BLOB (:1881-1899, written with the collocation points)."""
    w = np.zeros((Nx, Ny))
    for i in range(Nx):
        x = _x(i, Nx)
        for j in range(Ny):
            y = _x(j, Ny)
            if u0 == "TG_VORT":
                v = 2.0 * KAPPA * math.cos(KAPPA * x) * math.cos(KAPPA * y)
            elif u0 == "DOUBLE_SHEAR_LAYER":
                v = DELTA * math.cos(y) - SIGMA / math.pow(math.cosh(SIGMA * (x - 0.5 * math.pi)), 2.0)
                v += DELTA * math.cos(y) + SIGMA / math.pow(math.cosh(SIGMA * (1.5 * math.pi - x)), 2.0)
            else:
                v = -math.exp(-(math.pow(x - math.pi, 2.0) + BETA * math.pow(y - math.pi, 2.0))
                              / math.pow(2.0 * math.pi / S, 2.0))
            w[i, j] = v
    return np.array(r2c(w), dtype=complex)


def _tg_vel(Nx: int, Ny: int, filt: str, forcing: str, r2c) -> np.ndarray:
    """solver.c:1124-1160"""
    Nyf = Ny // 2 + 1
    u = np.zeros((Nx, Ny))
    v = np.zeros((Nx, Ny))
    for i in range(Nx):
        x = _x(i, Nx)
        for j in range(Ny):
            y = _x(j, Ny)
            ex, ey = _expi(KAPPA * x), _expi(KAPPA * y)
            u[i, j] = ex.real * ey.imag
            v[i, j] = -ex.imag * ey.real
    uh = np.array(r2c(u), dtype=complex)
    vh = np.array(r2c(v), dtype=complex)
    dealias(uh, Nx, Ny, filt, forcing)
    dealias(vh, Nx, Ny, filt, forcing)
    w = np.zeros((Nx, Nyf), complex)
    for i in range(Nx):
        k0 = _kx(i, Nx)
        for j in range(Nyf):
            if k0 == 0 or j != 0:
                a = complex(vh[i, j]); b = complex(uh[i, j])
                d = complex(k0 * a.real - j * b.real, k0 * a.imag - j * b.imag)
                w[i, j] = complex(-d.imag, d.real)
    return w


def _max_palin(Nx: int, Ny: int, psi: np.ndarray, r2c) -> np.ndarray:
    """solver.c:1766-1880 with the stream function already read"""
    ph = np.array(r2c(np.asarray(psi, float).reshape(Nx, Ny)), dtype=complex)
    Nyf = Ny // 2 + 1
    for i in range(Nx):
        k0 = _kx(i, Nx)
        for j in range(Nyf):
            ph[i, j] = _scale(-float(k0 * k0 + j * j), complex(ph[i, j]))
    return ph


def initial_condition(u0: str, Nx: int, Ny: int, filt: str = "23", forcing: str = "NONE",
                      r2c=None, c2r=None, psi: np.ndarray | None = None, ranks: int = 1) -> np.ndarray:
    """InitialConditions + ApplyDealiasing + ForceConjugacy (solver.c:1107-2000)
    as a run on ``ranks`` MPI processes forms it (``_Ranks``)."""
    if r2c is None:
        r2c = lambda a: np.fft.rfft2(a)
    if c2r is None:
        c2r = lambda a: np.fft.irfft2(a, s=(Nx, Ny), norm="forward")
    R = _Ranks(Nx, ranks)
    if u0 == "TG_VEL":
        w = _tg_vel(Nx, Ny, filt, forcing, r2c)
    elif u0 in ("TG_VORT", "DOUBLE_SHEAR_LAYER", "GAUSS_BLOB"):
        w = _real_space(Nx, Ny, u0, r2c)
    elif u0 in ("ZERO", "ZERO_KOLM"):
        w = _zero(Nx, Ny, u0 == "ZERO_KOLM")
    elif u0 in ("DECAY_TURB_BB", "DECAY_TURB_NB", "DECAY_TURB_EXP"):
        w = _decay_turb(Nx, Ny, u0, R)
    elif u0 == "DECAY_TURB_ALT":
        w = _decay_turb_alt(Nx, Ny, R, r2c, c2r)
    elif u0 == "GAUSS_DECAY_TURB":
        w = _gauss_decay_turb(Nx, Ny, R)
    elif u0 == "EXTRM_ENS":
        w = _extrm_ens(Nx, Ny, R)
    elif u0 == "RING":
        w = _ring(Nx, Ny, R)
    elif u0 == "UNIF":
        w = _unif(Nx, Ny, R)
    elif u0 == "MAX_PALIN":
        w = _max_palin(Nx, Ny, psi, r2c)
    elif u0 == "RANDOM":
        w = _random(Nx, Ny, R)
    elif u0 == "TESTING":
        w = _testing(Nx, Ny)
    else:
        w = _none(Nx, Ny, R)
    dealias(w, Nx, Ny, filt, forcing)
    force_conjugacy(w, Nx)
    return w


def mean_flows(w: np.ndarray, Nx: int, Ny: int):
    """One measurement's mean-flow sums, sys_msr.c:196-230 (y) and, with each
    mode's own ky and |k|^2, :339-347 (x)."""
    Nyf = Ny // 2 + 1
    my = np.zeros(Nx)
    mx = np.zeros(Nyf)
    for i in range(Nx):
        k0 = _kx(i, Nx)
        for j in range(Nyf):
            if k0 == 0 and j == 0:
                continue
            k2 = float(k0 * k0 + j * j)
            a = abs(complex(w[i, j]))
            vz = abs(k0) * a / k2
            uz = j * a / k2
            wgt = 1.0 if (j == 0 or j == Nyf - 1) else 2.0
            my[i] += wgt * vz
            mx[j] += wgt * uz
    return my, mx
