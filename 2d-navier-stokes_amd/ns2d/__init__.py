"""Python host interface to libns2d_gpu.so (the MI355X vorticity solver).

A thin ctypes layer over the C ABI in ``include/ns2d_gpu.h``.  Method names
follow the reference solver's functions (EndCar808/2D-Navier-Stokes,
``Solver/src``) so that tests read like the reference's own call sequence:

=====================  ===============================================
``Solver`` method       reference function
=====================  ===============================================
``spectral_step``       ``RK4Step`` / ``AB4Step`` (solver.c:625, :261)
``nonlinear_rhs``       ``NonlinearRHSBatch`` (solver.c:884-1032)
``system_measures``     ``ComputeSystemMeasurables`` (sys_msr.c:32-373)
``max_vort``            ``GetMaxData("VORT")`` (solver.c:2171-2295)
``real_vorticity``      the ``w`` dataset transform (hdf5_funcs.c:174)
``r2c``                 ``fftw_mpi_execute_dft_r2c`` on a real field
=====================  ===============================================

The library must be built (``__graft_entry__.build()``); importing raises if
it is missing -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import dataclasses
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libns2d_gpu.so")

NS2D_OK, NS2D_ERR_INVALID, NS2D_ERR_HIP, NS2D_ERR_BLOWUP, NS2D_ERR_STATE = 0, 1, 2, 3, 4
FILTERS = {"23": 0, "houli_ref": 1, "houli": 2, "none": 3}
FORCINGS = {"NONE": 0, "KOLM": 1}
INTEGRATORS = {"RK4CN": 0, "AB4": 1}
QUIRK_CONJ_GLOBAL, QUIRK_VISC_ASSIGN, QUIRK_KOLM_UNSCALED = 0x01, 0x02, 0x04
QUIRK_KOLM_KX0_KILL, QUIRK_ENST_DISS = 0x08, 0x10
QUIRKS_REF = 0x1F


class _Params(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nu", ctypes.c_double),
                ("hyper_flag", ctypes.c_int), ("hyper_pow", ctypes.c_double),
                ("alpha", ctypes.c_double), ("drag_flag", ctypes.c_int), ("drag_pow", ctypes.c_double),
                ("filter", ctypes.c_int), ("forcing", ctypes.c_int), ("force_k", ctypes.c_int),
                ("force_scale", ctypes.c_double), ("integrator", ctypes.c_int),
                ("quirks", ctypes.c_uint), ("check_every", ctypes.c_int), ("device", ctypes.c_int)]


class _Measures(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "energy", "enstrophy", "palinstrophy", "divergence", "enrg_diss", "enst_diss", "forcing",
        "d_enrg_dt_sbst", "enrg_flux_sbst", "enrg_diss_sbst",
        "d_enst_dt_sbst", "enst_flux_sbst", "enst_diss_sbst")]


class _Info(ctypes.Structure):
    _fields_ = [("iters", ctypes.c_long), ("first_blowup_iter", ctypes.c_long),
                ("last_max_vort", ctypes.c_double), ("hyper_flag", ctypes.c_int),
                ("n_spect", ctypes.c_int), ("ld", ctypes.c_int)]


EXPORTS = ["ns2d_params_default", "ns2d_create", "ns2d_destroy", "ns2d_set_state", "ns2d_get_state",
           "ns2d_set_state_device", "ns2d_get_state_device", "ns2d_step", "ns2d_rhs", "ns2d_measures",
           "ns2d_max_vort", "ns2d_cfl_rate", "ns2d_real_vort", "ns2d_r2c", "ns2d_c2r", "ns2d_get_info", "ns2d_sync",
           "ns2d_get_stream", "ns2d_profile_steps", "ns2d_variant_count", "ns2d_variant_info", "ns2d_variant_extra",
           "ns2d_default_variant", "ns2d_set_variant", "ns2d_ens_create", "ns2d_ens_member", "ns2d_ens_step", "ns2d_ens_autotune", "ns2d_ens_variants", "ns2d_ens_sync",
           "ns2d_ens_destroy", "ns2d_vjp", "ns2d_vjp_device", "ns2d_measures_noalloc", "ns2d_layout", "ns2d_last_error"]

_lib = None


def _one_hip_runtime() -> None:
    """A process gets one HIP/HSA runtime.  torch's wheel carries its own
    runtime (libamdhip64.so with the same soname as /opt/rocm's, and its own
    libhsa-runtime64.so); when libns2d_gpu.so loads /opt/rocm's first and torch
    is imported later, torch brings a second HSA runtime and one of the two
    sees no GPU.  Importing torch first makes this library bind to torch's
    runtime, so both work in one process.  NS2D_NO_TORCH=1 skips it."""
    if os.environ.get("NS2D_NO_TORCH") == "1":
        return
    try:
        import torch  # noqa: F401
    except Exception:
        pass


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libns2d_gpu.so (raises OSError when it has not been built).
    NS2D_LIB names another build of the library to load instead."""
    global _lib
    if _lib is not None:
        return _lib
    path = os.environ.get("NS2D_LIB", path)
    if not os.path.exists(path):
        raise OSError(f"{path} not built: run __graft_entry__.build()")
    _one_hip_runtime()
    lib = ctypes.CDLL(path)
    P = ctypes.POINTER
    vp, dp, i, d, l = ctypes.c_void_p, P(ctypes.c_double), ctypes.c_int, ctypes.c_double, ctypes.c_long
    sig = {
        "ns2d_params_default": ([P(_Params), i, i], i),
        "ns2d_create": ([P(_Params), P(vp)], i),
        "ns2d_destroy": ([vp], i),
        "ns2d_set_state": ([vp, dp], i),
        "ns2d_get_state": ([vp, dp], i),
        "ns2d_set_state_device": ([vp, vp], i),
        "ns2d_get_state_device": ([vp, vp], i),
        "ns2d_step": ([vp, d, l], i),
        "ns2d_rhs": ([vp, dp, dp], i),
        "ns2d_measures": ([vp, i, P(_Measures), dp, dp, dp], i),
        "ns2d_measures_noalloc": ([vp, i, P(_Measures), dp, dp, dp], i),
        "ns2d_max_vort": ([vp, dp], i),
        "ns2d_cfl_rate": ([vp, dp], i),
        "ns2d_real_vort": ([vp, dp], i),
        "ns2d_r2c": ([vp, dp, dp], i),
        "ns2d_c2r": ([vp, dp, dp], i),
        "ns2d_get_info": ([vp, P(_Info)], i),
        "ns2d_sync": ([vp], i),
        "ns2d_get_stream": ([vp, P(vp)], i),
        "ns2d_profile_steps": ([vp, d, l, dp, dp, dp], i),
        "ns2d_variant_count": ([i, i], i),
        "ns2d_variant_info": ([i, i, i, P(i)], i),
        "ns2d_variant_extra": ([i, i, i, P(i)], i),
        "ns2d_default_variant": ([i, i, i, P(i)], i),
        "ns2d_set_variant": ([vp, i, i], i),
        "ns2d_layout": ([P(i), P(i), P(i)], i),
        "ns2d_ens_create": ([P(_Params), i, P(vp)], i),
        "ns2d_ens_member": ([vp, i, P(vp)], i),
        "ns2d_ens_step": ([vp, dp, l], i),
        "ns2d_ens_autotune": ([vp, i], i),
        "ns2d_ens_variants": ([vp, i, P(i), P(i)], i),
        "ns2d_ens_sync": ([vp], i),
        "ns2d_ens_destroy": ([vp], i),
        "ns2d_vjp": ([vp, dp, d, l, dp, dp, l], i),
        "ns2d_vjp_device": ([vp, vp, d, l, vp, vp, l], i),
        "ns2d_last_error": ([], ctypes.c_char_p),
    }
    for name, (argt, rest) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = argt
        fn.restype = rest
    _lib = lib
    return lib


PASS_COLUMN, PASS_ROW = 0, 1


def variants(pass_kind: int, n: int) -> list:
    """Compiled shapes of one pass for an n-point transform.  Column pass:
    dicts with cols, per_thread, threads, splits; row pass: threads_per_pair,
    per_thread, threads, pairs."""
    lib = load_library()
    keys = (("cols", "per_thread", "threads", "splits") if pass_kind == PASS_COLUMN
            else ("threads_per_pair", "per_thread", "threads", "pairs"))
    out = []
    for i in range(lib.ns2d_variant_count(pass_kind, n)):
        v = (ctypes.c_int * 4)()
        _check(lib.ns2d_variant_info(pass_kind, n, i, v))
        d = dict(zip(keys, list(v)))
        x = (ctypes.c_int * 4)()
        _check(lib.ns2d_variant_extra(pass_kind, n, i, x))
        d.update(min_waves=x[0], pad_shift=x[1], line_offset=x[2], packed_pair=x[3] & 1, prefetch=(x[3] >> 1) & 1,
                 pair_in_registers=(x[3] >> 2) & 1, single_row=(x[3] >> 3) & 1, half_forward=(x[3] >> 4) & 1,
                 twiddle_powers=(x[3] >> 5) & 1)
        out.append(d)
    return out


def kpass_variants(nx: int) -> list:
    return variants(PASS_COLUMN, nx)


def rpass_variants(ny: int) -> list:
    return variants(PASS_ROW, ny)


def default_variant(pass_kind: int, n: int, integrator: str = "RK4CN") -> int:
    """Index of the shape a new solver with this integrator runs for an
    n-point transform (before NS2D_KPASS_VARIANT_<Nx> / NS2D_RPASS_VARIANT_<Ny>)."""
    lib = load_library()
    idx = ctypes.c_int()
    _check(lib.ns2d_default_variant(pass_kind, n, INTEGRATORS[integrator], ctypes.byref(idx)))
    return idx.value


def layout() -> tuple:
    """(k-space x-tile, column-to-row mixed-space x-tile, row-to-column
    mixed-space x-tile) the library stores its device arrays in; k-space 0 =
    column-major [Ny/2+1][Nx]."""
    lib = load_library()
    k, g, n = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _check(lib.ns2d_layout(ctypes.byref(k), ctypes.byref(g), ctypes.byref(n)))
    return k.value, g.value, n.value


class NS2DError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[ns2d error {code}] {msg}")
        self.code = code


def _check(rc: int) -> None:
    if rc != NS2D_OK:
        raise NS2DError(rc, _lib.ns2d_last_error().decode())


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@dataclasses.dataclass
class Config:
    """Solver configuration; field names follow the reference CLI (utils.c:34-461)."""
    Nx: int
    Ny: int
    nu: float = 1.0
    hyper_flag: int = 0
    hyper_pow: float = 2.0
    alpha: float = 1.0
    drag_flag: int = 0
    drag_pow: float = -2.0
    filter: str = "23"
    forcing: str = "NONE"
    force_k: int = 0
    force_scale: float = 1.0
    integrator: str = "RK4CN"
    quirks: int = QUIRKS_REF
    check_every: int = 1
    device: int = 0

    def _c(self) -> _Params:
        return _Params(self.Nx, self.Ny, self.nu, self.hyper_flag, self.hyper_pow, self.alpha,
                       self.drag_flag, self.drag_pow, FILTERS[self.filter], FORCINGS[self.forcing],
                       self.force_k, self.force_scale, INTEGRATORS[self.integrator], self.quirks,
                       self.check_every, self.device)


class Solver:
    """One solver instance on one GPU (owns device memory and a HIP stream)."""

    def __init__(self, cfg: Config, _handle=None):
        lib = load_library()
        self.cfg = cfg
        self.Nx, self.Ny, self.Nyf = cfg.Nx, cfg.Ny, cfg.Ny // 2 + 1
        self._owned = _handle is None
        if _handle is None:
            p = cfg._c()
            h = ctypes.c_void_p()
            _check(lib.ns2d_create(ctypes.byref(p), ctypes.byref(h)))
            _handle = h
        self._h = _handle
        self.n_spect = self.info()["n_spect"]

    def close(self) -> None:
        if getattr(self, "_h", None):
            if self._owned:
                _lib.ns2d_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- state -------------------------------------------------------------
    def set_state(self, w_hat: np.ndarray) -> None:
        w = np.ascontiguousarray(w_hat, dtype=np.complex128)
        assert w.shape == (self.Nx, self.Nyf), w.shape
        _check(_lib.ns2d_set_state(self._h, _dptr(w.view(np.float64))))

    def get_state(self) -> np.ndarray:
        out = np.empty((self.Nx, self.Nyf), dtype=np.complex128)
        _check(_lib.ns2d_get_state(self._h, _dptr(out.view(np.float64))))
        return out

    def set_state_device(self, ptr: int) -> None:
        _check(_lib.ns2d_set_state_device(self._h, ctypes.c_void_p(ptr)))

    def get_state_device(self, ptr: int) -> None:
        _check(_lib.ns2d_get_state_device(self._h, ctypes.c_void_p(ptr)))

    # -- the hot path --------------------------------------------------------
    def spectral_step(self, dt: float, nsteps: int = 1) -> None:
        """RK4Step / AB4Step x nsteps (+ SystemCheck when check_every)."""
        _check(_lib.ns2d_step(self._h, float(dt), int(nsteps)))

    def nonlinear_rhs(self, w_hat: np.ndarray) -> np.ndarray:
        """NonlinearRHSBatch(w_hat) incl. its ForceConjugacy write into the state."""
        w = np.ascontiguousarray(w_hat, dtype=np.complex128)
        out = np.empty_like(w)
        _check(_lib.ns2d_rhs(self._h, _dptr(w.view(np.float64)), _dptr(out.view(np.float64))))
        return out

    def system_measures(self, with_flux: bool = False) -> dict:
        m = _Measures()
        ek = np.zeros(self.n_spect)
        zk = np.zeros(self.n_spect)
        fl = np.zeros(6 * self.n_spect)
        _check(_lib.ns2d_measures(self._h, int(with_flux), ctypes.byref(m), _dptr(ek), _dptr(zk), _dptr(fl)))
        out = {n: getattr(m, n) for n, _ in _Measures._fields_}
        out["enrg_spect"] = ek
        out["enst_spect"] = zk
        if with_flux:
            ns = self.n_spect
            for j, n in enumerate(("d_enst_dt_spect", "enst_diss_spect", "enst_flux_spect",
                                   "d_enrg_dt_spect", "enrg_diss_spect", "enrg_flux_spect")):
                out[n] = fl[j * ns:(j + 1) * ns].copy()
        return out

    def max_vort(self) -> float:
        v = ctypes.c_double()
        _check(_lib.ns2d_max_vort(self._h, ctypes.byref(v)))
        return v.value

    def cfl_rate(self) -> float:
        """GetTimestep's convective rate max pi (|u|/dx + |v|/dy) of the
        state (solver.c:2381-2441), reduced on the device."""
        v = ctypes.c_double()
        _check(_lib.ns2d_cfl_rate(self._h, ctypes.byref(v)))
        return v.value

    def real_vorticity(self) -> np.ndarray:
        out = np.empty((self.Nx, self.Ny))
        _check(_lib.ns2d_real_vort(self._h, _dptr(out)))
        return out

    def r2c(self, w: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(w, dtype=np.float64)
        assert x.shape == (self.Nx, self.Ny)
        out = np.empty((self.Nx, self.Nyf), dtype=np.complex128)
        _check(_lib.ns2d_r2c(self._h, _dptr(x), _dptr(out.view(np.float64))))
        return out

    def c2r(self, w_hat: np.ndarray) -> np.ndarray:
        """Unnormalised inverse transform (FFTW c2r conventions); the state is untouched."""
        x = np.ascontiguousarray(w_hat, dtype=np.complex128)
        assert x.shape == (self.Nx, self.Nyf)
        out = np.empty((self.Nx, self.Ny), dtype=np.float64)
        _check(_lib.ns2d_c2r(self._h, _dptr(x.view(np.float64)), _dptr(out)))
        return out

    def info(self) -> dict:
        i = _Info()
        _check(_lib.ns2d_get_info(self._h, ctypes.byref(i)))
        return {n: getattr(i, n) for n, _ in _Info._fields_}

    def sync(self) -> None:
        _check(_lib.ns2d_sync(self._h))

    def stream(self) -> int:
        s = ctypes.c_void_p()
        _check(_lib.ns2d_get_stream(self._h, ctypes.byref(s)))
        return s.value or 0

    def vjp(self, w0: np.ndarray, dt: float, nsteps: int, lam_T: np.ndarray, ckpt_every: int = 0) -> np.ndarray:
        """dJ/dw_0 of nsteps RK4CN steps from w_0 given dJ/dw_n (packed dJ/dRe + i dJ/dIm)."""
        w = np.ascontiguousarray(w0, dtype=np.complex128)
        lt = np.ascontiguousarray(lam_T, dtype=np.complex128)
        out = np.empty_like(w)
        _check(_lib.ns2d_vjp(self._h, _dptr(w.view(np.float64)), float(dt), int(nsteps),
                             _dptr(lt.view(np.float64)), _dptr(out.view(np.float64)), int(ckpt_every)))
        return out

    def vjp_device(self, w0_ptr: int, dt: float, nsteps: int, lam_T_ptr: int, grad_ptr: int,
                   ckpt_every: int = 0) -> None:
        _check(_lib.ns2d_vjp_device(self._h, ctypes.c_void_p(w0_ptr), float(dt), int(nsteps),
                                    ctypes.c_void_p(lam_T_ptr), ctypes.c_void_p(grad_ptr), int(ckpt_every)))

    def set_kpass_variant(self, idx: int) -> None:
        _check(_lib.ns2d_set_variant(self._h, PASS_COLUMN, int(idx)))

    def set_rpass_variant(self, idx: int) -> None:
        _check(_lib.ns2d_set_variant(self._h, PASS_ROW, int(idx)))

    def profile_steps(self, dt: float, nsteps: int):
        k, r, t = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _check(_lib.ns2d_profile_steps(self._h, float(dt), int(nsteps), ctypes.byref(k), ctypes.byref(r),
                                       ctypes.byref(t)))
        return k.value, r.value, t.value


class Ensemble:
    """Many independent RK4CN solvers (any mix of grid sizes) stepped together:
    members of one size share batched kernel launches and a two-step period of
    all members replays as one hipGraph (ns2d_ens_* in the C ABI)."""

    def __init__(self, cfgs):
        lib = load_library()
        self.cfgs = list(cfgs)
        arr = (_Params * len(self.cfgs))(*[c._c() for c in self.cfgs])
        h = ctypes.c_void_p()
        _check(lib.ns2d_ens_create(arr, len(self.cfgs), ctypes.byref(h)))
        self._h = h
        self.members = []
        for i, c in enumerate(self.cfgs):
            mh = ctypes.c_void_p()
            _check(lib.ns2d_ens_member(self._h, i, ctypes.byref(mh)))
            self.members.append(Solver(c, _handle=mh))

    def __len__(self):
        return len(self.members)

    def __getitem__(self, i) -> Solver:
        return self.members[i]

    def step(self, dts, nsteps: int) -> None:
        d = np.ascontiguousarray(np.broadcast_to(np.asarray(dts, dtype=np.float64), (len(self.members),)))
        _check(_lib.ns2d_ens_step(self._h, _dptr(d), int(nsteps)))

    def autotune(self, reps: int = 3) -> None:
        _check(_lib.ns2d_ens_autotune(self._h, int(reps)))

    def variants(self, member: int) -> tuple:
        """(column-pass, row-pass) shape indices the member's size class uses."""
        k, r = ctypes.c_int(), ctypes.c_int()
        _check(_lib.ns2d_ens_variants(self._h, int(member), ctypes.byref(k), ctypes.byref(r)))
        return k.value, r.value

    def sync(self) -> None:
        _check(_lib.ns2d_ens_sync(self._h))

    def close(self) -> None:
        if getattr(self, "_h", None):
            for m in self.members:
                m._h = None
            _lib.ns2d_ens_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
