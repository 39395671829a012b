"""ctypes binding of libns2d_host.so: the driver's command line, output
naming, reference initial conditions and HDF5 writer (plain C, no GPU).

The initial conditions that start in real space need the caller's
transforms; ``initial_condition(..., solver=s)`` uses the GPU solver's
``r2c``/``c2r`` (the driver does the same through the C ABI), and explicit
``r2c=``/``c2r=`` callables may be given instead.
"""
from __future__ import annotations

import ctypes
import os
from typing import Callable, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libns2d_host.so")
SOLVER_EXE = os.path.join(os.path.dirname(_HERE), "ns2d_solver")


class RunCfg(ctypes.Structure):
    """ns2d_run_cfg of driver/ns2d_host.h (GetCMLArgs fields, utils.c:34-461)."""
    _fields_ = [
        ("input_file_name", ctypes.c_char * 512),
        ("input_dir", ctypes.c_char * 512),
        ("output_dir", ctypes.c_char * 512),
        ("output_tag", ctypes.c_char * 64),
        ("file_only", ctypes.c_int),
        ("N", ctypes.c_long * 2),
        ("t0", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("T", ctypes.c_double),
        ("cfl_const", ctypes.c_double),
        ("adapt_step_flag", ctypes.c_int),
        ("cfl_cond_flag", ctypes.c_int),
        ("u0", ctypes.c_char * 64),
        ("forcing", ctypes.c_char * 64),
        ("force_k", ctypes.c_int),
        ("force_scale_var", ctypes.c_double),
        ("trans_iters_flag", ctypes.c_int),
        ("trans_iters_frac", ctypes.c_double),
        ("nu", ctypes.c_double),
        ("hyper_visc_flag", ctypes.c_int),
        ("hyper_visc_pow", ctypes.c_double),
        ("ekmn_alpha", ctypes.c_double),
        ("ekmn_drag_flag", ctypes.c_int),
        ("ekmn_drag_pow", ctypes.c_double),
        ("save_every", ctypes.c_int),
        ("integrator", ctypes.c_char * 16),
        ("filter", ctypes.c_char * 16),
        ("quirks", ctypes.c_uint),
        ("testing", ctypes.c_int),
        ("ic_ranks", ctypes.c_int),
    ]

    def as_dict(self) -> dict:
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            if isinstance(v, bytes):
                v = v.decode()
            elif name == "N":
                v = (v[0], v[1])
            out[name] = v
        return out


class OutPaths(ctypes.Structure):
    _fields_ = [("dir", ctypes.c_char * 1200), ("main_file", ctypes.c_char * 1240),
                ("spectra_file", ctypes.c_char * 1240), ("sys_msr_file", ctypes.c_char * 1240)]


_XFORM = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                          ctypes.POINTER(ctypes.c_double))


class Transforms(ctypes.Structure):
    _fields_ = [("ctx", ctypes.c_void_p), ("r2c", _XFORM), ("c2r", _XFORM)]


_lib = None


def load_library(path: Optional[str] = None):
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"{p} is missing: build it with build_lib.py (or __graft_entry__.build())")
    lib = ctypes.CDLL(p)
    c, vp, sz = ctypes.c_char_p, ctypes.c_void_p, ctypes.c_size_t
    lib.ns2d_cli_defaults.argtypes = [ctypes.POINTER(RunCfg)]
    lib.ns2d_cli_defaults.restype = None
    lib.ns2d_cli_parse.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(RunCfg), c, sz]
    lib.ns2d_cli_parse.restype = ctypes.c_int
    lib.ns2d_output_paths.argtypes = [ctypes.POINTER(RunCfg), ctypes.c_int, ctypes.POINTER(OutPaths), c, sz]
    lib.ns2d_output_paths.restype = ctypes.c_int
    lib.ns2d_initial_condition.argtypes = [ctypes.POINTER(RunCfg), vp, ctypes.POINTER(Transforms), c, sz]
    lib.ns2d_initial_condition.restype = ctypes.c_int
    lib.ns2d_host_dealias.argtypes = [ctypes.POINTER(RunCfg), vp, ctypes.c_int]
    lib.ns2d_host_dealias.restype = None
    lib.ns2d_host_force_conjugacy.argtypes = [ctypes.POINTER(RunCfg), vp]
    lib.ns2d_host_force_conjugacy.restype = None
    lib.ns2d_mean_flows.argtypes = [ctypes.POINTER(RunCfg), vp, vp, vp]
    lib.ns2d_mean_flows.restype = None
    _lib = lib
    return lib


class HostError(RuntimeError):
    pass


def defaults() -> RunCfg:
    c = RunCfg()
    load_library().ns2d_cli_defaults(ctypes.byref(c))
    return c


def parse_cli(args: Sequence[str], prog: str = "ns2d_solver") -> RunCfg:
    """GetCMLArgs on an argument list (without the program name)."""
    c = defaults()
    argv = [prog.encode()] + [a.encode() for a in args]
    arr = (ctypes.c_char_p * (len(argv) + 1))(*argv, None)
    msg = ctypes.create_string_buffer(4096)
    if load_library().ns2d_cli_parse(len(argv), arr, ctypes.byref(c), msg, len(msg)):
        raise HostError(msg.value.decode())
    return c


def output_paths(c: RunCfg, mkdirs: bool = False) -> dict:
    p = OutPaths()
    msg = ctypes.create_string_buffer(4096)
    if load_library().ns2d_output_paths(ctypes.byref(c), int(mkdirs), ctypes.byref(p), msg, len(msg)):
        raise HostError(msg.value.decode())
    return {k: getattr(p, k).decode() for k, _ in OutPaths._fields_}


def initial_condition(c: RunCfg, solver=None, r2c: Optional[Callable] = None,
                      c2r: Optional[Callable] = None) -> np.ndarray:
    """The reference's initial condition c.u0 (after dealiasing and the
    conjugacy fix-up) as complex128 [Nx][Ny/2+1] in FFTW units."""
    Nx, Ny = int(c.N[0]), int(c.N[1])
    Nyf = Ny // 2 + 1
    if solver is not None:
        r2c = r2c or solver.r2c
        c2r = c2r or solver.c2r
    err = []

    def _r2c(_ctx, src, dst):
        try:
            x = np.ctypeslib.as_array(src, shape=(Nx * Ny,)).reshape(Nx, Ny).copy()
            out = np.ctypeslib.as_array(dst, shape=(2 * Nx * Nyf,)).view(np.complex128).reshape(Nx, Nyf)
            out[:] = r2c(x)
            return 0
        except Exception as e:  # reported after the C call returns
            err.append(e)
            return 1

    def _c2r(_ctx, src, dst):
        try:
            x = np.ctypeslib.as_array(src, shape=(2 * Nx * Nyf,)).view(np.complex128).reshape(Nx, Nyf).copy()
            out = np.ctypeslib.as_array(dst, shape=(Nx * Ny,)).reshape(Nx, Ny)
            out[:] = c2r(x)
            return 0
        except Exception as e:
            err.append(e)
            return 1

    def _missing(_ctx, _src, _dst):
        err.append(HostError(f"initial condition {c.u0.decode()} needs transforms: pass solver= or r2c=/c2r="))
        return 1

    tf = Transforms(None, _XFORM(_r2c if r2c else _missing), _XFORM(_c2r if c2r else _missing))
    w = np.zeros((Nx, Nyf), np.complex128)
    msg = ctypes.create_string_buffer(4096)
    rc = load_library().ns2d_initial_condition(ctypes.byref(c), w.ctypes.data, ctypes.byref(tf), msg, len(msg))
    if err:
        raise err[0]
    if rc:
        raise HostError(msg.value.decode())
    return w


def dealias(c: RunCfg, f_hat: np.ndarray) -> np.ndarray:
    f = np.ascontiguousarray(f_hat, dtype=np.complex128).copy()
    load_library().ns2d_host_dealias(ctypes.byref(c), f.ctypes.data, 1)
    return f


def mean_flows(c: RunCfg, w_hat: np.ndarray):
    Nx, Nyf = int(c.N[0]), int(c.N[1]) // 2 + 1
    w = np.ascontiguousarray(w_hat, dtype=np.complex128)
    my, mx = np.zeros(Nx), np.zeros(Nyf)
    load_library().ns2d_mean_flows(ctypes.byref(c), w.ctypes.data, my.ctypes.data, mx.ctypes.data)
    return my, mx


# ------------------------------------------------------------------ HDF5
_dp = ctypes.POINTER(ctypes.c_double)


class _Save(ctypes.Structure):
    _fields_ = [("t", ctypes.c_double), ("dt", ctypes.c_double), ("index", ctypes.c_int),
                ("w", _dp), ("w_hat", _dp), ("tg_soln", _dp),
                ("enst_spect", _dp), ("enrg_spect", _dp), ("flux_spect", _dp)]


class _Series(ctypes.Structure):
    _fields_ = [("n", ctypes.c_long)] + [(k, _dp) for k in (
        "time", "tot_energy", "tot_enstr", "tot_palin", "tot_forc", "tot_div", "enrg_diss", "enst_diss",
        "d_enst_dt_sbst", "enst_flux_sbst", "enst_diss_sbst", "d_enrg_dt_sbst", "enrg_flux_sbst",
        "enrg_diss_sbst", "mean_flow_y", "mean_flow_x")]


SERIES_KEYS = [k for k, _ in _Series._fields_[1:]]


def _ptr(a: Optional[np.ndarray], dtype) -> tuple:
    if a is None:
        return None, None
    arr = np.ascontiguousarray(a, dtype=dtype)
    return arr, arr.ctypes.data_as(_dp)


class H5Writer:
    """The reference's HDF5 output (hdf5_funcs.c) for a run configuration:
    Main/Spectra files with one /Iter_%05d group per save and the system
    measures file written by final()."""

    def __init__(self, c: RunCfg, paths: dict, n_spect: int):
        lib = load_library()
        lib.ns2d_h5_create.argtypes = [ctypes.POINTER(RunCfg), ctypes.POINTER(OutPaths), ctypes.c_int,
                                       ctypes.c_char_p, ctypes.c_size_t]
        lib.ns2d_h5_create.restype = ctypes.c_void_p
        lib.ns2d_h5_write_save.argtypes = [ctypes.c_void_p, ctypes.POINTER(_Save), ctypes.c_char_p, ctypes.c_size_t]
        lib.ns2d_h5_final.argtypes = [ctypes.c_void_p, ctypes.POINTER(_Series), ctypes.c_char_p, ctypes.c_size_t]
        lib.ns2d_h5_close.argtypes = [ctypes.c_void_p]
        lib.ns2d_h5_close.restype = None
        self._lib, self._c = lib, c
        p = OutPaths()
        for k, _ in OutPaths._fields_:
            setattr(p, k, paths[k].encode())
        msg = ctypes.create_string_buffer(4096)
        self._f = lib.ns2d_h5_create(ctypes.byref(c), ctypes.byref(p), int(n_spect), msg, len(msg))
        if not self._f:
            raise HostError(msg.value.decode())

    def save(self, index: int, t: float, dt: float, w=None, w_hat=None, tg_soln=None, enst_spect=None,
             enrg_spect=None, flux_spect=None) -> None:
        keep = []
        s = _Save(t=t, dt=dt, index=index)
        for name, a, dt_ in (("w", w, np.float64), ("tg_soln", tg_soln, np.float64),
                             ("enst_spect", enst_spect, np.float64), ("enrg_spect", enrg_spect, np.float64),
                             ("flux_spect", flux_spect, np.float64)):
            arr, ptr = _ptr(a, dt_)
            keep.append(arr)
            setattr(s, name, ptr)
        if w_hat is not None:
            arr = np.ascontiguousarray(w_hat, dtype=np.complex128)
            keep.append(arr)
            s.w_hat = arr.ctypes.data_as(_dp)
        msg = ctypes.create_string_buffer(4096)
        if self._lib.ns2d_h5_write_save(self._f, ctypes.byref(s), msg, len(msg)):
            raise HostError(msg.value.decode())

    def final(self, series: dict) -> None:
        keep = []
        s = _Series(n=len(series["time"]))
        for k in SERIES_KEYS:
            arr, ptr = _ptr(series[k], np.float64)
            keep.append(arr)
            setattr(s, k, ptr)
        msg = ctypes.create_string_buffer(4096)
        if self._lib.ns2d_h5_final(self._f, ctypes.byref(s), msg, len(msg)):
            raise HostError(msg.value.decode())

    def close(self) -> None:
        if self._f:
            self._lib.ns2d_h5_close(self._f)
            self._f = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
