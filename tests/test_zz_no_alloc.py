"""ns2d_measures_noalloc (the allocation-free variant of ns2d_measures, for
callers that cannot allocate device memory) against ns2d_measures, through the
C ABI: two calls in a row, the second on a smaller grid, with and without the
flux terms.  Both entry points run the same measures kernel and sum the bins in
the same order, so the results agree bit for bit."""
import ctypes

import numpy as np
import pytest

import ns2d_oracle as O

pytestmark = pytest.mark.gpu


def _measure(s, fn, with_flux):
    import ns2d
    m = ns2d._Measures()
    ek = np.full(s.n_spect, np.nan)
    zk = np.full(s.n_spect, np.nan)
    fl = np.full(6 * s.n_spect, np.nan)
    rc = fn(s._h, int(with_flux), ctypes.byref(m), ns2d._dptr(ek), ns2d._dptr(zk), ns2d._dptr(fl))
    assert rc == 0, ns2d._lib.ns2d_last_error().decode()
    tot = np.array([getattr(m, n) for n, _ in ns2d._Measures._fields_])
    return tot, ek, zk, (fl if with_flux else None)


@pytest.mark.parametrize("with_flux", [0, 1])
def test_noalloc_measures_equal_measures(with_flux):
    import ns2d
    lib = ns2d.load_library()
    # the second call runs on a smaller grid (fewer spectrum bins) than the first
    for n, forcing in ((256, "KOLM"), (64, "NONE")):
        print(f"noalloc measures {n}x{n} forcing={forcing} with_flux={with_flux}", flush=True)
        p = O.Params(n, n, nu=1e-3, filter="23")
        w = O.ic_random_spectrum(p, seed=n + 11, k0=4.0)
        cfg = ns2d.Config(Nx=n, Ny=n, nu=1e-3, hyper_flag=0, hyper_pow=1.0, alpha=0.1, drag_flag=1, drag_pow=0.0,
                          filter="23", forcing=forcing, force_k=4, force_scale=1.0)
        with ns2d.Solver(cfg) as s:
            s.set_state(w)
            s.spectral_step(1e-3, 3)
            w3 = s.get_state()
            ref = _measure(s, lib.ns2d_measures, with_flux)
            s.set_state(w3)   # the flux RHS applies the ForceConjugacy write to the state
            got = _measure(s, lib.ns2d_measures_noalloc, with_flux)
        for a, b in zip(got, ref):
            if b is None:
                continue
            assert np.array_equal(a, b), (n, with_flux, np.max(np.abs(a - b)))
        assert np.all(np.isfinite(got[1])) and got[0][0] > 0.0
