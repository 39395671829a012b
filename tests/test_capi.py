"""CPU-side checks of the C ABI and the host package (no GPU calls)."""
import ctypes
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "ns2d_gpu.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(ns2d_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    import ns2d
    lib = ns2d.load_library()
    declared = header_functions()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(ns2d.EXPORTS) == declared


def test_variant_tables_cover_every_size():
    import ns2d
    for n in (8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096):
        vs = ns2d.kpass_variants(n)
        assert vs, n
        for v in vs:
            assert v["threads"] % v["cols"] == 0 and 4 % v["splits"] == 0
            assert v["per_thread"] * v["threads"] == n * v["cols"]
        # every length keeps an unpacked shape (the packed pair is an option,
        # at 4096 among others), the prefetching epilogue only on unpacked
        # shapes, defaults unpacked (measured)
        assert any(not v["packed_pair"] for v in vs), n
        assert not any(v["packed_pair"] and v["prefetch"] for v in vs), n
        assert vs[0]["packed_pair"] == 0, n
        if n == 4096:
            assert any(v["packed_pair"] for v in vs)
            assert vs[0]["threads"] == 1024 and vs[0]["cols"] == 2 and vs[0]["pad_shift"] == 4
        assert all(v["min_waves"] >= 1 and v["pad_shift"] >= 2 for v in vs), n
        # defaults: index 0 of each length for both integrators
        for integ in ("RK4CN", "AB4"):
            assert ns2d.default_variant(ns2d.PASS_ROW, n, integ) == 0
            assert ns2d.default_variant(ns2d.PASS_COLUMN, n, integ) == 0
        rs = ns2d.rpass_variants(n)
        assert rs, n
        for v in rs:
            assert v["per_thread"] * v["threads_per_pair"] == n
            assert v["threads"] == v["threads_per_pair"] * v["pairs"]
            assert v["threads_per_pair"] > 64 and v["pairs"] == 1 or 64 % v["threads_per_pair"] == 0
            assert not (v["single_row"] and v["pairs"] != 1), n
        assert rs[0]["single_row"] == (n >= 256), n
        # the half-length forward transform is the default at 512 only, the
        # preloaded twiddle powers at 256 only
        assert rs[0]["half_forward"] == (n == 512), n
        assert rs[0]["twiddle_powers"] == (n == 256), n
        assert all(v["single_row"] and v["threads_per_pair"] <= 64 for v in rs if v["twiddle_powers"]), n
        assert all(v["single_row"] and v["per_thread"] >= 4 for v in rs if v["half_forward"]), n
        if n >= 256:
            assert any(v["single_row"] for v in rs), n


def test_ics_match_oracle_generator():
    """The product IC generator (ns2d.ics) and the oracle's produce identical w_hat."""
    import ns2d_oracle as O
    from ns2d import ics
    for nx, ny, f in ((64, 64, "23"), (128, 32, "houli_ref"), (32, 64, "houli")):
        a = O.ic_random_spectrum(O.Params(nx, ny, filter=f), seed=5, k0=6.0)
        b = ics.random_phase(nx, ny, seed=5, k0=6.0, filt=f)
        assert np.array_equal(a, b)


def test_invalid_params_rejected_without_gpu():
    """Argument validation happens before any device call."""
    import ctypes

    import ns2d
    lib = ns2d.load_library()
    p = ns2d.Config(Nx=100, Ny=64)._c()
    h = ctypes.c_void_p()
    rc = lib.ns2d_create(ctypes.byref(p), ctypes.byref(h))
    assert rc == ns2d.NS2D_ERR_INVALID
    assert b"powers of two" in lib.ns2d_last_error()


def test_layout_query_and_traffic_key():
    """ns2d_layout reports the compiled device layout without a GPU (k-space
    column-major, 2-row x-tiles for both mixed-space sets by default), rejects
    null pointers, and bench.py keys its traffic table by kernel signature
    and this layout."""
    import ns2d
    lib = ns2d.load_library()
    assert ns2d.layout() == (0, 2, 2)
    k = ctypes.c_int(-1)
    assert lib.ns2d_layout(ctypes.byref(k), None, None) == ns2d.NS2D_ERR_INVALID
    import bench
    key = bench.traffic_key("kpass", 1024)
    assert key.startswith("kpass_kernel<1024, ") and key.endswith(" layout k0g2n2")
    v = ns2d.kpass_variants(1024)[ns2d.default_variant(ns2d.PASS_COLUMN, 1024)]
    assert f"<1024, {v['cols']}, {v['threads']}, {v['splits']}, {v['pad_shift']}, {v['line_offset']}, " in key


def test_default_variant_rejects_bad_arguments():
    """ns2d_default_variant validates pass, length, integrator and its output pointer."""
    import ctypes

    import ns2d
    lib = ns2d.load_library()
    idx = ctypes.c_int(-7)
    assert lib.ns2d_default_variant(ns2d.PASS_COLUMN, 100, 0, ctypes.byref(idx)) == ns2d.NS2D_ERR_INVALID
    assert lib.ns2d_default_variant(ns2d.PASS_ROW, 8192, 0, ctypes.byref(idx)) == ns2d.NS2D_ERR_INVALID
    assert lib.ns2d_default_variant(2, 256, 0, ctypes.byref(idx)) == ns2d.NS2D_ERR_INVALID
    assert lib.ns2d_default_variant(ns2d.PASS_COLUMN, 256, 5, ctypes.byref(idx)) == ns2d.NS2D_ERR_INVALID
    assert lib.ns2d_default_variant(ns2d.PASS_COLUMN, 256, 0, None) == ns2d.NS2D_ERR_INVALID
    assert idx.value == -7
    assert lib.ns2d_default_variant(ns2d.PASS_COLUMN, 4096, 1, ctypes.byref(idx)) == ns2d.NS2D_OK
    assert idx.value == ns2d.default_variant(ns2d.PASS_COLUMN, 4096, "AB4") == 0
