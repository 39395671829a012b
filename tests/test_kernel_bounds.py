"""CPU audit of the kernels' index arithmetic (no GPU calls).

For every compiled pass shape (the library's own shape tables) and every grid
the C ABI accepts (Nx, Ny powers of two in [8, 4096]), this replays, with
numpy, the index expressions of ns2d_kernels.hpp -- thread -> (slot, line,
rows), work item -> column group, the register FFT's LDS exchange positions,
the packed pair's held columns, the ForceConjugacy hand-off, the batched
member tables with the XCD-ordered work-item map, the adjoint passes and the
measures bins -- and asserts that every global and LDS access stays inside
the buffer the host allocates for it (ns2d_capi.cpp: [Nx][ld] spectral
arrays, ld = roundup(Ny/2+1, 8); an 8*Nx packed-pair buffer; Nx-long
hand-off columns; NQ*n_spect measures partials; the 8192-entry twiddle
table; LDS arrays as declared).  It also checks that each FFT exchange writes
every line position exactly once and that the XCD map is a permutation."""
import math

import numpy as np
import pytest

TW_LEN = 8192
LENGTHS = [8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096]
NQ = 14
LDS_BYTES = 160 * 1024


def padx_s(j, ps=3):
    return j + (j >> ps)


def line_stride(n):
    return padx_s(n) + 1


def fft_reg_exchanges(N, TL, PS=3):
    """LDS positions written/read by fft_reg<N, TL, ., ., PS> (ns2d_fft_reg.hpp)
    for lanes l = 0..TL-1; checks twiddle indices and that each exchange is a
    permutation of the line.  Returns the largest LDS position used."""
    E = N // TL
    l = np.arange(TL)
    P, top = 1, 0
    while N // P > 1:
        REM = N // P
        R = min(min(E, 16), REM)
        NB = E // R
        last = P * R == N
        tws = TW_LEN // (P * R)
        written = []
        for m in range(NB):
            i = l + TL * m
            k = i & (P - 1)
            assert np.all(k * tws < TW_LEN)
            if not last:
                j0 = (i - k) * R + k
                for r in range(R):
                    written.append(j0 + r * P)
        if not last:
            w = np.concatenate(written)
            assert np.array_equal(np.sort(w), np.arange(N)), (N, TL, P)
            top = max(top, int(padx_s(w.max(), PS)), int(padx_s((l + TL * (E - 1)).max(), PS)))
        P *= R
    return top


def xcd_remap(b, n):
    NX = 8
    if n < NX:
        return b
    q, r = n // NX, n % NX
    x, j = b % NX, b // NX
    return np.where(x < r, x * (q + 1), r * (q + 1) + (x - r) * q) + j


def test_xcd_remap_is_a_permutation():
    for n in list(range(1, 70)) + [257, 514, 1024, 1025, 2049, 4100, 12345]:
        b = np.arange(n)
        assert np.array_equal(np.sort(xcd_remap(b, n)), b), n


def kshapes(n):
    import ns2d
    return ns2d.kpass_variants(n)


def rshapes(n):
    import ns2d
    return ns2d.rpass_variants(n)


def ld_of(ny):
    nyf = ny // 2 + 1
    return nyf, (nyf + 7) // 8 * 8


@pytest.mark.parametrize("L", LENGTHS)
def test_column_pass_indices(L):
    """kpass_body / kpass_kernel / kpass_batched for every shape of length L
    on every Ny."""
    for v in kshapes(L):
        C, T, KS, PS, OFF = v["cols"], v["threads"], v["splits"], v["pad_shift"], v["line_offset"]
        PK = 2 if v["pair_in_registers"] else (1 if v["packed_pair"] else 0)
        TL = T // C
        E = L // TL
        assert T % C == 0 and L % TL == 0 and E >= 2 and 4 % KS == 0
        LS = padx_s(L, PS) + OFF
        assert C * LS * 16 <= LDS_BYTES, v
        tid = np.arange(T)
        c, l = tid % C, tid // C
        # every LDS position a lane touches lies inside its slot's line
        top = fft_reg_exchanges(L, TL, PS)
        assert top < LS and c.max() * LS + top < C * LS
        rows = (l[:, None] + TL * np.arange(E)[None, :])          # x = l + TL*e
        assert rows.min() == 0 and rows.max() == L - 1
        assert np.array_equal(np.sort(rows[c == 0].ravel()), np.arange(L))
        # the packed pair's mirror reads buf[padx_s((L - x) & (L - 1))]
        assert padx_s(((L - rows) & (L - 1)).max(), PS) < LS
        for ny in LENGTHS:
            nyf, ld = ld_of(ny)
            H = nyf - 1
            nslots = H if PK else nyf
            groups = (nslots + C - 1) // C
            grid = groups * KS
            wi = np.arange(grid)
            g, split = wi // KS, wi % KS
            ky = g[:, None] * C + np.arange(C)[None, :]           # [work item][slot]
            colv = ky < nslots
            assert np.all(split * (4 // KS) + (4 // KS) - 1 < 4)
            # global element index x*ld + ky (row-major [Nx][ld]) of valid slots
            assert ky[colv].max() < ld and (not PK or H < ld)
            # every column (PK: every column but ky = 0 and H, which slot 0 of
            # work group 0 carries) is owned by exactly one slot per split
            own = np.sort(ky[split == 0][colv[split == 0]])
            assert np.array_equal(own, np.arange(nslots)), (v, ny)
            M = L * ld
            assert (L - 1) * ld + ky[colv].max() < M
            if PK == 1:
                # hold = pairbuf + split*2*L; accesses hold[L + x], hold[q*L + x]
                assert (KS - 1) * 2 * L + 2 * L <= 8 * L
            # ForceConjugacy hand-off: conj_src[-kx] for kx < 0, conj_dst[x]
            kx = np.where(rows <= L // 2, rows, rows - L)
            assert (-kx[kx < 0]).max(initial=0) < L and rows.max() < L
            # single-solver launch with last_split (launch_kpass: unpacked,
            # KS = 1, C > 1, one valid column in the last group): the grid has
            # one more work item; items groups-1 and groups are the last
            # group's halves with inverse fields {0, 1} and {2, 3}, split 0
            # the only writer, every other item unchanged
            if KS == 1 and PK == 0 and C > 1 and nyf % C == 1:
                n = groups + 1
                wi = np.arange(n)
                tail = wi >= groups - 1
                g2 = np.where(tail, groups - 1, wi)
                s2 = np.where(tail, wi - (groups - 1), 0)
                nf = np.where(tail, 2, 4)
                assert g2.max() == groups - 1 and s2.max() == 1
                f_lo, f_hi = s2 * nf, s2 * nf + nf - 1
                assert f_lo.min() == 0 and f_hi.max() == 3
                last = wi[tail]
                assert sorted((s2[last] * 2).tolist() + (s2[last] * 2 + 1).tolist()) == [0, 1, 2, 3]
                ky2 = g2[:, None] * C + np.arange(C)[None, :]
                own = np.sort(ky2[s2 == 0][(ky2 < nslots)[s2 == 0]])
                assert np.array_equal(own, np.arange(nslots)), (v, ny)
                b = xcd_remap(np.arange(n), n)
                assert np.array_equal(np.sort(b), np.arange(n))
            # batched launch: b -> xcd_remap -> member m = b // per
            for members in (1, 3, 24):
                n = grid * members
                b = xcd_remap(np.arange(n), n)
                assert np.array_equal(np.sort(b), np.arange(n))
                assert (b // grid).max() == members - 1


@pytest.mark.parametrize("NY", LENGTHS)
def test_row_pass_indices(NY):
    """rpass_body / rpass_kernel / rpass_batched for every shape of length NY
    on every Nx."""
    for v in rshapes(NY):
        TLR, PAIRS, T = v["threads_per_pair"], v["pairs"], v["threads"]
        SR = bool(v["single_row"])
        H, E = NY // 2, NY // TLR
        BLOCK = TLR > 64
        assert E >= 2 and (PAIRS == 1 if BLOCK else 64 % TLR == 0) and T == TLR * PAIRS
        LS = line_stride(NY)
        assert PAIRS * LS * 16 + ((T + 63) // 64) * 8 <= LDS_BYTES
        top = fft_reg_exchanges(NY, TLR)
        assert top < LS
        l = np.arange(TLR)
        ky = l[:, None] + TLR * np.arange(E)[None, :]
        kk = np.where(ky > H, NY - ky, ky)
        assert kk.min() >= 0 and kk.max() <= H
        # split of the forward transform: buf[padx((NY - k) & (NY - 1))]
        assert padx_s(((NY - ky) & (NY - 1)).max()) < LS
        nyf, ld = ld_of(NY)
        assert H < ld
        if v["twiddle_powers"]:
            # TwFull: w^r = tw[k*r*NS2D_TW_LEN/(P*R)] for k < P, r < R, every pass
            assert SR and TLR <= 64 and PAIRS == 1
            for n in ((NY, NY // 2) if v["half_forward"] else (NY,)):
                En, P = n // TLR, 1
                while P < n:
                    R = min(16, En, n // P)
                    k = (np.arange(TLR)[:, None] + TLR * np.arange(En // R)[None, :]) & (P - 1)
                    top = (k.max() * (R - 1)) * (8192 // (P * R))
                    assert top < 8192, (v, n, P)
                    P *= R
        if v["half_forward"]:
            # rpass_forward_half: lanes l and l ^ 1 swap values (one wave);
            # lane label lh, half line of M points, outputs k = lh + TLR*h
            assert SR and E >= 4 and TLR % 2 == 0
            M, EH = NY // 2, E // 2
            assert np.all((l ^ 1) // 64 == l // 64) and (l ^ 1).max() < TLR
            lh = (l >> 1) + (l & 1) * (TLR // 2)
            assert np.array_equal(np.sort(lh), np.arange(TLR))
            h = np.arange(EH)[None, :]
            m = lh[:, None] + TLR * h
            odd = (l & 1)[:, None] == 1
            y_even = np.where(odd, (l - 1)[:, None] + TLR * (2 * h + 1), l[:, None] + TLR * (2 * h))
            y_odd = np.where(odd, l[:, None] + TLR * (2 * h + 1), (l + 1)[:, None] + TLR * (2 * h))
            assert np.array_equal(y_even, 2 * m) and np.array_equal(y_odd, 2 * m + 1)
            assert np.array_equal(np.sort(m.ravel()), np.arange(M))
            assert fft_reg_exchanges(M, TLR) < LS
            assert padx_s(m.max()) < LS and padx_s(((M - m) & (M - 1)).max()) < LS
            assert (m.max()) * (8192 // NY) < 8192          # exp(-2 pi i k/NY) from the table
            written = np.concatenate([m.ravel(), [M]])      # k = M by the k = 0 lane
            assert np.array_equal(np.sort(written), np.arange(H + 1))
        for nx in LENGTHS:
            units = nx if SR else nx // 2
            grid = (units + PAIRS - 1) // PAIRS
            blk = np.arange(grid)
            q = np.arange(PAIRS)
            x0 = blk[:, None] if SR else 2 * (blk[:, None] * PAIRS + q[None, :])
            valid = x0 < nx
            # with BLOCK every thread of a work group runs the transforms: x0 must be valid
            if BLOCK:
                assert np.all(valid)
            rows_used = x0[valid] if SR else np.concatenate([x0[valid], x0[valid] + 1])
            assert rows_used.max() < nx
            assert np.array_equal(np.sort(rows_used), np.arange(nx))
            M = nx * ld
            assert rows_used.max() * ld + H < M
            # real-space vorticity rows (wreal[x*NY + y])
            assert rows_used.max() * NY + NY - 1 < nx * NY
            for members in (1, 5, 24):
                assert ((np.arange(grid * members)) // grid).max() == members - 1


def _adjoint_tables():
    """The adjoint shape lists of ns2d_adj.hip, parsed from the source."""
    import os
    import re
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = open(os.path.join(root, "2d-navier-stokes_amd", "csrc", "ns2d_adj.hip")).read()
    r = src[src.index("#define NS2D_RADJ_VARIANTS"):src.index("#define NS2D_KADJ_VARIANTS")]
    k = src[src.index("#define NS2D_KADJ_VARIANTS"):src.index("struct RAdjVariant")]
    radj = [tuple(map(int, t)) for t in re.findall(r"X\((\d+), (\d+), (\d+), (\d+)\)", r)]
    kadj = [tuple(map(int, t)) for t in re.findall(r"X\((\d+), (\d+), (\d+)\)", k)]
    return radj, kadj


def test_adjoint_pass_indices():
    radj, kadj = _adjoint_tables()
    assert {t[0] for t in radj} == set(LENGTHS) and {t[0] for t in kadj} == set(LENGTHS)
    for NY, TLR, PAIRS, SR in radj:
        E, H, BLOCK = NY // TLR, NY // 2, TLR > 64
        assert E >= 2 and (PAIRS == 1 if BLOCK else 64 % TLR == 0) and (not SR or PAIRS == 1)
        LS = line_stride(NY)
        assert PAIRS * LS * 16 <= LDS_BYTES
        assert fft_reg_exchanges(NY, TLR) < LS
        _, ld = ld_of(NY)
        for nx in LENGTHS:
            units = nx if SR else nx // 2
            grid = (units + PAIRS - 1) // PAIRS
            x0 = (np.arange(grid)[:, None] if SR else
                  2 * (np.arange(grid)[:, None] * PAIRS + np.arange(PAIRS)[None, :]))
            if BLOCK:
                assert np.all(x0 < nx)
            v = x0[x0 < nx]
            rows = v if SR else np.concatenate([v, v + 1])
            assert rows.max() < nx and np.array_equal(np.sort(rows), np.arange(nx))
            assert rows.max() * ld + H < nx * ld
    for L, C, T in kadj:
        TL, LS = T // C, line_stride(L)
        E = L // TL
        assert T % C == 0 and L % TL == 0 and E >= 2 and C * LS * 16 <= LDS_BYTES
        assert fft_reg_exchanges(L, TL) < LS
        for ny in LENGTHS:
            nyf, ld = ld_of(ny)
            grid = (nyf + C - 1) // C
            ky = np.arange(grid)[:, None] * C + np.arange(C)[None, :]
            assert np.array_equal(np.sort(ky[ky < nyf]), np.arange(nyf))
            assert ky[ky < nyf].max() < ld and (L - 1) * ld + nyf - 1 < L * ld
            b = xcd_remap(np.arange(grid), grid)
            assert np.array_equal(np.sort(b), np.arange(grid))


def isqrt_floor(v):
    """floor(sqrt(v)) for int64 arrays, -1 where v < 0 (isqrt_floor in the kernel)."""
    v = np.asarray(v, dtype=np.int64)
    r = np.floor(np.sqrt(np.maximum(v, 0).astype(np.float64))).astype(np.int64)
    r = np.where(r * r > v, r - 1, r)
    r = np.where((r + 1) * (r + 1) <= v, r + 1, r)
    return np.where(v < 0, -1, r)


@pytest.mark.parametrize("nx,ny", [(8, 8), (16, 64), (64, 16), (128, 128), (256, 32), (1024, 1024),
                                   (4096, 4096), (4096, 8), (8, 4096)])
def test_measures_bins(nx, ny):
    """measures_kernel: bin b sums the annulus b^2 - b < |k|^2 <= b^2 + b over
    kx in (-Nx/2, Nx/2], 0 <= ky <= Ny/2; every mode of the half spectrum
    except (0, 0) lands in exactly one bin b < n_spect, and every index
    i*ld + ky and out[b*NQ + q] is in bounds."""
    H = ny // 2
    _, ld = ld_of(ny)
    nspect = int(round(math.sqrt((nx / 2.0) ** 2 + (ny / 2.0) ** 2) + 1))
    count = np.zeros((nx, H + 2), dtype=np.int64)   # difference array over ky
    for b in range(1, nspect):
        lo, hi = b * b - b, b * b + b
        kxmax = b + 1
        kx = np.arange(-kxmax, kxmax + 1, dtype=np.int64)
        kx = kx[(kx >= -nx // 2 + 1) & (kx <= nx // 2)]
        kx2 = kx * kx
        ylo = np.where(lo - kx2 < 0, 0, isqrt_floor(lo - kx2) + 1)
        yhi = np.minimum(isqrt_floor(hi - kx2), H)
        i = np.where(kx >= 0, kx, kx + nx)
        use = yhi >= ylo
        if np.any(use):
            assert i[use].min() >= 0 and i[use].max() < nx
            assert np.all(i[use] * ld + yhi[use] < nx * ld)
            np.add.at(count, (i[use], ylo[use]), 1)
            np.add.at(count, (i[use], yhi[use] + 1), -1)
        assert b * NQ + NQ - 1 < nspect * NQ
    per_mode = np.cumsum(count, axis=1)[:, :H + 1]
    expect = np.ones_like(per_mode)
    expect[0, 0] = 0
    assert np.array_equal(per_mode, expect)


@pytest.mark.parametrize("nbins", [1, 5, 8, 255, 256, 257, 725, 2897, 5793])
def test_measures_workgroup_to_bin_map(nbins):
    """meas_bin / meas_grid (ns2d_internal.h): the measures grid is rounded up
    to 8 * MEAS_CHUNK workgroups and workgroup g computes bin
    ((j / CH) * 8 + g % 8) * CH + j % CH with j = g / 8 -- one to one on
    [0, grid), so every bin b < n_spect is summed by exactly one workgroup
    (the rest exit), consecutive bins of a chunk share an XCD (g % 8), and the
    XCDs' shares of bins differ by at most one chunk."""
    CH = 32
    grid = -(-nbins // (8 * CH)) * (8 * CH)
    g = np.arange(grid)
    j = g // 8
    b = ((j // CH) * 8 + g % 8) * CH + j % CH
    assert np.array_equal(np.sort(b), np.arange(grid))
    live = b < nbins
    assert np.array_equal(np.sort(b[live]), np.arange(nbins))
    xcd_of_bin = np.empty(grid, dtype=np.int64)
    xcd_of_bin[b] = g % 8
    chunks = xcd_of_bin.reshape(-1, CH)
    assert np.all(chunks == chunks[:, :1])
    share = np.bincount(xcd_of_bin[:nbins], minlength=8)
    assert share.max() - share.min() <= CH


@pytest.mark.parametrize("xt", [1, 2, 4, 8])
def test_spectral_tile_index_is_a_permutation(xt):
    """tile_idx (ns2d_internal.h, the kidx / gidx / nidx layouts): with x-tiles of XT
    rows, element (x, ky) of an [Nx][ld] array sits at ((x/XT)*ld + ky)*XT +
    x%XT -- a permutation of [0, Nx*ld) for every grid the ABI accepts (XT <= 8
    divides every Nx), so every spectral buffer keeps its size."""
    s = {1: 0, 2: 1, 4: 2, 8: 3}[xt]
    for nx in LENGTHS:
        for ny in LENGTHS:
            _, ld = ld_of(ny)
            x = np.arange(nx)[:, None]
            ky = np.arange(ld)[None, :]
            idx = (((x >> s) * ld + ky) << s) + (x & (xt - 1))
            assert np.array_equal(np.sort(idx.ravel()), np.arange(nx * ld)), (nx, ny)


def fft_slot_swizzled(p, P, R):
    """fft_slot with NS2D_FFT_SWIZZLE (ns2d_fft_reg.hpp)."""
    if P >= 16:
        return p
    PR = P * R
    RS = PR // 16 if PR >= 16 else 1
    M = 16 // P
    return p ^ ((((p >> 4) // RS) & (M - 1)) * P)


def test_fft_swizzle_is_a_row_permutation_without_bank_conflicts():
    """The optional XOR swizzle of the register FFT's LDS exchanges keeps
    every slot in its 16-slot row (a line needs no more than N slots) and is
    one to one; in a 64-bank, 16-B-slot model every group of sixteen lanes
    writes (Stockham positions of each pass with span P < 16) and reads
    (sixteen consecutive positions) sixteen different bank columns, for every
    line length and lane count the kernels use."""
    for N in LENGTHS:
        for TL in (16, 32, 64, 128, 256, 512, 1024):
            if TL > N or N // TL < 2:
                continue
            E = N // TL
            P = 1
            while N // P > 1:
                R = min(min(E, 16), N // P)
                if P * R == N:
                    break   # the last pass ends in registers
                perm = [fft_slot_swizzled(p, P, R) for p in range(N)]
                assert sorted(perm) == list(range(N)), (N, TL, P, R)
                assert all(a >> 4 == b >> 4 for a, b in zip(perm, range(N)))
                for m in range(E // R):
                    for r in range(R):
                        for g in range(0, TL, 16):
                            cols = set()
                            for t in range(g, g + 16):
                                i = t + TL * m
                                k = i & (P - 1)
                                cols.add(fft_slot_swizzled((i - k) * R + k + r * P, P, R) % 16)
                            assert len(cols) == 16, (N, TL, P, R, m, r)
                for e in range(E):
                    for g in range(0, TL, 16):
                        assert len({fft_slot_swizzled(t + TL * e, P, R) % 16 for t in range(g, g + 16)}) == 16
                P *= R
