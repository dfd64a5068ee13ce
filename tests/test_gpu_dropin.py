"""GPU parity of the S2kit drop-ins (host pointers, through the C ABI)
against the CPU oracle on the same inputs and the same tables.

Tolerances: FP64 transforms must agree with the oracle to 1e-11 relative
(normwise: max|d| / max|ref|), the bar the reference's FP64 path is held to;
tables and weights to ~1e-14 relative; index shuffles (TransposeCosPmlTable)
bit-exactly.
"""
import ctypes

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

_dp = ctypes.POINTER(ctypes.c_double)


@pytest.fixture(scope="module")
def S():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    import s2kit_amd
    s2kit_amd.lib()
    return s2kit_amd


def rel(a, b):
    b = np.asarray(b)
    return float(np.abs(np.asarray(a) - b).max() / max(np.abs(b).max(), 1e-300))


def crel(ar, ai, br, bi):
    return rel(np.asarray(ar) + 1j * np.asarray(ai), np.asarray(br) + 1j * np.asarray(bi))


def _p(a):
    return a.ctypes.data_as(_dp)


# ---------------------------------------------------------------------------
# weights and tables
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("bw", [8, 17, 64, 256])
def test_weights(S, O, bw):
    print(f"weights bw={bw}", flush=True)
    assert rel(S.weights(bw), O.weights(bw)) < 1e-14


@pytest.mark.parametrize("bw,cutoff", [(8, 8), (17, 17), (23, 5), (64, 64), (64, 16), (128, 128)])
def test_table_builders(S, O, bw, cutoff):
    print(f"tables bw={bw} cutoff={cutoff}", flush=True)
    ht = S.HostTables(bw, cutoff)
    ot = O.Tables(bw, cutoff)
    n = ot.fwd.size
    assert np.array_equal(ht.offsets(), ot.off)
    assert rel(ht.fwd[:n], ot.fwd) < 1e-13
    assert rel(ht.inv[:n], ot.inv) < 1e-13


@pytest.mark.parametrize("bw,m", [(8, 0), (8, 7), (17, 4), (17, 16), (64, 0), (64, 33), (64, 63)])
def test_single_order_builders(S, O, bw, m):
    t = S.cospml_table(bw, m)
    assert rel(t, O.cospml_table(bw, m)) < 1e-13
    ot = O.cospml_table(bw, m)
    # pure data movement: bit-exact against the oracle's transposition of the same input
    assert np.array_equal(S.transpose_cospml_table(bw, m, ot), O.transpose_cospml_table(bw, m, ot))
    assert rel(S.pml_table(bw, m), O.pml_table(bw, m)) < 1e-13


# ---------------------------------------------------------------------------
# transforms on the oracle's tables (same inputs, same tables)
# ---------------------------------------------------------------------------
def fst_with(S, rd, idd, bw, tabs, fmt, cutoff, w=None):
    L = S.lib()
    w = tabs_w = (S.weights(bw) if w is None else w)
    rc, ic = np.empty(bw * bw), np.empty(bw * bw)
    rd, idd = np.ascontiguousarray(rd), np.ascontiguousarray(idd)
    L.FSTSemiMemo(_p(rd), _p(idd), _p(rc), _p(ic), bw, tabs._pf, None, fmt, cutoff, None, None, _p(tabs_w))
    return rc, ic


def inv_with(S, rc, ic, bw, tabs, fmt, cutoff):
    L = S.lib()
    n = 2 * bw
    rd, idd = np.empty(n * n), np.empty(n * n)
    rc, ic = np.ascontiguousarray(rc), np.ascontiguousarray(ic)
    L.InvFSTSemiMemo(_p(rc), _p(ic), _p(rd), _p(idd), bw, tabs._pi, None, fmt, cutoff, None, None)
    return rd, idd


CASES = [(8, 8, 0), (17, 17, 0), (23, 23, 0), (64, 64, 0), (64, 16, 0), (64, 1, 0), (64, 0, 0), (64, 64, 1),
         (64, 20, 1), (128, 128, 0), (256, 256, 0), (31, 31, 1), (127, 40, 0)]


@pytest.mark.parametrize("bw,cutoff,fmt", CASES)
def test_forward_matches_oracle(S, O, bw, cutoff, fmt):
    print(f"forward bw={bw} cutoff={cutoff} fmt={fmt}", flush=True)
    rng = np.random.default_rng(bw * 100 + cutoff)
    n = 2 * bw
    rd = rng.uniform(-1, 1, n * n)
    idd = np.zeros(n * n) if fmt == 1 else rng.uniform(-1, 1, n * n)
    ot = O.tables(bw, cutoff)
    w = O.weights(bw)
    rc, ic = fst_with(S, rd, idd, bw, ot, fmt, cutoff, w)
    orc, oic = O.fst(rd, idd, bw, data_format=fmt, cutoff=cutoff)
    assert crel(rc, ic, orc, oic) < 1e-11


@pytest.mark.parametrize("bw,cutoff,fmt", CASES)
def test_inverse_matches_oracle(S, O, bw, cutoff, fmt):
    print(f"inverse bw={bw} cutoff={cutoff} fmt={fmt}", flush=True)
    rc, ic = O.make_coeffs(bw + 7 * cutoff, bw, 0)
    ot = O.tables(bw, cutoff)
    rd, idd = inv_with(S, rc, ic, bw, ot, fmt, cutoff)
    ord_, oid = O.inv_fst(rc, ic, bw, data_format=fmt, cutoff=cutoff)
    assert crel(rd, idd, ord_, oid) < 1e-11


def test_round_trip_config1(S, O):
    # B=64 on the 128x128 grid: coefficients U[-1,1], inverse then forward.
    bw = 64
    tabs = S.HostTables(bw)
    rc, ic = O.make_coeffs(1, bw, 0)
    rd, idd = S.inv_fst(rc, ic, bw, tabs)
    r2, i2 = S.fst(rd, idd, bw, tabs)
    d = np.abs((r2 - rc) + 1j * (i2 - ic))
    print(f"config1 round trip: max abs {d.max():.3e}", flush=True)
    assert d.max() < 1e-12


def test_hybrid_round_trip_like_test_s2_semi_memo(S, O):
    # test_s2_semi_memo.c:74-75,156-172: cutoff = bw/4, Hermitian coefficients.
    bw = 128
    tabs = S.HostTables(bw, bw // 4)
    rc, ic = O.make_coeffs(3, bw, 2)
    rd, idd = S.inv_fst(rc, ic, bw, tabs)
    assert np.abs(idd).max() < 1e-12
    r2, i2 = S.fst(rd, idd, bw, tabs)
    assert np.abs((r2 - rc) + 1j * (i2 - ic)).max() < 1e-11


KNOWN = [
    ("y20_bw8", 8, {(0, 2): 1.0}),
    ("y31_bw8", 8, {(1, 3): 1.0}),
    ("y43_bw23", 23, {(3, 4): np.sqrt(2) + np.pi * 1j}),
    ("yMix_bw17", 17, {(1, 1): 1.0, (-2, 5): 3 - 2j}),
]


@pytest.mark.parametrize("name,bw,expect", KNOWN)
def test_known_answers(S, O, name, bw, expect):
    re, im = golden(f"y_{name}.npy")
    tabs = S.HostTables(bw)
    rc, ic = S.fst(re, im, bw, tabs)
    ex = np.zeros(bw * bw, complex)
    for (m, l), v in expect.items():
        ex[O.index_of_coeff(m, l, bw)] = v
    assert np.abs(rc + 1j * ic - ex).max() < 1e-13
    # the inverse of the known coefficients reproduces the oracle's grid
    # (for odd bw both keep the reference's truncated transposed rows)
    rd, idd = S.inv_fst(ex.real.copy(), ex.imag.copy(), bw, tabs)
    ord_, oid = O.inv_fst(ex.real.copy(), ex.imag.copy(), bw)
    assert crel(rd, idd, ord_, oid) < 1e-11


@pytest.mark.parametrize("bw", [64, 128])
def test_conv_golden(S, bw):
    s = golden(f"conv{bw}_signal.npy")
    f = np.repeat(golden(f"conv{bw}_filter.npy"), 2 * bw)
    o = golden(f"conv{bw}_out.npy")
    z = np.zeros_like(s)
    rr, ir = S.conv(s, z, f, z, bw)
    d = np.abs(rr - o).max()
    print(f"conv golden bw={bw}: max |d| = {d:.3e} (reference test tolerance 1e-15)", flush=True)
    assert d < 1e-14
    assert np.abs(ir).max() < 1e-14


def test_conv_config2(S, O):
    # B=128: signal from U[-1,1] complex coefficients; zonal filter
    # h_l = exp(-l(l+1)/(2*20^2)) * U[0.5,1].
    bw = 128
    rc, ic = O.make_coeffs(22, bw, 0)
    sr, si = O.inv_fst(rc, ic, bw)
    u = O.uniform(23, bw)
    l = np.arange(bw)
    hl = np.exp(-l * (l + 1) / (2 * 20.0 ** 2)) * (0.5 + 0.5 * u)
    hc_r, hc_i = np.zeros(bw * bw), np.zeros(bw * bw)
    for ll in range(bw):
        hc_r[O.index_of_coeff(0, ll, bw)] = hl[ll]
    hr, hi = O.inv_fst(hc_r, hc_i, bw)
    rr, ir = S.conv(sr, si, hr, hi, bw)
    orr, oir = O.conv(sr, si, hr, hi, bw)
    e = crel(rr, ir, orr, oir)
    print(f"config2 conv rel err {e:.3e}", flush=True)
    assert e < 1e-11


@pytest.mark.parametrize("fmt", [0, 1])
def test_fzt(S, O, fmt):
    bw = 64
    rng = np.random.default_rng(5)
    n = 2 * bw
    rd, idd = rng.uniform(-1, 1, n * n), rng.uniform(-1, 1, n * n)
    tabs = S.HostTables(bw)
    zr, zi = S.fzt(rd, idd, bw, tabs, data_format=fmt)
    ozr, ozi = O.fzt(rd, idd, bw, data_format=fmt)
    assert crel(zr, zi, ozr, ozi) < 1e-11
    if fmt == 1:
        assert np.all(zi == 0.0)


def test_trans_mult_with_complex_filter(S, O):
    bw = 32
    rng = np.random.default_rng(4)
    a = [rng.uniform(-1, 1, bw * bw) for _ in range(2)] + [rng.uniform(-1, 1, bw) for _ in range(2)]
    rr, ir = S.trans_mult(*a, bw)
    orr, oir = O.trans_mult(*a, bw)
    assert crel(rr, ir, orr, oir) < 1e-15


def test_table_cache_sees_in_place_rewrite(S, O):
    # Same pointers, new contents: the cached device copy must be refreshed.
    bw = 16
    ot = O.Tables(bw)
    n = 2 * bw
    rng = np.random.default_rng(9)
    rd, idd = rng.uniform(-1, 1, n * n), rng.uniform(-1, 1, n * n)
    w = O.weights(bw)
    a = fst_with(S, rd, idd, bw, ot, 0, bw, w)
    ot.fwd *= 2.0
    b = fst_with(S, rd, idd, bw, ot, 0, bw, w)
    assert crel(b[0], b[1], 2 * a[0], 2 * a[1]) < 1e-14


@pytest.mark.parametrize("m", [0, 3, 40])
def test_table_cache_sees_one_entry_edited_in_place(S, O, m):
    # One entry of one order changed in place, away from any sparse sample
    # grid (not a multiple of len/64, not the last): both directions must
    # read the edited value, as the reference reads the caller's table on
    # every call.  The oracle transforms with the same edited table.
    print(f"table cache one-entry edit m={m}", flush=True)
    bw, fmt = 64, 0
    ot = O.Tables(bw)
    n = 2 * bw
    rng = np.random.default_rng(10 + m)
    rd, idd = rng.uniform(-1, 1, n * n), rng.uniform(-1, 1, n * n)
    rc, ic = rng.uniform(-1, 1, bw * bw), rng.uniform(-1, 1, bw * bw)
    w = O.weights(bw)
    length = O.table_size(m, bw)
    step = max(1, length // 64)
    k = next(k for k in range(1, length - 1) if k % step)
    fst_with(S, rd, idd, bw, ot, fmt, bw, w)  # cache both directions' device copies
    inv_with(S, rc, ic, bw, ot, fmt, bw)
    ot.fwd[ot.off[m] + k] += 0.5
    ot.inv[ot.off[m] + k] -= 0.25
    a = fst_with(S, rd, idd, bw, ot, fmt, bw, w)
    b = O.fst(rd, idd, bw, data_format=fmt, tabs=ot)
    assert crel(a[0], a[1], b[0], b[1]) < 1e-11
    c = inv_with(S, rc, ic, bw, ot, fmt, bw)
    d = O.inv_fst(rc, ic, bw, data_format=fmt, tabs=ot)
    assert crel(c[0], c[1], d[0], d[1]) < 1e-11


# ---- FST_semi_fly.h: the same transforms with library-made tables -----------
FLY_CASES = [(8, 8, 0), (17, 17, 1), (64, 22, 0), (64, 22, 1), (128, 54, 0), (64, 0, 0)]


@pytest.mark.parametrize("bw,cutoff,fmt", FLY_CASES)
def test_fly_matches_oracle(S, O, bw, cutoff, fmt):
    # cutoff = bw/2 - 10 is test_s2_semi_fly.c:128's choice
    print(f"fly bw={bw} cutoff={cutoff} fmt={fmt}", flush=True)
    kind = 2 if fmt == 1 else 0
    rc, ic = O.make_coeffs(bw + cutoff, bw, kind)
    rd, idd = S.inv_fst_fly(rc, ic, bw, cutoff, data_format=fmt)
    ord_, oid = O.inv_fst(rc, ic, bw, data_format=fmt, cutoff=cutoff)
    assert crel(rd, idd, ord_, oid) < 1e-11
    r2, i2 = S.fst_fly(ord_, oid, bw, cutoff, data_format=fmt)
    orc, oic = O.fst(ord_, oid, bw, data_format=fmt, cutoff=cutoff)
    assert crel(r2, i2, orc, oic) < 1e-11
    # band-limited data: the round trip returns the coefficients (even bw only:
    # for odd bw the reference's transposed rows are short, Transpose_RowSize)
    if bw % 2 == 0:
        assert crel(r2, i2, rc, ic) < 1e-11


@pytest.mark.parametrize("fmt", [0, 1])
def test_fzt_fly(S, O, fmt):
    bw = 64
    rng = np.random.default_rng(6)
    n = 2 * bw
    rd, idd = rng.uniform(-1, 1, n * n), rng.uniform(-1, 1, n * n)
    zr, zi = S.fzt_fly(rd, idd, bw, data_format=fmt)
    ozr, ozi = O.fzt(rd, idd, bw, data_format=fmt)
    assert crel(zr, zi, ozr, ozi) < 1e-11
    if fmt == 1:
        assert np.all(zi == 0.0)


def test_conv_fly_golden(S):
    # the conv goldens hold for ConvOn2SphereSemiFly too (same transform)
    bw = 64
    s = golden(f"conv{bw}_signal.npy")
    f = np.repeat(golden(f"conv{bw}_filter.npy"), 2 * bw)
    o = golden(f"conv{bw}_out.npy")
    z = np.zeros_like(s)
    rr, ir = S.conv(s, z, f, z, bw, fly=True)
    assert np.abs(rr - o).max() < 1e-14
    assert np.abs(ir).max() < 1e-14


# ---- seminaive.h, naive.h: one order at a time -------------------------------
def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300))


@pytest.mark.parametrize("bw,m", [(8, 0), (8, 3), (17, 4), (17, 16), (64, 0), (64, 33), (256, 1), (256, 200)])
def test_per_order_dlts_match_oracle(S, O, bw, m):
    print(f"per-order DLTs bw={bw} m={m}", flush=True)
    rng = np.random.default_rng(bw * 1000 + m)
    n = 2 * bw
    w = O.weights(bw)
    data = rng.uniform(-1, 1, n)
    coeffs = rng.uniform(-1, 1, bw - m)
    sinv = np.sin((2 * np.arange(n) + 1) * np.pi / (2 * n))
    ct = O.cospml_table(bw, m)
    tt = O.transpose_cospml_table(bw, m, ct)
    pml = O.pml_table(bw, m)
    assert rel(S.dlt_semi(data, bw, m, ct, w), O.dlt_semi(data, bw, m, ct, w)) < 1e-11
    assert rel(S.inv_dlt_semi(coeffs, bw, m, tt, sinv), O.inv_dlt_semi(coeffs, bw, m, tt, sinv)) < 1e-11
    assert rel(S.dlt_naive(data, bw, m, pml, w), O.dlt_naive(data, bw, m, w, pml)) < 1e-11
    assert rel(S.inv_dlt_naive(coeffs, bw, m, pml), O.inv_dlt_naive(coeffs, bw, m, pml)) < 1e-11
    # the two forward algorithms agree on the grid's samples of one order
    if bw % 2 == 0:
        assert rel(S.dlt_semi(data, bw, m, ct, w), S.dlt_naive(data, bw, m, pml, w)) < 1e-11
