"""Pins the CPU oracle before anything is checked against it.

Each check is chosen so a plausible slip in the oracle (a dropped term, a
wrong sign or index, a transposed operand) fails it:
  * the reference's own goldens (conv o64/o128, |d| <= 1e-15 as dist/test.sh:7-15)
    and known-answer grids (y*.dat, PDF p.9);
  * bit-exact agreement with the FFTW-free part of the reference, compiled
    from its sources into oracle/_ref (weights, Pmm, Pml tables, naive DLTs,
    index map, TransMult);
  * closed forms: DFT/DCT definitions as direct sums, quadrature exactness,
    normalised Legendre functions from scipy, brute-force O(B^4) SHT.
"""
import ctypes
import math
import os

import numpy as np
import pytest
from scipy import special

from conftest import REF_PARTIAL, golden

# ---------------------------------------------------------------------------
# layout arithmetic
# ---------------------------------------------------------------------------


@pytest.mark.parametrize("bw", [8, 9, 16, 17, 23, 64, 256])
def test_table_offsets_are_prefix_sums(O, bw):
    for m in range(bw):
        acc = 0
        for l in range(m, bw):
            assert O.table_offset(m, l) == acc
            acc += O.row_size(m, l)
        assert O.table_size(m, bw) == acc


def test_table_sizes_match_published_memory():
    # PDF p.10: seminaive tables per direction <1 / 3 / 22 / 171 MB at B = 64..512
    from oracle import oracle as O
    mb = {bw: 8 * O.reduced_sph_table_size(bw, bw) / 1e6 for bw in (64, 128, 256, 512)}
    assert mb[64] < 1 and int(mb[128]) == 2 and int(mb[256]) == 22 and int(mb[512] / 1.048576) == 171
    for bw in (8, 64, 256, 512):
        assert O.sph_table_size(bw) >= O.reduced_sph_table_size(bw, bw)
    assert O.reduced_sph_table_size(256, 256) == 2812672


@pytest.mark.parametrize("bw", [8, 17, 64])
def test_index_map_is_a_bijection(O, bw):
    seen = set()
    for m in range(-(bw - 1), bw):
        for l in range(abs(m), bw):
            seen.add(O.index_of_coeff(m, l, bw))
    assert seen == set(range(bw * bw))
    assert O.index_of_coeff(0, 0, bw) == 0 and O.index_of_coeff(-1, bw - 1, bw) == bw * bw - 1


@pytest.mark.parametrize("bw", [8, 16, 64])
def test_transpose_row_size_counts_nonzeros_even_bw(O, bw):
    for m in range(bw):
        for i in range(bw):
            deg_ok = [l for l in range(m, bw)
                      if (l if m % 2 == 0 else l - 1) >= i and ((l if m % 2 == 0 else l - 1) - i) % 2 == 0]
            assert O.transpose_row_size(i, m, bw) == len(deg_ok)


def test_transpose_row_size_odd_bw_quirk(O):
    # For odd B the reference's count is one short for even cosine indices
    # <= m of even orders and odd indices < m of odd orders, and order B-1
    # (even) keeps no entries at all.  True counts by brute force below.
    def true_count(i, m, bw):
        return len([l for l in range(m, bw)
                    if (l if m % 2 == 0 else l - 1) >= i and ((l if m % 2 == 0 else l - 1) - i) % 2 == 0])

    assert O.transpose_row_size(0, 0, 17) == 8  # true count is 9
    assert sum(O.transpose_row_size(i, 16, 17) for i in range(17)) == 0
    for bw in (9, 17, 23):
        for m in range(bw):
            for i in range(bw - (m % 2)):
                short = (i % 2 == m % 2) and (i <= m if m % 2 == 0 else i < m)
                assert O.transpose_row_size(i, m, bw) == true_count(i, m, bw) - short, (bw, m, i)


# ---------------------------------------------------------------------------
# FFTW-defined primitives
# ---------------------------------------------------------------------------


@pytest.mark.parametrize("n", [1, 2, 8, 16, 17, 23, 31, 34, 46, 128, 512, 1024])
def test_dft_matches_definition(O, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    k = np.arange(n)
    M = np.exp(-2j * np.pi * np.outer(k, k) / n)
    ref = M @ x
    got = O.dft(x, -1)
    assert np.max(np.abs(got - ref)) <= 1e-12 * max(1.0, np.abs(ref).max())
    got_inv = O.dft(x, +1)
    assert np.max(np.abs(got_inv - np.conj(M) @ x)) <= 1e-12 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("n", [8, 17, 34, 46, 128, 512])
def test_dct_pair_matches_definition(O, n):
    rng = np.random.default_rng(100 + n)
    x = rng.standard_normal(n)
    j = np.arange(n)
    C2 = 2 * np.cos(np.pi * np.outer(j, j + 0.5) / n)  # [k, j]
    assert np.max(np.abs(O.redft10(x) - C2 @ x)) <= 1e-12 * np.abs(x).sum()
    C3 = 2 * np.cos(np.pi * np.outer(j + 0.5, j) / n)  # [k, j]
    C3[:, 0] = 1.0
    assert np.max(np.abs(O.redft01(x) - C3 @ x)) <= 1e-12 * np.abs(x).sum()
    # REDFT01 is the inverse of REDFT10 up to 2n
    assert np.max(np.abs(O.redft01(O.redft10(x)) / (2 * n) - x)) <= 1e-13


# ---------------------------------------------------------------------------
# weights and tables
# ---------------------------------------------------------------------------


@pytest.mark.parametrize("bw", [8, 17, 64])
def test_weights_integrate_band_limited_polynomials(O, bw):
    w = O.weights(bw)
    n = 2 * bw
    th = np.pi * (2 * np.arange(n) + 1) / (4 * bw)
    for p in range(0, 2 * bw - 1):
        exact = (1 + (-1) ** p) / (p + 1)  # int_0^pi cos^p sin
        assert abs(np.dot(w[:n], np.cos(th) ** p) - exact) < 1e-13
    np.testing.assert_allclose(w[n:], w[:n] * np.sin(th), rtol=1e-15)


def _normalised_plm(m, l, x):
    """L2-normalised P_l^m on [-1,1] with the Condon-Shortley phase."""
    c = math.sqrt((2 * l + 1) / 2 * math.factorial(l - m) / math.factorial(l + m))
    return c * special.lpmv(m, l, x)


@pytest.mark.parametrize("bw", [8, 24])
def test_cospml_rows_are_cosine_series_of_plm(O, bw):
    # Independent restatement: sample normalised P_l^m (/sin for odd m) at
    # bw Chebyshev nodes, DCT-II by direct sum, scale, strip zeros.
    th = (2 * np.arange(bw) + 1) * np.pi / (2 * bw)
    x = np.cos(th)
    j = np.arange(bw)
    C = 2 * np.cos(np.pi * np.outer(j, j + 0.5) / bw)
    for m in range(bw):
        tab = O.cospml_table(bw, m)
        for l in range(m, bw):
            f = _normalised_plm(m, l, x)
            if m % 2:
                f = f / np.sin(th)
            c = C @ f / math.sqrt(bw)
            c[0] *= math.sqrt(0.5)
            tau = (l - m) % 2
            rs = O.row_size(m, l)
            got = tab[O.table_offset(m, l):O.table_offset(m, l) + rs]
            scale = max(1.0, np.abs(c).max())
            np.testing.assert_allclose(got, c[tau:tau + 2 * rs:2], rtol=0, atol=2e-14 * scale)
            # the stripped entries are (numerically) zero
            other = np.delete(c, np.arange(tau, tau + 2 * rs, 2))
            assert np.abs(other).max(initial=0.0) < 2e-14 * scale


@pytest.mark.parametrize("bw", [8, 16, 64])
def test_transposed_table_is_adjoint(O, bw):
    # <P c, a> = <c, P^T a>: forward projection against the stripped table and
    # the inverse's cosine sum against the transposed one are transposes.
    rng = np.random.default_rng(bw)
    n = 2 * bw
    for m in (0, 1, 2, bw // 2, bw // 2 + 1, bw - 2, bw - 1):
        t = O.cospml_table(bw, m)
        tt = O.transpose_cospml_table(bw, m, t)
        # dense matrix A[l-m, i] from the stripped table
        A = np.zeros((bw - m, bw))
        for l in range(m, bw):
            tau = (l - m) % 2
            for k in range(O.row_size(m, l)):
                A[l - m, 2 * k + tau] = t[O.table_offset(m, l) + k]
        # dense matrix from the transposed table, read as InvDLTSemi reads it
        At = np.zeros((bw - m, bw))
        pos = 0
        for i in range(bw):
            if i == bw - 1 and m % 2:
                break
            first = (i - m) + (m % 2) if i > m else i % 2
            rs = O.transpose_row_size(i, m, bw)
            for k in range(rs):
                At[first + 2 * k, i] = tt[pos + k]
            pos += rs
        if m % 2:
            A[:, bw - 1] = 0.0  # the inverse skips the last cosine index of odd orders
        np.testing.assert_array_equal(At, A)


def _ref_partial():
    if not os.path.exists(REF_PARTIAL):
        pytest.skip("oracle/_ref not built (needs the reference tree; `make -C oracle ref`)")
    return ctypes.CDLL(REF_PARTIAL)


_dp = ctypes.POINTER(ctypes.c_double)


def _p(a):
    return a.ctypes.data_as(_dp)


@pytest.mark.parametrize("bw", [8, 17, 64, 256])
def test_weights_bit_exact_vs_reference(O, bw):
    R = _ref_partial()
    w = np.empty(4 * bw)
    R.GenerateWeightsForDLT(ctypes.c_int(bw), _p(w))
    np.testing.assert_array_equal(O.weights(bw), w)


@pytest.mark.parametrize("bw,m", [(8, 0), (8, 3), (64, 0), (64, 17), (64, 63), (128, 100)])
def test_pml_table_bit_exact_vs_reference(O, bw, m):
    R = _ref_partial()
    t = np.empty(2 * bw * (bw - m))
    ws = np.empty(16 * bw)
    R.GeneratePmlTable(ctypes.c_int(bw), ctypes.c_int(m), _p(t), _p(ws))
    np.testing.assert_array_equal(O.pml_table(bw, m), t)


def test_recurrence_coefficients_bit_exact_vs_reference(O):
    R = _ref_partial()
    R.L2_an.restype = ctypes.c_double
    R.L2_cn.restype = ctypes.c_double
    lib = O.lib()
    for m in (0, 1, 5, 40):
        for l in range(m, 64):
            assert lib.so_l2_an(m, l) == R.L2_an(ctypes.c_int(m), ctypes.c_int(l))
            assert lib.so_l2_cn(m, l) == R.L2_cn(ctypes.c_int(m), ctypes.c_int(l))
    args = (2 * np.arange(64) + 1) * np.pi / 128
    for m in (0, 1, 2, 31, 63):
        out = np.empty(64)
        R.Pmm_L2(ctypes.c_int(m), _p(args), ctypes.c_int(64), _p(out))
        np.testing.assert_array_equal(O.pmm_l2(m, args), out)


@pytest.mark.parametrize("bw", [8, 64])
def test_naive_dlts_and_index_bit_exact_vs_reference(O, bw):
    R = _ref_partial()
    rng = np.random.default_rng(7)
    w = O.weights(bw)
    for m in (0, 3, bw - 1):
        pml = O.pml_table(bw, m)
        data = rng.uniform(-1, 1, 2 * bw)
        out = np.empty(bw - m)
        ws = np.empty(2 * bw)
        R.DLTNaive(_p(data), ctypes.c_int(bw), ctypes.c_int(m), _p(w), _p(out), _p(pml), _p(ws))
        np.testing.assert_array_equal(O.dlt_naive(data, bw, m, w, pml), out)
        coeffs = rng.uniform(-1, 1, bw - m)
        coeffs[0] = 0.0
        res = np.empty(2 * bw)
        R.InvDLTNaive(_p(coeffs), ctypes.c_int(bw), ctypes.c_int(m), _p(res), _p(pml))
        np.testing.assert_array_equal(O.inv_dlt_naive(coeffs, bw, m, pml), res)
    for m in range(-(bw - 1), bw):
        for l in range(abs(m), bw):
            assert O.index_of_coeff(m, l, bw) == R.IndexOfHarmonicCoeff(ctypes.c_int(m), ctypes.c_int(l),
                                                                       ctypes.c_int(bw))


def test_trans_mult_bit_exact_vs_reference_with_complex_filter(O):
    # A filter with an imaginary part exercises util.c:26 (x*v - y*u).
    R = _ref_partial()
    bw = 16
    rng = np.random.default_rng(3)
    a = [rng.uniform(-1, 1, bw * bw) for _ in range(2)] + [rng.uniform(-1, 1, bw) for _ in range(2)]
    rr, ir = np.empty(bw * bw), np.empty(bw * bw)
    R.TransMult(*[_p(x) for x in a], _p(rr), _p(ir), ctypes.c_int(bw))
    orr, oir = O.trans_mult(*a, bw)
    np.testing.assert_array_equal(orr, rr)
    np.testing.assert_array_equal(oir, ir)


# ---------------------------------------------------------------------------
# transforms: known answers, goldens, brute force, round trips
# ---------------------------------------------------------------------------

KNOWN = [
    ("y20_bw8", 8, {(0, 2): 1.0}),
    ("y31_bw8", 8, {(1, 3): 1.0}),
    ("y43_bw23", 23, {(3, 4): math.sqrt(2) + math.pi * 1j}),
    ("yMix_bw17", 17, {(1, 1): 1.0, (-2, 5): 3 - 2j}),
]


@pytest.mark.parametrize("name,bw,expect", KNOWN)
def test_known_answer_forward(O, name, bw, expect):
    re, im = golden(f"y_{name}.npy")
    rc, ic = O.fst(re, im, bw)
    ex = np.zeros(bw * bw, complex)
    for (m, l), v in expect.items():
        ex[O.index_of_coeff(m, l, bw)] = v
    assert np.abs(rc + 1j * ic - ex).max() < 1e-14


@pytest.mark.parametrize("name,bw,expect", KNOWN[:2])
def test_known_answer_inverse_even_bw(O, name, bw, expect):
    re, im = golden(f"y_{name}.npy")
    rc, ic = np.zeros(bw * bw), np.zeros(bw * bw)
    for (m, l), v in expect.items():
        rc[O.index_of_coeff(m, l, bw)] = v.real
        ic[O.index_of_coeff(m, l, bw)] = np.imag(v)
    rd, idd = O.inv_fst(rc, ic, bw)
    assert np.abs(rd - re).max() < 1e-14 and np.abs(idd - im).max() < 1e-14


@pytest.mark.parametrize("name,bw,expect", KNOWN[2:])
def test_known_answer_inverse_odd_bw(O, name, bw, expect):
    # The odd-bandwidth inverse against the reference's own y43_bw23 and
    # yMix_bw17 grids: none of their coefficients has degree bw-1, the only
    # degree the Transpose_RowSize quirk (cospml.c:270-288) touches.
    re, im = golden(f"y_{name}.npy")
    rc, ic = np.zeros(bw * bw), np.zeros(bw * bw)
    for (m, l), v in expect.items():
        assert l < bw - 1
        rc[O.index_of_coeff(m, l, bw)] = np.real(v)
        ic[O.index_of_coeff(m, l, bw)] = np.imag(v)
    rd, idd = O.inv_fst(rc, ic, bw)
    scale = max(np.abs(re).max(), np.abs(im).max())
    assert max(np.abs(rd - re).max(), np.abs(idd - im).max()) < 1e-14 * scale


@pytest.mark.parametrize("bw", [64, 128])
def test_conv_golden(O, bw):
    s = golden(f"conv{bw}_signal.npy")
    f = np.repeat(golden(f"conv{bw}_filter.npy"), 2 * bw)
    o = golden(f"conv{bw}_out.npy")
    z = np.zeros_like(s)
    rr, ir = O.conv(s, z, f, z, bw)
    assert np.abs(rr - o).max() <= 1e-15  # dist/test.sh:7-15 tolerance
    assert np.abs(ir).max() < 1e-15


def _synthesise(bw, c):
    n = 2 * bw
    th = np.pi * (2 * np.arange(n) + 1) / (4 * bw)
    ph = 2 * np.pi * np.arange(n) / n
    TH, PH = np.meshgrid(th, ph, indexing="ij")
    from oracle import oracle as O
    f = np.zeros((n, n), complex)
    for m in range(-(bw - 1), bw):
        for l in range(abs(m), bw):
            f += c[O.index_of_coeff(m, l, bw)] * special.sph_harm_y(l, m, TH, PH)
    return f


@pytest.mark.parametrize("bw", [4, 9])
def test_brute_force_sht(O, bw):
    # O(B^4) synthesis with scipy's orthonormal Condon-Shortley Y_l^m.
    rc, ic = O.make_coeffs(11 + bw, bw, 0)
    f = _synthesise(bw, rc + 1j * ic)
    rd, idd = O.inv_fst(rc, ic, bw)
    if bw % 2 == 0:
        assert np.abs(rd + 1j * idd - f.ravel()).max() < 1e-13
    r2, i2 = O.fst(f.real, f.imag, bw)
    assert np.abs(r2 + 1j * i2 - (rc + 1j * ic)).max() < 1e-13


@pytest.mark.parametrize("bw,cutoff,fmt", [(8, 8, 0), (16, 4, 0), (32, 32, 0), (64, 64, 0), (64, 16, 0),
                                           (32, 32, 1), (64, 16, 1)])
def test_round_trip(O, bw, cutoff, fmt):
    kind = 2 if fmt == 1 else 0
    rc, ic = O.make_coeffs(5, bw, kind)
    rd, idd = O.inv_fst(rc, ic, bw, data_format=fmt, cutoff=cutoff)
    if fmt == 1:
        assert np.abs(idd).max() < 1e-12  # Hermitian coefficients -> real grid
    r2, i2 = O.fst(rd, idd, bw, data_format=fmt, cutoff=cutoff)
    assert np.abs(r2 + 1j * i2 - (rc + 1j * ic)).max() < 1e-12


@pytest.mark.parametrize("bw", [8, 16])
def test_cutoff_zero_is_the_exact_transform(O, bw):
    # At cutoff 0 the reference's order-0 naive table is only partly written
    # (cospml.c:456,464), so its output there is undefined; the oracle (like
    # the GPU path) computes order 0 naively on a full Pml table and returns
    # the exact transform, checked here against the brute-force synthesis.
    rc, ic = O.make_coeffs(77 + bw, bw, 0)
    f = _synthesise(bw, rc + 1j * ic)
    rd, idd = O.inv_fst(rc, ic, bw, cutoff=0)
    assert np.abs(rd + 1j * idd - f.ravel()).max() < 1e-13
    r2, i2 = O.fst(f.real, f.imag, bw, cutoff=0)
    assert np.abs(r2 + 1j * i2 - (rc + 1j * ic)).max() < 1e-13
    assert len(O.tables(bw, 0).order(0)) == 2 * bw * bw


@pytest.mark.parametrize("bw", [9, 11, 17])
def test_brute_force_inverse_odd_bw(O, bw):
    # Odd bandwidths, every degree below bw-1: the inverse is the exact
    # synthesis (scipy's orthonormal Y_l^m summed on the grid).
    rc, ic = O.make_coeffs(31 + bw, bw, 0)
    deg = O.degrees(bw)
    rc[deg == bw - 1] = 0.0
    ic[deg == bw - 1] = 0.0
    f = _synthesise(bw, rc + 1j * ic)
    rd, idd = O.inv_fst(rc, ic, bw)
    assert np.abs(rd + 1j * idd - f.ravel()).max() < 1e-13


def _top_degree_quirk_model(bw, c):
    """Synthesis of degree-(bw-1) coefficients c[m] (|m| < bw) as the
    reference's odd-bw inverse performs it, built from the mathematics: per
    order, the cosine series (in theta) of P~_{bw-1}^{|m|} (divided by
    sin(theta) for odd |m|) with the entries at cosine indices i = |m|
    (mod 2), i <= |m| for even |m| and i < |m| for odd |m|, removed -- the
    rows whose length Transpose_RowSize (cospml.c:270-288) counts one short,
    losing their last term, degree bw-1 -- summed at the 2bw rings, times
    sin(theta) for odd
    |m|, times e^{i m phi}/sqrt(2 pi) and the Condon-Shortley sign of
    Y_l^m for m < 0."""
    n, l = 2 * bw, bw - 1
    th = np.pi * (2 * np.arange(n) + 1) / (4 * bw)
    ph = 2 * np.pi * np.arange(n) / n
    nodes = (2 * np.arange(bw) + 1) * np.pi / (2 * bw)  # bw Chebyshev nodes: exact for cos degree < bw
    k = np.arange(bw)
    out = np.zeros((n, n), complex)
    for m, v in c.items():
        am = abs(m)
        g = _normalised_plm(am, l, np.cos(nodes))
        if am % 2:
            g = g / np.sin(nodes)
        a = (2.0 / bw) * np.cos(np.outer(k, nodes)) @ g  # cos-series coefficients, a[0] doubled
        a[0] *= 0.5
        drop = ((k < am) | ((k == am) & (am % 2 == 0))) & ((k - am) % 2 == 0)
        a[drop] = 0.0
        theta_part = np.cos(np.outer(th, k)) @ a
        if am % 2:
            theta_part *= np.sin(th)
        sign = (-1) ** am if m < 0 else 1.0
        out += v * sign * np.outer(theta_part, np.exp(1j * m * ph)) / math.sqrt(2 * math.pi)
    return out


@pytest.mark.parametrize("bw", [9, 11, 17])
def test_odd_bw_top_degree_matches_quirk_model(O, bw):
    # Degree bw-1 alone at every order: the oracle's inverse equals the model
    # above, so the terms it drops are exactly the counted-out ones and no
    # other (a check of the oracle against the mathematics, not against a
    # run of the reference: the dropped set itself is read from cospml.c).
    rng = np.random.default_rng(bw)
    rc, ic = np.zeros(bw * bw), np.zeros(bw * bw)
    c = {}
    for m in range(-(bw - 1), bw):
        v = complex(rng.uniform(-1, 1), rng.uniform(-1, 1))
        c[m] = v
        rc[O.index_of_coeff(m, bw - 1, bw)] = v.real
        ic[O.index_of_coeff(m, bw - 1, bw)] = v.imag
    rd, idd = O.inv_fst(rc, ic, bw)
    model = _top_degree_quirk_model(bw, c)
    assert np.abs(rd + 1j * idd - model.ravel()).max() < 1e-13
    # and the dropped part is not empty: the exact synthesis differs
    assert np.abs(rd + 1j * idd - _synthesise(bw, rc + 1j * ic).ravel()).max() > 1e-3


def test_odd_bw_inverse_drops_terms_like_the_reference(O):
    # Consequence of the Transpose_RowSize quirk: at B=17 the top order
    # m=16 is lost by the inverse, while the forward stays exact.
    bw = 17
    rc, ic = np.zeros(bw * bw), np.zeros(bw * bw)
    rc[O.index_of_coeff(16, 16, bw)] = 1.0
    rd, idd = O.inv_fst(rc, ic, bw)
    assert np.abs(rd).max() == 0.0 and np.abs(idd).max() == 0.0


def test_fzt_matches_m0_of_forward(O):
    bw = 32
    rc, ic = O.make_coeffs(9, bw, 0)
    rd, idd = O.inv_fst(rc, ic, bw)
    fr, fi = O.fzt(rd, idd, bw, data_format=0)
    np.testing.assert_allclose(fr, rc[:bw], atol=1e-13)
    np.testing.assert_allclose(fi, ic[:bw], atol=1e-13)
    fr, fi = O.fzt(rd, idd, bw, data_format=1)
    assert np.all(fi == 0.0)


def test_generators_are_deterministic(O):
    a = O.make_coeffs(123, 16, 1)
    b = O.make_coeffs(123, 16, 1)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    u = O.uniform(42, 1000)
    assert 0.0 <= u.min() and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.05
    rc, ic = O.make_coeffs(7, 64, 1)
    deg = O.degrees(64)
    var = np.array([np.mean(rc[deg == l] ** 2 + ic[deg == l] ** 2) for l in range(64)])
    scaled = var * (np.arange(64) + 1) ** 2
    assert abs(scaled[8:].mean() - 1.0) < 0.1  # >= 17 samples per degree from l = 8 on
