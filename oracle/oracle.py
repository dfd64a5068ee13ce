"""ctypes wrapper over the CPU FP64 oracle (oracle/libs2oracle.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by ``__graft_entry__.smoke()``
and by bench.py's cpu_baseline leg, always as the reference the results are
compared with or the timed CPU port -- never by the product package ``s2kit_amd``.

Each wrapper names the reference function it restates; the C source carries
the file:line citations.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = ctypes.POINTER(ctypes.c_double)
_i = ctypes.c_int


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libs2oracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        L = ctypes.CDLL(path)
        for name in ("so_table_size", "so_reduced_sph_table_size", "so_reduced_naive_table_size",
                     "so_table_offset", "so_row_size", "so_index_of_coeff"):
            getattr(L, name).argtypes = [_i, _i]
            getattr(L, name).restype = _i
        L.so_sph_table_size.argtypes = [_i]
        L.so_sph_table_size.restype = _i
        L.so_transpose_row_size.argtypes = [_i, _i, _i]
        L.so_transpose_row_size.restype = _i
        L.so_dft.argtypes = [_i, _dp, _dp, _dp, _dp, _i]
        L.so_redft10.argtypes = [_i, _dp, _dp]
        L.so_redft01.argtypes = [_i, _dp, _dp]
        L.so_weights.argtypes = [_i, _dp]
        L.so_pmm_l2.argtypes = [_i, _dp, _i, _dp]
        L.so_l2_an.argtypes = [_i, _i]
        L.so_l2_an.restype = ctypes.c_double
        L.so_l2_cn.argtypes = [_i, _i]
        L.so_l2_cn.restype = ctypes.c_double
        L.so_cospml_table.argtypes = [_i, _i, _dp]
        L.so_transpose_cospml_table.argtypes = [_i, _i, _dp, _dp]
        L.so_pml_table.argtypes = [_i, _i, _dp]
        L.so_tables.argtypes = [_i, _i, _dp, _dp, ctypes.POINTER(ctypes.c_long)]
        L.so_dlt_semi.argtypes = [_dp, _i, _i, _dp, _dp, _dp]
        L.so_inv_dlt_semi.argtypes = [_dp, _i, _i, _dp, _dp, _dp]
        L.so_dlt_naive.argtypes = [_dp, _i, _i, _dp, _dp, _dp]
        L.so_inv_dlt_naive.argtypes = [_dp, _i, _i, _dp, _dp]
        pp = ctypes.POINTER(_dp)
        L.so_fst.argtypes = [_dp, _dp, _dp, _dp, _i, pp, _dp, _i, _i]
        L.so_inv_fst.argtypes = [_dp, _dp, _dp, _dp, _i, pp, _i, _i]
        L.so_fzt.argtypes = [_dp, _dp, _dp, _dp, _i, _dp, _dp, _i]
        L.so_trans_mult.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, _i]
        L.so_conv.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, _i]
        L.so_uniform.argtypes = [ctypes.c_ulonglong, ctypes.c_long, _dp]
        L.so_make_coeffs.argtypes = [ctypes.c_ulonglong, _i, _i, _dp, _dp]
        _LIB = L
    return _LIB


def _p(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---- layout arithmetic ------------------------------------------------------
def table_size(m, bw):
    return lib().so_table_size(m, bw)


def sph_table_size(bw):
    return lib().so_sph_table_size(bw)


def reduced_sph_table_size(bw, m):
    return lib().so_reduced_sph_table_size(bw, m)


def reduced_naive_table_size(bw, m):
    return lib().so_reduced_naive_table_size(bw, m)


def table_offset(m, l):
    return lib().so_table_offset(m, l)


def row_size(m, l):
    return lib().so_row_size(m, l)


def transpose_row_size(row, m, bw):
    return lib().so_transpose_row_size(row, m, bw)


def index_of_coeff(m, l, bw):
    return lib().so_index_of_coeff(m, l, bw)


# ---- primitives -------------------------------------------------------------
def dft(x, sign=-1):
    x = np.asarray(x, dtype=np.complex128)
    n = x.size
    re, im = f64(x.real), f64(x.imag)
    ore, oim = np.empty(n), np.empty(n)
    lib().so_dft(n, _p(re), _p(im), _p(ore), _p(oim), sign)
    return ore + 1j * oim


def redft10(x):
    x = f64(x)
    out = np.empty_like(x)
    lib().so_redft10(x.size, _p(x), _p(out))
    return out


def redft01(x):
    x = f64(x)
    out = np.empty_like(x)
    lib().so_redft01(x.size, _p(x), _p(out))
    return out


def weights(bw):
    w = np.empty(4 * bw)
    lib().so_weights(bw, _p(w))
    return w


def pmm_l2(m, args):
    args = f64(args)
    out = np.empty_like(args)
    lib().so_pmm_l2(m, _p(args), args.size, _p(out))
    return out


def cospml_table(bw, m):
    t = np.empty(table_size(m, bw))
    lib().so_cospml_table(bw, m, _p(t))
    return t


def transpose_cospml_table(bw, m, t):
    t = f64(t)
    tt = np.zeros(table_size(m, bw))  # odd bw: the transposed rows leave a tail unwritten
    lib().so_transpose_cospml_table(bw, m, _p(t), _p(tt))
    return tt


def pml_table(bw, m):
    t = np.empty(2 * bw * (bw - m))
    lib().so_pml_table(bw, m, _p(t))
    return t


class Tables:
    """Forward and transposed mixed tables, laid out as
    SemiNaive_Naive_Pml_Table / Transpose_SemiNaive_Naive_Pml_Table
    (cospml.c:451-518) lay them out; ``off[m]`` is order m's start."""

    def __init__(self, bw, cutoff=None):
        cutoff = bw if cutoff is None else cutoff
        self.bw, self.cutoff = bw, cutoff
        n = reduced_naive_table_size(bw, cutoff) + reduced_sph_table_size(bw, cutoff)
        self.fwd = np.zeros(n)
        self.inv = np.zeros(n)
        off = (ctypes.c_long * bw)()
        lib().so_tables(bw, cutoff, _p(self.fwd), _p(self.inv), off)
        self.off = np.array(off[:], dtype=np.int64)
        self._pf = (_dp * bw)(*[ctypes.cast(self.fwd.ctypes.data + 8 * int(o), _dp) for o in self.off])
        self._pi = (_dp * bw)(*[ctypes.cast(self.inv.ctypes.data + 8 * int(o), _dp) for o in self.off])

    def order(self, m, inverse=False):
        a = self.inv if inverse else self.fwd
        end = self.off[m + 1] if m + 1 < self.bw else a.size
        return a[self.off[m]:end]


_TCACHE = {}


def tables(bw, cutoff=None):
    key = (bw, bw if cutoff is None else cutoff)
    if key not in _TCACHE:
        _TCACHE[key] = Tables(bw, cutoff)
    return _TCACHE[key]


# ---- per-order transforms ---------------------------------------------------
def dlt_semi(data, bw, m, table, w):
    data, table, w = f64(data), f64(table), f64(w)
    out = np.empty(bw - m)
    lib().so_dlt_semi(_p(data), bw, m, _p(out), _p(table), _p(w))
    return out


def inv_dlt_semi(coeffs, bw, m, ttable, sinv):
    coeffs, ttable, sinv = f64(coeffs), f64(ttable), f64(sinv)
    out = np.empty(2 * bw)
    lib().so_inv_dlt_semi(_p(coeffs), bw, m, _p(out), _p(ttable), _p(sinv))
    return out


def dlt_naive(data, bw, m, w, pml):
    data, w, pml = f64(data), f64(w), f64(pml)
    out = np.empty(bw - m)
    lib().so_dlt_naive(_p(data), bw, m, _p(w), _p(out), _p(pml))
    return out


def inv_dlt_naive(coeffs, bw, m, pml):
    coeffs, pml = f64(coeffs), f64(pml)
    out = np.empty(2 * bw)
    lib().so_inv_dlt_naive(_p(coeffs), bw, m, _p(out), _p(pml))
    return out


# ---- spherical transforms ---------------------------------------------------
def fst(rdata, idata, bw, data_format=0, cutoff=None, tabs=None):
    """FSTSemiMemo (FST_semi_memo.c:68-202). Returns (rcoeffs, icoeffs)."""
    cutoff = bw if cutoff is None else cutoff
    T = tabs or tables(bw, cutoff)
    rdata, idata = f64(rdata).ravel(), f64(idata).ravel()
    rc, ic = np.empty(bw * bw), np.empty(bw * bw)
    w = weights(bw)
    lib().so_fst(_p(rdata), _p(idata), _p(rc), _p(ic), bw, T._pf, _p(w), data_format, cutoff)
    return rc, ic


def inv_fst(rcoeffs, icoeffs, bw, data_format=0, cutoff=None, tabs=None):
    """InvFSTSemiMemo (FST_semi_memo.c:228-351). Returns (rdata, idata)."""
    cutoff = bw if cutoff is None else cutoff
    T = tabs or tables(bw, cutoff)
    rc, ic = f64(rcoeffs), f64(icoeffs)
    n = 2 * bw
    rd, idd = np.empty(n * n), np.empty(n * n)
    lib().so_inv_fst(_p(rc), _p(ic), _p(rd), _p(idd), bw, T._pi, data_format, cutoff)
    return rd, idd


def fzt(rdata, idata, bw, data_format=0):
    """FZTSemiMemo (FST_semi_memo.c:374-407). Returns (rres, ires) of length bw."""
    T = tables(bw, bw)
    rdata, idata = f64(rdata).ravel(), f64(idata).ravel()
    rr, ir = np.empty(bw), np.empty(bw)
    w = weights(bw)
    lib().so_fzt(_p(rdata), _p(idata), _p(rr), _p(ir), bw, _p(f64(T.order(0))), _p(w), data_format)
    return rr, ir


def trans_mult(rdc, idc, rfc, ifc, bw):
    """TransMult (util.c:68-103), keeping its imaginary-part sign."""
    rdc, idc, rfc, ifc = f64(rdc), f64(idc), f64(rfc), f64(ifc)
    rr, ir = np.empty(bw * bw), np.empty(bw * bw)
    lib().so_trans_mult(_p(rdc), _p(idc), _p(rfc), _p(ifc), _p(rr), _p(ir), bw)
    return rr, ir


def conv(rdata, idata, rfilter, ifilter, bw):
    """ConvOn2SphereSemiMemo (FST_semi_memo.c:439-518). Returns (rres, ires)."""
    a = [f64(x).ravel() for x in (rdata, idata, rfilter, ifilter)]
    n = 2 * bw
    rr, ir = np.empty(n * n), np.empty(n * n)
    lib().so_conv(*[_p(x) for x in a], _p(rr), _p(ir), bw)
    return rr, ir


# ---- deterministic inputs ---------------------------------------------------
def uniform(seed, n):
    out = np.empty(n)
    lib().so_uniform(seed, n, _p(out))
    return out


def make_coeffs(seed, bw, kind=0):
    """kind 0: uniform[-1,1]; 1: complex Gaussian, var (l+1)^-2; 2: Hermitian uniform."""
    rc, ic = np.empty(bw * bw), np.empty(bw * bw)
    lib().so_make_coeffs(seed, bw, kind, _p(rc), _p(ic))
    return rc, ic


def degrees(bw):
    """Degree l of every coefficient slot, in IndexOfHarmonicCoeff order."""
    out = np.empty(bw * bw, dtype=np.int64)
    for m in range(-(bw - 1), bw):
        for l in range(abs(m), bw):
            out[index_of_coeff(m, l, bw)] = l
    return out


def orders(bw):
    out = np.empty(bw * bw, dtype=np.int64)
    for m in range(-(bw - 1), bw):
        for l in range(abs(m), bw):
            out[index_of_coeff(m, l, bw)] = m
    return out
