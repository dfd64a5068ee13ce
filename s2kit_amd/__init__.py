"""s2kit_amd -- MI355X-native memoised seminaive spherical harmonic transform.

Python host side over libs2kit_amd.so (HIP kernels for gfx950 behind a C ABI).

* ``FSTSemiMemo`` & co. are thin ctypes bindings of the library's S2kit
  drop-ins: same arguments, same host-memory layouts and table formats as
  S2kit 1.1's FST_semi_memo.h / cospml.h / util.h / weights.h.
* ``Plan`` binds the batched device API (s2kit_amd.h): many fields per call,
  device-resident data (torch tensors are used only as device allocations).

Nothing here computes on the CPU: if the library is missing the import of a
binding raises, and every transform runs on the GPU.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libs2kit_amd.so")

COMPLEX, REAL = 0, 1
FP64, FP32, MIXED, SPLIT_BF16 = 0, 1, 2, 3
NSTAGES = 10

_dp = ctypes.POINTER(ctypes.c_double)
_dpp = ctypes.POINTER(_dp)
_i = ctypes.c_int
_vp = ctypes.c_void_p

# every symbol the headers under include/ declare, with (restype, argtypes)
SIGNATURES = {
    # util.h
    "IndexOfHarmonicCoeff": (_i, [_i, _i, _i]),
    "TransMult": (None, [_dp, _dp, _dp, _dp, _dp, _dp, _i]),
    # weights.h
    "GenerateWeightsForDLT": (None, [_i, _dp]),
    # pml.h
    "GeneratePmlTable": (None, [_i, _i, _dp, _dp]),
    # cospml.h
    "TableSize": (_i, [_i, _i]),
    "Spharmonic_TableSize": (_i, [_i]),
    "Reduced_SpharmonicTableSize": (_i, [_i, _i]),
    "Reduced_Naive_TableSize": (_i, [_i, _i]),
    "TableOffset": (_i, [_i, _i]),
    "RowSize": (_i, [_i, _i]),
    "Transpose_RowSize": (_i, [_i, _i, _i]),
    "GenerateCosPmlTable": (None, [_i, _i, _dp, _dp]),
    "TransposeCosPmlTable": (None, [_i, _i, _dp, _dp]),
    "Spharmonic_Pml_Table": (_dpp, [_i, _dp, _dp]),
    "Transpose_Spharmonic_Pml_Table": (_dpp, [_dpp, _i, _dp]),
    "SemiNaive_Naive_Pml_Table": (_dpp, [_i, _i, _dp, _dp]),
    "Transpose_SemiNaive_Naive_Pml_Table": (_dpp, [_dpp, _i, _i, _dp, _dp]),
    # FST_semi_memo.h
    "FSTSemiMemo": (None, [_dp, _dp, _dp, _dp, _i, _dpp, _dp, _i, _i, _vp, _vp, _dp]),
    "InvFSTSemiMemo": (None, [_dp, _dp, _dp, _dp, _i, _dpp, _dp, _i, _i, _vp, _vp]),
    "FZTSemiMemo": (None, [_dp, _dp, _dp, _dp, _i, _dp, _dp, _i, _vp, _dp]),
    "ConvOn2SphereSemiMemo": (None, [_dp, _dp, _dp, _dp, _dp, _dp, _i, _dp]),
    # seminaive.h, naive.h
    "DLTSemi": (None, [_dp, _i, _i, _dp, _dp, _dp, _dp, _vp]),
    "InvDLTSemi": (None, [_dp, _i, _i, _dp, _dp, _dp, _dp, _vp]),
    "DLTNaive": (None, [_dp, _i, _i, _dp, _dp, _dp, _dp]),
    "InvDLTNaive": (None, [_dp, _i, _i, _dp, _dp]),
    # FST_semi_fly.h
    "FSTSemiFly": (None, [_dp, _dp, _dp, _dp, _i, _dp, _i, _i, _vp, _vp, _dp]),
    "InvFSTSemiFly": (None, [_dp, _dp, _dp, _dp, _i, _dp, _i, _i, _vp, _vp]),
    "FZTSemiFly": (None, [_dp, _dp, _dp, _dp, _i, _dp, _i, _vp, _dp]),
    "ConvOn2SphereSemiFly": (None, [_dp, _dp, _dp, _dp, _dp, _dp, _i, _dp]),
    # s2kit_amd.h
    "s2amd_plan_create": (_i, [ctypes.POINTER(_vp), _i, _i, _i, _i]),
    "s2amd_plan_destroy": (_i, [_vp]),
    "s2amd_plan_bytes": (ctypes.c_size_t, [_vp]),
    "s2amd_plan_tables": (_i, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(ctypes.c_long)]),
    "s2amd_forward": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _i, _vp]),
    "s2amd_forward_noalloc": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _i, _vp]),
    "s2amd_inverse": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _i, _vp]),
    "s2amd_zonal": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _i, _vp]),
    "s2amd_conv": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _vp]),
    "s2amd_conv_batch_create": (_i, [ctypes.POINTER(_vp), _vp, _i, _i, _vp]),
    "s2amd_conv_batch_run": (_i, [_vp, _vp]),
    "s2amd_conv_batch_info": (_i, [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i),
                                   ctypes.POINTER(_i)]),
    "s2amd_conv_batch_destroy": (_i, [_vp]),
    "s2amd_plan_timing": (_i, [_vp, _i]),
    "s2amd_plan_autotune": (_i, [_vp, _i, _vp]),
    "s2amd_plan_tuning": (_i, [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "s2amd_plan_stage_times": (_i, [_vp, _dp, ctypes.POINTER(ctypes.c_long)]),
    "s2amd_stage_name": (ctypes.c_char_p, [_i]),
    "s2amd_set_abort_on_error": (None, [_i]),
    "s2amd_last_error": (ctypes.c_char_p, []),
}

_LIB = None


def lib():
    """Load libs2kit_amd.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"{LIBPATH} is missing: build it with `make -C s2kit_amd/csrc` "
                               "or __graft_entry__.build()")
        L = ctypes.CDLL(LIBPATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


class S2Error(RuntimeError):
    pass


def last_error():
    return lib().s2amd_last_error().decode()


def set_abort_on_error(on):
    lib().s2amd_set_abort_on_error(1 if on else 0)


def _p(a):
    if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous):
        raise TypeError("expected a C-contiguous float64 numpy array")
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---------------------------------------------------------------------------
# S2kit drop-ins (host arrays)
# ---------------------------------------------------------------------------
class HostTables:
    """Tables produced by the drop-in builders into caller (numpy) memory,
    with the malloc'd pointer arrays the builders return."""

    def __init__(self, bw, cutoff=None):
        L = lib()
        cutoff = bw if cutoff is None else cutoff
        self.bw, self.cutoff = bw, cutoff
        n = L.Reduced_Naive_TableSize(bw, cutoff) + L.Reduced_SpharmonicTableSize(bw, cutoff)
        self.fwd = np.zeros(max(n, 1))
        self.inv = np.zeros(max(n, 1))
        ws = np.zeros(16 * bw)
        if cutoff == bw:
            self.pf = L.Spharmonic_Pml_Table(bw, _p(self.fwd), _p(ws))
            self.pi = L.Transpose_Spharmonic_Pml_Table(self.pf, bw, _p(self.inv))
        else:
            self.pf = L.SemiNaive_Naive_Pml_Table(bw, cutoff, _p(self.fwd), _p(ws))
            self.pi = L.Transpose_SemiNaive_Naive_Pml_Table(self.pf, bw, cutoff, _p(self.inv), _p(ws))
        if not self.pf or not self.pi:
            raise S2Error(last_error())

    def offsets(self):
        base = self.fwd.ctypes.data
        return np.array([ctypes.cast(self.pf[m], _vp).value - base for m in range(self.bw)]) // 8

    def __del__(self):
        libc = ctypes.CDLL(None)
        libc.free.argtypes = [_vp]
        for p in (getattr(self, "pf", None), getattr(self, "pi", None)):
            if p:
                libc.free(ctypes.cast(p, _vp))


def weights(bw):
    w = np.empty(4 * bw)
    lib().GenerateWeightsForDLT(bw, _p(w))
    return w


def fst(rdata, idata, bw, tables, w=None, data_format=COMPLEX, cutoff=None):
    """FSTSemiMemo on host arrays; returns (rcoeffs, icoeffs)."""
    cutoff = tables.cutoff if cutoff is None else cutoff
    rdata, idata = _f64(rdata).ravel(), _f64(idata).ravel()
    w = weights(bw) if w is None else _f64(w)
    rc, ic = np.empty(bw * bw), np.empty(bw * bw)
    lib().FSTSemiMemo(_p(rdata), _p(idata), _p(rc), _p(ic), bw, tables.pf, None, data_format, cutoff, None, None,
                      _p(w))
    return rc, ic


def inv_fst(rcoeffs, icoeffs, bw, tables, data_format=COMPLEX, cutoff=None):
    """InvFSTSemiMemo on host arrays; returns (rdata, idata)."""
    cutoff = tables.cutoff if cutoff is None else cutoff
    rc, ic = _f64(rcoeffs), _f64(icoeffs)
    n = 2 * bw
    rd, idd = np.empty(n * n), np.empty(n * n)
    lib().InvFSTSemiMemo(_p(rc), _p(ic), _p(rd), _p(idd), bw, tables.pi, None, data_format, cutoff, None, None)
    return rd, idd


def fzt(rdata, idata, bw, tables, w=None, data_format=COMPLEX):
    """FZTSemiMemo on host arrays; returns (rres, ires) of length bw."""
    rdata, idata = _f64(rdata).ravel(), _f64(idata).ravel()
    w = weights(bw) if w is None else _f64(w)
    rr, ir = np.empty(bw), np.empty(bw)
    lib().FZTSemiMemo(_p(rdata), _p(idata), _p(rr), _p(ir), bw, tables.pf[0], None, data_format, None, _p(w))
    return rr, ir


def conv(rdata, idata, rfilter, ifilter, bw, fly=False):
    """ConvOn2SphereSemiMemo (or ConvOn2SphereSemiFly) on host arrays; returns (rres, ires)."""
    a = [_f64(x).ravel() for x in (rdata, idata, rfilter, ifilter)]
    n = 2 * bw
    rr, ir = np.empty(n * n), np.empty(n * n)
    fn = lib().ConvOn2SphereSemiFly if fly else lib().ConvOn2SphereSemiMemo
    fn(*[_p(x) for x in a], _p(rr), _p(ir), bw, None)
    return rr, ir


def dlt_semi(data, bw, m, table, w=None):
    """DLTSemi of one order on host arrays; table from GenerateCosPmlTable(bw, m)."""
    data, table = _f64(data), _f64(table)
    w = weights(bw) if w is None else _f64(w)
    out = np.empty(bw - m)
    lib().DLTSemi(_p(data), bw, m, _p(out), None, _p(table), _p(w), None)
    return out


def inv_dlt_semi(coeffs, bw, m, ttable, sin_values):
    """InvDLTSemi of one order on host arrays; ttable from TransposeCosPmlTable."""
    coeffs, ttable, sin_values = _f64(coeffs), _f64(ttable), _f64(sin_values)
    out = np.empty(2 * bw)
    lib().InvDLTSemi(_p(coeffs), bw, m, _p(out), _p(ttable), _p(sin_values), None, None)
    return out


def dlt_naive(data, bw, m, pml, w=None):
    """DLTNaive of one order on host arrays; pml from GeneratePmlTable(bw, m)."""
    data, pml = _f64(data), _f64(pml)
    w = weights(bw) if w is None else _f64(w)
    out = np.empty(bw - m)
    lib().DLTNaive(_p(data), bw, m, _p(w), _p(out), _p(pml), None)
    return out


def inv_dlt_naive(coeffs, bw, m, pml):
    """InvDLTNaive of one order on host arrays."""
    coeffs, pml = _f64(coeffs), _f64(pml)
    out = np.empty(2 * bw)
    lib().InvDLTNaive(_p(coeffs), bw, m, _p(out), _p(pml))
    return out


def fst_fly(rdata, idata, bw, cutoff, w=None, data_format=COMPLEX):
    """FSTSemiFly on host arrays (the library makes the tables); returns (rcoeffs, icoeffs)."""
    rdata, idata = _f64(rdata).ravel(), _f64(idata).ravel()
    w = weights(bw) if w is None else _f64(w)
    rc, ic = np.empty(bw * bw), np.empty(bw * bw)
    lib().FSTSemiFly(_p(rdata), _p(idata), _p(rc), _p(ic), bw, None, data_format, cutoff, None, None, _p(w))
    return rc, ic


def inv_fst_fly(rcoeffs, icoeffs, bw, cutoff, data_format=COMPLEX):
    """InvFSTSemiFly on host arrays; returns (rdata, idata)."""
    rc, ic = _f64(rcoeffs), _f64(icoeffs)
    n = 2 * bw
    rd, idd = np.empty(n * n), np.empty(n * n)
    lib().InvFSTSemiFly(_p(rc), _p(ic), _p(rd), _p(idd), bw, None, data_format, cutoff, None, None)
    return rd, idd


def fzt_fly(rdata, idata, bw, w=None, data_format=COMPLEX):
    """FZTSemiFly on host arrays; returns (rres, ires) of length bw."""
    rdata, idata = _f64(rdata).ravel(), _f64(idata).ravel()
    w = weights(bw) if w is None else _f64(w)
    rr, ir = np.empty(bw), np.empty(bw)
    lib().FZTSemiFly(_p(rdata), _p(idata), _p(rr), _p(ir), bw, None, data_format, None, _p(w))
    return rr, ir


def trans_mult(rdc, idc, rfc, ifc, bw):
    a = [_f64(x) for x in (rdc, idc, rfc, ifc)]
    rr, ir = np.empty(bw * bw), np.empty(bw * bw)
    lib().TransMult(*[_p(x) for x in a], _p(rr), _p(ir), bw)
    return rr, ir


def cospml_table(bw, m):
    t = np.empty(lib().TableSize(m, bw))
    lib().GenerateCosPmlTable(bw, m, _p(t), None)
    return t


def transpose_cospml_table(bw, m, t):
    t = _f64(t)
    tt = np.empty(lib().TableSize(m, bw))
    lib().TransposeCosPmlTable(bw, m, _p(t), _p(tt))
    return tt


def pml_table(bw, m):
    t = np.empty(2 * bw * (bw - m))
    lib().GeneratePmlTable(bw, m, _p(t), None)
    return t


# ---------------------------------------------------------------------------
# batched device API
# ---------------------------------------------------------------------------
def _ptr(t):
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def _check_tensors(args, device):
    """Reject what the C ABI cannot see through a raw pointer.  `args` holds
    (name, tensor, values needed) triples: every tensor must be a float64,
    contiguous torch tensor holding at least the values needed (checked for
    all of them first), and live on CUDA device index `device`."""
    import torch
    for name, t, need in args:
        if not isinstance(t, torch.Tensor):
            raise S2Error(f"{name}: expected a torch tensor, got {type(t).__name__}")
        if t.dtype != torch.float64:
            raise S2Error(f"{name}: dtype must be torch.float64, got {t.dtype}")
        if not t.is_contiguous():
            raise S2Error(f"{name}: tensor must be contiguous")
        if t.numel() < need:
            raise S2Error(f"{name}: holds {t.numel()} values, the call needs {need}")
    for name, t, _ in args:
        if not t.is_cuda or t.device.index != device:
            raise S2Error(f"{name}: must be a CUDA tensor on cuda:{device}, got {t.device}")


class Plan:
    """Batched transforms of up to ``max_fields`` fields at bandwidth ``bw``
    on device memory.  Arguments are torch float64 CUDA tensors laid out per
    field like the drop-ins: grids [F, 2bw, 2bw], coefficients [F, bw*bw]."""

    def __init__(self, bw, max_fields, precision=FP64, cutoff=None):
        import torch
        cutoff = bw if cutoff is None else cutoff
        self.device = torch.cuda.current_device()
        h = ctypes.c_void_p()
        rc = lib().s2amd_plan_create(ctypes.byref(h), bw, cutoff, max_fields, precision)
        if rc != 0:
            raise S2Error(last_error())
        self.h, self.bw, self.cutoff, self.max_fields, self.precision = h, bw, cutoff, max_fields, precision

    def close(self):
        if getattr(self, "h", None):
            lib().s2amd_plan_destroy(self.h)
            self.h = None

    __del__ = close

    def nbytes(self):
        return lib().s2amd_plan_bytes(self.h)

    def autotune(self, nfields, stream=None):
        """Time the GEMM tile shapes at nfields fields and keep the fastest."""
        self._check(lib().s2amd_plan_autotune(self.h, nfields, self._stream(stream)))

    def tuning(self):
        """(columns per block, column tiles per table load, ring groups); 0: heuristic."""
        v = [ctypes.c_int() for _ in range(3)]
        self._check(lib().s2amd_plan_tuning(self.h, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    @staticmethod
    def _stream(stream):
        import torch
        s = torch.cuda.current_stream() if stream is None else stream
        return ctypes.c_void_p(s.cuda_stream)

    def set_timing(self, on):
        self._check(lib().s2amd_plan_timing(self.h, 1 if on else 0))

    def stage_times(self):
        """{stage: (total_ms, launches)} since the last call; waits for the events."""
        n = NSTAGES
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_long * n)()
        self._check(lib().s2amd_plan_stage_times(self.h, ms, cnt))
        return {lib().s2amd_stage_name(i).decode(): (ms[i], cnt[i]) for i in range(n) if cnt[i]}

    def _check(self, rc):
        if rc != 0:
            raise S2Error(last_error())

    def _fields(self, nfields, args):
        """Validate (name, tensor, values per field) triples for a call of
        nfields fields (default: the first tensor's leading dimension)."""
        first = args[0][1]
        if nfields is not None:
            n = int(nfields)
        elif hasattr(first, "dim") and first.dim() > 0:
            n = first.shape[0]
        else:
            n = 1
        if n < 0 or n > self.max_fields:
            raise S2Error(f"nfields must be in [0, {self.max_fields}], got {n}")
        _check_tensors([(name, t, n * per) for name, t, per in args], self.device)
        return n

    def forward(self, re, im, cre, cim, nfields=None, data_format=COMPLEX, stream=None):
        g, c = 4 * self.bw * self.bw, self.bw * self.bw
        n = self._fields(nfields, [("re", re, g), ("im", im, g), ("cre", cre, c), ("cim", cim, c)])
        self._check(lib().s2amd_forward(self.h, _ptr(re), _ptr(im), _ptr(cre), _ptr(cim), n, data_format,
                                        self._stream(stream)))

    def inverse(self, cre, cim, re, im, nfields=None, data_format=COMPLEX, stream=None):
        g, c = 4 * self.bw * self.bw, self.bw * self.bw
        n = self._fields(nfields, [("cre", cre, c), ("cim", cim, c), ("re", re, g), ("im", im, g)])
        self._check(lib().s2amd_inverse(self.h, _ptr(cre), _ptr(cim), _ptr(re), _ptr(im), n, data_format,
                                        self._stream(stream)))

    def zonal(self, re, im, zr, zi, nfields=None, data_format=COMPLEX, stream=None):
        g = 4 * self.bw * self.bw
        n = self._fields(nfields, [("re", re, g), ("im", im, g), ("zr", zr, self.bw), ("zi", zi, self.bw)])
        self._check(lib().s2amd_zonal(self.h, _ptr(re), _ptr(im), _ptr(zr), _ptr(zi), n, data_format,
                                      self._stream(stream)))

    def conv(self, sre, sim, hre, him, ore, oim, nfields=None, stream=None):
        g = 4 * self.bw * self.bw
        n = self._fields(nfields, [("sre", sre, g), ("sim", sim, g), ("hre", hre, g), ("him", him, g),
                                   ("ore", ore, g), ("oim", oim, g)])
        self._check(lib().s2amd_conv(self.h, _ptr(sre), _ptr(sim), _ptr(hre), _ptr(him), _ptr(ore), _ptr(oim), n,
                                     self._stream(stream)))


def _on(stream):
    """torch's elementwise work on the plan calls' stream."""
    import contextlib
    import torch
    return torch.cuda.stream(stream) if stream is not None else contextlib.nullcontext()


def forward_refined(lo, p64, re, im, cre, cim, nfields=None, data_format=COMPLEX, stream=None):
    """Forward transform by plan `lo` (mixed or split BF16) with one FP64
    residual refinement through plan `p64` (same bw, FP64):
        c0 = lo.forward(x);  c = c0 + lo.forward(x - p64.inverse(c0)).
    The residual is ~1e-7 of x, so its low-precision transform adds ~1e-14:
    FP64 accuracy, at the cost of an FP64 inverse plus two low-precision
    forwards (more than one FP64 forward).  Only the forward (analysis)
    admits this: the FP32 errors of a low-precision synthesis lie partly
    outside the band-limited space, where a coefficient-space residual
    cannot see them."""
    import torch
    with _on(stream):
        lo.forward(re, im, cre, cim, nfields, data_format, stream)
        gr, gi = torch.empty_like(re), torch.empty_like(im)
        p64.inverse(cre, cim, gr, gi, nfields, data_format, stream)
        gr.neg_().add_(re)
        gi.neg_().add_(im)
        dr, di = torch.empty_like(cre), torch.empty_like(cim)
        lo.forward(gr, gi, dr, di, nfields, data_format, stream)
        cre.add_(dr)
        cim.add_(di)


class ConvRequest(ctypes.Structure):
    """s2amd_conv_request: one ConvOn2SphereSemiMemo on device grids [2bw, 2bw]."""
    _fields_ = [("bw", ctypes.c_int), ("sre", ctypes.c_void_p), ("sim", ctypes.c_void_p),
                ("hre", ctypes.c_void_p), ("him", ctypes.c_void_p), ("ore", ctypes.c_void_p),
                ("oim", ctypes.c_void_p)]


class ConvBatch:
    """Convolutions of uneven bandwidth recorded once as a HIP graph.

    ``requests`` is a list of (bw, sre, sim, hre, him, ore, oim) with torch
    float64 CUDA tensors of shape [2bw, 2bw] (or any tensors holding 4bw^2
    values).  The tensors must stay alive while the batch is used; their
    contents are read at every run()."""

    def __init__(self, requests, precision=MIXED):
        import torch
        dev = torch.cuda.current_device()
        arr = (ConvRequest * len(requests))()
        for k, (bw, *ts) in enumerate(requests):
            if len(ts) != 6:
                raise S2Error(f"request {k}: expected (bw, sre, sim, hre, him, ore, oim)")
            _check_tensors([(f"request {k} {name}", t, 4 * bw * bw)
                            for name, t in zip(("sre", "sim", "hre", "him", "ore", "oim"), ts)], dev)
            arr[k] = ConvRequest(bw, *[t.data_ptr() for t in ts])
        h = ctypes.c_void_p()
        rc = lib().s2amd_conv_batch_create(ctypes.byref(h), ctypes.cast(arr, ctypes.c_void_p), len(requests),
                                           precision, None)
        if rc != 0:
            raise S2Error(last_error())
        self.h = h
        self._keep = [r[1:] for r in requests]

    def close(self):
        if getattr(self, "h", None):
            lib().s2amd_conv_batch_destroy(self.h)
            self.h = None

    __del__ = close

    def run(self, stream=None):
        rc = lib().s2amd_conv_batch_run(self.h, Plan._stream(stream))
        if rc != 0:
            raise S2Error(last_error())

    def groups(self):
        """[(bw, count, precision)] in ascending bandwidth."""
        n = ctypes.c_int()
        lib().s2amd_conv_batch_info(self.h, ctypes.byref(n), None, None, None)
        bws, cnt, prec = (ctypes.c_int * n.value)(), (ctypes.c_int * n.value)(), (ctypes.c_int * n.value)()
        rc = lib().s2amd_conv_batch_info(self.h, ctypes.byref(n), bws, cnt, prec)
        if rc != 0:
            raise S2Error(last_error())
        return [(bws[i], cnt[i], prec[i]) for i in range(n.value)]
