"""CPU checks of the C ABI: the library loads, exports every declared symbol,
and its host-side integer layout functions and argument checks behave like
the reference's (no GPU work is started here)."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT

HEADERS = [os.path.join(ROOT, "include", "s2kit", h) for h in
           ("FST_semi_memo.h", "FST_semi_fly.h", "cospml.h", "util.h", "weights.h", "pml.h",
            "seminaive.h", "naive.h")] + \
          [os.path.join(ROOT, "include", "s2kit_amd.h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", text, flags=re.M):
            if not m.group(0).lstrip().startswith(("typedef", "#", "return")):
                names.add(m.group(1))
    return names


@pytest.fixture(scope="module")
def S():
    import s2kit_amd
    s2kit_amd.lib()
    return s2kit_amd


def test_every_declared_symbol_is_exported(S):
    names = declared_functions()
    assert {"FSTSemiMemo", "InvFSTSemiMemo", "FZTSemiMemo", "ConvOn2SphereSemiMemo", "FSTSemiFly",
            "InvFSTSemiFly", "FZTSemiFly", "ConvOn2SphereSemiFly", "TableSize",
            "Transpose_RowSize", "s2amd_forward"} <= names
    L = S.lib()
    for n in sorted(names):
        assert hasattr(L, n), f"{n} declared in include/ but not exported"
        assert n in S.SIGNATURES, f"{n} has no binding in s2kit_amd.SIGNATURES"
    assert set(S.SIGNATURES) == names


def test_no_fftw_or_oracle_symbols(S):
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", S.LIBPATH], capture_output=True, text=True).stdout
    assert "fftw_" not in out and " so_" not in out
    und = subprocess.run(["nm", "-D", "--undefined-only", S.LIBPATH], capture_output=True, text=True).stdout
    assert "fftw" not in und and "s2oracle" not in und


@pytest.mark.parametrize("bw", [8, 9, 17, 23, 64, 256, 512])
def test_layout_functions_match_oracle(S, O, bw):
    L = S.lib()
    for m in range(bw):
        assert L.TableSize(m, bw) == O.table_size(m, bw)
        for l in range(m, bw, max(1, bw // 32)):
            assert L.TableOffset(m, l) == O.table_offset(m, l)
            assert L.RowSize(m, l) == O.row_size(m, l)
        for i in range(bw):
            assert L.Transpose_RowSize(i, m, bw) == O.transpose_row_size(i, m, bw)
    for m in range(0, bw + 1, max(1, bw // 16)):
        assert L.Reduced_SpharmonicTableSize(bw, m) == O.reduced_sph_table_size(bw, m)
        assert L.Reduced_Naive_TableSize(bw, m) == O.reduced_naive_table_size(bw, m)
    assert L.Spharmonic_TableSize(bw) == O.sph_table_size(bw)
    for m in range(-(bw - 1), bw, max(1, bw // 16)):
        for l in range(abs(m), bw):
            assert L.IndexOfHarmonicCoeff(m, l, bw) == O.index_of_coeff(m, l, bw)


def test_dropin_rejects_bad_arguments_before_touching_the_device(S):
    L = S.lib()
    S.set_abort_on_error(False)
    try:
        z = np.zeros(16)
        p = z.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        L.FSTSemiMemo(p, p, p, p, 0, None, None, 0, 0, None, None, p)
        assert "bw must be in [2, 1024]" in S.last_error()
        L.InvFSTSemiMemo(p, p, p, p, 8, None, None, 0, 8, None, None)
        assert "transpose_seminaive_naive_table is NULL" in S.last_error()
        L.FSTSemiMemo(p, p, p, p, 8, None, None, 3, 8, None, None, p)
        assert "data_format" in S.last_error()
        L.FSTSemiMemo(p, p, p, p, 8, None, None, 0, 9, None, None, p)
        assert "cutoff" in S.last_error()
        L.ConvOn2SphereSemiMemo(p, p, p, None, p, p, 8, None)
        assert "NULL" in S.last_error()
        L.GenerateCosPmlTable(8, 8, p, None)
        assert "m must be in [0, bw)" in S.last_error()
    finally:
        S.set_abort_on_error(True)


def test_plan_rejects_bad_arguments(S):
    L = S.lib()
    h = ctypes.c_void_p()
    assert L.s2amd_plan_create(ctypes.byref(h), 1, 1, 4, 0) == -1
    assert "bw" in S.last_error()
    assert L.s2amd_plan_create(ctypes.byref(h), 16, 17, 4, 0) == -1
    assert L.s2amd_plan_create(ctypes.byref(h), 16, 16, 0, 0) == -1
    assert L.s2amd_plan_create(ctypes.byref(h), 16, 16, 4, 9) == -1
    assert "precision" in S.last_error()
    # mixed: any bw in [32, 512]; split BF16: powers of two in [32, 512]
    for bw, prec, name in ((16, S.MIXED, "mixed"), (600, S.MIXED, "mixed"), (96, S.SPLIT_BF16, "split-BF16"),
                           (16, S.SPLIT_BF16, "split-BF16")):
        assert L.s2amd_plan_create(ctypes.byref(h), bw, bw, 4, prec) == -1
        assert name in S.last_error()
    assert L.s2amd_forward(None, None, None, None, None, 1, 0, None) == -1
    assert "plan is NULL" in S.last_error()
    assert L.s2amd_plan_destroy(None) == 0


def test_conv_batch_rejects_bad_arguments_before_touching_the_device(S):
    L = S.lib()
    h = ctypes.c_void_p()
    req = (S.ConvRequest * 1)()
    req[0] = S.ConvRequest(8, 256, 256, 256, 256, 256, 256)  # aligned dummy addresses, never dereferenced
    p = ctypes.cast(req, ctypes.c_void_p)
    assert L.s2amd_conv_batch_create(ctypes.byref(h), p, 0, 0, None) == -1
    assert "at least one request" in S.last_error()
    assert L.s2amd_conv_batch_create(ctypes.byref(h), p, 1, 1, None) == -1
    assert "precision" in S.last_error()
    req[0].bw = 2000
    assert L.s2amd_conv_batch_create(ctypes.byref(h), p, 1, 0, None) == -1
    assert "bw must be in" in S.last_error()
    req[0].bw = 8
    req[0].ore = None
    assert L.s2amd_conv_batch_create(ctypes.byref(h), p, 1, 0, None) == -1
    assert "NULL" in S.last_error()
    req[0].ore = 260
    assert L.s2amd_conv_batch_create(ctypes.byref(h), p, 1, 0, None) == -1
    assert "aligned" in S.last_error()
    assert L.s2amd_conv_batch_run(None, None) == -1
    assert L.s2amd_conv_batch_destroy(None) == 0


def _unbound_plan(S, bw=8, max_fields=4):
    """A Plan object with no library handle: enough to run the wrappers'
    argument checks, which must all fire before any library call."""
    p = S.Plan.__new__(S.Plan)
    p.h, p.bw, p.cutoff, p.max_fields, p.precision, p.device = None, bw, bw, max_fields, S.FP64, 0
    return p


def test_plan_wrappers_reject_tensors_the_abi_cannot_check(S):
    import torch
    p = _unbound_plan(S)
    g = lambda *shape, **kw: torch.zeros(*shape, dtype=kw.get("dtype", torch.float64))
    grid, coef = g(2, 16, 16), g(2, 64)
    cases = [
        ((grid.float(), grid, coef, coef), "dtype must be torch.float64"),
        ((g(2, 16, 32)[:, :, ::2], grid, coef, coef), "contiguous"),
        ((grid, grid, g(1, 64), coef), "holds 64 values, the call needs 128"),
        ((grid, grid, coef, coef), "must be a CUDA tensor"),  # CPU tensors: never passed as pointers
        (([1.0], grid, coef, coef), "expected a torch tensor"),
    ]
    for args, msg in cases:
        with pytest.raises(S.S2Error, match=msg):
            p.forward(*args)
    with pytest.raises(S.S2Error, match=r"nfields must be in \[0, 4\]"):
        p.forward(grid, grid, coef, coef, nfields=5)
    with pytest.raises(S.S2Error, match="cre: holds 64 values"):
        p.inverse(g(1, 64), coef, grid, grid, nfields=2)
    with pytest.raises(S.S2Error, match="zr: holds 8 values"):
        p.zonal(grid, grid, g(8), g(16))
    with pytest.raises(S.S2Error, match="oim: dtype"):
        p.conv(grid, grid, grid, grid, grid, grid.float())
    S.Plan.close(p)  # no handle: nothing to free


def test_conv_batch_wrapper_rejects_bad_tensors(S, monkeypatch):
    import torch
    monkeypatch.setattr(torch.cuda, "current_device", lambda: 0)
    t = torch.zeros(16, 16, dtype=torch.float64)
    with pytest.raises(S.S2Error, match="request 0 sre: must be a CUDA tensor"):
        S.ConvBatch([(8, t, t, t, t, t, t)])
    with pytest.raises(S.S2Error, match="request 0 hre: holds 64 values, the call needs 256"):
        S.ConvBatch([(8, t, t, torch.zeros(64, dtype=torch.float64), t, t, t)])
    with pytest.raises(S.S2Error, match="expected"):
        S.ConvBatch([(8, t, t, t)])
