"""s2amd_forward_noalloc (the forward on the library's static scratch array)
against s2amd_forward on the same plan and inputs, through the C ABI.

Both run the same kernels in the same order, so the outputs must be equal
bit for bit.  Calls follow each other with smaller inputs than the one
before (fewer fields, a smaller bandwidth), so every call after the first
finds the previous call's values in the shared array: a kernel that read
scratch it had not written would show here.
"""
import ctypes

import pytest

pytestmark = pytest.mark.gpu

NOALLOC_BYTES = 128 << 20  # S2AMD_NOALLOC_SCRATCH_BYTES


@pytest.fixture(scope="module")
def env():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    import s2kit_amd
    s2kit_amd.lib()
    return s2kit_amd, torch, torch.device("cuda:0")


def _grids(torch, dev, bw, F, seed, real):
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    n = 2 * bw
    re = torch.rand((F, n, n), generator=g, dtype=torch.float64, device=dev) * 2 - 1
    im = torch.zeros_like(re) if real else torch.rand((F, n, n), generator=g, dtype=torch.float64, device=dev) * 2 - 1
    return re, im


def _forward(S, torch, plan, re, im, F, fmt, noalloc):
    L = S.lib()
    bw = plan.bw
    cre = torch.full((F, bw * bw), float("nan"), dtype=torch.float64, device=re.device)
    cim = torch.full_like(cre, float("nan"))
    fn = L.s2amd_forward_noalloc if noalloc else L.s2amd_forward
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = fn(plan.h, ctypes.c_void_p(re.data_ptr()), ctypes.c_void_p(im.data_ptr()),
            ctypes.c_void_p(cre.data_ptr()), ctypes.c_void_p(cim.data_ptr()), F, fmt, stream)
    assert rc == 0, S.last_error()
    return cre, cim


# (precision, bw, fields, data format): each call needs less scratch than the
# one before (96, 32, 24, 18, 12, 6 MiB, ...); FP64 at bw=512 with one field
# takes the per-field path, the other FP64 and FP32 calls the GEMMs
SEQUENCE = [(2, 256, 48, 0), (3, 256, 16, 1), (0, 512, 1, 0), (0, 256, 3, 0), (1, 128, 8, 0), (2, 128, 12, 1),
            (0, 64, 6, 1), (3, 64, 3, 0), (2, 64, 2, 0), (2, 32, 3, 1), (3, 32, 2, 0), (0, 16, 1, 0)]


def test_noalloc_forward_matches_forward_on_successive_calls(env):
    S, torch, dev = env
    for k, (prec, bw, F, fmt) in enumerate(SEQUENCE):
        print(f"noalloc call {k}: precision {prec} bw={bw} fields={F} format={fmt}", flush=True)
        plan = S.Plan(bw, F, precision=prec)
        re, im = _grids(torch, dev, bw, F, 100 + k, fmt == 1)
        a = _forward(S, torch, plan, re, im, F, fmt, noalloc=True)
        b = _forward(S, torch, plan, re, im, F, fmt, noalloc=False)
        torch.cuda.synchronize()
        assert bool(torch.isfinite(a[0]).all()) and bool(torch.isfinite(a[1]).all())
        assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1]), (prec, bw, F, fmt)
        plan.close()


def test_noalloc_forward_repeated_call_on_the_same_plan(env):
    # a second call with fewer fields on the same plan reads only its own fields
    S, torch, dev = env
    plan = S.Plan(128, 32, precision=S.MIXED)
    re, im = _grids(torch, dev, 128, 32, 7, False)
    big = _forward(S, torch, plan, re, im, 32, 0, noalloc=True)
    small = _forward(S, torch, plan, re[5:8].contiguous(), im[5:8].contiguous(), 3, 0, noalloc=True)
    ref = _forward(S, torch, plan, re[5:8].contiguous(), im[5:8].contiguous(), 3, 0, noalloc=False)
    torch.cuda.synchronize()
    assert torch.equal(small[0], ref[0]) and torch.equal(small[1], ref[1])
    assert torch.equal(big[0][5:8], ref[0]) and torch.equal(big[1][5:8], ref[1])
    plan.close()


def test_noalloc_forward_rejects_more_fields_than_the_array_holds(env):
    S, torch, dev = env
    L = S.lib()
    bw = 512
    per_field = 8 * (2 * bw) ** 2  # mixed form: the FP32 intermediate
    fits = NOALLOC_BYTES // per_field
    plan = S.Plan(bw, fits + 1, precision=S.MIXED)
    p = ctypes.c_void_p(16)  # never dereferenced: the size check comes first
    rc = L.s2amd_forward_noalloc(plan.h, p, p, p, p, fits + 1, 0, None)
    assert rc == -1
    msg = S.last_error()
    assert "beyond the 134217728" in msg and f"({per_field} per field)" in msg, msg
    plan.close()
