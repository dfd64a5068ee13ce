"""Convolution batches of uneven bandwidth (s2amd_conv_batch_*): every
request against the CPU oracle's ConvOn2SphereSemiMemo
(FST_semi_memo.c:439-518), for both request layouts (separate buffers, which
go through the gather/scatter kernels, and back-to-back buffers, used in
place), and across graph replays with eager work in between.

FP64 batches: 1e-11 relative (normwise) per request.  Mixed and split-BF16
batches: 1e-6 for bandwidths those forms support (powers of two from bw =
32); other bandwidths run in FP64 and keep the FP64 bound.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    import s2kit_amd
    s2kit_amd.lib()
    return s2kit_amd, torch, torch.device("cuda:0")


def crel(ar, ai, br, bi):
    a = np.asarray(ar) + 1j * np.asarray(ai)
    b = np.asarray(br) + 1j * np.asarray(bi)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def zonal_filter(O, bw, sigma, rng):
    """Real zonal grid of h_l = exp(-l(l+1)/(2 sigma^2)) * U[0.5, 1]."""
    l = np.arange(bw)
    hc = np.zeros(bw * bw)
    hc[:bw] = np.exp(-l * (l + 1) / (2 * sigma ** 2)) * rng.uniform(0.5, 1.0, bw)
    return O.inv_fst(hc, np.zeros(bw * bw), bw)


def make_requests(S, torch, dev, O, bws, seed, contiguous_bw=None):
    """Requests (bw, sre, sim, hre, him, ore, oim) and their oracle outputs.
    Requests at contiguous_bw are views into one tensor per array, back to
    back in request order; the others are separate tensors."""
    rng = np.random.default_rng(seed)
    host = []
    for k, bw in enumerate(bws):
        s = O.inv_fst(*O.make_coeffs(seed * 100 + k, bw, 0), bw)
        h = zonal_filter(O, bw, 5 + rng.uniform() * (bw / 2 - 5), rng)
        host.append((bw, s, h, O.conv(s[0], s[1], h[0], h[1], bw)))
    reqs = [None] * len(bws)
    ks = [k for k, bw in enumerate(bws) if bw == contiguous_bw]
    if ks:
        n = 2 * contiguous_bw
        big = [torch.empty((len(ks), n, n), dtype=torch.float64, device=dev) for _ in range(6)]
        for j, k in enumerate(ks):
            _, s, h, _ = host[k]
            for a, v in enumerate((s[0], s[1], h[0], h[1])):
                big[a][j].copy_(torch.tensor(v.reshape(n, n), device=dev))
            reqs[k] = (contiguous_bw, *[t[j] for t in big])
    for k, (bw, s, h, _) in enumerate(host):
        if reqs[k] is not None:
            continue
        n = 2 * bw
        ts = [torch.tensor(v.reshape(n, n), dtype=torch.float64, device=dev) for v in (s[0], s[1], h[0], h[1])]
        reqs[k] = (bw, *ts, torch.empty((n, n), dtype=torch.float64, device=dev),
                   torch.empty((n, n), dtype=torch.float64, device=dev))
    expect = [h[3] for h in host]
    return reqs, expect


def check(reqs, expect, tol_of):
    worst = 0.0
    for (bw, sre, sim, hre, him, ore, oim), (er, ei) in zip(reqs, expect):
        e = crel(ore.cpu().numpy().ravel(), oim.cpu().numpy().ravel(), er, ei)
        assert e < tol_of(bw), (bw, e)
        worst = max(worst, e)
    return worst


@pytest.mark.parametrize("prec", [0, 2, 3])
def test_conv_batch_matches_oracle(env, O, prec):
    S, torch, dev = env
    bws = [32, 64, 128, 64, 32, 128, 64, 16]
    reqs, expect = make_requests(S, torch, dev, O, bws, seed=5 + prec, contiguous_bw=64)
    batch = S.ConvBatch(reqs, precision=prec)
    groups = batch.groups()
    assert [g[0] for g in groups] == [16, 32, 64, 128]
    assert [g[1] for g in groups] == [1, 2, 3, 2]
    # the mixed and split-BF16 forms cover bw >= 32; other groups run in FP64
    low = {0: 1 << 30, 2: 32, 3: 32}[prec]
    assert [g[2] for g in groups] == [prec if bw >= low else 0 for bw in (16, 32, 64, 128)]

    def tol(bw):
        return 1e-6 if bw >= low else 1e-11

    batch.run()
    torch.cuda.synchronize()
    check(reqs, expect, tol)
    # replay after eager work on the outputs: the graph rewrites every output
    for r in reqs:
        r[5].zero_()
        r[6].fill_(7.0)
    torch.cuda.synchronize()
    batch.run()
    torch.cuda.synchronize()
    check(reqs, expect, tol)
    # inputs are read at every replay: scaling the signals scales the outputs
    for r in reqs:
        r[1].mul_(2.0)
        r[2].mul_(2.0)
    torch.cuda.synchronize()
    batch.run()
    torch.cuda.synchronize()
    check(reqs, [(2 * er, 2 * ei) for er, ei in expect], tol)
    batch.close()


@pytest.mark.parametrize("prec", [0, 2])
def test_conv_batch_bandwidths_not_powers_of_two(env, O, prec):
    # groups of 3 and 4 requests at bw 24 and 50 run the FP64 GEMM path on
    # mixed-radix FFTs (48 = 16*3, 100 = 4*25 rings), the single bw=45 request
    # the per-field path; mixed batches run bw 46 and 50 in the mixed form
    # (mixed-radix FP32 FFTs) and bw 24 in FP64.  (An even bw for the mixed
    # batch: at odd bw the mixed form synthesises degree bw-1 whole, where the
    # reference's convolution keeps its short Transpose_RowSize rows.)
    S, torch, dev = env
    odd = 45 if prec == 0 else 46
    bws = [24, 50, 24, odd, 24, 50, 50, 50]
    reqs, expect = make_requests(S, torch, dev, O, bws, seed=11 + prec, contiguous_bw=50)
    batch = S.ConvBatch(reqs, precision=prec)
    groups = batch.groups()
    assert [g[0] for g in groups] == [24, odd, 50]
    assert [g[1] for g in groups] == [3, 1, 4]
    assert [g[2] for g in groups] == ([0, 0, 0] if prec == 0 else [0, prec, prec])
    for _ in range(2):  # the recorded graph, twice
        batch.run()
        torch.cuda.synchronize()
        check(reqs, expect, lambda bw: 1e-11 if (prec == 0 or bw < 32) else 1e-6)
        for r in reqs:
            r[5].zero_()
        torch.cuda.synchronize()
    batch.close()


def test_conv_batch_argument_errors(env):
    S, torch, dev = env
    t = torch.zeros((8, 8), dtype=torch.float64, device=dev)
    with pytest.raises(S.S2Error, match="precision"):
        S.ConvBatch([(4, t, t, t, t, t, t)], precision=1)
    with pytest.raises(S.S2Error, match="bw"):
        S.ConvBatch([(1, t, t, t, t, t, t)], precision=0)
    odd = torch.zeros(65, dtype=torch.float64, device=dev)[1:].reshape(8, 8)
    with pytest.raises(S.S2Error, match="aligned"):
        S.ConvBatch([(4, odd, t, t, t, t, t)], precision=0)
