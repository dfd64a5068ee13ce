"""GPU parity of the batched device API (s2kit_amd.h) against the CPU
oracle, plus size-independent properties at the full B=256 shape.

FP64 plans: 1e-11 relative (normwise) against the oracle, per field.
FP32 seminaive plans: 1e-4 (their measured error is 1e-6 .. 4e-5: the
cosine-space projections cancel, so FP32 rounding of the cosine
coefficients and FP32 accumulation are amplified).
Mixed plans (FP32 FFTs + naive DLT on FP32 MFMA): 1e-6, at power-of-two
bandwidths and others (mixed-radix FP32 FFTs, GEMMs over rings padded to
whole 64-ring groups).
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


def to_dev(torch, dev, arrs):
    return torch.tensor(np.stack(arrs), dtype=torch.float64, device=dev)


# cutoff < bw: the naive branch for orders >= cutoff (bw/4 is test_s2_semi_memo.c:74-75's
# choice, bw/2 - 10 test_s2_semi_fly.c:128's; 0 makes every order naive)
@pytest.mark.parametrize("bw,cutoff,fmt,prec", [(8, 8, 0, 0), (17, 17, 0, 0), (32, 32, 0, 0), (64, 64, 0, 0),
                                                (64, 16, 0, 0), (64, 64, 1, 0), (128, 128, 0, 0), (512, 512, 0, 0),
                                                (32, 0, 0, 0), (32, 31, 1, 0), (128, 54, 1, 0), (256, 64, 0, 0),
                                                (512, 128, 0, 0), (17, 4, 0, 0), (64, 16, 0, 1), (128, 0, 1, 1),
                                                (256, 64, 0, 1), (100, 100, 0, 0), (255, 255, 1, 0),
                                                (90, 30, 0, 0), (45, 45, 1, 0), (28, 28, 0, 0), (21, 7, 1, 0),
                                                (32, 32, 0, 1), (64, 64, 0, 1), (128, 128, 1, 1), (256, 256, 0, 1),
                                                (64, 64, 0, 2), (128, 128, 1, 2), (256, 256, 0, 2),
                                                (256, 256, 1, 2), (32, 32, 0, 2), (32, 32, 1, 2), (32, 16, 0, 2),
                                                (100, 100, 0, 2), (100, 100, 1, 2), (255, 255, 0, 2), (255, 255, 1, 2),
                                                (45, 45, 0, 2), (33, 33, 1, 2), (150, 150, 0, 2), (300, 75, 1, 2),
                                                (32, 32, 0, 3), (32, 32, 1, 3), (64, 64, 0, 3), (128, 128, 1, 3),
                                                (256, 256, 0, 3), (256, 256, 1, 3),
                                                # rings with a prime factor above 19 (Bluestein FFTs): 2*127,
                                                # 2*59, 2*37, 2*47, 2*263
                                                (127, 127, 0, 0), (59, 20, 1, 0), (37, 0, 0, 0), (127, 127, 0, 2),
                                                (47, 47, 1, 2), (37, 37, 0, 2), (263, 263, 0, 2)])
def test_plan_matches_oracle(env, O, bw, cutoff, fmt, prec):
    S, torch, dev = env
    print(f"plan bw={bw} cutoff={cutoff} fmt={fmt} prec={prec}", flush=True)
    # FP32 seminaive: cosine-space projections cancel, measured ~1e-6 .. 4e-5
    tol = {0: 1e-11, 1: 1e-4, 2: 1e-6, 3: 1e-6}[prec]
    F = 3 if bw <= 256 else 1
    plan = S.Plan(bw, F, precision=prec, cutoff=cutoff)
    kind = 2 if fmt == 1 else 0
    coeffs = [O.make_coeffs(100 + f, bw, kind) for f in range(F)]
    cre = to_dev(torch, dev, [c[0] for c in coeffs])
    cim = to_dev(torch, dev, [c[1] for c in coeffs])
    n = 2 * bw
    gre = torch.empty((F, n, n), dtype=torch.float64, device=dev)
    gim = torch.empty_like(gre)
    plan.inverse(cre, cim, gre, gim, data_format=fmt)
    c2r, c2i = torch.empty_like(cre), torch.empty_like(cim)
    plan.forward(gre, gim, c2r, c2i, data_format=fmt)
    torch.cuda.synchronize()
    gre_h, gim_h = gre.cpu().numpy(), gim.cpu().numpy()
    c2r_h, c2i_h = c2r.cpu().numpy(), c2i.cpu().numpy()
    # The mixed and split forms run the naive transform for every order, so at
    # odd bw they synthesise degree bw-1 whole, as the reference's naive branch
    # does; the oracle at cutoff 0 is that transform (at cutoff bw its seminaive
    # inverse keeps the reference's short Transpose_RowSize rows).
    ocut = 0 if (prec >= 2 and bw % 2) else cutoff
    for f in range(F):
        ord_, oid = O.inv_fst(coeffs[f][0], coeffs[f][1], bw, data_format=fmt, cutoff=ocut)
        assert crel(gre_h[f].ravel(), gim_h[f].ravel(), ord_, oid) < tol
        orc, oic = O.fst(gre_h[f].ravel(), gim_h[f].ravel(), bw, data_format=fmt, cutoff=ocut)
        assert crel(c2r_h[f], c2i_h[f], orc, oic) < tol
    plan.close()


@pytest.mark.parametrize("bw,cutoff", [(16, 16), (64, 64), (64, 16), (256, 256)])
def test_plan_tables_match_oracle(env, O, bw, cutoff):
    import ctypes
    S, torch, dev = env
    plan = S.Plan(bw, 1, cutoff=cutoff)
    f, i, n = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_long()
    assert S.lib().s2amd_plan_tables(plan.h, ctypes.byref(f), ctypes.byref(i), ctypes.byref(n)) == 0
    ot = O.Tables(bw, cutoff)
    assert n.value == ot.fwd.size
    hf, hi = np.empty(n.value), np.empty(n.value)
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    for host, src in ((hf, f), (hi, i)):
        assert hip.hipMemcpy(host.ctypes.data, src.value, 8 * n.value, 2) == 0  # device to host
    scale = np.abs(ot.fwd).max()
    assert np.abs(hf - ot.fwd).max() / scale < 1e-13
    assert np.abs(hi - ot.inv).max() / scale < 1e-13
    plan.close()


@pytest.mark.parametrize("bw", [64, 45, 23, 59])
def test_plan_conv_and_zonal(env, O, bw):
    # 45: 90 = 2*3^2*5 rings (register mixed-radix FFTs); 23 and 59: Bluestein
    # (46 rings, convolution length 96 = 2^5*3; 118 rings, length 256)
    S, torch, dev = env
    F = 2
    plan = S.Plan(bw, F)
    rng = np.random.default_rng(31)
    n = 2 * bw
    sig = [O.inv_fst(*O.make_coeffs(40 + f, bw, 0), bw) for f in range(F)]
    filt = []
    for f in range(F):
        sigma = 5 + rng.uniform() * (bw / 2 - 5)
        l = np.arange(bw)
        hc = np.zeros(bw * bw)
        hc[:bw] = np.exp(-l * (l + 1) / (2 * sigma ** 2))
        filt.append(O.inv_fst(hc, np.zeros(bw * bw), bw))
    sre = to_dev(torch, dev, [s[0].reshape(n, n) for s in sig])
    sim = to_dev(torch, dev, [s[1].reshape(n, n) for s in sig])
    hre = to_dev(torch, dev, [h[0].reshape(n, n) for h in filt])
    him = to_dev(torch, dev, [h[1].reshape(n, n) for h in filt])
    ore, oim = torch.empty_like(sre), torch.empty_like(sim)
    plan.conv(sre, sim, hre, him, ore, oim)
    zr = torch.empty((F, bw), dtype=torch.float64, device=dev)
    zi = torch.empty_like(zr)
    plan.zonal(hre, him, zr, zi, data_format=0)
    torch.cuda.synchronize()
    for f in range(F):
        orr, oir = O.conv(sig[f][0], sig[f][1], filt[f][0], filt[f][1], bw)
        assert crel(ore[f].cpu().numpy().ravel(), oim[f].cpu().numpy().ravel(), orr, oir) < 1e-11
        ozr, ozi = O.fzt(filt[f][0], filt[f][1], bw, data_format=0)
        assert crel(zr[f].cpu().numpy(), zi[f].cpu().numpy(), ozr, ozi) < 1e-11
    plan.close()


def test_plan_argument_errors(env):
    S, torch, dev = env
    plan = S.Plan(16, 2)
    x = torch.zeros((3, 32, 32), dtype=torch.float64, device=dev)
    c = torch.zeros((3, 256), dtype=torch.float64, device=dev)
    with pytest.raises(S.S2Error, match="nfields"):
        plan.forward(x, x, c, c)
    plan.forward(x, x, c, c, nfields=0)  # no-op
    plan.close()


@pytest.mark.parametrize("prec,tol", [(0, 1e-11), (1, 1e-4), (2, 1e-6), (3, 1e-6)])
def test_full_size_round_trip_bw256(env, O, prec, tol):
    # B=256, 8 power-law fields: inverse then forward is the identity on
    # band-limited coefficients; and the transform is linear.
    S, torch, dev = env
    bw, F = 256, 8
    plan = S.Plan(bw, F, precision=prec)
    coeffs = [O.make_coeffs(500 + f, bw, 1) for f in range(F)]
    cre = to_dev(torch, dev, [c[0] for c in coeffs])
    cim = to_dev(torch, dev, [c[1] for c in coeffs])
    n = 2 * bw
    gre = torch.empty((F, n, n), dtype=torch.float64, device=dev)
    gim = torch.empty_like(gre)
    plan.inverse(cre, cim, gre, gim)
    c2r, c2i = torch.empty_like(cre), torch.empty_like(cim)
    plan.forward(gre, gim, c2r, c2i)
    # linearity: forward(2 g0 - g1) = 2 c0 - c1
    lre = (2 * gre[0] - gre[1]).unsqueeze(0).contiguous()
    lim = (2 * gim[0] - gim[1]).unsqueeze(0).contiguous()
    l2r, l2i = torch.empty_like(cre[:1]), torch.empty_like(cim[:1])
    plan.forward(lre, lim, l2r, l2i, nfields=1)
    torch.cuda.synchronize()
    a = torch.complex(c2r, c2i)
    b = torch.complex(cre, cim)
    err = float((a - b).abs().max() / b.abs().max())
    lin = torch.complex(l2r[0], l2i[0]) - (2 * a[0] - a[1])
    lerr = float(lin.abs().max() / b.abs().max())
    print(f"bw=256 prec={prec}: round-trip rel err {err:.3e}, linearity {lerr:.3e}", flush=True)
    assert err < tol and lerr < tol
    plan.close()


@pytest.mark.parametrize("bw,prec,tol", [(512, 1, 1e-4), (512, 0, 1e-11), (512, 2, 1e-6), (512, 3, 1e-6), (1024, 0, 1e-11),
                                         (1024, 1, 1e-3)])
def test_large_bw_round_trip(env, bw, prec, tol):
    # Largest supported grids: inverse then forward on band-limited
    # power-law coefficients (size-independent identity; no oracle at this size).
    S, torch, dev = env
    from s2kit_amd import synth
    plan = S.Plan(bw, 1, precision=prec)
    cre, cim = synth.power_law_coeffs(bw, 1, 77, dev)
    n = 2 * bw
    gre = torch.empty((1, n, n), dtype=torch.float64, device=dev)
    gim = torch.empty_like(gre)
    plan.inverse(cre, cim, gre, gim)
    c2r, c2i = torch.empty_like(cre), torch.empty_like(cim)
    plan.forward(gre, gim, c2r, c2i)
    torch.cuda.synchronize()
    b = torch.complex(cre, cim)
    err = float((torch.complex(c2r, c2i) - b).abs().max() / b.abs().max())
    print(f"bw={bw} prec={prec}: round-trip rel err {err:.3e}", flush=True)
    assert err < tol
    plan.close()


@pytest.mark.parametrize("bw", [600, 1024])
def test_single_field_path_agrees_with_batched(env, bw):
    # An FP64 plan of one field at bw >= 512 runs the per-field kernels (row
    # GEMVs, 32 lanes per row in the forward above bw=542; mixed radix at
    # bw=600), a plan of two fields the MFMA GEMMs: the same transforms, FP64
    # both ways, so each direction pins the other path independently.
    S, torch, dev = env
    from s2kit_amd import synth
    n = 2 * bw
    cre, cim = synth.power_law_coeffs(bw, 2, 31, dev)
    p1, p2 = S.Plan(bw, 1, precision=0), S.Plan(bw, 2, precision=0)
    g2r = torch.empty((2, n, n), dtype=torch.float64, device=dev)
    g2i = torch.empty_like(g2r)
    p2.inverse(cre, cim, g2r, g2i)
    g1r, g1i = torch.empty_like(g2r[:1]), torch.empty_like(g2i[:1])
    p1.inverse(cre[:1].contiguous(), cim[:1].contiguous(), g1r, g1i)
    c2r, c2i = torch.empty_like(cre), torch.empty_like(cim)
    p2.forward(g2r, g2i, c2r, c2i)
    c1r, c1i = torch.empty_like(cre[:1]), torch.empty_like(cim[:1])
    p1.forward(g2r[:1].contiguous(), g2i[:1].contiguous(), c1r, c1i)
    torch.cuda.synchronize()
    ginv = float(torch.maximum((g1r - g2r[:1]).abs().max(), (g1i - g2i[:1]).abs().max())
                 / torch.maximum(g2r[:1].abs().max(), g2i[:1].abs().max()))
    cfwd = float(torch.maximum((c1r - c2r[:1]).abs().max(), (c1i - c2i[:1]).abs().max())
                 / torch.maximum(c2r[:1].abs().max(), c2i[:1].abs().max()))
    print(f"bw={bw}: one field vs two, inverse {ginv:.2e}, forward {cfwd:.2e}", flush=True)
    assert ginv < 1e-11 and cfwd < 1e-11
    p1.close()
    p2.close()


@pytest.mark.parametrize("bw,prec,tol", [(256, 0, 1e-11), (256, 1, 1e-4), (200, 0, 1e-11)])
@pytest.mark.parametrize("fmt", [0, 1])
def test_cutoff_plans_agree(env, bw, prec, tol, fmt):
    # Seminaive and naive orders compute the same quadrature sums, so plans at
    # cutoff bw/4 and 0 must reproduce the cutoff = bw plan (both directions)
    # at the full B=256 shape (and at B=200, a grid of 400 = 4*4*5*5 rings),
    # 16 fields.  (Odd bw is left out: there the reference's transposed
    # seminaive table is short by design, so its inverse differs from the
    # naive one.)
    S, torch, dev = env
    from s2kit_amd import synth
    F = 16
    n = 2 * bw
    cre, cim = synth.power_law_coeffs(bw, F, 91, dev)
    if fmt == 1:  # Hermitian coefficients of a real field
        g0 = torch.empty((F, n, n), dtype=torch.float64, device=dev)
        g1 = torch.empty_like(g0)
        full = S.Plan(bw, F)
        full.inverse(cre, cim, g0, g1)
        full.forward(g0, torch.zeros_like(g1), cre, cim)
        full.close()
    out = {}
    for cut in (bw, bw // 4, 0):
        plan = S.Plan(bw, F, precision=prec, cutoff=cut)
        gre = torch.empty((F, n, n), dtype=torch.float64, device=dev)
        gim = torch.empty_like(gre)
        plan.inverse(cre, cim, gre, gim, data_format=fmt)
        c2r, c2i = torch.empty_like(cre), torch.empty_like(cim)
        plan.forward(gre, gim, c2r, c2i, data_format=fmt)
        torch.cuda.synchronize()
        out[cut] = (torch.complex(gre, gim), torch.complex(c2r, c2i))
        plan.close()
    g_ref, c_ref = out[bw]
    for cut in (bw // 4, 0):
        g, c = out[cut]
        eg = float((g - g_ref).abs().max() / g_ref.abs().max())
        ec = float((c - c_ref).abs().max() / c_ref.abs().max())
        print(f"cutoff={cut} prec={prec} fmt={fmt}: grid {eg:.3e}, coefficients {ec:.3e}", flush=True)
        assert eg < tol and ec < tol


@pytest.mark.parametrize("bw,prec,tol", [(1000, 0, 1e-11), (1018, 0, 1e-11), (1021, 0, 1e-11), (509, 2, 1e-6)])
def test_large_bw_not_power_of_two_round_trip(env, bw, prec, tol):
    # bw=1000: 2000 = 4*4*5*5*5 rings (register mixed-radix FFTs and DCTs);
    # bw=1018 and 1021: 2036 = 4*509 and 2042 = 2*1021 rings (Bluestein,
    # M = 4096, one ring per block in 131 KiB of LDS); mixed form at bw=509:
    # 1018 = 2*509 rings in FP32 (Bluestein, M = 2048).  Two fields take the
    # GEMM path (1008- and 1024-double LDS rows, up to 160 KiB).  Inverse then
    # forward is the identity on band-limited coefficients; at odd bw the
    # FP64 (seminaive) inverse synthesises degree bw-1 only in part (the
    # reference's short Transpose_RowSize rows), so those coefficients are
    # zero here, while the mixed form (naive) synthesises every degree.
    S, torch, dev = env
    from s2kit_amd import synth
    F = 2
    print(f"round trip bw={bw} prec={prec}", flush=True)
    plan = S.Plan(bw, F, precision=prec)
    cre, cim = synth.power_law_coeffs(bw, F, 78, dev)
    if bw % 2 and prec == 0:
        top = torch.tensor(synth.degrees(bw) == bw - 1, device=dev)
        cre[:, top] = 0
        cim[:, top] = 0
    n = 2 * bw
    gre = torch.empty((F, n, n), dtype=torch.float64, device=dev)
    gim = torch.empty_like(gre)
    plan.inverse(cre, cim, gre, gim)
    c2r, c2i = torch.empty_like(cre), torch.empty_like(cim)
    plan.forward(gre, gim, c2r, c2i)
    torch.cuda.synchronize()
    b = torch.complex(cre, cim)
    err = float((torch.complex(c2r, c2i) - b).abs().max() / b.abs().max())
    print(f"bw={bw} F={F}: round-trip rel err {err:.3e}", flush=True)
    assert err < tol
    plan.close()


@pytest.mark.parametrize("prec", [2, 3])
@pytest.mark.parametrize("fmt", [0, 1])
def test_one_fp64_refinement_reaches_fp64(env, O, prec, fmt):
    # mixed / split-BF16 forward with one FP64 residual refinement
    # (s2kit_amd.forward_refined) against the oracle at the FP64 bound, where
    # the unrefined forward sits at ~3e-7
    S, torch, dev = env
    bw, F = 128, 2
    lo = S.Plan(bw, F, precision=prec)
    p64 = S.Plan(bw, F, precision=0)
    kind = 2 if fmt == 1 else 0
    coeffs = [O.make_coeffs(300 + f, bw, kind) for f in range(F)]
    cre = to_dev(torch, dev, [c[0] for c in coeffs])
    cim = to_dev(torch, dev, [c[1] for c in coeffs])
    n = 2 * bw
    gre = torch.empty((F, n, n), dtype=torch.float64, device=dev)
    gim = torch.empty_like(gre)
    p64.inverse(cre, cim, gre, gim, data_format=fmt)
    c2r, c2i = torch.empty_like(cre), torch.empty_like(cim)
    S.forward_refined(lo, p64, gre, gim, c2r, c2i, data_format=fmt)
    u2r, u2i = torch.empty_like(cre), torch.empty_like(cim)
    lo.forward(gre, gim, u2r, u2i, data_format=fmt)
    torch.cuda.synchronize()
    for f in range(F):
        orc, oic = O.fst(gre[f].cpu().numpy().ravel(), gim[f].cpu().numpy().ravel(), bw, data_format=fmt)
        ef = crel(c2r[f].cpu().numpy(), c2i[f].cpu().numpy(), orc, oic)
        eu = crel(u2r[f].cpu().numpy(), u2i[f].cpu().numpy(), orc, oic)
        print(f"prec={prec} fmt={fmt} field {f}: refined forward {ef:.2e}, unrefined {eu:.2e}", flush=True)
        assert ef < 1e-11 and eu > 1e-9
    lo.close()
    p64.close()


@pytest.mark.parametrize("bw,prec", [(256, 2), (64, 0), (100, 0)])
def test_plan_calls_replay_from_a_hip_graph(env, bw, prec):
    # forward + inverse recorded once in a HIP graph (torch.cuda.CUDAGraph) and
    # replayed three times, the inputs rescaled eagerly between replays; every
    # replay matches the eager calls on the same inputs
    S, torch, dev = env
    from s2kit_amd import synth
    F, n = 4, 2 * bw
    plan = S.Plan(bw, F, precision=prec)
    cre, cim = synth.power_law_coeffs(bw, F, 17, dev)
    gre = torch.empty((F, n, n), dtype=torch.float64, device=dev)
    gim = torch.empty_like(gre)
    plan.inverse(cre, cim, gre, gim)
    c2r, c2i = torch.empty_like(cre), torch.empty_like(cim)
    ore, oim = torch.empty_like(gre), torch.empty_like(gim)

    def step():
        plan.forward(gre, gim, c2r, c2i)
        plan.inverse(c2r, c2i, ore, oim)

    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        step()  # warm-up outside the capture (tables, LDS attributes)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        step()
    for k in range(3):
        gre.mul_(1.5)
        gim.mul_(-0.5)
        ore.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        gr_c, gi_c = c2r.clone(), c2i.clone()
        go_r, go_i = ore.clone(), oim.clone()
        step()  # eager on the same inputs
        torch.cuda.synchronize()
        assert torch.equal(gr_c, c2r) and torch.equal(gi_c, c2i), k
        assert torch.equal(go_r, ore) and torch.equal(go_i, oim), k
    plan.close()


@pytest.mark.parametrize("bw,prec", [(256, 0), (128, 1), (256, 2), (100, 0), (64, 3)])
def test_autotune_keeps_results(env, bw, prec):
    # s2amd_plan_autotune times the tile shapes and keeps the fastest; every
    # shape computes the same sums in the same order, so the results are
    # bit-identical before and after
    S, torch, dev = env
    from s2kit_amd import synth
    F, n = 8, 2 * bw
    plan = S.Plan(bw, F, precision=prec)
    cre, cim = synth.power_law_coeffs(bw, F, 23, dev)
    gre = torch.empty((F, n, n), dtype=torch.float64, device=dev)
    gim = torch.empty_like(gre)
    c2r, c2i = torch.empty_like(cre), torch.empty_like(cim)

    def pair():
        plan.inverse(cre, cim, gre, gim)
        plan.forward(gre, gim, c2r, c2i)
        torch.cuda.synchronize()
        return gre.clone(), gim.clone(), c2r.clone(), c2i.clone()

    before = pair()
    assert plan.tuning() == (0, 0, 0)
    plan.autotune(F)
    cols, tiles, rings = plan.tuning()
    print(f"bw={bw} prec={prec}: tuned cols={cols} col_tiles={tiles} ring_groups={rings}", flush=True)
    if prec in (0, 1):
        assert cols in (16, 32, 64) and tiles in (1, 2) and rings == 0
    elif prec == 2:
        assert cols == 0 and rings in (1, 2, 4)
    else:
        assert (cols, tiles, rings) == (0, 0, 0)  # split BF16: nothing to choose
    after = pair()
    for a, b in zip(before, after):
        assert torch.equal(a, b)
    plan.close()


@pytest.mark.parametrize("bw,prec", [(256, 0), (128, 0), (64, 0), (256, 1), (100, 0)])
def test_few_field_calls_match_batched(env, bw, prec):
    # Calls of few fields run the ring FFTs and DCTs on one-wave blocks and
    # slice each order's GEMM tiles over several blocks; the butterflies and
    # every tile's sums are the same, so one field per call reproduces the
    # batched call bit for bit (the batch of 8 at bw >= 256 runs the default
    # 512-thread blocks and unsliced GEMMs).
    S, torch, dev = env
    from s2kit_amd import synth
    F, n = 8, 2 * bw
    cre, cim = synth.power_law_coeffs(bw, F, 29, dev)
    outs = []
    for per_call in (F, 1):
        plan = S.Plan(bw, per_call, precision=prec)
        gre = torch.empty((F, n, n), dtype=torch.float64, device=dev)
        gim = torch.empty_like(gre)
        c2r, c2i = torch.empty_like(cre), torch.empty_like(cim)
        for f0 in range(0, F, per_call):
            sl = slice(f0, f0 + per_call)
            plan.inverse(cre[sl], cim[sl], gre[sl], gim[sl])
            plan.forward(gre[sl], gim[sl], c2r[sl], c2i[sl])
        torch.cuda.synchronize()
        plan.close()
        outs.append((gre, gim, c2r, c2i))
    for a, b in zip(*outs):
        assert torch.equal(a, b)


@pytest.mark.parametrize("bw,prec,cutoff", [(256, 0, 256), (128, 0, 128), (255, 0, 255), (256, 1, 256), (128, 1, 128),
                                            (512, 0, 512), (512, 1, 512), (300, 0, 300), (256, 0, 64), (255, 0, 100),
                                            (512, 0, 128), (128, 1, 32)])
@pytest.mark.parametrize("fmt", [0, 1])
def test_k_chunked_gemms_match_whole_row_gemms(env, O, bw, prec, cutoff, fmt):
    # Calls of 64 or more fields at bw in [128, 512] run the seminaive orders
    # on the K-chunked GEMMs (64 columns per block, data streamed through LDS
    # chunks) and, at cutoff < bw, the naive orders on the whole-row GEMMs in
    # a second launch; calls of 8 fields run every order on the whole-row
    # GEMMs.  Both take every accumulator's K steps in the same order, so the
    # results agree bit for bit; two fields are also checked against the
    # oracle (FP64).
    S, torch, dev = env
    F, per, n = 64, 8, 2 * bw
    print(f"k-chunked bw={bw} prec={prec} cutoff={cutoff} fmt={fmt}", flush=True)
    kind = 2 if fmt == 1 else 0
    coeffs = [O.make_coeffs(300 + f, bw, kind) for f in range(F)]
    cre = to_dev(torch, dev, [c[0] for c in coeffs])
    cim = to_dev(torch, dev, [c[1] for c in coeffs])
    outs = []
    for per_call in (F, per):
        plan = S.Plan(bw, per_call, precision=prec, cutoff=cutoff)
        gre = torch.empty((F, n, n), dtype=torch.float64, device=dev)
        gim = torch.empty_like(gre)
        c2r, c2i = torch.empty_like(cre), torch.empty_like(cim)
        for f0 in range(0, F, per_call):
            sl = slice(f0, f0 + per_call)
            plan.inverse(cre[sl], cim[sl], gre[sl], gim[sl], data_format=fmt)
            plan.forward(gre[sl], gim[sl], c2r[sl], c2i[sl], data_format=fmt)
        torch.cuda.synchronize()
        plan.close()
        outs.append((gre, gim, c2r, c2i))
    for a, b in zip(*outs):
        assert torch.equal(a, b)
    if prec == 0:
        gre, gim, c2r, c2i = outs[0]
        for f in (0, F - 1):
            og = O.inv_fst(coeffs[f][0], coeffs[f][1], bw, data_format=fmt, cutoff=cutoff)
            assert crel(gre[f].cpu().numpy().ravel(), gim[f].cpu().numpy().ravel(), *og) < 1e-11
            oc = O.fst(gre[f].cpu().numpy().ravel(), gim[f].cpu().numpy().ravel(), bw, data_format=fmt, cutoff=cutoff)
            assert crel(c2r[f].cpu().numpy(), c2i[f].cpu().numpy(), *oc) < 1e-11


@pytest.mark.parametrize("bw,F,prec", [(128, 64, 0), (128, 8, 0), (256, 64, 1), (255, 64, 0)])
def test_real_inverse_does_not_depend_on_earlier_calls(env, bw, F, prec):
    # Every call writes all the plan scratch its later stages read: a REAL
    # inverse gives the same bits on a fresh plan and after COMPLEX and REAL
    # calls on the same plan have left their intermediates behind (an order
    # whose odd parity has no degree still stores its zero cosine row).
    S, torch, dev = env
    from s2kit_amd import synth
    n = 2 * bw
    cre, cim = synth.power_law_coeffs(bw, F, 41, dev)
    p0 = S.Plan(bw, F, precision=prec)
    gre = torch.empty((F, n, n), dtype=torch.float64, device=dev)
    gim = torch.empty_like(gre)
    p0.inverse(cre, cim, gre, gim)
    grr, gri = gre.clone(), torch.zeros_like(gim)
    crr, cri = torch.empty_like(cre), torch.empty_like(cim)
    p0.forward(grr, gri, crr, cri, data_format=S.REAL)
    p0.close()

    def real_inverse(plan):
        a, b = torch.empty_like(gre), torch.empty_like(gim)
        plan.inverse(crr, cri, a, b, data_format=S.REAL)
        torch.cuda.synchronize()
        return a, b

    plan = S.Plan(bw, F, precision=prec)
    fresh = real_inverse(plan)
    c2r, c2i = torch.empty_like(cre), torch.empty_like(cim)
    plan.inverse(cre, cim, gre, gim)
    plan.forward(gre, gim, c2r, c2i)
    plan.forward(grr, gri, c2r, c2i, data_format=S.REAL)
    later = real_inverse(plan)
    plan.close()
    assert torch.equal(fresh[0], later[0]) and torch.equal(fresh[1], later[1])
