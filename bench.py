#!/usr/bin/env python3
"""Throughput of the batched forward + inverse SHT on one MI355X.

Workload: 1024 independent fields at bandwidth 256 (512x512 Driscoll-Healy
grids, complex FP64 in and out), power-law coefficients (E|c|^2 = (l+1)^-2)
synthesised on the GPU.  One step = forward transform of all fields (grids
-> coefficients) followed by the inverse (coefficients -> grids), inputs
already resident in HBM.  value = fwd+inv pairs per second = fields * steps /
timed seconds.

Also reported on the same JSON line:
  errors      the timed form's results against the FP64 GPU form on the same
              input (normwise max|d|/max|ref|), and the FP64 round trip;
  roofline    the dominant kernel's algorithmic HBM bytes per launch over its
              average duration (HIP events on the launch stream, recorded
              during the timed steps), against 8 TB/s;
  mfma        the Legendre GEMMs' useful flops per launch over their average
              duration, and the MFMA flops issued against the dense peak of
              their operand type;
  other_forms the FP64 form and the other FP32-accurate form (mixed: FP32
              MFMA GEMMs; split: BF16 MFMA with split operands) timed on the
              same fields, with their errors against FP64; and the FP64 form
              at cutoff = bw/4 (naive branch for orders >= 64);
  non_pow2    FP64 and mixed pairs/s at bw=255 (register mixed-radix FFTs)
              beside bw=256, bw=127 and bw=263 (Bluestein FFTs) beside bw=128,
              256 fields, and the bw=255 / 127 batches against their per-field path;
  cpu_baseline  the CPU FP64 oracle (a port of the reference path) timed on
              the host, one process per core, on a bounded sample.

With --workload conv: an uneven-bandwidth batch of 4096 convolutions (one
recorded HIP graph per form), mixed against FP64, errors per bandwidth.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload sht|conv]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s


def stage_bytes(stage, bw, nfields, elem, form):
    """Algorithmic HBM bytes of one launch of `stage` over nfields fields: the
    compulsory reads + writes of its inputs and outputs.  Grids and
    coefficients are complex FP64 (16 B); the order-major intermediates hold
    `elem` bytes per complex value.  Tables (shared by all fields of a launch)
    are left out."""
    n = 2 * bw
    rows = 2 * bw - 1  # order rows used (all but the Nyquist row)
    grid = 16 * n * n
    coef = 16 * bw * bw
    xrow = elem * n * rows  # phi-FFT output rows that the next stage reads
    crow = elem * bw * rows  # cosine coefficients (seminaive forms)
    if form in ("mixed", "split"):
        per_field = {"phi_fwd": grid + elem * n * n, "leg_fwd": xrow + coef, "leg_inv": coef + xrow,
                     "phi_inv": xrow + grid}[stage]
    else:
        per_field = {"phi_fwd": grid + elem * n * n, "dct_fwd": xrow + crow, "leg_fwd": crow + coef,
                     "leg_inv": coef + crow, "dct_inv": crow + xrow, "phi_inv": xrow + grid}[stage]
    return per_field * nfields


# kernel behind each timed stage, per form (names as rocprofv3 reports them;
# CT, the GEMMs' column tiles per table load, is 1 or 2 depending on bw)
KERNELS = {
    "mixed": {"phi_fwd": "k_fphi_fwd<float, 16, 512, true>", "leg_fwd": "k_mix_fwd", "leg_inv": "k_mix_inv",
              "phi_inv": "k_fphi_inv<float, 16, 512, true>"},
    "split": {"phi_fwd": "k_fphi_fwd<float, 16, 512, true>", "leg_fwd": "k_split_fwd", "leg_inv": "k_split_inv",
              "phi_inv": "k_fphi_inv<float, 16, 512, true>"},
    "fp32": {"phi_fwd": "k_fphi_fwd<float, 16, 512, true>", "dct_fwd": "k_fdct_fwd<float, 16>",
             "leg_fwd": "k_gemm_fwd<float, false, CT>", "leg_inv": "k_gemm_inv<float, false, CT>",
             "dct_inv": "k_fdct_inv<float, 16>", "phi_inv": "k_fphi_inv<float, 16, 512, false>"},
    "fp64": {"phi_fwd": "k_fphi_fwd<double, 8, 512, true>", "dct_fwd": "k_fdct_fwd<double, 8>",
             "leg_fwd": "k_gemm_fwd<double, false, CT>", "leg_inv": "k_gemm_inv<double, false, CT>",
             "dct_inv": "k_fdct_inv<double, 8>", "phi_inv": "k_fphi_inv<double, 8, 512, false>"},
}

def smooth(n, pmax):
    """n has no prime factor above pmax."""
    for p in range(2, pmax + 1):
        while n % p == 0:
            n //= p
    return n == 1


def ring_fft_instance(n):
    """Instance of the per-field FFT/DCT kernels for ring length n that is
    not a power of two (the host's choice in s2_kernels.hip, fft_cfg): 1 the
    register mixed-radix passes (prime factors up to 19), else Bluestein with
    a convolution of length M >= 2n-1, 2 for a power-of-two M (taken when
    within 5/4 of the smallest 7-smooth length), 3 for that 7-smooth M."""
    if smooth(n, 19):
        return 1
    p2 = 1
    while p2 < 2 * n - 1:
        p2 <<= 1
    ms = 2 * n - 1
    while not smooth(ms, 7):
        ms += 1
    return 2 if 4 * p2 <= 5 * ms else 3


def kernels_gen(bw, form):
    """Kernel behind each timed stage at a bandwidth that is not a power of
    two: the per-field FFT/DCT instances and the GEMMs."""
    mode = ring_fft_instance(2 * bw)
    if form in ("mixed", "split"):
        k = KERNELS[form]
        return {"phi_fwd": f"k_phi_fwd<float2, {mode}>", "leg_fwd": k["leg_fwd"], "leg_inv": k["leg_inv"],
                "phi_inv": f"k_phi_inv<float2, {mode}>"}
    return {"phi_fwd": f"k_phi_fwd<double2, {mode}>", "dct_fwd": f"k_dct_fwd<double2, {mode}>",
            "leg_fwd": "k_gemm_fwd<double, true, CT>", "leg_inv": "k_gemm_inv<double, true, CT>",
            "dct_inv": f"k_dct_inv<double2, {mode}>", "phi_inv": f"k_phi_inv<double2, {mode}>"}


def kernel_map(form, bw, nfields):
    """Kernel behind each timed stage of a batched call of nfields fields on
    an all-seminaive plan.  The FP64/FP32 GEMMs run K-chunked (64 columns per
    block; 512 threads up to bw = 256, 1024 above) at bw in [128, 512] with 64
    or more fields."""
    k = dict(KERNELS[form] if bw & (bw - 1) == 0 else kernels_gen(bw, form))
    if form in ("fp64", "fp32") and 128 <= bw <= 512 and nfields >= 64:
        r = "double" if form == "fp64" else "float"
        t = 512 if bw <= 256 else 1024
        k["leg_fwd"], k["leg_inv"] = f"k_gemm_fwd_kc<{r}, {t}>", f"k_gemm_inv_kc<{r}, {t}>"
    return k



def measured_traffic(form, bw, nfields, stage):
    """HBM bytes per launch of `stage`'s kernel from the committed rocprofv3
    counter summary of this workload (FETCH_SIZE and WRITE_SIZE passes, each
    scaled by a plain copy's reading), or None when there is none."""
    path = os.path.join(ROOT, "profiles", f"pmc_traffic_{form}_bw{bw}_f{nfields}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 3 and f[0] == "traffic_bytes" and f[1] == stage:
            return float(f[2])
    return None


def table_doubles(bw):
    """Doubles in one direction's table at cutoff = bw: sum over m of
    TableSize(m, bw) (cospml.c:39-59)."""
    tot = 0
    for m in range(bw):
        mm, ll = (m - 1, bw - 1) if m & 1 else (m, bw)
        hl, hm = ll // 2, mm // 2
        off = hl * (hl + 1) - hm * (hm + 1)
        if ll & 1:
            off += hl + 1
        tot += off
    return tot


def legendre_flops(bw, nfields, form):
    """Useful flops of one Legendre GEMM launch (either direction) over
    nfields fields: 2 flops per table MAC, 4 data columns per order m > 0
    (re/im x +-m) and 2 at m = 0.  Seminaive forms: TableSize(m) MACs per
    column (4 (2T - TableSize(0)) per field); naive forms (mixed, split): bw
    folded rings x (bw - m) degrees per column (4 bw^3 per field)."""
    if form in ("mixed", "split"):
        return 4 * bw ** 3 * nfields
    ts0 = bw // 2 * (bw // 2 + 1) + (bw // 2 + 1 if bw & 1 else 0)  # TableSize(0, bw)
    return 4 * (2 * table_doubles(bw) - ts0) * nfields


# dense MFMA peaks of MI355X (TFLOP/s) by the GEMM's operand type, and the MFMA
# flops issued per useful flop (split BF16: six piece products)
MFMA_PEAK = {"fp64": (78.6, 1), "fp32": (157.3, 1), "mixed": (157.3, 1), "split": (2500.0, 6)}


def non_pow2(dev, steps):
    """FP64 plans at bandwidths whose grids are not powers of two, beside the
    power of two next to them, 256 fields each: bw=255 (510 = 2*3*5*17 rings:
    register mixed-radix FFTs and DCTs), bw=127 (254 = 2*127: Bluestein, a
    convolution of length 512) and bw=263 (526 = 2*263: Bluestein, length
    1080 = 2^3 3^3 5); MFMA Legendre GEMMs throughout.  The bw=255 and bw=127
    batches are checked against the drop-ins (per-field kernels) on their
    first field.  The mixed form (FP32 ring FFTs, naive GEMMs over rings
    padded to whole 64-ring groups) is timed on the same fields, with its
    errors against the FP64 plan at cutoff 0 (the all-naive transform, which
    the mixed form computes: at odd bw it synthesises degree bw-1 whole)."""
    import torch
    import s2kit_amd as S
    from s2kit_amd import synth
    out = {"fields": 256}
    for bw in (255, 256, 127, 128, 263):
        F, n = 256, 2 * bw
        c0r, c0i = synth.power_law_coeffs(bw, F, 7, dev)
        gre = torch.empty((F, n, n), dtype=torch.float64, device=dev)
        gim = torch.empty_like(gre)
        cr, ci = torch.empty_like(c0r), torch.empty_like(c0i)
        p = S.Plan(bw, F, precision=S.FP64)
        for _ in range(2):
            p.inverse(c0r, c0i, gre, gim)
            p.forward(gre, gim, cr, ci)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            p.inverse(c0r, c0i, gre, gim)
            p.forward(gre, gim, cr, ci)
        torch.cuda.synchronize()
        out[f"bw{bw}_pairs_per_s"] = F * steps / (time.perf_counter() - t0)
        if bw in (255, 127):  # field 0 through the drop-ins (per-field kernels, host buffers)
            tabs = S.HostTables(bw)
            r1, i1 = S.fst(gre[0].cpu().numpy(), gim[0].cpu().numpy(), bw, tabs)
            g1r, g1i = S.inv_fst(c0r[0].cpu().numpy(), c0i[0].cpu().numpy(), bw, tabs)
            a = cr[0].cpu().numpy() + 1j * ci[0].cpu().numpy()
            out[f"bw{bw}_forward_vs_per_field"] = float(np.abs(a - (r1 + 1j * i1)).max() / np.abs(r1 + 1j * i1).max())
            g = (gre[0] + 1j * gim[0]).cpu().numpy().ravel()
            out[f"bw{bw}_inverse_vs_per_field"] = float(np.abs(g - (g1r + 1j * g1i)).max() / np.abs(g1r + 1j * g1i).max())
        p.close()
        # the mixed form on the same fields
        pr = S.Plan(bw, F, precision=S.FP64, cutoff=0)
        pr.inverse(c0r, c0i, gre, gim)
        pr.forward(gre, gim, cr, ci)
        pr.close()
        pm = S.Plan(bw, F, precision=S.MIXED)
        gmr, gmi = torch.empty_like(gre), torch.empty_like(gim)
        cmr, cmi = torch.empty_like(cr), torch.empty_like(ci)
        pm.inverse(c0r, c0i, gmr, gmi)
        pm.forward(gre, gim, cmr, cmi)
        torch.cuda.synchronize()
        out[f"bw{bw}_mixed_inverse_vs_fp64"] = float(torch.hypot(gmr - gre, gmi - gim).max() / torch.hypot(gre, gim).max())
        out[f"bw{bw}_mixed_forward_vs_fp64"] = float(torch.hypot(cmr - cr, cmi - ci).max() / torch.hypot(cr, ci).max())
        t0 = time.perf_counter()
        for _ in range(steps):
            pm.inverse(c0r, c0i, gmr, gmi)
            pm.forward(gmr, gmi, cmr, cmi)
        torch.cuda.synchronize()
        out[f"bw{bw}_mixed_pairs_per_s"] = F * steps / (time.perf_counter() - t0)
        pm.close()
        del gre, gim, gmr, gmi
        torch.cuda.empty_cache()
    return out


def graph_replay_time(fn, reps):
    """Seconds per call of fn (a sequence of plan calls on torch's current
    stream), recorded once in a HIP graph and replayed: the device time of
    the calls with their launch gaps, without the Python and ctypes cost of
    issuing each launch from a loop."""
    import torch
    fn()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        fn()
    g.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        g.replay()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    del g
    return dt


def single_field_bw512(dev, pairs):
    """One power-law field at bw=512 (1024x1024 grid): the FP64 form called
    per field streams both FP64 tables from HBM on every pair; the FP32, mixed
    and split-BF16 forms, and split BF16 with one FP64 residual refinement, are
    compared with it on the same field."""
    import torch
    import s2kit_amd as S
    from s2kit_amd import synth
    bw, n = 512, 1024
    c0r, c0i = synth.power_law_coeffs(bw, 1, 99, dev)
    gre = torch.empty((1, n, n), dtype=torch.float64, device=dev)
    gim = torch.empty_like(gre)
    p64 = S.Plan(bw, 1, precision=S.FP64)
    p64.inverse(c0r, c0i, gre, gim)
    cr, ci = torch.empty_like(c0r), torch.empty_like(c0i)
    for _ in range(2):
        p64.forward(gre, gim, cr, ci)
        p64.inverse(cr, ci, gre, gim)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(pairs):
        p64.forward(gre, gim, cr, ci)
        p64.inverse(cr, ci, gre, gim)
    torch.cuda.synchronize()
    dt_loop = (time.perf_counter() - t0) / pairs

    def pair():
        p64.forward(gre, gim, cr, ci)
        p64.inverse(cr, ci, gre, gim)
    dt = graph_replay_time(pair, max(pairs, 10))
    tbytes = 2 * 8 * table_doubles(bw)  # forward + transposed table, FP64
    out = {"bw": bw, "fp64_pair_ms": dt * 1e3, "fp64_pair_ms_python_loop": dt_loop * 1e3,
           "table_bytes_per_pair": tbytes, "table_stream_GBps": tbytes / dt / 1e9, "errors_vs_fp64": {}}
    p64.inverse(c0r, c0i, gre, gim)
    ref_r, ref_i = torch.empty_like(c0r), torch.empty_like(c0i)
    p64.forward(gre, gim, ref_r, ref_i)
    for name, prec in (("fp32", S.FP32), ("mixed", S.MIXED), ("split", S.SPLIT_BF16)):
        p = S.Plan(bw, 1, precision=prec)
        r, i = torch.empty_like(c0r), torch.empty_like(c0i)
        g2r, g2i = torch.empty_like(gre), torch.empty_like(gim)
        p.forward(gre, gim, r, i)
        p.inverse(c0r, c0i, g2r, g2i)
        torch.cuda.synchronize()
        fe = float((torch.complex(r, i) - torch.complex(ref_r, ref_i)).abs().max() /
                   torch.complex(ref_r, ref_i).abs().max())
        ie = float((torch.complex(g2r, g2i) - torch.complex(gre, gim)).abs().max() / torch.complex(gre, gim).abs().max())
        out["errors_vs_fp64"][name] = {"forward": fe, "inverse": ie}
        if name == "split":
            # split BF16 x3 forward with one FP64 residual refinement
            # (forward_refined): FP64 accuracy, for more than an FP64 forward costs
            S.forward_refined(p, p64, gre, gim, r, i)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(pairs):
                S.forward_refined(p, p64, gre, gim, r, i)
            torch.cuda.synchronize()
            tr = (time.perf_counter() - t0) / pairs
            fe = float((torch.complex(r, i) - torch.complex(ref_r, ref_i)).abs().max() /
                       torch.complex(ref_r, ref_i).abs().max())
            out["errors_vs_fp64"]["split_fp64_refined"] = {"forward": fe, "forward_ms": tr * 1e3}
        p.close()
    p64.close()
    return out


def conv_uneven(args):
    """Uneven-bandwidth convolution batch: 4096 ConvOn2SphereSemiMemo requests
    with bw drawn from {32, 64, 128, 256, 512} (p = .3, .3, .2, .15, .05),
    each a uniform[-1, 1] coefficient signal and its own zonal Gaussian filter
    h_l = exp(-l(l+1)/(2 sigma^2)), sigma ~ U[5, bw/2].  The whole batch is one
    recorded HIP graph (s2amd_conv_batch_*); a step is one replay.  Mixed and
    FP64 batches run on the same inputs; errors are per bandwidth, mixed
    against FP64, normwise per request."""
    import torch
    import s2kit_amd as S
    from s2kit_amd import synth
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    dev = torch.device("cuda:0")
    nreq = args.conv_requests
    bws, data = synth.conv_uneven_requests(nreq, 2024, dev)
    req64 = synth.conv_requests_of(bws, data, "o64")
    reqmx = synth.conv_requests_of(bws, data, "omx")
    out = {}
    for name, reqs, prec in (("fp64", req64, S.FP64), ("mixed", reqmx, S.MIXED)):
        b = S.ConvBatch(reqs, precision=prec)
        for _ in range(args.warmup):
            b.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            b.run()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        out[name] = {"requests_per_s": nreq / dt, "ms_per_batch": dt * 1e3, "groups": b.groups()}
        b.close()
    errs = {}
    for bw, t in data.items():
        d = torch.complex(t["omxr"], t["omxi"]) - torch.complex(t["o64r"], t["o64i"])
        ref = torch.complex(t["o64r"], t["o64i"]).abs().amax(dim=(1, 2))
        errs[str(bw)] = {"requests": int(t["sre"].shape[0]),
                         "max_rel_err_mixed_vs_fp64": float((d.abs().amax(dim=(1, 2)) / ref).max())}
    line = {
        "metric": "uneven-bandwidth conv requests/s (bw in {32..512}, 4096 requests): mixed vs FP64",
        "value": out["mixed"]["requests_per_s"],
        "unit": "requests/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": out["mixed"]["ms_per_batch"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{nreq} ConvOn2SphereSemiMemo requests, bw from {{32,64,128,256,512}} "
                               "p={.3,.3,.2,.15,.05}, uniform signals, zonal Gaussian filters, one HIP graph",
                   "requests": nreq, "parallelism": "single GPU"},
        "fp64_requests_per_s": out["fp64"]["requests_per_s"],
        "speedup_mixed_vs_fp64": out["mixed"]["requests_per_s"] / out["fp64"]["requests_per_s"],
        "groups": {"mixed": out["mixed"]["groups"], "fp64": out["fp64"]["groups"]},
        "errors_per_bw": errs,
    }
    print(json.dumps(line), flush=True)


def dropin_inputs(seed=3):
    """Host inputs of the two CPU-runnable cases: a bw=64 coefficient set
    (uniform[-1, 1]) and a bw=128 signal/filter pair (uniform signal
    coefficients; zonal filter h_l = exp(-l(l+1)/(2*20^2)) * U[0.5, 1])."""
    rng = np.random.default_rng(seed)
    c64 = (rng.uniform(-1, 1, 64 * 64), rng.uniform(-1, 1, 64 * 64))
    s128 = (rng.uniform(-1, 1, 128 * 128), rng.uniform(-1, 1, 128 * 128))
    l = np.arange(128)
    h = np.zeros(128 * 128)
    h[:128] = np.exp(-l * (l + 1) / (2 * 20.0 ** 2)) * rng.uniform(0.5, 1.0, 128)
    return c64, s128, h


def dropin_latency(reps):
    """Per-call time of the drop-ins on host buffers (PCIe copies included),
    on the two CPU-runnable cases: InvFSTSemiMemo + FSTSemiMemo at bw=64
    (cutoff = bw, tables built once by the drop-in builders), and
    ConvOn2SphereSemiMemo at bw=128 (which builds its tables per call, as the
    reference's does)."""
    import s2kit_amd as S
    (cr, ci), (sr, si), h = dropin_inputs()
    T = S.HostTables(64)
    w = S.weights(64)
    rd, idd = S.inv_fst(cr, ci, 64, T)  # first call uploads the tables
    S.fst(rd, idd, 64, T, w)
    t0 = time.perf_counter()
    for _ in range(reps):
        rd, idd = S.inv_fst(cr, ci, 64, T)
        S.fst(rd, idd, 64, T, w)
    pair = (time.perf_counter() - t0) / reps
    T128 = S.HostTables(128)
    gr, gi = S.inv_fst(sr, si, 128, T128)
    hr, hi = S.inv_fst(h, np.zeros_like(h), 128, T128)
    S.conv(gr, gi, hr, hi, 128)
    t0 = time.perf_counter()
    for _ in range(reps):
        S.conv(gr, gi, hr, hi, 128)
    conv = (time.perf_counter() - t0) / reps
    return {"bw64_inv_plus_fwd_ms": pair * 1e3, "bw128_conv_ms": conv * 1e3, "reps": reps}


def dropin_cpu(reps):
    """The same two cases on the CPU oracle (one core), tables precomputed for
    the transform pair, rebuilt per call for the convolution."""
    from oracle import oracle as O
    (cr, ci), (sr, si), h = dropin_inputs()
    O.tables(64)
    t0 = time.perf_counter()
    for _ in range(reps):
        rd, idd = O.inv_fst(cr, ci, 64)
        O.fst(rd, idd, 64)
    pair = (time.perf_counter() - t0) / reps
    gr, gi = O.inv_fst(sr, si, 128)
    hr, hi = O.inv_fst(h, np.zeros_like(h), 128)
    t0 = time.perf_counter()
    for _ in range(max(1, reps // 4)):
        O.conv(gr, gi, hr, hi, 128)
    conv = (time.perf_counter() - t0) / max(1, reps // 4)
    return {"bw64_inv_plus_fwd_ms": pair * 1e3, "bw128_conv_ms": conv * 1e3, "cores": 1}


def cpu_worker(args):
    """One CPU process: `pairs` oracle inverse+forward transforms at bw."""
    bw, pairs, seed = args
    from oracle import oracle as O
    O.tables(bw)  # precomputed and not timed, as test_s2_semi_memo.c:6-8
    rc, ic = O.make_coeffs(seed, bw, 1)
    t0 = time.perf_counter()
    for _ in range(pairs):
        rd, idd = O.inv_fst(rc, ic, bw)
        rc2, ic2 = O.fst(rd, idd, bw)
    return time.perf_counter() - t0


# A shared multi-GPU host allots this many cores to each GPU; the CPU baseline
# runs one process per core of that share (os.cpu_count() reports the whole
# host, which the other GPUs' jobs share).
CPU_SHARE_PER_GPU = 16


def host_cores():
    """(cores this process may run on, cores the host reports)."""
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return usable, os.cpu_count() or usable


def cpu_baseline(bw, pairs_per_proc, procs):
    import multiprocessing as mp
    ctx = mp.get_context("fork")  # before any GPU initialisation in this process
    with ctx.Pool(procs) as pool:
        t0 = time.perf_counter()
        times = pool.map(cpu_worker, [(bw, pairs_per_proc, 7 + i) for i in range(procs)])
        wall = time.perf_counter() - t0
    total = pairs_per_proc * procs
    usable, reported = host_cores()
    return {
        "value": total / max(times),
        "unit": "fwd+inv pairs/s",
        "cores": procs,
        "kind": "port",
        "host_cores_usable": usable,
        "host_cores_reported": reported,
        "sample": f"{total} fwd+inv pairs at bw={bw} (power-law coefficients), {procs} single-threaded "
                  f"processes x {pairs_per_proc} pairs (one per core of the {CPU_SHARE_PER_GPU}-core share of "
                  f"one GPU; {usable} cores usable by this process, {reported} reported by the host), tables "
                  f"precomputed; slowest process {max(times):.2f} s, pool wall {wall:.2f} s; own FFT/DCT "
                  f"(no FFTW on this host)",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--bw", type=int, default=256)
    ap.add_argument("--fields", type=int, default=1024)
    ap.add_argument("--precision", choices=["fp64", "fp32", "mixed", "split"], default="mixed")
    ap.add_argument("--cpu-pairs", type=int, default=150, help="oracle pairs per CPU process (0: skip)")
    ap.add_argument("--cpu-procs", type=int, default=0,
                    help="CPU processes (0: one per usable core, at most the 16-core share of one GPU)")
    ap.add_argument("--plain-fields", type=int, default=64, help="fields timed in the batch-1 FP64 form")
    ap.add_argument("--other-forms", type=int, default=1,
                    help="also time the other batched forms (FP64, mixed / split BF16) on the same fields (0: skip)")
    ap.add_argument("--dropin-reps", type=int, default=20,
                    help="calls timed per drop-in case (bw=64 pair, bw=128 conv); 0: skip")
    ap.add_argument("--single-pairs", type=int, default=10,
                    help="FP64 pairs timed on one bw=512 field (table streaming); 0: skip")
    ap.add_argument("--non-pow2-steps", type=int, default=3,
                    help="FP64 steps timed at bw=255 and bw=256, 256 fields (0: skip)")
    ap.add_argument("--workload", choices=["sht", "conv"], default="sht",
                    help="sht: batched fwd+inv SHT (default); conv: uneven-bandwidth convolution batch")
    ap.add_argument("--conv-requests", type=int, default=4096, help="requests in the conv workload")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one plain 1 GiB device copy (reference for rocprofv3 byte counters)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(fields would shard across GPUs with no exchange); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.workload == "conv":
        conv_uneven(args)
        return

    bw, F = args.bw, args.fields
    cpu = None
    if args.cpu_pairs > 0:
        procs = args.cpu_procs or min(CPU_SHARE_PER_GPU, host_cores()[0])
        cpu = cpu_baseline(bw, args.cpu_pairs, procs)
        if args.dropin_reps > 0:
            cpu["dropin_cases"] = dropin_cpu(args.dropin_reps)

    import torch

    import s2kit_amd as S
    from s2kit_amd import synth

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    dev = torch.device("cuda:0")
    n = 2 * bw
    prec = {"fp64": S.FP64, "fp32": S.FP32, "mixed": S.MIXED, "split": S.SPLIT_BF16}[args.precision]

    if args.copy_calibration:  # reads 1 GiB and writes 1 GiB, nothing else
        src = torch.ones(1 << 27, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    # ---- workload: band-limited grids from power-law coefficients ---------
    c0r, c0i = synth.power_law_coeffs(bw, F, 1234, dev)
    p64 = S.Plan(bw, F, precision=S.FP64)
    gre = torch.empty((F, n, n), dtype=torch.float64, device=dev)
    gim = torch.empty_like(gre)
    p64.inverse(c0r, c0i, gre, gim)
    plan = p64 if prec == S.FP64 else S.Plan(bw, F, precision=prec)
    cre, cim = torch.empty_like(c0r), torch.empty_like(c0i)
    ore, oim = torch.empty_like(gre), torch.empty_like(gim)

    def step():
        plan.forward(gre, gim, cre, cim)
        plan.inverse(cre, cim, ore, oim)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    plan.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    stages = plan.stage_times()
    plan.set_timing(False)
    value = F * args.steps / dt

    # ---- errors of the timed form against the FP64 form, same inputs --------
    # forward: grids gre -> coefficients; inverse: coefficients c0 -> grids
    ref_cr, ref_ci = torch.empty_like(c0r), torch.empty_like(c0i)
    p64.forward(gre, gim, ref_cr, ref_ci)
    plan.forward(gre, gim, cre, cim)
    plan.inverse(c0r, c0i, ore, oim)
    torch.cuda.synchronize()

    def rel(ar, ai, br, bi):
        return float((torch.complex(ar, ai) - torch.complex(br, bi)).abs().max() /
                     torch.complex(br, bi).abs().max())

    errors = {
        "forward_vs_fp64": rel(cre, cim, ref_cr, ref_ci),
        "inverse_vs_fp64": rel(ore, oim, gre, gim),
        "fp64_round_trip": rel(ref_cr, ref_ci, c0r, c0i),
    }
    # the timed form's own round trip, coefficients -> grids -> coefficients
    plan.forward(ore, oim, cre, cim)
    torch.cuda.synchronize()
    errors["round_trip"] = rel(cre, cim, c0r, c0i)

    # ---- the other batched forms on the same fields ---------------------------
    def pairs_per_s(pl, nsteps):
        pl.forward(gre, gim, cre, cim)
        pl.inverse(cre, cim, ore, oim)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(nsteps):
            pl.forward(gre, gim, cre, cim)
            pl.inverse(cre, cim, ore, oim)
        torch.cuda.synchronize()
        return F * nsteps / (time.perf_counter() - t1)

    forms = {}
    if args.other_forms:
        if prec != S.FP64:
            forms["fp64"] = {"pairs_per_s": pairs_per_s(p64, 2)}
        for name, fprec in (("mixed", S.MIXED), ("split", S.SPLIT_BF16)):
            if fprec == prec:
                continue
            pl = S.Plan(bw, F, precision=fprec)
            v = pairs_per_s(pl, args.steps)
            pl.forward(gre, gim, cre, cim)
            pl.inverse(c0r, c0i, ore, oim)
            torch.cuda.synchronize()
            forms[name] = {"pairs_per_s": v, "forward_vs_fp64": rel(cre, cim, ref_cr, ref_ci),
                           "inverse_vs_fp64": rel(ore, oim, gre, gim)}
            pl.close()
        # FP64 with the naive branch for orders >= bw/4 (the reference's own
        # round-trip cutoff); the same quadrature as cutoff = bw
        pl = S.Plan(bw, F, precision=S.FP64, cutoff=bw // 4)
        v = pairs_per_s(pl, 2)
        pl.forward(gre, gim, cre, cim)
        pl.inverse(c0r, c0i, ore, oim)
        torch.cuda.synchronize()
        forms["fp64_cutoff_bw/4"] = {"pairs_per_s": v, "forward_vs_fp64": rel(cre, cim, ref_cr, ref_ci),
                                     "inverse_vs_fp64": rel(ore, oim, gre, gim)}
        pl.close()

    # ---- batch-1 FP64 form: one field per call -----------------------------
    plain = plain_loop = None
    if args.plain_fields > 0:
        pf = min(args.plain_fields, F)
        p1 = S.Plan(bw, 1, precision=S.FP64)
        for f in range(2):
            p1.forward(gre[f:f + 1], gim[f:f + 1], cre[f:f + 1], cim[f:f + 1])
            p1.inverse(cre[f:f + 1], cim[f:f + 1], ore[f:f + 1], oim[f:f + 1])
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for f in range(pf):
            p1.forward(gre[f:f + 1], gim[f:f + 1], cre[f:f + 1], cim[f:f + 1])
            p1.inverse(cre[f:f + 1], cim[f:f + 1], ore[f:f + 1], oim[f:f + 1])
        torch.cuda.synchronize()
        plain_loop = pf / (time.perf_counter() - t1)
        views = [(gre[f:f + 1], gim[f:f + 1], cre[f:f + 1], cim[f:f + 1], ore[f:f + 1], oim[f:f + 1])
                 for f in range(pf)]

        def per_field():
            for g0, g1, c0, c1, o0, o1 in views:
                p1.forward(g0, g1, c0, c1)
                p1.inverse(c0, c1, o0, o1)
        plain = pf / graph_replay_time(per_field, 5)
        p1.close()

    # ---- the drop-ins on host buffers (PCIe included) ------------------------
    dropin = dropin_latency(args.dropin_reps) if args.dropin_reps > 0 else None

    # ---- one field at bw=512: table streaming and per-form error -------------
    single = single_field_bw512(dev, args.single_pairs) if args.single_pairs > 0 else None

    # ---- a grid that is not a power of two ------------------------------------
    npow2 = non_pow2(dev, args.non_pow2_steps) if args.non_pow2_steps > 0 else None

    # ---- roofline of the dominant kernel ----------------------------------
    elem = 16 if prec == S.FP64 else 8  # bytes per complex intermediate value
    dom = max(stages, key=lambda k: stages[k][0])
    dom_ms, dom_n = stages[dom]
    avg_s = dom_ms / dom_n / 1e3
    achieved = stage_bytes(dom, bw, F, elem, args.precision) / avg_s
    roofline = {
        "bound": "hbm",
        "kernel": kernel_map(args.precision, bw, F).get(dom, dom),
        "stage": dom,
        "achieved": achieved / 1e9,
        "peak": HBM_PEAK / 1e9,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK,
        "traffic": measured_traffic(args.precision, bw, F, dom),
        "algorithmic_bytes": stage_bytes(dom, bw, F, elem, args.precision),
        "avg_launch_ms": dom_ms / dom_n,
        "stages_ms_per_step": {k: v[0] / args.steps for k, v in stages.items()},
    }

    # ---- MFMA utilisation of the Legendre GEMMs ---------------------------
    peak, issue = MFMA_PEAK[args.precision]
    kmap = kernel_map(args.precision, bw, F)
    mfma = {}
    for st_ in ("leg_fwd", "leg_inv"):
        if st_ in stages:
            ms, cnt = stages[st_]
            useful = legendre_flops(bw, F, args.precision) / (ms / cnt / 1e3) / 1e12
            # the seminaive count (cosine-series table MACs) is the transform's
            # algorithmic work whatever the form; the naive forms do 4 bw^3 per
            # field instead, more MACs for the same coefficients
            algo = legendre_flops(bw, F, "fp64") / (ms / cnt / 1e3) / 1e12
            mfma[st_] = {"kernel": kmap.get(st_, st_), "useful_tflops": useful, "issued_tflops": useful * issue,
                         "peak_tflops": peak, "frac": useful * issue / peak,
                         "seminaive_algorithmic_tflops": algo, "seminaive_algorithmic_frac": algo / peak}

    out = {
        "metric": "fwd+inv SHT pairs/s, bw=256 x1024 batch: mixed vs FP64, rel err<=1e-6",
        "value": value,
        "unit": "fwd+inv pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32", "mixed": "f32", "split": "f32"}[args.precision],
        "data": "synthetic",
        "config": {"workload": f"batched fwd+inv SHT, bw={bw}, {F} fields, {n}x{n} complex FP64 grids, "
                               "power-law (l+1)^-2 coefficients", "bw": bw, "fields": F,
                   "precision": args.precision, "parallelism": "single GPU"},
        "errors": errors,
        "other_forms": forms,
        "speedup_vs_fp64_batched": (value / forms["fp64"]["pairs_per_s"]) if "fp64" in forms else None,
        "fp64_batch1_pairs_per_s": plain,
        "fp64_batch1_python_loop_pairs_per_s": plain_loop,
        "speedup_vs_fp64_batch1": (value / plain) if plain else None,
        "single_field_bw512": single,
        "non_pow2": npow2,
        "dropin_host_buffers": dropin,
        "roofline": roofline,
        "mfma": mfma,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    plan.close()
    p64.close()


if __name__ == "__main__":
    main()
