"""bench.py pieces that run without a GPU: the one-GPU refusal, the
algorithmic byte counts behind roofline.achieved, and the HBM traffic read
from the committed rocprofv3 summaries."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


@pytest.mark.parametrize("bw", [64, 256, 255])
def test_stage_bytes(bw):
    n, rows = 2 * bw, 2 * bw - 1
    # mixed form: FP64 complex grids (16 B) in and out, FP32 complex intermediate (8 B)
    assert bench.stage_bytes("phi_fwd", bw, 1, 8, "mixed") == 16 * n * n + 8 * n * n
    assert bench.stage_bytes("leg_fwd", bw, 3, 8, "mixed") == 3 * (8 * n * rows + 16 * bw * bw)
    assert bench.stage_bytes("phi_inv", bw, 1, 8, "mixed") == 8 * n * rows + 16 * n * n
    # FP64 seminaive form: the DCT rows hold bw complex values
    assert bench.stage_bytes("dct_fwd", bw, 1, 16, "fp64") == 16 * n * rows + 16 * bw * rows
    assert bench.stage_bytes("leg_inv", bw, 1, 16, "fp64") == 16 * bw * bw + 16 * bw * rows


def test_traffic_from_committed_profiles():
    # the default workload's dominant stages, within 10% of their algorithmic bytes
    for stage in ("phi_fwd", "phi_inv"):
        t = bench.measured_traffic("mixed", 256, 1024, stage)
        a = bench.stage_bytes(stage, 256, 1024, 8, "mixed")
        assert t is not None and abs(t / a - 1) < 0.1, (stage, t, a)
    assert bench.measured_traffic("mixed", 256, 7, "phi_fwd") is None  # no profile for that workload


@pytest.mark.parametrize("bw,pair_flops", [(8, 1504), (64, 707328), (128, 5625344), (256, 44870656),
                                           (512, 358438912)])
def test_legendre_flops_per_pair(bw, pair_flops):
    # seminaive Legendre flops per fwd+inv pair: 8 (2 T(B) - TableSize(0, B)),
    # tabulated from the TableSize formula (cospml.c:39-59) independently
    assert 2 * bench.legendre_flops(bw, 1, "fp64") == pair_flops
    assert bench.legendre_flops(bw, 3, "mixed") == 3 * 4 * bw ** 3


def test_kernel_map_names_the_launched_gemms():
    # the kernels the host launches per form, bandwidth and field count
    # (s2_fast.hip kc_ok / kc_threads, s2_kernels.hip fft_cfg), as the
    # roofline object and the profile summaries name them
    k = bench.kernel_map("mixed", 256, 1024)
    assert (k["leg_fwd"], k["leg_inv"]) == ("k_mix_fwd", "k_mix_inv")
    k = bench.kernel_map("fp64", 256, 1024)
    assert (k["leg_fwd"], k["leg_inv"]) == ("k_gemm_fwd_kc<double, 512>", "k_gemm_inv_kc<double, 512>")
    assert bench.kernel_map("fp32", 512, 128)["leg_fwd"] == "k_gemm_fwd_kc<float, 1024>"
    # fewer than 64 fields, or bw below 128: the whole-row GEMMs
    assert bench.kernel_map("fp64", 256, 32)["leg_fwd"].startswith("k_gemm_fwd<double, false")
    assert bench.kernel_map("fp64", 64, 1024)["leg_fwd"].startswith("k_gemm_fwd<double, false")
    # bw=255: 510 = 2*3*5*17 rings on the register mixed-radix instance;
    # bw=263: 526 = 2*263, Bluestein over 1080 = 2^3 3^3 5 points
    k = bench.kernel_map("fp64", 255, 256)
    assert k["phi_fwd"] == "k_phi_fwd<double2, 1>" and k["leg_fwd"] == "k_gemm_fwd_kc<double, 512>"
    assert bench.kernel_map("fp64", 263, 256)["phi_fwd"] == "k_phi_fwd<double2, 3>"
    assert bench.kernel_map("fp64", 263, 256)["leg_fwd"] == "k_gemm_fwd_kc<double, 1024>"
    # the committed profile summaries name the same kernels
    for form, bw, f in (("fp64", 256, 1024), ("fp64", 255, 256), ("fp64", 263, 256)):
        text = open(os.path.join(ROOT, "profiles", f"kernel_stats_{form}_bw{bw}_f{f}.txt")).read()
        for st in ("leg_fwd", "leg_inv"):
            assert bench.kernel_map(form, bw, f)[st] in text, (form, bw, st)
