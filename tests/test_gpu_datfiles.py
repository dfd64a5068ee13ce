"""The example-program drivers (python -m s2kit_amd.datfiles) on text files
made from the golden fixtures: the drop-ins' results through the reference's
file formats, against the oracle and the reference's conv goldens."""
import os

import numpy as np
import pytest

from s2kit_amd import datfiles

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def S():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    import s2kit_amd
    s2kit_amd.lib()
    return s2kit_amd


def crel(ar, ai, br, bi):
    a = np.asarray(ar) + 1j * np.asarray(ai)
    b = np.asarray(br) + 1j * np.asarray(bi)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("name,bw,cutoff", [("y31_bw8", 8, None), ("yMix_bw17", 17, None), ("y43_bw23", 23, 6)])
def test_fwd_then_inv_files(S, O, tmp_path, name, bw, cutoff):
    # test_s2_semi_memo_fwd / _inv on a known-answer grid: the coefficients
    # file matches the oracle's forward of the same samples (to the 15 written
    # decimals), and the inverse of that file gives the samples back
    print(f"datfiles fwd/inv {name} cutoff={cutoff}", flush=True)
    d = np.load(os.path.join(GOLDEN, f"y_{name}.npy"))
    src, coef, back = tmp_path / "in.dat", tmp_path / "coef.dat", tmp_path / "back.dat"
    datfiles.write_interleaved(src, d[0], d[1])
    extra = [] if cutoff is None else ["--cutoff", str(cutoff)]
    assert datfiles.main(["fwd", str(src), str(coef), str(bw)] + extra) == 0
    rc, ic = datfiles.read_interleaved(coef, bw * bw)
    rd, idd = datfiles.read_interleaved(src, 4 * bw * bw)
    orc, oic = O.fst(rd, idd, bw, cutoff=cutoff)
    assert np.abs((rc + 1j * ic) - (orc + 1j * oic)).max() <= 1e-14 + 1e-11 * np.abs(orc + 1j * oic).max()
    assert datfiles.main(["inv", str(coef), str(back), str(bw)] + extra) == 0
    br, bi = datfiles.read_interleaved(back, 4 * bw * bw)
    ordd, oidd = O.inv_fst(rc, ic, bw, cutoff=cutoff)
    assert crel(br, bi, ordd, oidd) < 1e-11
    if name == "y31_bw8":  # a band-limited grid: the round trip is the identity
        assert crel(br, bi, rd, idd) < 1e-12


def test_fwd_human_order(S, O, tmp_path):
    d = np.load(os.path.join(GOLDEN, "y_y31_bw8.npy"))
    src, out = tmp_path / "in.dat", tmp_path / "h.dat"
    datfiles.write_interleaved(src, d[0], d[1])
    assert datfiles.main(["fwd", str(src), str(out), "8", "1"]) == 0
    lines = out.read_text().splitlines()
    assert len(lines) == 64
    orc, oic = O.fst(d[0], d[1], 8)
    for line in lines:
        head, val = line.split("\t ", 2)[:2], line.split("\t ")[2]
        l, m = int(head[0].split("=")[1]), int(head[1].split("=")[1])
        re, im = val[:-2].split(" + ")
        k = O.index_of_coeff(m, l, 8)
        assert abs(float(re) - orc[k]) < 1e-13 and abs(float(im) - oic[k]) < 1e-13


@pytest.mark.parametrize("bw,cmd", [(64, "conv"), (64, "conv_fly"), (128, "conv")])
def test_conv_files_match_the_goldens(S, tmp_path, bw, cmd):
    # test_conv_semi_memo / _fly: signal and filter files -> result file; the
    # reference's result (o{bw}_conv_semi_memo_original.dat) to 1e-14
    print(f"datfiles {cmd} bw={bw}", flush=True)
    n = 2 * bw
    sig = np.load(os.path.join(GOLDEN, f"conv{bw}_signal.npy"))
    ring = np.load(os.path.join(GOLDEN, f"conv{bw}_filter.npy"))
    gold = np.load(os.path.join(GOLDEN, f"conv{bw}_out.npy"))
    s, f, o = tmp_path / "s.dat", tmp_path / "f.dat", tmp_path / "o.dat"
    datfiles.write_real(s, sig)
    datfiles.write_real(f, np.repeat(ring[:, None], n, axis=1))
    assert datfiles.main([cmd, str(s), str(f), str(o), str(bw)]) == 0
    out = datfiles.read_values(o, n * n)
    assert np.abs(out - gold).max() / np.abs(gold).max() < 1e-13
