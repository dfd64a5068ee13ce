"""Text data-file readers and writers of the example-program formats
(s2kit_amd/datfiles.py), on the CPU.

The "%.16f" writer is pinned byte for byte against the reference's own
convolution goldens (data/o64_conv_semi_memo_original.dat and
o128_..., written by test_conv_semi_memo.c:96-98): the golden fixtures hold
their parsed values, and the SHA-256 of each original file is recorded here.
"""
import hashlib
import os

import numpy as np
import pytest

from s2kit_amd import datfiles

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
CONV_SHA256 = {
    64: "0ec1ea7f3f146e736a6f2205abd52d45945e0a701eec66f8baa002ec7ee38cc0",
    128: "acc99d0b788b2ed213582c71ebb8f7ecc358cd064d9b3f8ab29f7ca042039750",
}


@pytest.mark.parametrize("bw", [64, 128])
def test_real_writer_reproduces_the_conv_goldens(bw):
    out = np.load(os.path.join(GOLDEN, f"conv{bw}_out.npy"))
    text = datfiles.format_real(out)
    assert text.count("\n") == (2 * bw) ** 2
    assert hashlib.sha256(text.encode()).hexdigest() == CONV_SHA256[bw]


def test_interleaved_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    re, im = rng.uniform(-2, 2, 50), rng.uniform(-2, 2, 50)
    re[0], im[0] = -0.0, 1e-17
    p = tmp_path / "c.dat"
    datfiles.write_interleaved(p, re, im)
    lines = p.read_text().splitlines()
    assert len(lines) == 100 and lines[0] == "-0.000000000000000" and lines[1] == "0.000000000000000"
    assert lines[2] == "%.15f" % re[1] and lines[3] == "%.15f" % im[1]
    r2, i2 = datfiles.read_interleaved(p, 50)
    # every value is what its text says; the text is the value to 15 decimals
    assert np.array_equal(r2, np.array([float(x) for x in lines[0::2]]))
    assert np.abs(r2 - re).max() <= 5e-16 and np.abs(i2 - im).max() <= 5e-16


def test_reader_takes_any_whitespace_and_rejects_short_files(tmp_path):
    p = tmp_path / "s.dat"
    p.write_text("1 2\n3\t4\n\n5")
    assert datfiles.read_values(p, 5).tolist() == [1, 2, 3, 4, 5]
    re, im = datfiles.read_interleaved(p, 2)
    assert re.tolist() == [1, 3] and im.tolist() == [2, 4]
    with pytest.raises(ValueError, match="6 needed"):
        datfiles.read_values(p, 6)


def test_known_answer_grids_parse(tmp_path):
    # the known-answer sample fixtures hold the parsed values of the
    # reference's y*.dat files; writing and reading them back keeps them to 1e-15
    d = np.load(os.path.join(GOLDEN, "y_y31_bw8.npy"))
    p = tmp_path / "y.dat"
    datfiles.write_interleaved(p, d[0], d[1])
    re, im = datfiles.read_interleaved(p, 16 * 16)
    assert np.abs(re - d[0]).max() <= 5e-16 and np.abs(im - d[1]).max() <= 5e-16


def test_human_order(O):
    bw = 4
    rc = np.arange(bw * bw, dtype=float)
    ic = -rc
    text = datfiles.format_human(rc, ic, bw, O.index_of_coeff)
    lines = text.splitlines()
    assert len(lines) == bw * bw
    assert lines[0] == "l = 0\t m = 0\t 0.000000000000000 + -0.000000000000000 I"
    k = O.index_of_coeff(-2, 3, bw)
    assert lines[9 + 1] == "l = 3\t m = -2\t %.15f + %.15f I" % (k, -k)
