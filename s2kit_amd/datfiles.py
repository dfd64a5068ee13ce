"""Text data files of S2kit's example programs, and command-line drivers
that run the drop-ins on them.

Formats (written by the reference's example programs):
  samples, "code"-order coefficients  interleaved re/im, one "%.15f" value per
      line (test_s2_semi_memo_fwd.c:145-147, test_s2_semi_memo_inv.c:145-147);
      samples are latitude-major (2bw x 2bw), coefficients in
      IndexOfHarmonicCoeff order
  "human"-order coefficients  "l = %d\\t m = %d\\t %.15f + %.15f I" per (l, m),
      l ascending, m from -l to l (test_s2_semi_memo_fwd.c:148-153)
  convolution signal, filter, result  one real value per line, the result as
      "%.16f" (test_conv_semi_memo.c:70-98, test_conv_semi_fly.c:68-96)
Readers take whitespace-separated numbers, as fscanf("%lf") does, and fail if
the file holds fewer values than the transform needs.

Drivers (same argument order as the reference's programs; the transforms run
through the GPU drop-ins):
  python -m s2kit_amd.datfiles fwd  sample_file output_file bw [output_format]
  python -m s2kit_amd.datfiles inv  coeffs_file output_file bw
  python -m s2kit_amd.datfiles conv signal_file filter_file output_file bw
  python -m s2kit_amd.datfiles conv_fly signal_file filter_file output_file bw
fwd and inv use cutoff = bw and COMPLEX data, as the reference's programs do;
--cutoff C picks another cutoff.
"""
import argparse
import sys

import numpy as np


def read_values(path, count):
    """The first `count` whitespace-separated numbers of a text file."""
    with open(path) as fh:
        vals = np.array(fh.read().split()[:count], dtype=np.float64)
    if vals.size < count:
        raise ValueError(f"{path}: {vals.size} values, {count} needed")
    return vals


def read_interleaved(path, n):
    """n complex values stored as re, im pairs; returns (re, im)."""
    v = read_values(path, 2 * n)
    return np.ascontiguousarray(v[0::2]), np.ascontiguousarray(v[1::2])


def format_interleaved(re, im):
    re, im = np.asarray(re, dtype=np.float64).ravel(), np.asarray(im, dtype=np.float64).ravel()
    return "".join("%.15f\n%.15f\n" % (a, b) for a, b in zip(re.tolist(), im.tolist()))


def write_interleaved(path, re, im):
    with open(path, "w") as fh:
        fh.write(format_interleaved(re, im))


def format_human(rc, ic, bw, index):
    """Coefficients by degree then order; index(m, l) is IndexOfHarmonicCoeff."""
    out = []
    for l in range(bw):
        for m in range(-l, l + 1):
            k = index(m, l, bw)
            out.append("l = %d\t m = %d\t %.15f + %.15f I\n" % (l, m, rc[k], ic[k]))
    return "".join(out)


def format_real(values):
    return "".join("%.16f\n" % v for v in np.asarray(values, dtype=np.float64).ravel().tolist())


def write_real(path, values):
    with open(path, "w") as fh:
        fh.write(format_real(values))


def main(argv=None):
    import s2kit_amd as S

    ap = argparse.ArgumentParser(prog="python -m s2kit_amd.datfiles",
                                 description="S2kit example programs on text data files (GPU drop-ins)")
    sub = ap.add_subparsers(dest="cmd", required=True)
    f = sub.add_parser("fwd", help="samples -> coefficients (test_s2_semi_memo_fwd)")
    f.add_argument("sample_file")
    f.add_argument("output_file")
    f.add_argument("bw", type=int)
    f.add_argument("output_format", type=int, nargs="?", default=0, help="0 code order, 1 human order")
    f.add_argument("--cutoff", type=int, default=None)
    i = sub.add_parser("inv", help="coefficients -> samples (test_s2_semi_memo_inv)")
    i.add_argument("coeffs_file")
    i.add_argument("output_file")
    i.add_argument("bw", type=int)
    i.add_argument("--cutoff", type=int, default=None)
    for name in ("conv", "conv_fly"):
        c = sub.add_parser(name, help="convolution (test_conv_semi_memo / test_conv_semi_fly)")
        c.add_argument("signal_file")
        c.add_argument("filter_file")
        c.add_argument("output_file")
        c.add_argument("bw", type=int)
    a = ap.parse_args(argv)
    bw = a.bw
    if bw < 2 or bw > 1024:
        ap.error("bw must be in [2, 1024]")
    n = 2 * bw
    if a.cmd == "fwd":
        rd, idd = read_interleaved(a.sample_file, n * n)
        tabs = S.HostTables(bw, a.cutoff)
        rc, ic = S.fst(rd, idd, bw, tabs)
        with open(a.output_file, "w") as fh:
            fh.write(format_interleaved(rc, ic) if a.output_format == 0 else
                     format_human(rc, ic, bw, lambda m, l, b: S.lib().IndexOfHarmonicCoeff(m, l, b)))
    elif a.cmd == "inv":
        rc, ic = read_interleaved(a.coeffs_file, bw * bw)
        tabs = S.HostTables(bw, a.cutoff)
        rd, idd = S.inv_fst(rc, ic, bw, tabs)
        write_interleaved(a.output_file, rd, idd)
    else:
        sig = read_values(a.signal_file, n * n)
        filt = read_values(a.filter_file, n * n)
        z = np.zeros(n * n)
        rr, _ = S.conv(sig, z, filt, z, bw, fly=(a.cmd == "conv_fly"))
        write_real(a.output_file, rr)
    return 0


if __name__ == "__main__":
    sys.exit(main())
