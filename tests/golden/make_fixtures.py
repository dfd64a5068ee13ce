"""Convert the reference's own data files into compact test fixtures.

Run in the build container (the reference tree is not on the GPU machine):

    python tests/golden/make_fixtures.py [/root/reference]

Sources (S2kit 1.1, data/ directory; described in its HowTo, PDF p.9):
  s64.dat, f64.dat, o64_conv_semi_memo_original.dat     conv golden, B=64
  s128.dat, f128.dat, o128_conv_semi_memo_original.dat  conv golden, B=128
  y20_bw8.dat, y31_bw8.dat, y43_bw23.dat, yMix_bw17.dat known-answer grids

Outputs are raw float64 arrays (.npy), bit-identical to the parsed text:
  conv{B}_signal.npy   (2B*2B) real signal samples
  conv{B}_filter.npy   (2B)    one value per latitude ring: the filter files
                               are zonal (every ring constant), checked here
  conv{B}_out.npy      (2B*2B) reference output rres ("%.16f" text)
  y_{name}.npy         (2, 2B*2B) re, im of the interleaved sample file
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref):
    data = os.path.join(ref, "data")
    for bw in (64, 128):
        n = 2 * bw
        s = np.loadtxt(os.path.join(data, f"s{bw}.dat"))
        f = np.loadtxt(os.path.join(data, f"f{bw}.dat")).reshape(n, n)
        o = np.loadtxt(os.path.join(data, f"o{bw}_conv_semi_memo_original.dat"))
        o_fly = np.loadtxt(os.path.join(data, f"o{bw}_conv_semi_fly_original.dat"))
        assert s.size == n * n and o.size == n * n
        assert np.array_equal(o, o_fly), "fly and memo goldens are expected to be identical"
        assert np.all(f == f[:, :1]), "filter is expected to be zonal (constant on every ring)"
        np.save(os.path.join(HERE, f"conv{bw}_signal.npy"), s)
        np.save(os.path.join(HERE, f"conv{bw}_filter.npy"), np.ascontiguousarray(f[:, 0]))
        np.save(os.path.join(HERE, f"conv{bw}_out.npy"), o)
    for name in ("y20_bw8", "y31_bw8", "y43_bw23", "yMix_bw17"):
        d = np.loadtxt(os.path.join(data, f"{name}.dat"))
        np.save(os.path.join(HERE, f"y_{name}.npy"), np.stack([d[0::2], d[1::2]]))
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
