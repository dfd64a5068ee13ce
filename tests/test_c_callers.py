"""C callers of the library, built with gcc against include/, and the
AddressSanitizer / UndefinedBehaviorSanitizer CPU builds (tests/c/Makefile).

* The integration guide's C example (its `round_trip` block) is
  tests/c/round_trip.c verbatim; it compiles as C11 with -Wall -Wextra
  -Werror and links against libs2kit_amd.so.  On the GPU its output (an
  InvFSTSemiMemo grid and the FSTSemiMemo coefficients of that grid) matches
  the oracle to 1e-11.
* The oracle (oracle/s2oracle.c) runs clean under ASan + UBSan at even and
  odd bandwidths, every cutoff kind and both data formats.
* The library's host code (argument checks, layout arithmetic) runs clean
  under ASan + UBSan, without a GPU.
"""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT

CDIR = os.path.join(ROOT, "tests", "c")
BUILD = os.path.join(CDIR, "_build")


def _make(*targets):
    r = subprocess.run(["make", "-C", CDIR, "-j8", *targets], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]


def test_integration_example_is_the_compiled_source():
    doc = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = [b for b in re.findall(r"```c\n(.*?)```", doc, flags=re.S) if "void round_trip" in b]
    assert len(blocks) == 1
    assert blocks[0] == open(os.path.join(CDIR, "round_trip.c")).read()


def test_c_example_builds_as_c11_with_warnings_as_errors():
    _make("all")
    assert os.access(os.path.join(BUILD, "round_trip"), os.X_OK)


def test_oracle_under_address_and_undefined_behaviour_sanitizers():
    _make(os.path.join(BUILD, "oracle_san"))
    r = subprocess.run([os.path.join(BUILD, "oracle_san")], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and "oracle sanitizer run: ok" in r.stdout, r.stdout[-3000:] + r.stderr[-3000:]


def test_library_host_code_under_address_and_undefined_behaviour_sanitizers():
    _make(os.path.join(BUILD, "host_checks_san"))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(BUILD, "host_checks_san")], capture_output=True, text=True, timeout=300,
                       env=env)
    assert r.returncode == 0 and "host checks: ok" in r.stdout, r.stdout[-3000:] + r.stderr[-3000:]


@pytest.mark.gpu
@pytest.mark.parametrize("bw,seed", [(64, 1), (23, 2), (128, 3)])
def test_c_round_trip_against_oracle(O, tmp_path, bw, seed):
    exe = os.path.join(BUILD, "round_trip")
    if not os.access(exe, os.X_OK):
        raise RuntimeError(f"{exe} is missing: build it with `make -C tests/c` (or __graft_entry__.build())")
    print(f"C round_trip bw={bw}", flush=True)
    rc, ic = O.make_coeffs(700 + seed, bw, 0)
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    np.concatenate([rc, ic]).tofile(inp)
    r = subprocess.run([exe, str(bw), str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    v = np.fromfile(out)
    ng, nc = 4 * bw * bw, bw * bw
    assert v.size == 2 * ng + 2 * nc
    rd, idd, rc2, ic2 = v[:ng], v[ng:2 * ng], v[2 * ng:2 * ng + nc], v[2 * ng + nc:]
    ord_, oid = O.inv_fst(rc, ic, bw)
    g = rd + 1j * idd
    og = ord_ + 1j * oid
    assert np.abs(g - og).max() < 1e-11 * np.abs(og).max()
    orc, oic = O.fst(rd, idd, bw)
    assert np.abs((rc2 + 1j * ic2) - (orc + 1j * oic)).max() < 1e-11 * np.abs(orc + 1j * oic).max()
