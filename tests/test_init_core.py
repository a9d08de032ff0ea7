"""Host-side check of the device initialisation's arithmetic
(zpic_amd/csrc/kernels/init_core.h), compiled with g++ from
tests/init_core_check.cpp:

* jump-ahead of the reference generator (em2d/random.c:48-53) equals
  stepping it call by call;
* the double-double log is correctly rounded on 3.4 M arguments (the polar
  method's own and every binade of (0, 1)); it differs from the C
  library's log in about 0.06 % of them, and every such case is one where
  the library's value is the farther one from the quadruple-precision log;
* deviates drawn attempt-parallel (counted in chunks, prefix-summed, then
  emitted as pairs, as the device kernels do) equal the sequential polar
  method (em2d/random.c:67-101) when both use the library's log.
"""
from __future__ import annotations

import json
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
KERNELS = os.path.join(os.path.dirname(HERE), "zpic_amd", "csrc", "kernels")


def test_init_arithmetic_matches_the_generator_and_is_correctly_rounded(tmp_path):
    exe = str(tmp_path / "init_core_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", f"-I{KERNELS}",
                    os.path.join(HERE, "init_core_check.cpp"), "-o", exe, "-lquadmath", "-lm"], check=True)
    r = json.loads(subprocess.run([exe], check=True, capture_output=True, text=True).stdout.strip())
    print(r)
    assert r["jump_fail"] == 0
    assert r["logs"] > 3_000_000
    assert r["log_not_nearest"] == 0
    assert r["log_ambiguous"] == 0
    assert r["log_differs_from_libc"] < 0.001 * r["logs"]
    assert r["deviates"] > 100_000 and r["deviate_mismatch"] == 0
