"""The closed-form repeated float addition that the synaptic-resource training kernels use for runs of equal
free-resource terms (knp_amd/csrc/float_repeat.h) against the additions done one by one, bit for bit, on the
host: random magnitudes, constructed ties to even, binade and sign crossings, zeros, subnormals, overflow and
NaN, and the chunk replay with contributing lanes (tests/cpp/float_repeat_test.hip)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "knp_amd", "bin", "float_repeat_test")


def test_repeat_add_matches_additions_one_by_one():
    if not os.path.exists(BIN):
        pytest.fail("knp_amd/bin/float_repeat_test is not built (python knp_amd/build.py)")
    r = subprocess.run([BIN, "100000"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.startswith("ok "), r.stdout
    assert int(r.stdout.split()[1]) > 1_000_000 // 4
