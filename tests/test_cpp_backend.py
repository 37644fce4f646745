"""Runs the C++ test program (tests/cpp/mi355x_backend_test.cpp), which exercises the MI355X backend through
the knp::core Backend / MessageBus / BackendLoader API the way the reference's GTest suite does."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "knp_amd", "bin", "mi355x_backend_test")
PLUGIN = os.path.join(ROOT, "knp_amd", "libknp-mi355x-backend.so")


def run(*extra, timeout=120):
    assert os.path.exists(BIN) and os.path.exists(PLUGIN), "run __graft_entry__.build() first"
    r = subprocess.run([BIN, PLUGIN, *extra], capture_output=True, text=True, timeout=timeout)
    print(r.stdout, r.stderr)
    return r


def test_cpp_cpu_cases():
    r = run("--cpu-only")
    assert r.returncode == 0, r.stdout
    assert r.stdout.count("PASS") == 3
    assert "PASS MultiThreadedTypesGetting" in r.stdout


@pytest.mark.gpu
def test_cpp_backend_cases():
    r = run()
    assert r.returncode == 0, r.stdout
    for name in ("SmallestNetwork", "AdditiveSTDPBLIFATNetwork", "ResourceSTDPBLIFATNetwork",
                 "NeuronsAndSynapsesGetting", "NetworkDataThroughConstBackend", "SelectDevice", "BackendLoader",
                 "StartWithStopPredicate", "MultiThreadedTypesGetting", "MultiThreadedSmallestNetwork",
                 "MultiThreadedSmallestResourceNetwork", "MultiThreadedRepeatedIndex"):
        assert f"PASS {name}" in r.stdout
