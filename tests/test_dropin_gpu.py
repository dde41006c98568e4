"""Runs the C++ drop-in API test program (tests/cpp/dropin_test.cpp, built by
__graft_entry__.build()): rawengine::ImageGraph / Renderer / nodes on the GPU
checked against the CPU oracle."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "dropin_test")


@pytest.mark.gpu
def test_cpp_dropin_api():
    assert os.path.exists(BIN), "tests/cpp/dropin_test is not built (run __graft_entry__.build())"
    p = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(p.stdout)
    print(p.stderr)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "ALL PASSED" in p.stdout


def test_cpp_dropin_binary_links_product_libraries():
    """The drop-in program links librawengine_amd.so -> librawops_hip.so (no GPU call)."""
    if not os.path.exists(BIN):
        pytest.skip("not built")
    out = subprocess.run(["ldd", BIN], capture_output=True, text=True).stdout
    assert "librawengine_amd.so" in out and "librawops_hip.so" in out
