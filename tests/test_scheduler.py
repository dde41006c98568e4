"""TileScheduler (include/rawengine/TileScheduler.hpp) through tests/cpp/scheduler_test.

CPU: queue semantics with a host-only probe node (priority order, FIFO,
pending budget, supersession, caller tokens, destruction) following the
reference's scheduler scenarios (edit_graph_tests.cpp:1773-1911).
GPU: batched dispatch of queued RAW renders equal to separate renders bit for
bit, four workers on four streams, TileCache clear racing a render.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "scheduler_test")


def run(mode):
    assert os.path.exists(BIN), "tests/cpp/scheduler_test is not built (run __graft_entry__.build())"
    p = subprocess.run([BIN, mode], capture_output=True, text=True, timeout=120)
    print(p.stdout)
    print(p.stderr)
    assert p.returncode == 0 and "ALL PASSED" in p.stdout, p.stdout + p.stderr


def test_scheduler_semantics_host_only():
    run("cpu")


@pytest.mark.gpu
def test_scheduler_device_batching_and_streams():
    run("gpu")
