"""The bench's command-line contract and its roofline arithmetic (no GPU
work): more than one GPU is refused, a machine without a GPU stops before
measuring, and the algorithmic bytes per launch add up as documented."""
import ctypes as C
import os
import subprocess
import sys

import pytest

from conftest import ROOT


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True,
                          text=True, timeout=300, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = _run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def test_no_gpu_stops_before_measuring(lib):
    n = C.c_int(0)
    if lib.sptolHipDeviceCount(C.byref(n)) == 0 and n.value > 0:
        pytest.skip("a GPU is visible")
    r = _run("--workload", "coo", "--steps", "1")
    assert r.returncode == 1
    assert "no GPU" in r.stderr
    assert r.stdout == ""


def test_algorithmic_bytes_per_launch():
    sys.path.insert(0, ROOT)
    import bench
    dims, U, nnz, R = [100, 50, 40], [90, 45, 30], 1000, 16
    # segmented kernel, mode 0: 16000 B of records, a 16 B row bitmap,
    # (45 + 30) distinct rows x 64 B gathered, 100 rows x 64 B written
    assert bench.alg_bytes("mttkrp_seg", nnz, dims, U, 0, R) == 16000 + 16 + 4800 + 6400
    # 4-D, mode 2, BF16 factors: 20-byte records, a 8 B bitmap, (90 + 45 + 20)
    # rows x 32 B gathered, 40 FP32 rows x 64 B written
    assert bench.alg_bytes("mttkrp_seg", nnz, [100, 50, 40, 30], [90, 45, 30, 20], 2, R,
                           factor_bytes=2) == 20000 + 8 + 4960 + 2560
    # atomic COO kernel, mode 1: u32 indices and a value per nonzero, the
    # touched rows written
    assert bench.alg_bytes("mttkrp_coo", nnz, dims, U, 1, R) == 16000 + 7680 + 2880
