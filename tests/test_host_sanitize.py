"""The library's host code under ThreadSanitizer and under AddressSanitizer
with UBSan, on the CPU: the block cache, pinned pools, CPD jobs, batched
launches, kernel timing and the CPD server's admission, compute and release
threads, built for the host only against a CPU stand-in of the HIP runtime
(tests/host_sanitize/hip_stub.cpp: streams are worker threads, kernels are
no-ops, use of destroyed streams or events is reported).  The server stress
driver runs 8 admission workers with 96 slots (the configuration the GPU
build refuses) and the default 4 workers with 64 slots, with requests from
several client threads, file submissions, an injected allocation failure,
budget rejections, concurrent device-resident solves and kernel timing
switched on.  Passing means: no data race, no memory or UB error, no leak
beyond the process-lifetime pinned pools, no misuse of runtime handles, and
every request ended as expected."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

HERE = os.path.join(ROOT, "tests", "host_sanitize")
REPORTS = ("WARNING: ThreadSanitizer", "ERROR: AddressSanitizer", "ERROR: LeakSanitizer", "runtime error:",
           "HIPSTUB misuse")


@pytest.mark.skipif(shutil.which("/opt/rocm/bin/hipcc") is None, reason="hipcc not available")
@pytest.mark.parametrize("variant", ["tsan", "asan"])
def test_server_host_code_under_sanitizer(variant):
    jobs = str(min(8, os.cpu_count() or 1))
    b = subprocess.run(["make", "-C", HERE, f"V={variant}", "-j", jobs], capture_output=True, text=True, timeout=900)
    assert b.returncode == 0, b.stderr[-3000:]
    env = dict(os.environ)
    env["TSAN_OPTIONS"] = "halt_on_error=1 second_deadlock_stack=1"
    env["ASAN_OPTIONS"] = "halt_on_error=1 detect_stack_use_after_return=1"
    env["UBSAN_OPTIONS"] = "halt_on_error=1 print_stacktrace=1"
    env["LSAN_OPTIONS"] = "suppressions=" + os.path.join(HERE, "lsan.supp")
    for admit, slots, nreq in (("8", "96", "300"), ("4", "64", "150")):
        r = subprocess.run([os.path.join(HERE, f"serve_stress_{variant}"), admit, slots, nreq],
                           capture_output=True, text=True, timeout=600, env=env)
        log = r.stdout + r.stderr
        found = [k for k in REPORTS if k in log]
        assert r.returncode == 0 and not found, (variant, admit, slots, found, log[-4000:])
        assert "0 unexpected, 0 runtime misuses" in log
