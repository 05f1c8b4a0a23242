"""sptolDevMTTKRPNoAlloc (the entry point for callers that cannot allocate
device memory) against sptolDevMTTKRP on the same prepared tensor: bit for
bit, on two calls in a row, the second on a smaller tensor, for 3- and
4-mode tensors, FP32 and BF16 factors; no device memory is allocated by
the call; a tensor needing more ranges than the fixed scratch holds is
rejected; an unprepared mode (no scratch) matches too.  The first call's
result is also held to the FP64 oracle within 1e-5."""
import ctypes as C

import numpy as np
import pytest

import oracle
from conftest import rel_fro
from gpu_util import DevFactors, DevOut
from sptol_amd import gen
from sptol_amd.parti import (SPTOL_BF16, SPTOL_F32, SPTERR_VALUE_ERROR, CooTensor, check, ptr, sptIndex,
                             sptolDevMatrix)

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _both(lib, d, inds, vals, R, mode, dtype, prepare=True):
    coo = CooTensor(d, inds, vals)
    h = C.c_void_p()
    check(lib, lib.sptolDevCooCreate(C.byref(h), coo.ref()), "DevCooCreate")
    if prepare:
        check(lib, lib.sptolDevCooPrepareMode(h, mode), "PrepareMode")
    A = gen.uniform_factors(d, R, seed=int(d[0]) + R)
    F = DevFactors(lib, A, dtype)
    o = np.array([(mode + k) % len(d) for k in range(len(d))], dtype=np.uint32)
    outs = []
    for fn in ("sptolDevMTTKRP", "sptolDevMTTKRPNoAlloc"):
        out = DevOut(lib, int(d[mode]), R)
        check(lib, lib.sptolHipStreamSynchronize(None), "sync")
        before = lib.sptolHipMemoryInUse()
        check(lib, getattr(lib, fn)(h, F.arr, ptr(o, sptIndex), mode, C.byref(out.d), None), fn)
        if fn.endswith("NoAlloc"):
            assert lib.sptolHipMemoryInUse() == before, "the no-alloc call allocated device memory"
        check(lib, lib.sptolHipStreamSynchronize(None), "sync")
        outs.append(out.fetch())
        out.free()
    F.free()
    lib.sptolDevCooFree(h)
    return outs, A, o


@pytest.mark.parametrize("dims,R,dtype", [([3000, 2000, 4000], 32, SPTOL_F32), ([500, 400, 300, 60], 16, SPTOL_F32),
                                          ([3000, 2000, 4000], 16, SPTOL_BF16)])
def test_no_alloc_matches_dev_mttkrp(gpu_lib, dims, R, dtype):
    big = gen.zipf_tensor(dims, 400000, 1.1, seed=3)
    small = gen.zipf_tensor([x // 3 for x in dims], 20000, 1.1, seed=4)
    for k, (d, inds, vals) in enumerate((big, small)):
        for mode in range(len(d)):
            _say(f"dims {list(d)} R {R} dtype {dtype} mode {mode} call {k}")
            (ref, got), A, o = _both(gpu_lib, d, inds, vals, R, mode, dtype)
            assert np.array_equal(ref.view(np.uint32), got.view(np.uint32)), (k, mode)
            if k == 0 and dtype == SPTOL_F32:
                want = oracle.mttkrp_coo(inds, vals, A, o, mode, dtype=np.float64)
                assert rel_fro(got[:, :R], want[:, :R]) < 1e-5
                assert np.all(got[:, R:] == 0)


def test_no_alloc_unprepared_mode(gpu_lib):
    d, inds, vals = gen.uniform_tensor([700, 500, 600], 50000, seed=8)
    (ref, got), A, o = _both(gpu_lib, d, inds, vals, 16, 1, SPTOL_F32, prepare=False)
    want = oracle.mttkrp_coo(inds, vals, A, o, 1, dtype=np.float64)
    assert rel_fro(got[:, :16], want[:, :16]) < 1e-5
    assert rel_fro(ref[:, :16], want[:, :16]) < 1e-5


def test_no_alloc_rejects_too_many_ranges(gpu_lib):
    """A mode of 2^28 + 1 rows needs more than 65536 ranges (one per 4096
    rows): rejected before any launch."""
    lib = gpu_lib
    d = np.array([(1 << 28) + 1, 50, 40], dtype=np.uint32)
    rng = np.random.default_rng(5)
    inds = [rng.integers(0, int(x), size=1000).astype(np.uint32) for x in d]
    vals = rng.random(1000, dtype=np.float32) + 0.1
    coo = CooTensor(d, inds, vals)
    h = C.c_void_p()
    check(lib, lib.sptolDevCooCreate(C.byref(h), coo.ref()), "DevCooCreate")
    check(lib, lib.sptolDevCooPrepareMode(h, 0), "PrepareMode")
    R = 8
    # device matrices of the right shapes, left unfilled: the call must fail
    # before any kernel reads them
    mats = (sptolDevMatrix * 3)()
    for m in range(3):
        check(lib, lib.sptolDevMatrixCreate(C.byref(mats[m]), int(d[m]), R, SPTOL_F32), "create")
    out = sptolDevMatrix()
    check(lib, lib.sptolDevMatrixCreate(C.byref(out), int(d[0]), R, SPTOL_F32), "create")
    o = np.array([0, 1, 2], dtype=np.uint32)
    assert lib.sptolDevMTTKRPNoAlloc(h, mats, ptr(o, sptIndex), 0, C.byref(out), None) == SPTERR_VALUE_ERROR
    lib.sptClearLastError()
    lib.sptolDevMatrixFree(C.byref(out))
    for m in range(3):
        lib.sptolDevMatrixFree(C.byref(mats[m]))
    lib.sptolDevCooFree(h)
