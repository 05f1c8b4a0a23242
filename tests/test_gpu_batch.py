"""GPU parity of the batched persistent MTTKRP (one launch over many
tensors), plain, autotuned and replayed from a hipGraph: every output
matches the FP64 oracle within 1e-5 relative Frobenius error on every
replay, on the segmented path (mode-sorted views), the atomic path, and
both in one launch.  Replays start from poisoned outputs, so a row that a
replay fails to write is caught."""
import ctypes as C

import numpy as np
import pytest

import oracle
from conftest import rel_fro
from gpu_util import DevFactors, DevOut, bf16_round, dev_coo
from sptol_amd import gen
from sptol_amd.parti import SPTOL_BF16, SPTOL_F32, check, ptr, sptIndex

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _make_batch(lib, specs, prepare, dtype=SPTOL_F32):
    items = []
    for k, (dims, nnz, R, mode, seed) in enumerate(specs):
        d, inds, vals = gen.uniform_tensor(dims, nnz, seed=seed)
        N = len(d)
        A = gen.uniform_factors(d, R, seed=seed + 1)
        order = np.array([mode] + [(mode + i) % N for i in range(1, N)], np.uint32)
        h = dev_coo(lib, d, inds, vals)
        if prepare is True or (prepare == "mixed" and k % 2 == 0):
            check(lib, lib.sptolDevCooPrepareMode(h, mode), "prepare")
        F = DevFactors(lib, A, dtype)
        out = DevOut(lib, int(d[mode]), R)
        Au = [bf16_round(a) for a in A] if dtype == SPTOL_BF16 else A
        want = oracle.mttkrp_coo(inds, vals, Au, order, mode, dtype=np.float64)
        items.append(dict(h=h, F=F, out=out, order=order, mode=mode, want=want, R=R))
    b = C.c_void_p()
    check(lib, lib.sptolBatchCreate(C.byref(b)), "BatchCreate")
    for it in items:
        check(lib, lib.sptolBatchAdd(b, it["h"], it["F"].arr, ptr(it["order"], sptIndex), it["mode"],
                                     C.byref(it["out"].d)), "BatchAdd")
    return b, items


def _verify(items):
    for i, it in enumerate(items):
        got = it["out"].fetch()[: it["want"].shape[0]]
        R = it["R"]
        assert np.all(got[:, R:] == 0)
        e = rel_fro(got[:, :R], it["want"][:, :R])
        assert e <= 1e-5, (i, e)


def _free(lib, b, items):
    lib.sptolBatchFree(b)
    for it in items:
        it["F"].free()
        it["out"].free()
        lib.sptolDevCooFree(it["h"])


SPECS = [([2000, 1500, 1000], 200000, 16, 0, 1), ([300, 200, 100], 10000, 8, 1, 2),
         ([5000, 40, 9000], 150000, 32, 2, 3), ([100, 100, 100], 50000, 64, 0, 4),
         ([60, 50, 40, 30], 40000, 16, 3, 5), ([1000, 1000, 1000], 1, 8, 1, 6),
         ([70000, 3000, 500], 300000, 32, 1, 7), ([400, 300, 200], 20000, 5, 2, 8)]


@pytest.mark.parametrize("prepare", [False, True, "mixed"])
def test_batch_plain_and_autotuned(gpu_lib, prepare):
    """prepare=True: every tensor on the segmented path; False: the atomic
    path; "mixed": both paths in one launch."""
    lib = gpu_lib
    b, items = _make_batch(lib, SPECS, prepare)
    _say("batch plain", prepare)
    check(lib, lib.sptolBatchRun(b, None), "BatchRun")
    check(lib, lib.sptolHipStreamSynchronize(None), "sync")
    _verify(items)
    for cfg in ((0.25, 2), (4.0, 16)):
        check(lib, lib.sptolBatchSetConfig(b, cfg[0], cfg[1]), "SetConfig")
        check(lib, lib.sptolBatchRun(b, None), "BatchRun")
        check(lib, lib.sptolHipStreamSynchronize(None), "sync")
        _verify(items)
    _say("batch autotune", prepare)
    check(lib, lib.sptolBatchAutotune(b, 2, None), "Autotune")
    sc, oc = C.c_double(), C.c_int()
    lib.sptolBatchGetConfig(b, C.byref(sc), C.byref(oc))
    _say(f"  tuned tile_scale={sc.value} blocks_per_cu={oc.value}")
    check(lib, lib.sptolBatchRun(b, None), "BatchRun")
    check(lib, lib.sptolHipStreamSynchronize(None), "sync")
    _verify(items)
    _free(lib, b, items)


@pytest.mark.parametrize("prepare", [True, "mixed"])
def test_batch_bf16_factors(gpu_lib, prepare):
    """BF16-stored factors (8 columns a lane on the segmented path, R from 5
    to 64): outputs match the FP64 oracle run on the same BF16-rounded
    factors."""
    lib = gpu_lib
    b, items = _make_batch(lib, SPECS, prepare, SPTOL_BF16)
    _say("batch bf16", prepare)
    check(lib, lib.sptolBatchRun(b, None), "BatchRun")
    check(lib, lib.sptolHipStreamSynchronize(None), "sync")
    _verify(items)
    _free(lib, b, items)


@pytest.mark.parametrize("prepare", [True, "mixed"])
def test_batch_graph_replays(gpu_lib, prepare):
    lib = gpu_lib
    b, items = _make_batch(lib, SPECS, prepare)
    s = C.c_void_p()
    check(lib, lib.sptolHipStreamCreate(C.byref(s)), "StreamCreate")
    check(lib, lib.sptolBatchCapture(b, s), "Capture")
    for rep in range(4):
        _say("graph replay", rep)
        # eager work in between: a plain single-tensor run on the default stream
        it = items[0]
        check(lib, lib.sptolDevMTTKRP(it["h"], it["F"].arr, ptr(it["order"], sptIndex), it["mode"],
                                      C.byref(it["out"].d), None), "eager")
        check(lib, lib.sptolHipStreamSynchronize(None), "sync")
        for x in items:   # poison outputs so a replay that skips the fill is caught
            x["out"].h.buf[:] = 7.0
            check(lib, lib.sptolDevMatrixUpload(C.byref(x["out"].d), C.byref(x["out"].h.m)), "poison")
        check(lib, lib.sptolBatchRun(b, s), "replay")
        check(lib, lib.sptolHipStreamSynchronize(s), "sync")
        _verify(items)
    lib.sptolHipStreamDestroy(s)
    _free(lib, b, items)
