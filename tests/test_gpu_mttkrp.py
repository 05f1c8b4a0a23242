"""GPU parity of COO and HiCOO MTTKRP through the C ABI.

Bar: FP32 results within 1e-5 relative Frobenius error of the reference's
sequential FP32 output (golden fixtures / bit-exact oracle restatement); the
error against the FP64 oracle is asserted alongside.  Padding columns must
be exactly zero.  BF16-stored factors are compared with the FP64 oracle run
on the same BF16-rounded factors (FP32 accumulation bar 1e-5)."""
import ctypes as C
import sys

import numpy as np
import pytest

import oracle
from conftest import rel_fro
from gpu_util import DevFactors, DevOut, bf16_round, dev_coo, dev_mttkrp_coo, dev_mttkrp_hicoo
from sptol_amd import gen
from sptol_amd.parti import (CooTensor, Matrix, SPTOL_BF16, check, mats_array, ptr, sptIndex, to_hicoo)

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _dropin(lib, d, inds, vals, A, order, mode):
    coo = CooTensor(d, inds, vals)
    mats = [Matrix(a) for a in A] + [Matrix(nrows=int(max(d)), ncols=A[0].shape[1])]
    o = np.ascontiguousarray(order, dtype=np.uint32)
    check(lib, lib.sptMTTKRP(coo.ref(), mats_array(mats), ptr(o, sptIndex), mode), "sptMTTKRP")
    return mats[-1].buf[: int(d[mode])].copy()


def _check(got, ref32, ref64, R):
    assert np.all(got[:, R:] == 0), "padding columns must be zero"
    e32 = rel_fro(got[:, :R], ref32[:, :R])
    e64 = rel_fro(got[:, :R], ref64[:, :R])
    assert e32 <= TOL and e64 <= TOL, (e32, e64)
    return e32, e64


def test_dropin_coo_vs_reference_golden(gpu_lib, golden_coo):
    g = golden_coo
    for case in g.cases():
        d, inds, vals = g.tensor(case)
        R = int(g.get(case, "R"))
        A = [g.get(case, f"A{m}") for m in range(len(d))]
        for mode in range(len(d)):
            for k in range(2):
                _say("dropin coo", case, mode, k)
                order = g.get(case, f"order{mode}_{k}")
                got = _dropin(gpu_lib, d, inds, vals, A, order, mode)
                ref64 = oracle.mttkrp_coo(inds, vals, A, order, mode, dtype=np.float64)
                _check(got, g.get(case, f"out{mode}_{k}"), ref64, R)


@pytest.mark.parametrize("prepare", [False, True])
def test_device_coo_all_modes(gpu_lib, golden_coo, prepare):
    g = golden_coo
    for case in g.cases():
        d, inds, vals = g.tensor(case)
        R = int(g.get(case, "R"))
        A = [g.get(case, f"A{m}") for m in range(len(d))]
        h = dev_coo(gpu_lib, d, inds, vals)
        for mode in range(len(d)):
            _say("device coo", case, mode, prepare)
            order = g.get(case, f"order{mode}_0")
            got = dev_mttkrp_coo(gpu_lib, d, inds, vals, A, order, mode, prepare=prepare, h=h)
            ref64 = oracle.mttkrp_coo(inds, vals, A, order, mode, dtype=np.float64)
            _check(got, g.get(case, f"out{mode}_0"), ref64, R)
        gpu_lib.sptolDevCooFree(h)


@pytest.mark.parametrize("R", [1, 5, 8, 16, 32, 64, 100, 300])
def test_ranks_and_padding(gpu_lib, R):
    d, inds, vals = gen.zipf_tensor([400, 300, 500], 30000, 1.1, seed=R)
    A = gen.uniform_factors(d, R, seed=R + 1)
    for mode in range(3):
        _say("rank", R, mode)
        order = [mode, (mode + 2) % 3, (mode + 1) % 3]
        ref32 = oracle.mttkrp_coo(inds, vals, A, order, mode)
        ref64 = oracle.mttkrp_coo(inds, vals, A, order, mode, dtype=np.float64)
        for prepare in (False, True):
            got = dev_mttkrp_coo(gpu_lib, d, inds, vals, A, order, mode, prepare=prepare)
            _check(got, ref32, ref64, R)


@pytest.mark.parametrize("R", [1, 5, 8, 16, 24, 32, 64, 100])
def test_ranks_and_padding_bf16(gpu_lib, R):
    """BF16-stored factors at every lane shape of the segmented kernel (1, 2,
    4 and 8 lanes of 8 columns, idle lanes at R=24; R above 64 takes the
    atomic kernel), 3-D and 4-D, every mode, sorted view and not, against
    the oracle run on the same BF16-rounded factors."""
    for dims in ([400, 300, 500], [60, 50, 40, 30]):
        d, inds, vals = gen.zipf_tensor(dims, 30000, 1.1, seed=R + len(dims))
        A = gen.uniform_factors(d, R, seed=R + 2)
        Au = [bf16_round(a) for a in A]
        N = len(d)
        for mode in range(N):
            _say("rank bf16", R, dims, mode)
            order = [mode] + [(mode + i) % N for i in range(1, N)]
            ref32 = oracle.mttkrp_coo(inds, vals, Au, order, mode)
            ref64 = oracle.mttkrp_coo(inds, vals, Au, order, mode, dtype=np.float64)
            for prepare in (False, True):
                got = dev_mttkrp_coo(gpu_lib, d, inds, vals, A, order, mode, prepare=prepare, dtype=SPTOL_BF16)
                _check(got, ref32, ref64, R)


def test_higher_order_and_edge_cases(gpu_lib):
    # 5-D tensor (generic-order kernel path)
    d, inds, vals = gen.uniform_tensor([20, 15, 10, 12, 9], 20000, seed=3)
    A = gen.uniform_factors(d, 16, seed=4)
    for mode in range(5):
        _say("5d", mode)
        order = [mode] + [(mode + i) % 5 for i in range(1, 5)]
        got = dev_mttkrp_coo(gpu_lib, d, inds, vals, A, order, mode, prepare=(mode % 2 == 0))
        _check(got, oracle.mttkrp_coo(inds, vals, A, order, mode),
               oracle.mttkrp_coo(inds, vals, A, order, mode, dtype=np.float64), 16)
    # 2-D tensor
    d, inds, vals = gen.uniform_tensor([300, 200], 5000, seed=5)
    A = gen.uniform_factors(d, 8, seed=6)
    for mode in range(2):
        got = dev_mttkrp_coo(gpu_lib, d, inds, vals, A, [mode, 1 - mode], mode)
        _check(got, oracle.mttkrp_coo(inds, vals, A, [mode, 1 - mode], mode),
               oracle.mttkrp_coo(inds, vals, A, [mode, 1 - mode], mode, dtype=np.float64), 8)
    # empty tensor: the output is all zeros
    d = np.array([10, 20, 30], np.uint32)
    e = [np.zeros(0, np.uint32)] * 3
    A = gen.uniform_factors(d, 16, seed=7)
    got = _dropin(gpu_lib, d, e, np.zeros(0, np.float32), A, [1, 2, 0], 1)
    assert got.shape == (20, 16) and np.all(got == 0)
    # one nonzero, and many repeats of the same coordinate (unsorted)
    one = [np.array([3], np.uint32), np.array([7], np.uint32), np.array([11], np.uint32)]
    got = _dropin(gpu_lib, d, one, np.array([2.5], np.float32), A, [0, 1, 2], 0)
    want = np.zeros((10, 16), np.float32)
    want[3] = (np.float32(2.5) * A[1][7]) * A[2][11]
    assert np.array_equal(got, want)
    rng = np.random.default_rng(0)
    rep = [np.full(5000, 4, np.uint32), rng.integers(0, 20, 5000).astype(np.uint32),
           rng.integers(0, 30, 5000).astype(np.uint32)]
    rv = rng.random(5000).astype(np.float32)
    got = _dropin(gpu_lib, d, rep, rv, A, [0, 1, 2], 0)
    _check(got, oracle.mttkrp_coo(rep, rv, A, [0, 1, 2], 0),
           oracle.mttkrp_coo(rep, rv, A, [0, 1, 2], 0, dtype=np.float64), 16)


def test_c1_full_size(gpu_lib):
    """2000x1500x1000, 1e6 nnz, R=16, mode 0 (the reference driver's case)."""
    d, inds, vals = gen.uniform_tensor([2000, 1500, 1000], 1_000_000, seed=2024)
    A = gen.uniform_factors(d, 16, seed=2025)
    order = [0, 1, 2]
    ref32 = oracle.mttkrp_coo(inds, vals, A, order, 0)
    ref64 = oracle.mttkrp_coo(inds, vals, A, order, 0, dtype=np.float64)
    for prepare in (False, True):
        _say("c1 full", prepare)
        got = dev_mttkrp_coo(gpu_lib, d, inds, vals, A, order, 0, prepare=prepare)
        e32, e64 = _check(got, ref32, ref64, 16)
        _say(f"  rel err vs FP32 reference {e32:.2e}, vs FP64 oracle {e64:.2e}; "
             f"reference FP32 vs FP64 {rel_fro(ref32, ref64):.2e}")


def _launches(lib, name):
    ms, n = C.c_double(0), C.c_uint64(0)
    lib.sptolProfileQuery(name.encode(), C.byref(ms), C.byref(n))
    return int(n.value)


@pytest.mark.parametrize("R", [16, 32])
def test_whole_row_ranges_have_no_carry(gpu_lib, R):
    """A uniform tensor (rows of nearly equal length) takes ranges of whole
    rows: no row continues across ranges, so no carry pass is launched; the
    result meets the bar in FP32 and BF16, is deterministic, and the
    allocation-free entry point gives the same bits.  A Zipf tensor (a few
    very long rows) keeps fixed-size ranges and its carry pass."""
    lib = gpu_lib
    cases = [("uniform", gen.uniform_tensor([2000, 1500, 1000], 1_000_000, seed=2024), 0),
             ("zipf", gen.zipf_tensor([3000, 2000, 1000], 400_000, 1.1, seed=5), 1)]
    for name, (d, inds, vals), carries in cases:
        A = gen.uniform_factors(d, R, seed=2025)
        h = dev_coo(lib, d, inds, vals)
        check(lib, lib.sptolDevCooPrepareMode(h, 0), "PrepareMode")
        order = np.array([0, 1, 2], np.uint32)
        for dtype in (0, SPTOL_BF16):
            _say("whole-row ranges", name, R, dtype)
            Au = [bf16_round(a) for a in A] if dtype else A
            lib.sptolProfileEnable(1)
            lib.sptolProfileReset()
            try:
                got = dev_mttkrp_coo(lib, d, inds, vals, A, order, 0, dtype=dtype, h=h)
                assert _launches(lib, "mttkrp_seg") == 1
                assert _launches(lib, "mttkrp_seg_carry") == carries, name
            finally:
                lib.sptolProfileEnable(0)
            again = dev_mttkrp_coo(lib, d, inds, vals, A, order, 0, dtype=dtype, h=h)
            assert np.array_equal(got.view(np.uint32), again.view(np.uint32))
            F = DevFactors(lib, A, dtype)
            out = DevOut(lib, int(d[0]), R)
            check(lib, lib.sptolDevMTTKRPNoAlloc(h, F.arr, ptr(order, sptIndex), 0, C.byref(out.d), None), "NoAlloc")
            check(lib, lib.sptolHipStreamSynchronize(None), "sync")
            noalloc = out.fetch()
            F.free()
            out.free()
            assert np.array_equal(got.view(np.uint32), noalloc.view(np.uint32))
            _check(got, oracle.mttkrp_coo(inds, vals, Au, order, 0),
                   oracle.mttkrp_coo(inds, vals, Au, order, 0, dtype=np.float64), R)
        lib.sptolDevCooFree(h)


@pytest.mark.parametrize("R,dtype", [(16, 0), (32, SPTOL_BF16)])
def test_lds_factor_3d_matches_global(gpu_lib, R, dtype):
    """2000 x 1500 x 1000 with 1M uniform nonzeros, mode 0: ranges of whole
    rows, and the 1000-row factor (64 KB) fits LDS, so each block of 8 waves
    gathers it from a copy in LDS.  Declaring the last mode 1100 rows long
    (the same nonzeros, 100 unused factor rows) puts that factor above
    64 KB, so every gather goes to global memory.  Both give the same bits,
    twice in a row, and meet the FP64 oracle bar."""
    lib = gpu_lib
    d, inds, vals = gen.uniform_tensor([2000, 1500, 1000], 1_000_000, seed=2024)
    A = gen.uniform_factors(d, R, seed=2025)
    extra = gen.uniform_factors([100], R, seed=2026)[0]
    dims2 = [int(d[0]), int(d[1]), int(d[2]) + 100]
    A2 = [A[0], A[1], np.vstack([A[2], extra])]
    order = np.array([0, 1, 2], np.uint32)
    res = []
    for dims, fac in ((d, A), (dims2, A2)):
        _say("lds factor 3-D", R, dtype, list(dims))
        h = dev_coo(lib, dims, inds, vals)
        check(lib, lib.sptolDevCooPrepareMode(h, 0), "PrepareMode")
        got = dev_mttkrp_coo(lib, dims, inds, vals, fac, order, 0, dtype=dtype, h=h)
        again = dev_mttkrp_coo(lib, dims, inds, vals, fac, order, 0, dtype=dtype, h=h)
        lib.sptolDevCooFree(h)
        assert np.array_equal(got.view(np.uint32), again.view(np.uint32))
        res.append(got)
    assert np.array_equal(res[0].view(np.uint32), res[1].view(np.uint32))
    Au = [bf16_round(a) for a in A] if dtype else A
    _check(res[0], oracle.mttkrp_coo(inds, vals, Au, order, 0),
           oracle.mttkrp_coo(inds, vals, Au, order, 0, dtype=np.float64), R)


@pytest.mark.parametrize("R", [16, 32])
def test_lds_factor_4d(gpu_lib, R):
    """A 4-D tensor of 2.5M nonzeros with a 50-row mode: the MTTKRPs of the
    other modes read that mode's factor from a copy in LDS, at each of the
    three positions of the Khatri-Rao order (mats order [m, m+1, m+2, m+3]),
    FP32 and BF16; two runs and the allocation-free entry point agree bit
    for bit.  The heaviest rows hold about 200k nonzeros, whose sequential
    FP32 sums are themselves 4e-5 off, so the bar is the FP64 oracle."""
    lib = gpu_lib
    d, inds, vals = gen.zipf_tensor([3000, 2000, 1500, 50], 2_500_000, 1.05, seed=61)
    A = gen.uniform_factors(d, R, seed=62)
    h = dev_coo(lib, d, inds, vals)
    for mode in range(4):
        check(lib, lib.sptolDevCooPrepareMode(h, mode), "PrepareMode")
        order = np.array([mode] + [(mode + k) % 4 for k in range(1, 4)], np.uint32)
        for dtype in (0, SPTOL_BF16):
            _say("lds factor", R, mode, dtype)
            Au = [bf16_round(a) for a in A] if dtype else A
            got = dev_mttkrp_coo(lib, d, inds, vals, A, order, mode, dtype=dtype, h=h)
            again = dev_mttkrp_coo(lib, d, inds, vals, A, order, mode, dtype=dtype, h=h)
            assert np.array_equal(got.view(np.uint32), again.view(np.uint32))
            F = DevFactors(lib, A, dtype)
            out = DevOut(lib, int(d[mode]), R)
            check(lib, lib.sptolDevMTTKRPNoAlloc(h, F.arr, ptr(order, sptIndex), mode, C.byref(out.d), None), "NoAlloc")
            check(lib, lib.sptolHipStreamSynchronize(None), "sync")
            noalloc = out.fetch()
            F.free()
            out.free()
            assert np.array_equal(got.view(np.uint32), noalloc.view(np.uint32))
            assert np.all(got[:, R:] == 0), "padding columns must be zero"
            e64 = rel_fro(got[:, :R], oracle.mttkrp_coo(inds, vals, Au, order, mode, dtype=np.float64)[:, :R])
            _say(f"  vs FP64 oracle {e64:.2e}")
            assert e64 <= TOL, e64
    lib.sptolDevCooFree(h)


def test_long_ranges_small_factors(gpu_lib):
    """20M nonzeros, R=32, factors small enough for L2 (at most 384 KB
    gathered per mode): the view's lane groups take about 616 (FP32) or 312
    (BF16) nonzeros each, above the 256 used when factors live in HBM. Every
    mode, FP32 and BF16."""
    d, inds, vals = gen.uniform_tensor([3000, 2000, 1000], 20_000_000, seed=77)
    A = gen.uniform_factors(d, 32, seed=78)
    h = dev_coo(gpu_lib, d, inds, vals)
    for mode in range(3):
        order = [mode, (mode + 1) % 3, (mode + 2) % 3]
        for dtype in (0, SPTOL_BF16):
            _say("long ranges", mode, dtype)
            Au = [bf16_round(a) for a in A] if dtype else A
            got = dev_mttkrp_coo(gpu_lib, d, inds, vals, A, order, mode, prepare=True, dtype=dtype, h=h)
            _check(got, oracle.mttkrp_coo(inds, vals, Au, order, mode),
                   oracle.mttkrp_coo(inds, vals, Au, order, mode, dtype=np.float64), 32)
    gpu_lib.sptolDevCooFree(h)


def test_carry_runs_across_batch_edges(gpu_lib):
    """Mode-0 rows that span from 1 to about 1,100 ranges of the segmented
    view (at R=32 and this size a range holds about 1,280 nonzeros: 8 lane
    groups of 160).  Their lengths are spread around multiples of that, so
    the carry pass sees runs that end inside and at the edges of its batches:
    first one range per lane group (32 ranges), then 16 per group (512).
    Every output row is compared with the FP64 oracle on its own.  A range
    dropped from a run, or added to it, is a relative error of at least
    1/1,100 in that row (8.9e-4 for the longest).  The GPU's partials hold
    at most a few hundred sequential adds, far below that bar; the
    sequential FP32 order is itself off by 6.5e-4 in the longest row, so the
    bar is the FP64 oracle."""
    rng = np.random.default_rng(91)
    unit = 1280
    spans = [1, 2, 31, 32, 33, 34, 63, 64, 65, 95, 96, 97, 543, 544, 545, 546, 1100]
    lens = [max(1, s * unit + int(rng.integers(-unit // 2, unit // 2))) for s in spans]
    lens += [int(x) for x in rng.integers(1, 300, size=3000)]
    lens = [lens[k] for k in rng.permutation(len(lens))]
    rows = np.repeat(np.arange(len(lens), dtype=np.uint32), lens)
    nnz = rows.size
    perm = rng.permutation(nnz)   # the view sorts by mode; the input order is arbitrary
    d = np.array([len(lens), 2000, 1000], dtype=np.uint32)
    inds = [rows[perm], rng.integers(0, 2000, nnz).astype(np.uint32), rng.integers(0, 1000, nnz).astype(np.uint32)]
    vals = gen.open_unit(rng, nnz)
    A = gen.uniform_factors(d, 32, seed=92)
    h = dev_coo(gpu_lib, d, inds, vals)
    for mode in (0, 1):
        _say("carry runs", mode, nnz)
        order = [mode, (mode + 1) % 3, (mode + 2) % 3]
        got = dev_mttkrp_coo(gpu_lib, d, inds, vals, A, order, mode, prepare=True, h=h)
        ref = oracle.mttkrp_coo(inds, vals, A, order, mode, dtype=np.float64)
        assert np.all(got[:, 32:] == 0), "padding columns must be zero"
        g, r = got[:, :32].astype(np.float64), ref[:, :32]
        err = np.linalg.norm(g - r, axis=1) / np.maximum(np.linalg.norm(r, axis=1), 1e-300)
        assert err.max() <= 1e-4, (mode, int(err.argmax()), float(err.max()))
    gpu_lib.sptolDevCooFree(h)


def test_bf16_factors(gpu_lib):
    d, inds, vals = gen.zipf_tensor([3000, 2000, 4000], 100000, 1.1, seed=9)
    A = gen.uniform_factors(d, 32, seed=10)
    Ab = [bf16_round(a) for a in A]
    for mode in range(3):
        _say("bf16", mode)
        order = [mode, (mode + 1) % 3, (mode + 2) % 3]
        ref64b = oracle.mttkrp_coo(inds, vals, Ab, order, mode, dtype=np.float64)
        ref64 = oracle.mttkrp_coo(inds, vals, A, order, mode, dtype=np.float64)
        for prepare in (False, True):
            got = dev_mttkrp_coo(gpu_lib, d, inds, vals, A, order, mode, prepare=prepare, dtype=SPTOL_BF16)
            assert rel_fro(got[:, :32], ref64b[:, :32]) <= TOL
            assert rel_fro(got[:, :32], ref64[:, :32]) <= 1e-2


def test_dropin_hicoo_vs_reference_golden(gpu_lib, golden_hicoo):
    g = golden_hicoo
    for key in g.cases():
        d, inds, vals = g.tensor(key, prefix="in_ind", vals="in_vals")
        sb, sk = (int(x) for x in g.get(key, "sb_sk"))
        coo = CooTensor(d, [a.copy() for a in inds], vals.copy())
        h, _ = to_hicoo(gpu_lib, coo, sb, sk)
        A = [g.get(key, f"hA{m}") for m in range(len(d))]
        for mode in range(len(d)):
            _say("dropin hicoo", key, mode)
            order = g.get(key, f"horder{mode}")
            mats = [Matrix(a) for a in A] + [Matrix(nrows=int(max(d)), ncols=8)]
            check(gpu_lib, gpu_lib.sptMTTKRPHiCOO(C.byref(h), mats_array(mats), ptr(order, sptIndex), mode),
                  "sptMTTKRPHiCOO")
            got = mats[-1].buf[: int(d[mode])]
            ref64 = oracle.mttkrp_coo(coo.inds, coo.vals, A, order, mode, dtype=np.float64)
            _check(got, g.get(key, f"hout{mode}"), ref64, 8)
        gpu_lib.sptFreeSparseTensorHiCOO(C.byref(h))


@pytest.mark.parametrize("R,dtype,prepare", [(32, 0, False), (16, 0, False), (32, 1, False), (8, 0, False),
                                            (32, 0, "tiled"), (16, 0, "tiled"), (32, 1, "tiled"), (8, 0, "tiled"),
                                            (5, 0, "tiled"), (64, 0, "tiled"), (100, 0, "tiled"),
                                            (32, 0, "seg"), (16, 0, "seg"), (32, 1, "seg"), (8, 0, "seg"),
                                            (5, 0, "seg"), (64, 0, "seg"), (100, 0, "seg"), (24, 1, "seg"),
                                            (8, 1, "seg"), (5, 1, "seg"), (16, 1, "seg"), (64, 1, "seg")])
def test_device_hicoo_zipf(gpu_lib, R, dtype, prepare):
    d, inds, vals = gen.zipf_tensor([120000, 90000, 280000], 300000, 1.1, seed=R + dtype)
    coo = CooTensor(d, [a.copy() for a in inds], vals.copy())
    h, _ = to_hicoo(gpu_lib, coo, 7, 14)
    o = oracle.Hicoo(d, inds, vals, 7, 14)
    A = gen.uniform_factors(d, R, seed=5)
    Au = [bf16_round(a) for a in A] if dtype else A
    for mode in range(3):
        _say("device hicoo", R, dtype, prepare, mode)
        order = [mode, (mode + 1) % 3, (mode + 2) % 3]
        got = dev_mttkrp_hicoo(gpu_lib, h, d, A, order, mode, dtype=dtype, prepare=prepare)
        ref32 = o.mttkrp(Au, order, mode)
        ref64 = o.mttkrp(Au, order, mode, dtype=np.float64)
        _check(got, ref32, ref64, R)
    gpu_lib.sptFreeSparseTensorHiCOO(C.byref(h))


@pytest.mark.parametrize("kind", ["tiled", "seg"])
@pytest.mark.parametrize("sb,sk", [(1, 3), (4, 6), (7, 14), (8, 12)])
def test_hicoo_tiled_block_sizes_4d(gpu_lib, sb, sk, kind):
    """HiCOO per-mode views across block sizes (tile rows 2..256), 3-D and 4-D,
    including a heavy output row shared by many work items (tiled) or
    spanning many waves (segmented)."""
    for dims, s in (([20000, 9000, 30000], 1.3), ([3000, 2000, 2500, 400], 1.05)):
        d, inds, vals = gen.zipf_tensor(dims, 200000, s, seed=sb * 7 + sk)
        coo = CooTensor(d, [a.copy() for a in inds], vals.copy())
        h, _ = to_hicoo(gpu_lib, coo, sb, sk)
        A = gen.uniform_factors(d, 16, seed=3)
        N = len(d)
        for mode in range(N):
            _say(kind, sb, sk, N, mode)
            order = [mode] + [(mode + i) % N for i in range(1, N)]
            got = dev_mttkrp_hicoo(gpu_lib, h, d, A, order, mode, prepare=kind)
            ref32 = oracle.mttkrp_coo(coo.inds, coo.vals, A, order, mode)
            ref64 = oracle.mttkrp_coo(coo.inds, coo.vals, A, order, mode, dtype=np.float64)
            _check(got, ref32, ref64, 16)
        gpu_lib.sptFreeSparseTensorHiCOO(C.byref(h))


def test_hicoo_default_view_dense_blocks(gpu_lib):
    """The default view kind is the segmented one for 3- and 4-mode tensors
    at any block density: on tensors of dense blocks (hundreds of nonzeros
    per 2^7 block) the default view's MTTKRP equals the segmented view's bit
    for bit and the FP64 oracle within the kernel's tolerance; the tiled view
    of the same tensor agrees with the oracle as well."""
    for dims, nnz in (([600, 500, 400], 400000), ([300, 250, 200, 40], 300000)):
        d, inds, vals = gen.uniform_tensor(dims, nnz, seed=17)
        coo = CooTensor(d, [a.copy() for a in inds], vals.copy())
        h, _ = to_hicoo(gpu_lib, coo, 7, 14)
        nb = int(h.bptr.len) - 1
        assert len(vals) >= 100 * nb, (len(vals), nb)
        A = gen.uniform_factors(d, 16, seed=4)
        N = len(d)
        for mode in range(N):
            _say("default view, dense blocks", N, mode)
            order = [mode] + [(mode + i) % N for i in range(1, N)]
            auto = dev_mttkrp_hicoo(gpu_lib, h, d, A, order, mode, prepare="auto")
            seg = dev_mttkrp_hicoo(gpu_lib, h, d, A, order, mode, prepare="seg")
            tiled = dev_mttkrp_hicoo(gpu_lib, h, d, A, order, mode, prepare="tiled")
            assert np.array_equal(auto.view(np.uint32), seg.view(np.uint32)), mode
            ref32 = oracle.mttkrp_coo(coo.inds, coo.vals, A, order, mode)
            ref64 = oracle.mttkrp_coo(coo.inds, coo.vals, A, order, mode, dtype=np.float64)
            _check(auto, ref32, ref64, 16)
            _check(tiled, ref32, ref64, 16)
        gpu_lib.sptFreeSparseTensorHiCOO(C.byref(h))


def test_segmented_kernel_deterministic_and_rowwise_exact(gpu_lib):
    """The segmented kernel (sorted copy, 3-D and 4-D) is deterministic: two
    runs agree bit for bit.  Rows whose nonzeros all fall inside one lane
    group's range are summed in the reference's order, so with short rows
    nearly all rows equal the sequential FP32 reference bit for bit."""
    for dims, nnz in (([50000, 400, 300], 60000), ([30000, 60, 50, 40], 50000)):
        d, inds, vals = gen.uniform_tensor(dims, nnz, seed=11)
        A = gen.uniform_factors(d, 32, seed=2)
        N = len(d)
        h = dev_coo(gpu_lib, d, inds, vals)
        order = [0] + list(range(1, N))
        _say("seg determinism", N)
        a = dev_mttkrp_coo(gpu_lib, d, inds, vals, A, order, 0, prepare=True, h=h)
        b = dev_mttkrp_coo(gpu_lib, d, inds, vals, A, order, 0, prepare=True, h=h)
        gpu_lib.sptolDevCooFree(h)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
        ref32 = oracle.mttkrp_coo(inds, vals, A, order, 0)
        exact = np.all(a[:, :32] == ref32[:, :32], axis=1)
        frac = exact.mean()
        _say("  rows bit-exact vs sequential reference:", frac)
        assert frac > 0.95, frac
        _check(a, ref32, oracle.mttkrp_coo(inds, vals, A, order, 0, dtype=np.float64), 32)


@pytest.mark.parametrize("R", [16, 32])
def test_segmented_hot_rows_short_modes(gpu_lib, R):
    """Segmented kernel on 4-D and 3-D Zipf tensors with very short modes
    (40-60 rows), so output rows are long and span many lane groups and
    ranges, every output mode, FP32 and BF16 factors."""
    for dims in ([500, 400, 300, 60], [3000, 40, 2000]):
        d, inds, vals = gen.zipf_tensor(dims, 150000, 1.05, seed=R + len(dims))
        A = gen.uniform_factors(d, R, seed=4)
        N = len(d)
        h = dev_coo(gpu_lib, d, inds, vals)
        for mode in range(N):
            order = [mode] + [(mode + i) % N for i in range(1, N)]
            for dtype in (0, SPTOL_BF16):
                _say("seg hot", dims, mode, dtype)
                got = dev_mttkrp_coo(gpu_lib, d, inds, vals, A, order, mode, prepare=True, dtype=dtype, h=h)
                Au = [bf16_round(a) for a in A] if dtype else A
                ref32 = oracle.mttkrp_coo(inds, vals, Au, order, mode)
                ref64 = oracle.mttkrp_coo(inds, vals, Au, order, mode, dtype=np.float64)
                _check(got, ref32, ref64, R)
        gpu_lib.sptolDevCooFree(h)


def test_sorted_and_shuffled_inputs(gpu_lib):
    """A lexicographically sorted tensor and a shuffled copy of it: the
    stable sort keeps each one's own order within a row, so both match the
    sequential reference on their own nonzero order (on mode 0, whose rows
    are short, bit for bit for the rows inside one lane group's range) and
    the FP64 oracle, on every mode."""
    for dims, nnz in (([60000, 300, 200], 80000), ([40000, 60, 50, 40], 60000)):
        d, inds, vals = gen.uniform_tensor(dims, nnz, seed=17)   # lexicographically sorted
        N = len(d)
        perm = np.random.default_rng(3).permutation(len(vals))
        sh_inds, sh_vals = [a[perm].copy() for a in inds], vals[perm].copy()
        A = gen.uniform_factors(d, 16, seed=5)
        for name, (ii, vv) in (("sorted", (inds, vals)), ("shuffled", (sh_inds, sh_vals))):
            for mode in range(N):
                order = [mode] + [(mode + k) % N for k in range(1, N)]
                got = dev_mttkrp_coo(gpu_lib, d, ii, vv, A, order, mode, prepare=True)
                ref32 = oracle.mttkrp_coo(ii, vv, A, order, mode)
                frac = np.all(got[:, :16] == ref32[:, :16], axis=1).mean()
                _say(name, N, mode, "rows bit-exact:", frac)
                if mode == 0:   # short rows (1.5 nonzeros on average), most inside one lane group
                    assert frac > 0.8, (name, mode, frac)
                _check(got, ref32, oracle.mttkrp_coo(ii, vv, A, order, mode, dtype=np.float64), 16)


@pytest.mark.parametrize("dims,nnz", [([60000, 3000, 2000], 200000), ([900, 4000, 3000, 50], 150000),
                                      ([70000, 80000, 90000], 250000)])
def test_in_order_mode_skips_sort_bitwise(gpu_lib, dims, nnz):
    """A mode whose indices are already non-decreasing is packed in the
    tensor's own order instead of being sorted.  A stable sort leaves such
    an order unchanged, so the shortcut must give the same view: a shuffled
    tensor (sorted on the device) and the same tensor stably pre-sorted by
    each mode on the host (taking the shortcut) give bit-identical outputs,
    for 3 and 4 modes, on modes of 50 to 90,000 rows."""
    d, inds, vals = gen.uniform_tensor(dims, nnz, seed=23)
    rng = np.random.default_rng(29)
    perm = rng.permutation(len(vals))
    sh_inds, sh_vals = [a[perm].copy() for a in inds], vals[perm].copy()
    N = len(d)
    A = gen.uniform_factors(d, 32, seed=31)
    for mode in range(N):
        order = [mode] + [(mode + k) % N for k in range(1, N)]
        s = np.argsort(sh_inds[mode], kind="stable")
        pre_inds, pre_vals = [a[s].copy() for a in sh_inds], sh_vals[s].copy()
        assert np.all(np.diff(pre_inds[mode].astype(np.int64)) >= 0)
        _say("in-order shortcut", N, mode)
        sorted_on_device = dev_mttkrp_coo(gpu_lib, d, sh_inds, sh_vals, A, order, mode, prepare=True)
        shortcut = dev_mttkrp_coo(gpu_lib, d, pre_inds, pre_vals, A, order, mode, prepare=True)
        assert np.array_equal(sorted_on_device.view(np.uint32), shortcut.view(np.uint32)), mode
        _check(shortcut, oracle.mttkrp_coo(pre_inds, pre_vals, A, order, mode),
               oracle.mttkrp_coo(pre_inds, pre_vals, A, order, mode, dtype=np.float64), 32)
