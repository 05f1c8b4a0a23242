"""MTTKRP at the full size of the 100M-nonzero Zipf workload: dims
12,000,000 x 9,000,000 x 28,000,000, 1e8 distinct nonzeros with Zipf(1.1)
indices over permuted ids, lognormal values, R = 32, every mode, through the
COO path (sptolDevMTTKRP on prepared modes) and the HiCOO path
(sptolDevMTTKRPHiCOO on the segmented views the benchmark uses).

The FP64 oracle cannot run the whole tensor in seconds, so the checks are
properties that hold at any size:
* all-ones factors turn every output row into its slice sum broadcast over
  the columns: the whole output against an FP64 bincount of the values
  (relative Frobenius error 1e-5), every column bitwise equal, untouched rows
  exactly 0;
* random U[0,1) factors: each mode's heaviest row (its nonzeros span about
  5,700 ranges, so the carry pass sums long runs), 300 random touched rows
  and 50 untouched rows, against FP64 sums over just those rows' nonzeros
  (relative Frobenius error 1e-5 over the sample; every row within 5e-5,
  the bound for sequential FP32 sums of the ~500 positive partials on a
  row's path);
* the HiCOO output against the COO output (relative Frobenius 1e-5).
"""
import ctypes as C
import time

import numpy as np
import pytest

from conftest import rel_fro
from sptol_amd import gen
from sptol_amd.parti import (CooTensor, Matrix, SPTOL_F32, check, ptr, sptIndex, sptolDevMatrix, to_hicoo_device)

pytestmark = pytest.mark.gpu

DIMS = [12_000_000, 9_000_000, 28_000_000]
NNZ = 100_000_000
R = 32


def _say(*a):
    print(*a, flush=True)


def _lib_sort(lib):
    def sort_fn(inds):
        coo = CooTensor([0xFFFFFFFF] * len(inds), inds, np.zeros(len(inds[0]), np.float32))
        coo.sortorder[0] = 1
        lib.sptSparseTensorSortIndex(coo.ref(), 1, 0)
        return coo.inds
    return sort_fn


class _Dev:
    """Device factors and outputs of the three modes."""

    def __init__(self, lib, factors):
        self.lib = lib
        self.dm = (sptolDevMatrix * 3)()
        self.out = (sptolDevMatrix * 3)()
        for m in range(3):
            hm = Matrix(factors[m])
            check(lib, lib.sptolDevMatrixCreate(C.byref(self.dm[m]), DIMS[m], R, SPTOL_F32), "create")
            check(lib, lib.sptolDevMatrixUpload(C.byref(self.dm[m]), C.byref(hm.m)), "upload")
            del hm
            check(lib, lib.sptolDevMatrixCreate(C.byref(self.out[m]), DIMS[m], R, SPTOL_F32), "create")

    def fetch(self, m):
        h = Matrix(nrows=DIMS[m], ncols=R)
        check(self.lib, self.lib.sptolDevMatrixDownload(C.byref(h.m), C.byref(self.out[m])), "download")
        return h.buf

    def free(self):
        for m in range(3):
            self.lib.sptolDevMatrixFree(C.byref(self.dm[m]))
            self.lib.sptolDevMatrixFree(C.byref(self.out[m]))


def _order(m):
    return np.array([m, (m + 1) % 3, (m + 2) % 3], dtype=np.uint32)


def _run_all(lib, fn, tensor, dev):
    for m in range(3):
        o = _order(m)
        check(lib, fn(tensor, dev.dm, ptr(o, sptIndex), m, C.byref(dev.out[m]), None), "mttkrp")
    check(lib, lib.sptolHipStreamSynchronize(None), "sync")


def _sampled_rows_fp64(inds, vals, factors, m, rows):
    """FP64 MTTKRP rows `rows` of mode m from just their nonzeros."""
    o = _order(m)
    pos = np.full(DIMS[m], -1, dtype=np.int64)
    pos[rows] = np.arange(len(rows))
    sel = np.nonzero(pos[inds[m]] >= 0)[0]
    want = np.zeros((len(rows), R))
    A1, A2 = factors[o[1]], factors[o[2]]
    for a in range(0, len(sel), 1 << 21):
        z = sel[a:a + (1 << 21)]
        p = vals[z].astype(np.float64)[:, None] * A1[inds[o[1]][z]].astype(np.float64)
        p *= A2[inds[o[2]][z]]
        where = pos[inds[m][z]]
        for r in range(R):
            want[:, r] += np.bincount(where, weights=p[:, r], minlength=len(rows))
    return want, len(sel)


def test_c2_full_size_properties(gpu_lib):
    lib = gpu_lib
    t0 = time.time()
    dims, inds, vals = gen.zipf_tensor(DIMS, NNZ, 1.1, seed=20260101, values="lognormal", sort_fn=_lib_sort(lib))
    assert len(vals) == NNZ
    _say(f"generated {NNZ} nonzeros in {time.time() - t0:.1f}s")
    coo = CooTensor(dims, inds, vals)
    dc = C.c_void_p()
    check(lib, lib.sptolDevCooCreate(C.byref(dc), coo.ref()), "DevCooCreate")
    for m in range(3):
        check(lib, lib.sptolDevCooPrepareMode(dc, m), "PrepareMode")
    counts = [np.bincount(inds[m], minlength=DIMS[m]) for m in range(3)]
    for m in range(3):
        _say(f"mode {m}: heaviest row {counts[m].max()} nonzeros, {int((counts[m] > 0).sum())} touched rows")

    # (a) all-ones factors: slice sums
    ones = [np.ones((d, R), dtype=np.float32) for d in DIMS]
    dev = _Dev(lib, ones)
    del ones
    _run_all(lib, lib.sptolDevMTTKRP, dc, dev)
    for m in range(3):
        got = dev.fetch(m)
        want = np.bincount(inds[m], weights=vals.astype(np.float64), minlength=DIMS[m])
        err = rel_fro(got[:, 0], want)
        _say(f"ones mode {m}: rel err {err:.2e}")
        assert err < 1e-5, m
        assert np.array_equal(got.view(np.uint32), np.repeat(got[:, :1], R, axis=1).view(np.uint32)), m
        assert not got[counts[m] == 0].any(), m
        del got, want
    dev.free()

    # (b) random factors: heaviest row + sampled rows against FP64
    factors = gen.uniform_factors(DIMS, R, seed=7)
    dev = _Dev(lib, factors)
    _run_all(lib, lib.sptolDevMTTKRP, dc, dev)
    rng = np.random.default_rng(11)
    samples, coo_out = [], []
    for m in range(3):
        touched = np.nonzero(counts[m])[0]
        empty = np.nonzero(counts[m] == 0)[0]
        rows = np.unique(np.concatenate([[int(np.argmax(counts[m]))], rng.choice(touched, 300, replace=False),
                                         rng.choice(empty, 50, replace=False)]))
        want, nsel = _sampled_rows_fp64(inds, vals, factors, m, rows)
        got = dev.fetch(m)
        coo_out.append(got[rows].copy())
        err = rel_fro(got[rows], want)
        per_row = np.linalg.norm(got[rows] - want, axis=1) / np.maximum(np.linalg.norm(want, axis=1), 1e-300)
        _say(f"random mode {m}: {len(rows)} rows ({nsel} nonzeros), rel err {err:.2e}, worst row {per_row.max():.2e}")
        assert err < 1e-5, m
        assert per_row.max() < 5e-5, m
        assert not got[rows][counts[m][rows] == 0].any(), m
        samples.append((rows, want))
        del got
    lib.sptolDevCooFree(dc)

    # (c) the HiCOO path of the benchmark (sb = 2^7, sk = 2^14), same factors
    t0 = time.time()
    hcoo = CooTensor(dims, [a.copy() for a in inds], vals.copy())
    h, _ = to_hicoo_device(lib, hcoo, 7, 14)
    del hcoo
    dh = C.c_void_p()
    check(lib, lib.sptolDevHicooCreate(C.byref(dh), C.byref(h)), "DevHicooCreate")
    for m in range(3):
        check(lib, lib.sptolDevHicooPrepareMode(dh, m), "HicooPrepareMode")
    _say(f"HiCOO built in {time.time() - t0:.1f}s")
    _run_all(lib, lib.sptolDevMTTKRPHiCOO, dh, dev)
    for m in range(3):
        rows, want = samples[m]
        got = dev.fetch(m)[rows]
        err = rel_fro(got, want)
        _say(f"hicoo mode {m}: rel err {err:.2e} vs FP64, {rel_fro(got, coo_out[m]):.2e} vs COO")
        assert err < 1e-5, m
        assert rel_fro(got, coo_out[m]) < 1e-5, m
    lib.sptolDevHicooFree(dh)
    lib.sptFreeSparseTensorHiCOO(C.byref(h))
    dev.free()
