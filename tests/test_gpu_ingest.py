"""GPU ingest: a device COO tensor streamed from a binary tensor file
(sptolDevCooLoadBinary) equals one uploaded from host arrays
(sptolDevCooCreate): the MTTKRP of every mode is bit-identical, since both
build the same mode-sorted views, and within 1e-5 of the FP64 oracle.  A
file with an index >= ndims is rejected.  The server's file submission
(sptolServerSubmitFile) returns the same result, bit for bit, as submitting
the host arrays, and its exact fit is within 1e-5 of the FP64 oracle's
(factors and weights within 1e-4)."""
import ctypes as C

import numpy as np
import pytest

import oracle
from conftest import rel_fro
from gpu_util import dev_mttkrp_coo
from sptol_amd import gen
from sptol_amd.parti import (CooTensor, Matrix, SPTERR_VALUE_ERROR, SPTOL_F32, SPTOL_REQ_DONE, check,
                             mats_array, ptr, seeded_uniform_factors, sptolRequestInfo, sptolServerConfig)

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def test_load_binary_matches_upload(gpu_lib, tmp_path):
    lib = gpu_lib
    # 5.5M nonzeros: each array spans two 16 MiB pinned chunks
    d, inds, vals = gen.zipf_tensor([30000, 20000, 9000], 5_500_000, 1.1, seed=12)
    coo = CooTensor(d, inds, vals)
    path = str(tmp_path / "x.bin").encode()
    check(lib, lib.sptolSaveSparseTensorBinary(coo.ref(), path), "save")
    h = C.c_void_p()
    _say("load binary", len(vals))
    check(lib, lib.sptolDevCooLoadBinary(C.byref(h), path, None), "DevCooLoadBinary")
    assert lib.sptolDevCooNnz(h) == len(vals)
    factors = gen.uniform_factors(d, 16, seed=3)
    for mode in range(3):
        order = [mode] + [(mode + k) % 3 for k in (1, 2)]
        a = dev_mttkrp_coo(lib, d, inds, vals, factors, order, mode, prepare=True, h=h)
        b = dev_mttkrp_coo(lib, d, inds, vals, factors, order, mode, prepare=True)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), mode
        want = oracle.mttkrp_coo(inds, vals, factors, np.asarray(order, np.uint32), mode, dtype=np.float64)
        err = rel_fro(a[:, :16], want[:, :16])
        _say(f"  mode {mode}: streamed tensor vs FP64 oracle {err:.2e}")
        assert err < 1e-5, mode
    lib.sptolDevCooFree(h)
    bad = CooTensor(d, [a.copy() for a in inds], vals)
    bad.inds[2][len(vals) - 1] = d[2]
    badpath = str(tmp_path / "bad.bin").encode()
    check(lib, lib.sptolSaveSparseTensorBinary(bad.ref(), badpath), "save")
    h2 = C.c_void_p()
    assert lib.sptolDevCooLoadBinary(C.byref(h2), badpath, None) == SPTERR_VALUE_ERROR
    assert lib.sptolDevCooLoadBinary(C.byref(h2), b"/nonexistent/x.bin", None) != 0


def test_server_submit_file(gpu_lib, tmp_path):
    lib = gpu_lib
    d, inds, vals, _ = gen.planted_tensor([400, 300, 200, 30], 60000, 8, 1.05, seed=77)
    coo = CooTensor(d, inds, vals)
    path = str(tmp_path / "r.bin").encode()
    check(lib, lib.sptolSaveSparseTensorBinary(coo.ref(), path), "save")
    cfg = sptolServerConfig()
    cfg.hbm_budget = 16 << 30
    cfg.max_active = 4
    cfg.batch_mttkrp = 1
    cfg.exact_fit = 1
    srv = C.c_void_p()
    check(lib, lib.sptolServerCreate(C.byref(srv), C.byref(cfg)), "ServerCreate")
    ids = []
    for k in range(2):
        rid = C.c_uint64()
        if k == 0:
            check(lib, lib.sptolServerSubmit(srv, coo.ref(), 8, 12, -1.0, 99, SPTOL_F32, C.byref(rid)), "Submit")
        else:
            check(lib, lib.sptolServerSubmitFile(srv, path, 8, 12, -1.0, 99, SPTOL_F32, C.byref(rid)), "SubmitFile")
        ids.append(rid.value)
    res = []
    for rid in ids:
        info = sptolRequestInfo()
        check(lib, lib.sptolServerWait(srv, rid, 600.0, C.byref(info)), "Wait")
        assert info.status == SPTOL_REQ_DONE
        mats = [Matrix(nrows=int(x), ncols=8) for x in d]
        lam = np.zeros(8, np.float32)
        check(lib, lib.sptolServerFetch(srv, rid, mats_array(mats), ptr(lam, C.c_float)), "Fetch")
        res.append((info.fit, lam, [m.array().copy() for m in mats]))
        lib.sptolServerRelease(srv, rid)
    rid = C.c_uint64()
    assert lib.sptolServerSubmitFile(srv, b"/nonexistent/x.bin", 8, 12, -1.0, 1, SPTOL_F32, C.byref(rid)) != 0
    lib.sptolServerDestroy(srv)
    (f0, l0, A0), (f1, l1, A1) = res
    _say("fits", f0, f1)
    assert f0 == f1
    assert np.array_equal(l0.view(np.uint32), l1.view(np.uint32))
    for x, y in zip(A0, A1):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))
    # the file-submitted solve against the FP64 oracle from the request's
    # seeded initial factors
    init = seeded_uniform_factors(99, d, 8)
    om, olam, ofits = oracle.cpd_als(d, inds, vals, init, 12)
    _say(f"  served fit {f1:.9f} oracle {ofits[-1]:.9f}")
    assert abs(f1 - ofits[-1]) < 1e-5
    assert rel_fro(l1, olam) < 1e-4
    for m in range(4):
        assert rel_fro(A1[m], om[m]) < 1e-4, m
