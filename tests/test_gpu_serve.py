"""GPU tests of the continuous-batching CPD-ALS server.

* Each request's result is the same alone (a server holding only that
  request) and in a mix of 3-D and 4-D requests of uneven sizes, ranks and
  factor storage that run concurrently: bit for bit.  Every kernel on the
  path partitions a request's work by that request's own sizes only and
  sums in a fixed order (the segmented MTTKRP, the batched solve / norm /
  Gram / fit steps), so the mix cannot change a result.
* Between paths that partition the MTTKRP differently (a borrowing and a
  copying server agree bit for bit; the served and the single-request
  device path do not), results agree to rounding; how far rounding carries
  depends on how well-posed the problem is:
    - well-posed requests (every entry of an exactly low-rank tensor present,
      so the CP solution is unique): fit within 1e-6, lambda and factors
      within 1e-4 relative Frobenius error;
    - sparse samples of planted models (fit ~0.06-0.3, many factorisations fit
      almost equally well): fit within 1e-4; their factors can drift apart by
      percent-level amounts, so they are not compared.
* Results match the single-request device path on the same seeded initial
  factors, and the FP64 oracle (fit within 1e-4).
* Failure isolation: a request over the HBM budget, one with an index >= ndims
  and one hit by an injected allocation failure each fail with their own
  error code while the others complete correctly; a non-finite value is
  rejected at submission, or, when the server borrows the caller's arrays
  instead of copying them, fails its request at admission.
* A borrowing server returns the same results as a copying one."""
import ctypes as C

import numpy as np
import pytest

import oracle
from conftest import rel_fro
from sptol_amd import gen
from sptol_amd.parti import (CooTensor, Matrix, SPTERR_VALUE_ERROR, SPTOL_BF16, SPTOL_F32, SPTOL_REQ_DONE,
                             SPTOL_REQ_FAILED, check, mats_array, ptr, seeded_uniform_factors,
                             sptolCpdStats, sptolDevMatrix, sptolRequestInfo, sptolServerConfig)

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _server(lib, budget=64 << 30, slots=16, batch=1, borrow=0):
    cfg = sptolServerConfig()
    cfg.hbm_budget = budget
    cfg.max_active = slots
    cfg.batch_mttkrp = batch
    cfg.exact_fit = 1
    cfg.borrow_inputs = borrow
    s = C.c_void_p()
    check(lib, lib.sptolServerCreate(C.byref(s), C.byref(cfg)), "ServerCreate")
    return s


def _dense_lowrank(dims, rank, seed, noise=0.0):
    """Every entry of a planted rank-`rank` tensor (values in (0, rank])."""
    import itertools
    rng = np.random.default_rng(seed)
    U = [rng.random((d, rank)) + 0.1 for d in dims]
    idx = np.array(list(itertools.product(*(range(d) for d in dims))), dtype=np.uint32).T
    v = np.ones((idx.shape[1], rank))
    for m, Um in enumerate(U):
        v *= Um[idx[m]]
    vals = v.sum(axis=1)
    if noise:
        vals += rng.normal(0, noise * np.sqrt(np.mean(vals ** 2)), size=vals.shape)
    return np.asarray(dims, np.uint32), list(idx), vals.astype(np.float32)


def _well_posed():
    out = []
    for k, (dims, R) in enumerate((([30, 25, 20], 3), ([12, 10, 9, 8], 2), ([40, 30, 25], 4), ([15, 12, 10, 6], 3))):
        d, inds, vals = _dense_lowrank(dims, R, seed=50 + k)
        out.append(dict(d=d, inds=inds, vals=vals, R=R, dtype=SPTOL_F32, seed=3000 + k, iters=40, tol=-1.0,
                        well_posed=True))
    return out


def _requests():
    specs = [([300, 200, 100], 20000, 8, 1.1, SPTOL_F32), ([80, 70, 60, 20], 15000, 16, 1.05, SPTOL_F32),
             ([2000, 1500, 1000], 120000, 16, 1.1, SPTOL_F32), ([500, 400, 300, 50], 80000, 32, 1.05, SPTOL_BF16),
             ([100, 100, 100], 5000, 4, 1.2, SPTOL_F32), ([3000, 200, 40], 50000, 32, 1.1, SPTOL_F32),
             ([60, 60, 60, 60], 30000, 8, 1.1, SPTOL_BF16), ([1000, 900, 800], 60000, 16, 1.1, SPTOL_F32)]
    reqs = []
    for k, (dims, nnz, R, s, dt) in enumerate(specs):
        d, inds, vals, _ = gen.planted_tensor(dims, nnz, R, s, seed=100 + k)
        reqs.append(dict(d=d, inds=inds, vals=vals, R=R, dtype=dt, seed=1000 + k, iters=15, tol=1e-6,
                         well_posed=False))
    return reqs


def _submit(lib, srv, r, keep=None):
    coo = CooTensor(r["d"], r["inds"], r["vals"])
    rid = C.c_uint64()
    rc = lib.sptolServerSubmit(srv, coo.ref(), r["R"], r["iters"], r["tol"], r["seed"], r["dtype"], C.byref(rid))
    if keep is not None:
        keep.append(coo)   # a borrowing server reads these arrays until the request is waited for
    return rc, rid.value


def _result(lib, srv, rid, r):
    info = sptolRequestInfo()
    check(lib, lib.sptolServerWait(srv, rid, 600.0, C.byref(info)), "Wait")
    out = dict(info=info)
    if info.status == SPTOL_REQ_DONE:
        mats = [Matrix(nrows=int(x), ncols=r["R"]) for x in r["d"]]
        lam = np.zeros(r["R"], np.float32)
        check(lib, lib.sptolServerFetch(srv, rid, mats_array(mats), ptr(lam, C.c_float)), "Fetch")
        out["A"] = [m.array().copy() for m in mats]
        out["lam"] = lam
    lib.sptolServerRelease(srv, rid)
    return out


def _same(a, b):
    assert a["info"].iterations == b["info"].iterations
    assert a["info"].fit == b["info"].fit, (a["info"].fit, b["info"].fit)
    assert np.array_equal(a["lam"].view(np.uint32), b["lam"].view(np.uint32))
    for x, y in zip(a["A"], b["A"]):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))


def _close(a, b, well_posed):
    if well_posed:
        assert abs(a["info"].fit - b["info"].fit) < 1e-6, (a["info"].fit, b["info"].fit)
        assert rel_fro(a["lam"], b["lam"]) < 1e-4
        for x, y in zip(a["A"], b["A"]):
            assert rel_fro(x, y) < 1e-4
    else:
        assert abs(a["info"].fit - b["info"].fit) < 1e-4, (a["info"].fit, b["info"].fit)


@pytest.mark.parametrize("batch", [1, 0])
def test_alone_equals_mixed(gpu_lib, batch):
    lib = gpu_lib
    reqs = _requests() + _well_posed()
    alone = []
    for k, r in enumerate(reqs):
        _say("alone", k, batch)
        srv = _server(lib, batch=batch)
        rc, rid = _submit(lib, srv, r)
        check(lib, rc, "Submit")
        alone.append(_result(lib, srv, rid, r))
        lib.sptolServerDestroy(srv)
        assert alone[-1]["info"].status == SPTOL_REQ_DONE
    _say("mixed", batch)
    srv = _server(lib, batch=batch, slots=6)
    ids = []
    for r in reqs:
        rc, rid = _submit(lib, srv, r)
        check(lib, rc, "Submit")
        ids.append(rid)
    mixed = [_result(lib, srv, rid, r) for rid, r in zip(ids, reqs)]
    assert lib.sptolServerTicks(srv) > 0
    lib.sptolServerDestroy(srv)
    for k, (a, b, r) in enumerate(zip(alone, mixed, reqs)):
        assert b["info"].status == SPTOL_REQ_DONE
        _say(f"  request {k} ({'well-posed' if r['well_posed'] else 'sparse'}): fit alone "
             f"{a['info'].fit:.9f} mixed {b['info'].fit:.9f}")
        if r["well_posed"]:
            _say(f"    lambda rel diff {rel_fro(a['lam'], b['lam']):.2e}, factor rel diff "
                 f"{max(rel_fro(x, y) for x, y in zip(a['A'], b['A'])):.2e}")
            assert a["info"].fit > 0.99
        _same(a, b)


def test_served_matches_device_path_and_oracle(gpu_lib):
    lib = gpu_lib
    r = _requests()[1]
    srv = _server(lib)
    rc, rid = _submit(lib, srv, r)
    check(lib, rc, "Submit")
    got = _result(lib, srv, rid, r)
    lib.sptolServerDestroy(srv)
    init = seeded_uniform_factors(r["seed"], r["d"], r["R"])
    # the single-request device path on the same initial factors
    coo = CooTensor(r["d"], r["inds"], r["vals"])
    h = C.c_void_p()
    check(lib, lib.sptolDevCooCreate(C.byref(h), coo.ref()), "DevCooCreate")
    dm = (sptolDevMatrix * 4)()
    hm = [Matrix(a) for a in init]
    for m in range(4):
        check(lib, lib.sptolDevMatrixCreate(C.byref(dm[m]), int(r["d"][m]), r["R"], SPTOL_F32), "create")
        check(lib, lib.sptolDevMatrixUpload(C.byref(dm[m]), C.byref(hm[m].m)), "upload")
    lam = np.zeros(r["R"], np.float32)
    st = sptolCpdStats()
    st.exact_fit = 1
    check(lib, lib.sptolCpdAlsDevice(h, r["R"], r["iters"], r["tol"], dm, ptr(lam, C.c_float), C.byref(st), None), "cpd")
    for m in range(4):
        lib.sptolDevMatrixFree(C.byref(dm[m]))
    lib.sptolDevCooFree(h)
    assert got["info"].iterations == st.iterations
    assert abs(got["info"].fit - st.fit) < 1e-5
    assert rel_fro(got["lam"], lam) < 1e-4
    _, olam, ofits = oracle.cpd_als(r["d"], r["inds"], r["vals"], init, r["iters"], tol=r["tol"])
    assert abs(got["info"].fit - ofits[-1]) < 1e-4


def test_failure_isolation(gpu_lib):
    lib = gpu_lib
    reqs = _requests()[:3]
    big = dict(reqs[2])
    srv = _server(lib, budget=8 << 20)             # 8 MiB: the two small requests fit, the 120k-nnz one does not
    rc, rid_big = _submit(lib, srv, big)
    check(lib, rc, "Submit")
    bad = dict(reqs[0])
    bad_inds = [a.copy() for a in bad["inds"]]
    bad_inds[1][5] = bad["d"][1] + 3                # index >= ndims
    bad["inds"] = bad_inds
    rc, rid_bad = _submit(lib, srv, bad)
    check(lib, rc, "Submit")
    nan = dict(reqs[0])
    nv = nan["vals"].copy()
    nv[7] = np.nan
    nan["vals"] = nv
    rc, _ = _submit(lib, srv, nan)
    assert rc == SPTERR_VALUE_ERROR
    rc, rid_ok = _submit(lib, srv, reqs[0])
    check(lib, rc, "Submit")
    rc, rid_ok2 = _submit(lib, srv, reqs[1])
    check(lib, rc, "Submit")
    r_big = _result(lib, srv, rid_big, big)
    r_bad = _result(lib, srv, rid_bad, bad)
    r_ok = _result(lib, srv, rid_ok, reqs[0])
    r_ok2 = _result(lib, srv, rid_ok2, reqs[1])
    assert r_big["info"].status == SPTOL_REQ_FAILED and r_big["info"].error != 0
    assert r_bad["info"].status == SPTOL_REQ_FAILED and r_bad["info"].error == SPTERR_VALUE_ERROR
    assert r_ok["info"].status == SPTOL_REQ_DONE and r_ok2["info"].status == SPTOL_REQ_DONE
    # injected allocation failure hits the next admitted request only
    lib.sptolHipInjectMallocFailure(1)
    rc, rid_inj = _submit(lib, srv, reqs[0])
    check(lib, rc, "Submit")
    r_inj = _result(lib, srv, rid_inj, reqs[0])
    assert r_inj["info"].status == SPTOL_REQ_FAILED
    rc, rid_after = _submit(lib, srv, reqs[0])
    check(lib, rc, "Submit")
    r_after = _result(lib, srv, rid_after, reqs[0])
    assert r_after["info"].status == SPTOL_REQ_DONE
    _close(r_ok, r_after, False)
    lib.sptolServerDestroy(srv)


def test_borrowed_inputs(gpu_lib):
    lib = gpu_lib
    reqs = _requests()[:4] + _well_posed()[:2]
    out = {}
    for borrow in (0, 1):
        _say("borrow", borrow)
        srv = _server(lib, slots=4, borrow=borrow)
        keep = []
        ids = []
        for r in reqs:
            rc, rid = _submit(lib, srv, r, keep)
            check(lib, rc, "Submit")
            ids.append(rid)
        nan = dict(reqs[0])
        nv = nan["vals"].copy()
        nv[7] = np.inf
        nan["vals"] = nv
        rc, rid_nan = _submit(lib, srv, nan, keep)
        if borrow:
            check(lib, rc, "Submit")
            r_nan = _result(lib, srv, rid_nan, nan)
            assert r_nan["info"].status == SPTOL_REQ_FAILED and r_nan["info"].error == SPTERR_VALUE_ERROR
        else:
            assert rc == SPTERR_VALUE_ERROR
        out[borrow] = [_result(lib, srv, rid, r) for rid, r in zip(ids, reqs)]
        lib.sptolServerDestroy(srv)
        del keep
    for a, b, r in zip(out[0], out[1], reqs):
        assert a["info"].status == SPTOL_REQ_DONE and b["info"].status == SPTOL_REQ_DONE
        _same(a, b)
