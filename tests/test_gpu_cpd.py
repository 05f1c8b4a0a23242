"""GPU parity of CPD-ALS through the C ABI.

Against the compiled reference's CpdAlsStep (golden factors/lambda after a
fixed iteration count from committed initial factors, FP32 LAPACK): relative
Frobenius error within 2e-3 (the reference's FP32 solve is the looser side,
see the oracle test of the same bound).  Against the FP64 oracle on the same
inputs: factors and lambda within 1e-4, fit within 1e-6 (FP32 factors and
MTTKRP, FP64 solve/norms/fit).  BF16-stored factors: fit within 1e-4 of the
FP64 oracle on a planted low-rank tensor."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, rel_fro
from sptol_amd import gen
from sptol_amd.parti import (CooTensor, Matrix, SPTOL_BF16, SPTOL_F32, check, mats_array, ptr,
                             sptKruskalTensor, sptMatrix, sptIndex, sptolCpdStats, sptolDevMatrix, sptRankKruskalTensor,
                             sptRankMatrix, to_hicoo)

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _cpd_step(lib, d, inds, vals, init, iters, tol=-1.0):
    R = init[0].shape[1]
    coo = CooTensor(d, inds, vals)
    mats = [Matrix(a) for a in init] + [Matrix(nrows=int(max(d)), ncols=R)]
    lam = np.zeros(R, np.float32)
    fit = lib.CpdAlsStep(coo.ref(), R, iters, tol, mats_array(mats), ptr(lam, C.c_float))
    assert np.isfinite(fit)
    return [m.array().copy() for m in mats[:-1]], lam, fit, mats[-1]


def test_cpd_step_vs_reference_and_oracle(gpu_lib, golden_cpd):
    """CpdAlsStep returns the fit with <X,K> from an FP64 pass over the
    nonzeros (FP64 products and sums of the FP32 factors), so its error is
    the FP32 factors' own rounding (~6e-8 relative per entry), which bounds
    ||X - K|| / ||X|| near 1e-6 on an exactly low-rank tensor (the FP64
    oracle reaches 1 - 1e-14 on the rank-2 test_cpd case): held to 1e-5.
    The per-iteration fit that drives tol reuses the last FP32 MTTKRP for
    <X,K>, as the reference does; near a fit of 1, ||X||^2 + ||K||^2 -
    2<X,K> cancels and a relative error e of <X,K> becomes an absolute fit
    error of about sqrt(2e), ~5e-4 for FP32: that fit is never returned."""
    g = golden_cpd
    for case in g.cases():
        d, inds, vals = g.tensor(case)
        R = int(g.get(case, "R"))
        iters = int(g.get(case, "iters"))
        init = [g.get(case, f"init{m}") for m in range(len(d))]
        _say("cpd", case)
        A, lam, fit, _ = _cpd_step(gpu_lib, d, inds, vals, init, iters)
        om, olam, ofits = oracle.cpd_als(d, inds, vals, init, iters)
        assert rel_fro(lam, g.get(case, "lambda")) < 2e-3, case
        assert rel_fro(lam, olam) < 1e-4, case
        for m in range(len(d)):
            assert rel_fro(A[m], g.get(case, f"A{m}")) < 2e-3, (case, m)
            assert rel_fro(A[m], om[m]) < 1e-4, (case, m)
        _say(f"  fit gpu {fit:.12f} oracle {ofits[-1]:.12f}")
        assert abs(fit - ofits[-1]) < 1e-5, (case, fit, ofits[-1])
        _, _, st = _device_cpd(gpu_lib, d, inds, vals, init, iters, -1.0, SPTOL_F32, exact_fit=1)
        assert abs(st.fit - ofits[-1]) < 1e-5, (case, st.fit, ofits[-1])
        assert st.fit == fit, case


def _dense_planted(dims, R, seed, noise=1e-4):
    """Every coordinate of the grid, values from a planted U[0,1) rank-R model
    plus noise*rms Gaussian noise: exactly low rank up to the noise, so ALS
    reaches a fit near 1 - noise."""
    rng = np.random.default_rng(seed)
    grid = np.stack(np.meshgrid(*[np.arange(x) for x in dims], indexing="ij"), -1).reshape(-1, len(dims))
    inds = [grid[:, m].astype(np.uint32) for m in range(len(dims))]
    factors = [rng.random((x, R)) for x in dims]
    vals, _ = gen.planted_values(inds, factors, noise, rng)
    return np.asarray(dims, dtype=np.uint32), inds, vals


def test_exact_fit_order5_and_hicoo(gpu_lib):
    """The exact fit beyond the 3-/4-mode COO kernel: a 5-mode COO tensor
    (CpdAlsStep) and a HiCOO tensor of both view kinds
    (sptolCpdAlsHicooDevice), both dense and exactly low rank up to noise
    of 1e-4 of the rms, so that the reference-method fit would cancel; each within
    1e-5 of the FP64 oracle, and the HiCOO fit within 1e-6 of the COO one
    on the same tensor in the same order (their MTTKRPs sum in different
    orders, so the FP32 factors differ in their last bits)."""
    from gpu_util import VIEW_KIND
    lib = gpu_lib
    d, inds, vals = _dense_planted([6, 5, 4, 3, 3], 3, seed=51)
    init = gen.uniform_factors(d, 3, seed=52)
    iters = 30
    _say("order-5 CpdAlsStep")
    A, lam, fit, _ = _cpd_step(lib, d, inds, vals, init, iters)
    om, olam, ofits = oracle.cpd_als(d, inds, vals, init, iters)
    _say(f"  fit gpu {fit:.12f} oracle {ofits[-1]:.12f}")
    assert ofits[-1] > 0.99
    assert abs(fit - ofits[-1]) < 1e-5
    assert rel_fro(lam, olam) < 1e-4
    d3, inds3, vals3 = _dense_planted([16, 12, 10], 2, seed=53)
    coo = CooTensor(d3, [a.copy() for a in inds3], vals3.copy())
    h, _ = to_hicoo(lib, coo, 3, 7)            # coo now in HiCOO order
    init3 = gen.uniform_factors(d3, 2, seed=54)
    _, _, ofits3 = oracle.cpd_als(d3, coo.inds, coo.vals, init3, iters)
    _, _, st_coo = _device_cpd(lib, d3, coo.inds, coo.vals, init3, iters, -1.0, SPTOL_F32, exact_fit=1)
    for kind in ("tiled", "seg"):
        check(lib, lib.sptolSetHicooViewKind(VIEW_KIND[kind]), "view kind")
        try:
            t = C.c_void_p()
            check(lib, lib.sptolDevHicooCreate(C.byref(t), C.byref(h)), "DevHicooCreate")
            dm = (sptolDevMatrix * 3)()
            hm = [Matrix(a) for a in init3]
            for m in range(3):
                check(lib, lib.sptolDevMatrixCreate(C.byref(dm[m]), int(d3[m]), 2, SPTOL_F32), "create")
                check(lib, lib.sptolDevMatrixUpload(C.byref(dm[m]), C.byref(hm[m].m)), "upload")
            lam3 = np.zeros(2, np.float32)
            st = sptolCpdStats()
            st.exact_fit = 1
            check(lib, lib.sptolCpdAlsHicooDevice(t, 2, iters, -1.0, dm, ptr(lam3, C.c_float), C.byref(st), None),
                  "cpd hicoo")
            for m in range(3):
                lib.sptolDevMatrixFree(C.byref(dm[m]))
            lib.sptolDevHicooFree(t)
        finally:
            lib.sptolSetHicooViewKind(0)
        _say(f"  hicoo {kind}: fit {st.fit:.12f} coo {st_coo.fit:.12f} oracle {ofits3[-1]:.12f}")
        assert ofits3[-1] > 0.99
        assert abs(st.fit - ofits3[-1]) < 1e-5, kind
        assert abs(st.fit - st_coo.fit) < 1e-6, kind
    lib.sptFreeSparseTensorHiCOO(C.byref(h))


def test_sptcpdals_dropin(gpu_lib):
    d, inds, vals = gen.uniform_tensor([60, 50, 40], 8000, seed=31)
    R, iters = 8, 6
    gpu_lib.sptolSetSeed(1234)
    kt = sptKruskalTensor()
    nd = np.asarray(d, np.uint32)
    gpu_lib.sptNewKruskalTensor.argtypes = [C.POINTER(sptKruskalTensor), sptIndex, C.POINTER(sptIndex), sptIndex]
    gpu_lib.sptFreeKruskalTensor.argtypes = [C.POINTER(sptKruskalTensor)]
    gpu_lib.sptRandomizeMatrix.argtypes = [C.c_void_p, sptIndex, sptIndex]
    check(gpu_lib, gpu_lib.sptNewKruskalTensor(C.byref(kt), 3, ptr(nd, sptIndex), R), "new ktensor")
    coo = CooTensor(d, inds, vals)
    check(gpu_lib, gpu_lib.sptCpdAls(coo.ref(), R, iters, 0.0, C.byref(kt)), "sptCpdAls")
    # the same seeded initial factors, regenerated
    gpu_lib.sptolSetSeed(1234)
    init = []
    for m in range(3):
        mm = Matrix(nrows=int(d[m]), ncols=R)
        gpu_lib.sptRandomizeMatrix(C.byref(mm.m), int(d[m]), R)
        init.append(mm.array().copy())
    assert all(np.all((a >= 0) & (a < 1)) for a in init)
    om, olam, ofits = oracle.cpd_als(d, inds, vals, init, iters)
    lam = np.ctypeslib.as_array(kt.lambda_, shape=(R,)).copy()
    assert rel_fro(lam, olam) < 1e-4
    for m in range(3):
        f = kt.factors[m].contents
        got = np.ctypeslib.as_array(f.values, shape=(f.nrows, f.stride))[:, :R]
        assert rel_fro(got, om[m]) < 1e-4
    assert abs(kt.fit - ofits[-1]) < 1e-6
    gpu_lib.sptFreeKruskalTensor(C.byref(kt))


def _device_cpd(lib, d, inds, vals, init, iters, tol, dtype, exact_fit=0):
    R = init[0].shape[1]
    coo = CooTensor(d, inds, vals)
    h = C.c_void_p()
    check(lib, lib.sptolDevCooCreate(C.byref(h), coo.ref()), "DevCooCreate")
    dm = (sptolDevMatrix * len(d))()
    hm = [Matrix(a) for a in init]
    for m in range(len(d)):
        check(lib, lib.sptolDevMatrixCreate(C.byref(dm[m]), int(d[m]), R, SPTOL_F32), "create")
        check(lib, lib.sptolDevMatrixUpload(C.byref(dm[m]), C.byref(hm[m].m)), "upload")
    lam = np.zeros(R, np.float32)
    st = sptolCpdStats()
    st.factor_dtype = dtype
    st.exact_fit = exact_fit
    check(lib, lib.sptolCpdAlsDevice(h, R, iters, tol, dm, ptr(lam, C.c_float), C.byref(st), None), "CpdAlsDevice")
    out = []
    for m in range(len(d)):
        check(lib, lib.sptolDevMatrixDownload(C.byref(hm[m].m), C.byref(dm[m])), "download")
        out.append(hm[m].array().copy())
        lib.sptolDevMatrixFree(C.byref(dm[m]))
    lib.sptolDevCooFree(h)
    return out, lam, st


@pytest.mark.parametrize("R", [16, 32])
def test_planted_fp32_and_bf16(gpu_lib, R):
    d, inds, vals, planted = gen.planted_tensor([500, 400, 300, 60], 200000, R, 1.05, seed=R)
    init = gen.uniform_factors(d, R, seed=R + 1)
    iters = 20
    om, olam, ofits = oracle.cpd_als(d, inds, vals, init, iters)
    for dtype in (SPTOL_F32, SPTOL_BF16):
        _say("planted", R, dtype)
        A, lam, st = _device_cpd(gpu_lib, d, inds, vals, init, iters, -1.0, dtype, exact_fit=1)
        assert st.iterations == iters
        _say(f"  fit gpu {st.fit:.6f} oracle {ofits[-1]:.6f} diff {abs(st.fit - ofits[-1]):.2e}")
        assert abs(st.fit - ofits[-1]) < 1e-4
        if dtype == SPTOL_F32:
            assert abs(st.fit - ofits[-1]) < 1e-5


def test_tolerance_stops_early(gpu_lib):
    d, inds, vals, _ = gen.planted_tensor([80, 70, 60], 20000, 4, 1.1, seed=5)
    init = gen.uniform_factors(d, 4, seed=6)
    _, _, st = _device_cpd(gpu_lib, d, inds, vals, init, 200, 1e-4, SPTOL_F32)
    _, _, ofits = oracle.cpd_als(d, inds, vals, init, 200, tol=1e-4)
    assert st.iterations == len(ofits)
    assert st.iterations < 200


def test_cpd_rejects_bad_rank(gpu_lib):
    d, inds, vals = gen.uniform_tensor([20, 20, 20], 500, seed=1)
    init = gen.uniform_factors(d, 65, seed=2)
    with pytest.raises(RuntimeError):
        _device_cpd(gpu_lib, d, inds, vals, init, 2, -1.0, SPTOL_F32)


def test_block_cache_reuse(gpu_lib):
    """The workspace of a device CPD goes to the block cache, not back to HIP:
    memory in use returns to where it was, a second identical solve reuses
    the cached blocks and gives bit-identical results, and releasing the
    cache empties it."""
    lib = gpu_lib
    lib.sptolHipReleaseCache()
    assert lib.sptolHipMemoryCached() == 0
    base = lib.sptolHipMemoryInUse()
    d, inds, vals, _ = gen.planted_tensor([300, 250, 200, 40], 60000, 16, 1.05, seed=21)
    init = gen.uniform_factors(d, 16, seed=22)
    A1, lam1, st1 = _device_cpd(lib, d, inds, vals, init, 6, -1.0, SPTOL_F32)
    assert lib.sptolHipMemoryInUse() == base
    cached = lib.sptolHipMemoryCached()
    _say("cached after one solve", cached)
    assert cached > 0
    A2, lam2, st2 = _device_cpd(lib, d, inds, vals, init, 6, -1.0, SPTOL_F32)
    assert lib.sptolHipMemoryInUse() == base
    assert lib.sptolHipMemoryCached() <= cached + (1 << 20)
    assert st1.fit == st2.fit
    assert np.array_equal(lam1.view(np.uint32), lam2.view(np.uint32))
    for x, y in zip(A1, A2):
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))
    lib.sptolHipReleaseCache()
    assert lib.sptolHipMemoryCached() == 0


@pytest.mark.parametrize("kind", ["tiled", "seg"])
def test_sptcpdalshicoo_dropin(gpu_lib, kind):
    """sptCpdAlsHiCOO (hicoo/cpd.c:149) on the device, over both view kinds:
    the same ALS as sptCpdAls on the tensor in HiCOO order, against the FP64
    oracle from the same seeded initial factors."""
    from gpu_util import VIEW_KIND
    lib = gpu_lib
    d, inds, vals = gen.uniform_tensor([60, 50, 40], 8000, seed=33)
    R, iters = 8, 6
    coo = CooTensor(d, [a.copy() for a in inds], vals.copy())
    h, _ = to_hicoo(lib, coo, 3, 5)            # permutes coo into the HiCOO order
    lib.sptolSetSeed(4321)
    kt = sptRankKruskalTensor()
    nd = np.asarray(d, np.uint32)
    check(lib, lib.sptNewRankKruskalTensor(C.byref(kt), 3, ptr(nd, sptIndex), R), "new rank ktensor")
    check(lib, lib.sptolSetHicooViewKind(VIEW_KIND[kind]), "view kind")
    try:
        check(lib, lib.sptCpdAlsHiCOO(C.byref(h), R, iters, 0.0, C.byref(kt)), "sptCpdAlsHiCOO")
    finally:
        lib.sptolSetHicooViewKind(0)
    lib.sptolSetSeed(4321)
    init = []
    for m in range(3):
        rm = sptRankMatrix()
        check(lib, lib.sptNewRankMatrix(C.byref(rm), int(d[m]), R), "new rank matrix")
        check(lib, lib.sptRandomizeRankMatrix(C.byref(rm), int(d[m]), R), "randomize")
        init.append(np.ctypeslib.as_array(rm.values, shape=(int(d[m]), rm.stride))[:, :R].copy())
        lib.sptFreeRankMatrix(C.byref(rm))
    om, olam, ofits = oracle.cpd_als(d, coo.inds, coo.vals, init, iters)
    lam = np.ctypeslib.as_array(kt.lambda_, shape=(R,)).copy()
    _say(f"hicoo cpd ({kind}): fit {kt.fit:.9f} oracle {ofits[-1]:.9f}")
    assert rel_fro(lam, olam) < 1e-4
    for m in range(3):
        f = kt.factors[m].contents
        got = np.ctypeslib.as_array(f.values, shape=(f.nrows, f.stride))[:, :R]
        assert rel_fro(got, om[m]) < 1e-4
    assert abs(kt.fit - ofits[-1]) < 1e-6
    lib.sptFreeRankKruskalTensor(C.byref(kt))
    lib.sptFreeSparseTensorHiCOO(C.byref(h))


def test_hicoo_device_cpd_matches_coo(gpu_lib):
    """sptolCpdAlsHicooDevice and sptolCpdAlsDevice on the same 4-D tensor
    (HiCOO order) and initial factors reach the same fits, and both match
    the FP64 oracle."""
    lib = gpu_lib
    d, inds, vals, _ = gen.planted_tensor([300, 250, 200, 40], 80000, 16, 1.05, seed=41)
    coo = CooTensor(d, [a.copy() for a in inds], vals.copy())
    h, _ = to_hicoo(lib, coo, 4, 8)
    R, iters = 16, 8
    init = gen.uniform_factors(d, R, seed=42)
    out = {}
    for kind in ("coo", "hicoo"):
        dm = (sptolDevMatrix * 4)()
        hm = [Matrix(a) for a in init]
        for m in range(4):
            check(lib, lib.sptolDevMatrixCreate(C.byref(dm[m]), int(d[m]), R, SPTOL_F32), "create")
            check(lib, lib.sptolDevMatrixUpload(C.byref(dm[m]), C.byref(hm[m].m)), "upload")
        lam = np.zeros(R, np.float32)
        st = sptolCpdStats()
        t = C.c_void_p()
        if kind == "coo":
            check(lib, lib.sptolDevCooCreate(C.byref(t), coo.ref()), "DevCooCreate")
            check(lib, lib.sptolCpdAlsDevice(t, R, iters, -1.0, dm, ptr(lam, C.c_float), C.byref(st), None), "cpd")
            lib.sptolDevCooFree(t)
        else:
            check(lib, lib.sptolDevHicooCreate(C.byref(t), C.byref(h)), "DevHicooCreate")
            check(lib, lib.sptolCpdAlsHicooDevice(t, R, iters, -1.0, dm, ptr(lam, C.c_float), C.byref(st), None),
                  "cpd hicoo")
            lib.sptolDevHicooFree(t)
        A = []
        for m in range(4):
            check(lib, lib.sptolDevMatrixDownload(C.byref(hm[m].m), C.byref(dm[m])), "download")
            A.append(hm[m].array().copy())
            lib.sptolDevMatrixFree(C.byref(dm[m]))
        out[kind] = (A, lam, st.fit, st.iterations)
    _, olam, ofits = oracle.cpd_als(d, coo.inds, coo.vals, init, iters)
    _say(f"fit coo {out['coo'][2]:.9f} hicoo {out['hicoo'][2]:.9f} oracle {ofits[-1]:.9f}")
    assert out["coo"][3] == out["hicoo"][3] == iters
    assert abs(out["coo"][2] - out["hicoo"][2]) < 1e-5
    assert abs(out["hicoo"][2] - ofits[-1]) < 1e-5
    assert rel_fro(out["hicoo"][1], olam) < 1e-3
    lib.sptFreeSparseTensorHiCOO(C.byref(h))


def _solve_normals(lib, mode, grams, rhs):
    R = grams[0].shape[0]
    mats = [Matrix(g) for g in grams] + [Matrix(nrows=R, ncols=R)]
    b = Matrix(rhs)
    lib.sptMatrixSolveNormals.argtypes = [C.c_uint32, C.c_uint32, C.POINTER(C.POINTER(sptMatrix)),
                                          C.POINTER(sptMatrix)]
    check(lib, lib.sptMatrixSolveNormals(mode, len(grams), mats_array(mats), C.byref(b.m)), "SolveNormals")
    return b.array().copy(), mats[-1].array().copy()


def _hadamard_upper(grams, mode):
    N, R = len(grams), grams[0].shape[0]
    V = np.ones((R, R))
    for k in range(1, N):
        g = np.triu(grams[(mode + k) % N].astype(np.float64))
        V *= g + np.triu(g, 1).T
    return V


def test_solve_normals_cholesky_lu_and_singular(gpu_lib):
    """sptMatrixSolveNormals (matrix_solver.c:32) on the device: SPD Grams
    (Cholesky), a symmetric indefinite V (the LU-with-pivoting fallback) and
    exactly singular V (rows left as they are).  Against numpy's FP64 solve
    of the same V within 1e-6, against the compiled reference's FP32 LAPACK
    (golden) within 1e-3, and bit for bit on the singular cases, where the
    reference leaves the rows too."""
    z = np.load(os.path.join(GOLDEN, "solve.npz"), allow_pickle=False)
    for name in ("spd", "indefinite", "singular", "zero_column"):
        mode = int(z[f"{name}/mode"])
        grams = []
        while f"{name}/gram{len(grams)}" in z:
            grams.append(z[f"{name}/gram{len(grams)}"])
        rhs = z[f"{name}/rhs_in"]
        got, V = _solve_normals(gpu_lib, mode, grams, rhs)
        want_V = _hadamard_upper(grams, mode)
        assert np.array_equal(V, want_V.astype(np.float32)), name
        ref = z[f"{name}/rhs_out"]
        if name in ("singular", "zero_column"):
            assert np.linalg.matrix_rank(want_V) < want_V.shape[0]
            assert np.array_equal(got.view(np.uint32), rhs.view(np.uint32)), name
            assert np.array_equal(ref.view(np.uint32), rhs.view(np.uint32)), name
            continue
        exact = np.linalg.solve(want_V, rhs.astype(np.float64).T).T
        _say(f"  {name}: vs numpy {rel_fro(got, exact):.2e}, vs reference {rel_fro(got, ref):.2e}")
        assert rel_fro(got, exact) < 1e-6, name
        assert rel_fro(got, ref) < 1e-3, name
    if name == "indefinite":
        assert np.min(np.linalg.eigvalsh(want_V)) < 0


@pytest.mark.parametrize("dims", [[16, 12, 10], [16, 12, 10, 8]])
def test_cpd_singular_normal_equations(gpu_lib, dims):
    """Initial factors with two equal columns of small dyadic values make the
    first mode update's V exactly singular (Cholesky meets an exact zero,
    LU an exact zero pivot): the solve reports it in singular_solves and
    the run stays finite.  The 4-D case forms V^-1 inside the MTTKRP launch
    (the 3-D one in its own kernel)."""
    d, inds, vals = gen.uniform_tensor(dims, 800, seed=61)
    R = 4
    rng = np.random.default_rng(62)
    init = []
    for n in d:
        a = (rng.integers(1, 5, size=(int(n), R)) * 0.5).astype(np.float32)
        # column 0: n - 1 ones and (n - 2) / 2, whose squares sum to (n / 2)^2,
        # a perfect square, so its norm is exact; column 1 repeats it
        a[:, 0] = 1.0
        a[0, 0] = (int(n) - 2) / 2
        a[:, 1] = a[:, 0]
        assert np.sqrt((a[:, 0].astype(np.float64) ** 2).sum()) == int(n) / 2
        init.append(a)
    A, lam, st = _device_cpd(gpu_lib, d, inds, vals, init, 2, -1.0, SPTOL_F32)
    _say(f"  singular solves {st.singular_solves}, fit {st.fit:.6f}")
    assert st.singular_solves >= 1
    assert all(np.isfinite(x).all() for x in A) and np.isfinite(lam).all() and np.isfinite(st.fit)
