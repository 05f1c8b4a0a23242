"""C callers of the library on the GPU, checked against the oracle.

* tests/c_driver/sptol_driver: this repository's C program written in the
  reference drivers' calling style (sptAssert, PARTI_SCN_* option parsing,
  status reports, timers).  It runs sptMTTKRP and sptMTTKRPHiCOO on every
  mode and sptCpdAls and writes every input and result with full FP32
  precision: MTTKRP within 1e-5 relative Frobenius error of the FP64 oracle,
  CPD fit within 1e-5, factors and lambda within 1e-4.
* oracle/_ref/drivers/{mttkrp,cpd}: the reference's own examples/mttkrp.c
  and examples/cpd.c, compiled unchanged against include/ParTI.h and linked
  with libsptol_amd.so (built where the reference tree exists).  They print
  their results with the reference's dump formats ("%.2f" matrices, "%lf"
  fit), so they are checked at that print precision.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle
from conftest import ROOT, rel_fro
from sptol_amd import gen
from sptol_amd.parti import Matrix, sptMatrix

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(ROOT, "tests", "c_driver", "sptol_driver")
REF_DRIVERS = os.path.join(ROOT, "oracle", "_ref", "drivers")
DEFAULT_SEED = 0x5EED5EED5EED5EED   # the library's generator state in a new process


def _say(*a):
    print(*a, flush=True)


def _write_tns(path, ndims, inds, vals):
    with open(path, "w") as f:
        f.write(f"{len(ndims)}\n")
        f.write(" ".join(str(int(d)) for d in ndims) + "\n")
        cols = [np.asarray(i, dtype=np.int64) + 1 for i in inds]
        for z in range(len(vals)):
            f.write(" ".join(str(int(c[z])) for c in cols) + f" {float(vals[z]):.9e}\n")


def _randomized(lib, ndims, R, seed):
    """The factors sptRandomizeMatrix draws, mode by mode, after
    sptolSetSeed(seed)."""
    lib.sptRandomizeMatrix.argtypes = [C.POINTER(sptMatrix), C.c_uint32, C.c_uint32]
    lib.sptolSetSeed(seed)
    out = []
    for d in ndims:
        m = Matrix(nrows=int(d), ncols=R)
        lib.sptRandomizeMatrix(C.byref(m.m), int(d), R)
        out.append(m.array().copy())
    return out


def _run(cmd, cwd):
    _say("run", " ".join(os.path.basename(c) if i == 0 else c for i, c in enumerate(cmd)))
    r = subprocess.run(["timeout", "-k", "10", "120"] + cmd, cwd=cwd, capture_output=True, text=True)
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-2000:])
    return r.stdout


@pytest.mark.parametrize("ndims,R", [([300, 200, 250], 16), ([120, 90, 150, 30], 8)])
def test_c_driver_vs_oracle(gpu_lib, tmp_path, ndims, R):
    if not os.path.exists(DRIVER):
        pytest.fail("tests/c_driver/sptol_driver not built (run build())")
    d, inds, vals = gen.zipf_tensor(ndims, 20000, 1.1, seed=len(ndims))
    tns = tmp_path / "x.tns"
    _write_tns(tns, d, inds, vals)
    out = tmp_path / "out"
    out.mkdir()
    seed, iters = 2024, 6
    log = _run([DRIVER, "-i", str(tns), "-o", str(out), "-r", str(R), "-n", str(iters), "-s", str(seed),
                "-b", "4", "-k", "7"], tmp_path)
    assert "COO Sparse Tensor information" in log and "HiCOO Sparse Tensor information" in log
    assert "driver done" in log
    N = len(d)
    t = np.loadtxt(out / "tensor.txt", ndmin=2)
    sinds = [t[:, m].astype(np.uint32) for m in range(N)]
    svals = t[:, N].astype(np.float32)
    order = np.lexsort(inds[::-1])
    for m in range(N):
        assert np.array_equal(sinds[m], inds[m][order])
    assert np.array_equal(svals.view(np.uint32), vals[order].view(np.uint32))
    fac = np.loadtxt(out / "factors.txt", ndmin=2).astype(np.float32)
    A, off = [], 0
    for m in range(N):
        A.append(fac[off:off + int(d[m])])
        off += int(d[m])
    want_A = _randomized(gpu_lib, d, R, seed)
    for m in range(N):
        assert np.array_equal(A[m], want_A[m])
    for mode in range(N):
        o = np.array([(mode + k) % N for k in range(N)], dtype=np.uint32)
        want = oracle.mttkrp_coo(sinds, svals, A, o, mode, dtype=np.float64)[:, :R]
        for kind in ("coo", "hicoo"):
            got = np.loadtxt(out / f"{kind}_mode{mode}.txt", ndmin=2)
            err = rel_fro(got, want)
            _say(f"  {kind} mode {mode}: rel err {err:.2e}")
            assert err < 1e-5, (kind, mode, err)
    lines = (out / "cpd.txt").read_text().splitlines()
    fit = float(lines[0])
    lam = np.array(lines[1].split(), dtype=np.float64)
    rows = np.array([l.split() for l in lines[2:]], dtype=np.float64)
    init = _randomized(gpu_lib, d, R, seed + 1)
    om, olam, ofits = oracle.cpd_als(d, sinds, svals, init, iters)
    _say(f"  cpd fit {fit:.9f} oracle {ofits[-1]:.9f}")
    assert abs(fit - ofits[-1]) < 1e-5
    assert rel_fro(lam, olam) < 1e-4
    off = 0
    for m in range(N):
        assert rel_fro(rows[off:off + int(d[m])], om[m]) < 1e-4, m
        off += int(d[m])


def _need_ref_driver(name):
    path = os.path.join(REF_DRIVERS, name)
    if not os.path.exists(path):
        pytest.skip("reference drivers not built (the reference tree was absent at build time)")
    return path


def _read_dumped_matrix(lines, at):
    rows, cols = (int(x) for x in lines[at].split(" matrix")[0].split(" x "))
    vals = np.array([l.split() for l in lines[at + 1:at + 1 + rows]], dtype=np.float64)
    return vals.reshape(rows, cols), at + 1 + rows + 1


def test_reference_mttkrp_driver(gpu_lib, tmp_path):
    """examples/mttkrp.c, unchanged, mode 1, sequential path (-d -2): it
    sorts the tensor, draws U[0..N-1] with sptRandomizeMatrix (the library's
    seeded generator from its initial state), runs sptMTTKRP and dumps the
    output matrix ("%.2f")."""
    exe = _need_ref_driver("mttkrp")
    d, inds, vals = gen.uniform_tensor([400, 300, 200], 20000, seed=17)
    tns = tmp_path / "x.tns"
    _write_tns(tns, d, inds, vals)
    R, mode = 8, 1
    log = _run([exe, "-i", str(tns), "-m", str(mode), "-d", "-2", "-r", str(R), "-o", str(tmp_path / "out.txt")],
               tmp_path)
    assert "COO Sparse Tensor information" in log
    A = _randomized(gpu_lib, d, R, DEFAULT_SEED)
    order = np.lexsort(inds[::-1])
    si = [a[order] for a in inds]
    o = np.array([mode, (mode + 1) % 3, (mode + 2) % 3], dtype=np.uint32)
    want = oracle.mttkrp_coo(si, vals[order], A, o, mode, dtype=np.float64)[:, :R]
    got, _ = _read_dumped_matrix((tmp_path / "out.txt").read_text().splitlines(), 0)
    assert got.shape == (int(max(d)), R)
    assert np.all(got[int(d[mode]):] == 0)
    diff = np.abs(got[: int(d[mode])] - want)
    _say(f"  max |printed - oracle| = {diff.max():.4f} (print step 0.01)")
    assert np.all(diff <= 0.005 + 1e-5 * np.abs(want))


def test_reference_cpd_driver(gpu_lib, tmp_path):
    """examples/cpd.c, unchanged: sptCpdAls with rank 4, 5 iterations and
    tol 1e-5 from factors drawn by the library's seeded generator; its dump
    of the Kruskal tensor against the FP64 oracle from the same factors."""
    exe = _need_ref_driver("cpd")
    d, inds, vals = gen.uniform_tensor([60, 50, 40], 6000, seed=19)
    tns = tmp_path / "x.tns"
    _write_tns(tns, d, inds, vals)
    R = 4
    log = _run([exe, "-i", str(tns), "-r", str(R), "-o", str(tmp_path / "kt.txt")], tmp_path)
    assert "COO Sparse Tensor information" in log
    init = _randomized(gpu_lib, d, R, DEFAULT_SEED)
    om, olam, ofits = oracle.cpd_als(d, inds, vals, init, 5, tol=1e-5)
    lines = (tmp_path / "kt.txt").read_text().splitlines()
    assert lines[0] == f"nmodes: 3, rank: {R}"
    fit = float(lines[2].split(":")[1])
    lam = np.array(lines[4].split(), dtype=np.float64)
    _say(f"  fit printed {fit:.6f} oracle {ofits[-1]:.9f} after {len(ofits)} iterations")
    assert abs(fit - ofits[-1]) < 1e-5
    assert rel_fro(lam, olam) < 1e-4
    at = 6
    for m in range(3):
        got, at = _read_dumped_matrix(lines, at)
        assert np.all(np.abs(got - om[m]) <= 0.005 + 1e-4)
