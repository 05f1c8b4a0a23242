"""Host-side product logic on the CPU: the C-ABI library loads and exports
its header's symbols; containers, reader, orders and the COO->HiCOO converter
match the reference (golden fixtures) and the oracle.  No GPU calls."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, ROOT
from sptol_amd import gen
from sptol_amd.parti import (CooTensor, Matrix, hicoo_to_numpy, mats_array, ptr, sptIndex,
                             sptSparseTensor, to_hicoo, SPTERR_SHAPE_MISMATCH, SPTERR_VALUE_ERROR,
                             last_error)


def _declared_functions():
    names = set()
    for h in ("ParTI.h", "sptol_hip.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*(?:[A-Za-z_][\w\s\*]*?)\b([A-Za-z_]\w*)\s*\(", txt, flags=re.M):
            name = m.group(1)
            if name in ("if", "for", "while", "return", "sizeof"):
                continue
            names.add(name)
    return names


def test_library_exports_every_declared_symbol(lib):
    so = os.path.join(ROOT, "sptol_amd", "libsptol_amd.so")
    out = subprocess.run(["nm", "-D", "--defined-only", so], capture_output=True, text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    declared = _declared_functions()
    assert len(declared) > 60
    missing = sorted(declared - exported)
    assert not missing, missing


def test_library_links_no_oracle():
    so = os.path.join(ROOT, "sptol_amd", "libsptol_amd.so")
    out = subprocess.run(["nm", "-D", so], capture_output=True, text=True, check=True).stdout
    assert "oracle_" not in out
    ldd = subprocess.run(["ldd", so], capture_output=True, text=True).stdout
    assert "oracle" not in ldd and "parti_ref" not in ldd


def _coo_from_npz(path):
    z = np.load(path, allow_pickle=False)
    d = z["ndims"]
    return d, [z[f"ind{m}"] for m in range(len(d))], z["vals"]


def test_reader_matches_reference(lib):
    libc = C.CDLL("libc.so.6")
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    lib.sptLoadSparseTensor.argtypes = [C.POINTER(sptSparseTensor), sptIndex, C.c_void_p]
    lib.sptFreeSparseTensor.argtypes = [C.POINTER(sptSparseTensor)]
    for name in ("3d_3_8", "3d_4_12", "3d_dense"):
        fp = libc.fopen(os.path.join(GOLDEN, f"{name}.tns").encode(), b"r")
        t = sptSparseTensor()
        assert lib.sptLoadSparseTensor(C.byref(t), 1, fp) == 0
        libc.fclose(fp)
        d, inds, vals = _coo_from_npz(os.path.join(GOLDEN, f"tns_{name}.npz"))
        n = int(t.nnz)
        assert n == len(vals)
        assert np.array_equal(np.ctypeslib.as_array(t.ndims, shape=(len(d),)), d)
        for m in range(len(d)):
            assert np.array_equal(np.ctypeslib.as_array(t.inds[m].data, shape=(n,)), inds[m])
        assert np.array_equal(np.ctypeslib.as_array(t.values.data, shape=(n,)), vals)
        lib.sptFreeSparseTensor(C.byref(t))


def test_lexicographic_sort_matches_oracle(lib):
    for ndims in ([300, 200, 100], [2**31, 2**31, 2**31, 7]):
        rng = np.random.default_rng(1)
        n = 20000
        inds = [rng.integers(0, d, size=n).astype(np.uint32) for d in ndims]
        vals = rng.random(n).astype(np.float32)
        coo = CooTensor(ndims, [a.copy() for a in inds], vals.copy())
        coo.sortorder[0] = 1   # force a sort
        lib.sptSparseTensorSortIndex(coo.ref(), 1, 4)
        oi, ov = oracle.sort_lex(inds, vals)
        for m in range(len(ndims)):
            assert np.array_equal(coo.inds[m], oi[m])


def _check_against_golden(arr, coo, g, key, N):
    for f in ("kptr", "cptr", "bptr", "nkiters"):
        assert np.array_equal(arr[f], g.get(key, f)), (key, f)
    assert np.array_equal(arr["values"].view(np.uint32), g.get(key, "values").view(np.uint32))
    for m in range(N):
        assert np.array_equal(arr["binds"][m], g.get(key, f"binds{m}")), (key, m)
        assert np.array_equal(arr["einds"][m], g.get(key, f"einds{m}")), (key, m)
        assert np.array_equal(coo.inds[m], g.get(key, f"perm_ind{m}")), (key, m)
        ks = arr["kschr"][m]
        p = np.concatenate([[0], np.cumsum([len(x) for x in ks])]).astype(np.uint64)
        assert np.array_equal(p, g.get(key, f"kschr_ptr{m}"))
        idx = np.concatenate(ks).astype(np.uint32) if ks else np.zeros(0, np.uint32)
        assert np.array_equal(idx, g.get(key, f"kschr_idx{m}"))


def test_hicoo_converter_matches_reference(lib, golden_hicoo):
    g = golden_hicoo
    for key in g.cases():
        d, inds, vals = g.tensor(key, prefix="in_ind", vals="in_vals")
        sb, sk = (int(x) for x in g.get(key, "sb_sk"))
        coo = CooTensor(d, [a.copy() for a in inds], vals.copy())
        h, mx = to_hicoo(lib, coo, sb, sk)
        arr = hicoo_to_numpy(h)
        _check_against_golden(arr, coo, g, key, len(d))
        assert mx == int(g.get(key, "max_nnzb"))
        lib.sptFreeSparseTensorHiCOO(C.byref(h))


def test_hicoo_dump_text_matches_reference(lib, golden_hicoo, tmp_path):
    libc = C.CDLL("libc.so.6")
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    lib.sptDumpSparseTensorHiCOO.argtypes = [C.c_void_p, C.c_void_p]
    g = golden_hicoo
    for fname in sorted(os.listdir(GOLDEN)):
        if not fname.endswith(".dump.txt"):
            continue
        key = fname[: -len(".dump.txt")]
        d, inds, vals = g.tensor(key, prefix="in_ind", vals="in_vals")
        sb, sk = (int(x) for x in g.get(key, "sb_sk"))
        coo = CooTensor(d, [a.copy() for a in inds], vals.copy())
        h, _ = to_hicoo(lib, coo, sb, sk)
        out = tmp_path / fname
        fp = libc.fopen(str(out).encode(), b"w")
        lib.sptDumpSparseTensorHiCOO(C.addressof(h), fp)
        libc.fclose(fp)
        assert out.read_text() == open(os.path.join(GOLDEN, fname)).read(), fname
        lib.sptFreeSparseTensorHiCOO(C.byref(h))


@pytest.mark.parametrize("ndims,sb,sk,s", [([12000, 9000, 28000], 7, 14, 1.1),
                                           ([5000, 5000, 5000, 500], 4, 8, 1.05),
                                           ([70000, 3, 900000], 2, 9, 1.2)])
def test_hicoo_converter_matches_oracle_zipf(lib, ndims, sb, sk, s):
    d, inds, vals = gen.zipf_tensor(ndims, 60000, s, seed=sb * 31 + sk)
    coo = CooTensor(d, [a.copy() for a in inds], vals.copy())
    h, mx = to_hicoo(lib, coo, sb, sk)
    arr = hicoo_to_numpy(h)
    o = oracle.Hicoo(d, inds, vals, sb, sk)
    oa = o.arrays()
    for f in ("kptr", "cptr", "bptr", "nkiters"):
        assert np.array_equal(arr[f], oa[f]), f
    for m in range(len(d)):
        assert np.array_equal(arr["binds"][m], oa["binds"][m])
        assert np.array_equal(arr["einds"][m], oa["einds"][m])
        assert np.array_equal(coo.inds[m], o.inds[m])
    assert mx == oa["max_nnzb"]
    lib.sptFreeSparseTensorHiCOO(C.byref(h))


def test_hicoo_converter_rejects_bad_input(lib):
    d = np.array([10, 10, 10], np.uint32)
    inds = [np.array([1, 2], np.uint32), np.array([3, 12], np.uint32), np.array([0, 0], np.uint32)]
    coo = CooTensor(d, inds, np.array([1, 2], np.float32))
    with pytest.raises(RuntimeError):
        to_hicoo(lib, coo, 1, 2)        # index 12 >= ndims[1]
    coo = CooTensor(d, [a % 10 for a in inds], np.array([1, 2], np.float32))
    with pytest.raises(RuntimeError):
        to_hicoo(lib, coo, 3, 2)        # sk_bits < sb_bits


def test_mttkrp_shape_mismatch_is_reported(lib):
    """Shape checks run on the host before any device work (mttkrp.c:58-65)."""
    d, inds, vals = gen.uniform_tensor([20, 10, 5], 100, seed=1)
    coo = CooTensor(d, inds, vals)
    A = gen.uniform_factors([20, 10, 6], 8, seed=2)   # wrong rows in mode 2
    mats = [Matrix(a) for a in A] + [Matrix(nrows=20, ncols=8)]
    order = np.array([0, 1, 2], np.uint32)
    lib.sptClearLastError()
    rc = lib.sptMTTKRP(coo.ref(), mats_array(mats), ptr(order, sptIndex), 0)
    assert rc == SPTERR_SHAPE_MISMATCH
    code, msg = last_error(lib)
    assert code == SPTERR_SHAPE_MISMATCH and "nrows" in msg
    A = gen.uniform_factors([20, 10, 5], 8, seed=2)
    mats = [Matrix(a) for a in A[:2]] + [Matrix(gen.uniform_factors([5], 4, seed=3)[0])] + [Matrix(nrows=20, ncols=8)]
    assert lib.sptMTTKRP(coo.ref(), mats_array(mats), ptr(order, sptIndex), 0) == SPTERR_SHAPE_MISMATCH


def _tensor_arrays(t, nmodes):
    n = int(t.nnz)
    d = np.ctypeslib.as_array(t.ndims, shape=(nmodes,)).copy()
    inds = [np.ctypeslib.as_array(t.inds[m].data, shape=(n,)).copy() if n else np.zeros(0, np.uint32)
            for m in range(nmodes)]
    vals = np.ctypeslib.as_array(t.values.data, shape=(n,)).copy() if n else np.zeros(0, np.float32)
    return d, inds, vals


def test_merge_duplicates(lib):
    """Duplicated coordinates merge into one nonzero whose value is the FP64
    sum of the run (in input order), rounded once; the result is sorted
    lexicographically.  Checked against a numpy restatement (unique rows of
    the stacked coordinates, float64 bincount of the values)."""
    rng = np.random.default_rng(5)
    for dims, n in (([40, 30, 20], 30000), ([7, 5, 3, 4], 2000), ([2**31, 3], 500)):
        inds = [rng.integers(0, d, size=n).astype(np.uint32) for d in dims]
        vals = rng.standard_normal(n).astype(np.float32)
        coo = CooTensor(dims, [a.copy() for a in inds], vals.copy())
        assert lib.sptolSparseTensorMergeDuplicates(coo.ref(), 4) == 0
        m = int(coo.st.nnz)
        keys = np.stack(inds[::-1])          # lexsort: last key is primary
        order = np.lexsort(keys)
        coords = np.stack(inds, axis=1)[order]
        first = np.ones(n, bool)
        first[1:] = np.any(coords[1:] != coords[:-1], axis=1)
        grp = np.cumsum(first) - 1
        want = np.bincount(grp, weights=vals[order].astype(np.float64)).astype(np.float32)
        assert m == int(first.sum())
        assert m < n or dims[0] == 2**31      # the wide case draws no duplicates
        for q in range(len(dims)):
            assert np.array_equal(coo.inds[q][:m], coords[first][:, q])
        assert np.array_equal(coo.vals[:m].view(np.uint32), want.view(np.uint32))


def test_binary_roundtrip_and_rejects(lib, tmp_path):
    d, inds, vals = gen.zipf_tensor([5000, 300, 70000, 12], 50000, 1.1, seed=9)
    coo = CooTensor(d, inds, vals)
    path = str(tmp_path / "t.bin").encode()
    assert lib.sptolSaveSparseTensorBinary(coo.ref(), path) == 0
    t = sptSparseTensor()
    assert lib.sptolLoadSparseTensorBinary(C.byref(t), path) == 0
    d2, i2, v2 = _tensor_arrays(t, 4)
    lib.sptFreeSparseTensor(C.byref(t))
    assert np.array_equal(d2, d)
    for m in range(4):
        assert np.array_equal(i2[m], coo.inds[m])
    assert np.array_equal(v2.view(np.uint32), coo.vals.view(np.uint32))
    raw = open(path, "rb").read()
    hdr_bytes = (24 + 4 * 4 + 7) // 8 * 8
    assert raw[:8] == b"SPTOLBT1" and len(raw) == hdr_bytes + 5 * 4 * len(vals)
    # truncated file, bad magic, out-of-range index
    bad = str(tmp_path / "bad.bin").encode()
    open(bad, "wb").write(raw[:-4])
    assert lib.sptolLoadSparseTensorBinary(C.byref(sptSparseTensor()), bad) == SPTERR_VALUE_ERROR
    open(bad, "wb").write(b"XXXXXXXX" + raw[8:])
    assert lib.sptolLoadSparseTensorBinary(C.byref(sptSparseTensor()), bad) == SPTERR_VALUE_ERROR
    hdr = bytearray(raw)
    hdr[hdr_bytes:hdr_bytes + 4] = np.uint32(d[0]).tobytes()    # inds[0][0] = ndims[0]
    open(bad, "wb").write(bytes(hdr))
    assert lib.sptolLoadSparseTensorBinary(C.byref(sptSparseTensor()), bad) == SPTERR_VALUE_ERROR


def test_server_rejects_more_than_four_admission_threads(lib):
    """sptolServerCreate refuses admit_threads above 4 before touching the
    device (no stream or thread is created)."""
    from sptol_amd.parti import SPTERR_VALUE_ERROR, sptolServerConfig
    cfg = sptolServerConfig()
    cfg.hbm_budget = 1 << 30
    cfg.max_active = 4
    cfg.admit_threads = 8
    srv = C.c_void_p()
    assert lib.sptolServerCreate(C.byref(srv), C.byref(cfg)) == SPTERR_VALUE_ERROR
    assert not srv.value
