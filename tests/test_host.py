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
    # nnz whose byte count wraps to 0 in 64 bits: 2^62 * 4 * (4 + 1) = 5 * 2^64
    for nnz in (1 << 62, (1 << 64) - 1, len(vals) + 1):
        hdr = bytearray(raw)
        hdr[16:24] = np.uint64(nnz).tobytes()
        open(bad, "wb").write(bytes(hdr))
        assert lib.sptolLoadSparseTensorBinary(C.byref(sptSparseTensor()), bad) == SPTERR_VALUE_ERROR, nnz


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


# ---------------------------------------------------------------------------
# Driver-support API: the calls the reference's own drivers make around the
# MTTKRP/CPD entry points (examples/mttkrp.c, cpd.c, mttkrp_hicoo.c).
# ---------------------------------------------------------------------------

def _libc():
    libc = C.CDLL("libc.so.6")
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    libc.free.argtypes = [C.c_void_p]
    return libc


def _golden_driver():
    return np.load(os.path.join(GOLDEN, "driver.npz"), allow_pickle=False)


def test_mode_orders_match_reference(lib):
    """sptGetBestModeOrder / sptGetWorstModeOrder against the compiled
    reference on dimension lists with ties (where its quicksort decides which
    equal modes trade places), for every mode."""
    g = _golden_driver()
    k = 0
    while f"dims/{k}" in g:
        nd = g[f"dims/{k}"]
        N = len(nd)
        for mode in range(N):
            for which, fn in (("best", lib.sptGetBestModeOrder), ("worst", lib.sptGetWorstModeOrder)):
                o = np.zeros(N, dtype=np.uint32)
                fn(ptr(o, sptIndex), mode, ptr(nd, sptIndex), N)
                assert np.array_equal(o, g[f"{which}/{k}/{mode}"]), (which, list(nd), mode, o)
                assert sorted(o) == list(range(N))
        k += 1
    assert k >= 9


def test_random_shuffle_matches_reference(lib):
    g = _golden_driver()
    d = g["shuffle/ndims"]
    coo = CooTensor(d, [g[f"shuffle/ind_in{m}"].copy() for m in range(len(d))], g["shuffle/vals_in"].copy())
    lib.sptGetRandomShuffleElements(coo.ref())
    for m in range(len(d)):
        assert np.array_equal(coo.inds[m], g[f"shuffle/ind_out{m}"])
    assert np.array_equal(coo.vals.view(np.uint32), g["shuffle/vals_out"].view(np.uint32))


def test_custom_order_and_at_mode_sorts(lib):
    """Custom-order and at-mode sorts are lexicographic orders over a mode
    permutation: pinned by np.lexsort on distinct coordinates; sortorder
    becomes the order used, and an unforced sort in that order is a no-op."""
    d, inds, vals = gen.uniform_tensor([300, 7, 5000, 40], 20000, seed=12)
    rng = np.random.default_rng(3)
    perm = rng.permutation(len(vals))
    inds = [a[perm] for a in inds]
    vals = vals[perm]
    for order in ([2, 0, 3, 1], [1, 3, 2, 0], [0, 1, 2, 3]):
        coo = CooTensor(d, [a.copy() for a in inds], vals.copy())
        o = np.asarray(order, dtype=np.uint32)
        lib.sptSparseTensorSortIndexCustomOrder(coo.ref(), ptr(o, sptIndex), 1, 4)
        want = np.lexsort([inds[m] for m in order[::-1]])
        for m in range(4):
            assert np.array_equal(coo.inds[m], inds[m][want])
        assert np.array_equal(coo.vals, vals[want])
        so = coo.sortorder
        assert list(so) == order
    for mode in range(4):
        coo = CooTensor(d, [a.copy() for a in inds], vals.copy())
        # unforced: the sort runs only when sortorder changes (mode 3 keeps
        # the identity order and so leaves the tensor alone)
        lib.sptSparseTensorSortIndexAtMode(coo.ref(), mode, 0, 4)
        if mode == 3:
            assert np.array_equal(coo.inds[0], inds[0])
            lib.sptSparseTensorSortIndexAtMode(coo.ref(), mode, 1, 4)
        order = [m for m in range(4) if m != mode] + [mode]
        want = np.lexsort([inds[m] for m in order[::-1]])
        for m in range(4):
            assert np.array_equal(coo.inds[m], inds[m][want])
        so = coo.sortorder
        assert list(so) == order


def test_mixed_order_is_hicoo_order(lib, golden_hicoo):
    """sptSparseTensorMixedOrder leaves the tensor in the order the
    reference's sptSparseTensorToHiCOO does (golden permuted tensors)."""
    g = golden_hicoo
    for key in g.cases():
        d, inds, vals = g.tensor(key, prefix="in_ind", vals="in_vals")
        sb, sk = (int(x) for x in g.get(key, "sb_sk"))
        coo = CooTensor(d, [a.copy() for a in inds], vals.copy())
        assert lib.sptSparseTensorMixedOrder(coo.ref(), sb, sk, 4) == 0
        for m in range(len(d)):
            assert np.array_equal(coo.inds[m], g.get(key, f"perm_ind{m}")), key
        assert np.array_equal(coo.vals.view(np.uint32), g.get(key, "perm_vals").view(np.uint32)), key


def test_status_text_matches_reference(lib, tmp_path):
    """sptSparseTensorStatus and sptSparseTensorStatusHiCOO print the
    reference's text byte for byte (golden text from the compiled
    reference)."""
    libc = _libc()
    g = np.load(os.path.join(GOLDEN, "coo_mttkrp.npz"), allow_pickle=False)
    h5 = np.load(os.path.join(GOLDEN, "hicoo.npz"), allow_pickle=False)
    names = sorted(f for f in os.listdir(GOLDEN) if f.startswith("status_"))
    assert len(names) == 5
    for fname in names:
        out = tmp_path / fname
        fp = libc.fopen(str(out).encode(), b"w")
        if fname.startswith("status_coo_"):
            case = fname[len("status_coo_"):-4]
            d = g[f"{case}/ndims"]
            coo = CooTensor(d, [g[f"{case}/ind{m}"] for m in range(len(d))], g[f"{case}/vals"])
            lib.sptSparseTensorStatus(coo.ref(), fp)
        else:
            key = fname[len("status_hicoo_"):-4]
            d = h5[f"{key}/ndims"]
            coo = CooTensor(d, [h5[f"{key}/in_ind{m}"].copy() for m in range(len(d))], h5[f"{key}/in_vals"].copy())
            sb, sk = (int(x) for x in h5[f"{key}/sb_sk"])
            h, _ = to_hicoo(lib, coo, sb, sk)
            lib.sptSparseTensorStatusHiCOO(C.byref(h), fp)
            lib.sptFreeSparseTensorHiCOO(C.byref(h))
        libc.fclose(fp)
        assert out.read_text() == open(os.path.join(GOLDEN, fname)).read(), fname


def test_small_print_helpers(lib, tmp_path):
    libc = _libc()
    for n, want in ((0, "0.00B"), (1024, "1024.00B"), (1025, "1.00KB"), (3 << 30, "3.00GB"),
                    (5 << 50, "5120.00TB")):
        p = lib.sptBytesString(n)
        assert C.string_at(p).decode() == want
        libc.free(p)
    out = tmp_path / "idx.txt"
    fp = libc.fopen(str(out).encode(), b"w")
    a = np.array([3, 1, 4], dtype=np.uint32)
    assert lib.sptDumpIndexArray(ptr(a, sptIndex), 3, fp) == 0
    libc.fclose(fp)
    assert out.read_text() == "sptIndex array length: 3\n3\t1\t4\t\n"


REF_EXAMPLES = "/root/reference/examples"


@pytest.mark.skipif(not os.path.isdir(REF_EXAMPLES), reason="reference tree absent")
@pytest.mark.parametrize("driver", ["cpd", "mttkrp", "mttkrp_hicoo"])
def test_reference_driver_builds_against_this_library(driver, tmp_path):
    """The reference's own example drivers, unchanged, compile against
    include/ParTI.h and link against libsptol_amd.so with no undefined
    symbol.  (examples/cpd_hicoo.c also needs the index-reordering
    subsystem, src/sptensor/reorder.c, which is not on this path.)"""
    exe = tmp_path / driver
    r = subprocess.run(["gcc", "-std=gnu99", "-O1", "-fopenmp", "-D_GNU_SOURCE", "-DPARTI_USE_OPENMP",
                        "-Werror=implicit-function-declaration", "-I", os.path.join(ROOT, "include"),
                        os.path.join(REF_EXAMPLES, driver + ".c"), "-o", str(exe),
                        "-L", os.path.join(ROOT, "sptol_amd"), "-lsptol_amd", "-Wl,--no-undefined", "-lm"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
