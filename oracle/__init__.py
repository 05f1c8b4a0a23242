"""ctypes bindings of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU
baseline, never as the product path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
P = C.POINTER


class OracleHicoo(C.Structure):
    _fields_ = [("nmodes", C.c_uint32), ("sb_bits", C.c_uint32), ("sk_bits", C.c_uint32),
                ("sc_bits", C.c_uint32), ("nnz", C.c_uint64), ("nb", C.c_uint64), ("nk", C.c_uint64),
                ("nc", C.c_uint64), ("max_nnzb", C.c_uint64),
                ("kptr", P(C.c_uint64)), ("cptr", P(C.c_uint64)), ("bptr", P(C.c_uint64)),
                ("binds", P(C.c_uint32)), ("einds", P(C.c_uint8)), ("values", P(C.c_float)),
                ("nkiters", P(C.c_uint32)), ("nkd", P(C.c_uint32)),
                ("kschr_ptr", P(C.c_uint64)), ("kschr_idx", P(C.c_uint32))]


_LIB = None


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run `make -C oracle oracle`")
        L = C.CDLL(path)
        u32p, f32p = P(C.c_uint32), P(C.c_float)
        common = [C.c_uint32, C.c_uint64, P(u32p), f32p, P(f32p), C.c_uint32, C.c_uint32, u32p,
                  C.c_uint32, C.c_uint32]
        L.oracle_mttkrp_coo_f32.argtypes = common + [f32p]
        L.oracle_mttkrp_coo_f64.argtypes = common + [P(C.c_double)]
        L.oracle_mttkrp_coo_omp_f32.argtypes = common + [f32p, C.c_int]
        L.oracle_mttkrp_coo_omp_reduce_f32.argtypes = common + [f32p, C.c_int]
        L.oracle_sort_lex.argtypes = [C.c_uint32, C.c_uint64, P(u32p), f32p]
        L.oracle_morton_cmp.argtypes = [C.c_uint32, u32p, u32p]
        L.oracle_hicoo_convert.argtypes = [C.c_uint32, u32p, C.c_uint64, P(u32p), f32p, C.c_uint32,
                                           C.c_uint32, P(OracleHicoo)]
        L.oracle_hicoo_free.argtypes = [P(OracleHicoo)]
        hc = [P(OracleHicoo), P(f32p), C.c_uint32, C.c_uint32, u32p, C.c_uint32, C.c_uint32]
        L.oracle_mttkrp_hicoo_f32.argtypes = hc + [f32p]
        L.oracle_mttkrp_hicoo_f64.argtypes = hc + [P(C.c_double)]
        L.oracle_mttkrp_hicoo_sched_omp_f32.argtypes = hc + [f32p, C.c_int]
        L.oracle_mttkrp_hicoo_sched_reduce_omp_f32.argtypes = hc + [f32p, C.c_int]
        L.oracle_cpd_als_f64.argtypes = [C.c_uint32, u32p, C.c_uint64, P(u32p), f32p, C.c_uint32,
                                         C.c_uint32, C.c_double, P(P(C.c_double)), P(C.c_double),
                                         P(C.c_double)]
        L.oracle_cpd_als_f64.restype = C.c_int
        _LIB = L
    return _LIB


def _u32pp(arrs):
    a = (P(C.c_uint32) * len(arrs))()
    for i, x in enumerate(arrs):
        a[i] = x.ctypes.data_as(P(C.c_uint32))
    return a


def _f32pp(arrs):
    a = (P(C.c_float) * len(arrs))()
    for i, x in enumerate(arrs):
        a[i] = x.ctypes.data_as(P(C.c_float)) if x is not None else None
    return a


def _padded(factors, stride):
    out = []
    for f in factors:
        b = np.zeros((f.shape[0], stride), dtype=np.float32)
        b[:, : f.shape[1]] = f
        out.append(b)
    return out


def _stride(R):
    return ((R - 1) // 8 + 1) * 8


def mttkrp_coo(inds, vals, factors, order, mode, out_rows=None, dtype=np.float32, threads=0, reduce=False):
    """Reference-order MTTKRP; returns (out_rows, stride) array (padding zero).
    dtype float32 -> sequential FP32 restatement (threads=0), the OpenMP
    atomic restatement (threads>0) or, with reduce, the private-copy
    restatement of sptOmpMTTKRP_Reduce; float64 -> FP64 oracle."""
    L = lib()
    N = len(inds)
    R = factors[0].shape[1]
    S = _stride(R)
    rows = int(out_rows if out_rows is not None else factors[mode].shape[0])
    inds = [np.ascontiguousarray(i, dtype=np.uint32) for i in inds]
    vals = np.ascontiguousarray(vals, dtype=np.float32)
    pf = _padded(factors, S)
    o = np.ascontiguousarray(order, dtype=np.uint32)
    out = np.zeros((max(rows, 1), S), dtype=dtype)
    args = [N, len(vals), _u32pp(inds), vals.ctypes.data_as(P(C.c_float)), _f32pp(pf), S, R,
            o.ctypes.data_as(P(C.c_uint32)), mode, rows]
    if dtype == np.float64:
        L.oracle_mttkrp_coo_f64(*args, out.ctypes.data_as(P(C.c_double)))
    elif threads > 0 and reduce:
        L.oracle_mttkrp_coo_omp_reduce_f32(*args, out.ctypes.data_as(P(C.c_float)), threads)
    elif threads > 0:
        L.oracle_mttkrp_coo_omp_f32(*args, out.ctypes.data_as(P(C.c_float)), threads)
    else:
        L.oracle_mttkrp_coo_f32(*args, out.ctypes.data_as(P(C.c_float)))
    return out[:rows]


def sort_lex(inds, vals):
    inds = [np.ascontiguousarray(i, dtype=np.uint32).copy() for i in inds]
    vals = np.ascontiguousarray(vals, dtype=np.float32).copy()
    lib().oracle_sort_lex(len(inds), len(vals), _u32pp(inds), vals.ctypes.data_as(P(C.c_float)))
    return inds, vals


def morton_cmp(a, b):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(b, dtype=np.uint32)
    return lib().oracle_morton_cmp(len(a), a.ctypes.data_as(P(C.c_uint32)), b.ctypes.data_as(P(C.c_uint32)))


class Hicoo:
    """Oracle HiCOO tensor (owns the C struct; arrays exposed as numpy views)."""

    def __init__(self, ndims, inds, vals, sb_bits, sk_bits):
        self.ndims = np.ascontiguousarray(ndims, dtype=np.uint32)
        self.inds = [np.ascontiguousarray(i, dtype=np.uint32).copy() for i in inds]
        self.vals = np.ascontiguousarray(vals, dtype=np.float32).copy()
        self.h = OracleHicoo()
        rc = lib().oracle_hicoo_convert(len(self.ndims), self.ndims.ctypes.data_as(P(C.c_uint32)),
                                        len(self.vals), _u32pp(self.inds),
                                        self.vals.ctypes.data_as(P(C.c_float)), sb_bits, sk_bits,
                                        C.byref(self.h))
        if rc != 0:
            raise ValueError("oracle_hicoo_convert failed")

    def __del__(self):
        try:
            lib().oracle_hicoo_free(C.byref(self.h))
        except Exception:
            pass

    def arrays(self):
        h = self.h
        N, nb, nnz, nk = h.nmodes, int(h.nb), int(h.nnz), int(h.nk)
        A = np.ctypeslib.as_array
        out = {
            "kptr": A(h.kptr, shape=(nk + 1,)).copy(),
            "cptr": A(h.cptr, shape=(int(h.nc),)).copy(),
            "bptr": A(h.bptr, shape=(nb + 1,)).copy(),
            "binds": [A(h.binds, shape=(N * nb,))[m * nb:(m + 1) * nb].copy() for m in range(N)],
            "einds": [A(h.einds, shape=(N * nnz,))[m * nnz:(m + 1) * nnz].copy() for m in range(N)],
            "values": A(h.values, shape=(nnz,)).copy(),
            "nkiters": A(h.nkiters, shape=(N,)).copy(),
            "max_nnzb": int(h.max_nnzb),
        }
        nkd = A(h.nkd, shape=(N,)).copy()
        ptrs = A(h.kschr_ptr, shape=(int(nkd.sum()) + N,))
        idx = A(h.kschr_idx, shape=(N * nk,)) if nk else np.zeros(0, dtype=np.uint32)
        ks, off = [], 0
        for m in range(N):
            p = ptrs[off: off + nkd[m] + 1].copy()
            ks.append((p, idx[m * nk:(m + 1) * nk].copy()))
            off += nkd[m] + 1
        out["kschr"] = ks
        return out

    def mttkrp(self, factors, order, mode, dtype=np.float32, threads=0, reduce=False):
        L = lib()
        R = factors[0].shape[1]
        S = _stride(R)
        pf = _padded(factors, S)
        o = np.ascontiguousarray(order, dtype=np.uint32)
        rows = int(self.ndims[mode])
        out = np.zeros((max(rows, 1), S), dtype=dtype)
        args = [C.byref(self.h), _f32pp(pf), S, R, o.ctypes.data_as(P(C.c_uint32)), mode, rows]
        if dtype == np.float64:
            L.oracle_mttkrp_hicoo_f64(*args, out.ctypes.data_as(P(C.c_double)))
        elif threads > 0 and reduce:
            L.oracle_mttkrp_hicoo_sched_reduce_omp_f32(*args, out.ctypes.data_as(P(C.c_float)), threads)
        elif threads > 0:
            L.oracle_mttkrp_hicoo_sched_omp_f32(*args, out.ctypes.data_as(P(C.c_float)), threads)
        else:
            L.oracle_mttkrp_hicoo_f32(*args, out.ctypes.data_as(P(C.c_float)))
        return out[:rows]


def cpd_als(ndims, inds, vals, init_factors, niters, tol=-1.0):
    """FP64 CPD-ALS oracle.  Returns (factors, lambda, fits)."""
    L = lib()
    N = len(ndims)
    R = init_factors[0].shape[1]
    nd = np.ascontiguousarray(ndims, dtype=np.uint32)
    inds = [np.ascontiguousarray(i, dtype=np.uint32) for i in inds]
    vals = np.ascontiguousarray(vals, dtype=np.float32)
    mats = [np.ascontiguousarray(f, dtype=np.float64).copy() for f in init_factors]
    mp = (P(C.c_double) * N)()
    for m in range(N):
        mp[m] = mats[m].ctypes.data_as(P(C.c_double))
    lam = np.zeros(R, dtype=np.float64)
    fits = np.zeros(max(niters, 1), dtype=np.float64)
    it = L.oracle_cpd_als_f64(N, nd.ctypes.data_as(P(C.c_uint32)), len(vals), _u32pp(inds),
                              vals.ctypes.data_as(P(C.c_float)), R, niters, tol, mp,
                              lam.ctypes.data_as(P(C.c_double)), fits.ctypes.data_as(P(C.c_double)))
    return mats, lam, fits[:it]
