#!/usr/bin/env python3
"""Benchmark: HiCOO MTTKRP over all three modes on one MI355X.

Workload (the headline configuration): a third-order tensor of dims
12,000,000 x 9,000,000 x 28,000,000 with 1e8 distinct nonzeros, each mode's
index drawn from a bounded Zipf(1.1) over randomly permuted ids, values
lognormal(0,1) float32, HiCOO block 2^7 and superblock 2^14, float32
factors U[0,1) with R=32.  One step = one MTTKRP on each of the 3 modes
(mats_order = [n, n+1, n+2] mod 3), tensor, factors and outputs resident in
HBM.  The same tensor in COO form (one mode-sorted copy per mode) is timed
beside it.

Prints one JSON line (see DESIGN.md for the field definitions).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from sptol_amd import gen  # noqa: E402
from sptol_amd.parti import (CooTensor, Matrix, SPTOL_F32, check, load_library, ptr,  # noqa: E402
                             sptIndex, sptolDevMatrix, to_hicoo)

HBM_PEAK = 8.0e12
DIMS = [12_000_000, 9_000_000, 28_000_000]


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def lib_sort(lib):
    def sort_fn(inds):
        coo = CooTensor([0xFFFFFFFF] * len(inds), inds, np.zeros(len(inds[0]), np.float32))
        coo.sortorder[0] = 1
        lib.sptSparseTensorSortIndex(coo.ref(), 1, 0)
        return coo.inds
    return sort_fn


def distinct_count(idx, dim):
    flag = np.zeros(int(dim), dtype=bool)
    flag[idx] = True
    return int(flag.sum())


def sync(lib):
    check(lib, lib.sptolHipStreamSynchronize(None), "sync")


def prof(lib, name):
    ms, n = C.c_double(0), C.c_uint64(0)
    lib.sptolProfileQuery(name.encode(), C.byref(ms), C.byref(n))
    return ms.value, int(n.value)


def cpu_baseline(lib, coo_inds, vals, dims, factors, order_of, budget_s=20.0):
    """The reference's OpenMP superblock-scheduled HiCOO MTTKRP
    (hicoo/mttkrp_omp.c:158, compiled from the reference sources into
    oracle/_ref) on a random sample of the same tensor; falls back to the
    oracle's restatement of it when oracle/_ref is not present."""
    nnz = len(vals)
    ncpu = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    tk = max(2, min(16, ncpu))
    rng = np.random.default_rng(99)
    sample = min(nnz, 4_000_000)
    pick = np.sort(rng.choice(nnz, size=sample, replace=False))
    inds = [a[pick].copy() for a in coo_inds]
    sv = vals[pick].copy()
    R = factors[0].shape[1]
    ref_path = os.path.join(ROOT, "oracle", "_ref", "libparti_ref.so")
    kind = "reference"
    try:
        ref = C.CDLL(ref_path)
    except OSError:
        ref = None
        kind = "port"
    if ref is not None:
        from sptol_amd.parti import sptRankMatrix, sptSparseTensorHiCOO
        coo = CooTensor(dims, inds, sv)
        h, _ = to_hicoo(lib, coo, 7, 14)
        rms = []
        for f in factors:
            rms.append(sptRankMatrix(f.shape[0], R, f.shape[0], R, ptr(f, C.c_float)))
        out = np.zeros((max(dims), R), dtype=np.float32)
        rms.append(sptRankMatrix(max(dims), R, max(dims), R, ptr(out, C.c_float)))
        arr = (C.POINTER(sptRankMatrix) * 4)(*[C.pointer(r) for r in rms])
        fn = ref.sptOmpMTTKRPHiCOO_MatrixTiling_Scheduled
        fn.argtypes = [C.POINTER(sptSparseTensorHiCOO), C.POINTER(C.POINTER(sptRankMatrix)),
                       C.POINTER(sptIndex), sptIndex, C.c_int, C.c_int]
        done, t_tot = 0, 0.0
        reps = 0
        while reps < 1 or (t_tot < budget_s and reps < 5):
            for mode in range(3):
                o = order_of(mode)
                t0 = time.perf_counter()
                rc = fn(C.byref(h), arr, ptr(o, sptIndex), mode, tk, 0)
                t_tot += time.perf_counter() - t0
                assert rc == 0
                done += sample
            reps += 1
        lib.sptFreeSparseTensorHiCOO(C.byref(h))
        what = "sptOmpMTTKRPHiCOO_MatrixTiling_Scheduled (hicoo/mttkrp_omp.c:158)"
    else:
        import oracle
        oh = oracle.Hicoo(dims, inds, sv, 7, 14)
        done, t_tot = 0, 0.0
        for mode in range(3):
            t0 = time.perf_counter()
            oh.mttkrp(factors, order_of(mode), mode, threads=tk)
            t_tot += time.perf_counter() - t0
            done += sample
        what = "oracle restatement of the superblock-scheduled HiCOO MTTKRP"
    return {"value": done / t_tot / 1e9, "unit": "Gnnz/s", "cores": tk, "kind": kind,
            "sample": f"{what}, tk={tk}, {sample} nnz drawn at random from the benchmark tensor, "
                      f"HiCOO sb=2^7 sk=2^14, R={R}, all 3 modes x {done // (3 * sample)} passes, "
                      f"{t_tot:.1f} s of CPU time"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nnz", type=float, default=1e8)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-coo", action="store_true", help="skip the COO comparison")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (see DESIGN.md, "
              "'How the path would shard'); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    lib = load_library()
    n = C.c_int(0)
    if lib.sptolHipDeviceCount(C.byref(n)) != 0 or n.value < 1:
        print("no GPU visible to HIP", file=sys.stderr)
        sys.exit(1)
    nnz = int(args.nnz)
    R = 32
    t0 = time.time()
    dims, inds, vals = gen.zipf_tensor(DIMS, nnz, 1.1, seed=20260101, values="lognormal",
                                       sort_fn=lib_sort(lib))
    nnz = len(vals)
    log(f"generated {nnz} nnz in {time.time() - t0:.1f}s")
    U = [distinct_count(inds[m], dims[m]) for m in range(3)]
    factors = gen.uniform_factors(dims, R, seed=7)
    coo_keep = [a.copy() for a in inds]

    t0 = time.time()
    hcoo = CooTensor(dims, inds, vals.copy())
    h, max_nnzb = to_hicoo(lib, hcoo, 7, 14)
    nb = int(h.bptr.len) - 1
    nk = int(h.kptr.len) - 1
    log(f"HiCOO: nb={nb} nk={nk} max_nnzb={max_nnzb} in {time.time() - t0:.1f}s")

    dh = C.c_void_p()
    check(lib, lib.sptolDevHicooCreate(C.byref(dh), C.byref(h)), "DevHicooCreate")
    t0 = time.time()
    for m in range(3):
        check(lib, lib.sptolDevHicooPrepareMode(dh, m), "DevHicooPrepareMode")
    log(f"HiCOO per-mode views built in {time.time() - t0:.1f}s")
    dm = (sptolDevMatrix * 3)()
    hm = [Matrix(f) for f in factors]
    for m in range(3):
        check(lib, lib.sptolDevMatrixCreate(C.byref(dm[m]), int(dims[m]), R, SPTOL_F32), "DevMatrixCreate")
        check(lib, lib.sptolDevMatrixUpload(C.byref(dm[m]), C.byref(hm[m].m)), "DevMatrixUpload")
    del hm
    outs = (sptolDevMatrix * 3)()
    for m in range(3):
        check(lib, lib.sptolDevMatrixCreate(C.byref(outs[m]), int(dims[m]), R, SPTOL_F32), "DevMatrixCreate")
    orders = [np.array([m, (m + 1) % 3, (m + 2) % 3], dtype=np.uint32) for m in range(3)]

    def step_hicoo():
        for m in range(3):
            check(lib, lib.sptolDevMTTKRPHiCOO(dh, dm, ptr(orders[m], sptIndex), m, C.byref(outs[m]), None),
                  "DevMTTKRPHiCOO")

    def timed(step, kname):
        for _ in range(args.warmup):
            step()
        sync(lib)
        lib.sptolProfileReset()
        lib.sptolProfileEnable(1)
        sync(lib)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        sync(lib)
        dt = time.perf_counter() - t0
        lib.sptolProfileEnable(0)
        kms, kn = prof(lib, kname)
        fms, fn = prof(lib, "fill_zero")
        return dt, kms / max(kn, 1), fms / max(fn, 1)

    dt, k_ms, fill_ms = timed(step_hicoo, "mttkrp_hicoo")
    ms_step = dt / args.steps * 1e3
    value = 3 * nnz * args.steps / dt / 1e9
    log(f"HiCOO: {ms_step:.3f} ms/step, {value:.2f} Gnnz/s, kernel avg {k_ms:.3f} ms, fill avg {fill_ms:.3f} ms")

    # algorithmic bytes per launch (averaged over the 3 modes)
    S_h = nnz * (3 + 4) + nb * (4 * 3 + 8) + (nk + 1) * 8 + nk * 4
    B = []
    for mode in range(3):
        G = sum(U[m] * R * 4 for m in range(3) if m != mode)
        W = int(dims[mode]) * R * 4
        B.append(S_h + G + W)
    B_avg = float(np.mean(B))
    achieved = B_avg / (k_ms * 1e-3)

    result = {
        "metric": "MTTKRP Gnnz/s & %HBM roofline",
        "value": round(value, 4),
        "unit": "Gnnz/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "hicoo_mttkrp_all_modes", "dims": DIMS, "nnz": nnz, "zipf_s": 1.1,
                   "values": "lognormal(0,1)", "R": R, "sb_bits": 7, "sk_bits": 14,
                   "nb": nb, "nk": nk, "distinct_per_mode": U},
        "roofline": {"bound": "hbm", "achieved": round(achieved / 1e9, 2), "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 4), "traffic": None,
                     "kernel": "k_mttkrp_hicoo", "kernel_avg_ms": round(k_ms, 4),
                     "algorithmic_bytes_per_launch": int(B_avg), "fill_zero_avg_ms": round(fill_ms, 4)},
    }

    if not args.no_coo:
        dc = C.c_void_p()
        ccoo = CooTensor(dims, coo_keep, vals)
        check(lib, lib.sptolDevCooCreate(C.byref(dc), ccoo.ref()), "DevCooCreate")
        for m in range(3):
            check(lib, lib.sptolDevCooPrepareMode(dc, m), "PrepareMode")

        def step_coo():
            for m in range(3):
                check(lib, lib.sptolDevMTTKRP(dc, dm, ptr(orders[m], sptIndex), m, C.byref(outs[m]), None),
                      "DevMTTKRP")

        dt_c, kc_ms, _ = timed(step_coo, "mttkrp_coo")
        S_c = nnz * (4 * 3 + 4)
        Bc = float(np.mean([S_c + sum(U[m] * R * 4 for m in range(3) if m != mode) + int(dims[mode]) * R * 4
                            for mode in range(3)]))
        result["coo_same_shape"] = {
            "value": round(3 * nnz * args.steps / dt_c / 1e9, 4), "unit": "Gnnz/s",
            "ms_per_step": round(dt_c / args.steps * 1e3, 4), "kernel_avg_ms": round(kc_ms, 4),
            "roofline_frac": round(Bc / (kc_ms * 1e-3) / HBM_PEAK, 4),
            "layout": "one copy per mode, sorted by that mode on the device"}
        lib.sptolDevCooFree(dc)
        log(f"COO: {result['coo_same_shape']}")

    lib.sptolDevHicooFree(dh)
    for m in range(3):
        lib.sptolDevMatrixFree(C.byref(dm[m]))
        lib.sptolDevMatrixFree(C.byref(outs[m]))
    lib.sptFreeSparseTensorHiCOO(C.byref(h))

    if not args.no_cpu:
        t0 = time.time()
        result["cpu_baseline"] = cpu_baseline(lib, coo_keep, vals, dims, factors, lambda m: orders[m])
        log(f"CPU baseline in {time.time() - t0:.1f}s: {result['cpu_baseline']}")
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
