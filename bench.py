#!/usr/bin/env python3
"""Benchmark: HiCOO MTTKRP over all three modes on one MI355X.

Workload (the headline configuration): a third-order tensor of dims
12,000,000 x 9,000,000 x 28,000,000 with 1e8 distinct nonzeros, each mode's
index drawn from a bounded Zipf(1.1) over randomly permuted ids, values
lognormal(0,1) float32, HiCOO block 2^7 and superblock 2^14, float32
factors U[0,1) with R=32.  One step = one MTTKRP on each of the 3 modes
(mats_order = [n, n+1, n+2] mod 3), tensor, factors and outputs resident in
HBM.  The same tensor in COO form (one mode-sorted view per mode) is timed
beside it.  The same line carries the served half of the metric
(`served_cpd`): 400 requests of the served CPD-ALS stream at t=0, solves/s
with p50/p99 latency, against the same requests through the plain
one-at-a-time device path.  Other workloads: --workload coo | cpd | batch |
serve | ingest | copy.

Prints one JSON line: metric, value, unit, timing fields, the workload's
config, `roofline` for the dominant kernel (algorithmic bytes per launch
over its average duration measured with HIP events on its launch stream,
against 8 TB/s; `traffic` = measured HBM bytes per launch from the committed
rocprofv3 counter summary), `cpu_baseline` (the compiled reference's OpenMP
COO MTTKRP on the full tensor, on this host's cores; the oracle's
restatement where the reference build is absent) and the reference's
superblock-scheduled HiCOO kernel on a 4M-nonzero sample beside it.
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
from sptol_amd.parti import (CooTensor, Matrix, SPTOL_BF16, SPTOL_F32, check, load_library, ptr,  # noqa: E402
                             sptIndex, sptSparseTensor, sptolDevMatrix, to_hicoo, to_hicoo_device)

HBM_PEAK = 8.0e12
DIMS = [12_000_000, 9_000_000, 28_000_000]
HICOO_NNZ = 100_000_000
SERVE_REQUESTS_DEFAULT = 400
METRIC = "MTTKRP Gnnz/s & %HBM roofline; CPD-ALS solves/s + p99 vs plain, 1 GPU"
# counter summary of k_mttkrp_seg on the full-size HiCOO workload
TRAFFIC_PROFILE = "k_mttkrp_seg_hicoo_12Mx9Mx28M_100Mnnz_R32"


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


def cpu_baseline_full(coo_inds, vals, dims, factors, order_of):
    """The reference's OpenMP COO MTTKRP sptOmpMTTKRP (mttkrp_omp.c:125,
    compiled from the reference sources into oracle/_ref) on the full
    benchmark tensor, every mode once, on the box's cores (at most 16); the
    oracle's restatement of it when oracle/_ref is absent.  On this tensor it
    is the reference's fastest CPU path (its superblock-scheduled HiCOO
    kernel loops over every superblock row for each of nkiters steps)."""
    tk = host_threads()
    ref = ref_library()
    nnz = len(vals)
    R = factors[0].shape[1]
    t_tot = 0.0
    if ref is not None:
        from sptol_amd.parti import mats_array
        coo = CooTensor(dims, coo_inds, vals)
        out = Matrix(nrows=int(max(dims)), ncols=R)
        mats = [Matrix(f) for f in factors] + [out]
        arr = mats_array(mats)
        ref.sptOmpMTTKRP.argtypes = [C.POINTER(sptSparseTensor), C.c_void_p, C.POINTER(sptIndex), sptIndex, C.c_int]
        with quiet_stdout():
            for mode in range(3):
                o = np.ascontiguousarray(order_of(mode), dtype=np.uint32)
                out.m.nrows = int(dims[mode])
                t0 = time.perf_counter()
                rc = ref.sptOmpMTTKRP(coo.ref(), C.cast(arr, C.c_void_p), ptr(o, sptIndex), mode, tk)
                t_tot += time.perf_counter() - t0
                assert rc == 0
        what, kind = "the compiled reference's sptOmpMTTKRP (COO, atomic, mttkrp_omp.c:125)", "reference"
    else:
        import oracle
        for mode in range(3):
            t0 = time.perf_counter()
            oracle.mttkrp_coo(coo_inds, vals, factors, order_of(mode), mode, threads=tk)
            t_tot += time.perf_counter() - t0
        what, kind = "oracle restatement of sptOmpMTTKRP (COO, atomic, mttkrp_omp.c:125)", "port"
    return {"value": 3 * nnz / t_tot / 1e9, "unit": "Gnnz/s", "cores": tk, "kind": kind,
            "sample": f"{what}, tk={tk}, the full {nnz}-nonzero benchmark tensor, all 3 modes once, "
                      f"R={R}, {t_tot:.1f} s of CPU time"}


def cpu_baseline(lib, coo_inds, vals, dims, factors, order_of, budget_s=8.0):
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


KERNEL_SYMBOL = {"mttkrp_seg": "k_mttkrp_seg", "mttkrp_hicoo": "k_mttkrp_hicoo_tiled", "mttkrp_coo": "k_mttkrp_coo"}


def profile_traffic(name):
    """HBM bytes per launch measured by rocprofv3 FETCH_SIZE / WRITE_SIZE runs
    (scaled by the copy calibration), read from the committed summary
    profiles/<name>.md (its `traffic_bytes_per_launch:` line); None when the
    summary is absent."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", name + ".md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def alg_bytes(kname, nnz, dims, U, mode, R, nb=0, nk=0, factor_bytes=4):
    """Algorithmic HBM bytes of one MTTKRP launch on `mode`: the nonzero
    stream, one read of each distinct factor row gathered (U[m] rows of R
    values), and the output rows the kernel writes.
      mttkrp_seg:   16-byte records (N=3; 20 with the value array for N=4)
                    + the row bitmap; writes all I_n rows (zeros included).
      mttkrp_hicoo: HiCOO stream (u8 element offsets, values, per-block u32
                    block coordinates and u64 pointers, kernel pointers);
                    writes the U[mode] touched rows (the zero fill is a
                    separate launch).
      mttkrp_coo:   u32 indices + values; writes the touched rows."""
    N = len(dims)
    gather = sum(U[m] * R * factor_bytes for m in range(N) if m != mode)
    if kname == "mttkrp_seg":
        stream = nnz * (16 if N == 3 else 20) + (int(dims[mode]) + 31) // 32 * 4
        write = int(dims[mode]) * R * 4
    elif kname == "mttkrp_hicoo":
        stream = nnz * (N + 4) + nb * (4 * N + 8) + (nk + 1) * 8 + nk * 4
        write = U[mode] * R * 4
    else:
        stream = nnz * (4 * N + 4)
        write = U[mode] * R * 4
    return stream + gather + write


def timed_steps(lib, step, kname, steps, warmup, extra=("fill_zero", "mttkrp_seg_carry")):
    """Warm up, then time exactly `steps` calls of step() between device
    synchronisations (the throughput; no instrumentation).  A second pass of
    the same `steps` calls records hipEvents around every launch on its
    stream (sptolProfile*) for the kernel durations."""
    for _ in range(warmup):
        step()
    sync(lib)
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    sync(lib)
    dt = time.perf_counter() - t0
    lib.sptolProfileReset()
    lib.sptolProfileEnable(1)
    sync(lib)
    for _ in range(steps):
        step()
    sync(lib)
    lib.sptolProfileEnable(0)
    kms, kn = prof(lib, kname)
    if kn == 0:   # the segmented kernel ran instead
        kname = "mttkrp_seg"
        kms, kn = prof(lib, kname)
    out = {"dt": dt, "kernel_ms": kms / max(kn, 1), "launches": kn, "kernel": kname}
    for e in extra:
        ems, en = prof(lib, e)
        out[e + "_ms"] = ems / max(en, 1)
        out[e + "_launches"] = en
    return out


def run_hicoo(args, lib):
    nnz = int(args.nnz)
    full = nnz == HICOO_NNZ
    R = 32
    t0 = time.time()
    dims, inds, vals = gen.zipf_tensor(DIMS, nnz, 1.1, seed=20260101, values="lognormal",
                                       sort_fn=lib_sort(lib))
    nnz = len(vals)
    log(f"generated {nnz} nnz in {time.time() - t0:.1f}s")
    U = [distinct_count(inds[m], dims[m]) for m in range(3)]
    factors = gen.uniform_factors(dims, R, seed=7)
    coo_keep = [a.copy() for a in inds]

    hcoo = CooTensor(dims, inds, vals.copy())
    t0 = time.perf_counter()
    h, max_nnzb = to_hicoo(lib, hcoo, 7, 14)
    t_host = time.perf_counter() - t0
    nb = int(h.bptr.len) - 1
    nk = int(h.kptr.len) - 1
    log(f"HiCOO: nb={nb} nk={nk} max_nnzb={max_nnzb} in {t_host:.1f}s")
    # the GPU converter on a copy of the same input: COO upload, sorts and
    # block arrays on the device, HiCOO arrays (and the permuted COO) back
    dcoo = CooTensor(dims, [a.copy() for a in coo_keep], vals.copy())
    sync(lib)
    t0 = time.perf_counter()
    hd, _ = to_hicoo_device(lib, dcoo, 7, 14)
    t_dev = time.perf_counter() - t0
    lib.sptFreeSparseTensorHiCOO(C.byref(hd))
    del dcoo
    log(f"HiCOO on the GPU in {t_dev:.2f}s")

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

    hargs = [(dh, dm, ptr(orders[m], sptIndex), m, C.byref(outs[m]), None) for m in range(3)]

    def step_hicoo():
        for a in hargs:
            check(lib, lib.sptolDevMTTKRPHiCOO(*a), "DevMTTKRPHiCOO")

    t = timed_steps(lib, step_hicoo, "mttkrp_hicoo", args.steps, args.warmup)
    dt, k_ms, kname = t["dt"], t["kernel_ms"], t["kernel"]
    ms_step = dt / args.steps * 1e3
    value = 3 * nnz * args.steps / dt / 1e9
    log(f"HiCOO: {ms_step:.3f} ms/step, {value:.2f} Gnnz/s, {kname} avg {k_ms:.3f} ms, "
        f"carry avg {t['mttkrp_seg_carry_ms']:.4f} ms, fill avg {t['fill_zero_ms']:.3f} ms")
    B_avg = float(np.mean([alg_bytes(kname, nnz, dims, U, mode, R, nb=nb, nk=nk) for mode in range(3)]))
    achieved = B_avg / (k_ms * 1e-3)

    result = {
        "metric": METRIC,
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
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 4),
                     "traffic": profile_traffic(TRAFFIC_PROFILE) if full else None,
                     "traffic_source": f"profiles/{TRAFFIC_PROFILE}.md" if full else None,
                     # measured bytes (FETCH + WRITE) over the same duration: the counter-based fraction
                     "traffic_frac": (round(profile_traffic(TRAFFIC_PROFILE) / (k_ms * 1e-3) / HBM_PEAK, 4)
                                      if full and profile_traffic(TRAFFIC_PROFILE) else None),
                     "kernel": KERNEL_SYMBOL[kname], "kernel_avg_ms": round(k_ms, 4),
                     "algorithmic_bytes_per_launch": int(B_avg),
                     "carry_kernel_avg_ms": round(t["mttkrp_seg_carry_ms"], 4),
                     "fill_zero_avg_ms": round(t["fill_zero_ms"], 4)},
        # sptSparseTensorToHiCOO (convert.c:300) on the host and
        # sptolSparseTensorToHiCOODevice, host arrays in and out, same layout
        "hicoo_conversion": {"host_s": round(t_host, 3), "device_s": round(t_dev, 3),
                             "host_threads": host_threads()},
    }

    if not args.no_coo:
        dc = C.c_void_p()
        ccoo = CooTensor(dims, coo_keep, vals)
        check(lib, lib.sptolDevCooCreate(C.byref(dc), ccoo.ref()), "DevCooCreate")
        for m in range(3):
            check(lib, lib.sptolDevCooPrepareMode(dc, m), "PrepareMode")

        cargs = [(dc, dm, ptr(orders[m], sptIndex), m, C.byref(outs[m]), None) for m in range(3)]

        def step_coo():
            for a in cargs:
                check(lib, lib.sptolDevMTTKRP(*a), "DevMTTKRP")

        tc = timed_steps(lib, step_coo, "mttkrp_coo", args.steps, args.warmup)
        dt_c, kc_ms = tc["dt"], tc["kernel_ms"]
        Bc = float(np.mean([alg_bytes(tc["kernel"], nnz, dims, U, mode, R) for mode in range(3)]))
        result["coo_same_shape"] = {
            "value": round(3 * nnz * args.steps / dt_c / 1e9, 4), "unit": "Gnnz/s",
            "ms_per_step": round(dt_c / args.steps * 1e3, 4), "kernel": KERNEL_SYMBOL[tc["kernel"]],
            "kernel_avg_ms": round(kc_ms, 4), "roofline_frac": round(Bc / (kc_ms * 1e-3) / HBM_PEAK, 4),
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
        result["cpu_baseline"] = cpu_baseline_full(coo_keep, vals, dims, factors, lambda m: orders[m])
        log(f"CPU baseline in {time.time() - t0:.1f}s: {result['cpu_baseline']}")
        t0 = time.time()
        result["cpu_hicoo_scheduled_sample"] = cpu_baseline(lib, coo_keep, vals, dims, factors, lambda m: orders[m])
        log(f"CPU HiCOO scheduled sample in {time.time() - t0:.1f}s: {result['cpu_hicoo_scheduled_sample']}")
    else:
        result["cpu_baseline"] = None
    del coo_keep, vals, factors
    if not args.no_serve:
        # the served half of the metric: 400 requests of the served CPD stream
        # at t=0, and the same requests through the plain one-at-a-time path
        sargs = argparse.Namespace(nnz=None, requests=SERVE_REQUESTS_DEFAULT, rate=0.0, no_plain=False,
                                   steps=args.steps, warmup=args.warmup)
        sv = run_serve(sargs, lib)
        result["served_cpd"] = {"value": sv["value"], "unit": "solves/s", "requests": SERVE_REQUESTS_DEFAULT,
                                "latency_s": sv["latency_s"], "plain_one_at_a_time": sv.get("plain_one_at_a_time"),
                                "speedup_vs_plain": sv.get("speedup_vs_plain"),
                                "mean_iterations": sv["mean_iterations"], "config": sv["config"]}
    return result


# ---------------------------------------------------------------------------
# Other workloads (parity-test configurations measured on request)
# ---------------------------------------------------------------------------
def base_result(args, metric, value, unit, ms_step, config, higher=True, dtype="f32"):
    return {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": higher, "scaling": "weak",
            "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": config}


def dev_factors(lib, factors, dtype=SPTOL_F32):
    arr = (sptolDevMatrix * len(factors))()
    for m, f in enumerate(factors):
        hm = Matrix(f)
        check(lib, lib.sptolDevMatrixCreate(C.byref(arr[m]), f.shape[0], f.shape[1], dtype), "create")
        check(lib, lib.sptolDevMatrixUpload(C.byref(arr[m]), C.byref(hm.m)), "upload")
    return arr


def free_factors(lib, arr):
    for m in range(len(arr)):
        lib.sptolDevMatrixFree(C.byref(arr[m]))


def dropin_rate(lib, coo, factors, dims, R, order, nnz, reps=20):
    """The reference-named sptMTTKRP (mttkrp.c:39) on host arrays, as a
    reference caller uses it: every call uploads the tensor and the factors,
    runs the MTTKRP on the GPU and copies the output back, so the rate
    includes PCIe and the per-call upload (not the `value` of any line)."""
    from sptol_amd.parti import mats_array
    out = Matrix(nrows=int(dims[0]), ncols=R)
    mats = [Matrix(f) for f in factors] + [out]
    arr = mats_array(mats)
    o = ptr(order, sptIndex)
    check(lib, lib.sptMTTKRP(coo.ref(), arr, o, 0), "sptMTTKRP")   # warm-up (allocator, module load)
    t0 = time.perf_counter()
    for _ in range(reps):
        check(lib, lib.sptMTTKRP(coo.ref(), arr, o, 0), "sptMTTKRP")
    dt = time.perf_counter() - t0
    return {"value": round(nnz * reps / dt / 1e9, 4), "unit": "Gnnz/s", "ms_per_call": round(dt / reps * 1e3, 4),
            "calls": reps, "what": "sptMTTKRP with host arrays in and out (upload, GPU MTTKRP, download per call)"}


def run_coo(args, lib):
    """3-D COO MTTKRP, mode 0: 2000x1500x1000, 1e6 uniform nonzeros, R=16
    (the reference driver's configuration, examples/mttkrp.c)."""
    dims = [2000, 1500, 1000]
    nnz = int(args.nnz or 1_000_000)
    R = 16
    d, inds, vals = gen.uniform_tensor(dims, nnz, seed=2024, sort_fn=lib_sort(lib))
    nnz = len(vals)
    factors = gen.uniform_factors(d, R, seed=2025)
    coo = CooTensor(d, inds, vals)
    h = C.c_void_p()
    check(lib, lib.sptolDevCooCreate(C.byref(h), coo.ref()), "DevCooCreate")
    check(lib, lib.sptolDevCooPrepareMode(h, 0), "prepare")
    dm = dev_factors(lib, factors)
    out = sptolDevMatrix()
    check(lib, lib.sptolDevMatrixCreate(C.byref(out), dims[0], R, SPTOL_F32), "create")
    order = np.array([0, 1, 2], np.uint32)
    steps = max(args.steps, 200)

    sargs = (h, dm, ptr(order, sptIndex), 0, C.byref(out), None)

    def step():
        check(lib, lib.sptolDevMTTKRP(*sargs), "mttkrp")

    t = timed_steps(lib, step, "mttkrp_coo", steps, args.warmup)
    U = [distinct_count(inds[m], d[m]) for m in range(3)]
    B = alg_bytes(t["kernel"], nnz, dims, U, 0, R)
    k_ms = t["kernel_ms"]
    res = base_result(args, "MTTKRP Gnnz/s & %HBM roofline", round(nnz * steps / t["dt"] / 1e9, 4), "Gnnz/s",
                      round(t["dt"] / steps * 1e3, 5),
                      {"workload": "coo_mttkrp_mode0", "dims": dims, "nnz": nnz, "R": R, "values": "U(0,1)"})
    res["steps"] = steps
    res["roofline"] = {"bound": "hbm", "achieved": round(B / (k_ms * 1e-3) / 1e9, 2), "peak": HBM_PEAK / 1e9,
                       "unit": "GB/s", "frac": round(B / (k_ms * 1e-3) / HBM_PEAK, 4), "traffic": None,
                       "kernel": KERNEL_SYMBOL[t["kernel"]], "kernel_avg_ms": round(k_ms, 5),
                       "algorithmic_bytes_per_launch": B, "carry_kernel_avg_ms": round(t["mttkrp_seg_carry_ms"], 5),
                       "fill_zero_avg_ms": round(t["fill_zero_ms"], 5)}
    res["dropin_sptMTTKRP"] = dropin_rate(lib, coo, factors, dims, R, order, nnz)
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline_coo(dims, inds, vals, factors, order)
    lib.sptolDevMatrixFree(C.byref(out))
    free_factors(lib, dm)
    lib.sptolDevCooFree(h)
    return res


def ref_library():
    """The reference library compiled from its sources into oracle/_ref, or
    None where it was not built."""
    try:
        return C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libparti_ref.so"))
    except OSError:
        return None


class quiet_stdout:
    """Send what C code prints to stdout to /dev/null (the reference's
    drivers print per call; this program's stdout is its one JSON line)."""

    def __enter__(self):
        sys.stdout.flush()
        self.libc = C.CDLL(None)
        self.libc.fflush(None)
        self.saved = os.dup(1)
        nul = os.open(os.devnull, os.O_WRONLY)
        os.dup2(nul, 1)
        os.close(nul)
        return self

    def __exit__(self, *exc):
        self.libc.fflush(None)
        os.dup2(self.saved, 1)
        os.close(self.saved)
        return False


def host_threads():
    ncpu = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return max(1, min(16, ncpu))


def cpu_baseline_coo(dims, inds, vals, factors, order):
    """The reference's OpenMP COO MTTKRP sptOmpMTTKRP (mttkrp_omp.c:125) from
    oracle/_ref, one warm-up and the mean of 5 calls (examples/mttkrp.c:240-283);
    the oracle's restatement when oracle/_ref is absent."""
    tk = host_threads()
    ref = ref_library()
    nnz = len(vals)
    ts = []
    if ref is not None:
        from sptol_amd.parti import mats_array
        coo = CooTensor(dims, inds, vals)
        mats = [Matrix(f) for f in factors] + [Matrix(nrows=int(dims[0]), ncols=factors[0].shape[1],
                                                      cap=int(max(dims)))]
        arr = mats_array(mats)
        o = np.ascontiguousarray(order, dtype=np.uint32)
        ref.sptOmpMTTKRP.argtypes = [C.POINTER(sptSparseTensor), C.c_void_p, C.POINTER(sptIndex), sptIndex, C.c_int]
        with quiet_stdout():
            for k in range(6):
                t0 = time.perf_counter()
                rc = ref.sptOmpMTTKRP(coo.ref(), C.cast(arr, C.c_void_p), ptr(o, sptIndex), 0, tk)
                if k:
                    ts.append(time.perf_counter() - t0)
                assert rc == 0
        what, kind = "the compiled reference's sptOmpMTTKRP (atomic, mttkrp_omp.c:125)", "reference"
    else:
        import oracle
        for k in range(6):
            t0 = time.perf_counter()
            oracle.mttkrp_coo(inds, vals, factors, order, 0, threads=tk)
            if k:
                ts.append(time.perf_counter() - t0)
        what, kind = "oracle restatement of sptOmpMTTKRP (atomic, mttkrp_omp.c:125)", "port"
    return {"value": nnz / float(np.mean(ts)) / 1e9, "unit": "Gnnz/s", "cores": tk, "kind": kind,
            "sample": f"{what}, tk={tk}, the full {nnz}-nonzero tensor, mean of 5 calls after one warm-up"}


def cpu_baseline_cpd(dims, inds, vals, init, R):
    """One ALS iteration of the reference's OpenMP CPD (OmpCpdAlsStep,
    cpd_omp.c:32, atomic MTTKRP, FP32 LAPACK solve) from oracle/_ref on the
    full tensor, from the same initial factors; None without oracle/_ref."""
    ref = ref_library()
    if ref is None:
        return None
    from sptol_amd.parti import mats_array, sptValue
    tk = host_threads()
    coo = CooTensor(dims, inds, vals)
    mats = [Matrix(a) for a in init] + [Matrix(nrows=int(max(dims)), ncols=R)]
    lam = np.zeros(R, np.float32)
    ref.OmpCpdAlsStep.argtypes = [C.POINTER(sptSparseTensor), sptIndex, sptIndex, C.c_double, C.c_int, C.c_int,
                                  C.c_void_p, C.c_void_p, C.POINTER(sptValue)]
    ref.OmpCpdAlsStep.restype = C.c_double
    iters = 1
    with quiet_stdout():
        t0 = time.perf_counter()
        ref.OmpCpdAlsStep(coo.ref(), R, iters, 0.0, tk, 0, C.cast(mats_array(mats), C.c_void_p), None,
                          ptr(lam, C.c_float))
        dt = time.perf_counter() - t0
    return {"value": dt / iters, "unit": "s/iter", "cores": tk, "kind": "reference",
            "sample": f"OmpCpdAlsStep (cpd_omp.c:32), tk={tk}, use_reduce=0, {iters} iteration on the full "
                      f"{len(vals)}-nonzero tensor from the same initial factors, {dt:.1f} s of CPU time"}


def run_cpd(args, lib):
    """4-D CPD-ALS, 20 iterations, R=32: 5000x5000x5000x500, 5e7 nonzeros,
    Zipf(1.05) indices, values from a planted rank-32 model plus 1% noise;
    FP32 factors and BF16-stored factors (FP32 accumulation), FP64 solve and
    fit, then FP32 factors over the device HiCOO form of the same tensor.
    Reports seconds per iteration and the fit of each variant."""
    dims = [5000, 5000, 5000, 500]
    nnz = int(args.nnz or 50_000_000)
    R, iters = 32, 20
    t0 = time.time()
    d, inds, vals, _ = gen.planted_tensor(dims, nnz, R, 1.05, seed=33, sort_fn=lib_sort(lib))
    nnz = len(vals)
    log(f"cpd: generated {nnz} nnz in {time.time() - t0:.1f}s")
    init = gen.uniform_factors(d, R, seed=34)
    coo = CooTensor(d, inds, vals)
    h = C.c_void_p()
    check(lib, lib.sptolDevCooCreate(C.byref(h), coo.ref()), "DevCooCreate")
    for m in range(4):
        check(lib, lib.sptolDevCooPrepareMode(h, m), "prepare")
    out = {}
    from sptol_amd.parti import sptolCpdStats
    for name, dt_ in (("f32", SPTOL_F32), ("bf16", SPTOL_BF16)):
        dm = dev_factors(lib, init)
        lam = np.zeros(R, np.float32)
        st = sptolCpdStats()
        st.factor_dtype = dt_
        st.exact_fit = 1
        sync(lib)
        t0 = time.perf_counter()
        check(lib, lib.sptolCpdAlsDevice(h, R, iters, -1.0, dm, ptr(lam, C.c_float), C.byref(st), None), "cpd")
        sync(lib)
        dt = time.perf_counter() - t0
        out[name] = {"seconds": dt, "s_per_iter": dt / st.iterations, "fit": st.fit, "iterations": st.iterations}
        free_factors(lib, dm)
        log(f"cpd {name}: {out[name]}")
    lib.sptolDevCooFree(h)
    # the same ALS over the device HiCOO tensor (sb=2^7, sk=2^14), FP32
    # factors: sptolCpdAlsHicooDevice, hicoo/cpd.c:34-135; every mode's view
    # is prepared before the timed loop, as for the COO tensor above
    hcoo = CooTensor(d, [a.copy() for a in inds], vals.copy())
    t0 = time.time()
    hh, _ = to_hicoo(lib, hcoo, 7, 14)
    del hcoo
    dh = C.c_void_p()
    check(lib, lib.sptolDevHicooCreate(C.byref(dh), C.byref(hh)), "DevHicooCreate")
    lib.sptFreeSparseTensorHiCOO(C.byref(hh))
    for m in range(4):
        check(lib, lib.sptolDevHicooPrepareMode(dh, m), "DevHicooPrepareMode")
    log(f"cpd: device HiCOO tensor and views in {time.time() - t0:.1f}s")
    dm = dev_factors(lib, init)
    lam = np.zeros(R, np.float32)
    st = sptolCpdStats()
    st.factor_dtype = SPTOL_F32
    st.exact_fit = 1
    sync(lib)
    t0 = time.perf_counter()
    check(lib, lib.sptolCpdAlsHicooDevice(dh, R, iters, -1.0, dm, ptr(lam, C.c_float), C.byref(st), None),
          "cpd hicoo")
    sync(lib)
    dt = time.perf_counter() - t0
    out["hicoo_f32"] = {"seconds": dt, "s_per_iter": dt / st.iterations, "fit": st.fit,
                        "iterations": st.iterations}
    free_factors(lib, dm)
    lib.sptolDevHicooFree(dh)
    log(f"cpd hicoo_f32: {out['hicoo_f32']}")
    res = base_result(args, "CPD-ALS seconds per iteration", round(out["f32"]["s_per_iter"], 5), "s/iter",
                      round(out["f32"]["s_per_iter"] * 1e3, 4),
                      {"workload": "cpd_als_4d", "dims": dims, "nnz": nnz, "R": R, "iterations": iters,
                       "zipf_s": 1.05, "planted_rank": R, "noise": 0.01}, higher=False)
    res["variants"] = out
    res["bf16_minus_f32_fit"] = out["bf16"]["fit"] - out["f32"]["fit"]
    res["hicoo_minus_coo_fit"] = out["hicoo_f32"]["fit"] - out["f32"]["fit"]
    res["cpu_baseline"] = None if args.no_cpu else cpu_baseline_cpd(d, inds, vals, init, R)
    if res["cpu_baseline"]:
        log(f"cpd CPU baseline: {res['cpu_baseline']}")
    return res


def run_batch(args, lib):
    """256 independent 3-D COO tensors (nnz log-uniform in [1e4, 1e7], dims
    log-uniform in [1e3, 1e6], R in {8, 16, 32, 64}, random mode): one
    autotuned persistent launch replayed from a hipGraph versus 256 plain
    launches in sequence."""
    rng = np.random.default_rng(404)
    ntens = 256
    scale = float(args.nnz or 1.0)   # --nnz scales the nonzero counts (1.0 = as specified)
    specs = []
    for t in range(ntens):
        nnz = int(np.exp(rng.uniform(np.log(1e4), np.log(1e7))) * scale)
        dims = [int(np.exp(rng.uniform(np.log(1e3), np.log(1e6)))) for _ in range(3)]
        specs.append((dims, max(nnz, 1), int(rng.choice([8, 16, 32, 64])), int(rng.integers(0, 3)), t))
    t0 = time.time()
    items = []
    total = 0
    for dims, nnz, R, mode, t in specs:
        d, inds, vals = gen.uniform_tensor(dims, nnz, seed=5000 + t, sort_fn=lib_sort(lib))
        coo = CooTensor(d, inds, vals)
        h = C.c_void_p()
        check(lib, lib.sptolDevCooCreate(C.byref(h), coo.ref()), "DevCooCreate")
        check(lib, lib.sptolDevCooPrepareMode(h, mode), "prepare")
        dm = dev_factors(lib, gen.uniform_factors(d, R, seed=6000 + t))
        out = sptolDevMatrix()
        check(lib, lib.sptolDevMatrixCreate(C.byref(out), int(d[mode]), R, SPTOL_F32), "create")
        order = np.array([mode, (mode + 1) % 3, (mode + 2) % 3], np.uint32)
        items.append((h, dm, out, order, mode, len(vals)))
        total += len(vals)
    log(f"batch: {ntens} tensors, {total} nnz prepared in {time.time() - t0:.1f}s")

    pargs = [(h, dm, ptr(order, sptIndex), mode, C.byref(out), None) for h, dm, out, order, mode, _ in items]

    def plain():
        for a in pargs:
            check(lib, lib.sptolDevMTTKRP(*a), "mttkrp")

    tp = timed_steps(lib, plain, "mttkrp_coo", args.steps, args.warmup)
    b = C.c_void_p()
    check(lib, lib.sptolBatchCreate(C.byref(b)), "BatchCreate")
    for h, dm, out, order, mode, _ in items:
        check(lib, lib.sptolBatchAdd(b, h, dm, ptr(order, sptIndex), mode, C.byref(out)), "BatchAdd")
    s = C.c_void_p()
    check(lib, lib.sptolHipStreamCreate(C.byref(s)), "stream")
    check(lib, lib.sptolBatchAutotune(b, 3, s), "autotune")
    sc, oc = C.c_double(), C.c_int()
    lib.sptolBatchGetConfig(b, C.byref(sc), C.byref(oc))
    check(lib, lib.sptolBatchCapture(b, s), "capture")

    def batched():
        check(lib, lib.sptolBatchRun(b, s), "BatchRun")

    def sync_s():
        check(lib, lib.sptolHipStreamSynchronize(s), "sync")

    for _ in range(args.warmup):
        batched()
    sync_s()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batched()
    sync_s()
    dtb = time.perf_counter() - t0
    gb = total * args.steps / dtb / 1e9
    gp = total * args.steps / tp["dt"] / 1e9
    res = base_result(args, "batched MTTKRP Gnnz/s", round(gb, 4), "Gnnz/s", round(dtb / args.steps * 1e3, 4),
                      {"workload": "batched_coo_mttkrp", "tensors": ntens, "total_nnz": total,
                       "ranks": [8, 16, 32, 64], "nnz_range": [1e4 * scale, 1e7 * scale],
                       "dims_range": [1e3, 1e6]})
    # one MTTKRP launch per tensor, plus a carry launch for each tensor whose
    # view keeps fixed-size ranges (ranges of whole rows need none)
    plain_launches = (tp["launches"] + tp["mttkrp_seg_carry_launches"]) / max(args.steps, 1)
    res["plain_sequential"] = {"value": round(gp, 4), "unit": "Gnnz/s", "ms_per_step": round(tp["dt"] / args.steps * 1e3, 4),
                               "launches_per_step": plain_launches}
    res["speedup_vs_plain"] = round(gb / gp, 3)
    res["autotuned"] = {"tile_scale": sc.value, "blocks_per_cu": oc.value}
    lib.sptolBatchFree(b)
    lib.sptolHipStreamDestroy(s)
    for h, dm, out, _, _, _ in items:
        lib.sptolDevMatrixFree(C.byref(out))
        free_factors(lib, dm)
        lib.sptolDevCooFree(h)
    return res


def run_serve(args, lib):
    """Served CPD-ALS stream: requests are 3-D or 4-D (50/50) tensors with nnz
    log-uniform in [1e5, 5e7] (drawn from a pool of 24 pre-generated tensors
    spanning that range), Zipf(1.1) indices, planted rank-16 values plus 1%
    noise, R in {16, 32}, stopping at a fit change below 1e-4 or 50
    iterations; 200 GB admission budget.  The same stream is then run through
    the plain one-request-at-a-time device path."""
    from sptol_amd.parti import SPTOL_REQ_DONE, sptolCpdStats, sptolRequestInfo, sptolServerConfig
    rng = np.random.default_rng(505)
    npool = 24
    scale = float(args.nnz or 1.0)
    pool = []
    t0 = time.time()
    for k in range(npool):
        nnz = int(np.exp(np.log(1e5) + (np.log(5e7) - np.log(1e5)) * (k + 0.5) / npool) * scale)
        three = k % 2 == 0
        dims = [20000, 15000, 10000] if three else [5000, 4000, 3000, 400]
        d, inds, vals, _ = gen.planted_tensor(dims, max(nnz, 1000), 16, 1.1, seed=700 + k, sort_fn=lib_sort(lib))
        pool.append(CooTensor(d, inds, vals))
    log(f"serve: pool of {npool} tensors generated in {time.time() - t0:.1f}s")
    nreq = args.requests
    reqs = [(int(rng.integers(0, npool)), int(rng.choice([16, 32])), 9000 + i) for i in range(nreq)]
    gaps = rng.exponential(1.0 / args.rate, size=nreq) if args.rate > 0 else np.zeros(nreq)
    cfg = sptolServerConfig()
    cfg.hbm_budget = 200 << 30
    cfg.max_active = 64
    cfg.batch_mttkrp = 1
    cfg.exact_fit = 0
    cfg.borrow_inputs = 1   # the pool tensors outlive every request
    srv = C.c_void_p()
    check(lib, lib.sptolServerCreate(C.byref(srv), C.byref(cfg)), "ServerCreate")
    ids = []
    t0 = time.perf_counter()
    tnext = t0
    for (p, R, seed), gap in zip(reqs, gaps):
        tnext += gap
        now = time.perf_counter()
        if tnext > now:
            time.sleep(tnext - now)
        rid = C.c_uint64()
        check(lib, lib.sptolServerSubmit(srv, pool[p].ref(), R, 50, 1e-4, seed, SPTOL_F32, C.byref(rid)), "submit")
        ids.append(rid.value)
    lat, iters, fits = [], [], []
    for rid in ids:
        info = sptolRequestInfo()
        check(lib, lib.sptolServerWait(srv, rid, -1.0, C.byref(info)), "wait")
        if info.status == SPTOL_REQ_DONE:
            lat.append(info.latency_s)
            iters.append(info.iterations)
            fits.append(info.fit)
        lib.sptolServerRelease(srv, rid)
    t_served = time.perf_counter() - t0
    lib.sptolServerDestroy(srv)
    done = len(lat)
    log(f"serve: {done}/{nreq} done in {t_served:.1f}s")
    # the plain path: one request at a time (upload, prepare, solve, download)
    t0 = time.perf_counter()
    plat = []
    for (p, R, seed) in ([] if args.no_plain else reqs):
        ts = time.perf_counter()
        coo = pool[p]
        h = C.c_void_p()
        check(lib, lib.sptolDevCooCreate(C.byref(h), coo.ref()), "DevCooCreate")
        init = gen.uniform_factors(coo.ndims, R, seed=seed)
        dm = dev_factors(lib, init)
        lam = np.zeros(R, np.float32)
        st = sptolCpdStats()
        check(lib, lib.sptolCpdAlsDevice(h, R, 50, 1e-4, dm, ptr(lam, C.c_float), C.byref(st), None), "cpd")
        for m in range(len(coo.ndims)):
            hm = Matrix(nrows=int(coo.ndims[m]), ncols=R)
            check(lib, lib.sptolDevMatrixDownload(C.byref(hm.m), C.byref(dm[m])), "download")
        free_factors(lib, dm)
        lib.sptolDevCooFree(h)
        plat.append(time.perf_counter() - ts)
    t_plain = time.perf_counter() - t0
    sps = done / t_served
    pps = len(plat) / t_plain
    res = base_result(args, "served CPD-ALS solves/s", round(sps, 4), "solves/s", round(t_served / max(done, 1) * 1e3, 3),
                      {"workload": "served_cpd_als", "requests": nreq, "pool": npool,
                       "nnz_range": [1e5 * scale, 5e7 * scale], "ranks": [16, 32], "max_iters": 50, "tol": 1e-4,
                       "arrival_rate": args.rate or "all at t=0", "hbm_budget_gb": 200, "slots": cfg.max_active})
    res["steps"] = nreq
    res["latency_s"] = {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))} if lat else None
    if plat:
        res["plain_one_at_a_time"] = {"value": round(pps, 4), "unit": "solves/s",
                                      "latency_s": {"p50": float(np.percentile(plat, 50)),
                                                    "p99": float(np.percentile(plat, 99))}}
        res["speedup_vs_plain"] = round(sps / pps, 3)
    res["mean_iterations"] = float(np.mean(iters)) if iters else None
    return res


def _libc():
    libc = C.CDLL(None)
    libc.fopen.restype = C.c_void_p
    libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
    libc.fclose.argtypes = [C.c_void_p]
    return libc


def _load_tns(L, path):
    """sptLoadSparseTensor (load.c:31) of library L on a .tns file with
    1-based indices; returns (seconds, ndims, inds, values) as numpy copies."""
    libc = _libc()
    L.sptLoadSparseTensor.argtypes = [C.POINTER(sptSparseTensor), sptIndex, C.c_void_p]
    L.sptFreeSparseTensor.argtypes = [C.POINTER(sptSparseTensor)]
    fp = libc.fopen(path, b"r")
    assert fp, path
    t = sptSparseTensor()
    t0 = time.perf_counter()
    rc = L.sptLoadSparseTensor(C.byref(t), 1, fp)
    dt = time.perf_counter() - t0
    libc.fclose(fp)
    assert rc == 0, rc
    n, nm = int(t.nnz), int(t.nmodes)
    d = np.ctypeslib.as_array(t.ndims, shape=(nm,)).copy()
    inds = [np.ctypeslib.as_array(t.inds[m].data, shape=(n,)).copy() for m in range(nm)]
    vals = np.ctypeslib.as_array(t.values.data, shape=(n,)).copy()
    L.sptFreeSparseTensor(C.byref(t))
    return dt, d, inds, vals


def run_ingest(args, lib):
    """Ingest of the 100M-nonzero Zipf(1.1) tensor 12M x 9M x 28M: one step =
    sptolDevCooLoadBinary, which streams the binary tensor file (u32 index
    arrays, f32 values) through pinned chunks into a device COO tensor and
    checks every index against its dimension on the device.  The file sits
    in the host's page cache (as after a first read).  Beside it: the host
    binary load followed by sptolDevCooCreate; duplicate merging
    (sptolSparseTensorMergeDuplicates) of the loaded tensor on the host
    threads; and the .tns text reader, this library's and the compiled
    reference's sptLoadSparseTensor (load.c:31), on a 2M-nonzero text file of
    the same tensor, whose loaded arrays must be identical.  cpu_baseline is
    the reference's text reader, the reference's only way in."""
    import tempfile
    nnz = int(args.nnz or HICOO_NNZ)
    t0 = time.time()
    dims, inds, vals = gen.zipf_tensor(DIMS, nnz, 1.1, seed=20260101, values="lognormal",
                                       sort_fn=lib_sort(lib))
    nnz = len(vals)
    log(f"ingest: generated {nnz} nnz in {time.time() - t0:.1f}s")
    coo = CooTensor(dims, inds, vals)
    tmp = tempfile.mkdtemp(prefix="sptol_ingest_")
    binpath = os.path.join(tmp, "x.bin").encode()
    tnspath = os.path.join(tmp, "x.tns").encode()
    res = None
    try:
        check(lib, lib.sptolSaveSparseTensorBinary(coo.ref(), binpath), "save binary")
        fbytes = os.path.getsize(binpath)
        ts = []
        for k in range(args.warmup + args.steps):
            h = C.c_void_p()
            sync(lib)
            t0 = time.perf_counter()
            check(lib, lib.sptolDevCooLoadBinary(C.byref(h), binpath, None), "DevCooLoadBinary")
            sync(lib)
            if k >= args.warmup:
                ts.append(time.perf_counter() - t0)
            assert lib.sptolDevCooNnz(h) == nnz
            lib.sptolDevCooFree(h)
        dt = float(np.mean(ts))
        log(f"ingest: file -> device {dt * 1e3:.1f} ms ({fbytes / dt / 1e9:.2f} GB/s)")

        lib.sptolLoadSparseTensorBinary.argtypes = [C.POINTER(sptSparseTensor), C.c_char_p]
        lib.sptFreeSparseTensor.argtypes = [C.POINTER(sptSparseTensor)]
        two = []
        for k in range(2):
            t = sptSparseTensor()
            t0 = time.perf_counter()
            check(lib, lib.sptolLoadSparseTensorBinary(C.byref(t), binpath), "LoadSparseTensorBinary")
            t1 = time.perf_counter()
            h = C.c_void_p()
            check(lib, lib.sptolDevCooCreate(C.byref(h), C.byref(t)), "DevCooCreate")
            sync(lib)
            two.append((t1 - t0, time.perf_counter() - t1))
            lib.sptolDevCooFree(h)
            lib.sptFreeSparseTensor(C.byref(t))
        host_load, upload = two[-1]
        log(f"ingest: host binary load {host_load:.3f} s, then sptolDevCooCreate {upload:.3f} s")

        tk = host_threads()
        mcoo = CooTensor(dims, [a.copy() for a in inds], vals.copy())
        t0 = time.perf_counter()
        check(lib, lib.sptolSparseTensorMergeDuplicates(mcoo.ref(), tk), "MergeDuplicates")
        t_merge = time.perf_counter() - t0
        merged = int(mcoo.st.nnz)
        del mcoo
        log(f"ingest: merge duplicates {t_merge:.2f} s ({nnz} -> {merged} nonzeros)")

        ns = min(nnz, 2_000_000)
        sample = CooTensor(dims, [a[:ns].copy() for a in inds], vals[:ns].copy())
        libc = _libc()
        lib.sptDumpSparseTensor.argtypes = [C.POINTER(sptSparseTensor), sptIndex, C.c_void_p]
        fp = libc.fopen(tnspath, b"w")
        check(lib, lib.sptDumpSparseTensor(sample.ref(), 1, fp), "DumpSparseTensor")
        libc.fclose(fp)
        t_ours, d1, i1, v1 = _load_tns(lib, tnspath)
        text = {"nnz": ns, "file_bytes": os.path.getsize(tnspath), "this_library_s": t_ours,
                "this_library_gnnz_s": ns / t_ours / 1e9}
        ref = None if args.no_cpu else ref_library()
        cpu = None
        if ref is not None:
            t_ref, d2, i2, v2 = _load_tns(ref, tnspath)
            same = (np.array_equal(d1, d2) and all(np.array_equal(a, b) for a, b in zip(i1, i2))
                    and np.array_equal(v1.view(np.uint32), v2.view(np.uint32)))
            assert same, "the .tns readers disagree"
            text.update({"reference_s": t_ref, "reference_gnnz_s": ns / t_ref / 1e9, "identical_arrays": same})
            cpu = {"value": ns / t_ref / 1e9, "unit": "Gnnz/s", "cores": 1, "kind": "reference",
                   "sample": f"the compiled reference's sptLoadSparseTensor (load.c:31, one thread) on a "
                             f".tns text file of the first {ns} nonzeros of the tensor, {t_ref:.1f} s"}
        log(f"ingest: .tns text {text}")
    finally:
        for p in (binpath, tnspath):
            if os.path.exists(p):
                os.remove(p)
        os.rmdir(tmp)
    res = base_result(args, "ingest Gnnz/s (binary file to device COO tensor)", round(nnz / dt / 1e9, 4),
                      "Gnnz/s", round(dt * 1e3, 4),
                      {"workload": "ingest_binary_to_device", "dims": DIMS, "nnz": nnz, "zipf_s": 1.1,
                       "values": "lognormal(0,1)", "file_bytes": fbytes, "file": "in the page cache"},
                      dtype="u32")
    res["file_gb_s"] = round(fbytes / dt / 1e9, 3)
    res["host_load_then_upload_s"] = {"host_binary_load": host_load, "sptolDevCooCreate": upload}
    res["merge_duplicates"] = {"seconds": t_merge, "threads": tk, "nnz_in": nnz, "nnz_out": merged}
    res["tns_text"] = text
    res["cpu_baseline"] = cpu
    return res


def run_copy(args, lib):
    """Device-to-device copy of 1 GiB with hipMemcpy (the runtime's blit
    kernel): the known-size reference for scaling rocprofv3 FETCH_SIZE /
    WRITE_SIZE readings to bytes.  Not a product path."""
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    n = 1 << 30
    src, dst = C.c_void_p(), C.c_void_p()
    check(lib, lib.sptolHipMalloc(C.byref(src), n), "malloc")
    check(lib, lib.sptolHipMalloc(C.byref(dst), n), "malloc")
    for _ in range(args.warmup):
        assert hip.hipMemcpy(dst, src, n, 3) == 0
    sync(lib)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        assert hip.hipMemcpy(dst, src, n, 3) == 0
    sync(lib)
    dt = time.perf_counter() - t0
    lib.sptolHipFree(src)
    lib.sptolHipFree(dst)
    gbs = 2 * n * args.steps / dt / 1e9
    return base_result(args, "device copy GB/s", round(gbs, 2), "GB/s", round(dt / args.steps * 1e3, 4),
                       {"workload": "copy_1GiB", "bytes_read_per_launch": n, "bytes_written_per_launch": n},
                       dtype="u8")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="hicoo", choices=["hicoo", "coo", "cpd", "batch", "serve", "ingest", "copy"],
                    help="hicoo: the headline HiCOO MTTKRP (default); coo: 3-D COO mode 0 (2000x1500x1000, 1e6 nnz); "
                         "cpd: 4-D CPD-ALS FP32 and BF16; batch: 256 tensors in one persistent launch; "
                         "serve: CPD request stream through the server; ingest: binary tensor file to the device; "
                         "copy: 1 GiB device copy (counter scale)")
    ap.add_argument("--nnz", type=float, default=None, help="override the workload's nonzero count")
    ap.add_argument("--requests", type=int, default=2000, help="serve: number of requests")
    ap.add_argument("--rate", type=float, default=0.0, help="serve: Poisson arrival rate (req/s); 0 = all at once")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-plain", action="store_true", help="serve: skip the plain one-at-a-time comparison")
    ap.add_argument("--no-coo", action="store_true", help="hicoo: skip the COO comparison")
    ap.add_argument("--no-serve", action="store_true",
                    help="hicoo: skip the served CPD leg (400 requests and the plain path)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path (sharding is a "
              "design note, not code); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    lib = load_library()
    n = C.c_int(0)
    if lib.sptolHipDeviceCount(C.byref(n)) != 0 or n.value < 1:
        print("no GPU visible to HIP", file=sys.stderr)
        sys.exit(1)
    if args.workload == "hicoo" and args.nnz is None:
        args.nnz = 1e8
    result = {"hicoo": run_hicoo, "coo": run_coo, "cpd": run_cpd, "batch": run_batch,
              "serve": run_serve, "ingest": run_ingest, "copy": run_copy}[args.workload](args, lib)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
