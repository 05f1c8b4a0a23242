"""Helpers for the GPU parity tests: run the product's device-resident
MTTKRP through the C ABI from numpy inputs."""
import ctypes as C

import numpy as np

from sptol_amd.parti import (CooTensor, Matrix, SPTOL_BF16, SPTOL_F32, check, ptr, sptIndex,
                             sptolDevMatrix, to_hicoo)


class DevFactors:
    def __init__(self, lib, factors, dtype=SPTOL_F32):
        self.lib = lib
        self.n = len(factors)
        self.arr = (sptolDevMatrix * self.n)()
        self.host = [Matrix(f) for f in factors]
        for m, hm in enumerate(self.host):
            check(lib, lib.sptolDevMatrixCreate(C.byref(self.arr[m]), hm.m.nrows, hm.m.ncols, dtype), "DevMatrixCreate")
            check(lib, lib.sptolDevMatrixUpload(C.byref(self.arr[m]), C.byref(hm.m)), "DevMatrixUpload")

    def free(self):
        for m in range(self.n):
            self.lib.sptolDevMatrixFree(C.byref(self.arr[m]))


class DevOut:
    def __init__(self, lib, rows, R):
        self.lib = lib
        self.d = sptolDevMatrix()
        check(lib, lib.sptolDevMatrixCreate(C.byref(self.d), rows, R, SPTOL_F32), "DevMatrixCreate")
        self.h = Matrix(nrows=rows, ncols=R)

    def fetch(self):
        check(self.lib, self.lib.sptolDevMatrixDownload(C.byref(self.h.m), C.byref(self.d)), "Download")
        return self.h.buf.copy()

    def free(self):
        self.lib.sptolDevMatrixFree(C.byref(self.d))


def dev_coo(lib, ndims, inds, vals):
    coo = CooTensor(ndims, inds, vals)
    h = C.c_void_p()
    check(lib, lib.sptolDevCooCreate(C.byref(h), coo.ref()), "DevCooCreate")
    return h


def dev_mttkrp_coo(lib, ndims, inds, vals, factors, order, mode, prepare=False, dtype=SPTOL_F32, h=None):
    own = h is None
    if own:
        h = dev_coo(lib, ndims, inds, vals)
    if prepare:
        check(lib, lib.sptolDevCooPrepareMode(h, mode), "PrepareMode")
    F = DevFactors(lib, factors, dtype)
    out = DevOut(lib, int(ndims[mode]), factors[0].shape[1])
    o = np.ascontiguousarray(order, dtype=np.uint32)
    check(lib, lib.sptolDevMTTKRP(h, F.arr, ptr(o, sptIndex), mode, C.byref(out.d), None), "DevMTTKRP")
    check(lib, lib.sptolHipStreamSynchronize(None), "sync")
    res = out.fetch()
    F.free()
    out.free()
    if own:
        lib.sptolDevCooFree(h)
    return res


VIEW_KIND = {True: 0, "auto": 0, "tiled": 1, "seg": 2}


def dev_mttkrp_hicoo(lib, hicoo_struct, ndims, factors, order, mode, dtype=SPTOL_F32, prepare=False):
    """prepare: False (block-order kernel), True/"auto", "tiled" or "seg"
    (per-mode view of that kind)."""
    h = C.c_void_p()
    check(lib, lib.sptolDevHicooCreate(C.byref(h), C.byref(hicoo_struct)), "DevHicooCreate")
    if prepare:
        check(lib, lib.sptolSetHicooViewKind(VIEW_KIND[prepare]), "SetHicooViewKind")
        try:
            check(lib, lib.sptolDevHicooPrepareMode(h, mode), "DevHicooPrepareMode")
        finally:
            lib.sptolSetHicooViewKind(0)
    F = DevFactors(lib, factors, dtype)
    out = DevOut(lib, int(ndims[mode]), factors[0].shape[1])
    o = np.ascontiguousarray(order, dtype=np.uint32)
    check(lib, lib.sptolDevMTTKRPHiCOO(h, F.arr, ptr(o, sptIndex), mode, C.byref(out.d), None), "DevMTTKRPHiCOO")
    check(lib, lib.sptolHipStreamSynchronize(None), "sync")
    res = out.fetch()
    F.free()
    out.free()
    lib.sptolDevHicooFree(h)
    return res


def bf16_round(a):
    """Round-to-nearest-even to bfloat16, returned as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32)
