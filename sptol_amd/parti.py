"""ctypes mirror of the ParTI! C ABI (include/ParTI.h, include/sptol_hip.h).

The struct layouts are the reference's (include/includes/structs.h), so the
same helpers drive this library and, in tests only, a compiled reference
library.  Tensors built here borrow numpy buffers; the Python objects keep
those buffers alive and must never be handed to a *Free* function of either
library.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

sptIndex = C.c_uint32
sptNnzIndex = C.c_uint64
sptValue = C.c_float
sptElementIndex = C.c_uint8

SPTERR_SHAPE_MISMATCH = 2
SPTERR_VALUE_ERROR = 3
SPTERR_CUDA_ERROR = 0x20000


def _vec(elem):
    class V(C.Structure):
        _fields_ = [("len", sptNnzIndex), ("cap", sptNnzIndex), ("data", C.POINTER(elem))]
    return V


sptValueVector = _vec(sptValue)
sptIndexVector = _vec(sptIndex)
sptElementIndexVector = _vec(sptElementIndex)
sptBlockIndexVector = _vec(C.c_uint32)
sptNnzIndexVector = _vec(sptNnzIndex)


class sptMatrix(C.Structure):
    _fields_ = [("nrows", sptIndex), ("ncols", sptIndex), ("cap", sptIndex),
                ("stride", sptIndex), ("values", C.POINTER(sptValue))]


class sptRankMatrix(C.Structure):
    _fields_ = [("nrows", sptIndex), ("ncols", sptElementIndex), ("cap", sptIndex),
                ("stride", sptElementIndex), ("values", C.POINTER(sptValue))]


class sptSparseTensor(C.Structure):
    _fields_ = [("nmodes", sptIndex), ("sortorder", C.POINTER(sptIndex)),
                ("ndims", C.POINTER(sptIndex)), ("nnz", sptNnzIndex),
                ("inds", C.POINTER(sptIndexVector)), ("values", sptValueVector)]


class sptSparseTensorHiCOO(C.Structure):
    _fields_ = [("nmodes", sptIndex), ("sortorder", C.POINTER(sptIndex)),
                ("ndims", C.POINTER(sptIndex)), ("nnz", sptNnzIndex),
                ("sb_bits", sptElementIndex), ("sk_bits", sptElementIndex),
                ("sc_bits", sptElementIndex),
                ("kptr", sptNnzIndexVector), ("kschr", C.POINTER(C.POINTER(sptIndexVector))),
                ("nkiters", C.POINTER(sptIndex)), ("cptr", sptNnzIndexVector),
                ("kschr_balanced", C.POINTER(C.POINTER(sptIndexVector))),
                ("kschr_balanced_pos", C.POINTER(C.POINTER(sptIndexVector))),
                ("nkpars", C.POINTER(sptIndex)), ("kschr_rest", C.POINTER(sptIndexVector)),
                ("knnzs", sptNnzIndexVector), ("bptr", sptNnzIndexVector),
                ("binds", C.POINTER(sptBlockIndexVector)),
                ("einds", C.POINTER(sptElementIndexVector)), ("values", sptValueVector)]


class sptKruskalTensor(C.Structure):
    _fields_ = [("nmodes", sptIndex), ("rank", sptIndex), ("ndims", C.POINTER(sptIndex)),
                ("lambda_", C.POINTER(sptValue)), ("fit", C.c_double),
                ("factors", C.POINTER(C.POINTER(sptMatrix)))]


class sptRankKruskalTensor(C.Structure):
    _fields_ = [("nmodes", sptIndex), ("rank", sptElementIndex), ("ndims", C.POINTER(sptIndex)),
                ("lambda_", C.POINTER(sptValue)), ("fit", C.c_double),
                ("factors", C.POINTER(C.POINTER(sptRankMatrix)))]


class sptolDevMatrix(C.Structure):
    _fields_ = [("nrows", sptIndex), ("ncols", sptIndex), ("stride", sptIndex),
                ("dtype", C.c_int), ("values", C.c_void_p)]


class sptolCpdStats(C.Structure):
    _fields_ = [("factor_dtype", C.c_int), ("exact_fit", C.c_int), ("iterations", C.c_int),
                ("singular_solves", C.c_int), ("fit", C.c_double), ("seconds", C.c_double)]


class sptolServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("max_active", C.c_int), ("batch_mttkrp", C.c_int),
                ("exact_fit", C.c_int), ("blocks_per_cu", C.c_int), ("tile_scale", C.c_double),
                ("admit_threads", C.c_int), ("borrow_inputs", C.c_int)]


class sptolRequestInfo(C.Structure):
    _fields_ = [("status", C.c_int), ("error", C.c_int), ("iterations", C.c_int), ("rank", C.c_int),
                ("fit", C.c_double), ("latency_s", C.c_double), ("queue_s", C.c_double)]


SPTOL_F32 = 0
SPTOL_BF16 = 1
SPTOL_REQ_DONE = 3
SPTOL_REQ_FAILED = 4
SPTERR_NO_MORE = 99


def seeded_uniform_factors(seed, ndims, R):
    """The factor initialisation of a served request: one splitmix64 stream
    from `seed`, U[0,1) with 24 random bits, rows of mode 0, then mode 1, ..."""
    mask = (1 << 64) - 1
    st = seed & mask
    out = []
    for d in ndims:
        n = int(d) * R
        a = np.empty(n, dtype=np.float32)
        for i in range(n):
            st = (st + 0x9E3779B97F4A7C15) & mask
            z = st
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
            z ^= z >> 31
            a[i] = np.float32((z >> 40) * (1.0 / 16777216.0))
        out.append(a.reshape(int(d), R))
    return out


def ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def stride_for(ncols: int) -> int:
    return ((ncols - 1) // 8 + 1) * 8 if ncols else 8


class CooTensor:
    """An sptSparseTensor over numpy buffers (index arrays uint32, values float32)."""

    def __init__(self, ndims, inds, vals):
        self.ndims = np.ascontiguousarray(ndims, dtype=np.uint32)
        self.nmodes = len(self.ndims)
        self.inds = [np.ascontiguousarray(i, dtype=np.uint32) for i in inds]
        self.vals = np.ascontiguousarray(vals, dtype=np.float32)
        self.sortorder = np.arange(self.nmodes, dtype=np.uint32)
        n = len(self.vals)
        self._ivecs = (sptIndexVector * self.nmodes)()
        for m in range(self.nmodes):
            assert len(self.inds[m]) == n
            self._ivecs[m] = sptIndexVector(n, n, ptr(self.inds[m], sptIndex))
        self.st = sptSparseTensor(self.nmodes, ptr(self.sortorder, sptIndex), ptr(self.ndims, sptIndex),
                                  n, self._ivecs, sptValueVector(n, n, ptr(self.vals, sptValue)))

    @property
    def nnz(self):
        return int(self.st.nnz)

    def ref(self):
        return C.byref(self.st)


class Matrix:
    """An sptMatrix over a numpy (rows, stride) float32 buffer."""

    def __init__(self, a: np.ndarray | None = None, nrows=None, ncols=None, cap=None):
        if a is not None:
            a = np.asarray(a, dtype=np.float32)
            nrows, ncols = a.shape
        stride = stride_for(ncols)
        cap = max(cap or nrows, 1)
        self.buf = np.zeros((cap, stride), dtype=np.float32)
        if a is not None:
            self.buf[:nrows, :ncols] = a
        self.m = sptMatrix(nrows, ncols, cap, stride, ptr(self.buf, sptValue))

    def array(self):
        return self.buf[: self.m.nrows, : self.m.ncols]

    def ref(self):
        return C.pointer(self.m)


def mats_array(mats):
    arr = (C.POINTER(sptMatrix) * len(mats))()
    for i, m in enumerate(mats):
        arr[i] = m.ref()
    return arr


def hicoo_to_numpy(h: sptSparseTensorHiCOO) -> dict:
    """Copy every populated member of an sptSparseTensorHiCOO into numpy arrays."""
    N = h.nmodes
    ndims = np.ctypeslib.as_array(h.ndims, shape=(N,)).copy()
    nnz = int(h.nnz)

    def vec(v, dtype):
        if v.len == 0:
            return np.zeros(0, dtype=dtype)
        return np.ctypeslib.as_array(v.data, shape=(int(v.len),)).astype(dtype).copy()

    out = {
        "nmodes": N, "ndims": ndims, "nnz": nnz,
        "sb_bits": int(h.sb_bits), "sk_bits": int(h.sk_bits), "sc_bits": int(h.sc_bits),
        "kptr": vec(h.kptr, np.uint64), "cptr": vec(h.cptr, np.uint64), "bptr": vec(h.bptr, np.uint64),
        "nkiters": np.ctypeslib.as_array(h.nkiters, shape=(N,)).copy(),
        "binds": [vec(h.binds[m], np.uint32) for m in range(N)],
        "einds": [vec(h.einds[m], np.uint8) for m in range(N)],
        "values": vec(h.values, np.float32),
    }
    sk = 1 << int(h.sk_bits)
    kschr = []
    for m in range(N):
        kd = (int(ndims[m]) + sk - 1) // sk
        kschr.append([vec(h.kschr[m][i], np.uint32) for i in range(kd)])
    out["kschr"] = kschr
    return out


def _here():
    return os.path.dirname(os.path.abspath(__file__))


_LIB = None


def load_library(path: str | None = None) -> C.CDLL:
    """Load libsptol_amd.so built in-tree; raise if it has not been built."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or os.environ.get("SPTOL_LIBRARY") or os.path.join(_here(), "libsptol_amd.so")
    if not os.path.exists(p):
        raise RuntimeError(f"{p} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                           "or `make -C sptol_amd/csrc`")
    lib = C.CDLL(p)
    _declare(lib)
    if path is None:
        _LIB = lib
    return lib


def _declare(lib):
    P = C.POINTER
    sig = {
        "sptMTTKRP": (C.c_int, [P(sptSparseTensor), P(P(sptMatrix)), P(sptIndex), sptIndex]),
        "sptMTTKRPHiCOO": (C.c_int, [P(sptSparseTensorHiCOO), P(P(sptMatrix)), P(sptIndex), sptIndex]),
        "sptSparseTensorToHiCOO": (C.c_int, [P(sptSparseTensorHiCOO), P(sptNnzIndex), P(sptSparseTensor),
                                             sptElementIndex, sptElementIndex, C.c_int]),
        "sptFreeSparseTensorHiCOO": (None, [P(sptSparseTensorHiCOO)]),
        "sptSparseTensorSortIndex": (None, [P(sptSparseTensor), C.c_int, C.c_int]),
        "sptGetLastError": (C.c_int, [P(C.c_char_p), P(C.c_char_p), P(C.c_uint), P(C.c_char_p)]),
        "sptClearLastError": (None, []),
        "sptolSetSeed": (None, [C.c_uint64]),
        "sptolSetVerbose": (None, [C.c_int]),
        "sptolHipSetMemoryBudget": (C.c_int, [C.c_size_t]),
        "sptolHipMemoryInUse": (C.c_size_t, []),
        "sptolHipMemoryCached": (C.c_size_t, []),
        "sptolHipReleaseCache": (C.c_int, []),
        "sptolHipInjectMallocFailure": (C.c_int, [C.c_int]),
        "sptolHipStreamSynchronize": (C.c_int, [C.c_void_p]),
        "sptolHipDeviceCount": (C.c_int, [P(C.c_int)]),
        "sptolProfileEnable": (None, [C.c_int]),
        "sptolSetHicooViewKind": (C.c_int, [C.c_int]),
        "sptolSparseTensorToHiCOODevice": (C.c_int, [P(sptSparseTensorHiCOO), P(sptNnzIndex), P(sptSparseTensor),
                                                     C.c_uint8, C.c_uint8]),
        "sptolHipMalloc": (C.c_int, [P(C.c_void_p), C.c_size_t]),
        "sptolHipFree": (C.c_int, [C.c_void_p]),
        "sptolProfileReset": (None, []),
        "sptolProfileQuery": (C.c_int, [C.c_char_p, P(C.c_double), P(C.c_uint64)]),
        "sptolDevMatrixCreate": (C.c_int, [P(sptolDevMatrix), sptIndex, sptIndex, C.c_int]),
        "sptolDevMatrixUpload": (C.c_int, [P(sptolDevMatrix), P(sptMatrix)]),
        "sptolDevMatrixDownload": (C.c_int, [P(sptMatrix), P(sptolDevMatrix)]),
        "sptolDevMatrixFree": (None, [P(sptolDevMatrix)]),
        "sptolDevCooCreate": (C.c_int, [P(C.c_void_p), P(sptSparseTensor)]),
        "sptolDevCooPrepareMode": (C.c_int, [C.c_void_p, sptIndex]),
        "sptolDevCooFree": (None, [C.c_void_p]),
        "sptolDevMTTKRP": (C.c_int, [C.c_void_p, P(sptolDevMatrix), P(sptIndex), sptIndex,
                                     P(sptolDevMatrix), C.c_void_p]),
        "sptolDevMTTKRPNoAlloc": (C.c_int, [C.c_void_p, P(sptolDevMatrix), P(sptIndex), sptIndex,
                                            P(sptolDevMatrix), C.c_void_p]),
        "sptolDevHicooCreate": (C.c_int, [P(C.c_void_p), P(sptSparseTensorHiCOO)]),
        "sptolDevHicooFree": (None, [C.c_void_p]),
        "sptolDevHicooPrepareMode": (C.c_int, [C.c_void_p, sptIndex]),
        "sptolDevMTTKRPHiCOO": (C.c_int, [C.c_void_p, P(sptolDevMatrix), P(sptIndex), sptIndex,
                                          P(sptolDevMatrix), C.c_void_p]),
        "CpdAlsStep": (C.c_double, [P(sptSparseTensor), sptIndex, sptIndex, C.c_double,
                                    P(P(sptMatrix)), P(sptValue)]),
        "sptCpdAls": (C.c_int, [P(sptSparseTensor), sptIndex, sptIndex, C.c_double, P(sptKruskalTensor)]),
        "sptolCpdAlsDevice": (C.c_int, [C.c_void_p, sptIndex, sptIndex, C.c_double, P(sptolDevMatrix),
                                        P(sptValue), P(sptolCpdStats), C.c_void_p]),
        "sptolCpdAlsHicooDevice": (C.c_int, [C.c_void_p, sptIndex, sptIndex, C.c_double, P(sptolDevMatrix),
                                             P(sptValue), P(sptolCpdStats), C.c_void_p]),
        "sptCpdAlsHiCOO": (C.c_int, [P(sptSparseTensorHiCOO), sptIndex, sptIndex, C.c_double,
                                     P(sptRankKruskalTensor)]),
        "sptolSparseTensorMergeDuplicates": (C.c_int, [P(sptSparseTensor), C.c_int]),
        "sptolSaveSparseTensorBinary": (C.c_int, [P(sptSparseTensor), C.c_char_p]),
        "sptolLoadSparseTensorBinary": (C.c_int, [P(sptSparseTensor), C.c_char_p]),
        "sptFreeSparseTensor": (None, [P(sptSparseTensor)]),
        "sptolDevCooLoadBinary": (C.c_int, [P(C.c_void_p), C.c_char_p, C.c_void_p]),
        "sptolDevCooNnz": (sptNnzIndex, [C.c_void_p]),
        "sptolServerSubmitFile": (C.c_int, [C.c_void_p, C.c_char_p, sptIndex, sptIndex, C.c_double, C.c_uint64,
                                            C.c_int, P(C.c_uint64)]),
        "sptNewRankKruskalTensor": (C.c_int, [P(sptRankKruskalTensor), sptIndex, P(sptIndex), sptElementIndex]),
        "sptFreeRankKruskalTensor": (None, [P(sptRankKruskalTensor)]),
        "sptRandomizeRankMatrix": (C.c_int, [P(sptRankMatrix), sptIndex, sptElementIndex]),
        "sptNewRankMatrix": (C.c_int, [P(sptRankMatrix), sptIndex, sptElementIndex]),
        "sptFreeRankMatrix": (None, [P(sptRankMatrix)]),
        "sptolBatchCreate": (C.c_int, [P(C.c_void_p)]),
        "sptolBatchAdd": (C.c_int, [C.c_void_p, C.c_void_p, P(sptolDevMatrix), P(sptIndex), sptIndex,
                                    P(sptolDevMatrix)]),
        "sptolBatchSetConfig": (C.c_int, [C.c_void_p, C.c_double, C.c_int]),
        "sptolBatchGetConfig": (C.c_int, [C.c_void_p, P(C.c_double), P(C.c_int)]),
        "sptolBatchAutotune": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
        "sptolBatchCapture": (C.c_int, [C.c_void_p, C.c_void_p]),
        "sptolBatchRun": (C.c_int, [C.c_void_p, C.c_void_p]),
        "sptolBatchSize": (C.c_int, [C.c_void_p]),
        "sptolBatchFree": (None, [C.c_void_p]),
        "sptolHipStreamCreate": (C.c_int, [P(C.c_void_p)]),
        "sptolHipStreamDestroy": (C.c_int, [C.c_void_p]),
        "sptolHipSetDevice": (C.c_int, [C.c_int]),
        "sptolServerCreate": (C.c_int, [P(C.c_void_p), P(sptolServerConfig)]),
        "sptolServerSubmit": (C.c_int, [C.c_void_p, P(sptSparseTensor), sptIndex, sptIndex, C.c_double,
                                        C.c_uint64, C.c_int, P(C.c_uint64)]),
        "sptolServerWait": (C.c_int, [C.c_void_p, C.c_uint64, C.c_double, P(sptolRequestInfo)]),
        "sptolServerFetch": (C.c_int, [C.c_void_p, C.c_uint64, P(P(sptMatrix)), P(sptValue)]),
        "sptolServerRelease": (C.c_int, [C.c_void_p, C.c_uint64]),
        "sptolServerDrain": (C.c_int, [C.c_void_p]),
        "sptolServerTicks": (C.c_uint64, [C.c_void_p]),
        "sptolServerDestroy": (None, [C.c_void_p]),
        "sptGetBestModeOrder": (None, [P(sptIndex), sptIndex, P(sptIndex), sptIndex]),
        "sptGetWorstModeOrder": (None, [P(sptIndex), sptIndex, P(sptIndex), sptIndex]),
        "sptGetRandomShuffleElements": (None, [P(sptSparseTensor)]),
        "sptSparseTensorSortIndexCustomOrder": (None, [P(sptSparseTensor), P(sptIndex), C.c_int, C.c_int]),
        "sptSparseTensorSortIndexAtMode": (None, [P(sptSparseTensor), sptIndex, C.c_int, C.c_int]),
        "sptSparseTensorMixedOrder": (C.c_int, [P(sptSparseTensor), sptElementIndex, sptElementIndex, C.c_int]),
        "sptSparseTensorStatus": (None, [P(sptSparseTensor), C.c_void_p]),
        "sptSparseTensorStatusHiCOO": (None, [P(sptSparseTensorHiCOO), C.c_void_p]),
        "sptBytesString": (C.c_void_p, [C.c_uint64]),
        "sptDumpIndexArray": (C.c_int, [P(sptIndex), sptNnzIndex, C.c_void_p]),
        "sptDumpRankMatrix": (C.c_int, [P(sptRankMatrix), C.c_void_p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name, None)
        if f is None:
            continue
        f.restype = res
        f.argtypes = args


def last_error(lib) -> tuple[int, str]:
    mod, fil, reason = C.c_char_p(), C.c_char_p(), C.c_char_p()
    line = C.c_uint()
    code = lib.sptGetLastError(C.byref(mod), C.byref(fil), C.byref(line), C.byref(reason))
    msg = (reason.value or b"").decode(errors="replace")
    return code, msg


def check(lib, rc: int, what: str = "call"):
    if rc != 0:
        code, msg = last_error(lib)
        raise RuntimeError(f"{what} failed: rc=0x{rc:x} ({msg})")


def to_hicoo_device(lib, coo: CooTensor, sb_bits: int, sk_bits: int):
    """Convert on the GPU through lib.sptolSparseTensorToHiCOODevice (same
    arrays as to_hicoo; permutes coo in place)."""
    h = sptSparseTensorHiCOO()
    mx = sptNnzIndex(0)
    rc = lib.sptolSparseTensorToHiCOODevice(C.byref(h), C.byref(mx), coo.ref(), sb_bits, sk_bits)
    check(lib, rc, "sptolSparseTensorToHiCOODevice")
    return h, int(mx.value)


def to_hicoo(lib, coo: CooTensor, sb_bits: int, sk_bits: int, tk: int = 0):
    """Convert through lib.sptSparseTensorToHiCOO (permutes coo in place).
    Returns (hicoo struct, max_nnzb); free with lib.sptFreeSparseTensorHiCOO."""
    h = sptSparseTensorHiCOO()
    mx = sptNnzIndex(0)
    rc = lib.sptSparseTensorToHiCOO(C.byref(h), C.byref(mx), coo.ref(), sb_bits, sk_bits, tk)
    check(lib, rc, "sptSparseTensorToHiCOO")
    return h, int(mx.value)
