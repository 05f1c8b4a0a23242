/*
 * sptol_amd -- device-resident C ABI (additive to ParTI.h).
 *
 * The reference reaches the GPU only through sptCudaMTTKRP*, which copies the
 * whole tensor and all factors host-to-device on every call
 * (src/sptensor/mttkrp_cuda_onekernel.cu:100-149), and through the memory
 * helpers of src/cudawrap.{h,cu}.  This header replaces those helpers with a
 * plain C "hipwrap" and adds handles that keep tensors and factors resident
 * in device memory, which is what the reference's own device-pointer entry
 * sptCudaMTTKRPDevice (src/sptensor/mttkrp_cuda.cu:378-389) points toward.
 *
 * All functions return 0 or an SptError code (HIP errors as
 * SPTERR_CUDA_ERROR + hipError_t) and record it in the thread-local
 * last-error slot read by sptGetLastError.  `stream` arguments are
 * hipStream_t values passed as void* (NULL = the library's default stream).
 */
#ifndef SPTOL_AMD_HIP_H
#define SPTOL_AMD_HIP_H

#include "ParTI.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------- */
/* hipwrap: device selection, memory, copies (replaces src/cudawrap.cu)    */
/* ---------------------------------------------------------------------- */
int    sptolHipSetDevice(int device);                       /* cudawrap.cu:26 sptCudaSetDevice */
int    sptolHipGetLastError(void);                          /* cudawrap.cu:30 sptCudaGetLastError */
int    sptolHipDeviceCount(int *count);
/* Device allocation counted against the library's memory budget. */
int    sptolHipMalloc(void **ptr, size_t bytes);
int    sptolHipFree(void *ptr);
/* Allocate + copy host -> device (cudawrap.cu:54-72 spt_CudaDuplicateMemoryGenerics). */
int    sptolHipDuplicateMemory(void **dst, const void *src, size_t bytes);
int    sptolHipMemcpyH2D(void *dst, const void *src, size_t bytes, void *stream);
int    sptolHipMemcpyD2H(void *dst, const void *src, size_t bytes, void *stream);
int    sptolHipStreamSynchronize(void *stream);
/* Budget accounting: allocations beyond the budget fail with
 * SPTERR_CUDA_ERROR + hipErrorOutOfMemory before reaching the runtime. */
int    sptolHipSetMemoryBudget(size_t bytes);               /* 0 = unlimited */
size_t sptolHipMemoryInUse(void);
/* Fault injection: the n-th sptolHipMalloc from now fails (n >= 1); 0 disables. */
int    sptolHipInjectMallocFailure(int n);
/* Seed of the library's generator (sptRandomizeMatrix, sptCpdAls init). */
void   sptolSetSeed(uint64_t seed);
/* 1 = print the reference's per-call timing and per-iteration CPD lines. */
void   sptolSetVerbose(int verbose);
/* Kernel timing: when enabled, every library kernel launch is bracketed by
 * hipEvents on its own stream.  Query returns the summed milliseconds and
 * launch count per kernel name ("mttkrp_coo", "mttkrp_hicoo", "fill_zero",
 * ...), synchronising on the recorded events. */
void   sptolProfileEnable(int on);
void   sptolProfileReset(void);
int    sptolProfileQuery(const char *name, double *total_ms, uint64_t *count);

/* ---------------------------------------------------------------------- */
/* Device-resident factor matrices                                        */
/* ---------------------------------------------------------------------- */
#define SPTOL_F32  0
#define SPTOL_BF16 1
/* Row-major nrows x stride, stride = ncols rounded up to 8 (as sptMatrix). */
typedef struct {
    sptIndex nrows;
    sptIndex ncols;
    sptIndex stride;
    int      dtype;      /* SPTOL_F32 or SPTOL_BF16 */
    void    *values;     /* device pointer */
} sptolDevMatrix;

int  sptolDevMatrixCreate(sptolDevMatrix *m, sptIndex nrows, sptIndex ncols, int dtype);
/* Upload converts to the matrix's dtype (round-to-nearest-even for BF16). */
int  sptolDevMatrixUpload(sptolDevMatrix *m, const sptMatrix *src);
/* Download converts to FP32 into dst (dst->nrows >= m->nrows, same stride). */
int  sptolDevMatrixDownload(sptMatrix *dst, const sptolDevMatrix *m);
void sptolDevMatrixFree(sptolDevMatrix *m);

/* ---------------------------------------------------------------------- */
/* Device-resident COO tensors and MTTKRP                                 */
/* ---------------------------------------------------------------------- */
typedef struct sptolDevCoo_s *sptolDevCoo;

int  sptolDevCooCreate(sptolDevCoo *h, const sptSparseTensor *X);
/* Build, on the device, a copy of the tensor stably sorted by the mode-th
 * index; later MTTKRPs on that mode reduce runs of equal output rows in
 * registers instead of issuing one atomic per nonzero. */
int  sptolDevCooPrepareMode(sptolDevCoo h, sptIndex mode);
void sptolDevCooFree(sptolDevCoo h);
sptNnzIndex sptolDevCooNnz(sptolDevCoo h);

/* mats[0..nmodes-1] are device factors; out (rows >= ndims[mode], same
 * stride) receives the result, fully overwritten including padding.
 * Semantics of mats_order/mode as sptMTTKRP. */
int  sptolDevMTTKRP(sptolDevCoo X, const sptolDevMatrix *mats, const sptIndex mats_order[],
                    sptIndex mode, sptolDevMatrix *out, void *stream);

/* ---------------------------------------------------------------------- */
/* Device-resident HiCOO tensors and MTTKRP                               */
/* ---------------------------------------------------------------------- */
typedef struct sptolDevHicoo_s *sptolDevHicoo;

int  sptolDevHicooCreate(sptolDevHicoo *h, const sptSparseTensorHiCOO *X);
void sptolDevHicooFree(sptolDevHicoo h);
int  sptolDevMTTKRPHiCOO(sptolDevHicoo X, const sptolDevMatrix *mats, const sptIndex mats_order[],
                         sptIndex mode, sptolDevMatrix *out, void *stream);

/* ---------------------------------------------------------------------- */
/* Batched MTTKRP: many device tensors in one persistent launch           */
/* ---------------------------------------------------------------------- */
typedef struct sptolBatch_s *sptolBatch;

int  sptolBatchCreate(sptolBatch *b);
/* Queue one MTTKRP (3-D or 4-D tensor, stride <= 256, all factors of a batch
 * share one dtype).  Uses the tensor's mode-sorted copy when prepared.  The
 * handles and matrices must outlive the batch. */
int  sptolBatchAdd(sptolBatch b, sptolDevCoo X, const sptolDevMatrix *mats, const sptIndex mats_order[],
                   sptIndex mode, sptolDevMatrix *out);
/* Tile length multiplier (base: 131072/stride nonzeros per tile) and
 * resident blocks per CU of the persistent kernel. */
int  sptolBatchSetConfig(sptolBatch b, double tile_scale, int blocks_per_cu);
int  sptolBatchGetConfig(sptolBatch b, double *tile_scale, int *blocks_per_cu);
/* Time candidate configurations on this batch and keep the fastest; results
 * are cached per batch signature for the life of the process. */
int  sptolBatchAutotune(sptolBatch b, int reps, void *stream);
/* Record the fill + persistent launch into a hipGraph on `stream` (a created
 * stream, not NULL); sptolBatchRun on that stream then replays the graph. */
int  sptolBatchCapture(sptolBatch b, void *stream);
/* Zero every output and run all queued MTTKRPs (asynchronous on `stream`). */
int  sptolBatchRun(sptolBatch b, void *stream);
int  sptolBatchSize(sptolBatch b);
void sptolBatchFree(sptolBatch b);

/* Streams at the default priority (hipStreamCreate). */
int  sptolHipStreamCreate(void **stream);
int  sptolHipStreamDestroy(void *stream);

/* ---------------------------------------------------------------------- */
/* Device-resident CPD-ALS                                                */
/* ---------------------------------------------------------------------- */
typedef struct {
    int      factor_dtype;   /* SPTOL_F32 or SPTOL_BF16 storage for MTTKRP gathers */
    int      iterations;     /* out: iterations executed                            */
    double   fit;            /* out: last fit (FP64, correct ||K||^2)               */
    double   seconds;        /* out: wall time of the ALS loop                      */
} sptolCpdStats;

/* ALS on a device tensor.  mats[0..nmodes-1] (FP32 device matrices) hold the
 * initial factors and receive the final ones; lambda (host, rank floats)
 * receives the weights.  Follows src/sptensor/cpd.c:32-137 with the solve,
 * norms and fit in FP64.  tol <= 0 runs all niters. */
int  sptolCpdAlsDevice(sptolDevCoo X, sptIndex rank, sptIndex niters, double tol,
                       sptolDevMatrix *mats, sptValue *lambda, sptolCpdStats *stats, void *stream);

#ifdef __cplusplus
}
#endif
#endif
