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
/* Device blocks held for reuse by the library's block cache (not in use), and
 * a call that returns them all to HIP (it waits for the device to drain).
 * SPTOL_CACHE_BYTES caps the cache (default 64 GiB; 0 disables it). */
size_t sptolHipMemoryCached(void);
int    sptolHipReleaseCache(void);
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
/* A device COO tensor read straight from a binary tensor file
 * (sptolSaveSparseTensorBinary), streamed through pinned chunks so that the
 * file read overlaps the copies; indices >= ndims are rejected. */
int  sptolDevCooLoadBinary(sptolDevCoo *h, const char *path, void *stream);
void sptolDevCooFree(sptolDevCoo h);
sptNnzIndex sptolDevCooNnz(sptolDevCoo h);

/* mats[0..nmodes-1] are device factors; out (rows >= ndims[mode], same
 * stride) receives the result, fully overwritten including padding.
 * Semantics of mats_order/mode as sptMTTKRP. */
int  sptolDevMTTKRP(sptolDevCoo X, const sptolDevMatrix *mats, const sptIndex mats_order[],
                    sptIndex mode, sptolDevMatrix *out, void *stream);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): the same
 * result as sptolDevMTTKRP, bit for bit, with no device allocation.  A mode
 * prepared by sptolDevCooPrepareMode (3- and 4-mode tensors) runs the
 * segmented kernels with their carry space and work counter in a fixed
 * library array of 16.3 MiB per GPU, held until the process ends; a tensor
 * that needs more than 65536 ranges (about 134M nonzeros at rank 32) is
 * rejected with SPTERR_VALUE_ERROR.  An unprepared mode runs the unsorted
 * kernel, which needs no scratch.  Calls that overlap on one GPU share the
 * array and race: order them on one stream. */
int  sptolDevMTTKRPNoAlloc(sptolDevCoo X, const sptolDevMatrix *mats, const sptIndex mats_order[],
                           sptIndex mode, sptolDevMatrix *out, void *stream);

/* ---------------------------------------------------------------------- */
/* Device-resident HiCOO tensors and MTTKRP                               */
/* ---------------------------------------------------------------------- */
typedef struct sptolDevHicoo_s *sptolDevHicoo;

int  sptolDevHicooCreate(sptolDevHicoo *h, const sptSparseTensorHiCOO *X);
/* Build, on the device, a per-mode view that later MTTKRPs on that mode use
 * (built once; the HiCOO arrays themselves are not changed).  Two kinds:
 *  - tiled: blocks stably ordered by their mode-th block coordinate
 *    (entries move with their blocks) plus the work list of the LDS-tiled
 *    kernel, which accumulates each 2^sb_bits x R output tile in LDS.
 *    Needs sb_bits <= 8, stride <= 64.  Built only on request (kind 1) or
 *    for tensors of an order the segmented kernel does not take.
 *  - segmented: coordinates rebuilt from block coordinates and element
 *    offsets, sorted by the mode, packed 16 bytes per nonzero, for the
 *    segmented kernel (one plain store per output row, no atomics,
 *    deterministic).  For 3- and 4-mode tensors with stride <= 64; the
 *    default for them at any block density (7-14x faster than tiled from
 *    15 to 651 nonzeros per block).
 * Shapes that neither kind handles keep the block-order kernel. */
int  sptolDevHicooPrepareMode(sptolDevHicoo h, sptIndex mode);
/* View kind for later sptolDevHicooPrepareMode calls (process-wide):
 * 0 = automatic (default: segmented for 3 and 4 modes, tiled otherwise),
 * 1 = tiled, 2 = segmented. */
#define SPTOL_HICOO_VIEW_AUTO 0
#define SPTOL_HICOO_VIEW_TILED 1
#define SPTOL_HICOO_VIEW_SEG 2
int  sptolSetHicooViewKind(int kind);
void sptolDevHicooFree(sptolDevHicoo h);
int  sptolDevMTTKRPHiCOO(sptolDevHicoo X, const sptolDevMatrix *mats, const sptIndex mats_order[],
                         sptIndex mode, sptolDevMatrix *out, void *stream);

/* COO -> HiCOO conversion on the GPU.  Same contract and the same arrays,
 * bit for bit, as sptSparseTensorToHiCOO (convert.c:300, ParTI.h): fills
 * `hitsr` (host arrays), permutes `tsr` in place into HiCOO order, returns
 * the largest block's nonzero count in *max_nnzb (may be NULL).  Needs
 * nnz < 2^31; the work runs on the default stream and is synchronous. */
int  sptolSparseTensorToHiCOODevice(sptSparseTensorHiCOO *hitsr, sptNnzIndex *max_nnzb, sptSparseTensor *tsr,
                                    sptElementIndex sb_bits, sptElementIndex sk_bits);

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
    int      factor_dtype;   /* in: SPTOL_F32 or SPTOL_BF16 storage for MTTKRP gathers       */
    int      exact_fit;      /* in: 1 = the returned fit takes <X,K> from an FP64 pass over the
                                nonzeros (COO or HiCOO, any order), once, after the last
                                iteration; the per-iteration fit that drives tol always comes
                                from the last mode's FP32 MTTKRP, as the reference
                                (ktensor.c:173-257), which loses digits to cancellation when
                                the fit nears 1.  sptCpdAls, CpdAlsStep and sptCpdAlsHiCOO
                                always return the exact fit                                 */
    int      iterations;     /* out: iterations executed                                    */
    int      singular_solves;/* out: mode updates whose normal equations were exactly singular
                                (a zero pivot after partial pivoting): their factors keep the
                                MTTKRP rows unsolved, as LAPACK gesv leaves its right-hand
                                sides when info > 0 (matrix_solver.c:72-77)                 */
    double   fit;            /* out: last fit (FP64, correct ||K||^2)                       */
    double   seconds;        /* out: wall time of the ALS loop                              */
} sptolCpdStats;

/* ALS on a device tensor.  mats[0..nmodes-1] (FP32 device matrices) hold the
 * initial factors and receive the final ones; lambda (host, rank floats)
 * receives the weights.  Follows src/sptensor/cpd.c:32-137 with the solve,
 * norms and fit in FP64.  tol <= 0 runs all niters. */
int  sptolCpdAlsDevice(sptolDevCoo X, sptIndex rank, sptIndex niters, double tol,
                       sptolDevMatrix *mats, sptValue *lambda, sptolCpdStats *stats, void *stream);
/* The same loop over a device HiCOO tensor (src/sptensor/hicoo/cpd.c:34-135):
 * every mode's view is prepared first, the MTTKRP is the HiCOO one. */
int  sptolCpdAlsHicooDevice(sptolDevHicoo X, sptIndex rank, sptIndex niters, double tol,
                            sptolDevMatrix *mats, sptValue *lambda, sptolCpdStats *stats, void *stream);

/* ---------------------------------------------------------------------- */
/* Continuous-batching CPD-ALS server                                     */
/* ---------------------------------------------------------------------- */
typedef struct sptolServer_s *sptolServer;

typedef struct {
    size_t hbm_budget;      /* bytes of device memory admitted requests may reserve         */
    int    max_active;      /* requests on the device at once (slots)                       */
    int    batch_mttkrp;    /* 1 = one persistent launch per step for every active request  */
    int    exact_fit;       /* 1 = FP64 pass for <X,K> (see sptolCpdStats.exact_fit)         */
    int    blocks_per_cu;   /* persistent-kernel residency (0 = default 8)                  */
    double tile_scale;      /* batched tile length multiplier (0 = default 1.0)             */
    int    admit_threads;   /* parallel admission workers, each with its own stream: 1..4 (0 = 4) */
    int    borrow_inputs;   /* 1 = Submit keeps pointers to the caller's index and value
                               arrays, which must stay valid and unchanged until the
                               request has been waited for; 0 = Submit copies them     */
} sptolServerConfig;

#define SPTOL_REQ_QUEUED    0
#define SPTOL_REQ_PREPARING 1
#define SPTOL_REQ_RUNNING   2
#define SPTOL_REQ_DONE      3
#define SPTOL_REQ_FAILED    4

typedef struct {
    int    status;          /* SPTOL_REQ_*                                   */
    int    error;           /* SptError of a failed request                  */
    int    iterations;
    int    rank;
    double fit;
    double latency_s;       /* submit -> result available                    */
    double queue_s;         /* submit -> admitted to the device              */
} sptolRequestInfo;

/* Start the admission and compute threads on the current device. */
int  sptolServerCreate(sptolServer *s, const sptolServerConfig *cfg);
/* Queue a CPD-ALS request (copies the tensor, so the caller may free it on
 * return, unless the server borrows its inputs).  Factors start U[0,1) from `seed`; the solve stops after niters
 * iterations or when the fit changes by less than tol (tol <= 0: never). */
int  sptolServerSubmit(sptolServer s, const sptSparseTensor *X, sptIndex rank, sptIndex niters, double tol,
                       uint64_t seed, int factor_dtype, uint64_t *request_id);
/* Block until the request finished (timeout_s < 0: no limit).  Returns
 * SPTERR_NO_MORE on timeout. */
/* The same for a tensor in a binary tensor file (sptolSaveSparseTensorBinary):
 * the header is read now, the arrays are streamed to the device at
 * admission, so the file must stay unchanged until then. */
int  sptolServerSubmitFile(sptolServer s, const char *path, sptIndex rank, sptIndex niters, double tol,
                           uint64_t seed, int factor_dtype, uint64_t *request_id);
int  sptolServerWait(sptolServer s, uint64_t request_id, double timeout_s, sptolRequestInfo *info);
/* Copy a finished request's factors (mats[m]: ndims[m] x rank, stride as
 * sptMatrix) and lambda (rank floats); either may be NULL. */
int  sptolServerFetch(sptolServer s, uint64_t request_id, sptMatrix **mats, sptValue *lambda);
int  sptolServerRelease(sptolServer s, uint64_t request_id);
int  sptolServerDrain(sptolServer s);
uint64_t sptolServerTicks(sptolServer s);
void sptolServerDestroy(sptolServer s);

#ifdef __cplusplus
}
#endif
#endif
