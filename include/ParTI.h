/*
 * sptol_amd -- drop-in C ABI for ParTI!'s MTTKRP / CPD-ALS path on MI355X.
 *
 * A caller written against ParTI! (hpcgarage/SpTOL) includes this header in
 * place of the reference's include/ParTI.h and links libsptol_amd.so.  The
 * type widths, struct layouts and function signatures below are the
 * reference's (widths: include/includes/types.h:29-36,49-52,61-67,90 forced to
 * 32/32/8 by include/ParTI.h:43-50; structs: include/includes/structs.h), so
 * existing object code and ctypes/FFI bindings keep working.  What runs behind
 * the MTTKRP and CPD names is HIP code for gfx950; everything else here is the
 * host-side container and conversion API those calls need.
 *
 * Only functions this library implements are declared.  Each declaration names
 * the reference interface it replaces.
 */
#ifndef SPTOL_AMD_PARTI_H
#define SPTOL_AMD_PARTI_H

#include <stddef.h>
#include <stdio.h>
#include <stdint.h>
#include <inttypes.h>
#include <math.h>
#include <stdbool.h>
#ifdef PARTI_USE_OPENMP
#include <omp.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------- */
/* Scalar types (include/includes/types.h, 32-bit index / 32-bit value /   */
/* 8-bit element index configuration)                                     */
/* ---------------------------------------------------------------------- */
typedef uint32_t sptIndex;
typedef uint32_t sptBlockIndex;
typedef uint8_t  sptElementIndex;
typedef uint16_t sptBlockMatrixIndex;
typedef sptBlockIndex sptBlockNnzIndex;
typedef uint64_t sptNnzIndex;
typedef float    sptValue;

#define PARTI_INDEX_TYPEWIDTH 32
#define PARTI_VALUE_TYPEWIDTH 32
#define PARTI_ELEMENT_INDEX_TYPEWIDTH 8
typedef unsigned __int128 sptMortonIndex;

/* printf/scanf conversions of the scalar types (types.h:34-92); callers'
 * option parsing reads them, e.g. sscanf(arg, "%"PARTI_SCN_INDEX, &mode). */
#define PARTI_INDEX_MAX UINT32_MAX
#define PARTI_PRI_INDEX PRIu32
#define PARTI_SCN_INDEX SCNu32
#define PARTI_PRI_BLOCK_INDEX PRIu32
#define PARTI_SCN_BLOCK_INDEX SCNu32
#define PARTI_PRI_BLOCKNNZ_INDEX PARTI_PRI_BLOCK_INDEX
#define PARTI_SCN_BLOCKNNZ_INDEX PARTI_SCN_BLOCK_INDEX
#define PARTI_PRI_ELEMENT_INDEX PRIu8
#define PARTI_SCN_ELEMENT_INDEX SCNu8
#define PARTI_PRI_BLOCKMATRIX_INDEX PRIu16
#define PARTI_SCN_BLOCKMATRIX_INDEX SCNu16
#define PARTI_PRI_NNZ_INDEX PRIu64
#define PARTI_SCN_NNZ_INDEX SCNu64
#define PARTI_PRI_VALUE "f"
#define PARTI_SCN_VALUE "f"

/* The reference's system-configuration constants (macros.h:23-36).  Callers
 * use them to pick OpenMP schedules and privatisation and print them as
 * tuning hints; the GPU path does not depend on them. */
#define NUM_CORES 32
#define L1_SIZE 32000
#define LLC_SIZE 40960000
#define PAR_DEGREE_REDUCE 20
#define PAR_MAX_DEGREE 100
#define PAR_MIN_DEGREE 4
#ifndef PARTI_DEFAULT_NLOCKS
#define PARTI_DEFAULT_NLOCKS 1024
#endif
#ifndef PARTI_DEFAULT_LOCK_PAD_SIZE
#define PARTI_DEFAULT_LOCK_PAD_SIZE 16
#endif

/* Error codes (include/includes/macros.h:43-52).  HIP runtime errors are
 * reported as SPTERR_CUDA_ERROR + hipError_t, in the slot the reference
 * reserves for device errors. */
typedef enum {
    SPTERR_NO_ERROR       = 0,
    SPTERR_UNKNOWN        = 1,
    SPTERR_SHAPE_MISMATCH = 2,
    SPTERR_VALUE_ERROR    = 3,
    SPTERR_ZERO_DIVISION  = 4,
    SPTERR_NO_MORE        = 99,
    SPTERR_OS_ERROR       = 0x10000,
    SPTERR_CUDA_ERROR     = 0x20000
} SptError;

/* ---------------------------------------------------------------------- */
/* Containers (include/includes/structs.h:27-156,196-214)                 */
/* ---------------------------------------------------------------------- */
typedef struct { sptNnzIndex len; sptNnzIndex cap; sptValue        *data; } sptValueVector;
typedef struct { sptNnzIndex len; sptNnzIndex cap; sptIndex        *data; } sptIndexVector;
typedef struct { sptNnzIndex len; sptNnzIndex cap; sptElementIndex *data; } sptElementIndexVector;
typedef struct { sptNnzIndex len; sptNnzIndex cap; sptBlockIndex   *data; } sptBlockIndexVector;
typedef struct { sptNnzIndex len; sptNnzIndex cap; sptNnzIndex     *data; } sptNnzIndexVector;

/* Row-major dense matrix; stride = ncols rounded up to a multiple of 8.
 * Columns [ncols, stride) are padding and are never read. */
typedef struct {
    sptIndex nrows;
    sptIndex ncols;
    sptIndex cap;
    sptIndex stride;
    sptValue *values;
} sptMatrix;

/* Dense matrix whose column count fits in 8 bits (HiCOO CPD). */
typedef struct {
    sptIndex        nrows;
    sptElementIndex ncols;
    sptIndex        cap;
    sptElementIndex stride;
    sptValue       *values;
} sptRankMatrix;

/* COO sparse tensor, structure-of-arrays indices. */
typedef struct {
    sptIndex        nmodes;
    sptIndex       *sortorder;
    sptIndex       *ndims;
    sptNnzIndex     nnz;
    sptIndexVector *inds;
    sptValueVector  values;
} sptSparseTensor;

/* HiCOO sparse tensor.  The "balanced scheduler" members are kept for layout
 * compatibility; like the reference, this library never fills them. */
typedef struct {
    sptIndex             nmodes;
    sptIndex            *sortorder;
    sptIndex            *ndims;
    sptNnzIndex          nnz;

    sptElementIndex      sb_bits;
    sptElementIndex      sk_bits;
    sptElementIndex      sc_bits;

    sptNnzIndexVector    kptr;
    sptIndexVector     **kschr;
    sptIndex            *nkiters;
    sptNnzIndexVector    cptr;

    sptIndexVector     **kschr_balanced;
    sptIndexVector     **kschr_balanced_pos;
    sptIndex            *nkpars;
    sptIndexVector      *kschr_rest;
    sptNnzIndexVector    knnzs;

    sptNnzIndexVector    bptr;
    sptBlockIndexVector *binds;
    sptElementIndexVector *einds;
    sptValueVector       values;
} sptSparseTensorHiCOO;

/* Kruskal tensor: CPD result (structs.h:196-203). */
typedef struct {
    sptIndex    nmodes;
    sptIndex    rank;
    sptIndex   *ndims;
    sptValue   *lambda;
    double      fit;
    sptMatrix **factors;
} sptKruskalTensor;

typedef struct {
    sptIndex         nmodes;
    sptElementIndex  rank;
    sptIndex        *ndims;
    sptValue        *lambda;
    double           fit;
    sptRankMatrix  **factors;
} sptRankKruskalTensor;

typedef struct sptTagTimer *sptTimer;

/* ---------------------------------------------------------------------- */
/* Errors (src/error/error.c:28-103)                                      */
/* ---------------------------------------------------------------------- */
int  sptGetLastError(const char **module, const char **file, unsigned *line, const char **err_str); /* error.c:89 */
void sptClearLastError(void);                                                                      /* error.c:110 */

/* Assertion that stays on under NDEBUG (helper_funcs.h:29-35): prints
 * "file:line: assertion error: "expr"" to stderr and aborts. */
void spt_Panic(const char *file, unsigned line, const char *expr);                             /* error.c:75 */
#define sptAssert(expr) ((expr) ? (void)0 : spt_Panic(__FILE__, __LINE__, #expr))
#if !defined(__cplusplus) && !defined(max)
/* helper_funcs.h:22-25 (C callers only: a function-like `max` would break C++) */
#define max(a, b) ({ __typeof__(a) sptol_a_ = (a); __typeof__(b) sptol_b_ = (b); sptol_a_ > sptol_b_ ? sptol_a_ : sptol_b_; })
#endif

/* "%0.2f" followed by B/KB/MB/GB/TB (powers of 1024); the caller frees the
 * string.  NULL when out of memory. */
char *sptBytesString(uint64_t const bytes);                                                    /* base.c:25 */

/* ---------------------------------------------------------------------- */
/* Timers (src/timer.c:29-96)                                             */
/* ---------------------------------------------------------------------- */
int    sptNewTimer(sptTimer *timer, int use_cuda);           /* timer.c:35 */
int    sptStartTimer(sptTimer timer);                         /* timer.c:44 */
int    sptStopTimer(sptTimer timer);                          /* timer.c:55 */
double sptElapsedTime(const sptTimer timer);                  /* timer.c:66 */
double sptPrintElapsedTime(const sptTimer timer, const char *name); /* timer.c:75 */
double sptPrintAverageElapsedTime(const sptTimer timer, const int niters, const char *name); /* timer.c:82 */
int    sptFreeTimer(sptTimer timer);                          /* timer.c:89 */

/* ---------------------------------------------------------------------- */
/* Vectors (src/vector/vector.c)                                          */
/* ---------------------------------------------------------------------- */
int  sptNewValueVector(sptValueVector *vec, sptNnzIndex len, sptNnzIndex cap);          /* vector.c:34 */
int  sptAppendValueVector(sptValueVector *vec, sptValue value);                         /* vector.c:93 */
void sptFreeValueVector(sptValueVector *vec);
int  sptNewIndexVector(sptIndexVector *vec, sptNnzIndex len, sptNnzIndex cap);
int  sptAppendIndexVector(sptIndexVector *vec, sptIndex value);
void sptFreeIndexVector(sptIndexVector *vec);
int  sptNewElementIndexVector(sptElementIndexVector *vec, sptNnzIndex len, sptNnzIndex cap);
int  sptAppendElementIndexVector(sptElementIndexVector *vec, sptElementIndex value);
void sptFreeElementIndexVector(sptElementIndexVector *vec);
int  sptNewBlockIndexVector(sptBlockIndexVector *vec, sptNnzIndex len, sptNnzIndex cap);
int  sptAppendBlockIndexVector(sptBlockIndexVector *vec, sptBlockIndex value);
void sptFreeBlockIndexVector(sptBlockIndexVector *vec);
int  sptNewNnzIndexVector(sptNnzIndexVector *vec, sptNnzIndex len, sptNnzIndex cap);
int  sptAppendNnzIndexVector(sptNnzIndexVector *vec, sptNnzIndex value);
void sptFreeNnzIndexVector(sptNnzIndexVector *vec);
sptIndex sptMaxIndexArray(sptIndex const *indices, sptNnzIndex len);                   /* vector/array.c:39 */
int  sptDumpIndexArray(sptIndex *array, sptNnzIndex const n, FILE *fp);                 /* vector/dump.c:143 */

/* ---------------------------------------------------------------------- */
/* Dense matrices (src/matrix/matrix.c, rankmatrix.c, dump.c)             */
/* ---------------------------------------------------------------------- */
int  sptNewMatrix(sptMatrix *mtx, sptIndex const nrows, sptIndex const ncols);          /* matrix.c:37 */
/* Seeded replacement for the reference's srand(time(NULL)) fill (matrix.c:68-75):
 * values are U[0,1) from the library's seeded generator (see sptolSetSeed). */
int  sptRandomizeMatrix(sptMatrix *mtx, sptIndex const nrows, sptIndex const ncols);    /* matrix.c:68 */
int  sptConstantMatrix(sptMatrix * const mtx, sptValue const val);                      /* matrix.c:107 */
int  sptCopyMatrix(sptMatrix *dest, const sptMatrix *src);                              /* matrix.c:147 */
void sptFreeMatrix(sptMatrix *mtx);                                                     /* matrix.c:231 */
int  sptDumpMatrix(sptMatrix *mtx, FILE *fp);                                           /* matrix/dump.c:34 */
int  sptNewRankMatrix(sptRankMatrix *mtx, sptIndex const nrows, sptElementIndex const ncols); /* rankmatrix.c:38 */
int  sptConstantRankMatrix(sptRankMatrix *mtx, sptValue const val);                         /* rankmatrix.c:87 */
void sptFreeRankMatrix(sptRankMatrix *mtx);                                              /* rankmatrix.c:127 */
int  sptDumpRankMatrix(sptRankMatrix *mtx, FILE *fp);                                     /* matrix/dump.c:60 */
int  sptRandomizeRankMatrix(sptRankMatrix *mtx, sptIndex const nrows, sptElementIndex const ncols); /* rankmatrix.c:70 */
int  sptMatrix2Norm(sptMatrix * const A, sptValue * const lambda);                      /* matrix.c:390 */
int  sptMatrixMaxNorm(sptMatrix * const A, sptValue * const lambda);                    /* matrix.c:473 */

/* ---------------------------------------------------------------------- */
/* COO sparse tensors (src/sptensor/sptensor.c, load.c, dump.c, sort.c)   */
/* ---------------------------------------------------------------------- */
int    sptNewSparseTensor(sptSparseTensor *tsr, sptIndex nmodes, const sptIndex ndims[]);   /* sptensor.c:30 */
int    sptCopySparseTensor(sptSparseTensor *dest, const sptSparseTensor *src, int const nt); /* sptensor.c:58 */
void   sptFreeSparseTensor(sptSparseTensor *tsr);                                            /* sptensor.c:83 */
double SparseTensorFrobeniusNormSquared(sptSparseTensor const * const spten);                /* sptensor.c:96 */
int    sptLoadSparseTensor(sptSparseTensor *tsr, sptIndex start_index, FILE *fp);            /* load.c:31 */
int    sptDumpSparseTensor(const sptSparseTensor *tsr, sptIndex start_index, FILE *fp);      /* sptensor/dump.c:29 */
void   sptSparseTensorSortIndex(sptSparseTensor *tsr, int force, int tk);                    /* sort.c:559 */
/* Lexicographic order over mode_order[0], mode_order[1], ... (most
 * significant first); sortorder becomes mode_order.  Without force, a tensor
 * whose sortorder already equals mode_order is left alone. */
void   sptSparseTensorSortIndexCustomOrder(sptSparseTensor *tsr, sptIndex const *mode_order,
           int force, int tk);                                                               /* sort.c:523 */
/* Lexicographic order over every mode but `mode`, then `mode` last. */
void   sptSparseTensorSortIndexAtMode(sptSparseTensor *tsr, sptIndex const mode, int force, int tk); /* sort.c:588 */
/* mode_order = [mode, the other modes by increasing dimension] (best) or
 * [the other modes by decreasing dimension, mode] (worst); equal dimensions
 * keep increasing mode number. */
void   sptGetBestModeOrder(sptIndex *mode_order, sptIndex const mode, sptIndex const *ndims,
           sptIndex const nmodes);                                                           /* sort.c:193 */
void   sptGetWorstModeOrder(sptIndex *mode_order, sptIndex const mode, sptIndex const *ndims,
           sptIndex const nmodes);                                                           /* sort.c:239 */
/* HiCOO order without building HiCOO arrays: superblocks (2^sk_bits) row
 * major, Z-Morton inside each. */
int    sptSparseTensorMixedOrder(sptSparseTensor *tsr, const sptElementIndex sb_bits,
           const sptElementIndex sk_bits, int const tk);                                     /* sort.c:283 */
/* The reference's deterministic shuffle: for z = 0..nnz-1, srand(z+1) and
 * swap nonzero z with nonzero (rand()/RAND_MAX * nnz) % nnz. */
void   sptGetRandomShuffleElements(sptSparseTensor *tsr);                                    /* sort.c:371 */
/* Dimensions, nnz, density, average slice lengths and COO storage. */
void   sptSparseTensorStatus(sptSparseTensor *tsr, FILE *fp);                                /* status.c:37 */
double sptSparseTensorDensity(sptSparseTensor const * const tsr);                           /* status.c:24 */
/* Ingest additions (no counterpart in the reference, which keeps duplicated
 * coordinates as separate nonzeros, load.c:31-86):
 *   sptolSparseTensorMergeDuplicates: sort lexicographically and merge each
 *     run of equal coordinates into one nonzero (FP64 sum, rounded once);
 *   sptolSave/LoadSparseTensorBinary: a binary file of the COO arrays
 *     ("SPTOLBT1" header, then u32 index arrays and f32 values), read or
 *     written with one sequential transfer per array; the loader rejects
 *     indices >= ndims. */
int    sptolSparseTensorMergeDuplicates(sptSparseTensor *tsr, int tk);
int    sptolSaveSparseTensorBinary(const sptSparseTensor *tsr, const char *path);
int    sptolLoadSparseTensorBinary(sptSparseTensor *tsr, const char *path);
void   sptSparseTensorSortIndexRowBlock(sptSparseTensor *tsr, int force,
           const sptNnzIndex begin, const sptNnzIndex end,
           const sptElementIndex sk_bits, int const tk);                                     /* sort.c:462 */
void   sptSparseTensorSortIndexMorton(sptSparseTensor *tsr, int force,
           const sptNnzIndex begin, const sptNnzIndex end,
           const sptElementIndex sb_bits, int tk);                                           /* sort.c:410 */

/* ---------------------------------------------------------------------- */
/* HiCOO (src/sptensor/hicoo/hicoo.c, convert.c, dump.c)                  */
/* ---------------------------------------------------------------------- */
int  sptNewSparseTensorHiCOO(sptSparseTensorHiCOO *hitsr, const sptIndex nmodes,
         const sptIndex ndims[], const sptNnzIndex nnz, const sptElementIndex sb_bits,
         const sptElementIndex sk_bits, const sptElementIndex sc_bits);                     /* hicoo.c:29 */
void sptFreeSparseTensorHiCOO(sptSparseTensorHiCOO *hitsr);                                  /* hicoo.c:136 */
/* Permutes *tsr in place into HiCOO order, as the reference does. */
int  sptSparseTensorToHiCOO(sptSparseTensorHiCOO *hitsr, sptNnzIndex *max_nnzb,
         sptSparseTensor *tsr, const sptElementIndex sb_bits,
         const sptElementIndex sk_bits, int const tk);                                       /* convert.c:300 */
int  sptSetKernelPointers(sptNnzIndexVector *kptr, sptSparseTensor *tsr,
         const sptElementIndex sk_bits);                                                     /* convert.c:237 */
int  sptDumpSparseTensorHiCOO(sptSparseTensorHiCOO * const hitsr, FILE *fp);                 /* hicoo/dump.c:28 */
double sptSparseTensorFrobeniusNormSquaredHiCOO(sptSparseTensorHiCOO const * const hitsr);   /* hicoo.c:179 */
/* Block/kernel/chunk counts, storage, schedule sizes and block-density
 * statistics of a HiCOO tensor (nb >= 1). */
void sptSparseTensorStatusHiCOO(sptSparseTensorHiCOO *hitsr, FILE *fp);                       /* hicoo/status.c:28 */

/* ---------------------------------------------------------------------- */
/* MTTKRP: the hot path.  All of these run on the GPU.                    */
/* mats[0..nmodes-1] are the factors, mats[nmodes] receives the result    */
/* (rows = mats[mode]->nrows, fully overwritten including padding).       */
/* ---------------------------------------------------------------------- */
int sptMTTKRP(sptSparseTensor const * const X, sptMatrix * mats[],
        sptIndex const mats_order[], sptIndex const mode);                                  /* mttkrp.c:39 */
int sptOmpMTTKRP(sptSparseTensor const * const X, sptMatrix * mats[],
        sptIndex const mats_order[], sptIndex const mode, const int tk);                    /* mttkrp_omp.c:125 */
int sptMTTKRPHiCOO(sptSparseTensorHiCOO const * const hitsr, sptMatrix * mats[],
        sptIndex const mats_order[], sptIndex const mode);                                  /* hicoo/mttkrp.c:53 */
int sptMTTKRPHiCOO_MatrixTiling(sptSparseTensorHiCOO const * const hitsr, sptRankMatrix * mats[],
        sptIndex const mats_order[], sptIndex const mode);                                  /* hicoo/mttkrp.c:158 */
int sptOmpMTTKRPHiCOO(sptSparseTensorHiCOO const * const hitsr, sptMatrix * mats[],
        sptIndex const mats_order[], sptIndex const mode, const int nt);                    /* hicoo/mttkrp_omp.c:110 */
/* OpenMP-named variants keep their signatures; thread counts, private copies
 * and the balanced flag are ignored because the work runs on the GPU. */
int sptOmpMTTKRP_Reduce(sptSparseTensor const * const X, sptMatrix * mats[], sptMatrix * copy_mats[],
        sptIndex const mats_order[], sptIndex const mode, const int tk);                    /* mttkrp_omp.c:401 */
int sptOmpMTTKRPHiCOO_MatrixTiling(sptSparseTensorHiCOO const * const hitsr, sptRankMatrix * mats[],
        sptIndex const mats_order[], sptIndex const mode, const int nt);                    /* hicoo/mttkrp_omp.c:134 */
int sptOmpMTTKRPHiCOO_MatrixTiling_Scheduled(sptSparseTensorHiCOO const * const hitsr, sptRankMatrix * mats[],
        sptIndex const mats_order[], sptIndex const mode, const int tk, int balanced);       /* hicoo/mttkrp_omp.c:158 */
int sptOmpMTTKRPHiCOO_MatrixTiling_Scheduled_Reduce(sptSparseTensorHiCOO const * const hitsr,
        sptRankMatrix * mats[], sptRankMatrix * copy_mats[], sptIndex const mats_order[],
        sptIndex const mode, const int tk, int balanced);                                   /* hicoo/mttkrp_omp.c:190 */

/* ---------------------------------------------------------------------- */
/* Kruskal tensors and CPD-ALS (src/ktensor/ktensor.c, src/sptensor/cpd.c) */
/* ---------------------------------------------------------------------- */
int    sptNewKruskalTensor(sptKruskalTensor *ktsr, sptIndex nmodes, const sptIndex ndims[], sptIndex rank); /* ktensor.c:33 */
void   sptFreeKruskalTensor(sptKruskalTensor *ktsr);                                          /* ktensor.c:79 */
int    sptNewRankKruskalTensor(sptRankKruskalTensor *ktsr, sptIndex nmodes, const sptIndex ndims[],
                               sptElementIndex rank);                                        /* rankktensor.c:33 */
void   sptFreeRankKruskalTensor(sptRankKruskalTensor *ktsr);                                  /* rankktensor.c:78 */
int    sptDumpKruskalTensor(sptKruskalTensor *ktsr, FILE *fp);                                /* ktensor/dump.c:24 */
int    sptDumpRankKruskalTensor(sptRankKruskalTensor *ktsr, FILE *fp);                        /* ktensor/dump.c:57 */
/* fit = 1 - ||X - K|| / ||X||, with ||K||^2 from the full symmetric Gram
 * Hadamard product (the reference reads an unwritten triangle, ktensor.c:151). */
double sptKruskalTensorFit(sptSparseTensor const * const spten, sptValue const * const lambda,
           sptMatrix ** mats, sptMatrix ** ata);                                             /* ktensor.c:102 */

/* ALS on the GPU: MTTKRP, FP64 Gram-Hadamard Cholesky solve, normalisation
 * and fit, with tensor and factors resident in device memory.  Factors are
 * initialised U[0,1) from the library's seeded generator. */
int sptCpdAls(sptSparseTensor const * const spten, sptIndex const rank, sptIndex const niters,
        double const tol, sptKruskalTensor * ktensor);                                      /* cpd.c:148 */
int sptOmpCpdAls(sptSparseTensor const * const spten, sptIndex const rank, sptIndex const niters,
        double const tol, const int tk, const int use_reduce, sptKruskalTensor * ktensor);  /* cpd_omp.c:166 */
/* rhs[i] := rhs[i] V^{-1} for every row of rhs, V = the Hadamard product of
 * the Grams aTa[(mode + k) % nmodes], k = 1..nmodes-1 (their upper
 * triangles), by an FP64 Cholesky factorisation on the GPU with an LU (partial
 * pivoting) fallback when V is not positive definite; rows are left unsolved
 * when V is exactly singular, as gesv leaves them.  aTa[nmodes] receives V.
 * rank <= 64; rhs has rank columns and the Grams' stride. */
int sptMatrixSolveNormals(sptIndex const mode, sptIndex const nmodes, sptMatrix ** aTa, sptMatrix * rhs); /* matrix_solver.c:32 */
/* The reference's exported inner loop (cpd.c:32-137), with caller-supplied
 * factors mats[0..nmodes] (mats[nmodes] is scratch with >= max_dim rows).
 * Returns the last fit. */
double CpdAlsStep(sptSparseTensor const * const spten, sptIndex const rank, sptIndex const niters,
        double const tol, sptMatrix ** mats, sptValue * const lambda);                       /* cpd.c:32 */
/* ALS over a HiCOO tensor on the GPU (rank <= 64), as sptCpdAls; the factors
 * are nmodes+1 sptRankMatrix structs, the last one emptied. */
int sptCpdAlsHiCOO(sptSparseTensorHiCOO const * const hitsr, sptIndex const rank, sptIndex const niters,
        double const tol, sptRankKruskalTensor * ktensor);                                  /* hicoo/cpd.c:149 */
int sptOmpCpdAlsHiCOO(sptSparseTensorHiCOO const * const hitsr, sptIndex const rank, sptIndex const niters,
        double const tol, const int tk, int balanced, sptRankKruskalTensor * ktensor);      /* hicoo/cpd_omp.c:190 */

#ifdef __cplusplus
}
#endif

#endif
