/*
 * Host-side base layer: last-error slot, timers, dynamic vectors, dense
 * matrices, Kruskal tensor containers and the seeded generator.
 *
 * Behaviour follows the reference's L1 utilities (src/error/error.c,
 * src/timer.c, src/vector/vector.c, src/matrix/matrix.c, src/ktensor/ktensor.c);
 * the code is this library's own.
 */
#define _GNU_SOURCE
#include "host_internal.h"
#include <stdio.h>
#include <math.h>
#include <time.h>

/* ------------------------------------------------------------------------ */
/* Last error (thread-local, src/error/error.c:28-34)                        */
/* ------------------------------------------------------------------------ */
/* The reason is copied into a per-thread buffer (truncated to 255 bytes):
 * nothing is allocated, so a worker thread that exits leaves nothing behind. */
static __thread struct {
    const char *module;
    int errcode;
    const char *file;
    unsigned line;
    char *reason;          /* t_err.buf, or NULL when no reason was given */
    char buf[256];
} t_err = { NULL, 0, NULL, 0, NULL, {0} };

void sptol_complain(const char *module, int errcode, const char *file, unsigned line,
                    const char *reason)
{
    t_err.module = module;
    t_err.errcode = errcode;
    t_err.file = file;
    t_err.line = line;
    if (reason) {
        snprintf(t_err.buf, sizeof(t_err.buf), "%s", reason);
        t_err.reason = t_err.buf;
    }
#ifndef NDEBUG
    if (t_err.reason && t_err.reason[0])
        fprintf(stderr, "[%s] error 0x%08x at %s:%u, %s\n", module, errcode, file, line, t_err.reason);
    else
        fprintf(stderr, "[%s] error 0x%08x at %s:%u\n", module, errcode, file, line);
#endif
}

int sptGetLastError(const char **module, const char **file, unsigned *line, const char **err_str)
{
    if (module) *module = t_err.module;
    if (file) *file = t_err.file;
    if (line) *line = t_err.line;
    if (err_str) *err_str = t_err.reason;
    return t_err.errcode;
}

void sptClearLastError(void)
{
    t_err.module = NULL;
    t_err.errcode = 0;
    t_err.file = NULL;
    t_err.line = 0;
    t_err.buf[0] = 0;
    t_err.reason = NULL;
}

/* ------------------------------------------------------------------------ */
/* Timers (CLOCK_MONOTONIC, src/timer.c:29-96).  use_cuda != 0 is rejected  */
/* as in a reference build without CUDA; device timing uses hipEvents.       */
/* ------------------------------------------------------------------------ */
struct sptTagTimer {
    int use_cuda;
    struct timespec t0, t1;
};

int sptNewTimer(sptTimer *timer, int use_cuda)
{
    *timer = (sptTimer)malloc(sizeof(**timer));
    SPTOL_CHECK_OS(!*timer, "Timer New");
    (*timer)->use_cuda = use_cuda;
    if (use_cuda) SPTOL_CHECK(3 + SPTERR_CUDA_ERROR, "Timer New", "device timers use hipEvents");
    return 0;
}

int sptStartTimer(sptTimer timer)
{
    SPTOL_CHECK_OS(clock_gettime(CLOCK_MONOTONIC, &timer->t0) != 0, "Timer Start");
    return 0;
}

int sptStopTimer(sptTimer timer)
{
    SPTOL_CHECK_OS(clock_gettime(CLOCK_MONOTONIC, &timer->t1) != 0, "Timer Stop");
    return 0;
}

double sptElapsedTime(const sptTimer timer)
{
    return (double)(timer->t1.tv_sec - timer->t0.tv_sec) + (timer->t1.tv_nsec - timer->t0.tv_nsec) * 1e-9;
}

double sptPrintElapsedTime(const sptTimer timer, const char *name)
{
    double t = sptElapsedTime(timer);
    fprintf(stdout, "[%s]: %.9lf s\n", name, t);
    return t;
}

int sptFreeTimer(sptTimer timer)
{
    free(timer);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Dynamic vectors (src/vector/vector.c): New zero-fills `len` elements of a */
/* buffer of max(len, cap, 2) elements; Append grows the buffer by 1.5x.     */
/* ------------------------------------------------------------------------ */
#define SPTOL_VECTOR_IMPL(VT, ET, NEWF, APPF, FREEF, MOD)                       \
    int NEWF(VT *vec, sptNnzIndex len, sptNnzIndex cap)                         \
    {                                                                           \
        if (cap < len) cap = len;                                               \
        if (cap < 2) cap = 2;                                                   \
        vec->len = len;                                                         \
        vec->cap = cap;                                                         \
        vec->data = (ET *)malloc(cap * sizeof(ET));                             \
        SPTOL_CHECK_OS(!vec->data, MOD " New");                                 \
        memset(vec->data, 0, cap * sizeof(ET));                                 \
        return 0;                                                               \
    }                                                                           \
    int APPF(VT *vec, ET value)                                                 \
    {                                                                           \
        if (vec->cap <= vec->len) {                                             \
            sptNnzIndex ncap = vec->cap + vec->cap / 2 + 2;                     \
            ET *nd = (ET *)realloc(vec->data, ncap * sizeof(ET));               \
            SPTOL_CHECK_OS(!nd, MOD " Append");                                 \
            vec->cap = ncap;                                                    \
            vec->data = nd;                                                     \
        }                                                                       \
        vec->data[vec->len++] = value;                                          \
        return 0;                                                               \
    }                                                                           \
    void FREEF(VT *vec)                                                         \
    {                                                                           \
        free(vec->data);                                                        \
        vec->data = NULL;                                                       \
        vec->len = 0;                                                           \
        vec->cap = 0;                                                           \
    }

SPTOL_VECTOR_IMPL(sptValueVector, sptValue, sptNewValueVector, sptAppendValueVector, sptFreeValueVector, "ValVec")
SPTOL_VECTOR_IMPL(sptIndexVector, sptIndex, sptNewIndexVector, sptAppendIndexVector, sptFreeIndexVector, "IdxVec")
SPTOL_VECTOR_IMPL(sptElementIndexVector, sptElementIndex, sptNewElementIndexVector, sptAppendElementIndexVector, sptFreeElementIndexVector, "EleIdxVec")
SPTOL_VECTOR_IMPL(sptBlockIndexVector, sptBlockIndex, sptNewBlockIndexVector, sptAppendBlockIndexVector, sptFreeBlockIndexVector, "BlkIdxVec")
SPTOL_VECTOR_IMPL(sptNnzIndexVector, sptNnzIndex, sptNewNnzIndexVector, sptAppendNnzIndexVector, sptFreeNnzIndexVector, "NnzIdxVec")

sptIndex sptMaxIndexArray(sptIndex const *indices, sptNnzIndex len)
{
    sptIndex mx = 0;
    for (sptNnzIndex i = 0; i < len; ++i) if (indices[i] > mx) mx = indices[i];
    return mx;
}

/* ------------------------------------------------------------------------ */
/* Seeded generator: splitmix64.  sptolSetSeed (declared in sptol_hip.h)     */
/* resets it; the default seed is fixed so runs are reproducible.            */
/* ------------------------------------------------------------------------ */
/* The state only ever advances by a constant, so concurrent callers each
 * take a distinct step with one atomic add; one thread alone sees the plain
 * splitmix64 sequence. */
static uint64_t g_rng_state = 0x5EED5EED5EED5EEDull;

void sptolSetSeed(uint64_t seed) { __atomic_store_n(&g_rng_state, seed, __ATOMIC_RELAXED); }

uint64_t sptol_rng_next(void)
{
    uint64_t z = __atomic_add_fetch(&g_rng_state, 0x9E3779B97F4A7C15ull, __ATOMIC_RELAXED);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

float sptol_rng_uniform01(void)
{
    /* 24 random mantissa bits -> [0, 1) */
    return (float)(sptol_rng_next() >> 40) * (1.0f / 16777216.0f);
}

/* ------------------------------------------------------------------------ */
/* Dense matrices (src/matrix/matrix.c:37-247, rankmatrix.c)                 */
/* ------------------------------------------------------------------------ */
int sptNewMatrix(sptMatrix *mtx, sptIndex const nrows, sptIndex const ncols)
{
    mtx->nrows = nrows;
    mtx->ncols = ncols;
    mtx->cap = nrows != 0 ? nrows : 1;
    mtx->stride = ncols ? ((ncols - 1) / 8 + 1) * 8 : 8;
    size_t bytes = (size_t)mtx->cap * mtx->stride * sizeof(sptValue);
    if (posix_memalign((void **)&mtx->values, 64, bytes) != 0) mtx->values = NULL;
    SPTOL_CHECK_OS(!mtx->values, "Mtx New");
    return 0;
}

int sptRandomizeMatrix(sptMatrix *mtx, sptIndex const nrows, sptIndex const ncols)
{
    for (sptIndex i = 0; i < nrows; ++i)
        for (sptIndex j = 0; j < ncols; ++j)
            mtx->values[(size_t)i * mtx->stride + j] = sptol_rng_uniform01();
    return 0;
}

int sptConstantMatrix(sptMatrix * const mtx, sptValue const val)
{
    size_t n = (size_t)mtx->nrows * mtx->stride;
    for (size_t i = 0; i < n; ++i) mtx->values[i] = val;
    return 0;
}

int sptCopyMatrix(sptMatrix *dest, const sptMatrix *src)
{
    int r = sptNewMatrix(dest, src->nrows, src->ncols);
    if (r) return r;
    memcpy(dest->values, src->values, (size_t)dest->nrows * dest->stride * sizeof(sptValue));
    return 0;
}

void sptFreeMatrix(sptMatrix *mtx)
{
    free(mtx->values);
    mtx->values = NULL;
    mtx->nrows = mtx->ncols = mtx->cap = mtx->stride = 0;
}

int sptDumpMatrix(sptMatrix *mtx, FILE *fp)
{
    int iores = fprintf(fp, "%" PRIu32 " x %" PRIu32 " matrix\n", mtx->nrows, mtx->ncols);
    SPTOL_CHECK_OS(iores < 0, "Mtx Dump");
    for (sptIndex i = 0; i < mtx->nrows; ++i) {
        for (sptIndex j = 0; j < mtx->ncols; ++j) {
            iores = fprintf(fp, "%.2f\t", mtx->values[(size_t)i * mtx->stride + j]);
            SPTOL_CHECK_OS(iores < 0, "Mtx Dump");
        }
        fprintf(fp, "\n");
    }
    fprintf(fp, "\n");
    return 0;
}

int sptNewRankMatrix(sptRankMatrix *mtx, sptIndex const nrows, sptElementIndex const ncols)
{
    mtx->nrows = nrows;
    mtx->ncols = ncols;
    mtx->cap = nrows != 0 ? nrows : 1;
    mtx->stride = ncols ? (sptElementIndex)(((ncols - 1) / 8 + 1) * 8) : 8;
    size_t bytes = (size_t)mtx->cap * mtx->stride * sizeof(sptValue);
    if (posix_memalign((void **)&mtx->values, 64, bytes) != 0) mtx->values = NULL;
    SPTOL_CHECK_OS(!mtx->values, "RankMtx New");
    return 0;
}

int sptConstantRankMatrix(sptRankMatrix *mtx, sptValue const val)
{
    size_t n = (size_t)mtx->nrows * mtx->stride;
    for (size_t i = 0; i < n; ++i) mtx->values[i] = val;
    return 0;
}

/* Seeded U[0,1) entries, the stream of sptRandomizeMatrix (the reference
 * reseeds with time(NULL) per call, rankmatrix.c:70-78). */
int sptRandomizeRankMatrix(sptRankMatrix *mtx, sptIndex const nrows, sptElementIndex const ncols)
{
    for (sptIndex i = 0; i < nrows; ++i)
        for (sptElementIndex j = 0; j < ncols; ++j)
            mtx->values[(size_t)i * mtx->stride + j] = sptol_rng_uniform01();
    return 0;
}

void sptFreeRankMatrix(sptRankMatrix *mtx)
{
    free(mtx->values);
    mtx->values = NULL;
    mtx->nrows = mtx->cap = 0;
    mtx->ncols = mtx->stride = 0;
}

/* Column 2-norms into lambda, then divide each column by its norm
 * (src/matrix/matrix.c:390-470, sequential form). */
int sptMatrix2Norm(sptMatrix * const A, sptValue * const lambda)
{
    sptIndex R = A->ncols, st = A->stride;
    for (sptIndex j = 0; j < R; ++j) lambda[j] = 0;
    for (sptIndex i = 0; i < A->nrows; ++i)
        for (sptIndex j = 0; j < R; ++j) lambda[j] += A->values[(size_t)i * st + j] * A->values[(size_t)i * st + j];
    for (sptIndex j = 0; j < R; ++j) lambda[j] = sqrtf(lambda[j]);
    for (sptIndex i = 0; i < A->nrows; ++i)
        for (sptIndex j = 0; j < R; ++j) A->values[(size_t)i * st + j] /= lambda[j];
    return 0;
}

/* Column maxima floored at 1 into lambda, then divide (matrix.c:473-554). */
int sptMatrixMaxNorm(sptMatrix * const A, sptValue * const lambda)
{
    sptIndex R = A->ncols, st = A->stride;
    for (sptIndex j = 0; j < R; ++j) lambda[j] = 0;
    for (sptIndex i = 0; i < A->nrows; ++i)
        for (sptIndex j = 0; j < R; ++j)
            if (A->values[(size_t)i * st + j] > lambda[j]) lambda[j] = A->values[(size_t)i * st + j];
    for (sptIndex j = 0; j < R; ++j) if (lambda[j] < 1) lambda[j] = 1;
    for (sptIndex i = 0; i < A->nrows; ++i)
        for (sptIndex j = 0; j < R; ++j) A->values[(size_t)i * st + j] /= lambda[j];
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Kruskal tensors (src/ktensor/ktensor.c:33-118, dump.c:24-55)              */
/* ------------------------------------------------------------------------ */
int sptNewKruskalTensor(sptKruskalTensor *ktsr, sptIndex nmodes, const sptIndex ndims[], sptIndex rank)
{
    ktsr->nmodes = nmodes;
    ktsr->rank = rank;
    ktsr->ndims = (sptIndex *)malloc(nmodes * sizeof(sptIndex));
    SPTOL_CHECK_OS(!ktsr->ndims, "KruskalTns New");
    memcpy(ktsr->ndims, ndims, nmodes * sizeof(sptIndex));
    ktsr->lambda = (sptValue *)malloc((rank ? rank : 1) * sizeof(sptValue));
    SPTOL_CHECK_OS(!ktsr->lambda, "KruskalTns New");
    ktsr->fit = 0.0;
    ktsr->factors = NULL;
    return 0;
}

/* Frees the factor matrices and their pointer array when present.  Factors
 * set by sptCpdAls are nmodes+1 heap-allocated sptMatrix structs (the last is
 * the MTTKRP scratch), all owned by the Kruskal tensor. */
void sptFreeKruskalTensor(sptKruskalTensor *ktsr)
{
    if (ktsr->factors) {
        for (sptIndex m = 0; m < ktsr->nmodes + 1; ++m) {
            if (ktsr->factors[m]) {
                sptFreeMatrix(ktsr->factors[m]);
                free(ktsr->factors[m]);
            }
        }
        free(ktsr->factors);
    }
    free(ktsr->ndims);
    free(ktsr->lambda);
    ktsr->ndims = NULL;
    ktsr->lambda = NULL;
    ktsr->factors = NULL;
    ktsr->rank = 0;
    ktsr->fit = 0.0;
    ktsr->nmodes = 0;
}

/* Rank Kruskal tensor (rankktensor.c:33-44, 78-88). */
int sptNewRankKruskalTensor(sptRankKruskalTensor *ktsr, sptIndex nmodes, const sptIndex ndims[], sptElementIndex rank)
{
    ktsr->nmodes = nmodes;
    ktsr->rank = rank;
    ktsr->ndims = (sptIndex *)malloc(nmodes * sizeof(sptIndex));
    SPTOL_CHECK_OS(!ktsr->ndims, "RankKruskalTns New");
    memcpy(ktsr->ndims, ndims, nmodes * sizeof(sptIndex));
    ktsr->lambda = (sptValue *)malloc((rank ? rank : 1) * sizeof(sptValue));
    SPTOL_CHECK_OS(!ktsr->lambda, "RankKruskalTns New");
    ktsr->fit = 0.0;
    ktsr->factors = NULL;
    return 0;
}

/* Frees the factors as sptFreeKruskalTensor does: sptCpdAlsHiCOO sets
 * nmodes+1 heap-allocated sptRankMatrix structs (the last is the emptied
 * MTTKRP scratch). */
void sptFreeRankKruskalTensor(sptRankKruskalTensor *ktsr)
{
    if (ktsr->factors) {
        for (sptIndex m = 0; m < ktsr->nmodes + 1; ++m) {
            if (ktsr->factors[m]) {
                sptFreeRankMatrix(ktsr->factors[m]);
                free(ktsr->factors[m]);
            }
        }
        free(ktsr->factors);
    }
    free(ktsr->ndims);
    free(ktsr->lambda);
    ktsr->ndims = NULL;
    ktsr->lambda = NULL;
    ktsr->factors = NULL;
    ktsr->rank = 0;
    ktsr->fit = 0.0;
    ktsr->nmodes = 0;
}

/* Prints `rank` lambdas and the nmodes factor matrices (the reference prints
 * nmodes lambdas and also the scratch matrix, src/ktensor/dump.c:42-52). */
int sptDumpKruskalTensor(sptKruskalTensor *ktsr, FILE *fp)
{
    int iores = fprintf(fp, "nmodes: %" PRIu32 ", rank: %" PRIu32 "\n", ktsr->nmodes, ktsr->rank);
    SPTOL_CHECK_OS(iores < 0, "KruskalTns Dump");
    for (sptIndex m = 0; m < ktsr->nmodes; ++m) fprintf(fp, m ? " %" PRIu32 : "%" PRIu32, ktsr->ndims[m]);
    fputs("\n", fp);
    fprintf(fp, "fit: %lf\n", ktsr->fit);
    fprintf(fp, "lambda:\n");
    for (sptIndex r = 0; r < ktsr->rank; ++r) fprintf(fp, "%f ", ktsr->lambda[r]);
    fprintf(fp, "\nFactor matrices:\n");
    if (ktsr->factors)
        for (sptIndex m = 0; m < ktsr->nmodes; ++m) {
            int r = sptDumpMatrix(ktsr->factors[m], fp);
            if (r) return r;
        }
    return 0;
}

/* fit = 1 - ||X - K|| / ||X|| (src/ktensor/ktensor.c:102-118).
 *   ||K||^2 = sum_{a,b} lambda_a lambda_b prod_m (A_m^T A_m)[a][b], over the
 *   full symmetric Grams (ata[m] hold full R x R Grams, row stride
 *   ata[m]->stride); <X,K> = sum_r lambda_r sum_i A_last[i][r] M[i][r] with
 *   M = mats[nmodes], the last mode's MTTKRP.  Accumulated in double. */
double sptKruskalTensorFit(sptSparseTensor const * const spten, sptValue const * const lambda,
                           sptMatrix ** mats, sptMatrix ** ata)
{
    sptIndex N = spten->nmodes, R = mats[0]->ncols;
    double normX = SparseTensorFrobeniusNormSquared(spten);
    double normK = 0.0;
    for (sptIndex a = 0; a < R; ++a)
        for (sptIndex b = 0; b < R; ++b) {
            double g = 1.0;
            for (sptIndex m = 0; m < N; ++m) g *= (double)ata[m]->values[(size_t)a * ata[m]->stride + b];
            normK += g * lambda[a] * lambda[b];
        }
    normK = fabs(normK);
    const sptMatrix *A = mats[N - 1], *M = mats[N];
    double inner = 0.0;
    for (sptIndex r = 0; r < R; ++r) {
        double acc = 0.0;
        for (sptIndex i = 0; i < A->nrows; ++i)
            acc += (double)A->values[(size_t)i * A->stride + r] * (double)M->values[(size_t)i * M->stride + r];
        inner += acc * lambda[r];
    }
    double residual = normX + normK - 2.0 * inner;
    if (residual > 0.0) residual = sqrt(residual);
    return 1.0 - residual / sqrt(normX);
}
