/*
 * Driver-support API of the reference: the assertion hook, byte-size
 * strings, averaged timer prints, dumps of index arrays and rank matrices,
 * mode orders for custom sorts, and the COO / HiCOO status reports.
 *
 * Callers written against ParTI! (examples/mttkrp.c, cpd.c, mttkrp_hicoo.c)
 * print these reports and parse none of them back, but the text is kept
 * line for line, so a run of a reference driver against this library reads
 * like one against the reference (src/sptensor/status.c,
 * src/sptensor/hicoo/status.c).  The code is this library's own.
 */
#define _GNU_SOURCE
#include "host_internal.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

/* ------------------------------------------------------------------------ */
/* Assertions and small helpers (src/error/error.c:75, src/base.c:25,        */
/* src/timer.c:82, src/vector/dump.c:143)                                    */
/* ------------------------------------------------------------------------ */
void spt_Panic(const char *file, unsigned line, const char *expr)
{
    fprintf(stderr, "%s:%u: assertion error: \"%s\"\n", file, line, expr);
    abort();
}

char *sptBytesString(uint64_t const bytes)
{
    static const char *const unit[] = {"B", "KB", "MB", "GB", "TB"};
    double v = (double)bytes;
    int u = 0;
    while (v > 1024 && u < 4) {
        v /= 1024.;
        ++u;
    }
    char *s = NULL;
    if (asprintf(&s, "%0.2f%s", v, unit[u]) < 0) s = NULL;
    return s;
}

double sptPrintAverageElapsedTime(const sptTimer timer, const int niters, const char *name)
{
    double t = sptElapsedTime(timer) / niters;
    fprintf(stdout, "[%s]: %.9lf s\n", name, t);
    return t;
}

int sptDumpIndexArray(sptIndex *array, sptNnzIndex const n, FILE *fp)
{
    SPTOL_CHECK_OS(fprintf(fp, "sptIndex array length: %" PRIu64 "\n", n) < 0, "IdxArray Dump");
    for (sptNnzIndex i = 0; i < n; ++i)
        SPTOL_CHECK_OS(fprintf(fp, "%" PRIu32 "\t", array[i]) < 0, "IdxArray Dump");
    fprintf(fp, "\n");
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Rank matrices and Kruskal tensors (src/matrix/dump.c:60,                 */
/* src/ktensor/dump.c:57)                                                    */
/* ------------------------------------------------------------------------ */
int sptDumpRankMatrix(sptRankMatrix *mtx, FILE *fp)
{
    SPTOL_CHECK_OS(fprintf(fp, "%" PRIu32 " x %" PRIu8 " matrix\n", mtx->nrows, mtx->ncols) < 0, "RankMtx Dump");
    for (sptIndex i = 0; i < mtx->nrows; ++i) {
        const sptValue *row = mtx->values + (size_t)i * mtx->stride;
        for (sptElementIndex j = 0; j < mtx->ncols; ++j)
            SPTOL_CHECK_OS(fprintf(fp, "%.2f\t", row[j]) < 0, "RankMtx Dump");
        fprintf(fp, "\n");
    }
    fprintf(fp, "\n");
    return 0;
}

/* `rank` lambdas and the nmodes factors, like sptDumpKruskalTensor (the
 * reference prints nmodes lambdas and also the MTTKRP scratch matrix,
 * src/ktensor/dump.c:77-86, which sptCpdAlsHiCOO leaves empty). */
int sptDumpRankKruskalTensor(sptRankKruskalTensor *ktsr, FILE *fp)
{
    SPTOL_CHECK_OS(fprintf(fp, "nmodes: %" PRIu32 ", rank: %" PRIu8 "\n", ktsr->nmodes, ktsr->rank) < 0,
                   "KruskalTns Dump");
    for (sptIndex m = 0; m < ktsr->nmodes; ++m) fprintf(fp, m ? " %" PRIu32 : "%" PRIu32, ktsr->ndims[m]);
    fputs("\n", fp);
    fprintf(fp, "fit: %lf\n", ktsr->fit);
    fprintf(fp, "lambda:\n");
    for (sptIndex r = 0; r < ktsr->rank; ++r) fprintf(fp, "%f ", ktsr->lambda[r]);
    fprintf(fp, "\nFactor matrices:\n");
    if (ktsr->factors)
        for (sptIndex m = 0; m < ktsr->nmodes; ++m) {
            int r = sptDumpRankMatrix(ktsr->factors[m], fp);
            if (r) return r;
        }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Mode orders (src/sptensor/sort.c:193-278)                                 */
/* ------------------------------------------------------------------------ */

/* Modes by increasing dimension.  The sort is a Hoare partition around the
 * middle slot, following that slot's element as swaps move it (the
 * reference's sptPairArraySort, src/vector/array.c:76-106): modes of equal
 * dimension may trade places, and which do decides the custom sort order a
 * driver then applies, so the same partition steps are kept. */
static void modes_by_dim(sptIndex *mode, sptIndex *dim, sptIndex lo, sptIndex hi)
{
    while (hi - lo >= 2) {
        sptIndex p = (lo + hi) / 2, i = lo, j = hi - 1;
        for (;; ++i, --j) {
            while (dim[i] < dim[p]) ++i;
            while (dim[p] < dim[j]) --j;
            if (i >= j) break;
            sptIndex t = mode[i]; mode[i] = mode[j]; mode[j] = t;
            t = dim[i]; dim[i] = dim[j]; dim[j] = t;
            if (i == p) p = j;
            else if (j == p) p = i;
        }
        modes_by_dim(mode, dim, lo, i);
        lo = i;
    }
}

static int sorted_modes(sptIndex *out, sptIndex const *ndims, sptIndex nmodes)
{
    sptIndex *dim = (sptIndex *)malloc((size_t)nmodes * sizeof(sptIndex));
    if (!dim) return -1;
    for (sptIndex m = 0; m < nmodes; ++m) { out[m] = m; dim[m] = ndims[m]; }
    modes_by_dim(out, dim, 0, nmodes);
    free(dim);
    return 0;
}

void sptGetBestModeOrder(sptIndex *mode_order, sptIndex const mode, sptIndex const *ndims, sptIndex const nmodes)
{
    if (sorted_modes(mode_order, ndims, nmodes)) return;
    /* move `mode` to the front, the modes before it one slot back */
    sptIndex at = 0;
    for (sptIndex m = 0; m < nmodes; ++m) if (mode_order[m] == mode) at = m;
    for (sptIndex m = at; m >= 1; --m) mode_order[m] = mode_order[m - 1];
    mode_order[0] = mode;
}

void sptGetWorstModeOrder(sptIndex *mode_order, sptIndex const mode, sptIndex const *ndims, sptIndex const nmodes)
{
    sptIndex *inc = (sptIndex *)malloc((size_t)nmodes * sizeof(sptIndex));
    if (!inc || sorted_modes(inc, ndims, nmodes)) { free(inc); return; }
    for (sptIndex m = 0; m < nmodes; ++m) mode_order[m] = inc[nmodes - 1 - m];
    free(inc);
    /* move `mode` to the back, the modes after it one slot forward */
    sptIndex at = 0;
    for (sptIndex m = 0; m < nmodes; ++m) if (mode_order[m] == mode) at = m;
    for (sptIndex m = at; m + 1 < nmodes; ++m) mode_order[m] = mode_order[m + 1];
    mode_order[nmodes - 1] = mode;
}

/* ------------------------------------------------------------------------ */
/* Status reports (src/sptensor/status.c:24-59, hicoo/status.c:28-113)       */
/* ------------------------------------------------------------------------ */
double sptSparseTensorDensity(sptSparseTensor const * const tsr)
{
    /* product over modes of nnz^(1/N) / I_m: the N-th root of the filled
     * fraction, one factor per mode */
    double root = pow((double)tsr->nnz, 1. / (double)tsr->nmodes);
    double d = 1.0;
    for (sptIndex m = 0; m < tsr->nmodes; ++m) d *= root / (double)tsr->ndims[m];
    return d;
}

void sptSparseTensorStatus(sptSparseTensor *tsr, FILE *fp)
{
    const sptIndex N = tsr->nmodes;
    fprintf(fp, "COO Sparse Tensor information (use sptIndex, sptValue))---------\n");
    for (sptIndex m = 0; m < N; ++m) fprintf(fp, m ? "x%" PRIu32 : "DIMS=%" PRIu32, tsr->ndims[m]);
    fprintf(fp, " NNZ=%" PRIu64 " DENSITY=%e\n", tsr->nnz, sptSparseTensorDensity(tsr));
    fprintf(fp, "Average slice length (c): ");
    for (sptIndex m = 0; m < N; ++m)
        fprintf(fp, m + 1 < N ? "%.2lf , " : "%.2lf\n", (double)tsr->nnz / tsr->ndims[m]);
    char *b = sptBytesString(tsr->nnz * (sizeof(sptIndex) * N + sizeof(sptValue)));
    fprintf(fp, "COO-STORAGE=%s\n\n", b ? b : "?");
    free(b);
}

static int cmp_u64(const void *a, const void *b)
{
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : x > y;
}

void sptSparseTensorStatusHiCOO(sptSparseTensorHiCOO *hitsr, FILE *fp)
{
    const sptIndex N = hitsr->nmodes;
    const sptIndex sb = (sptIndex)1 << hitsr->sb_bits, sk = (sptIndex)1 << hitsr->sk_bits;
    const sptNnzIndex nb = hitsr->bptr.len - 1;
    fprintf(fp, "HiCOO Sparse Tensor information ---------\n");
    for (sptIndex m = 0; m < N; ++m) fprintf(fp, m ? "x%" PRIu32 : "DIMS=%" PRIu32, hitsr->ndims[m]);
    fprintf(fp, " NNZ=%" PRIu64 "\n", hitsr->nnz);
    fprintf(fp, "sb=%" PRIu32 " sk=%" PRIu32 "\n", sb, sk);
    fprintf(fp, "nb=%" PRIu64 " nk=%" PRIu64 " nc=%" PRIu64 "\n", nb, hitsr->kptr.len - 1, hitsr->cptr.len - 1);

    /* storage: values, element offsets, block coordinates, the three
     * pointer arrays, the kernel schedule lists with their row pointers,
     * and nkiters */
    sptNnzIndex bytes = hitsr->nnz * (sizeof(sptValue) + N * sizeof(sptElementIndex));
    bytes += hitsr->binds[0].len * N * sizeof(sptBlockIndex);
    bytes += (hitsr->bptr.len + hitsr->kptr.len + hitsr->cptr.len) * sizeof(sptNnzIndex);
    for (sptIndex m = 0; m < N; ++m) {
        const sptIndex rows = (hitsr->ndims[m] + sk - 1) / sk;
        for (sptIndex i = 0; i < rows; ++i) bytes += hitsr->kschr[m][i].len * sizeof(sptIndex);
        bytes += rows * sizeof(sptIndexVector *);
    }
    bytes += N * sizeof(sptIndexVector **) + N * sizeof(sptIndex);
    char *b = sptBytesString(bytes);
    fprintf(fp, "HiCOO-STORAGE=%s\n", b ? b : "?");
    free(b);

    fprintf(fp, "SCHEDULE INFO [KERNEL]: \n");
    for (sptIndex m = 0; m < N; ++m)
        fprintf(fp, "SCHEDULE MODE %" PRIu32 " : %" PRIu32 " x %" PRIu32 "\n", m,
                (hitsr->ndims[m] + sk - 1) / sk, hitsr->nkiters[m]);

    fprintf(fp, "block nnzs:\n");
    sptNnzIndex *cnt = (sptNnzIndex *)malloc((nb ? nb : 1) * sizeof(sptNnzIndex));
    if (!cnt) return;
    sptNnzIndex mx = 0, mn = 0, sum = 0;
    double geo = 1;   /* running product of (nnzb / sb)^(1/nb), block order */
    for (sptNnzIndex i = 0; i < nb; ++i) {
        const sptNnzIndex c = hitsr->bptr.data[i + 1] - hitsr->bptr.data[i];
        if (i == 0 || c > mx) mx = c;
        if (i == 0 || c < mn) mn = c;
        sum += c;
        geo *= pow((double)c / sb, 1.0 / nb);
        cnt[i] = c;
    }
    const sptNnzIndex aver = nb ? sum / nb : 0;
    qsort(cnt, nb, sizeof(sptNnzIndex), cmp_u64);
    const sptNnzIndex median = nb ? cnt[(nb + 1) / 2 - 1] : 0;
    free(cnt);
    fprintf(fp, "Nnzb: Max=%" PRIu64 ", Min=%" PRIu64 ", Aver=%" PRIu64 "\n", mx, mn, aver);
    fprintf(fp, "cb: Max=%.3lf, Min=%.3lf, Aver=%.3lf\n", (double)mx / sb, (double)mn / sb, (double)aver / sb);
    fprintf(fp, "median cb: %.3lf, geometric mean cb: %.3lf\n", (double)median / sb, geo);
    fprintf(fp, "alpha_b: %lf\n", (double)nb / hitsr->nnz);
    fprintf(fp, "\nParameter configuration --------\n");
    fprintf(fp, "Suggest B (sb) <= %.2lf / R. For cache efficiency\n", (double)L1_SIZE / N / sizeof(sptValue));
    fprintf(fp, "Suggest alpha_b in (0,1], small is better. For tensor storage\n");
    fprintf(fp, "Suggest cb > 1, large is better. For MTTKRP performance\n");
    fprintf(fp, "Suggest num_tasks should in [%d, %d] PAR_DEGREE: [%d, %d]. For parallel efficiency\n",
            PAR_MIN_DEGREE * NUM_CORES, PAR_MAX_DEGREE * NUM_CORES, PAR_MIN_DEGREE, PAR_MAX_DEGREE);
    fprintf(fp, "\n\n");
}
