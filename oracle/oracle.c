/*
 * CPU oracle: a plain-C restatement of the reference's sequential MTTKRP,
 * HiCOO conversion and CPD-ALS.  TEST INFRASTRUCTURE ONLY -- the product
 * library never links or calls this file.  Each function cites the reference
 * code it follows; the pinning checks are listed in oracle.h.
 */
#define _GNU_SOURCE
#include "oracle.h"
#include <stdlib.h>
#include <string.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* COO MTTKRP                                                               */
/* ------------------------------------------------------------------------ */

/* src/sptensor/mttkrp.c:39-177.  Per nonzero x and column r the reference
 * forms s = v*A_{o1}[i_{o1}][r], then s *= A_{o2}[..][r], ... (the 3-D form
 * at :168-170 is the same association), and adds s to out[i_mode][r], in nnz
 * order.  The output is zero-filled first (:72,:146). */
void oracle_mttkrp_coo_f32(uint32_t nmodes, uint64_t nnz, const uint32_t *const *inds,
                           const float *vals, const float *const *mats, uint32_t stride,
                           uint32_t R, const uint32_t *mats_order, uint32_t mode,
                           uint32_t out_rows, float *out)
{
    memset(out, 0, (size_t)out_rows * stride * sizeof(float));
    float *s = (float *)malloc((R ? R : 1) * sizeof(float));
    for (uint64_t x = 0; x < nnz; ++x) {
        uint32_t m1 = mats_order[1];
        const float *row = mats[m1] + (size_t)inds[m1][x] * stride;
        float v = vals[x];
        for (uint32_t r = 0; r < R; ++r) s[r] = v * row[r];
        for (uint32_t i = 2; i < nmodes; ++i) {
            uint32_t mi = mats_order[i];
            const float *rw = mats[mi] + (size_t)inds[mi][x] * stride;
            for (uint32_t r = 0; r < R; ++r) s[r] *= rw[r];
        }
        float *o = out + (size_t)inds[mode][x] * stride;
        for (uint32_t r = 0; r < R; ++r) o[r] += s[r];
    }
    free(s);
}

void oracle_mttkrp_coo_f64(uint32_t nmodes, uint64_t nnz, const uint32_t *const *inds,
                           const float *vals, const float *const *mats, uint32_t stride,
                           uint32_t R, const uint32_t *mats_order, uint32_t mode,
                           uint32_t out_rows, double *out)
{
    memset(out, 0, (size_t)out_rows * stride * sizeof(double));
    for (uint64_t x = 0; x < nnz; ++x) {
        double *o = out + (size_t)inds[mode][x] * stride;
        for (uint32_t r = 0; r < R; ++r) {
            double s = (double)vals[x];
            for (uint32_t i = 1; i < nmodes; ++i) {
                uint32_t mi = mats_order[i];
                s *= (double)mats[mi][(size_t)inds[mi][x] * stride + r];
            }
            o[r] += s;
        }
    }
}

/* src/sptensor/mttkrp_omp.c:125-252: parallel over nonzeros, one atomic
 * update per (nnz, r).  Result equals the sequential one up to the order of
 * floating-point additions. */
void oracle_mttkrp_coo_omp_f32(uint32_t nmodes, uint64_t nnz, const uint32_t *const *inds,
                               const float *vals, const float *const *mats, uint32_t stride,
                               uint32_t R, const uint32_t *mats_order, uint32_t mode,
                               uint32_t out_rows, float *out, int tk)
{
    memset(out, 0, (size_t)out_rows * stride * sizeof(float));
    #pragma omp parallel num_threads(tk)
    {
        float *s = (float *)malloc((R ? R : 1) * sizeof(float));
        #pragma omp for schedule(static)
        for (uint64_t x = 0; x < nnz; ++x) {
            uint32_t m1 = mats_order[1];
            const float *row = mats[m1] + (size_t)inds[m1][x] * stride;
            float v = vals[x];
            for (uint32_t r = 0; r < R; ++r) s[r] = v * row[r];
            for (uint32_t i = 2; i < nmodes; ++i) {
                uint32_t mi = mats_order[i];
                const float *rw = mats[mi] + (size_t)inds[mi][x] * stride;
                for (uint32_t r = 0; r < R; ++r) s[r] *= rw[r];
            }
            float *o = out + (size_t)inds[mode][x] * stride;
            for (uint32_t r = 0; r < R; ++r) {
                #pragma omp atomic update
                o[r] += s[r];
            }
        }
        free(s);
    }
}

/* src/sptensor/mttkrp_omp.c:401-563 (sptOmpMTTKRP_Reduce and its 3-D form):
 * the nonzeros split statically over tk threads, each adding into its own
 * copy of the output, then every output row summed over the copies in
 * thread order.  With the same OpenMP runtime and tk the partition is the
 * reference's, so the result is bit-identical to it. */
void oracle_mttkrp_coo_omp_reduce_f32(uint32_t nmodes, uint64_t nnz, const uint32_t *const *inds,
                                      const float *vals, const float *const *mats, uint32_t stride,
                                      uint32_t R, const uint32_t *mats_order, uint32_t mode,
                                      uint32_t out_rows, float *out, int tk)
{
    const size_t per = (size_t)out_rows * stride;
    float *copies = (float *)calloc((size_t)tk * per + 1, sizeof(float));
    memset(out, 0, per * sizeof(float));
    #pragma omp parallel num_threads(tk)
    {
        float *s = (float *)malloc((R ? R : 1) * sizeof(float));
        float *mine = copies + (size_t)omp_get_thread_num() * per;
        #pragma omp for schedule(static)
        for (uint64_t x = 0; x < nnz; ++x) {
            uint32_t m1 = mats_order[1];
            const float *row = mats[m1] + (size_t)inds[m1][x] * stride;
            float v = vals[x];
            for (uint32_t r = 0; r < R; ++r) s[r] = v * row[r];
            for (uint32_t i = 2; i < nmodes; ++i) {
                uint32_t mi = mats_order[i];
                const float *rw = mats[mi] + (size_t)inds[mi][x] * stride;
                for (uint32_t r = 0; r < R; ++r) s[r] *= rw[r];
            }
            float *o = mine + (size_t)inds[mode][x] * stride;
            for (uint32_t r = 0; r < R; ++r) o[r] += s[r];
        }
        free(s);
    }
    #pragma omp parallel for schedule(static) num_threads(tk)
    for (uint32_t i = 0; i < out_rows; ++i)
        for (int t = 0; t < tk; ++t)
            for (uint32_t r = 0; r < R; ++r)
                out[(size_t)i * stride + r] += copies[(size_t)t * per + (size_t)i * stride + r];
    free(copies);
}

/* ------------------------------------------------------------------------ */
/* Orders                                                                   */
/* ------------------------------------------------------------------------ */

typedef struct {
    uint32_t nmodes;
    uint32_t bits;           /* row-block shift */
    uint32_t *const *inds;
} sort_ctx;

/* sort.c:629-642: lexicographic, mode 0 most significant. */
static int cmp_lex(const void *pa, const void *pb, void *pc)
{
    const sort_ctx *c = (const sort_ctx *)pc;
    uint64_t a = *(const uint64_t *)pa, b = *(const uint64_t *)pb;
    for (uint32_t m = 0; m < c->nmodes; ++m) {
        uint32_t x = c->inds[m][a], y = c->inds[m][b];
        if (x != y) return x < y ? -1 : 1;
    }
    return 0;
}

/* sort.c:713-736: lexicographic on (i_m >> sk_bits). */
static int cmp_rowblock(const void *pa, const void *pb, void *pc)
{
    const sort_ctx *c = (const sort_ctx *)pc;
    uint64_t a = *(const uint64_t *)pa, b = *(const uint64_t *)pb;
    for (uint32_t m = 0; m < c->nmodes; ++m) {
        uint32_t x = c->inds[m][a] >> c->bits, y = c->inds[m][b] >> c->bits;
        if (x != y) return x < y ? -1 : 1;
    }
    return 0;
}

/* Z-Morton order of sort.c:748-805 (3-D) and :816-995 (4-D): bits are
 * interleaved with mode 0 the most significant bit of each group, so two
 * coordinates compare on the mode holding the highest differing bit, the
 * lower mode winning a tie in bit position.  The 3-D key keeps only bits
 * 0-23 of each index (its top byte is shifted out of the 128-bit key); the
 * 4-D key keeps all 32 bits.  Restated here as a most-significant-differing-
 * bit comparison instead of a key build. */
static int msb_less(uint32_t a, uint32_t b) { return a < b && a < (a ^ b); }

int oracle_morton_cmp(uint32_t nmodes, const uint32_t *a, const uint32_t *b)
{
    uint32_t mask = nmodes == 3 ? 0x00FFFFFFu : 0xFFFFFFFFu;
    uint32_t best = 0, dbest = (a[0] ^ b[0]) & mask;
    for (uint32_t m = 1; m < nmodes; ++m) {
        uint32_t d = (a[m] ^ b[m]) & mask;
        if (msb_less(dbest, d)) { best = m; dbest = d; }
    }
    if (dbest == 0) return 0;
    return (a[best] & mask) < (b[best] & mask) ? -1 : 1;
}

static int cmp_morton(const void *pa, const void *pb, void *pc)
{
    const sort_ctx *c = (const sort_ctx *)pc;
    uint64_t a = *(const uint64_t *)pa, b = *(const uint64_t *)pb;
    uint32_t ca[8], cb[8];
    for (uint32_t m = 0; m < c->nmodes; ++m) { ca[m] = c->inds[m][a]; cb[m] = c->inds[m][b]; }
    return oracle_morton_cmp(c->nmodes, ca, cb);
}

/* Apply the permutation perm[begin..end) to inds/vals in place (via copy). */
static void apply_perm(uint32_t nmodes, uint64_t begin, uint64_t end, const uint64_t *perm,
                       uint32_t **inds, float *vals)
{
    uint64_t n = end - begin;
    if (n == 0) return;
    uint32_t *ti = (uint32_t *)malloc(n * sizeof(uint32_t));
    float *tv = (float *)malloc(n * sizeof(float));
    for (uint32_t m = 0; m < nmodes; ++m) {
        for (uint64_t i = 0; i < n; ++i) ti[i] = inds[m][perm[i]];
        memcpy(inds[m] + begin, ti, n * sizeof(uint32_t));
    }
    for (uint64_t i = 0; i < n; ++i) tv[i] = vals[perm[i]];
    memcpy(vals + begin, tv, n * sizeof(float));
    free(ti);
    free(tv);
}

static void sort_range(uint32_t nmodes, uint64_t begin, uint64_t end, uint32_t **inds,
                       float *vals, uint32_t bits,
                       int (*cmp)(const void *, const void *, void *))
{
    uint64_t n = end - begin;
    if (n < 2) return;
    uint64_t *perm = (uint64_t *)malloc(n * sizeof(uint64_t));
    for (uint64_t i = 0; i < n; ++i) perm[i] = begin + i;
    sort_ctx c = { nmodes, bits, inds };
    qsort_r(perm, n, sizeof(uint64_t), cmp, &c);
    apply_perm(nmodes, begin, end, perm, inds, vals);
    free(perm);
}

void oracle_sort_lex(uint32_t nmodes, uint64_t nnz, uint32_t **inds, float *vals)
{
    sort_range(nmodes, 0, nnz, inds, vals, 0, cmp_lex);
}

/* ------------------------------------------------------------------------ */
/* HiCOO conversion (src/sptensor/hicoo/convert.c:103-497)                  */
/* ------------------------------------------------------------------------ */

typedef struct { uint64_t len, cap; uint64_t *d; } u64vec;
static void u64push(u64vec *v, uint64_t x)
{
    if (v->len == v->cap) { v->cap = v->cap ? v->cap * 2 : 16; v->d = (uint64_t *)realloc(v->d, v->cap * 8); }
    v->d[v->len++] = x;
}

int oracle_hicoo_convert(uint32_t nmodes, const uint32_t *ndims, uint64_t nnz,
                         uint32_t **inds, float *vals, uint32_t sb_bits, uint32_t sk_bits,
                         oracle_hicoo *h)
{
    memset(h, 0, sizeof(*h));
    if (nmodes < 2 || nmodes > 8 || nnz == 0 || sk_bits < sb_bits) return -1;
    h->nmodes = nmodes; h->sb_bits = sb_bits; h->sk_bits = sk_bits; h->sc_bits = 14;
    h->nnz = nnz;

    /* 1. row-block sort on (i >> sk) (convert.c:173 -> sort.c:462). */
    sort_range(nmodes, 0, nnz, inds, vals, sk_bits, cmp_rowblock);

    /* 2. superblock (kernel) pointers as nonzero offsets (convert.c:237-288). */
    u64vec kp = {0, 0, NULL};
    u64push(&kp, 0);
    for (uint64_t z = 1; z < nnz; ++z) {
        int same = 1;
        for (uint32_t m = 0; m < nmodes; ++m)
            if ((inds[m][z] >> sk_bits) != (inds[m][z - 1] >> sk_bits)) { same = 0; break; }
        if (!same) u64push(&kp, z);
    }
    u64push(&kp, nnz);
    uint64_t nk = kp.len - 1;

    /* 3. kernel scheduler: kschr[m][i_m >> sk] lists kernels in ascending
     *    order, nkiters[m] = longest list (convert.c:103-146). */
    h->nkd = (uint32_t *)malloc(nmodes * sizeof(uint32_t));
    h->nkiters = (uint32_t *)calloc(nmodes, sizeof(uint32_t));
    uint64_t sk = 1ull << sk_bits;
    uint64_t tot_ptr = 0;
    for (uint32_t m = 0; m < nmodes; ++m) { h->nkd[m] = (uint32_t)((ndims[m] + sk - 1) / sk); tot_ptr += h->nkd[m] + 1; }
    h->kschr_ptr = (uint64_t *)calloc(tot_ptr, sizeof(uint64_t));
    h->kschr_idx = (uint32_t *)malloc((nk ? nk : 1) * nmodes * sizeof(uint32_t));
    {
        uint64_t poff = 0;
        for (uint32_t m = 0; m < nmodes; ++m) {
            uint64_t *ptr = h->kschr_ptr + poff;
            uint32_t *idx = h->kschr_idx + (uint64_t)m * nk;
            for (uint64_t k = 0; k < nk; ++k) ptr[(inds[m][kp.d[k]] >> sk_bits) + 1]++;
            for (uint32_t i = 0; i < h->nkd[m]; ++i) {
                if (ptr[i + 1] > h->nkiters[m]) h->nkiters[m] = (uint32_t)ptr[i + 1];
                ptr[i + 1] += ptr[i];
            }
            uint64_t *fill = (uint64_t *)malloc((h->nkd[m] + 1) * sizeof(uint64_t));
            memcpy(fill, ptr, (h->nkd[m] + 1) * sizeof(uint64_t));
            for (uint64_t k = 0; k < nk; ++k) idx[fill[inds[m][kp.d[k]] >> sk_bits]++] = (uint32_t)k;
            free(fill);
            poff += h->nkd[m] + 1;
        }
    }

    /* 4. Morton sort inside every superblock (convert.c:205-219). */
    for (uint64_t k = 0; k < nk; ++k) sort_range(nmodes, kp.d[k], kp.d[k + 1], inds, vals, 0, cmp_morton);

    /* 5. one pass emitting blocks, chunks and element offsets
     *    (convert.c:358-472). */
    u64vec bptr = {0, 0, NULL}, cptr = {0, 0, NULL};
    uint64_t bcap = 1024, nb = 0;
    uint32_t *binds = (uint32_t *)malloc(bcap * nmodes * sizeof(uint32_t));
    h->einds = (uint8_t *)malloc(nnz * nmodes);
    h->values = (float *)malloc(nnz * sizeof(float));
    h->kptr = (uint64_t *)malloc((nk + 1) * sizeof(uint64_t));
    uint64_t sc = 1ull << h->sc_bits;
    uint64_t ne = 0, chunk_size = 0;
    uint32_t prior[8];
    for (uint64_t k = 0; k < nk; ++k) {
        h->kptr[k] = nb;                  /* block offset of this superblock */
        u64push(&cptr, nb);               /* a chunk starts with each kernel  */
        chunk_size = 0;
        for (uint64_t z = kp.d[k]; z < kp.d[k + 1]; ++z) {
            int newblock = (z == 0);
            if (!newblock)
                for (uint32_t m = 0; m < nmodes; ++m)
                    if ((inds[m][z] >> sb_bits) != prior[m]) { newblock = 1; break; }
            for (uint32_t m = 0; m < nmodes; ++m)
                h->einds[(uint64_t)m * nnz + z] = (uint8_t)(inds[m][z] - ((inds[m][z] >> sb_bits) << sb_bits));
            h->values[z] = vals[z];
            if (!newblock) { ++ne; continue; }
            if (nb == bcap) { bcap *= 2; binds = (uint32_t *)realloc(binds, bcap * nmodes * sizeof(uint32_t)); }
            for (uint32_t m = 0; m < nmodes; ++m) { prior[m] = inds[m][z] >> sb_bits; binds[nb * nmodes + m] = prior[m]; }
            u64push(&bptr, z);
            if (z != 0) {
                /* the chunk rule reads the previous block's count (:445-452) */
                if (chunk_size + ne >= sc) { u64push(&cptr, nb); chunk_size = 0; }
                else chunk_size += ne;
            }
            ++nb;
            ne = 1;
        }
    }
    h->kptr[nk] = nb;
    u64push(&cptr, nb);
    u64push(&bptr, nnz);
    h->nb = nb; h->nk = nk; h->nc = cptr.len;
    h->bptr = bptr.d; h->cptr = cptr.d;
    /* transpose block coordinates to mode-major */
    h->binds = (uint32_t *)malloc((nb ? nb : 1) * nmodes * sizeof(uint32_t));
    for (uint64_t b = 0; b < nb; ++b)
        for (uint32_t m = 0; m < nmodes; ++m) h->binds[(uint64_t)m * nb + b] = binds[b * nmodes + m];
    free(binds);
    h->max_nnzb = 0;
    for (uint64_t b = 0; b < nb; ++b) {
        uint64_t c = h->bptr[b + 1] - h->bptr[b];
        if (c > h->max_nnzb) h->max_nnzb = c;
    }
    free(kp.d);
    return 0;
}

void oracle_hicoo_free(oracle_hicoo *h)
{
    free(h->kptr); free(h->cptr); free(h->bptr); free(h->binds); free(h->einds);
    free(h->values); free(h->nkiters); free(h->nkd); free(h->kschr_ptr); free(h->kschr_idx);
    memset(h, 0, sizeof(*h));
}

/* ------------------------------------------------------------------------ */
/* HiCOO MTTKRP (src/sptensor/hicoo/mttkrp.c:53-144, 251-329)               */
/* ------------------------------------------------------------------------ */

/* Kernels -> blocks -> entries; coordinate = (binds << sb_bits) + einds.
 * Same per-entry association as the COO loop. */
void oracle_mttkrp_hicoo_f32(const oracle_hicoo *h, const float *const *mats, uint32_t stride,
                             uint32_t R, const uint32_t *mats_order, uint32_t mode,
                             uint32_t out_rows, float *out)
{
    uint32_t N = h->nmodes;
    memset(out, 0, (size_t)out_rows * stride * sizeof(float));
    float *s = (float *)malloc((R ? R : 1) * sizeof(float));
    uint32_t c[8];
    for (uint64_t k = 0; k < h->nk; ++k)
        for (uint64_t b = h->kptr[k]; b < h->kptr[k + 1]; ++b)
            for (uint64_t z = h->bptr[b]; z < h->bptr[b + 1]; ++z) {
                for (uint32_t m = 0; m < N; ++m)
                    c[m] = (h->binds[(uint64_t)m * h->nb + b] << h->sb_bits) + h->einds[(uint64_t)m * h->nnz + z];
                uint32_t m1 = mats_order[1];
                const float *row = mats[m1] + (size_t)c[m1] * stride;
                float v = h->values[z];
                for (uint32_t r = 0; r < R; ++r) s[r] = v * row[r];
                for (uint32_t i = 2; i < N; ++i) {
                    uint32_t mi = mats_order[i];
                    const float *rw = mats[mi] + (size_t)c[mi] * stride;
                    for (uint32_t r = 0; r < R; ++r) s[r] *= rw[r];
                }
                float *o = out + (size_t)c[mode] * stride;
                for (uint32_t r = 0; r < R; ++r) o[r] += s[r];
            }
    free(s);
}

void oracle_mttkrp_hicoo_f64(const oracle_hicoo *h, const float *const *mats, uint32_t stride,
                             uint32_t R, const uint32_t *mats_order, uint32_t mode,
                             uint32_t out_rows, double *out)
{
    uint32_t N = h->nmodes;
    memset(out, 0, (size_t)out_rows * stride * sizeof(double));
    uint32_t c[8];
    for (uint64_t b = 0; b < h->nb; ++b)
        for (uint64_t z = h->bptr[b]; z < h->bptr[b + 1]; ++z) {
            for (uint32_t m = 0; m < N; ++m)
                c[m] = (h->binds[(uint64_t)m * h->nb + b] << h->sb_bits) + h->einds[(uint64_t)m * h->nnz + z];
            double *o = out + (size_t)c[mode] * stride;
            for (uint32_t r = 0; r < R; ++r) {
                double s = (double)h->values[z];
                for (uint32_t i = 1; i < N; ++i) {
                    uint32_t mi = mats_order[i];
                    s *= (double)mats[mi][(size_t)c[mi] * stride + r];
                }
                o[r] += s;
            }
        }
}

/* src/sptensor/hicoo/mttkrp_omp.c:555-820: iteration i processes, in
 * parallel over superblock rows k of the output mode, the i-th superblock
 * listed in kschr[mode][k].  Superblock rows own disjoint output rows, so no
 * atomics are needed. */
void oracle_mttkrp_hicoo_sched_omp_f32(const oracle_hicoo *h, const float *const *mats,
                                       uint32_t stride, uint32_t R, const uint32_t *mats_order,
                                       uint32_t mode, uint32_t out_rows, float *out, int tk)
{
    uint32_t N = h->nmodes;
    memset(out, 0, (size_t)out_rows * stride * sizeof(float));
    uint64_t poff = 0;
    for (uint32_t m = 0; m < mode; ++m) poff += h->nkd[m] + 1;
    const uint64_t *kptr_m = h->kschr_ptr + poff;
    const uint32_t *kidx_m = h->kschr_idx + (uint64_t)mode * h->nk;
    uint32_t nkd = h->nkd[mode];
    for (uint32_t it = 0; it < h->nkiters[mode]; ++it) {
        #pragma omp parallel num_threads(tk)
        {
            float *s = (float *)malloc((R ? R : 1) * sizeof(float));
            uint32_t c[8];
            #pragma omp for schedule(dynamic, 1)
            for (uint32_t kr = 0; kr < nkd; ++kr) {
                if (kptr_m[kr] + it >= kptr_m[kr + 1]) continue;
                uint64_t k = kidx_m[kptr_m[kr] + it];
                for (uint64_t b = h->kptr[k]; b < h->kptr[k + 1]; ++b)
                    for (uint64_t z = h->bptr[b]; z < h->bptr[b + 1]; ++z) {
                        for (uint32_t m = 0; m < N; ++m)
                            c[m] = (h->binds[(uint64_t)m * h->nb + b] << h->sb_bits) + h->einds[(uint64_t)m * h->nnz + z];
                        uint32_t m1 = mats_order[1];
                        const float *row = mats[m1] + (size_t)c[m1] * stride;
                        float v = h->values[z];
                        for (uint32_t r = 0; r < R; ++r) s[r] = v * row[r];
                        for (uint32_t i = 2; i < N; ++i) {
                            uint32_t mi = mats_order[i];
                            const float *rw = mats[mi] + (size_t)c[mi] * stride;
                            for (uint32_t r = 0; r < R; ++r) s[r] *= rw[r];
                        }
                        float *o = out + (size_t)c[mode] * stride;
                        for (uint32_t r = 0; r < R; ++r) o[r] += s[r];
                    }
            }
            free(s);
        }
    }
}

/* src/sptensor/hicoo/mttkrp_omp.c:1214-1503 (Scheduled_Reduce): iteration i
 * of the kernel schedule is the parallel loop (dynamic, chunk 1); it walks
 * the i-th kernel of every superblock row of the output mode, adding into
 * the running thread's own copy of the output; the copies are then summed
 * in thread order.  Which thread takes an iteration is up to the runtime,
 * so results agree with the sequential kernel up to the order of additions. */
void oracle_mttkrp_hicoo_sched_reduce_omp_f32(const oracle_hicoo *h, const float *const *mats,
                                              uint32_t stride, uint32_t R, const uint32_t *mats_order,
                                              uint32_t mode, uint32_t out_rows, float *out, int tk)
{
    uint32_t N = h->nmodes;
    const size_t per = (size_t)out_rows * stride;
    float *copies = (float *)calloc((size_t)tk * per + 1, sizeof(float));
    memset(out, 0, per * sizeof(float));
    uint64_t poff = 0;
    for (uint32_t m = 0; m < mode; ++m) poff += h->nkd[m] + 1;
    const uint64_t *kptr_m = h->kschr_ptr + poff;
    const uint32_t *kidx_m = h->kschr_idx + (uint64_t)mode * h->nk;
    uint32_t nkd = h->nkd[mode];
    #pragma omp parallel num_threads(tk)
    {
        float *s = (float *)malloc((R ? R : 1) * sizeof(float));
        float *mine = copies + (size_t)omp_get_thread_num() * per;
        uint32_t c[8];
        #pragma omp for schedule(dynamic, 1)
        for (uint32_t it = 0; it < h->nkiters[mode]; ++it)
            for (uint32_t kr = 0; kr < nkd; ++kr) {
                if (kptr_m[kr] + it >= kptr_m[kr + 1]) continue;
                uint64_t k = kidx_m[kptr_m[kr] + it];
                for (uint64_t b = h->kptr[k]; b < h->kptr[k + 1]; ++b)
                    for (uint64_t z = h->bptr[b]; z < h->bptr[b + 1]; ++z) {
                        for (uint32_t m = 0; m < N; ++m)
                            c[m] = (h->binds[(uint64_t)m * h->nb + b] << h->sb_bits) + h->einds[(uint64_t)m * h->nnz + z];
                        uint32_t m1 = mats_order[1];
                        const float *row = mats[m1] + (size_t)c[m1] * stride;
                        float v = h->values[z];
                        for (uint32_t r = 0; r < R; ++r) s[r] = v * row[r];
                        for (uint32_t i = 2; i < N; ++i) {
                            uint32_t mi = mats_order[i];
                            const float *rw = mats[mi] + (size_t)c[mi] * stride;
                            for (uint32_t r = 0; r < R; ++r) s[r] *= rw[r];
                        }
                        float *o = mine + (size_t)c[mode] * stride;
                        for (uint32_t r = 0; r < R; ++r) o[r] += s[r];
                    }
            }
        free(s);
    }
    #pragma omp parallel for schedule(static) num_threads(tk)
    for (uint32_t i = 0; i < out_rows; ++i)
        for (int t = 0; t < tk; ++t)
            for (uint32_t r = 0; r < R; ++r)
                out[(size_t)i * stride + r] += copies[(size_t)t * per + (size_t)i * stride + r];
    free(copies);
}

/* ------------------------------------------------------------------------ */
/* CPD-ALS in FP64 (src/sptensor/cpd.c:32-137, src/matrix/matrix.c:341-573, */
/* src/matrix/matrix_solver.c:32-83, src/ktensor/ktensor.c:102-257)         */
/* ------------------------------------------------------------------------ */

static void gram(const double *A, uint32_t I, uint32_t R, double *G)
{
    memset(G, 0, (size_t)R * R * sizeof(double));
    for (uint32_t i = 0; i < I; ++i)
        for (uint32_t a = 0; a < R; ++a)
            for (uint32_t b = 0; b < R; ++b) G[a * R + b] += A[(size_t)i * R + a] * A[(size_t)i * R + b];
}

/* Cholesky V = L L^T; returns 0 if V is SPD. */
static int cholesky(double *L, const double *V, uint32_t R)
{
    memcpy(L, V, (size_t)R * R * sizeof(double));
    for (uint32_t j = 0; j < R; ++j) {
        double d = L[j * R + j];
        for (uint32_t k = 0; k < j; ++k) d -= L[j * R + k] * L[j * R + k];
        if (!(d > 0.0)) return -1;
        d = sqrt(d);
        L[j * R + j] = d;
        for (uint32_t i = j + 1; i < R; ++i) {
            double s = L[i * R + j];
            for (uint32_t k = 0; k < j; ++k) s -= L[i * R + k] * L[j * R + k];
            L[i * R + j] = s / d;
        }
        for (uint32_t k = j + 1; k < R; ++k) L[j * R + k] = 0.0;
    }
    return 0;
}

/* Gaussian elimination with partial pivoting on [V | rhs^T]: the gesv
 * fallback of matrix_solver.c:66-79. */
static void gesv_rows(const double *V, uint32_t R, double *M, uint32_t I)
{
    double *LU = (double *)malloc((size_t)R * R * sizeof(double));
    uint32_t *piv = (uint32_t *)malloc(R * sizeof(uint32_t));
    memcpy(LU, V, (size_t)R * R * sizeof(double));
    for (uint32_t j = 0; j < R; ++j) {
        uint32_t p = j;
        for (uint32_t i = j + 1; i < R; ++i) if (fabs(LU[i * R + j]) > fabs(LU[p * R + j])) p = i;
        piv[j] = p;
        if (p != j) for (uint32_t k = 0; k < R; ++k) { double t = LU[j * R + k]; LU[j * R + k] = LU[p * R + k]; LU[p * R + k] = t; }
        for (uint32_t i = j + 1; i < R; ++i) {
            LU[i * R + j] /= LU[j * R + j];
            for (uint32_t k = j + 1; k < R; ++k) LU[i * R + k] -= LU[i * R + j] * LU[j * R + k];
        }
    }
    double *x = (double *)malloc(R * sizeof(double));
    for (uint32_t row = 0; row < I; ++row) {
        double *b = M + (size_t)row * R;
        memcpy(x, b, R * sizeof(double));
        for (uint32_t j = 0; j < R; ++j) if (piv[j] != j) { double t = x[j]; x[j] = x[piv[j]]; x[piv[j]] = t; }
        for (uint32_t i = 0; i < R; ++i) for (uint32_t k = 0; k < i; ++k) x[i] -= LU[i * R + k] * x[k];
        for (int i = (int)R - 1; i >= 0; --i) {
            for (uint32_t k = i + 1; k < R; ++k) x[i] -= LU[i * R + k] * x[k];
            x[i] /= LU[i * R + i];
        }
        memcpy(b, x, R * sizeof(double));
    }
    free(x); free(piv); free(LU);
}

int oracle_cpd_als_f64(uint32_t nmodes, const uint32_t *ndims, uint64_t nnz,
                       const uint32_t *const *inds, const float *vals, uint32_t R,
                       uint32_t niters, double tol, double **mats, double *lambda,
                       double *fit_hist)
{
    uint32_t N = nmodes, maxI = 0;
    for (uint32_t m = 0; m < N; ++m) if (ndims[m] > maxI) maxI = ndims[m];
    double **ata = (double **)malloc(N * sizeof(double *));
    for (uint32_t m = 0; m < N; ++m) { ata[m] = (double *)malloc((size_t)R * R * sizeof(double)); gram(mats[m], ndims[m], R, ata[m]); }
    double *V = (double *)malloc((size_t)R * R * sizeof(double));
    double *L = (double *)malloc((size_t)R * R * sizeof(double));
    double *M = (double *)malloc((size_t)maxI * R * sizeof(double));
    double *tmp = (double *)malloc(R * sizeof(double));
    double normX = 0.0;
    for (uint64_t x = 0; x < nnz; ++x) normX += (double)vals[x] * (double)vals[x];   /* sptensor.c:96-108 */
    double fit = 0.0, oldfit = 0.0;
    int done = 0;
    for (uint32_t it = 0; it < niters; ++it) {
        for (uint32_t m = 0; m < N; ++m) {
            /* mats_order = [m, m+1, ...] mod N (cpd.c:85-87); the product over
             * the other modes is taken in that order. */
            uint32_t I = ndims[m];
            memset(M, 0, (size_t)I * R * sizeof(double));
            for (uint64_t x = 0; x < nnz; ++x) {
                double *o = M + (size_t)inds[m][x] * R;
                for (uint32_t r = 0; r < R; ++r) {
                    double s = (double)vals[x];
                    for (uint32_t k = 1; k < N; ++k) {
                        uint32_t p = (m + k) % N;
                        s *= mats[p][(size_t)inds[p][x] * R + r];
                    }
                    o[r] += s;
                }
            }
            /* V = Hadamard of the other Grams (matrix.c:341-386) */
            for (uint32_t q = 0; q < R * R; ++q) V[q] = 1.0;
            for (uint32_t k = 1; k < N; ++k) {
                uint32_t p = (m + k) % N;
                for (uint32_t q = 0; q < R * R; ++q) V[q] *= ata[p][q];
            }
            /* A_m = M V^{-1} by Cholesky, else gesv (matrix_solver.c:32-83) */
            if (cholesky(L, V, R) == 0) {
                for (uint32_t i = 0; i < I; ++i) {
                    double *b = M + (size_t)i * R;
                    for (uint32_t a = 0; a < R; ++a) {
                        double s = b[a];
                        for (uint32_t k = 0; k < a; ++k) s -= L[a * R + k] * b[k];
                        b[a] = s / L[a * R + a];
                    }
                    for (int a = (int)R - 1; a >= 0; --a) {
                        double s = b[a];
                        for (uint32_t k = a + 1; k < R; ++k) s -= L[k * R + a] * b[k];
                        b[a] = s / L[a * R + a];
                    }
                }
            } else {
                gesv_rows(V, R, M, I);
            }
            double *A = mats[m];
            memcpy(A, M, (size_t)I * R * sizeof(double));
            /* normalise: column 2-norms on the first iteration, column max
             * (floored at 1) afterwards (cpd.c:99-103, matrix.c:390-554) */
            for (uint32_t r = 0; r < R; ++r) lambda[r] = 0.0;
            if (it == 0) {
                for (uint32_t i = 0; i < I; ++i) for (uint32_t r = 0; r < R; ++r) lambda[r] += A[(size_t)i * R + r] * A[(size_t)i * R + r];
                for (uint32_t r = 0; r < R; ++r) lambda[r] = sqrt(lambda[r]);
            } else {
                for (uint32_t i = 0; i < I; ++i) for (uint32_t r = 0; r < R; ++r) if (A[(size_t)i * R + r] > lambda[r]) lambda[r] = A[(size_t)i * R + r];
                for (uint32_t r = 0; r < R; ++r) if (lambda[r] < 1.0) lambda[r] = 1.0;
            }
            for (uint32_t i = 0; i < I; ++i) for (uint32_t r = 0; r < R; ++r) A[(size_t)i * R + r] /= lambda[r];
            gram(A, I, R, ata[m]);
        }
        /* fit (ktensor.c:102-118) with the full symmetric ||K||^2 */
        double normK = 0.0;
        for (uint32_t a = 0; a < R; ++a)
            for (uint32_t b = 0; b < R; ++b) {
                double g = 1.0;
                for (uint32_t m = 0; m < N; ++m) g *= ata[m][a * R + b];
                normK += g * lambda[a] * lambda[b];
            }
        normK = fabs(normK);
        /* <X, K> = sum_r lambda_r sum_i A_last[i][r] * M_last[i][r], where
         * M_last is the last mode's MTTKRP (ktensor.c:173-257).  The solve
         * overwrote that buffer, so recompute it: it does not depend on
         * A_last, hence equals the value the reference reuses. */
        uint32_t last = N - 1, I = ndims[last];
        memset(M, 0, (size_t)I * R * sizeof(double));
        for (uint64_t x = 0; x < nnz; ++x) {
            double *o = M + (size_t)inds[last][x] * R;
            for (uint32_t r = 0; r < R; ++r) {
                double s = (double)vals[x];
                for (uint32_t k = 1; k < N; ++k) {
                    uint32_t p = (last + k) % N;
                    s *= mats[p][(size_t)inds[p][x] * R + r];
                }
                o[r] += s;
            }
        }
        double inner = 0.0;
        for (uint32_t r = 0; r < R; ++r) tmp[r] = 0.0;
        for (uint32_t i = 0; i < I; ++i) for (uint32_t r = 0; r < R; ++r) tmp[r] += mats[last][(size_t)i * R + r] * M[(size_t)i * R + r];
        for (uint32_t r = 0; r < R; ++r) inner += tmp[r] * lambda[r];
        double residual = normX + normK - 2.0 * inner;
        if (residual > 0.0) residual = sqrt(residual);
        fit = 1.0 - residual / sqrt(normX);
        fit_hist[it] = fit;
        done = (int)it + 1;
        if (it > 0 && fabs(fit - oldfit) < tol) break;
        oldfit = fit;
    }
    /* GetFinalLambda (matrix.c:557-573): renormalise every factor by column
     * 2-norms and fold the norms into lambda. */
    for (uint32_t m = 0; m < N; ++m) {
        uint32_t I = ndims[m];
        double *A = mats[m];
        for (uint32_t r = 0; r < R; ++r) tmp[r] = 0.0;
        for (uint32_t i = 0; i < I; ++i) for (uint32_t r = 0; r < R; ++r) tmp[r] += A[(size_t)i * R + r] * A[(size_t)i * R + r];
        for (uint32_t r = 0; r < R; ++r) tmp[r] = sqrt(tmp[r]);
        for (uint32_t i = 0; i < I; ++i) for (uint32_t r = 0; r < R; ++r) A[(size_t)i * R + r] /= tmp[r];
        for (uint32_t r = 0; r < R; ++r) lambda[r] *= tmp[r];
    }
    for (uint32_t m = 0; m < N; ++m) free(ata[m]);
    free(ata); free(V); free(L); free(M); free(tmp);
    return done;
}
