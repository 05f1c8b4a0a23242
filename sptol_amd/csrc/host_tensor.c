/*
 * COO and HiCOO containers, text I/O, orders and the COO -> HiCOO converter.
 *
 * The orders are the reference's (lexicographic: src/sptensor/sort.c:629-642;
 * row-block: :713-736; Z-Morton: :748-995; HiCOO layout: hicoo/convert.c).
 * The reference sorts with recursive comparison quicksorts; this file instead
 * builds integer sort keys and runs a parallel stable LSD radix sort, which
 * gives the same total order on distinct coordinates and stays fast at 1e8
 * nonzeros.  The HiCOO layout pass reproduces convert.c:358-472 exactly,
 * including its chunk rule.
 */
#define _GNU_SOURCE
#include "host_internal.h"
#include <stdio.h>
#include <ctype.h>
#include <math.h>
#include <unistd.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int max_threads(int tk)
{
#ifdef _OPENMP
    int n = omp_get_max_threads();
    if (tk > 0 && tk < n) n = tk;
    return n > 0 ? n : 1;
#else
    (void)tk;
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Parallel stable LSD radix sort on (key, idx)                             */
/* ------------------------------------------------------------------------ */
#define RADIX_BITS 11
#define RADIX_BUCKETS (1u << RADIX_BITS)

int sptol_radix_sort_u64(uint64_t *keys, uint64_t *idx, uint64_t n, unsigned bits, int nthreads)
{
    if (n < 2 || bits == 0) return 0;
    int nt = max_threads(nthreads);
    if ((uint64_t)nt * 65536 > n) nt = (int)(n / 65536) + 1;
    uint64_t *k2 = (uint64_t *)malloc(n * sizeof(uint64_t));
    uint64_t *i2 = (uint64_t *)malloc(n * sizeof(uint64_t));
    uint64_t *hist = (uint64_t *)malloc((size_t)nt * RADIX_BUCKETS * sizeof(uint64_t));
    if (!k2 || !i2 || !hist) { free(k2); free(i2); free(hist); errno = ENOMEM; return SPTERR_OS_ERROR + ENOMEM; }
    uint64_t *ks = keys, *is = idx, *kd = k2, *id = i2;
    for (unsigned shift = 0; shift < bits; shift += RADIX_BITS) {
        memset(hist, 0, (size_t)nt * RADIX_BUCKETS * sizeof(uint64_t));
        #pragma omp parallel num_threads(nt)
        {
#ifdef _OPENMP
            int t = omp_get_thread_num();
#else
            int t = 0;
#endif
            uint64_t lo = n * (uint64_t)t / nt, hi = n * (uint64_t)(t + 1) / nt;
            uint64_t *h = hist + (size_t)t * RADIX_BUCKETS;
            for (uint64_t i = lo; i < hi; ++i) h[(ks[i] >> shift) & (RADIX_BUCKETS - 1)]++;
        }
        /* a pass where every key has the same digit changes nothing */
        int trivial = 0;
        for (unsigned b = 0; b < RADIX_BUCKETS; ++b) {
            uint64_t tot = 0;
            for (int t = 0; t < nt; ++t) tot += hist[(size_t)t * RADIX_BUCKETS + b];
            if (tot == n) { trivial = 1; break; }
            if (tot) break;
        }
        if (trivial) continue;
        uint64_t run = 0;
        for (unsigned b = 0; b < RADIX_BUCKETS; ++b)
            for (int t = 0; t < nt; ++t) {
                uint64_t c = hist[(size_t)t * RADIX_BUCKETS + b];
                hist[(size_t)t * RADIX_BUCKETS + b] = run;
                run += c;
            }
        #pragma omp parallel num_threads(nt)
        {
#ifdef _OPENMP
            int t = omp_get_thread_num();
#else
            int t = 0;
#endif
            uint64_t lo = n * (uint64_t)t / nt, hi = n * (uint64_t)(t + 1) / nt;
            uint64_t *h = hist + (size_t)t * RADIX_BUCKETS;
            for (uint64_t i = lo; i < hi; ++i) {
                uint64_t pos = h[(ks[i] >> shift) & (RADIX_BUCKETS - 1)]++;
                kd[pos] = ks[i];
                id[pos] = is[i];
            }
        }
        uint64_t *tk = ks; ks = kd; kd = tk;
        uint64_t *ti = is; is = id; id = ti;
    }
    if (ks != keys) {
        memcpy(keys, ks, n * sizeof(uint64_t));
        memcpy(idx, is, n * sizeof(uint64_t));
    }
    free(k2); free(i2); free(hist);
    return 0;
}

void sptol_parallel_memcpy(void *dst, const void *src, size_t n)
{
    const size_t piece = (size_t)1 << 20;
    const int64_t np = (int64_t)((n + piece - 1) / piece);
    if (np <= 1) {
        if (n) memcpy(dst, src, n);
        return;
    }
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < np; ++i) {
        const size_t off = (size_t)i * piece;
        const size_t len = n - off < piece ? n - off : piece;
        memcpy((char *)dst + off, (const char *)src + off, len);
    }
}

int sptol_parallel_pread(int fd, void *dst, size_t n, uint64_t off)
{
    const size_t piece = (size_t)1 << 20;
    const int64_t np = (int64_t)((n + piece - 1) / piece);
    int err = 0;
    #pragma omp parallel for schedule(static) reduction(max : err) if (np > 1)
    for (int64_t i = 0; i < np; ++i) {
        size_t done = (size_t)i * piece;
        const size_t end = n - done < piece ? n : done + piece;
        while (done < end) {
            const ssize_t got = pread(fd, (char *)dst + done, end - done, (off_t)(off + done));
            if (got <= 0) {
                err = got < 0 && errno ? errno : EIO;
                break;
            }
            done += (size_t)got;
        }
    }
    return err;
}

int sptol_permute_coo(sptSparseTensor *tsr, uint64_t begin, const uint64_t *perm, uint64_t n)
{
    if (n == 0) return 0;
    void *tmp = malloc(n * sizeof(uint32_t));
    SPTOL_CHECK_OS(!tmp, "SpTns Permute");
    for (sptIndex m = 0; m < tsr->nmodes; ++m) {
        uint32_t *src = tsr->inds[m].data, *t = (uint32_t *)tmp;
        #pragma omp parallel for schedule(static)
        for (uint64_t i = 0; i < n; ++i) t[i] = src[perm[i]];
        memcpy(src + begin, t, n * sizeof(uint32_t));
    }
    {
        float *src = tsr->values.data, *t = (float *)tmp;
        #pragma omp parallel for schedule(static)
        for (uint64_t i = 0; i < n; ++i) t[i] = src[perm[i]];
        memcpy(src + begin, t, n * sizeof(float));
    }
    free(tmp);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Sort keys.  A total order is applied as a sequence of stable radix passes */
/* over 64-bit key chunks, least significant chunk first.                    */
/* ------------------------------------------------------------------------ */
typedef enum { KEY_LEX, KEY_ROWBLOCK, KEY_MORTON_LO, KEY_MORTON_HI } key_kind;

typedef struct {
    key_kind kind;
    sptIndex m_lo, m_hi;      /* KEY_LEX / KEY_ROWBLOCK: modes [m_lo, m_hi) packed */
    unsigned shift;           /* KEY_ROWBLOCK: right shift of each index            */
    unsigned mbits;           /* KEY_MORTON_*: low bits of each index interleaved   */
    unsigned bits[8];         /* per-mode widths (packing)                          */
    unsigned total;           /* total key bits                                     */
} key_chunk;

/* Morton key of the low `mbits` bits of each index, mode 0 the most
 * significant bit of every group of nmodes bits.  Returned as two 64-bit
 * halves of a (nmodes*mbits)-bit number. */
static inline void morton_key(const sptSparseTensor *t, uint64_t x, unsigned mbits,
                              uint64_t *lo, uint64_t *hi)
{
    unsigned N = t->nmodes;
    unsigned __int128 k = 0;
    for (unsigned b = mbits; b-- > 0;)
        for (unsigned m = 0; m < N; ++m)
            k = (k << 1) | ((t->inds[m].data[x] >> b) & 1u);
    *lo = (uint64_t)k;
    *hi = (uint64_t)(k >> 64);
}

static inline uint64_t chunk_key(const sptSparseTensor *t, const key_chunk *c, uint64_t x)
{
    uint64_t k = 0, lo, hi;
    switch (c->kind) {
    case KEY_LEX:
    case KEY_ROWBLOCK:
        for (sptIndex m = c->m_lo; m < c->m_hi; ++m)
            k = (k << c->bits[m - c->m_lo]) | (uint64_t)(t->inds[m].data[x] >> c->shift);
        return k;
    case KEY_MORTON_LO:
        morton_key(t, x, c->mbits, &lo, &hi);
        return lo;
    case KEY_MORTON_HI:
        morton_key(t, x, c->mbits, &lo, &hi);
        return hi;
    }
    return 0;
}

/* Split a lexicographic order over modes into 64-bit chunks (most
 * significant mode first inside a chunk).  Returns the chunk count, listed
 * least significant first. */
static int lex_chunks(const sptSparseTensor *t, unsigned shift, key_kind kind, key_chunk *out)
{
    int nc = 0;
    sptIndex hi = t->nmodes;
    while (hi > 0) {
        key_chunk c;
        memset(&c, 0, sizeof(c));
        c.kind = kind;
        c.shift = shift;
        c.m_hi = hi;
        sptIndex lo = hi;
        unsigned tot = 0;
        while (lo > 0) {
            unsigned w = sptol_bits_for(((uint64_t)t->ndims[lo - 1] >> shift) + 1);
            if (w == 0) w = 1;
            if (tot + w > 64) break;
            tot += w;
            --lo;
        }
        if (lo == hi) lo = hi - 1, tot = 32; /* cannot happen for 32-bit indices */
        c.m_lo = lo;
        for (sptIndex m = lo; m < hi; ++m) {
            unsigned w = sptol_bits_for(((uint64_t)t->ndims[m] >> shift) + 1);
            c.bits[m - lo] = w ? w : 1;
        }
        c.total = tot;
        out[nc++] = c;
        hi = lo;
    }
    return nc;
}

/* Sort tensor rows [begin, end) by the chunk list (least significant first). */
static int sort_by_chunks(sptSparseTensor *t, uint64_t begin, uint64_t end,
                          const key_chunk *chunks, int nchunks, int tk)
{
    uint64_t n = end - begin;
    if (n < 2) return 0;
    uint64_t *keys = (uint64_t *)malloc(n * sizeof(uint64_t));
    uint64_t *perm = (uint64_t *)malloc(n * sizeof(uint64_t));
    if (!keys || !perm) { free(keys); free(perm); SPTOL_CHECK_OS(1, "SpTns Sort"); }
    for (uint64_t i = 0; i < n; ++i) perm[i] = begin + i;
    for (int c = 0; c < nchunks; ++c) {
        #pragma omp parallel for schedule(static) num_threads(max_threads(tk))
        for (uint64_t i = 0; i < n; ++i) keys[i] = chunk_key(t, &chunks[c], perm[i]);
        int r = sptol_radix_sort_u64(keys, perm, n, chunks[c].total, tk);
        if (r) { free(keys); free(perm); SPTOL_CHECK(r, "SpTns Sort", "radix sort"); }
    }
    free(keys);
    int r = sptol_permute_coo(t, begin, perm, n);
    free(perm);
    return r;
}

static int morton_bits_for(const sptSparseTensor *t)
{
    /* The reference's 3-D key keeps bits 0-23 of each index, the 4-D key all
     * 32 (sort.c:765-795, 835-985). */
    return t->nmodes == 3 ? 24 : 32;
}

static int morton_chunks(unsigned nmodes, unsigned mbits, key_chunk *out)
{
    unsigned tot = nmodes * mbits;
    int nc = 0;
    key_chunk c;
    memset(&c, 0, sizeof(c));
    c.mbits = mbits;
    c.kind = KEY_MORTON_LO;
    c.total = tot > 64 ? 64 : tot;
    out[nc++] = c;
    if (tot > 64) {
        c.kind = KEY_MORTON_HI;
        c.total = tot - 64;
        out[nc++] = c;
    }
    return nc;
}

static void set_identity_sortorder(sptSparseTensor *tsr, int *needsort)
{
    for (sptIndex m = 0; m < tsr->nmodes; ++m)
        if (tsr->sortorder[m] != m) { tsr->sortorder[m] = m; *needsort = 1; }
}

void sptSparseTensorSortIndex(sptSparseTensor *tsr, int force, int tk)
{
    int needsort = 0;
    set_identity_sortorder(tsr, &needsort);
    if (!(needsort || force)) return;
    key_chunk ch[16];
    int nc = lex_chunks(tsr, 0, KEY_LEX, ch);
    sort_by_chunks(tsr, 0, tsr->nnz, ch, nc, tk);
}

/* Lexicographic sort over the modes in `order` (most significant first):
 * the chunks are built over a view of the tensor whose i-th mode is mode
 * order[i], and the view's arrays are the tensor's, permuted in place. */
static int sort_lex_order(sptSparseTensor *tsr, const sptIndex *order, int tk)
{
    const sptIndex N = tsr->nmodes;
    sptSparseTensor v = *tsr;
    v.ndims = (sptIndex *)malloc((size_t)N * sizeof(sptIndex));
    v.inds = (sptIndexVector *)malloc((size_t)N * sizeof(sptIndexVector));
    key_chunk *ch = (key_chunk *)malloc(((size_t)N + 1) * sizeof(key_chunk));
    if (!v.ndims || !v.inds || !ch) {
        free(v.ndims); free(v.inds); free(ch);
        SPTOL_CHECK_OS(1, "SpTns Sort");
    }
    for (sptIndex m = 0; m < N; ++m) {
        v.ndims[m] = tsr->ndims[order[m]];
        v.inds[m] = tsr->inds[order[m]];
    }
    int nc = lex_chunks(&v, 0, KEY_LEX, ch);
    int r = sort_by_chunks(&v, 0, v.nnz, ch, nc, tk);
    free(v.ndims); free(v.inds); free(ch);
    return r;
}

void sptSparseTensorSortIndexCustomOrder(sptSparseTensor *tsr, sptIndex const *mode_order, int force, int tk)
{
    const sptIndex N = tsr->nmodes;
    if (!force && memcmp(tsr->sortorder, mode_order, (size_t)N * sizeof(sptIndex)) == 0) return;
    if (sort_lex_order(tsr, mode_order, tk)) return;
    memcpy(tsr->sortorder, mode_order, (size_t)N * sizeof(sptIndex));
}

void sptSparseTensorSortIndexAtMode(sptSparseTensor *tsr, sptIndex const mode, int force, int tk)
{
    const sptIndex N = tsr->nmodes;
    int needsort = 0;
    /* sortorder: the other modes in increasing order, then `mode` */
    for (sptIndex m = 0, k = 0; m < N; ++m) {
        if (m == mode) continue;
        if (tsr->sortorder[k] != m) { tsr->sortorder[k] = m; needsort = 1; }
        ++k;
    }
    if (tsr->sortorder[N - 1] != mode) { tsr->sortorder[N - 1] = mode; needsort = 1; }
    if (!(needsort || force)) return;
    sort_lex_order(tsr, tsr->sortorder, tk);
}

int sptSparseTensorMixedOrder(sptSparseTensor *tsr, const sptElementIndex sb_bits, const sptElementIndex sk_bits,
                              int const tk)
{
    /* superblocks row major, then Z-Morton inside each superblock: the
     * order sptSparseTensorToHiCOO leaves the tensor in */
    sptSparseTensorSortIndexRowBlock(tsr, 1, 0, tsr->nnz, sk_bits, tk);
    sptNnzIndexVector kptr;
    SPTOL_CHECK(sptNewNnzIndexVector(&kptr, 0, 0), "HiSpTns New", NULL);
    int r = sptSetKernelPointers(&kptr, tsr, sk_bits);
    if (r) { sptFreeNnzIndexVector(&kptr); return r; }
    for (sptNnzIndex k = 0; k + 1 < kptr.len; ++k)
        sptSparseTensorSortIndexMorton(tsr, 1, kptr.data[k], kptr.data[k + 1], sb_bits, tk);
    sptFreeNnzIndexVector(&kptr);
    return 0;
}

void sptGetRandomShuffleElements(sptSparseTensor *tsr)
{
    /* The reference's order: a pass of swaps whose targets come from C's
     * rand() reseeded with z + 1 at every step (sort.c:371-381), so the same
     * C library yields the same permutation. */
    const sptNnzIndex nnz = tsr->nnz;
    const sptIndex N = tsr->nmodes;
    for (sptNnzIndex z = 0; z < nnz; ++z) {
        srand((unsigned)(z + 1));
        const sptValue u = (sptValue)rand() / (sptValue)RAND_MAX;
        const sptNnzIndex t = (sptNnzIndex)(u * nnz) % nnz;
        for (sptIndex m = 0; m < N; ++m) {
            sptIndex *d = tsr->inds[m].data, x = d[z];
            d[z] = d[t];
            d[t] = x;
        }
        sptValue *v = tsr->values.data, x = v[z];
        v[z] = v[t];
        v[t] = x;
    }
}

void sptSparseTensorSortIndexRowBlock(sptSparseTensor *tsr, int force, const sptNnzIndex begin,
                                      const sptNnzIndex end, const sptElementIndex sk_bits, int const tk)
{
    int needsort = 0;
    set_identity_sortorder(tsr, &needsort);
    if (!(needsort || force)) return;
    key_chunk ch[16];
    int nc = lex_chunks(tsr, sk_bits, KEY_ROWBLOCK, ch);
    sort_by_chunks(tsr, begin, end, ch, nc, tk);
}

void sptSparseTensorSortIndexMorton(sptSparseTensor *tsr, int force, const sptNnzIndex begin,
                                    const sptNnzIndex end, const sptElementIndex sb_bits, int tk)
{
    (void)sb_bits;
    int needsort = 0;
    set_identity_sortorder(tsr, &needsort);
    if (!(needsort || force)) return;
    if (tsr->nmodes != 3 && tsr->nmodes != 4) {
        /* the reference only orders 3-D and 4-D tensors (sort.c:430-450) */
        fprintf(stdout, "No support for more than 4th-order tensors yet.\n");
        return;
    }
    key_chunk ch[2];
    int nc = morton_chunks(tsr->nmodes, morton_bits_for(tsr), ch);
    sort_by_chunks(tsr, begin, end, ch, nc, tk);
}

/* ------------------------------------------------------------------------ */
/* COO tensors                                                              */
/* ------------------------------------------------------------------------ */
int sptNewSparseTensor(sptSparseTensor *tsr, sptIndex nmodes, const sptIndex ndims[])
{
    tsr->nmodes = nmodes;
    tsr->sortorder = (sptIndex *)malloc(nmodes * sizeof(sptIndex));
    SPTOL_CHECK_OS(!tsr->sortorder, "SpTns New");
    for (sptIndex i = 0; i < nmodes; ++i) tsr->sortorder[i] = i;
    tsr->ndims = (sptIndex *)malloc(nmodes * sizeof(sptIndex));
    SPTOL_CHECK_OS(!tsr->ndims, "SpTns New");
    memcpy(tsr->ndims, ndims, nmodes * sizeof(sptIndex));
    tsr->nnz = 0;
    tsr->inds = (sptIndexVector *)malloc(nmodes * sizeof(sptIndexVector));
    SPTOL_CHECK_OS(!tsr->inds, "SpTns New");
    for (sptIndex i = 0; i < nmodes; ++i) {
        int r = sptNewIndexVector(&tsr->inds[i], 0, 0);
        if (r) return r;
    }
    return sptNewValueVector(&tsr->values, 0, 0);
}

int sptCopySparseTensor(sptSparseTensor *dest, const sptSparseTensor *src, int const nt)
{
    (void)nt;
    dest->nmodes = src->nmodes;
    dest->sortorder = (sptIndex *)malloc(src->nmodes * sizeof(sptIndex));
    dest->ndims = (sptIndex *)malloc(src->nmodes * sizeof(sptIndex));
    dest->inds = (sptIndexVector *)malloc(src->nmodes * sizeof(sptIndexVector));
    SPTOL_CHECK_OS(!dest->sortorder || !dest->ndims || !dest->inds, "SpTns Copy");
    memcpy(dest->sortorder, src->sortorder, src->nmodes * sizeof(sptIndex));
    memcpy(dest->ndims, src->ndims, src->nmodes * sizeof(sptIndex));
    dest->nnz = src->nnz;
    for (sptIndex m = 0; m < src->nmodes; ++m) {
        int r = sptNewIndexVector(&dest->inds[m], src->inds[m].len, src->inds[m].len);
        if (r) return r;
        memcpy(dest->inds[m].data, src->inds[m].data, src->inds[m].len * sizeof(sptIndex));
    }
    int r = sptNewValueVector(&dest->values, src->values.len, src->values.len);
    if (r) return r;
    memcpy(dest->values.data, src->values.data, src->values.len * sizeof(sptValue));
    return 0;
}

void sptFreeSparseTensor(sptSparseTensor *tsr)
{
    for (sptIndex m = 0; m < tsr->nmodes; ++m) sptFreeIndexVector(&tsr->inds[m]);
    free(tsr->sortorder);
    free(tsr->ndims);
    free(tsr->inds);
    sptFreeValueVector(&tsr->values);
    tsr->sortorder = NULL;
    tsr->ndims = NULL;
    tsr->inds = NULL;
    tsr->nmodes = 0;
    tsr->nnz = 0;
}

double SparseTensorFrobeniusNormSquared(sptSparseTensor const * const spten)
{
    double norm = 0;
    const sptValue *v = spten->values.data;
    for (sptNnzIndex n = 0; n < spten->nnz; ++n) norm += v[n] * v[n];
    return norm;
}

/* Move every exact zero to the tail by swapping in the last element, then
 * truncate: the order rule of src/sptensor/collect.c:29-49. */
static void collect_zeros(sptSparseTensor *tsr)
{
    sptNnzIndex i = 0, nnz = tsr->nnz;
    while (i < nnz) {
        if (tsr->values.data[i] == 0) {
            for (sptIndex m = 0; m < tsr->nmodes; ++m) tsr->inds[m].data[i] = tsr->inds[m].data[nnz - 1];
            tsr->values.data[i] = tsr->values.data[nnz - 1];
            --nnz;
        } else {
            ++i;
        }
    }
    tsr->nnz = nnz;
    for (sptIndex m = 0; m < tsr->nmodes; ++m) tsr->inds[m].len = nnz;
    tsr->values.len = nnz;
}

/* Text reader for the .tns format of src/sptensor/load.c:31-86: nmodes,
 * ndims, then one "i_0 ... i_{N-1} value" record per nonzero.  Indices are
 * shifted by start_index; values are read as double and stored as float;
 * exact zeros are dropped afterwards.  sortorder is zero-filled as in the
 * reference (load.c:39). */
int sptLoadSparseTensor(sptSparseTensor *tsr, sptIndex start_index, FILE *fp)
{
    unsigned nm = 0;
    memset(tsr, 0, sizeof(*tsr));
    SPTOL_CHECK_OS(fscanf(fp, "%u", &nm) != 1, "SpTns Load");
    tsr->nmodes = nm;
    tsr->sortorder = (sptIndex *)calloc(nm, sizeof(sptIndex));
    tsr->ndims = (sptIndex *)malloc(nm * sizeof(sptIndex));
    tsr->inds = (sptIndexVector *)malloc(nm * sizeof(sptIndexVector));
    SPTOL_CHECK_OS(!tsr->sortorder || !tsr->ndims || !tsr->inds, "SpTns Load");
    for (sptIndex m = 0; m < nm; ++m) SPTOL_CHECK_OS(fscanf(fp, "%u", &tsr->ndims[m]) != 1, "SpTns Load");
    for (sptIndex m = 0; m < nm; ++m) {
        int r = sptNewIndexVector(&tsr->inds[m], 0, 0);
        if (r) return r;
    }
    int r = sptNewValueVector(&tsr->values, 0, 0);
    if (r) return r;
    for (;;) {
        sptIndex m;
        unsigned idx;
        for (m = 0; m < nm; ++m) {
            if (fscanf(fp, "%u", &idx) != 1) break;
            if (idx < start_index) SPTOL_CHECK(SPTERR_VALUE_ERROR, "SpTns Load", "index < start_index");
            r = sptAppendIndexVector(&tsr->inds[m], idx - start_index);
            if (r) return r;
        }
        if (m < nm) {
            for (sptIndex q = 0; q < m; ++q) tsr->inds[q].len--;  /* drop a partial record */
            break;
        }
        double value;
        if (fscanf(fp, "%lf", &value) != 1) {
            for (sptIndex q = 0; q < nm; ++q) tsr->inds[q].len--;
            break;
        }
        r = sptAppendValueVector(&tsr->values, (sptValue)value);
        if (r) return r;
        ++tsr->nnz;
    }
    for (sptIndex m = 0; m < nm; ++m) tsr->inds[m].len = tsr->nnz;
    collect_zeros(tsr);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Ingest: duplicate merging and the binary tensor file                     */
/* ------------------------------------------------------------------------ */

/* The reference keeps duplicated coordinates as separate nonzeros
 * (load.c:31-86 only drops exact zeros).  Here they are merged: a stable
 * lexicographic sort (sptSparseTensorSortIndex), then each run of equal
 * coordinates becomes one nonzero whose value is the FP64 sum of the run
 * in input order, rounded once.  The tensor ends sorted. */
int sptolSparseTensorMergeDuplicates(sptSparseTensor *tsr, int tk)
{
    if (tsr->nnz < 2) return 0;
    sptSparseTensorSortIndex(tsr, 1, tk);
    const sptIndex N = tsr->nmodes;
    sptNnzIndex w = 0;
    for (sptNnzIndex r = 0; r < tsr->nnz;) {
        sptNnzIndex e = r + 1;
        for (; e < tsr->nnz; ++e) {
            sptIndex m = 0;
            while (m < N && tsr->inds[m].data[e] == tsr->inds[m].data[r]) ++m;
            if (m < N) break;
        }
        double acc = 0.0;
        for (sptNnzIndex k = r; k < e; ++k) acc += (double)tsr->values.data[k];
        for (sptIndex m = 0; m < N; ++m) tsr->inds[m].data[w] = tsr->inds[m].data[r];
        tsr->values.data[w] = (sptValue)acc;
        ++w;
        r = e;
    }
    tsr->nnz = w;
    for (sptIndex m = 0; m < N; ++m) tsr->inds[m].len = w;
    tsr->values.len = w;
    return 0;
}

static const char kBinMagic[8] = {'S', 'P', 'T', 'O', 'L', 'B', 'T', '1'};

/* Header: 8-byte magic "SPTOLBT1", u32 version (1), u32 nmodes, u64 nnz,
 * u32 ndims[nmodes], zero padding to a multiple of 8 bytes; then
 * inds[0..nmodes-1] (u32 x nnz each) and the values (f32 x nnz), all
 * little-endian.  One sequential read or write per array. */
static size_t binary_header_bytes(uint32_t nmodes) { return (24 + 4 * (size_t)nmodes + 7) / 8 * 8; }

int sptolSaveSparseTensorBinary(const sptSparseTensor *tsr, const char *path)
{
    if (tsr->nmodes < 1 || tsr->nmodes > 8) SPTOL_CHECK(SPTERR_VALUE_ERROR, "SpTns Save", "nmodes must be in [1, 8]");
    FILE *fp = fopen(path, "wb");
    SPTOL_CHECK_OS(!fp, "SpTns Save");
    unsigned char hdr[64];
    memset(hdr, 0, sizeof(hdr));
    const uint32_t version = 1, nm = tsr->nmodes;
    const uint64_t nnz = tsr->nnz;
    memcpy(hdr, kBinMagic, 8);
    memcpy(hdr + 8, &version, 4);
    memcpy(hdr + 12, &nm, 4);
    memcpy(hdr + 16, &nnz, 8);
    for (uint32_t m = 0; m < nm; ++m) memcpy(hdr + 24 + 4 * m, &tsr->ndims[m], 4);
    int ok = fwrite(hdr, 1, binary_header_bytes(nm), fp) == binary_header_bytes(nm);
    for (uint32_t m = 0; m < nm && ok; ++m)
        ok = fwrite(tsr->inds[m].data, 4, nnz, fp) == nnz;
    if (ok) ok = fwrite(tsr->values.data, 4, nnz, fp) == nnz;
    if (fclose(fp) != 0) ok = 0;
    SPTOL_CHECK_OS(!ok, "SpTns Save");
    return 0;
}

int sptol_read_binary_header(FILE *fp, uint32_t *nmodes, uint64_t *nnz, uint32_t *ndims)
{
    unsigned char hdr[64];
    if (fread(hdr, 1, 24, fp) != 24 || memcmp(hdr, kBinMagic, 8) != 0)
        SPTOL_CHECK(SPTERR_VALUE_ERROR, "SpTns Load Binary", "not a binary tensor file");
    uint32_t version, nm;
    memcpy(&version, hdr + 8, 4);
    memcpy(&nm, hdr + 12, 4);
    memcpy(nnz, hdr + 16, 8);
    if (version != 1) SPTOL_CHECK(SPTERR_VALUE_ERROR, "SpTns Load Binary", "unknown format version");
    if (nm < 1 || nm > 8) SPTOL_CHECK(SPTERR_VALUE_ERROR, "SpTns Load Binary", "nmodes must be in [1, 8]");
    const size_t rest = binary_header_bytes(nm) - 24;
    if (fread(hdr + 24, 1, rest, fp) != rest) SPTOL_CHECK(SPTERR_VALUE_ERROR, "SpTns Load Binary", "truncated header");
    for (uint32_t m = 0; m < nm; ++m) memcpy(&ndims[m], hdr + 24 + 4 * m, 4);
    *nmodes = nm;
    /* the file must hold every array */
    long here = ftell(fp);
    if (here < 0 || fseek(fp, 0, SEEK_END) != 0) SPTOL_CHECK_OS(1, "SpTns Load Binary");
    long end = ftell(fp);
    if (end < 0 || fseek(fp, here, SEEK_SET) != 0) SPTOL_CHECK_OS(1, "SpTns Load Binary");
    /* by division: a crafted nnz must not wrap the product of sizes */
    if (end < here || *nnz > (uint64_t)(end - here) / (4ull * (nm + 1ull)))
        SPTOL_CHECK(SPTERR_VALUE_ERROR, "SpTns Load Binary", "file shorter than its arrays");
    return 0;
}

int sptolLoadSparseTensorBinary(sptSparseTensor *tsr, const char *path)
{
    memset(tsr, 0, sizeof(*tsr));
    FILE *fp = fopen(path, "rb");
    SPTOL_CHECK_OS(!fp, "SpTns Load Binary");
    uint32_t nm = 0, nd[8];
    uint64_t nnz = 0;
    int r = sptol_read_binary_header(fp, &nm, &nnz, nd);
    /* the arrays follow the header back to back; each is read in 1 MiB
     * pieces on the OpenMP threads (which also first-touch its pages) */
    const long data0 = r ? 0 : ftell(fp);
    if (!r && data0 < 0) r = SPTERR_OS_ERROR + (errno ? errno : EIO);
    const int fd = fileno(fp);
    if (!r) r = sptNewSparseTensor(tsr, nm, nd);
    for (uint32_t m = 0; m < nm && !r; ++m) {
        sptFreeIndexVector(&tsr->inds[m]);
        r = sptNewIndexVector(&tsr->inds[m], nnz, nnz);
        if (!r) {
            const int en = sptol_parallel_pread(fd, tsr->inds[m].data, nnz * 4, (uint64_t)data0 + (uint64_t)m * nnz * 4);
            if (en) r = SPTERR_OS_ERROR + en;
        }
    }
    if (!r) {
        sptFreeValueVector(&tsr->values);
        r = sptNewValueVector(&tsr->values, nnz, nnz);
    }
    if (!r) {
        const int en = sptol_parallel_pread(fd, tsr->values.data, nnz * 4, (uint64_t)data0 + (uint64_t)nm * nnz * 4);
        if (en) r = SPTERR_OS_ERROR + en;
    }
    fclose(fp);
    if (r) {
        sptol_complain("SpTns Load Binary", r, __FILE__, __LINE__, "read failed");
        return r;
    }
    tsr->nnz = nnz;
    int bad = 0;
    for (uint32_t m = 0; m < nm && !bad; ++m) {
        const uint32_t *p = tsr->inds[m].data, lim = nd[m];
        #pragma omp parallel for schedule(static) reduction(| : bad)
        for (int64_t z = 0; z < (int64_t)nnz; ++z) bad |= p[z] >= lim;
    }
    if (bad) SPTOL_CHECK(SPTERR_VALUE_ERROR, "SpTns Load Binary", "index >= ndims");
    return 0;
}

int sptDumpSparseTensor(const sptSparseTensor *tsr, sptIndex start_index, FILE *fp)
{
    int iores = fprintf(fp, "%" PRIu32 "\n", tsr->nmodes);
    SPTOL_CHECK_OS(iores < 0, "SpTns Dump");
    for (sptIndex m = 0; m < tsr->nmodes; ++m) fprintf(fp, m ? " %" PRIu32 : "%" PRIu32, tsr->ndims[m]);
    fputs("\n", fp);
    for (sptNnzIndex i = 0; i < tsr->nnz; ++i) {
        for (sptIndex m = 0; m < tsr->nmodes; ++m) {
            iores = fprintf(fp, "%" PRIu32 "\t", tsr->inds[m].data[i] + start_index);
            SPTOL_CHECK_OS(iores < 0, "SpTns Dump");
        }
        iores = fprintf(fp, "%f\n", (double)tsr->values.data[i]);
        SPTOL_CHECK_OS(iores < 0, "SpTns Dump");
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* HiCOO                                                                    */
/* ------------------------------------------------------------------------ */
static sptIndex kernel_rows(sptIndex ndim, unsigned sk_bits)
{
    uint64_t sk = 1ull << sk_bits;
    return (sptIndex)(((uint64_t)ndim + sk - 1) / sk);
}

int sptNewSparseTensorHiCOO(sptSparseTensorHiCOO *hitsr, const sptIndex nmodes, const sptIndex ndims[],
                            const sptNnzIndex nnz, const sptElementIndex sb_bits,
                            const sptElementIndex sk_bits, const sptElementIndex sc_bits)
{
    memset(hitsr, 0, sizeof(*hitsr));
    hitsr->nmodes = nmodes;
    hitsr->sortorder = (sptIndex *)malloc(nmodes * sizeof(sptIndex));
    hitsr->ndims = (sptIndex *)malloc(nmodes * sizeof(sptIndex));
    SPTOL_CHECK_OS(!hitsr->sortorder || !hitsr->ndims, "HiSpTns New");
    for (sptIndex m = 0; m < nmodes; ++m) hitsr->sortorder[m] = m;
    memcpy(hitsr->ndims, ndims, nmodes * sizeof(sptIndex));
    hitsr->nnz = nnz;
    hitsr->sb_bits = sb_bits;
    hitsr->sk_bits = sk_bits;
    hitsr->sc_bits = sc_bits;
    hitsr->kschr = (sptIndexVector **)malloc(nmodes * sizeof(sptIndexVector *));
    SPTOL_CHECK_OS(!hitsr->kschr, "HiSpTns New");
    for (sptIndex m = 0; m < nmodes; ++m) {
        sptIndex kd = kernel_rows(ndims[m], sk_bits);
        hitsr->kschr[m] = (sptIndexVector *)malloc((kd ? kd : 1) * sizeof(sptIndexVector));
        SPTOL_CHECK_OS(!hitsr->kschr[m], "HiSpTns New");
        for (sptIndex i = 0; i < kd; ++i) {
            int r = sptNewIndexVector(&hitsr->kschr[m][i], 0, 0);
            if (r) return r;
        }
    }
    hitsr->nkiters = (sptIndex *)calloc(nmodes, sizeof(sptIndex));
    SPTOL_CHECK_OS(!hitsr->nkiters, "HiSpTns New");
    int r;
    if ((r = sptNewNnzIndexVector(&hitsr->kptr, 0, 0))) return r;
    if ((r = sptNewNnzIndexVector(&hitsr->cptr, 0, 0))) return r;
    if ((r = sptNewNnzIndexVector(&hitsr->knnzs, 0, 0))) return r;
    if ((r = sptNewNnzIndexVector(&hitsr->bptr, 0, 0))) return r;
    hitsr->binds = (sptBlockIndexVector *)malloc(nmodes * sizeof(sptBlockIndexVector));
    hitsr->einds = (sptElementIndexVector *)malloc(nmodes * sizeof(sptElementIndexVector));
    SPTOL_CHECK_OS(!hitsr->binds || !hitsr->einds, "HiSpTns New");
    for (sptIndex m = 0; m < nmodes; ++m) {
        if ((r = sptNewBlockIndexVector(&hitsr->binds[m], 0, 0))) return r;
        if ((r = sptNewElementIndexVector(&hitsr->einds[m], 0, 0))) return r;
    }
    return sptNewValueVector(&hitsr->values, 0, 0);
}

void sptFreeSparseTensorHiCOO(sptSparseTensorHiCOO *hitsr)
{
    sptIndex N = hitsr->nmodes;
    if (hitsr->kschr) {
        for (sptIndex m = 0; m < N; ++m) {
            sptIndex kd = kernel_rows(hitsr->ndims[m], hitsr->sk_bits);
            for (sptIndex i = 0; i < kd; ++i) sptFreeIndexVector(&hitsr->kschr[m][i]);
            free(hitsr->kschr[m]);
        }
        free(hitsr->kschr);
    }
    free(hitsr->nkiters);
    sptFreeNnzIndexVector(&hitsr->kptr);
    sptFreeNnzIndexVector(&hitsr->cptr);
    sptFreeNnzIndexVector(&hitsr->knnzs);
    sptFreeNnzIndexVector(&hitsr->bptr);
    if (hitsr->binds) for (sptIndex m = 0; m < N; ++m) sptFreeBlockIndexVector(&hitsr->binds[m]);
    if (hitsr->einds) for (sptIndex m = 0; m < N; ++m) sptFreeElementIndexVector(&hitsr->einds[m]);
    free(hitsr->binds);
    free(hitsr->einds);
    sptFreeValueVector(&hitsr->values);
    free(hitsr->sortorder);
    free(hitsr->ndims);
    memset(hitsr, 0, sizeof(*hitsr));
}

double sptSparseTensorFrobeniusNormSquaredHiCOO(sptSparseTensorHiCOO const * const hitsr)
{
    double norm = 0;
    for (sptNnzIndex n = 0; n < hitsr->nnz; ++n) norm += hitsr->values.data[n] * hitsr->values.data[n];
    return norm;
}

/* Superblock pointers as nonzero offsets of a row-block-sorted tensor
 * (convert.c:237-288): appends 0, one offset per superblock change, nnz. */
int sptSetKernelPointers(sptNnzIndexVector *kptr, sptSparseTensor *tsr, const sptElementIndex sk_bits)
{
    int r = sptAppendNnzIndexVector(kptr, 0);
    if (r) return r;
    for (sptNnzIndex z = 1; z < tsr->nnz; ++z) {
        for (sptIndex m = 0; m < tsr->nmodes; ++m)
            if ((tsr->inds[m].data[z] >> sk_bits) != (tsr->inds[m].data[z - 1] >> sk_bits)) {
                r = sptAppendNnzIndexVector(kptr, z);
                if (r) return r;
                break;
            }
    }
    return sptAppendNnzIndexVector(kptr, tsr->nnz);
}

static int exact_vec_nnz(sptNnzIndexVector *v, sptNnzIndex len)
{
    free(v->data);
    v->data = (sptNnzIndex *)malloc((len ? len : 1) * sizeof(sptNnzIndex));
    SPTOL_CHECK_OS(!v->data, "HiSpTns Convert");
    v->len = len;
    v->cap = len ? len : 1;
    return 0;
}

/* COO -> HiCOO (convert.c:300-497).  The tensor is permuted in place into
 * superblock row-major order with Z-Morton order inside each superblock;
 * then a single pass emits blocks (a new block whenever some i_m >> sb_bits
 * changes), element offsets, superblock pointers rewritten to block offsets,
 * the chunk pointers and the per-mode superblock schedule. */
int sptSparseTensorToHiCOO(sptSparseTensorHiCOO *hitsr, sptNnzIndex *max_nnzb, sptSparseTensor *tsr,
                           const sptElementIndex sb_bits, const sptElementIndex sk_bits, int const tk)
{
    const sptElementIndex sc_bits = 14;
    if (sk_bits < sb_bits) SPTOL_CHECK(SPTERR_VALUE_ERROR, "HiSpTns Convert", "sk_bits < sb_bits");
    if (sb_bits > 8) SPTOL_CHECK(SPTERR_VALUE_ERROR, "HiSpTns Convert", "sb_bits > 8 does not fit 8-bit element offsets");
    if (sk_bits > 31) SPTOL_CHECK(SPTERR_VALUE_ERROR, "HiSpTns Convert", "sk_bits > 31");
    if (tsr->nnz == 0) SPTOL_CHECK(SPTERR_VALUE_ERROR, "HiSpTns Convert", "empty tensor");
    const sptIndex N = tsr->nmodes;
    const sptNnzIndex nnz = tsr->nnz;
    if (N < 2 || N > 8) SPTOL_CHECK(SPTERR_VALUE_ERROR, "HiSpTns Convert", "nmodes must be in [2, 8]");
    /* The reference trusts its input; an index >= ndims would index past the
     * superblock schedule, so reject it here. */
    {
        int bad = 0;
        for (sptIndex m = 0; m < N; ++m) {
            const uint32_t *a = tsr->inds[m].data;
            const uint32_t lim = tsr->ndims[m];
            #pragma omp parallel for schedule(static) reduction(|:bad) num_threads(max_threads(tk))
            for (sptNnzIndex z = 0; z < nnz; ++z) bad |= a[z] >= lim;
        }
        if (bad) SPTOL_CHECK(SPTERR_VALUE_ERROR, "HiSpTns Convert", "index >= ndims");
    }
    int r = sptNewSparseTensorHiCOO(hitsr, N, tsr->ndims, nnz, sb_bits, sk_bits, sc_bits);
    if (r) return r;

    /* 1. order: superblock coordinates lexicographic, then Morton inside */
    for (sptIndex m = 0; m < N; ++m) tsr->sortorder[m] = m;
    {
        key_chunk ch[20];
        int nc = 0;
        if (N == 3 || N == 4) {
            unsigned mb = (unsigned)morton_bits_for(tsr);
            if (sk_bits < mb) mb = sk_bits;   /* higher bits are constant in a superblock */
            nc = morton_chunks(N, mb, ch);
        }
        nc += lex_chunks(tsr, sk_bits, KEY_ROWBLOCK, ch + nc);
        r = sort_by_chunks(tsr, 0, nnz, ch, nc, tk);
        if (r) return r;
    }

    /* 2. superblock boundaries (nonzero offsets) */
    const uint32_t *const *I = NULL;
    uint32_t *ip[8];
    if (N > 8) SPTOL_CHECK(SPTERR_VALUE_ERROR, "HiSpTns Convert", "more than 8 modes");
    for (sptIndex m = 0; m < N; ++m) ip[m] = tsr->inds[m].data;
    I = (const uint32_t *const *)ip;
    sptNnzIndex nk = 1, nb = 1;
    for (sptNnzIndex z = 1; z < nnz; ++z) {
        int newk = 0, newb = 0;
        for (sptIndex m = 0; m < N; ++m) {
            uint32_t a = I[m][z], b = I[m][z - 1];
            if ((a >> sb_bits) != (b >> sb_bits)) newb = 1;
            if ((a >> sk_bits) != (b >> sk_bits)) { newk = 1; break; }
        }
        nk += newk;
        nb += newb | newk;
    }

    /* 3. arrays */
    if ((r = exact_vec_nnz(&hitsr->kptr, nk + 1))) return r;
    if ((r = exact_vec_nnz(&hitsr->bptr, nb + 1))) return r;
    for (sptIndex m = 0; m < N; ++m) {
        free(hitsr->binds[m].data);
        hitsr->binds[m].data = (sptBlockIndex *)malloc(nb * sizeof(sptBlockIndex));
        hitsr->binds[m].len = hitsr->binds[m].cap = nb;
        free(hitsr->einds[m].data);
        hitsr->einds[m].data = (sptElementIndex *)malloc(nnz);
        hitsr->einds[m].len = hitsr->einds[m].cap = nnz;
        SPTOL_CHECK_OS(!hitsr->binds[m].data || !hitsr->einds[m].data, "HiSpTns Convert");
    }
    free(hitsr->values.data);
    hitsr->values.data = (sptValue *)malloc(nnz * sizeof(sptValue));
    SPTOL_CHECK_OS(!hitsr->values.data, "HiSpTns Convert");
    hitsr->values.len = hitsr->values.cap = nnz;
    memcpy(hitsr->values.data, tsr->values.data, nnz * sizeof(sptValue));
    #pragma omp parallel for schedule(static) num_threads(max_threads(tk))
    for (sptNnzIndex z = 0; z < nnz; ++z)
        for (sptIndex m = 0; m < N; ++m)
            hitsr->einds[m].data[z] = (sptElementIndex)(I[m][z] & ((1u << sb_bits) - 1u));

    /* 4. blocks, superblocks (as block offsets) and chunks in one pass */
    sptNnzIndexVector *cp = &hitsr->cptr;
    const uint64_t sc = 1ull << sc_bits;
    uint64_t b = 0, k = 0, ne = 0, chunk_size = 0;
    sptNnzIndex *kstart_nnz = (sptNnzIndex *)malloc(nk * sizeof(sptNnzIndex));
    SPTOL_CHECK_OS(!kstart_nnz, "HiSpTns Convert");
    for (sptNnzIndex z = 0; z < nnz; ++z) {
        int newk = (z == 0), newb = (z == 0);
        if (z) {
            for (sptIndex m = 0; m < N; ++m) {
                uint32_t a = I[m][z], c = I[m][z - 1];
                if ((a >> sb_bits) != (c >> sb_bits)) newb = 1;
                if ((a >> sk_bits) != (c >> sk_bits)) { newk = 1; newb = 1; break; }
            }
        }
        if (newk) {
            hitsr->kptr.data[k] = b;          /* index of the first block of this superblock */
            kstart_nnz[k] = z;
            ++k;
            if ((r = sptAppendNnzIndexVector(cp, b))) { free(kstart_nnz); return r; }
            chunk_size = 0;
        }
        if (newb) {
            hitsr->bptr.data[b] = z;
            for (sptIndex m = 0; m < N; ++m) hitsr->binds[m].data[b] = I[m][z] >> sb_bits;
            if (z) {
                /* the previous block's count decides whether a chunk ends */
                if (chunk_size + ne >= sc) {
                    if ((r = sptAppendNnzIndexVector(cp, b))) { free(kstart_nnz); return r; }
                    chunk_size = 0;
                } else {
                    chunk_size += ne;
                }
            }
            ++b;
            ne = 1;
        } else {
            ++ne;
        }
    }
    hitsr->kptr.data[nk] = nb;
    hitsr->bptr.data[nb] = nnz;
    if ((r = sptAppendNnzIndexVector(cp, nb))) { free(kstart_nnz); return r; }

    /* 5. per-mode superblock schedule: kschr[m][i_m >> sk] lists superblock
     *    ids in ascending order; nkiters[m] is the longest list
     *    (convert.c:103-146). */
    for (sptIndex m = 0; m < N; ++m) {
        sptIndex kd = kernel_rows(tsr->ndims[m], sk_bits);
        uint64_t *cnt = (uint64_t *)calloc(kd ? kd : 1, sizeof(uint64_t));
        SPTOL_CHECK_OS(!cnt, "HiSpTns Convert");
        for (sptNnzIndex q = 0; q < nk; ++q) cnt[I[m][kstart_nnz[q]] >> sk_bits]++;
        sptIndex mx = 0;
        for (sptIndex i = 0; i < kd; ++i) {
            sptIndexVector *v = &hitsr->kschr[m][i];
            free(v->data);
            v->cap = cnt[i] < 2 ? 2 : cnt[i];
            v->data = (sptIndex *)malloc(v->cap * sizeof(sptIndex));
            v->len = 0;
            if (cnt[i] > mx) mx = (sptIndex)cnt[i];
        }
        for (sptNnzIndex q = 0; q < nk; ++q) {
            sptIndexVector *v = &hitsr->kschr[m][I[m][kstart_nnz[q]] >> sk_bits];
            v->data[v->len++] = (sptIndex)q;
        }
        hitsr->nkiters[m] = mx;
        free(cnt);
    }
    free(kstart_nnz);

    if (max_nnzb) {
        sptNnzIndex mx = 0;
        for (sptNnzIndex q = 0; q < nb; ++q) {
            sptNnzIndex c = hitsr->bptr.data[q + 1] - hitsr->bptr.data[q];
            if (c > mx) mx = c;
        }
        *max_nnzb = mx;
    }
    return 0;
}

/* Text dump in the reference's layout (hicoo/dump.c:28-72, vector/dump.c). */
static void dump_u32(FILE *fp, const char *label, const uint32_t *d, uint64_t n)
{
    fprintf(fp, "%s length: %" PRIu64 "\n", label, n);
    for (uint64_t i = 0; i < n; ++i) fprintf(fp, "%" PRIu32 "\t", d[i]);
    fprintf(fp, "\n");
}

static void dump_u64(FILE *fp, const sptNnzIndexVector *v)
{
    fprintf(fp, "sptNnzIndexVector length: %" PRIu64 "\n", v->len);
    for (uint64_t i = 0; i < v->len; ++i) fprintf(fp, "%" PRIu64 "\t", v->data[i]);
    fprintf(fp, "\n");
}

int sptDumpSparseTensorHiCOO(sptSparseTensorHiCOO * const hitsr, FILE *fp)
{
    sptIndex N = hitsr->nmodes;
    int iores = fprintf(fp, "%" PRIu32 "\n", N);
    SPTOL_CHECK_OS(iores < 0, "HiSpTns Dump");
    for (sptIndex m = 0; m < N; ++m) fprintf(fp, m ? " %" PRIu32 : "%" PRIu32, hitsr->ndims[m]);
    fputs("\n", fp);
    fprintf(fp, "nkiters:\n");
    dump_u32(fp, "sptIndex array", hitsr->nkiters, N);
    fprintf(fp, "kschr:\n");
    for (sptIndex m = 0; m < N; ++m) {
        fprintf(fp, "mode %" PRIu32 "\n", m);
        sptIndex kd = kernel_rows(hitsr->ndims[m], hitsr->sk_bits);
        for (sptIndex i = 0; i < kd; ++i) dump_u32(fp, "sptIndexVector", hitsr->kschr[m][i].data, hitsr->kschr[m][i].len);
    }
    fprintf(fp, "kptr:\n");
    dump_u64(fp, &hitsr->kptr);
    fprintf(fp, "cptr:\n");
    dump_u64(fp, &hitsr->cptr);
    fprintf(fp, "bptr:\n");
    dump_u64(fp, &hitsr->bptr);
    fprintf(fp, "binds:\n");
    for (sptIndex m = 0; m < N; ++m) dump_u32(fp, "sptBlockIndexVector", hitsr->binds[m].data, hitsr->binds[m].len);
    fprintf(fp, "einds:\n");
    for (sptIndex m = 0; m < N; ++m) {
        fprintf(fp, "sptElementIndexVector length: %" PRIu64 "\n", hitsr->einds[m].len);
        for (uint64_t i = 0; i < hitsr->einds[m].len; ++i) fprintf(fp, "%u\t", (unsigned)hitsr->einds[m].data[i]);
        fprintf(fp, "\n");
    }
    fprintf(fp, "values:\n");
    fprintf(fp, "sptValueVector length: %" PRIu64 "\n", hitsr->values.len);
    for (uint64_t i = 0; i < hitsr->values.len; ++i) fprintf(fp, "%f\t", hitsr->values.data[i]);
    fprintf(fp, "\n");
    return 0;
}
