/*
 * CPU oracle for the MTTKRP / HiCOO / CPD-ALS path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code; only tests/, the smoke check and the CPU-baseline leg of
 * bench.py use it, and only as the checker or the timed CPU baseline.
 *
 * Plain arrays in, plain arrays out (no ParTI structs), so the tests can call
 * it through ctypes.  Index arrays are passed as an array of nmodes pointers
 * (structure of arrays), factor matrices as an array of row-major pointers
 * with a common row stride.
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   oracle_mttkrp_coo_f32     bit-exact vs the compiled reference sptMTTKRP on
 *                             golden fixtures; closed forms (all-ones factors,
 *                             single nonzero, dense rank-1)
 *   oracle_mttkrp_coo_f64     dense einsum brute force (numpy float64)
 *   oracle_hicoo_convert      bit-exact vs the reference's HiCOO dumps; order
 *                             invariants (row-block then Morton)
 *   oracle_mttkrp_hicoo_*     bit-exact vs the compiled reference
 *                             sptMTTKRPHiCOO; equals COO MTTKRP on the same
 *                             permuted tensor
 *   oracle_cpd_als_f64        reference CpdAlsStep factors/lambda (FP32
 *                             LAPACK, tolerance); exact low-rank recovery
 *                             (fit -> 1); fit closed form on a dense tensor
 *   oracle_*_omp_*            equal (tolerance) to the sequential functions
 *   oracle_mttkrp_coo_omp_reduce_f32
 *                             bit-exact vs the compiled reference
 *                             sptOmpMTTKRP_Reduce with the same thread count
 */
#ifndef SPTOL_ORACLE_H
#define SPTOL_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* out[out_rows*stride] = MTTKRP of mode `mode`, FP32, reference loop order. */
void oracle_mttkrp_coo_f32(uint32_t nmodes, uint64_t nnz, const uint32_t *const *inds,
                           const float *vals, const float *const *mats, uint32_t stride,
                           uint32_t R, const uint32_t *mats_order, uint32_t mode,
                           uint32_t out_rows, float *out);
/* Same, every product and sum in double. */
void oracle_mttkrp_coo_f64(uint32_t nmodes, uint64_t nnz, const uint32_t *const *inds,
                           const float *vals, const float *const *mats, uint32_t stride,
                           uint32_t R, const uint32_t *mats_order, uint32_t mode,
                           uint32_t out_rows, double *out);
/* OpenMP restatement of sptOmpMTTKRP (atomic update per (nnz, r)). */
void oracle_mttkrp_coo_omp_f32(uint32_t nmodes, uint64_t nnz, const uint32_t *const *inds,
                               const float *vals, const float *const *mats, uint32_t stride,
                               uint32_t R, const uint32_t *mats_order, uint32_t mode,
                               uint32_t out_rows, float *out, int tk);
/* OpenMP restatement of sptOmpMTTKRP_Reduce (static partition, private
 * copies summed in thread order: bit-identical to the reference per tk). */
void oracle_mttkrp_coo_omp_reduce_f32(uint32_t nmodes, uint64_t nnz, const uint32_t *const *inds,
                                      const float *vals, const float *const *mats, uint32_t stride,
                                      uint32_t R, const uint32_t *mats_order, uint32_t mode,
                                      uint32_t out_rows, float *out, int tk);

/* Lexicographic COO sort (sptSparseTensorSortIndex), in place. */
void oracle_sort_lex(uint32_t nmodes, uint64_t nnz, uint32_t **inds, float *vals);

/* Three-way compare of two coordinates in the HiCOO Morton order. */
int oracle_morton_cmp(uint32_t nmodes, const uint32_t *a, const uint32_t *b);

typedef struct {
    uint32_t nmodes;
    uint32_t sb_bits, sk_bits, sc_bits;
    uint64_t nnz, nb, nk, nc;
    uint64_t max_nnzb;
    uint64_t *kptr;      /* nk+1 block offsets of superblocks            */
    uint64_t *cptr;      /* nc chunk starts (block offsets), last = nb   */
    uint64_t *bptr;      /* nb+1 nonzero offsets of blocks               */
    uint32_t *binds;     /* nmodes*nb, mode-major block coordinates      */
    uint8_t  *einds;     /* nmodes*nnz, mode-major element offsets       */
    float    *values;    /* nnz                                          */
    uint32_t *nkiters;   /* nmodes                                       */
    uint32_t *nkd;       /* nmodes: number of superblock rows per mode   */
    uint64_t *kschr_ptr; /* per mode m: nkd[m]+1 offsets, concatenated   */
    uint32_t *kschr_idx; /* per mode m: nk kernel ids, concatenated      */
} oracle_hicoo;

/* sptSparseTensorToHiCOO: permutes inds/vals in place and fills *h. */
int  oracle_hicoo_convert(uint32_t nmodes, const uint32_t *ndims, uint64_t nnz,
                          uint32_t **inds, float *vals, uint32_t sb_bits, uint32_t sk_bits,
                          oracle_hicoo *h);
void oracle_hicoo_free(oracle_hicoo *h);

void oracle_mttkrp_hicoo_f32(const oracle_hicoo *h, const float *const *mats, uint32_t stride,
                             uint32_t R, const uint32_t *mats_order, uint32_t mode,
                             uint32_t out_rows, float *out);
void oracle_mttkrp_hicoo_f64(const oracle_hicoo *h, const float *const *mats, uint32_t stride,
                             uint32_t R, const uint32_t *mats_order, uint32_t mode,
                             uint32_t out_rows, double *out);
/* OpenMP restatement of the superblock-row scheduled HiCOO MTTKRP. */
void oracle_mttkrp_hicoo_sched_omp_f32(const oracle_hicoo *h, const float *const *mats,
                                       uint32_t stride, uint32_t R, const uint32_t *mats_order,
                                       uint32_t mode, uint32_t out_rows, float *out, int tk);
/* Scheduled_Reduce restatement (thread-private copies, ordered sum). */
void oracle_mttkrp_hicoo_sched_reduce_omp_f32(const oracle_hicoo *h, const float *const *mats,
                                              uint32_t stride, uint32_t R, const uint32_t *mats_order,
                                              uint32_t mode, uint32_t out_rows, float *out, int tk);

/* CPD-ALS in double precision with the reference's update rules and a
 * correct fit.  mats[m] (ndims[m] x R, row stride R) hold the initial factors
 * on entry and the final factors on exit.  fit_hist receives one fit per
 * executed iteration (capacity niters).  Returns iterations executed. */
int oracle_cpd_als_f64(uint32_t nmodes, const uint32_t *ndims, uint64_t nnz,
                       const uint32_t *const *inds, const float *vals, uint32_t R,
                       uint32_t niters, double tol, double **mats, double *lambda,
                       double *fit_hist);

#ifdef __cplusplus
}
#endif
#endif
