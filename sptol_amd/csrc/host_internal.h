/* Internal declarations shared by the host C files and the HIP layer. */
#ifndef SPTOL_HOST_INTERNAL_H
#define SPTOL_HOST_INTERNAL_H

#include "../../include/ParTI.h"
#include <errno.h>
#include <string.h>
#include <stdlib.h>
#include <stdio.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Record an error in the thread-local last-error slot and print it (the
 * reference's spt_ComplainError, src/error/error.c:46-71). */
void sptol_complain(const char *module, int errcode, const char *file, unsigned line,
                    const char *reason);

#define SPTOL_CHECK(errcode, module, reason)                                   \
    do {                                                                       \
        int sptol_e_ = (errcode);                                              \
        if (sptol_e_ != 0) {                                                   \
            sptol_complain((module), sptol_e_, __FILE__, __LINE__, (reason));  \
            return sptol_e_;                                                   \
        }                                                                      \
    } while (0)

#define SPTOL_CHECK_OS(cond, module)                                           \
    do {                                                                       \
        if (cond) {                                                            \
            int sptol_en_ = errno ? errno : ENOMEM;                            \
            sptol_complain((module), SPTERR_OS_ERROR + sptol_en_, __FILE__,    \
                           __LINE__, strerror(sptol_en_));                     \
            return SPTERR_OS_ERROR + sptol_en_;                                \
        }                                                                      \
    } while (0)

/* Seeded generator used by sptRandomizeMatrix (splitmix64 stream). */
uint64_t sptol_rng_next(void);
float    sptol_rng_uniform01(void);

/* Stable LSD radix sort of (key, idx) pairs on the low `bits` bits of key.
 * keys/idx are permuted in place; tmp buffers are allocated internally.
 * Returns 0 or an OS error code. */
int sptol_radix_sort_u64(uint64_t *keys, uint64_t *idx, uint64_t n, unsigned bits, int nthreads);

/* dst[0, n) = src[0, n), the copy split in 1 MiB pieces over the OpenMP
 * threads (so a fresh destination's pages are first touched in parallel). */
void sptol_parallel_memcpy(void *dst, const void *src, size_t n);

/* dst[0, n) = bytes [off, off+n) of the file open as fd, read with pread in
 * 1 MiB pieces over the OpenMP threads.  Returns 0, or an errno value (EIO
 * when the file ends early). */
int sptol_parallel_pread(int fd, void *dst, size_t n, uint64_t off);

/* Apply perm (new position i takes old element perm[i]) to tensor rows
 * [begin, begin+n) of a COO tensor's index and value arrays. */
int sptol_permute_coo(sptSparseTensor *tsr, uint64_t begin, const uint64_t *perm, uint64_t n);

/* Binary tensor file (sptolSaveSparseTensorBinary): read and check the
 * header; on return fp is positioned at inds[0].  ndims holds 8 entries. */
int sptol_read_binary_header(FILE *fp, uint32_t *nmodes, uint64_t *nnz, uint32_t *ndims);

/* Number of bits needed to hold values in [0, n). */
static inline unsigned sptol_bits_for(uint64_t n)
{
    unsigned b = 0;
    while (b < 64 && (n > (1ull << b))) ++b;
    return b;
}

#ifdef __cplusplus
}
#endif
#endif
