// MTTKRP on gfx950: COO and HiCOO kernels, device tensor handles and the
// ParTI!-compatible entry points sptMTTKRP / sptMTTKRPHiCOO / ... .
//
// Reference semantics (src/sptensor/mttkrp.c:39-177, hicoo/mttkrp.c:53-408):
//   out[i_mode][r] = sum_x v_x * prod_{k=1..N-1} A_{order[k]}[i_{order[k]}][r]
// with the output zero-filled first.  Per nonzero the product is formed in
// the reference's association ((v*A_o1)*A_o2)*...; the sum over nonzeros is
// reassociated (register runs + atomics), so FP32 results agree with the
// sequential reference to rounding, not bit for bit.
//
// Kernel design (HBM/L2 bound, no MFMA):
//  * A "lane group" of LPN = stride/4 lanes (capped at 64) owns one nonzero
//    at a time; lane l handles columns 4l..4l+3 with one 16-byte gather per
//    factor row, so a row gather is one contiguous 4*stride-byte access.
//  * Each group walks a contiguous range of nonzeros, U at a time: all index
//    and value loads of the U nonzeros, then all U*(N-1) row gathers, are
//    issued before any use, so U gathers per group are in flight.
//  * Products are summed in registers while the output row repeats and
//    flushed with one float atomic per column when it changes.  On a copy
//    sorted by the output mode this turns nnz*R atomics into about
//    (rows + groups)*R; on unsorted input it degrades to one atomic per
//    nonzero and column.
//  * HiCOO: the group finds the block of each of its U nonzeros from a
//    window of U block pointers (at most U block boundaries can fall inside
//    U consecutive nonzeros), then forms coordinates as
//    (binds << sb_bits) + einds from 1-byte element offsets.
#include "dev_common.hpp"
#include <algorithm>
#include <cerrno>
#include <cstdlib>
#include <hipcub/hipcub.hpp>
#include <vector>
#include <cstdio>
#include <cstring>
#include <chrono>

namespace sptol {

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ float4 ld4(const float *p)
{
    return *reinterpret_cast<const float4 *>(p);
}

__device__ __forceinline__ float4 ld4(const uint16_t *p)
{
    const uint2 u = *reinterpret_cast<const uint2 *>(p);
    float4 r;
    r.x = __uint_as_float(u.x << 16);
    r.y = __uint_as_float(u.x & 0xFFFF0000u);
    r.z = __uint_as_float(u.y << 16);
    r.w = __uint_as_float(u.y & 0xFFFF0000u);
    return r;
}

__device__ __forceinline__ void mul4(float4 &a, const float4 b)
{
    a.x *= b.x; a.y *= b.y; a.z *= b.z; a.w *= b.w;
}

__device__ __forceinline__ void add4(float4 &a, const float4 b)
{
    a.x += b.x; a.y += b.y; a.z += b.z; a.w += b.w;
}

__device__ __forceinline__ void flush4(float *out, uint32_t row, uint32_t S, uint32_t c0, uint32_t R,
                                       const float4 acc)
{
    float *o = out + (size_t)row * S + c0;
    if (c0 + 3 < R) {
        unsafeAtomicAdd(o + 0, acc.x);
        unsafeAtomicAdd(o + 1, acc.y);
        unsafeAtomicAdd(o + 2, acc.z);
        unsafeAtomicAdd(o + 3, acc.w);
    } else {
        if (c0 + 0 < R) unsafeAtomicAdd(o + 0, acc.x);
        if (c0 + 1 < R) unsafeAtomicAdd(o + 1, acc.y);
        if (c0 + 2 < R) unsafeAtomicAdd(o + 2, acc.z);
    }
}

__global__ void k_fill_zero(float4 *p, uint64_t n4)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n4; i += step) p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// ---------------------------------------------------------------------------
// COO kernel
// ---------------------------------------------------------------------------
template <int LPN, int NM, int U, typename FT>
__global__ __launch_bounds__(256) void k_mttkrp_coo(const MttkrpArgs a)
{
    const uint64_t tid = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    const uint64_t group = tid / LPN;
    const uint32_t l = threadIdx.x % LPN;
    const uint64_t z0 = group * a.per_group;
    if (z0 >= a.nnz) return;
    const uint64_t z1 = min(z0 + a.per_group, a.nnz);
    const uint32_t S = a.stride, R = a.R;
    const uint32_t nm = NM > 0 ? NM : a.nmodes;
    const uint32_t *__restrict__ rowi = a.inds[a.mode];
    const uint32_t o1 = a.order[1];
    const uint32_t *__restrict__ i1 = a.inds[o1];
    const FT *__restrict__ A1 = (const FT *)a.mats[o1];
    const float *__restrict__ vals = a.vals;

    for (uint32_t c0 = l * 4; c0 < S; c0 += 4 * LPN) {
        if (c0 >= R) break;
        float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
        uint32_t cur = 0xFFFFFFFFu;
        uint64_t z = z0;
        for (; z + U <= z1; z += U) {
            uint32_t row[U];
            float4 p[U];
            float v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                row[u] = rowi[z + u];
                v[u] = vals[z + u];
                p[u] = ld4(A1 + (size_t)i1[z + u] * S + c0);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                p[u].x = v[u] * p[u].x; p[u].y = v[u] * p[u].y;
                p[u].z = v[u] * p[u].z; p[u].w = v[u] * p[u].w;
            }
#pragma unroll
            for (int k = 2; k < (NM > 0 ? NM : kMaxModes); ++k) {
                if (NM == 0 && k >= (int)nm) break;
                const uint32_t ok = a.order[k];
                const uint32_t *__restrict__ ik = a.inds[ok];
                const FT *__restrict__ Ak = (const FT *)a.mats[ok];
                float4 q[U];
#pragma unroll
                for (int u = 0; u < U; ++u) q[u] = ld4(Ak + (size_t)ik[z + u] * S + c0);
#pragma unroll
                for (int u = 0; u < U; ++u) mul4(p[u], q[u]);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (row[u] != cur) {
                    if (cur != 0xFFFFFFFFu) flush4(a.out, cur, S, c0, R, acc);
                    cur = row[u];
                    acc = p[u];
                } else {
                    add4(acc, p[u]);
                }
            }
        }
        for (; z < z1; ++z) {
            const uint32_t r = rowi[z];
            const float vv = vals[z];
            float4 pp = ld4(A1 + (size_t)i1[z] * S + c0);
            pp.x = vv * pp.x; pp.y = vv * pp.y; pp.z = vv * pp.z; pp.w = vv * pp.w;
            for (uint32_t k = 2; k < nm; ++k) {
                const uint32_t ok = a.order[k];
                mul4(pp, ld4((const FT *)a.mats[ok] + (size_t)a.inds[ok][z] * S + c0));
            }
            if (r != cur) {
                if (cur != 0xFFFFFFFFu) flush4(a.out, cur, S, c0, R, acc);
                cur = r;
                acc = pp;
            } else {
                add4(acc, pp);
            }
        }
        if (cur != 0xFFFFFFFFu) flush4(a.out, cur, S, c0, R, acc);
    }
}

// ---------------------------------------------------------------------------
// HiCOO kernel
// ---------------------------------------------------------------------------
// First block b with bptr[b] <= z < bptr[b+1].
__device__ __forceinline__ uint64_t find_block(const uint64_t *__restrict__ bptr, uint64_t nb, uint64_t z)
{
    uint64_t lo = 0, hi = nb;   // invariant: bptr[lo] <= z < bptr[hi]
    while (hi - lo > 1) {
        uint64_t mid = (lo + hi) >> 1;
        if (bptr[mid] <= z) lo = mid; else hi = mid;
    }
    return lo;
}

template <int LPN, int NM, int U, typename FT>
__global__ __launch_bounds__(256) void k_mttkrp_hicoo(const MttkrpArgs a)
{
    const uint64_t tid = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    const uint64_t group = tid / LPN;
    const uint32_t l = threadIdx.x % LPN;
    const uint64_t z0 = group * a.per_group;
    if (z0 >= a.nnz) return;
    const uint64_t z1 = min(z0 + a.per_group, a.nnz);
    const uint32_t S = a.stride, R = a.R, sb = a.sb_bits;
    const uint32_t nm = NM > 0 ? NM : a.nmodes;
    const uint64_t *__restrict__ bptr = a.bptr;
    const uint64_t nb = a.nb;
    const uint32_t mode = a.mode;

    for (uint32_t c0 = l * 4; c0 < S; c0 += 4 * LPN) {
        if (c0 >= R) break;
        float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
        uint32_t cur = 0xFFFFFFFFu;
        uint64_t b = find_block(bptr, nb, z0);
        for (uint64_t z = z0; z < z1; z += U) {
            // block ids of z..z+U-1 from a window of U block pointers
            uint64_t bnext[U];
#pragma unroll
            for (int t = 0; t < U; ++t) bnext[t] = bptr[min(b + 1 + t, nb)];
            uint64_t blk[U];
            bool valid[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                uint64_t c = 0;
#pragma unroll
                for (int t = 0; t < U; ++t) c += (bnext[t] <= z + u) ? 1u : 0u;
                blk[u] = b + c;
                valid[u] = (z + u) < z1;
            }
            uint32_t coord[U][NM > 0 ? NM : kMaxModes];
            float v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                // padding lanes repeat the chunk's first nonzero z (block blk[0]):
                // a block coordinate and an element offset of the same nonzero
                const uint64_t zz = valid[u] ? z + u : z;
                const uint64_t bb = valid[u] ? blk[u] : blk[0];
                v[u] = a.vals[zz];
#pragma unroll
                for (int m = 0; m < (NM > 0 ? NM : kMaxModes); ++m) {
                    if (NM == 0 && m >= (int)nm) break;
                    coord[u][m] = (a.binds[m][bb] << sb) + (uint32_t)a.einds[m][zz];
                }
            }
            float4 p[U];
            const uint32_t o1 = a.order[1];
            const FT *__restrict__ A1 = (const FT *)a.mats[o1];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                p[u] = ld4(A1 + (size_t)coord[u][o1] * S + c0);
                p[u].x = v[u] * p[u].x; p[u].y = v[u] * p[u].y;
                p[u].z = v[u] * p[u].z; p[u].w = v[u] * p[u].w;
            }
#pragma unroll
            for (int k = 2; k < (NM > 0 ? NM : kMaxModes); ++k) {
                if (NM == 0 && k >= (int)nm) break;
                const uint32_t ok = a.order[k];
                const FT *__restrict__ Ak = (const FT *)a.mats[ok];
                float4 q[U];
#pragma unroll
                for (int u = 0; u < U; ++u) q[u] = ld4(Ak + (size_t)coord[u][ok] * S + c0);
#pragma unroll
                for (int u = 0; u < U; ++u) mul4(p[u], q[u]);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (!valid[u]) break;
                const uint32_t row = coord[u][mode];
                if (row != cur) {
                    if (cur != 0xFFFFFFFFu) flush4(a.out, cur, S, c0, R, acc);
                    cur = row;
                    acc = p[u];
                } else {
                    add4(acc, p[u]);
                }
            }
            b = blk[U - 1];
        }
        if (cur != 0xFFFFFFFFu) flush4(a.out, cur, S, c0, R, acc);
    }
}

// ---------------------------------------------------------------------------
// Launch logic
// ---------------------------------------------------------------------------
int num_cus()
{
    // read once, by whichever thread first asks (a thread-safe static)
    static const int n = [] {
        int dev = 0, c = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
            c = prop.multiProcessorCount;
        return c > 0 ? c : 256;
    }();
    return n;
}

int launch_fill_zero(float *p, uint64_t n, hipStream_t s)
{
    if (n == 0) return 0;
    if (n % 4) SPTOL_CHECK(SPTERR_VALUE_ERROR, "Fill", "length not a multiple of 4");
    uint64_t n4 = n / 4;
    uint64_t blocks = (n4 + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    int tok = prof_start("fill_zero", s);
    hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)blocks), dim3(256), 0, s, (float4 *)p, n4);
    prof_stop(tok, s);
    SPTOL_HIP(hipGetLastError(), "Fill");
    return 0;
}

static uint32_t lanes_per_nnz(uint32_t stride)
{
    uint32_t l = stride / 4;
    if (l > 64) l = 64;
    uint32_t p = 1;
    while (p < l) p <<= 1;
    return p;
}

#define SPTOL_DISPATCH_NM(KERNEL, LPN, U, FT, nm, grid, s, a)                                  \
    do {                                                                                       \
        if ((nm) == 3) hipLaunchKernelGGL((KERNEL<LPN, 3, U, FT>), grid, dim3(256), 0, s, a);  \
        else if ((nm) == 4) hipLaunchKernelGGL((KERNEL<LPN, 4, U, FT>), grid, dim3(256), 0, s, a); \
        else hipLaunchKernelGGL((KERNEL<LPN, 0, U, FT>), grid, dim3(256), 0, s, a);            \
    } while (0)

#define SPTOL_DISPATCH_LPN(KERNEL, FT, lpn, nm, grid, s, a)                                    \
    do {                                                                                       \
        switch (lpn) {                                                                         \
        case 1: case 2: SPTOL_DISPATCH_NM(KERNEL, 2, 8, FT, nm, grid, s, a); break;            \
        case 4: SPTOL_DISPATCH_NM(KERNEL, 4, 8, FT, nm, grid, s, a); break;                    \
        case 8: SPTOL_DISPATCH_NM(KERNEL, 8, 8, FT, nm, grid, s, a); break;                    \
        case 16: SPTOL_DISPATCH_NM(KERNEL, 16, 8, FT, nm, grid, s, a); break;                  \
        case 32: SPTOL_DISPATCH_NM(KERNEL, 32, 4, FT, nm, grid, s, a); break;                  \
        default: SPTOL_DISPATCH_NM(KERNEL, 64, 4, FT, nm, grid, s, a); break;                  \
        }                                                                                      \
    } while (0)

static void plan(MttkrpArgs &a, uint32_t lpn, uint64_t *blocks)
{
    const uint32_t groups_per_wave = 64 / (lpn < 2 ? 2 : lpn);
    const uint64_t target_groups = (uint64_t)num_cus() * 16 * groups_per_wave;
    uint64_t per = (a.nnz + target_groups - 1) / target_groups;
    if (per < 8) per = 8;
    per = (per + 7) / 8 * 8;
    a.per_group = per;
    const uint64_t groups = (a.nnz + per - 1) / per;
    const uint64_t threads = groups * (lpn < 2 ? 2 : lpn);
    *blocks = (threads + 255) / 256;
}

int launch_mttkrp_coo(const MttkrpArgs &in, int dtype, hipStream_t s)
{
    MttkrpArgs a = in;
    if (a.nnz == 0) return 0;
    const uint32_t lpn = lanes_per_nnz(a.stride);
    uint64_t blocks;
    plan(a, lpn, &blocks);
    if (blocks > 0x7FFFFFFFull) SPTOL_CHECK(SPTERR_VALUE_ERROR, "MTTKRP", "grid too large");
    dim3 grid((unsigned)blocks);
    int tok = prof_start("mttkrp_coo", s);
    if (dtype == SPTOL_BF16) SPTOL_DISPATCH_LPN(k_mttkrp_coo, uint16_t, lpn, a.nmodes, grid, s, a);
    else SPTOL_DISPATCH_LPN(k_mttkrp_coo, float, lpn, a.nmodes, grid, s, a);
    prof_stop(tok, s);
    SPTOL_HIP(hipGetLastError(), "MTTKRP COO launch");
    return 0;
}

int launch_mttkrp_hicoo(const MttkrpArgs &in, int dtype, hipStream_t s)
{
    MttkrpArgs a = in;
    if (a.nnz == 0) return 0;
    const uint32_t lpn = lanes_per_nnz(a.stride);
    uint64_t blocks;
    plan(a, lpn, &blocks);
    if (blocks > 0x7FFFFFFFull) SPTOL_CHECK(SPTERR_VALUE_ERROR, "MTTKRP", "grid too large");
    dim3 grid((unsigned)blocks);
    int tok = prof_start("mttkrp_hicoo", s);
    if (dtype == SPTOL_BF16) SPTOL_DISPATCH_LPN(k_mttkrp_hicoo, uint16_t, lpn, a.nmodes, grid, s, a);
    else SPTOL_DISPATCH_LPN(k_mttkrp_hicoo, float, lpn, a.nmodes, grid, s, a);
    prof_stop(tok, s);
    SPTOL_HIP(hipGetLastError(), "MTTKRP HiCOO launch");
    return 0;
}

// ---------------------------------------------------------------------------
// Device tensor handles
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_gather(T *__restrict__ dst, const T *__restrict__ src, const uint32_t *__restrict__ perm,
                         uint64_t n)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += step) dst[i] = src[perm[i]];
}

__global__ void k_iota(uint32_t *p, uint64_t n)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += step) p[i] = (uint32_t)i;
}

static unsigned grid_for(uint64_t n)
{
    uint64_t b = (n + 255) / 256;
    if (b > 65536) b = 65536;
    if (b == 0) b = 1;
    return (unsigned)b;
}

}  // namespace sptol

using namespace sptol;

extern "C" {

}  // extern "C"

namespace sptol {

int dup_on(void **dst, const void *src, size_t bytes, hipStream_t s, const char *module)
{
    SPTOL_TRY(dmalloc(dst, bytes, module));
    if (!bytes) return 0;
    const int r = upload_pageable(*dst, src, bytes, s);
    if (r) {
        dfree(*dst);
        *dst = nullptr;
        sptol_complain(module, r, __FILE__, __LINE__, "upload failed");
        return r;
    }
    return 0;
}


struct CheckArgs {
    const uint32_t *inds[kMaxModes];
    uint32_t ndims[kMaxModes];
    const float *vals;     // null: values not checked
    uint64_t nnz;
    uint32_t nmodes;
};

// bad[0] |= 1 for an index >= ndims, |= 2 for a NaN or infinite value,
// |= 4 << m where mode m's index decreases from one nonzero to the next:
// one pass over every array of the tensor.
__global__ void k_check_tensor(const CheckArgs a, uint32_t *__restrict__ bad)
{
    uint32_t b = 0;
    for (uint64_t z = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; z < a.nnz; z += (uint64_t)gridDim.x * blockDim.x) {
        for (uint32_t m = 0; m < a.nmodes; ++m) {
            const uint32_t i = a.inds[m][z];
            b |= i >= a.ndims[m] ? 1u : 0u;
            b |= (z > 0 && i < a.inds[m][z - 1]) ? (4u << m) : 0u;
        }
        if (a.vals) b |= (__float_as_uint(a.vals[z]) & 0x7F800000u) == 0x7F800000u ? 2u : 0u;
    }
    // Out-of-order modes set bits in most threads: merge them in LDS and
    // let one thread per block add only bits the flag does not hold yet, so
    // the global atomics stay few (one per block at most, none once every
    // bit is known).
    __shared__ uint32_t sb;
    if (threadIdx.x == 0) sb = 0u;
    __syncthreads();
    if (b) atomicOr(&sb, b);
    __syncthreads();
    if (threadIdx.x == 0 && sb && (sb & ~__hip_atomic_load(bad, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)))
        atomicOr(bad, sb);
}

// Reject any index >= ndims (and, with check_vals, any non-finite value)
// on the device (t's arrays are on s).
static int check_inds_dev(sptolDevCoo t, hipStream_t s, const char *module, bool check_vals = false)
{
    if (!t->nnz) return 0;
    CheckArgs a;
    std::memset(&a, 0, sizeof(a));
    for (uint32_t m = 0; m < t->nmodes; ++m) { a.inds[m] = t->inds[m]; a.ndims[m] = t->ndims[m]; }
    a.vals = check_vals ? t->vals : nullptr;
    a.nnz = t->nnz;
    a.nmodes = t->nmodes;
    uint32_t *bad = nullptr;
    SPTOL_TRY(dmalloc((void **)&bad, 4, module));
    uint32_t hb = 0;
    hipError_t e = hipMemsetAsync(bad, 0, 4, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_check_tensor, dim3(grid_for(t->nnz)), dim3(256), 0, s, a, bad);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&hb, bad, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    dfree(bad);
    SPTOL_HIP(e, module);
    if (hb & 1u) SPTOL_CHECK(SPTERR_VALUE_ERROR, module, "index >= ndims");
    if (hb & 2u) SPTOL_CHECK(SPTERR_VALUE_ERROR, module, "non-finite value");
    for (uint32_t m = 0; m < t->nmodes; ++m) t->in_order[m] = !(hb & (4u << m));
    return 0;
}

int coo_load_binary(sptolDevCoo *h, const char *path, hipStream_t s, bool check_vals)
{
    *h = nullptr;
    StreamScope scope(s);   // everything freed here was used on s only
    FILE *fp = fopen(path, "rb");
    if (!fp) {
        const int en = errno ? errno : ENOENT;
        sptol_complain("DevCoo Load", SPTERR_OS_ERROR + en, __FILE__, __LINE__, strerror(en));
        return SPTERR_OS_ERROR + en;
    }
    uint32_t nm = 0, nd[8] = {};
    uint64_t nnz = 0;
    int r = sptol_read_binary_header(fp, &nm, &nnz, nd);
    if (!r && nm < 2) {
        sptol_complain("DevCoo Load", SPTERR_VALUE_ERROR, __FILE__, __LINE__, "nmodes must be in [2, 8]");
        r = SPTERR_VALUE_ERROR;
    }
    if (r) { fclose(fp); return r; }
    sptolDevCoo t = new sptolDevCoo_s();
    t->nmodes = nm;
    t->nnz = nnz;
    for (uint32_t m = 0; m < nm; ++m) t->ndims[m] = nd[m];
    for (uint32_t m = 0; m <= nm && !r; ++m)
        r = dmalloc(m < nm ? (void **)&t->inds[m] : (void **)&t->vals, nnz * 4, "DevCoo Load");
    // every array in file order, chunk by chunk: the read of a chunk (its
    // pieces read in parallel) overlaps the copy of the previous one
    const long data0 = r ? -1 : ftell(fp);
    if (!r && data0 < 0) {
        const int en = errno ? errno : EIO;
        sptol_complain("DevCoo Load", SPTERR_OS_ERROR + en, __FILE__, __LINE__, strerror(en));
        r = SPTERR_OS_ERROR + en;
    }
    const int fd = fileno(fp);
    void *buf[2] = {};
    hipEvent_t ev[2] = {};
    bool used[2] = {false, false};
    hipError_t e = hipSuccess;
    for (int k = 0; k < 2 && !r; ++k) {
        r = pinned_chunk(&buf[k]);
        if (!r && (e = hipEventCreateWithFlags(&ev[k], hipEventDisableTiming)) != hipSuccess)
            r = SPTERR_CUDA_ERROR + (int)e;
    }
    int k = 0;
    for (uint32_t m = 0; m <= nm && !r; ++m) {
        char *dst = (char *)(m < nm ? (void *)t->inds[m] : (void *)t->vals);
        for (uint64_t off = 0; off < nnz * 4 && !r; off += kPinnedChunkBytes) {
            const size_t n = (size_t)std::min<uint64_t>(kPinnedChunkBytes, nnz * 4 - off);
            if (used[k] && (e = hipEventSynchronize(ev[k])) != hipSuccess) { r = SPTERR_CUDA_ERROR + (int)e; break; }
            const int en = sptol_parallel_pread(fd, buf[k], n, (uint64_t)data0 + (uint64_t)m * nnz * 4 + off);
            if (en) {
                sptol_complain("DevCoo Load", SPTERR_OS_ERROR + en, __FILE__, __LINE__,
                               en == EIO ? "short read" : strerror(en));
                r = SPTERR_OS_ERROR + en;
                break;
            }
            e = hipMemcpyAsync(dst + off, buf[k], n, hipMemcpyHostToDevice, s);
            if (e == hipSuccess) e = hipEventRecord(ev[k], s);
            if (e != hipSuccess) { r = SPTERR_CUDA_ERROR + (int)e; break; }
            used[k] = true;
            k ^= 1;
        }
    }
    fclose(fp);
    e = hipStreamSynchronize(s);
    if (!r && e != hipSuccess) r = SPTERR_CUDA_ERROR + (int)e;
    for (int q = 0; q < 2; ++q) {
        pinned_chunk_free(buf[q]);
        if (ev[q]) (void)hipEventDestroy(ev[q]);
    }
    if (!r) r = check_inds_dev(t, s, "DevCoo Load", check_vals);
    if (r) { sptolDevCooFree(t); return r; }
    *h = t;
    return 0;
}

int coo_create(sptolDevCoo *h, const sptSparseTensor *X, hipStream_t s, bool check_vals)
{
    *h = nullptr;
    StreamScope scope(s);   // everything freed here was used on s only
    if (X->nmodes < 2 || X->nmodes > kMaxModes)
        SPTOL_CHECK(SPTERR_VALUE_ERROR, "DevCoo New", "nmodes must be in [2, 8]");
    sptolDevCoo t = new sptolDevCoo_s();
    t->nmodes = X->nmodes;
    t->nnz = X->nnz;
    for (uint32_t m = 0; m < t->nmodes; ++m) t->ndims[m] = X->ndims[m];
    int r = 0;
    for (uint32_t m = 0; m < t->nmodes && !r; ++m)
        r = dup_on((void **)&t->inds[m], X->inds[m].data, X->nnz * sizeof(uint32_t), s, "DevCoo New");
    if (!r) r = dup_on((void **)&t->vals, X->values.data, X->nnz * sizeof(float), s, "DevCoo New");
    if (r) { sptolDevCooFree(t); return r; }
    // The kernels index factor rows with these values, so reject any index
    // >= ndims before any kernel uses them (the reference trusts its input;
    // here an out-of-range index would be a device fault).
    if ((r = check_inds_dev(t, s, "DevCoo New", check_vals))) { sptolDevCooFree(t); return r; }
    *h = t;
    return 0;
}

static int coo_prepare_from(sptolDevCoo h, uint32_t mode, hipStream_t s, const uint4 *rec0, const uint32_t *iota,
                            const uint16_t *keys16);

int coo_prepare(sptolDevCoo h, uint32_t mode, hipStream_t s)
{
    return coo_prepare_from(h, mode, s, nullptr, nullptr, nullptr);
}

int coo_prepare_all(sptolDevCoo h, hipStream_t s)
{
    bool todo = false, sorts = false;
    for (uint32_t m = 0; m < h->nmodes; ++m) {
        todo = todo || !h->sorted[m].ready;
        sorts = sorts || (!h->sorted[m].ready && !h->in_order[m]);
    }
    if (!todo) return 0;
    if ((h->nmodes != 3 && h->nmodes != 4) || h->nnz == 0 || h->nnz > kMaxSortItems || !sorts) {
        for (uint32_t m = 0; m < h->nmodes; ++m) SPTOL_TRY(coo_prepare(h, m, s));
        return 0;
    }
    StreamScope scope(s);   // the packed records, 0..nnz-1 and the 16-bit keys are used on s only
    uint4 *rec0 = nullptr;
    uint32_t *iota = nullptr;
    // modes of at most 2^16 rows sort 16-bit keys (6 bytes a pass instead of 8)
    uint16_t *k16[kMaxModes] = {}, *k16_all = nullptr;
    uint32_t n16 = 0;
    for (uint32_t m = 0; m < h->nmodes; ++m)
        n16 += (!h->sorted[m].ready && !h->in_order[m] && h->ndims[m] <= 65536u) ? 1u : 0u;
    int r = dmalloc((void **)&rec0, h->nnz * sizeof(uint4), "DevCoo Prepare");
    if (!r) r = dmalloc((void **)&iota, h->nnz * 4, "DevCoo Prepare");
    if (!r && n16) r = dmalloc((void **)&k16_all, h->nnz * 2 * n16, "DevCoo Prepare");
    for (uint32_t m = 0, k = 0; m < h->nmodes && !r && n16; ++m)
        if (!h->sorted[m].ready && !h->in_order[m] && h->ndims[m] <= 65536u) k16[m] = k16_all + (size_t)(k++) * h->nnz;
    if (!r) r = seg_pack_unsorted(rec0, h->nmodes, h->nnz, h->inds, h->vals, s, k16);
    if (!r) hipLaunchKernelGGL(k_iota, dim3(grid_for(h->nnz)), dim3(256), 0, s, iota, h->nnz);
    for (uint32_t m = 0; m < h->nmodes && !r; ++m) r = coo_prepare_from(h, m, s, rec0, iota, k16[m]);
    (void)hipStreamSynchronize(s);
    dfree(rec0);
    dfree(iota);
    dfree(k16_all);
    return r;
}

// iota: 0..nnz-1 on the device (the sort's input values), or null to make it;
// keys16: the mode's indices as 16-bit keys (with rec0; at most 2^16 rows),
// or null to sort the 32-bit indices
static int coo_prepare_from(sptolDevCoo h, uint32_t mode, hipStream_t s, const uint4 *rec0, const uint32_t *iota,
                            const uint16_t *keys16)
{
    if (mode >= h->nmodes) SPTOL_CHECK(SPTERR_VALUE_ERROR, "DevCoo Prepare", "mode out of range");
    SortedCopy &sc = h->sorted[mode];
    if (sc.ready || h->nnz == 0) { sc.ready = true; return 0; }
    StreamScope scope(s);   // everything freed here was used on s only
    if (h->nnz > kMaxSortItems)
        SPTOL_CHECK(SPTERR_VALUE_ERROR, "DevCoo Prepare", "nnz > 2^31 - 1 (the device sort counts items in 32-bit ints)");
    const uint64_t n = h->nnz;
    uint32_t *perm = nullptr, *perm2 = nullptr, *keys2 = nullptr;
    void *tmp = nullptr;
    size_t tmp_bytes = 0;
    int r = 0;
    if ((h->nmodes == 3 || h->nmodes == 4) && h->in_order[mode]) {
        // already in mode order: the records are packed in the tensor's own
        // order, which a stable sort would leave unchanged
        SPTOL_TRY(seg_build(sc.seg, h->nmodes, mode, h->ndims[mode], n, h->inds, h->vals, true, s, h->ndims, nullptr,
                            nullptr, nullptr, nullptr, !h->batch_only));
        sc.ready = true;
        return 0;
    }
    // 16-bit keys need the packed records (the view is built by gathering them)
    const bool k16 = keys16 && rec0 && (h->nmodes == 3 || h->nmodes == 4) && h->ndims[mode] <= 65536u;
    auto cleanup = [&]() { dfree(perm); dfree(perm2); dfree(keys2); dfree(tmp); };
    if ((!iota && (r = dmalloc((void **)&perm, n * 4, "DevCoo Prepare"))) ||
        (r = dmalloc((void **)&perm2, n * 4, "DevCoo Prepare")) ||
        (r = dmalloc((void **)&keys2, n * (k16 ? 2 : 4), "DevCoo Prepare"))) { cleanup(); return r; }
    if (!iota) hipLaunchKernelGGL(k_iota, dim3(grid_for(n)), dim3(256), 0, s, perm, n);
    const uint32_t *vin = iota ? iota : perm;
    uint16_t *keys2_16 = reinterpret_cast<uint16_t *>(keys2);
    int endbit = 1;
    while (endbit < 32 && (1ull << endbit) < h->ndims[mode]) ++endbit;
    // stable LSD radix sort of (row, position): the two calls size, then run
    auto sort = [&](void *t, size_t &bytes) {
        return k16 ? hipcub::DeviceRadixSort::SortPairs(t, bytes, keys16, keys2_16, vin, perm2, (int)n, 0, endbit, s)
                   : hipcub::DeviceRadixSort::SortPairs(t, bytes, h->inds[mode], keys2, vin, perm2, (int)n, 0,
                                                        endbit, s);
    };
    hipError_t e = sort(nullptr, tmp_bytes);
    if (e == hipSuccess) {
        if ((r = dmalloc(&tmp, tmp_bytes ? tmp_bytes : 4, "DevCoo Prepare"))) { cleanup(); return r; }
        e = sort(tmp, tmp_bytes);
    }
    if (e != hipSuccess) { cleanup(); SPTOL_HIP(e, "DevCoo Prepare"); }
    if (h->nmodes == 3 || h->nmodes == 4) {
        // 3- and 4-mode tensors: the packed records are gathered straight
        // through the permutation; no sorted SoA copy is kept (the atomic
        // fallback kernel reads the tensor in its own order)
        r = seg_build(sc.seg, h->nmodes, mode, h->ndims[mode], n, h->inds, h->vals, true, s, h->ndims, perm2,
                      k16 ? nullptr : keys2, rec0, k16 ? keys2_16 : nullptr, !h->batch_only);
        cleanup();
        if (r) return r;
        sc.ready = true;
        return 0;
    }
    for (uint32_t m = 0; m < h->nmodes && !r; ++m) {
        r = dmalloc((void **)&sc.inds[m], n * 4, "DevCoo Prepare");
        if (!r) hipLaunchKernelGGL(k_gather<uint32_t>, dim3(grid_for(n)), dim3(256), 0, s, sc.inds[m], h->inds[m], perm2, n);
    }
    if (!r) r = dmalloc((void **)&sc.vals, n * 4, "DevCoo Prepare");
    if (!r) hipLaunchKernelGGL(k_gather<float>, dim3(grid_for(n)), dim3(256), 0, s, sc.vals, h->vals, perm2, n);
    hipError_t e2 = hipStreamSynchronize(s);
    cleanup();
    if (r) {
        for (uint32_t m = 0; m < h->nmodes; ++m) { dfree(sc.inds[m]); sc.inds[m] = nullptr; }
        dfree(sc.vals); sc.vals = nullptr;
        sc.seg.release();
        return r;
    }
    SPTOL_HIP(e2, "DevCoo Prepare");
    sc.ready = true;
    return 0;
}

}  // namespace sptol

extern "C" {

int sptolDevCooCreate(sptolDevCoo *h, const sptSparseTensor *X) { return coo_create(h, X, nullptr); }

int sptolDevCooLoadBinary(sptolDevCoo *h, const char *path, void *stream)
{
    return coo_load_binary(h, path, stream_of(stream));
}

int sptolDevCooPrepareMode(sptolDevCoo h, sptIndex mode) { return coo_prepare(h, mode, nullptr); }

void sptolDevCooFree(sptolDevCoo h)
{
    if (!h) return;
    for (uint32_t m = 0; m < kMaxModes; ++m) {
        dfree(h->inds[m]);
        for (uint32_t q = 0; q < kMaxModes; ++q) dfree(h->sorted[m].inds[q]);
        dfree(h->sorted[m].vals);
        h->sorted[m].seg.release();
    }
    dfree(h->vals);
    delete h;
}

sptNnzIndex sptolDevCooNnz(sptolDevCoo h) { return h ? h->nnz : 0; }

static int check_dev_mats(uint32_t nmodes, const uint32_t *ndims, const sptolDevMatrix *mats,
                          const sptIndex mats_order[], sptIndex mode, const sptolDevMatrix *out,
                          const char *module)
{
    if (mode >= nmodes) SPTOL_CHECK(SPTERR_VALUE_ERROR, module, "mode out of range");
    const int dt = mats[0].dtype;
    for (uint32_t m = 0; m < nmodes; ++m) {
        if (mats[m].ncols != out->ncols) SPTOL_CHECK(SPTERR_SHAPE_MISMATCH, module, "mats[i]->cols != mats[nmodes]->ncols");
        if (mats[m].nrows != ndims[m]) SPTOL_CHECK(SPTERR_SHAPE_MISMATCH, module, "mats[i]->nrows != ndims[i]");
        if (mats[m].stride != out->stride) SPTOL_CHECK(SPTERR_SHAPE_MISMATCH, module, "factor strides differ");
        if (mats[m].dtype != dt) SPTOL_CHECK(SPTERR_VALUE_ERROR, module, "factors mix dtypes");
    }
    if (out->dtype != SPTOL_F32) SPTOL_CHECK(SPTERR_VALUE_ERROR, module, "output must be FP32");
    if (out->nrows < ndims[mode]) SPTOL_CHECK(SPTERR_SHAPE_MISMATCH, module, "output has too few rows");
    uint32_t seen = 0;
    for (uint32_t k = 1; k < nmodes; ++k) {
        if (mats_order[k] >= nmodes || mats_order[k] == mode || (seen >> mats_order[k]) & 1u)
            SPTOL_CHECK(SPTERR_VALUE_ERROR, module, "mats_order[1..N-1] must list the other modes");
        seen |= 1u << mats_order[k];
    }
    return 0;
}

}  // extern "C"

int sptol::coo_mttkrp(sptolDevCoo X, const sptolDevMatrix *mats, const uint32_t *mats_order, uint32_t mode,
                      sptolDevMatrix *out, hipStream_t s, const AlsFactorReq *fr, bool *factored)
{
    if (factored) *factored = false;
    SortedCopy &sc = X->sorted[mode];
    if (sc.ready && sc.seg.ready && seg_ok(X->nmodes, out->stride)) {
        const void *mp[kMaxModes];
        for (uint32_t m = 0; m < X->nmodes; ++m) mp[m] = mats[m].values;
        return seg_mttkrp(sc.seg, mp, mats_order, out->ncols, out->stride, mats[0].dtype, (float *)out->values, s,
                          fr, factored);
    }
    MttkrpArgs a;
    std::memset(&a, 0, sizeof(a));
    a.nnz = X->nnz;
    a.nmodes = X->nmodes;
    a.mode = mode;
    a.R = out->ncols;
    a.stride = out->stride;
    for (uint32_t m = 0; m < X->nmodes; ++m) {
        a.order[m] = mats_order[m];
        a.inds[m] = sc.vals ? sc.inds[m] : X->inds[m];
        a.mats[m] = mats[m].values;
    }
    a.vals = sc.vals ? sc.vals : X->vals;
    a.out = (float *)out->values;
    SPTOL_TRY(launch_fill_zero(a.out, (uint64_t)X->ndims[mode] * out->stride, s));
    return launch_mttkrp_coo(a, mats[0].dtype, s);
}

extern "C" {

int sptolDevMTTKRP(sptolDevCoo X, const sptolDevMatrix *mats, const sptIndex mats_order[], sptIndex mode,
                   sptolDevMatrix *out, void *stream)
{
    SPTOL_TRY(check_dev_mats(X->nmodes, X->ndims, mats, mats_order, mode, out, "DevCoo MTTKRP"));
    return coo_mttkrp(X, mats, mats_order, mode, out, stream_of(stream));
}

// sptolDevMTTKRP without any device allocation: a prepared mode runs the
// segmented kernels with their carry space and work counter in the
// library's fixed scratch array; an unprepared one runs the unsorted kernel,
// which needs no scratch.
int sptolDevMTTKRPNoAlloc(sptolDevCoo X, const sptolDevMatrix *mats, const sptIndex mats_order[], sptIndex mode,
                          sptolDevMatrix *out, void *stream)
{
    SPTOL_TRY(check_dev_mats(X->nmodes, X->ndims, mats, mats_order, mode, out, "DevCoo MTTKRP no-alloc"));
    hipStream_t s = stream_of(stream);
    SortedCopy &sc = X->sorted[mode];
    if (sc.ready && sc.seg.ready && seg_ok(X->nmodes, out->stride)) {
        const void *mp[kMaxModes];
        for (uint32_t m = 0; m < X->nmodes; ++m) mp[m] = mats[m].values;
        return seg_mttkrp_noalloc(sc.seg, mp, mats_order, out->ncols, out->stride, mats[0].dtype,
                                  (float *)out->values, s);
    }
    return coo_mttkrp(X, mats, mats_order, mode, out, s);
}

int sptolDevHicooCreate(sptolDevHicoo *h, const sptSparseTensorHiCOO *X)
{
    *h = nullptr;
    if (X->nmodes < 2 || X->nmodes > kMaxModes)
        SPTOL_CHECK(SPTERR_VALUE_ERROR, "DevHiCOO New", "nmodes must be in [2, 8]");
    if (X->bptr.len < 1) SPTOL_CHECK(SPTERR_VALUE_ERROR, "DevHiCOO New", "empty block pointer");
    {
        // Structural and range checks: block pointers monotone and ending at
        // nnz, and every reconstructed coordinate below ndims.
        const uint64_t nb = X->bptr.len - 1;
        if (X->bptr.data[0] != 0 || X->bptr.data[nb] != X->nnz)
            SPTOL_CHECK(SPTERR_VALUE_ERROR, "DevHiCOO New", "bptr must run from 0 to nnz");
        for (uint32_t m = 0; m < X->nmodes; ++m)
            if (X->binds[m].len < nb || X->einds[m].len < X->nnz)
                SPTOL_CHECK(SPTERR_VALUE_ERROR, "DevHiCOO New", "binds/einds shorter than the tensor");
        for (uint64_t b = 0; b < nb; ++b) {
            const uint64_t z0 = X->bptr.data[b], z1 = X->bptr.data[b + 1];
            if (z1 <= z0) SPTOL_CHECK(SPTERR_VALUE_ERROR, "DevHiCOO New", "empty or unordered block");
            for (uint32_t m = 0; m < X->nmodes; ++m) {
                const uint64_t base = (uint64_t)X->binds[m].data[b] << X->sb_bits;
                uint32_t emax = 0;
                for (uint64_t z = z0; z < z1; ++z) emax = X->einds[m].data[z] > emax ? X->einds[m].data[z] : emax;
                if (emax >> X->sb_bits) SPTOL_CHECK(SPTERR_VALUE_ERROR, "DevHiCOO New", "element offset >= 2^sb_bits");
                if (base + emax >= X->ndims[m]) SPTOL_CHECK(SPTERR_VALUE_ERROR, "DevHiCOO New", "index >= ndims");
            }
        }
    }
    sptolDevHicoo t = new sptolDevHicoo_s();
    t->nmodes = X->nmodes;
    t->nnz = X->nnz;
    t->nb = X->bptr.len - 1;
    t->nk = X->kptr.len ? X->kptr.len - 1 : 0;
    t->sb_bits = X->sb_bits;
    t->sk_bits = X->sk_bits;
    for (uint32_t m = 0; m < t->nmodes; ++m) t->ndims[m] = X->ndims[m];
    int r = sptolHipDuplicateMemory((void **)&t->bptr, X->bptr.data, X->bptr.len * sizeof(uint64_t));
    for (uint32_t m = 0; m < t->nmodes && !r; ++m) {
        r = sptolHipDuplicateMemory((void **)&t->binds[m], X->binds[m].data, t->nb * sizeof(uint32_t));
        if (!r) r = sptolHipDuplicateMemory((void **)&t->einds[m], X->einds[m].data, X->nnz);
    }
    if (!r) r = sptolHipDuplicateMemory((void **)&t->vals, X->values.data, X->nnz * sizeof(float));
    if (r) { sptolDevHicooFree(t); return r; }
    *h = t;
    return 0;
}

void sptolDevHicooFree(sptolDevHicoo h)
{
    if (!h) return;
    free_hicoo_views(h);
    dfree(h->bptr);
    for (uint32_t m = 0; m < kMaxModes; ++m) { dfree(h->binds[m]); dfree(h->einds[m]); }
    dfree(h->vals);
    delete h;
}

int sptolDevMTTKRPHiCOO(sptolDevHicoo X, const sptolDevMatrix *mats, const sptIndex mats_order[], sptIndex mode,
                        sptolDevMatrix *out, void *stream)
{
    SPTOL_TRY(check_dev_mats(X->nmodes, X->ndims, mats, mats_order, mode, out, "DevHiCOO MTTKRP"));
    return hicoo_mttkrp(X, mats, mats_order, mode, out, stream_of(stream));
}

}  // extern "C"

int sptol::hicoo_mttkrp(sptolDevHicoo X, const sptolDevMatrix *mats, const uint32_t *mats_order, uint32_t mode,
                        sptolDevMatrix *out, hipStream_t s, const AlsFactorReq *fr, bool *factored)
{
    if (factored) *factored = false;
    MttkrpArgs a;
    std::memset(&a, 0, sizeof(a));
    a.nnz = X->nnz;
    a.nmodes = X->nmodes;
    a.mode = mode;
    a.R = out->ncols;
    a.stride = out->stride;
    for (uint32_t m = 0; m < X->nmodes; ++m) {
        a.order[m] = mats_order[m];
        a.mats[m] = mats[m].values;
        a.binds[m] = X->binds[m];
        a.einds[m] = X->einds[m];
    }
    a.vals = X->vals;
    a.bptr = X->bptr;
    a.nb = X->nb;
    a.sb_bits = X->sb_bits;
    a.out = (float *)out->values;
    int used = 0;
    SPTOL_TRY(launch_hicoo_view(X, a, mats[0].dtype, s, &used, fr, factored));
    if (used) return 0;
    SPTOL_TRY(launch_fill_zero(a.out, (uint64_t)X->ndims[mode] * out->stride, s));
    return launch_mttkrp_hicoo(a, mats[0].dtype, s);
}

// ---------------------------------------------------------------------------
// ParTI!-compatible host entry points: upload, run on the device, download.
// ---------------------------------------------------------------------------
static int check_host_mats(sptIndex nmodes, const sptIndex *ndims, sptMatrix *mats[], sptIndex mode,
                           const char *module)
{
    if (nmodes < 2 || nmodes > kMaxModes) SPTOL_CHECK(SPTERR_VALUE_ERROR, module, "nmodes must be in [2, 8]");
    if (mode >= nmodes) SPTOL_CHECK(SPTERR_VALUE_ERROR, module, "mode out of range");
    for (sptIndex i = 0; i < nmodes; ++i) {
        if (mats[i]->ncols != mats[nmodes]->ncols)
            SPTOL_CHECK(SPTERR_SHAPE_MISMATCH, module, "mats[i]->cols != mats[nmodes]->ncols");
        if (mats[i]->nrows != ndims[i])
            SPTOL_CHECK(SPTERR_SHAPE_MISMATCH, module, "mats[i]->nrows != ndims[i]");
        if (mats[i]->stride != mats[0]->stride)
            SPTOL_CHECK(SPTERR_SHAPE_MISMATCH, module, "factor strides differ");
    }
    if (mats[nmodes]->stride != mats[0]->stride || mats[nmodes]->cap < ndims[mode])
        SPTOL_CHECK(SPTERR_SHAPE_MISMATCH, module, "output matrix too small");
    return 0;
}

// Upload factors, run `body(dev_mats, dev_out)`, download the result into
// mats[nmodes] (rows = ndims[mode]).
template <typename F>
static int with_device_factors(sptIndex nmodes, const sptIndex *ndims, sptMatrix *mats[], sptIndex mode,
                               F body)
{
    sptolDevMatrix dm[kMaxModes];
    sptolDevMatrix dout;
    std::memset(dm, 0, sizeof(dm));
    std::memset(&dout, 0, sizeof(dout));
    int r = 0;
    for (sptIndex m = 0; m < nmodes && !r; ++m) {
        r = sptolDevMatrixCreate(&dm[m], ndims[m], mats[m]->ncols, SPTOL_F32);
        if (!r && m != mode) r = sptolDevMatrixUpload(&dm[m], mats[m]);
    }
    if (!r) r = sptolDevMatrixCreate(&dout, ndims[mode], mats[nmodes]->ncols, SPTOL_F32);
    if (!r) r = body(dm, &dout);
    if (!r) r = sptolHipStreamSynchronize(nullptr);
    if (!r) {
        sptIndex keep = mats[nmodes]->nrows;
        mats[nmodes]->nrows = ndims[mode];
        dout.nrows = ndims[mode];
        r = sptolDevMatrixDownload(mats[nmodes], &dout);
        mats[nmodes]->nrows = keep;
    }
    for (sptIndex m = 0; m < nmodes; ++m) sptolDevMatrixFree(&dm[m]);
    sptolDevMatrixFree(&dout);
    return r;
}

extern "C" {

int sptMTTKRP(sptSparseTensor const * const X, sptMatrix * mats[], sptIndex const mats_order[], sptIndex const mode)
{
    SPTOL_TRY(check_host_mats(X->nmodes, X->ndims, mats, mode, "GPU SpTns MTTKRP"));
    auto t0 = std::chrono::steady_clock::now();
    sptolDevCoo dx = nullptr;
    SPTOL_TRY(sptolDevCooCreate(&dx, X));
    int r = with_device_factors(X->nmodes, X->ndims, mats, mode, [&](sptolDevMatrix *dm, sptolDevMatrix *dout) {
        return sptolDevMTTKRP(dx, dm, mats_order, mode, dout, nullptr);
    });
    sptolDevCooFree(dx);
    if (!r && verbose()) {
        double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        fprintf(stdout, "[GPU SpTns MTTKRP]: %.9lf s\n", sec);
    }
    return r;
}

int sptOmpMTTKRP(sptSparseTensor const * const X, sptMatrix * mats[], sptIndex const mats_order[],
                 sptIndex const mode, const int tk)
{
    (void)tk;
    return sptMTTKRP(X, mats, mats_order, mode);
}

int sptMTTKRPHiCOO(sptSparseTensorHiCOO const * const hitsr, sptMatrix * mats[], sptIndex const mats_order[],
                   sptIndex const mode)
{
    SPTOL_TRY(check_host_mats(hitsr->nmodes, hitsr->ndims, mats, mode, "GPU HiCOO SpTns MTTKRP"));
    sptolDevHicoo dx = nullptr;
    SPTOL_TRY(sptolDevHicooCreate(&dx, hitsr));
    // the mode's view is built with device sorts (at most 2^31 - 1 items
    // each); a larger tensor runs the block-order kernel instead
    const uint64_t nb = hitsr->bptr.len ? hitsr->bptr.len - 1 : 0;
    int r = 0;
    if (hitsr->nnz <= kMaxSortItems && nb <= kMaxSortItems) r = sptolDevHicooPrepareMode(dx, mode);
    if (r) { sptolDevHicooFree(dx); return r; }
    r = with_device_factors(hitsr->nmodes, hitsr->ndims, mats, mode, [&](sptolDevMatrix *dm, sptolDevMatrix *dout) {
        return sptolDevMTTKRPHiCOO(dx, dm, mats_order, mode, dout, nullptr);
    });
    sptolDevHicooFree(dx);
    return r;
}

int sptOmpMTTKRPHiCOO(sptSparseTensorHiCOO const * const hitsr, sptMatrix * mats[], sptIndex const mats_order[],
                      sptIndex const mode, const int nt)
{
    (void)nt;
    return sptMTTKRPHiCOO(hitsr, mats, mats_order, mode);
}

// sptRankMatrix carries 8-bit ncols/stride; view each as an sptMatrix.
int sptMTTKRPHiCOO_MatrixTiling(sptSparseTensorHiCOO const * const hitsr, sptRankMatrix * mats[],
                                sptIndex const mats_order[], sptIndex const mode)
{
    sptIndex N = hitsr->nmodes;
    if (N < 2 || N > kMaxModes) SPTOL_CHECK(SPTERR_VALUE_ERROR, "GPU HiCOO SpTns MTTKRP", "nmodes must be in [2, 8]");
    sptMatrix view[kMaxModes + 1];
    sptMatrix *pv[kMaxModes + 1];
    for (sptIndex m = 0; m <= N; ++m) {
        view[m].nrows = mats[m]->nrows;
        view[m].ncols = mats[m]->ncols;
        view[m].cap = mats[m]->cap;
        view[m].stride = mats[m]->stride;
        view[m].values = mats[m]->values;
        pv[m] = &view[m];
    }
    return sptMTTKRPHiCOO(hitsr, pv, mats_order, mode);
}

// The reference's OpenMP variants keep their signatures and run the same
// device path; thread counts, reduction buffers and the balanced flag have
// no meaning on the GPU (mttkrp_omp.c:401, hicoo/mttkrp_omp.c:134-207).
int sptOmpMTTKRP_Reduce(sptSparseTensor const * const X, sptMatrix * mats[], sptMatrix * copy_mats[],
                        sptIndex const mats_order[], sptIndex const mode, const int tk)
{
    (void)copy_mats;
    (void)tk;
    return sptMTTKRP(X, mats, mats_order, mode);
}

int sptOmpMTTKRPHiCOO_MatrixTiling(sptSparseTensorHiCOO const * const hitsr, sptRankMatrix * mats[],
                                   sptIndex const mats_order[], sptIndex const mode, const int nt)
{
    (void)nt;
    return sptMTTKRPHiCOO_MatrixTiling(hitsr, mats, mats_order, mode);
}

int sptOmpMTTKRPHiCOO_MatrixTiling_Scheduled(sptSparseTensorHiCOO const * const hitsr, sptRankMatrix * mats[],
                                             sptIndex const mats_order[], sptIndex const mode, const int tk,
                                             int balanced)
{
    (void)tk;
    (void)balanced;
    return sptMTTKRPHiCOO_MatrixTiling(hitsr, mats, mats_order, mode);
}

int sptOmpMTTKRPHiCOO_MatrixTiling_Scheduled_Reduce(sptSparseTensorHiCOO const * const hitsr, sptRankMatrix * mats[],
                                                    sptRankMatrix * copy_mats[], sptIndex const mats_order[],
                                                    sptIndex const mode, const int tk, int balanced)
{
    (void)copy_mats;
    (void)tk;
    (void)balanced;
    return sptMTTKRPHiCOO_MatrixTiling(hitsr, mats, mats_order, mode);
}

}  // extern "C"
