// CPD-ALS with tensor and factors resident on the GPU.
//
// Follows the reference loop src/sptensor/cpd.c:32-137 step for step:
//   for it: for m: M = MTTKRP_m(X, A) with mats_order = [m, m+1, ...] mod N
//               A_m = M V^{-1},  V = Hadamard_{p != m} A_p^T A_p   (matrix_solver.c:32-83)
//               lambda = column 2-norms (it == 0) or max(column max, 1) (cpd.c:99-103)
//               A_m /= lambda; ata[m] = A_m^T A_m
//           fit = 1 - sqrt(||X||^2 + ||K||^2 - 2<X,K>) / ||X||          (ktensor.c:102-118)
//   final: every A_m renormalised by its column 2-norms, folded into lambda (matrix.c:557-573)
// Differences, all deliberate:
//   * Gram matrices, the Hadamard product, the Cholesky factorisation, the
//     row solves, the norms and the fit are FP64 (the reference uses FP32
//     LAPACK); factors stay FP32, optionally with a BF16 copy for the MTTKRP
//     gathers.
//   * ||K||^2 uses the full symmetric Hadamard of the Grams; the reference
//     reads an unwritten triangle (ktensor.c:151).
//   * Reductions are two-stage with a fixed block count, so the solve, norm
//     and fit steps are bitwise reproducible for a given MTTKRP result.
#include "cpd_job.hpp"
#include "als_factor.hpp"
#include <algorithm>
#include <vector>
#include <cmath>
#include <cstring>
#include <cstdio>
#include <chrono>

namespace sptol {


// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// The kernels below are written as bodies taking the block's index and the
// block count of its request (blk, nblk), so that the single-request
// launches and the batched launches of the server (one grid over many
// requests) run the same arithmetic in the same order.

// partial[blk][a*R+b] = sum over this block's rows of A[i][a]*A[i][b] (FP64).
__device__ __forceinline__ void gram_partial_body(const float *__restrict__ A, uint32_t I, uint32_t R, uint32_t S,
                                                  double *__restrict__ partial, uint32_t blk, uint32_t nblk)
{
    __shared__ float tile[16][kMaxRank];
    const uint32_t RR = R * R;
    double acc[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = 0.0;
    const uint64_t rows_per = ((uint64_t)I + nblk - 1) / nblk;
    const uint64_t r0 = (uint64_t)blk * rows_per;
    const uint64_t r1 = min<uint64_t>(r0 + rows_per, I);
    for (uint64_t base = r0; base < r1; base += 16) {
        for (uint32_t t = threadIdx.x; t < 16 * R; t += 256) {
            const uint32_t rr = t / R, c = t % R;
            tile[rr][c] = (base + rr < r1) ? A[(base + rr) * S + c] : 0.0f;
        }
        __syncthreads();
        const uint32_t nrow = (uint32_t)min<uint64_t>(16, r1 - base);
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const uint32_t e = threadIdx.x + 256u * q;
            if (e < RR) {
                const uint32_t a = e / R, b = e % R;
                double s = acc[q];
                for (uint32_t rr = 0; rr < nrow; ++rr) s += (double)tile[rr][a] * (double)tile[rr][b];
                acc[q] = s;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const uint32_t e = threadIdx.x + 256u * q;
        if (e < RR) partial[(size_t)blk * RR + e] = acc[q];
    }
}

__global__ __launch_bounds__(256) void k_gram_partial(const float *__restrict__ A, uint32_t I, uint32_t R,
                                                      uint32_t S, double *__restrict__ partial)
{
    gram_partial_body(A, I, R, S, partial, blockIdx.x, gridDim.x);
}

// Wave-wide reduction of one value per lane in a fixed pattern.
__device__ __forceinline__ double wave_sum(double v)
{
#pragma clang fp contract(off)
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ double wave_max(double v)
{
    for (int o = 32; o > 0; o >>= 1) {
        const double y = __shfl_xor(v, o);
        v = y > v ? y : v;
    }
    return v;
}

// out[e] = sum over blk of partial[blk][e]: one wave per output, lanes
// stride the blocks, then a fixed butterfly.  Four outputs per block.
__device__ __forceinline__ void sum_partials_body(const double *__restrict__ partial, uint32_t nblk, uint32_t n,
                                                  double *__restrict__ out, uint32_t blk)
{
    const uint32_t e = blk * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63u;
    if (e >= n) return;   // whole wave
    double s = 0.0;
    for (uint32_t b = lane; b < nblk; b += 64) s += partial[(size_t)b * n + e];
    s = wave_sum(s);
    if (lane == 0) out[e] = s;
}

__global__ __launch_bounds__(256) void k_sum_partials(const double *__restrict__ partial, uint32_t nblk, uint32_t n,
                                                      double *__restrict__ out)
{
    sum_partials_body(partial, nblk, n, out, blockIdx.x);
}

static unsigned sum_grid(uint32_t n) { return (n + 3) / 4; }
// Column reduction partials: kind 0 = sum of squares, 1 = max (from 0).
__device__ __forceinline__ void col_partial_body(const float *__restrict__ A, uint32_t I, uint32_t R, uint32_t S,
                                                 int kind, double *__restrict__ partial, uint32_t blk, uint32_t nblk)
{
    __shared__ double red[256];
    const uint32_t lanes_per_col = 256 / R;          // R <= 64 so >= 4
    const uint32_t c = threadIdx.x % R;
    const uint32_t sub = threadIdx.x / R;
    const bool active = sub < lanes_per_col;
    const uint64_t rows_per = ((uint64_t)I + nblk - 1) / nblk;
    const uint64_t r0 = (uint64_t)blk * rows_per;
    const uint64_t r1 = min<uint64_t>(r0 + rows_per, I);
    double v = 0.0;
    if (active) {
        for (uint64_t i = r0 + sub; i < r1; i += lanes_per_col) {
            const double x = (double)A[i * S + c];
            if (kind == 0) v += x * x;
            else v = x > v ? x : v;
        }
    }
    red[threadIdx.x] = v;
    __syncthreads();
    if (threadIdx.x < R) {
        double s = red[threadIdx.x];
        for (uint32_t k = 1; k < lanes_per_col; ++k) {
            const double y = red[threadIdx.x + k * R];
            if (kind == 0) s += y;
            else s = y > s ? y : s;
        }
        partial[(size_t)blk * R + threadIdx.x] = s;
    }
}

__global__ __launch_bounds__(256) void k_col_partial(const float *__restrict__ A, uint32_t I, uint32_t R, uint32_t S,
                                                     int kind, double *__restrict__ partial)
{
    col_partial_body(A, I, R, S, kind, partial, blockIdx.x, gridDim.x);
}

// lambda[c] from the partials: kind 0 -> sqrt(sum), 1 -> max(max, 1).
// mult != 0 multiplies the existing lambda by the new norm instead
// (GetFinalLambda); the norms are also written to norms[].  One wave per
// column, four columns per block.
__device__ __forceinline__ void col_final_body(const double *__restrict__ partial, uint32_t nblk, uint32_t R, int kind,
                                               int mult, double *__restrict__ lambda, double *__restrict__ norms,
                                               uint32_t blk)
{
    const uint32_t c = blk * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63u;
    if (c >= R) return;   // whole wave
    double s = 0.0;
    for (uint32_t b = lane; b < nblk; b += 64) {
        const double y = partial[(size_t)b * R + c];
        if (kind == 0) s += y;
        else s = y > s ? y : s;
    }
    s = kind == 0 ? wave_sum(s) : wave_max(s);
    if (lane == 0) {
        const double nrm = kind == 0 ? sqrt(s) : (s < 1.0 ? 1.0 : s);
        norms[c] = nrm;
        lambda[c] = mult ? lambda[c] * nrm : nrm;
    }
}

__global__ __launch_bounds__(256) void k_col_final(const double *__restrict__ partial, uint32_t nblk, uint32_t R, int kind,
                                                   int mult, double *__restrict__ lambda, double *__restrict__ norms)
{
    col_final_body(partial, nblk, R, kind, mult, lambda, norms, blockIdx.x);
}

// A[i][c] /= norms[c] (FP64 divide, FP32 store); optional BF16 copy.
__device__ __forceinline__ void scale_body(float *__restrict__ A, uint16_t *__restrict__ Abf, uint32_t I, uint32_t R,
                                           uint32_t S, const double *__restrict__ norms, uint32_t blk, uint32_t nblk)
{
    const uint64_t n = (uint64_t)I * S;
    uint64_t e = (uint64_t)blk * blockDim.x + threadIdx.x;
    const uint64_t step = (uint64_t)nblk * blockDim.x;
    for (; e < n; e += step) {
        const uint32_t c = (uint32_t)(e % S);
        float x = A[e];
        if (c < R) {
            x = (float)((double)x / norms[c]);
            A[e] = x;
        }
        if (Abf) {
            uint32_t u = __float_as_uint(x);
            u += 0x7FFFu + ((u >> 16) & 1u);
            Abf[e] = (uint16_t)(u >> 16);
        }
    }
}

__global__ void k_scale(float *__restrict__ A, uint16_t *__restrict__ Abf, uint32_t I, uint32_t R, uint32_t S,
                        const double *__restrict__ norms)
{
    scale_body(A, Abf, I, R, S, norms, blockIdx.x, gridDim.x);
}

// One launch (single request: k_norm_scale_gram; batched: k_norm_scale_gram_b)
// in place of k_col_final, k_scale and k_gram_partial.  Every block first
// forms all R column norms from k_col_partial's ncb partials (thread t sums
// partials t / R, t / R + 256 / R, ... of column t % R, then the column's
// sums are added in thread order: a fixed order, all loads issued at once;
// block 0 also stores lambda and norms), then scales its rows of A (the
// rows gram_partial_body gives it; FP64 divide, FP32 store, BF16 copy of
// every column, as scale_body) and sums the Gram partials of the scaled
// rows (as gram_partial_body).
__device__ __forceinline__ void norm_scale_gram_body(float *__restrict__ A, uint16_t *__restrict__ Abf, uint32_t I,
                                                     uint32_t R, uint32_t S, int kind,
                                                     const double *__restrict__ colpart, uint32_t ncb,
                                                     double *__restrict__ lambda, double *__restrict__ norms,
                                                     double *__restrict__ gpart, uint32_t blk, uint32_t nblk)
{
    __shared__ double red[256];
    __shared__ double nrm[kMaxRank];
    __shared__ float tile[16][kMaxRank];
    {
        const uint32_t nsub = 256 / R, c = threadIdx.x % R, sub = threadIdx.x / R;   // R <= 64: nsub >= 4
        double v = 0.0;
        if (sub < nsub)
            for (uint32_t b = sub; b < ncb; b += nsub) {
                const double y = colpart[(size_t)b * R + c];
                if (kind == 0) v += y;
                else v = y > v ? y : v;
            }
        red[threadIdx.x] = v;
        __syncthreads();
        if (threadIdx.x < R) {
            double t = red[threadIdx.x];
            for (uint32_t k = 1; k < nsub; ++k) {
                const double y = red[threadIdx.x + k * R];
                if (kind == 0) t += y;
                else t = y > t ? y : t;
            }
            const double n = kind == 0 ? sqrt(t) : (t < 1.0 ? 1.0 : t);
            nrm[threadIdx.x] = n;
            if (blk == 0) {
                norms[threadIdx.x] = n;
                lambda[threadIdx.x] = n;
            }
        }
        __syncthreads();
    }
    const uint32_t RR = R * R;
    double acc[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = 0.0;
    const uint64_t rows_per = ((uint64_t)I + nblk - 1) / nblk;
    const uint64_t r0 = (uint64_t)blk * rows_per;
    const uint64_t r1 = min<uint64_t>(r0 + rows_per, I);
    for (uint64_t base = r0; base < r1; base += 16) {
        for (uint32_t t = threadIdx.x; t < 16 * S; t += 256) {
            const uint32_t rr = t / S, c = t % S;
            float x = 0.0f;
            if (base + rr < r1) {
                const uint64_t e = (base + rr) * S + c;
                x = A[e];
                if (c < R) {
                    x = (float)((double)x / nrm[c]);
                    A[e] = x;
                }
                if (Abf) {
                    uint32_t u = __float_as_uint(x);
                    u += 0x7FFFu + ((u >> 16) & 1u);
                    Abf[e] = (uint16_t)(u >> 16);
                }
            }
            if (c < R) tile[rr][c] = x;
        }
        __syncthreads();
        const uint32_t nrow = (uint32_t)min<uint64_t>(16, r1 - base);
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const uint32_t e = threadIdx.x + 256u * q;
            if (e < RR) {
                const uint32_t a = e / R, b = e % R;
                double sum = acc[q];
                for (uint32_t rr = 0; rr < nrow; ++rr) sum += (double)tile[rr][a] * (double)tile[rr][b];
                acc[q] = sum;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const uint32_t e = threadIdx.x + 256u * q;
        if (e < RR) gpart[(size_t)blk * RR + e] = acc[q];
    }
}

__global__ __launch_bounds__(256) void k_norm_scale_gram(float *__restrict__ A, uint16_t *__restrict__ Abf, uint32_t I,
                                                         uint32_t R, uint32_t S, int kind,
                                                         const double *__restrict__ colpart, uint32_t ncb,
                                                         double *__restrict__ lambda, double *__restrict__ norms,
                                                         double *__restrict__ gpart)
{
    norm_scale_gram_body(A, Abf, I, R, S, kind, colpart, ncb, lambda, norms, gpart, blockIdx.x, gridDim.x);
}

template <int RM>
__global__ __launch_bounds__(64) void k_factor(const double *__restrict__ ata, uint32_t N, uint32_t m, uint32_t R,
                                               double *__restrict__ fac, int *__restrict__ piv, int *__restrict__ kind)
{
    factor_body<RM>(ata, N, m, R, fac, piv, kind);
}

// A[i] = M[i] V^{-1} for every row: V^{-1} (FP64, from k_factor) times each
// row of M, one thread per (row, column), a block holding 256 / RM rows;
// every product and sum in FP64 in a fixed column order.  When V was
// exactly singular (kind 2) the rows stay the MTTKRP's, as gesv leaves its
// right-hand sides.
template <int RM>
__device__ __forceinline__ void solve_rows_body(const float *__restrict__ M, float *__restrict__ A, uint32_t I,
                                                uint32_t R, uint32_t S, const double *__restrict__ fac,
                                                const int *__restrict__ kind, uint32_t blk)
{
#pragma clang fp contract(off)
    constexpr uint32_t RPB = 256 / RM;
    __shared__ double FT[RM * RM];   // FT[k * RM + c] = V^{-1}[c][k]
    __shared__ float Ms[RPB * RM];
    const int kd = *kind;
    const uint32_t t = threadIdx.x, r = t / RM, c = t % RM;
    const uint64_t i = (uint64_t)blk * RPB + r;
    if (kd != 2)
        for (uint32_t e = t; e < R * R; e += blockDim.x) FT[(e % R) * RM + e / R] = fac[e];
    Ms[t] = (i < I && c < R) ? M[i * S + c] : 0.f;
    __syncthreads();
    if (i >= I || c >= R) return;
    if (kd == 2) {
        A[i * S + c] = Ms[t];
        return;
    }
    double x = 0.0;
#pragma unroll
    for (int k = 0; k < RM; ++k)
        if (k < (int)R) x += FT[k * RM + c] * (double)Ms[r * RM + k];
    A[i * S + c] = (float)x;
}

template <int RM>
__global__ __launch_bounds__(256) void k_solve_rows(const float *__restrict__ M, float *__restrict__ A, uint32_t I,
                                                    uint32_t R, uint32_t S, const double *__restrict__ fac,
                                                    const int *__restrict__ kind)
{
    solve_rows_body<RM>(M, A, I, R, S, fac, kind, blockIdx.x);
}

// partial[blk][r] = sum over this block's rows of A[i][r] * M[i][r].
__device__ __forceinline__ void inner_partial_body(const float *__restrict__ A, const float *__restrict__ M, uint32_t I,
                                                   uint32_t R, uint32_t S, double *__restrict__ partial, uint32_t blk,
                                                   uint32_t nblk)
{
    __shared__ double red[256];
    const uint32_t lanes_per_col = 256 / R;
    const uint32_t c = threadIdx.x % R;
    const uint32_t sub = threadIdx.x / R;
    const uint64_t rows_per = ((uint64_t)I + nblk - 1) / nblk;
    const uint64_t r0 = (uint64_t)blk * rows_per;
    const uint64_t r1 = min<uint64_t>(r0 + rows_per, I);
    double v = 0.0;
    if (sub < lanes_per_col)
        for (uint64_t i = r0 + sub; i < r1; i += lanes_per_col) v += (double)A[i * S + c] * (double)M[i * S + c];
    red[threadIdx.x] = v;
    __syncthreads();
    if (threadIdx.x < R) {
        double s = red[threadIdx.x];
        for (uint32_t k = 1; k < lanes_per_col; ++k) s += red[threadIdx.x + k * R];
        partial[(size_t)blk * R + threadIdx.x] = s;
    }
}

__global__ __launch_bounds__(256) void k_inner_partial(const float *__restrict__ A, const float *__restrict__ M, uint32_t I,
                                                       uint32_t R, uint32_t S, double *__restrict__ partial)
{
    inner_partial_body(A, M, I, R, S, partial, blockIdx.x, gridDim.x);
}

// <X, K> = sum_nnz x * sum_r lambda_r prod_m A_m[i_m][r], every product and
// sum in FP64 from the FP32 factors: no cancellation-prone reuse of an FP32
// MTTKRP.  One thread per nonzero; per-block partial sums (fixed order).
struct InnerArgs {   // 3- and 4-mode tensors
    const uint32_t *inds[4];
    const float *vals;
    const float *mats[4];
    uint64_t nnz;
    uint32_t N, R, S;
};

// Lane groups of LPN lanes take nonzeros kInnerU at a time (interleaved
// over the block's contiguous range, all gathers of a step issued before
// use); lane l holds columns 4l..4l+3 and gathers them with one 16-byte load
// per factor row.  Per lane: sum over its nonzeros of
// x * sum_{its columns} lambda_c * prod_m A_m[i_m][c], all in FP64; then a
// fixed-order block reduction into partial[block].
constexpr int kInnerU = 4;
constexpr uint32_t kInnerBlocks = 4096;

template <int LPN>
__global__ __launch_bounds__(256) void k_inner_exact_partial(const InnerArgs a, const double *__restrict__ lambda,
                                                             double *__restrict__ partial)
{
#pragma clang fp contract(off)
    constexpr uint32_t G = 256 / LPN;
    constexpr int U = kInnerU;
    __shared__ double red[256];
    const uint32_t g = threadIdx.x / LPN, l = threadIdx.x % LPN, c0 = 4 * l;
    double lam[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) lam[q] = c0 + q < a.R ? lambda[c0 + q] : 0.0;
    const uint32_t cc = c0 < a.S ? c0 : 0;
    const uint64_t per = (a.nnz + gridDim.x - 1) / gridDim.x;
    const uint64_t z0 = (uint64_t)blockIdx.x * per, z1 = min<uint64_t>(z0 + per, a.nnz);
    double s = 0.0;
    for (uint64_t zb = z0 + (uint64_t)g * U; zb < z1; zb += (uint64_t)G * U) {
        float4 f[U][4];
        float v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t z = zb + u < z1 ? zb + u : zb;   // zb < z1: in bounds
            v[u] = zb + u < z1 ? a.vals[z] : 0.0f;
#pragma unroll
            for (int m = 0; m < 4; ++m)
                if (m < (int)a.N) f[u][m] = *reinterpret_cast<const float4 *>(a.mats[m] + (size_t)a.inds[m][z] * a.S + cc);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            double p[4] = {lam[0], lam[1], lam[2], lam[3]};
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                if (m < (int)a.N) {
                    p[0] *= (double)f[u][m].x; p[1] *= (double)f[u][m].y;
                    p[2] *= (double)f[u][m].z; p[3] *= (double)f[u][m].w;
                }
            }
            // padding columns (>= R) may hold unspecified values: drop them
#pragma unroll
            for (int q = 0; q < 4; ++q) p[q] = c0 + q < a.R ? p[q] : 0.0;
            s += (double)v[u] * (((p[0] + p[1]) + p[2]) + p[3]);
        }
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// The same FP64 <X,K> from the packed records of one of the tensor's mode
// views (each record holds every mode's index: {i0, i1, i2, bits(v)} for 3
// modes, {i0, i1, i2, i3} with the view's value array for 4).  A lane group
// takes 8 records a step (against 4: 893 against 1052 us on a 50M-nonzero
// 4-D tensor at R = 32): its lanes load them with one 16-byte access
// each and pass indices and values to the group by shuffles, so a step
// issues one record load (and one value load) where the kernel above issues
// one index load per mode and nonzero.  Same per-lane sums and block
// reduction as above; the nonzeros come in the view's order.
template <int LPN, int NM>
__global__ __launch_bounds__(256) void k_inner_exact_rec(const uint4 *__restrict__ rec, const float *__restrict__ vals4,
                                                         const InnerArgs a, const double *__restrict__ lambda,
                                                         double *__restrict__ partial)
{
#pragma clang fp contract(off)
    constexpr uint32_t G = 256 / LPN;
    constexpr int U = 2 * kInnerU;
    constexpr int KL = (U + LPN - 1) / LPN;   // records a lane loads per step
    __shared__ double red[256];
    const uint32_t g = threadIdx.x / LPN, l = threadIdx.x % LPN, c0 = 4 * l;
    const int gbase = (int)((threadIdx.x & 63u) - l);   // the group's first lane in the wave
    double lam[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) lam[q] = c0 + q < a.R ? lambda[c0 + q] : 0.0;
    const uint32_t cc = c0 < a.S ? c0 : 0;
    const uint64_t per = (a.nnz + gridDim.x - 1) / gridDim.x;
    const uint64_t z0 = (uint64_t)blockIdx.x * per, z1 = min<uint64_t>(z0 + per, a.nnz);
    double s = 0.0;
    for (uint64_t zb = z0 + (uint64_t)g * U; zb < z1; zb += (uint64_t)G * U) {
        uint4 rk[KL];
        float vk[KL];
#pragma unroll
        for (int k = 0; k < KL; ++k) {
            // lanes past the step's U records load nothing
            const uint32_t idx = l + (uint32_t)k * LPN;
            const uint64_t z = min<uint64_t>(zb + idx, z1 - 1);   // zb < z1: in bounds
            rk[k] = idx < (uint32_t)U ? rec[z] : make_uint4(0u, 0u, 0u, 0u);
            vk[k] = idx < (uint32_t)U ? (NM == 4 ? vals4[z] : __uint_as_float(rk[k].w)) : 0.0f;
        }
        uint32_t j[U][4];
        float v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int src = gbase + u % LPN, k = u / LPN;
            j[u][0] = (uint32_t)__shfl((int)rk[k].x, src);
            j[u][1] = (uint32_t)__shfl((int)rk[k].y, src);
            j[u][2] = (uint32_t)__shfl((int)rk[k].z, src);
            j[u][3] = NM == 4 ? (uint32_t)__shfl((int)rk[k].w, src) : 0u;
            v[u] = zb + u < z1 ? __shfl(vk[k], src) : 0.0f;
        }
        float4 f[U][NM];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int m = 0; m < NM; ++m)
                f[u][m] = *reinterpret_cast<const float4 *>(a.mats[m] + (size_t)j[u][m] * a.S + cc);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            double p[4] = {lam[0], lam[1], lam[2], lam[3]};
#pragma unroll
            for (int m = 0; m < NM; ++m) {
                p[0] *= (double)f[u][m].x; p[1] *= (double)f[u][m].y;
                p[2] *= (double)f[u][m].z; p[3] *= (double)f[u][m].w;
            }
            // padding columns (>= R) may hold unspecified values: drop them
#pragma unroll
            for (int q = 0; q < 4; ++q) p[q] = c0 + q < a.R ? p[q] : 0.0;
            s += (double)v[u] * (((p[0] + p[1]) + p[2]) + p[3]);
        }
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

template <int NM>
static void launch_inner_exact_rec(const SegView &v, const InnerArgs &a, const double *lambda, double *partial,
                                   hipStream_t s)
{
    const dim3 g(kInnerBlocks), b(256);
    const float *v4 = NM == 4 ? v.vals : nullptr;
    if (a.S <= 8) hipLaunchKernelGGL((k_inner_exact_rec<2, NM>), g, b, 0, s, v.rec, v4, a, lambda, partial);
    else if (a.S <= 16) hipLaunchKernelGGL((k_inner_exact_rec<4, NM>), g, b, 0, s, v.rec, v4, a, lambda, partial);
    else if (a.S <= 32) hipLaunchKernelGGL((k_inner_exact_rec<8, NM>), g, b, 0, s, v.rec, v4, a, lambda, partial);
    else hipLaunchKernelGGL((k_inner_exact_rec<16, NM>), g, b, 0, s, v.rec, v4, a, lambda, partial);
}

// The same FP64 <X,K> for any order N <= kMaxModes, over a COO tensor (inds)
// or a HiCOO tensor (coordinate = (binds[m][b] << sb_bits) + einds[m][z] for
// the block b holding nonzero z).  Each lane group takes a contiguous run of
// its block's nonzeros (HiCOO: the run's first block by binary search over
// bptr, then walking forward), kInnerGenU nonzeros per step with all their
// gathers issued before use.
struct InnerGenArgs {
    const uint32_t *inds[kMaxModes];    // COO, or unused
    const uint64_t *bptr;               // HiCOO: nb + 1 offsets, or null
    const uint32_t *binds[kMaxModes];
    const uint8_t *einds[kMaxModes];
    uint64_t nb;
    uint32_t sb_bits;
    const float *vals;
    const float *mats[kMaxModes];
    uint64_t nnz;
    uint32_t N, R, S;
};
constexpr int kInnerGenU = 2;

template <int LPN, bool HICOO>
__global__ __launch_bounds__(256) void k_inner_exact_gen(const InnerGenArgs a, const double *__restrict__ lambda,
                                                         double *__restrict__ partial)
{
#pragma clang fp contract(off)
    constexpr uint32_t G = 256 / LPN;
    constexpr int U = kInnerGenU;
    __shared__ double red[256];
    const uint32_t g = threadIdx.x / LPN, l = threadIdx.x % LPN, c0 = 4 * l;
    double lam[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) lam[q] = c0 + q < a.R ? lambda[c0 + q] : 0.0;
    const uint32_t cc = c0 < a.S ? c0 : 0;
    const uint64_t per = (a.nnz + gridDim.x - 1) / gridDim.x;
    const uint64_t z0 = min<uint64_t>((uint64_t)blockIdx.x * per, a.nnz), z1 = min<uint64_t>(z0 + per, a.nnz);
    const uint64_t run = (z1 - z0 + G - 1) / G;
    const uint64_t y0 = min<uint64_t>(z0 + (uint64_t)g * run, z1), y1 = min<uint64_t>(y0 + run, z1);
    uint64_t b = 0;
    if (HICOO && y0 < y1) {   // last block whose first nonzero is <= y0
        uint64_t lo = 0, hi = a.nb;   // bptr[lo] <= y0 < bptr[hi]
        while (hi - lo > 1) {
            const uint64_t mid = lo + (hi - lo) / 2;
            if (a.bptr[mid] <= y0) lo = mid; else hi = mid;
        }
        b = lo;
    }
    double s = 0.0;
    for (uint64_t zb = y0; zb < y1; zb += U) {
        float4 f[U][kMaxModes];
        float v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const bool in = zb + u < y1;
            const uint64_t z = in ? zb + u : zb;   // zb < y1: in bounds
            v[u] = in ? a.vals[z] : 0.0f;
            if (HICOO) {
                while (a.bptr[b + 1] <= z) ++b;   // z < nnz = bptr[nb]: stops at b < nb
            }
#pragma unroll
            for (int m = 0; m < kMaxModes; ++m) {
                if (m < (int)a.N) {
                    const uint32_t i = HICOO ? (a.binds[m][b] << a.sb_bits) + (uint32_t)a.einds[m][z] : a.inds[m][z];
                    f[u][m] = *reinterpret_cast<const float4 *>(a.mats[m] + (size_t)i * a.S + cc);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            double p[4] = {lam[0], lam[1], lam[2], lam[3]};
#pragma unroll
            for (int m = 0; m < kMaxModes; ++m) {
                if (m < (int)a.N) {
                    p[0] *= (double)f[u][m].x; p[1] *= (double)f[u][m].y;
                    p[2] *= (double)f[u][m].z; p[3] *= (double)f[u][m].w;
                }
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) p[q] = c0 + q < a.R ? p[q] : 0.0;
            s += (double)v[u] * (((p[0] + p[1]) + p[2]) + p[3]);
        }
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

template <bool HICOO>
static void launch_inner_exact_gen(const InnerGenArgs &a, const double *lambda, double *partial, hipStream_t s)
{
    const dim3 g(kInnerBlocks), b(256);
    if (a.S <= 8) hipLaunchKernelGGL((k_inner_exact_gen<2, HICOO>), g, b, 0, s, a, lambda, partial);
    else if (a.S <= 16) hipLaunchKernelGGL((k_inner_exact_gen<4, HICOO>), g, b, 0, s, a, lambda, partial);
    else if (a.S <= 32) hipLaunchKernelGGL((k_inner_exact_gen<8, HICOO>), g, b, 0, s, a, lambda, partial);
    else hipLaunchKernelGGL((k_inner_exact_gen<16, HICOO>), g, b, 0, s, a, lambda, partial);
}

// Sum of squares of the values (||X||^2), two-stage.
__global__ __launch_bounds__(256) void k_sqnorm_partial(const float *__restrict__ v, uint64_t n, double *__restrict__ partial)
{
    __shared__ double red[256];
    const uint64_t per = (n + gridDim.x - 1) / gridDim.x;
    const uint64_t a = (uint64_t)blockIdx.x * per, b = min<uint64_t>(a + per, n);
    double s = 0.0;
    uint64_t i = a + threadIdx.x;
    // 8 loads in flight, added in the same order as one at a time
    for (; i + 7 * 256 < b; i += 8 * 256) {
        float x[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) x[k] = v[i + (uint64_t)k * 256];
#pragma unroll
        for (int k = 0; k < 8; ++k) s += (double)x[k] * (double)x[k];
    }
    for (; i < b; i += 256) s += (double)v[i] * (double)v[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// fit from ||X||^2, the Grams, lambda and the inner-product partials.
// inner_scalar = 1: inner_partial holds nblk scalar partials of <X,K>
// (exact pass); 0: nblk x R partials of sum_i A_last[i][r] M[i][r].
__device__ __forceinline__ void fit_body(const double *__restrict__ ata, uint32_t N, uint32_t R,
                                         const double *__restrict__ lambda, const double *__restrict__ inner_partial,
                                         uint32_t nblk, int inner_scalar, const double *__restrict__ normx,
                                         double *__restrict__ fit)
{
    __shared__ double red[256];
    const uint32_t RR = R * R;
    double s = 0.0;
    for (uint32_t e = threadIdx.x; e < RR; e += 256) {
        double g = 1.0;
        for (uint32_t m = 0; m < N; ++m) g *= ata[(size_t)m * RR + e];
        s += g * lambda[e / R] * lambda[e % R];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    const double normk = fabs(red[0]);
    __syncthreads();
    // inner product: per column r (or the scalar), its partials in block order
    __shared__ double col[kMaxRank];
    if (inner_scalar) {
        if (threadIdx.x == 0) {
            double acc = 0.0;
            for (uint32_t b = 0; b < nblk; ++b) acc += inner_partial[b];
            col[0] = acc;
        }
    } else if (threadIdx.x < R) {
        double acc = 0.0;
        for (uint32_t b = 0; b < nblk; ++b) acc += inner_partial[(size_t)b * R + threadIdx.x];
        col[threadIdx.x] = acc * lambda[threadIdx.x];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double inner = 0.0;
        if (inner_scalar) inner = col[0];
        else
            for (uint32_t r = 0; r < R; ++r) inner += col[r];
        double residual = *normx + normk - 2.0 * inner;
        if (residual > 0.0) residual = sqrt(residual);
        *fit = 1.0 - residual / sqrt(*normx);
    }
}

__global__ __launch_bounds__(256) void k_fit(const double *__restrict__ ata, uint32_t N, uint32_t R,
                                             const double *__restrict__ lambda, const double *__restrict__ inner_partial,
                                             uint32_t nblk, int inner_scalar, const double *__restrict__ normx,
                                             double *__restrict__ fit)
{
    fit_body(ata, N, R, lambda, inner_partial, nblk, inner_scalar, normx, fit);
}

__global__ void k_set(double *p, uint32_t n, double v)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_zero_i32(int *p, uint32_t n)
{
    for (uint32_t e = threadIdx.x; e < n; e += blockDim.x) p[e] = 0;
}

// Blocks of the two-stage row reductions: one per `per` rows, at most
// kRedBlocks (a function of the row count only, so results are reproducible).
static uint32_t red_blocks(uint64_t rows, uint32_t per)
{
    uint64_t b = (rows + per - 1) / per;
    if (b < 1) b = 1;
    if (b > (uint64_t)kRedBlocks) b = kRedBlocks;
    return (uint32_t)b;
}

static unsigned grid_rows(uint64_t n, uint32_t per_block)
{
    uint64_t b = (n + per_block - 1) / per_block;
    if (b < 1) b = 1;
    if (b > 65535) b = 65535;
    return (unsigned)b;
}

// Row solves: one thread per row, 256 rows per block.
// Blocks of the solve over I rows at rank R: 256 / RM rows a block, RM the
// rank class (8, 16, 32 or 64).
static unsigned solve_blocks(uint32_t I, uint32_t R)
{
    const uint64_t rm = R <= 8 ? 8 : (R <= 16 ? 16 : (R <= 32 ? 32 : 64)), rpb = 256 / rm;
    return (unsigned)(((uint64_t)I + rpb - 1) / rpb);
}

// ---------------------------------------------------------------------------
// Batched post-MTTKRP step: one grid per phase over many requests.  A block
// finds its request by binary search over the requests' first blocks of the
// phase and runs the same body as the single-request kernel with its block
// index within the request.
// ---------------------------------------------------------------------------
enum { kPhSolve = 0, kPhColPartial, kPhGramPartial, kPhGramSum, kPhInner, kPhases };

struct AlsStepDesc {
    const float *M;            // MTTKRP output, I x S
    float *A;                  // factor of the mode, I x S
    uint16_t *Abf;             // its BF16 copy, or null
    double *ata;               // N x R x R Grams
    double *partial;
    double *fac;
    int *piv, *kind;
    double *lambda, *norms;
    const double *normx;
    double *fit;
    double *host_fit;          // the job's pinned slot: the fit is also stored there
    uint32_t I, R, S, N, m, norm_kind, end_iter, pad;
    uint32_t nb_col, nb_gram, nb_inner, pad1;
    uint32_t blk[kPhases];     // first block in each phase's grid (solve: within the rank class)
    uint32_t pad2;
};

__device__ __forceinline__ const AlsStepDesc &als_find(const AlsStepDesc *__restrict__ d, uint32_t n, int ph,
                                                       uint32_t b)
{
    uint32_t lo = 0, hi = n;   // last desc with blk[ph] <= b
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (d[mid].blk[ph] <= b) lo = mid; else hi = mid;
    }
    return d[lo];
}

template <int RM>
__global__ __launch_bounds__(64) void k_factor_b(const AlsStepDesc *__restrict__ d)
{
    const AlsStepDesc &q = d[blockIdx.x];
    factor_body<RM>(q.ata, q.N, q.m, q.R, q.fac, q.piv, q.kind);
}

template <int RM>
__global__ __launch_bounds__(256) void k_solve_b(const AlsStepDesc *__restrict__ d, uint32_t n)
{
    const AlsStepDesc &q = als_find(d, n, kPhSolve, blockIdx.x);
    solve_rows_body<RM>(q.M, q.A, q.I, q.R, q.S, q.fac, q.kind, blockIdx.x - q.blk[kPhSolve]);
}

// Column partials beside the Gram partials of the request's workspace.
__device__ __forceinline__ double *als_colpart(const AlsStepDesc &q)
{
    return q.partial + (size_t)kRedBlocks * q.R * q.R;
}

__global__ __launch_bounds__(256) void k_col_partial_b(const AlsStepDesc *__restrict__ d, uint32_t n)
{
    const AlsStepDesc &q = als_find(d, n, kPhColPartial, blockIdx.x);
    col_partial_body(q.A, q.I, q.R, q.S, (int)q.norm_kind, als_colpart(q), blockIdx.x - q.blk[kPhColPartial], q.nb_col);
}

__global__ __launch_bounds__(256) void k_norm_scale_gram_b(const AlsStepDesc *__restrict__ d, uint32_t n)
{
    const AlsStepDesc &q = als_find(d, n, kPhGramPartial, blockIdx.x);
    norm_scale_gram_body(q.A, q.Abf, q.I, q.R, q.S, (int)q.norm_kind, als_colpart(q), q.nb_col, q.lambda, q.norms,
                         q.partial, blockIdx.x - q.blk[kPhGramPartial], q.nb_gram);
}

__global__ __launch_bounds__(256) void k_gram_sum_b(const AlsStepDesc *__restrict__ d, uint32_t n)
{
    const AlsStepDesc &q = als_find(d, n, kPhGramSum, blockIdx.x);
    sum_partials_body(q.partial, q.nb_gram, q.R * q.R, q.ata + (size_t)q.m * q.R * q.R, blockIdx.x - q.blk[kPhGramSum]);
}

__global__ __launch_bounds__(256) void k_inner_b(const AlsStepDesc *__restrict__ d, uint32_t n)
{
    const AlsStepDesc &q = als_find(d, n, kPhInner, blockIdx.x);
    inner_partial_body(q.A, q.M, q.I, q.R, q.S, q.partial, blockIdx.x - q.blk[kPhInner], q.nb_inner);
}

__global__ __launch_bounds__(256) void k_fit_b(const AlsStepDesc *__restrict__ d)
{
    const AlsStepDesc &q = d[blockIdx.x];
    if (!q.end_iter) return;   // whole block
    fit_body(q.ata, q.N, q.R, q.lambda, q.partial, q.nb_inner, 0, q.normx, q.fit);
    // thread 0 wrote the fit; the host reads the slot once the stream has
    // passed this kernel (no copy is queued per job)
    if (threadIdx.x == 0) *q.host_fit = *q.fit;
}

// ---------------------------------------------------------------------------
// One ALS solve as a resumable job (used directly below and by the server)
// ---------------------------------------------------------------------------
CpdWork::~CpdWork()
{
    dfree(M);
    for (auto p : bf) dfree(p);
    dfree(ata); dfree(partial); dfree(fac); dfree(piv); dfree(kind);
    dfree(lambda); dfree(norms); dfree(normx); dfree(fit);
}

static int gram(const float *A, uint32_t I, uint32_t R, uint32_t S, double *out, double *partial, hipStream_t s)
{
    int tok = prof_start("cpd_gram", s);
    const uint32_t nb = red_blocks(I, 32);
    hipLaunchKernelGGL(k_gram_partial, dim3(nb), dim3(256), 0, s, A, I, R, S, partial);
    hipLaunchKernelGGL(k_sum_partials, dim3(sum_grid(R * R)), dim3(256), 0, s, partial, nb, R * R, out);
    prof_stop(tok, s);
    SPTOL_HIP(hipGetLastError(), "CPD gram");
    return 0;
}

static int colnorm(float *A, uint16_t *Abf, uint32_t I, uint32_t R, uint32_t S, int kind, int mult, CpdWork &w,
                   hipStream_t s)
{
    int tok = prof_start("cpd_norm", s);
    const uint32_t nb = red_blocks(I, 64);
    hipLaunchKernelGGL(k_col_partial, dim3(nb), dim3(256), 0, s, A, I, R, S, kind, w.partial);
    hipLaunchKernelGGL(k_col_final, dim3(sum_grid(R)), dim3(256), 0, s, w.partial, nb, R, kind, mult, w.lambda,
                       w.norms);
    hipLaunchKernelGGL(k_scale, dim3(grid_rows((uint64_t)I * S, 256)), dim3(256), 0, s, A, Abf, I, R, S, w.norms);
    prof_stop(tok, s);
    SPTOL_HIP(hipGetLastError(), "CPD norm");
    return 0;
}

// The step after a mode's solve: column norms (kind as in colnorm, lambda
// set to them), the scaled factor and its BF16 copy, and the mode's Gram
// into ata_m.  Three launches; the same results as colnorm + gram.
static int norm_scale_gram(float *A, uint16_t *Abf, uint32_t I, uint32_t R, uint32_t S, int kind, CpdWork &w,
                           double *ata_m, hipStream_t s)
{
    int tok = prof_start("cpd_norm", s);
    const uint32_t ncb = red_blocks(I, 64), ngb = red_blocks(I, 32);
    double *colpart = w.partial + (size_t)kRedBlocks * R * R;   // beside the Gram partials
    hipLaunchKernelGGL(k_col_partial, dim3(ncb), dim3(256), 0, s, A, I, R, S, kind, colpart);
    hipLaunchKernelGGL(k_norm_scale_gram, dim3(ngb), dim3(256), 0, s, A, Abf, I, R, S, kind, colpart, ncb, w.lambda,
                       w.norms, w.partial);
    hipLaunchKernelGGL(k_sum_partials, dim3(sum_grid(R * R)), dim3(256), 0, s, w.partial, ngb, R * R, ata_m);
    prof_stop(tok, s);
    SPTOL_HIP(hipGetLastError(), "CPD norm");
    return 0;
}

// factored: the MTTKRP launch already ran the factor step.
static int solve(const float *M, float *A, uint32_t I, uint32_t R, uint32_t S, CpdWork &w, uint32_t N, uint32_t m,
                 hipStream_t s, bool factored = false)
{
    int tok = prof_start("cpd_solve", s);
    if (!factored) {
        if (R <= 8) hipLaunchKernelGGL(k_factor<8>, dim3(1), dim3(64), 0, s, w.ata, N, m, R, w.fac, w.piv, w.kind);
        else if (R <= 16) hipLaunchKernelGGL(k_factor<16>, dim3(1), dim3(64), 0, s, w.ata, N, m, R, w.fac, w.piv, w.kind);
        else if (R <= 32) hipLaunchKernelGGL(k_factor<32>, dim3(1), dim3(64), 0, s, w.ata, N, m, R, w.fac, w.piv, w.kind);
        else hipLaunchKernelGGL(k_factor<64>, dim3(1), dim3(64), 0, s, w.ata, N, m, R, w.fac, w.piv, w.kind);
    }
    const unsigned g = solve_blocks(I, R);
    if (R <= 8) hipLaunchKernelGGL(k_solve_rows<8>, dim3(g), dim3(256), 0, s, M, A, I, R, S, w.fac, w.kind);
    else if (R <= 16) hipLaunchKernelGGL(k_solve_rows<16>, dim3(g), dim3(256), 0, s, M, A, I, R, S, w.fac, w.kind);
    else if (R <= 32) hipLaunchKernelGGL(k_solve_rows<32>, dim3(g), dim3(256), 0, s, M, A, I, R, S, w.fac, w.kind);
    else hipLaunchKernelGGL(k_solve_rows<64>, dim3(g), dim3(256), 0, s, M, A, I, R, S, w.fac, w.kind);
    prof_stop(tok, s);
    SPTOL_HIP(hipGetLastError(), "CPD solve");
    return 0;
}

AlsBatchBuffers::~AlsBatchBuffers() { dfree(dev); }

static int rank_class(uint32_t R) { return R <= 8 ? 0 : (R <= 16 ? 1 : (R <= 32 ? 2 : 3)); }

int als_step_batch(const std::vector<CpdJob *> &jobs, AlsBatchBuffers &bb, hipStream_t s)
{
    if (jobs.empty()) return 0;
    StreamScope scope(s);   // the table buffer is used on s only
    std::vector<AlsStepDesc> d;
    d.reserve(jobs.size());
    uint32_t c0[5] = {};
    for (int c = 0; c < 4; ++c) {
        c0[c] = (uint32_t)d.size();
        uint32_t sb = 0;
        for (CpdJob *j : jobs) {
            if (rank_class(j->R) != c) continue;
            AlsStepDesc q;
            std::memset(&q, 0, sizeof(q));
            const uint32_t m = j->m, I = j->ndims[m];
            q.M = j->w.M;
            q.A = (float *)j->mats[m].values;
            q.Abf = j->fdt == SPTOL_BF16 ? j->w.bf[m] : nullptr;
            q.ata = j->w.ata;
            q.partial = j->w.partial;
            q.fac = j->w.fac; q.piv = j->w.piv; q.kind = j->w.kind;
            q.lambda = j->w.lambda; q.norms = j->w.norms;
            q.normx = j->w.normx; q.fit = j->w.fit; q.host_fit = j->h_fit;
            q.I = I; q.R = j->R; q.S = j->S; q.N = j->N; q.m = m;
            q.norm_kind = j->it == 0 ? 0 : 1;
            q.end_iter = m + 1 == j->N ? 1 : 0;
            q.nb_col = red_blocks(I, 64);
            q.nb_gram = red_blocks(I, 32);
            q.nb_inner = q.end_iter ? red_blocks(I, 256) : 0;
            q.blk[kPhSolve] = sb;
            sb += solve_blocks(I, j->R);
            d.push_back(q);
        }
    }
    c0[4] = (uint32_t)d.size();
    const uint32_t n = (uint32_t)d.size();
    uint64_t tot[kPhases] = {};
    for (AlsStepDesc &q : d) {
        const uint32_t nb[kPhases] = {0, q.nb_col, q.nb_gram, sum_grid(q.R * q.R), q.nb_inner};
        for (int p = kPhColPartial; p < kPhases; ++p) {
            q.blk[p] = (uint32_t)tot[p];
            tot[p] += nb[p];
        }
    }
    for (int p = 0; p < kPhases; ++p)
        if (tot[p] > 0x7FFFFFFFull) SPTOL_CHECK(SPTERR_VALUE_ERROR, "CPD batch", "too many blocks in one step");
    if (bb.cap < n) {
        dfree(bb.dev);
        bb.dev = nullptr;
        bb.cap = 0;
        SPTOL_TRY(dmalloc((void **)&bb.dev, (size_t)n * 2 * sizeof(AlsStepDesc), "CPD batch"));
        bb.cap = (size_t)n * 2;
    }
    SPTOL_TRY(bb.stage.upload(bb.dev, d.data(), (size_t)n * sizeof(AlsStepDesc), s));
    int tok = prof_start("cpd_batch_step", s);
    for (int c = 0; c < 4; ++c) {
        const uint32_t k = c0[c], nc = c0[c + 1] - c0[c];
        if (!nc) continue;
        const AlsStepDesc *dc = bb.dev + k;
        switch (c) {
        case 0: hipLaunchKernelGGL(k_factor_b<8>, dim3(nc), dim3(64), 0, s, dc); break;
        case 1: hipLaunchKernelGGL(k_factor_b<16>, dim3(nc), dim3(64), 0, s, dc); break;
        case 2: hipLaunchKernelGGL(k_factor_b<32>, dim3(nc), dim3(64), 0, s, dc); break;
        default: hipLaunchKernelGGL(k_factor_b<64>, dim3(nc), dim3(64), 0, s, dc); break;
        }
    }
    for (int c = 0; c < 4; ++c) {
        const uint32_t k = c0[c], nc = c0[c + 1] - c0[c];
        if (!nc) continue;
        const AlsStepDesc &last = d[c0[c + 1] - 1];
        const unsigned g = last.blk[kPhSolve] + solve_blocks(last.I, last.R);
        const AlsStepDesc *dc = bb.dev + k;
        switch (c) {
        case 0: hipLaunchKernelGGL(k_solve_b<8>, dim3(g), dim3(256), 0, s, dc, nc); break;
        case 1: hipLaunchKernelGGL(k_solve_b<16>, dim3(g), dim3(256), 0, s, dc, nc); break;
        case 2: hipLaunchKernelGGL(k_solve_b<32>, dim3(g), dim3(256), 0, s, dc, nc); break;
        default: hipLaunchKernelGGL(k_solve_b<64>, dim3(g), dim3(256), 0, s, dc, nc); break;
        }
    }
    hipLaunchKernelGGL(k_col_partial_b, dim3((unsigned)tot[kPhColPartial]), dim3(256), 0, s, bb.dev, n);
    hipLaunchKernelGGL(k_norm_scale_gram_b, dim3((unsigned)tot[kPhGramPartial]), dim3(256), 0, s, bb.dev, n);
    hipLaunchKernelGGL(k_gram_sum_b, dim3((unsigned)tot[kPhGramSum]), dim3(256), 0, s, bb.dev, n);
    if (tot[kPhInner]) {
        hipLaunchKernelGGL(k_inner_b, dim3((unsigned)tot[kPhInner]), dim3(256), 0, s, bb.dev, n);
        hipLaunchKernelGGL(k_fit_b, dim3(n), dim3(256), 0, s, bb.dev);
    }
    prof_stop(tok, s);
    SPTOL_HIP(hipGetLastError(), "CPD batch");
    for (CpdJob *j : jobs) j->advance();
    return 0;
}

CpdJob::~CpdJob()
{
    pinned_slot_free(h_fit);
}

size_t CpdJob::workspace_bytes(const uint32_t *ndims, uint32_t N, uint32_t R, int fdt)
{
    const uint32_t S = R ? ((R - 1) / 8 + 1) * 8 : 8;
    uint64_t maxI = 1, sumI = 0;
    for (uint32_t m = 0; m < N; ++m) { maxI = ndims[m] > maxI ? ndims[m] : maxI; sumI += ndims[m]; }
    size_t b = (size_t)maxI * S * 4 + (size_t)N * R * R * 8 +
               std::max<size_t>((size_t)kRedBlocks * R * (R + 1), kInnerBlocks + 1) * 8 + (size_t)R * R * 8 + 4096;
    if (fdt == SPTOL_BF16) b += (size_t)sumI * S * 2;
    return b;
}

int CpdJob::setup(sptolDevCoo X_, uint32_t R_, uint32_t niters_, double tol_, sptolDevMatrix *mats_, int fdt_,
                  hipStream_t s)
{
    X = X_;
    H = nullptr;
    N = X->nmodes;
    for (uint32_t m = 0; m < N; ++m) ndims[m] = X->ndims[m];
    nnz = X->nnz;
    vals = X->vals;
    return init(R_, niters_, tol_, mats_, fdt_, s);
}

int CpdJob::setup_hicoo(sptolDevHicoo H_, uint32_t R_, uint32_t niters_, double tol_, sptolDevMatrix *mats_, int fdt_,
                        hipStream_t s)
{
    X = nullptr;
    H = H_;
    N = H->nmodes;
    for (uint32_t m = 0; m < N; ++m) ndims[m] = H->ndims[m];
    nnz = H->nnz;
    vals = H->vals;
    return init(R_, niters_, tol_, mats_, fdt_, s);
}

int CpdJob::init(uint32_t R_, uint32_t niters_, double tol_, sptolDevMatrix *mats_, int fdt_, hipStream_t s)
{
    R = R_; niters = niters_; tol = tol_; mats = mats_; fdt = fdt_;
    if (R < 1 || R > (uint32_t)kMaxRank) SPTOL_CHECK(SPTERR_VALUE_ERROR, "CPD", "rank must be in [1, 64]");
    if (fdt != SPTOL_F32 && fdt != SPTOL_BF16) SPTOL_CHECK(SPTERR_VALUE_ERROR, "CPD", "bad factor dtype");
    maxI = 0;
    for (uint32_t m = 0; m < N; ++m) {
        if (mats[m].dtype != SPTOL_F32 || mats[m].ncols != R || mats[m].nrows != ndims[m] ||
            mats[m].stride != mats[0].stride)
            SPTOL_CHECK(SPTERR_SHAPE_MISMATCH, "CPD", "factors must be FP32 ndims[m] x rank");
        maxI = ndims[m] > maxI ? ndims[m] : maxI;
    }
    S = mats[0].stride;
    SPTOL_TRY(dmalloc((void **)&w.M, (size_t)(maxI ? maxI : 1) * S * 4, "CPD"));
    SPTOL_TRY(dmalloc((void **)&w.ata, (size_t)N * R * R * 8, "CPD"));
    SPTOL_TRY(dmalloc((void **)&w.partial,
                      std::max<size_t>((size_t)kRedBlocks * R * (R + 1), kInnerBlocks + 1) * 8, "CPD"));
    SPTOL_TRY(dmalloc((void **)&w.fac, (size_t)R * R * 8, "CPD"));
    SPTOL_TRY(dmalloc((void **)&w.piv, (size_t)R * 4, "CPD"));
    SPTOL_TRY(dmalloc((void **)&w.kind, 8, "CPD"));   // [0] factor kind, [1] singular solves so far
    SPTOL_TRY(dmalloc((void **)&w.lambda, (size_t)R * 8, "CPD"));
    SPTOL_TRY(dmalloc((void **)&w.norms, (size_t)R * 8, "CPD"));
    SPTOL_TRY(dmalloc((void **)&w.normx, 8, "CPD"));
    SPTOL_TRY(dmalloc((void **)&w.fit, 8, "CPD"));
    if (fdt == SPTOL_BF16)
        for (uint32_t m = 0; m < N; ++m)
            SPTOL_TRY(dmalloc((void **)&w.bf[m], (size_t)(ndims[m] ? ndims[m] : 1) * S * 2, "CPD"));
    for (uint32_t m = 0; m < N; ++m) {
        mt[m] = mats[m];
        if (fdt == SPTOL_BF16) { mt[m].dtype = SPTOL_BF16; mt[m].values = w.bf[m]; }
    }
    Mout.nrows = maxI; Mout.ncols = R; Mout.stride = S; Mout.dtype = SPTOL_F32; Mout.values = w.M;
    static_assert((kMaxRank + 2) * sizeof(double) <= kPinnedSlotBytes,
                  "fit, weights and the singular-solve count fit one pinned slot");
    SPTOL_TRY(pinned_slot((void **)&h_fit));
    // ||X||^2 and the initial Grams (cpd.c:65-70)
    hipLaunchKernelGGL(k_sqnorm_partial, dim3(kRedBlocks), dim3(256), 0, s, vals, nnz, w.partial);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, w.partial, (uint32_t)kRedBlocks, 1u, w.normx);
    // (k_sqnorm_partial runs kRedBlocks blocks over the nonzeros)
    hipLaunchKernelGGL(k_set, dim3(1), dim3(kMaxRank), 0, s, w.lambda, R, 1.0);
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, s, w.kind, 2u);
    SPTOL_HIP(hipGetLastError(), "CPD init");
    for (uint32_t m = 0; m < N; ++m) {
        SPTOL_TRY(gram((const float *)mats[m].values, ndims[m], R, S, w.ata + (size_t)m * R * R, w.partial, s));
        if (fdt == SPTOL_BF16) {
            hipLaunchKernelGGL(k_set, dim3(1), dim3(kMaxRank), 0, s, w.norms, R, 1.0);
            hipLaunchKernelGGL(k_scale, dim3(grid_rows((uint64_t)ndims[m] * S, 256)), dim3(256), 0, s,
                               (float *)mats[m].values, w.bf[m], ndims[m], R, S, w.norms);
        }
    }
    it = 0; m = 0; fit = oldfit = 0.0; iters = 0;
    done = niters == 0;
    fit_pending = false;
    t_iter = std::chrono::steady_clock::now();
    return 0;
}

void CpdJob::order(uint32_t *o) const
{
    o[0] = m;
    for (uint32_t k = 1; k < N; ++k) o[k] = (m + k) % N;
}

int CpdJob::mttkrp(hipStream_t s)
{
    uint32_t o[kMaxModes];
    order(o);
    Mout.nrows = ndims[m];
    factored = false;
    // a 4-mode launch forms this mode's V^{-1} beside its ranges (the
    // factor step needs only the other modes' Grams)
    const AlsFactorReq fr = {w.ata, w.fac, w.piv, w.kind, N, m, R, 0u};
    if (H) return hicoo_mttkrp(H, mt, o, m, &Mout, s, &fr, &factored);
    return coo_mttkrp(X, mt, o, m, &Mout, s, &fr, &factored);
}

int CpdJob::after_mttkrp(hipStream_t s)
{
    const uint32_t I = ndims[m];
    SPTOL_TRY(solve(w.M, (float *)mats[m].values, I, R, S, w, N, m, s, factored));
    SPTOL_TRY(norm_scale_gram((float *)mats[m].values, fdt == SPTOL_BF16 ? w.bf[m] : nullptr, I, R, S, it == 0 ? 0 : 1,
                              w, w.ata + (size_t)m * R * R, s));
    if (++m < N) return 0;
    // end of an iteration: fit from the last mode's MTTKRP, still in w.M
    const uint32_t last = N - 1;
    const uint32_t nfit = red_blocks(ndims[last], 256);
    hipLaunchKernelGGL(k_inner_partial, dim3(nfit), dim3(256), 0, s, (const float *)mats[last].values, w.M,
                       ndims[last], R, S, w.partial);
    hipLaunchKernelGGL(k_fit, dim3(1), dim3(256), 0, s, w.ata, N, R, w.lambda, w.partial, nfit, 0, w.normx, w.fit);
    SPTOL_HIP(hipGetLastError(), "CPD fit");
    SPTOL_HIP(hipMemcpyAsync(h_fit, w.fit, sizeof(double), hipMemcpyDeviceToHost, s), "CPD fit");
    fit_pending = true;
    return 0;
}

void CpdJob::advance()
{
    if (++m < N) return;
    fit_pending = true;
}

void CpdJob::fit_skipped()
{
    fit_pending = false;
    iters = (int)it + 1;
    if (it + 1 >= niters) {
        done = true;
    } else {
        ++it;
        m = 0;
    }
}

void CpdJob::fit_ready()
{
    fit = *h_fit;
    fit_pending = false;
    iters = (int)it + 1;
    if (verbose()) {
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_iter).count();
        printf("  its = %u ( %.3lf s ) fit = %0.5f  delta = %+0.4e\n", it + 1, sec, fit, fit - oldfit);
    }
    t_iter = std::chrono::steady_clock::now();
    if ((it > 0 && std::fabs(fit - oldfit) < tol) || it + 1 >= niters) {
        done = true;
    } else {
        oldfit = fit;
        ++it;
        m = 0;
    }
}

// Queue the final results: with exact_fit (3- and 4-mode tensors) the fit
// of the current model with <X,K> by an FP64 pass over the nonzeros (before
// the renormalisation below, which leaves the model unchanged), then the
// final weights (GetFinalLambda, matrix.c:557-573); both are copied to the
// pinned slot.
int CpdJob::finish_enqueue(hipStream_t s)
{
    // a HiCOO tensor's segmented view packs the same records as a COO view
    const SegView *hv = (H && (N == 3 || N == 4)) ? hicoo_seg_records(H) : nullptr;
    if (exact_fit) {
        if (hv) {
            InnerArgs ia;
            std::memset(&ia, 0, sizeof(ia));
            for (uint32_t q = 0; q < N; ++q) ia.mats[q] = (const float *)mats[q].values;
            ia.nnz = hv->nnz;
            ia.N = N; ia.R = R; ia.S = S;
            if (N == 3) launch_inner_exact_rec<3>(*hv, ia, w.lambda, w.partial, s);
            else launch_inner_exact_rec<4>(*hv, ia, w.lambda, w.partial, s);
        } else if (X && N <= 4) {
            InnerArgs ia;
            std::memset(&ia, 0, sizeof(ia));
            for (uint32_t q = 0; q < N; ++q) { ia.inds[q] = X->inds[q]; ia.mats[q] = (const float *)mats[q].values; }
            ia.vals = X->vals;
            ia.nnz = X->nnz;
            ia.N = N; ia.R = R; ia.S = S;
            // a mode view's packed records when one exists (the ALS loop
            // built them), else the tensor's own index arrays
            const SegView *sv = nullptr;
            for (uint32_t q = 0; q < N && !sv; ++q)
                if (X->sorted[q].ready && X->sorted[q].seg.ready && X->sorted[q].seg.rec && X->sorted[q].seg.nnz == X->nnz)
                    sv = &X->sorted[q].seg;
            const dim3 g(kInnerBlocks);
            if (sv && N == 3) launch_inner_exact_rec<3>(*sv, ia, w.lambda, w.partial, s);
            else if (sv && N == 4) launch_inner_exact_rec<4>(*sv, ia, w.lambda, w.partial, s);
            else if (S <= 8) hipLaunchKernelGGL(k_inner_exact_partial<2>, g, dim3(256), 0, s, ia, w.lambda, w.partial);
            else if (S <= 16) hipLaunchKernelGGL(k_inner_exact_partial<4>, g, dim3(256), 0, s, ia, w.lambda, w.partial);
            else if (S <= 32) hipLaunchKernelGGL(k_inner_exact_partial<8>, g, dim3(256), 0, s, ia, w.lambda, w.partial);
            else hipLaunchKernelGGL(k_inner_exact_partial<16>, g, dim3(256), 0, s, ia, w.lambda, w.partial);
        } else {
            InnerGenArgs ga;
            std::memset(&ga, 0, sizeof(ga));
            for (uint32_t q = 0; q < N; ++q) ga.mats[q] = (const float *)mats[q].values;
            ga.N = N; ga.R = R; ga.S = S;
            if (X) {
                for (uint32_t q = 0; q < N; ++q) ga.inds[q] = X->inds[q];
                ga.vals = X->vals;
                ga.nnz = X->nnz;
                launch_inner_exact_gen<false>(ga, w.lambda, w.partial, s);
            } else {
                for (uint32_t q = 0; q < N; ++q) { ga.binds[q] = H->binds[q]; ga.einds[q] = H->einds[q]; }
                ga.bptr = H->bptr;
                ga.nb = H->nb;
                ga.sb_bits = H->sb_bits;
                ga.vals = H->vals;
                ga.nnz = H->nnz;
                launch_inner_exact_gen<true>(ga, w.lambda, w.partial, s);
            }
        }
        // block partials -> one scalar (kept after the partials)
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, w.partial, kInnerBlocks, 1u,
                           w.partial + kInnerBlocks);
        hipLaunchKernelGGL(k_fit, dim3(1), dim3(256), 0, s, w.ata, N, R, w.lambda, w.partial + kInnerBlocks, 1u, 1,
                           w.normx, w.fit);
        SPTOL_HIP(hipGetLastError(), "CPD exact fit");
        SPTOL_HIP(hipMemcpyAsync(h_fit, w.fit, sizeof(double), hipMemcpyDeviceToHost, s), "CPD exact fit");
    }
    for (uint32_t q = 0; q < N; ++q)
        SPTOL_TRY(colnorm((float *)mats[q].values, nullptr, ndims[q], R, S, 0, 1, w, s));
    SPTOL_HIP(hipMemcpyAsync(h_fit + 1, w.lambda, (size_t)R * 8, hipMemcpyDeviceToHost, s), "CPD");
    SPTOL_HIP(hipMemcpyAsync(h_fit + 1 + kMaxRank, w.kind + 1, sizeof(int), hipMemcpyDeviceToHost, s), "CPD");
    return 0;
}

void CpdJob::finish_read(sptValue *lambda_out)
{
    if (exact_fit) fit = h_fit[0];
    for (uint32_t r = 0; r < R; ++r) lambda_out[r] = (sptValue)h_fit[1 + r];
    std::memcpy(&singular_solves, h_fit + 1 + kMaxRank, sizeof(int));
}

int CpdJob::finish(hipStream_t s, sptValue *lambda_out)
{
    SPTOL_TRY(finish_enqueue(s));
    SPTOL_HIP(hipStreamSynchronize(s), "CPD");
    finish_read(lambda_out);
    return 0;
}

}  // namespace sptol

using namespace sptol;

extern "C" {

}  // extern "C"

namespace sptol {

// The ALS loop of a set-up job, the final weights, and the statistics.
static int run_job(CpdJob &job, sptValue *lambda_out, sptolCpdStats *stats, hipStream_t s,
                   std::chrono::steady_clock::time_point t0)
{
    // Without a positive tolerance (and without per-iteration prints) only
    // the iteration count ends the loop: every iteration is queued without
    // waiting for its fit, and the last fit is read after the final
    // synchronisation.
    const bool wait_fits = job.tol > 0.0 || verbose();
    while (!job.done) {
        SPTOL_TRY(job.mttkrp(s));
        SPTOL_TRY(job.after_mttkrp(s));
        if (job.fit_pending) {
            if (wait_fits) {
                SPTOL_HIP(hipStreamSynchronize(s), "CPD");
                job.fit_ready();
            } else {
                job.fit_skipped();
            }
        }
    }
    SPTOL_TRY(job.finish(s, lambda_out));
    if (!wait_fits) job.fit = job.h_fit[0];   // the last iteration's fit, or the exact one
    if (stats) {
        stats->iterations = job.iters;
        stats->singular_solves = job.singular_solves;
        stats->fit = job.fit;
        stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return 0;
}

}  // namespace sptol

extern "C" {

int sptolCpdAlsDevice(sptolDevCoo X, sptIndex rank, sptIndex niters, double tol, sptolDevMatrix *mats,
                      sptValue *lambda_out, sptolCpdStats *stats, void *stream)
{
    hipStream_t s = stream_of(stream);
    // the workspace below is freed after the final synchronisation
    StreamScope scope(s);
    auto t0 = std::chrono::steady_clock::now();
    // mode views need device sorts (at most 2^31 - 1 items); a larger tensor
    // runs the unsorted (atomic) MTTKRP
    if (sptolDevCooNnz(X) <= kMaxSortItems) SPTOL_TRY(coo_prepare_all(X, s));
    CpdJob job;
    job.exact_fit = stats ? stats->exact_fit != 0 : false;
    SPTOL_TRY(job.setup(X, rank, niters, tol, mats, stats ? stats->factor_dtype : SPTOL_F32, s));
    return run_job(job, lambda_out, stats, s, t0);
}

int sptolCpdAlsHicooDevice(sptolDevHicoo X, sptIndex rank, sptIndex niters, double tol, sptolDevMatrix *mats,
                           sptValue *lambda_out, sptolCpdStats *stats, void *stream)
{
    hipStream_t s = stream_of(stream);
    for (uint32_t m = 0; m < X->nmodes; ++m) SPTOL_TRY(hicoo_prepare(X, m, s));
    StreamScope scope(s);   // the workspace below is freed after the final synchronisation
    auto t0 = std::chrono::steady_clock::now();
    CpdJob job;
    job.exact_fit = stats ? stats->exact_fit != 0 : false;
    SPTOL_TRY(job.setup_hicoo(X, rank, niters, tol, mats, stats ? stats->factor_dtype : SPTOL_F32, s));
    return run_job(job, lambda_out, stats, s, t0);
}

// sptMatrixSolveNormals (matrix_solver.c:32-83): rhs[i] := rhs[i] V^{-1}
// for every row, V = Hadamard of aTa[(mode + k) % nmodes], k = 1..nmodes-1,
// read from each Gram's upper triangle (sptMatrixDotMulSeqTriangle,
// matrix.c:341-386).  FP64 Cholesky with the LU-with-pivoting fallback on
// the device; rows stay as they are when V is exactly singular, as gesv
// leaves its right-hand sides.  aTa[nmodes] receives V (the reference leaves
// its factor there).  The reference's notices are printed to stdout.
int sptMatrixSolveNormals(sptIndex const mode, sptIndex const nmodes, sptMatrix **aTa, sptMatrix *rhs)
{
    if (nmodes < 1 || mode >= nmodes) SPTOL_CHECK(SPTERR_VALUE_ERROR, "SolveNormals", "mode >= nmodes");
    const uint32_t R = aTa[0]->ncols, S = aTa[0]->stride;
    if (R < 1 || R > (uint32_t)kMaxRank) SPTOL_CHECK(SPTERR_VALUE_ERROR, "SolveNormals", "rank must be in [1, 64]");
    for (uint32_t m = 0; m <= nmodes; ++m)
        if (aTa[m]->nrows < R || aTa[m]->ncols != R || aTa[m]->stride != S)
            SPTOL_CHECK(SPTERR_SHAPE_MISMATCH, "SolveNormals", "aTa[m] must be rank x rank with one stride");
    if (rhs->ncols != R || rhs->stride != S)
        SPTOL_CHECK(SPTERR_SHAPE_MISMATCH, "SolveNormals", "rhs must have rank columns and the Grams' stride");
    const uint32_t N = nmodes, I = rhs->nrows;
    std::vector<double> g((size_t)N * R * R), V((size_t)R * R, 1.0);
    for (uint32_t m = 0; m < N; ++m)
        for (uint32_t i = 0; i < R; ++i)
            for (uint32_t j = 0; j < R; ++j) {
                const uint32_t a = i < j ? i : j, b = i < j ? j : i;   // upper triangle
                g[((size_t)m * R + i) * R + j] = (double)aTa[m]->values[(size_t)a * S + b];
            }
    for (uint32_t q = 1; q < N; ++q)
        for (size_t e = 0; e < (size_t)R * R; ++e) V[e] *= g[(size_t)((mode + q) % N) * R * R + e];
    for (uint32_t i = 0; i < R; ++i)
        for (uint32_t j = 0; j < R; ++j) aTa[N]->values[(size_t)i * S + j] = (sptValue)V[(size_t)i * R + j];
    hipStream_t s = stream_of(nullptr);
    StreamScope scope(s);
    CpdWork w;
    float *M = nullptr, *A = nullptr;
    const size_t rb = (size_t)(I ? I : 1) * S * 4;
    SPTOL_TRY(dmalloc((void **)&w.ata, g.size() * 8, "SolveNormals"));
    SPTOL_TRY(dmalloc((void **)&w.fac, (size_t)R * R * 8, "SolveNormals"));
    SPTOL_TRY(dmalloc((void **)&w.piv, (size_t)R * 4, "SolveNormals"));
    SPTOL_TRY(dmalloc((void **)&w.kind, 8, "SolveNormals"));
    int r = dmalloc((void **)&M, rb, "SolveNormals");
    if (!r) r = dmalloc((void **)&A, rb, "SolveNormals");
    int kind[2] = {0, 0};
    auto run = [&]() -> int {
        SPTOL_HIP(hipMemcpyAsync(w.ata, g.data(), g.size() * 8, hipMemcpyHostToDevice, s), "SolveNormals");
        SPTOL_HIP(hipMemcpyAsync(w.kind, kind, sizeof(kind), hipMemcpyHostToDevice, s), "SolveNormals");
        if (I) SPTOL_HIP(hipMemcpyAsync(M, rhs->values, (size_t)I * S * 4, hipMemcpyHostToDevice, s), "SolveNormals");
        SPTOL_TRY(solve(M, A, I, R, S, w, N, mode, s));
        if (I) SPTOL_HIP(hipMemcpyAsync(rhs->values, A, (size_t)I * S * 4, hipMemcpyDeviceToHost, s), "SolveNormals");
        SPTOL_HIP(hipMemcpyAsync(kind, w.kind, sizeof(kind), hipMemcpyDeviceToHost, s), "SolveNormals");
        SPTOL_HIP(hipStreamSynchronize(s), "SolveNormals");
        return 0;
    };
    if (!r) r = run();
    dfree(M);
    dfree(A);
    if (r) return r;
    if (kind[0] != 0) printf("Gram matrix is not SPD. Trying `gesv`.\n");
    if (kind[0] == 2) printf("gesv: the normal equations are singular; the right-hand sides are left unsolved.\n");
    return 0;
}

static void report_singular(int n)
{
    if (n > 0)
        printf("CPD-ALS: the normal equations were singular in %d mode update%s; those factors keep the "
               "unsolved MTTKRP rows, as gesv leaves its right-hand sides.\n", n, n == 1 ? "" : "s");
}

// The reference's exported ALS body (cpd.c:32-137) with caller-supplied
// factors: mats[0..N-1] in/out, mats[N] receives the last mode's MTTKRP.
double CpdAlsStep(sptSparseTensor const * const spten, sptIndex const rank, sptIndex const niters,
                  double const tol, sptMatrix ** mats, sptValue * const lambda)
{
    const uint32_t N = spten->nmodes;
    for (uint32_t m = 0; m < N; ++m)
        if (spten->ndims[m] != mats[m]->nrows || mats[m]->ncols != rank) {
            sptol_complain("CPD", SPTERR_SHAPE_MISMATCH, __FILE__, __LINE__, "factor shape");
            return NAN;
        }
    sptolDevCoo X = nullptr;
    if (sptolDevCooCreate(&X, spten)) return NAN;
    sptolDevMatrix dm[kMaxModes];
    std::memset(dm, 0, sizeof(dm));
    int r = 0;
    for (uint32_t m = 0; m < N && !r; ++m) {
        r = sptolDevMatrixCreate(&dm[m], mats[m]->nrows, rank, SPTOL_F32);
        if (!r) r = sptolDevMatrixUpload(&dm[m], mats[m]);
    }
    sptolCpdStats st;
    std::memset(&st, 0, sizeof(st));
    st.factor_dtype = SPTOL_F32;
    st.exact_fit = 1;   // the returned fit: <X,K> by an FP64 pass over the nonzeros
    if (!r) r = sptolCpdAlsDevice(X, rank, niters, tol, dm, lambda, &st, nullptr);
    if (!r) report_singular(st.singular_solves);
    for (uint32_t m = 0; m < N && !r; ++m) r = sptolDevMatrixDownload(mats[m], &dm[m]);
    // mats[N]: the last mode's MTTKRP with the final factors, as the
    // reference leaves its scratch matrix
    if (!r && mats[N] && mats[N]->cap >= spten->ndims[N - 1]) {
        sptolDevMatrix out;
        r = sptolDevMatrixCreate(&out, spten->ndims[N - 1], rank, SPTOL_F32);
        uint32_t order[kMaxModes];
        order[0] = N - 1;
        for (uint32_t k = 1; k < N; ++k) order[k] = (N - 1 + k) % N;
        if (!r) r = sptolDevMTTKRP(X, dm, order, N - 1, &out, nullptr);
        if (!r) r = sptolHipStreamSynchronize(nullptr);
        if (!r) {
            sptIndex keep = mats[N]->nrows;
            mats[N]->nrows = spten->ndims[N - 1];
            r = sptolDevMatrixDownload(mats[N], &out);
            mats[N]->nrows = keep;
        }
        sptolDevMatrixFree(&out);
    }
    for (uint32_t m = 0; m < N; ++m) sptolDevMatrixFree(&dm[m]);
    sptolDevCooFree(X);
    return r ? NAN : st.fit;
}

int sptCpdAls(sptSparseTensor const * const spten, sptIndex const rank, sptIndex const niters, double const tol,
              sptKruskalTensor * ktensor)
{
    const uint32_t N = spten->nmodes;
    if (N < 2 || N > (uint32_t)kMaxModes) SPTOL_CHECK(SPTERR_VALUE_ERROR, "CPU SpTns CPD-ALS", "nmodes must be in [2, 8]");
    if (!ktensor->lambda) SPTOL_CHECK(SPTERR_VALUE_ERROR, "CPU SpTns CPD-ALS", "ktensor not created by sptNewKruskalTensor");
    sptIndex max_dim = sptMaxIndexArray(spten->ndims, N);
    sptMatrix **mats = (sptMatrix **)calloc(N + 1, sizeof(sptMatrix *));
    SPTOL_CHECK_OS(!mats, "CPU SpTns CPD-ALS");
    int r = 0;
    for (uint32_t m = 0; m <= N && !r; ++m) {
        mats[m] = (sptMatrix *)malloc(sizeof(sptMatrix));
        if (!mats[m]) { r = SPTERR_OS_ERROR + ENOMEM; break; }
        if (m < N) {
            r = sptNewMatrix(mats[m], spten->ndims[m], rank);
            if (!r) r = sptRandomizeMatrix(mats[m], spten->ndims[m], rank);
        } else {
            r = sptNewMatrix(mats[m], max_dim, rank);
            if (!r) r = sptConstantMatrix(mats[m], 0);
        }
    }
    if (!r) {
        auto t0 = std::chrono::steady_clock::now();
        double fit = CpdAlsStep(spten, rank, niters, tol, mats, ktensor->lambda);
        if (std::isnan(fit)) r = sptGetLastError(nullptr, nullptr, nullptr, nullptr);
        if (!r && std::isnan(fit)) r = SPTERR_UNKNOWN;
        ktensor->fit = fit;
        if (verbose())
            fprintf(stdout, "[GPU  SpTns CPD-ALS]: %.9lf s\n",
                    std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    }
    if (r) {
        for (uint32_t m = 0; m <= N; ++m)
            if (mats[m]) { sptFreeMatrix(mats[m]); free(mats[m]); }
        free(mats);
        return r;
    }
    if (ktensor->factors) {
        for (uint32_t m = 0; m <= ktensor->nmodes; ++m)
            if (ktensor->factors[m]) { sptFreeMatrix(ktensor->factors[m]); free(ktensor->factors[m]); }
        free(ktensor->factors);
    }
    ktensor->factors = mats;   // owned by the Kruskal tensor (freed by sptFreeKruskalTensor)
    return 0;
}

int sptOmpCpdAls(sptSparseTensor const * const spten, sptIndex const rank, sptIndex const niters, double const tol,
                 const int tk, const int use_reduce, sptKruskalTensor * ktensor)
{
    (void)tk;
    (void)use_reduce;
    return sptCpdAls(spten, rank, niters, tol, ktensor);
}

// sptCpdAlsHiCOO (hicoo/cpd.c:149-188 with spt_CpdAlsStepHiCOO, :34-135):
// seeded U[0,1) factors, the ALS loop on the device over the HiCOO tensor,
// factors and weights back in ktensor.  factors[nmodes] is the emptied
// MTTKRP scratch, as the reference leaves it.
int sptCpdAlsHiCOO(sptSparseTensorHiCOO const * const hitsr, sptIndex const rank, sptIndex const niters,
                   double const tol, sptRankKruskalTensor * ktensor)
{
    const uint32_t N = hitsr->nmodes;
    if (N < 2 || N > (uint32_t)kMaxModes) SPTOL_CHECK(SPTERR_VALUE_ERROR, "CPU HiCOO SpTns CPD-ALS", "nmodes must be in [2, 8]");
    if (rank < 1 || rank > (uint32_t)kMaxRank) SPTOL_CHECK(SPTERR_VALUE_ERROR, "CPU HiCOO SpTns CPD-ALS", "rank must be in [1, 64]");
    if (!ktensor->lambda) SPTOL_CHECK(SPTERR_VALUE_ERROR, "CPU HiCOO SpTns CPD-ALS", "ktensor not created by sptNewRankKruskalTensor");
    sptIndex max_dim = sptMaxIndexArray(hitsr->ndims, N);
    sptRankMatrix **mats = (sptRankMatrix **)calloc(N + 1, sizeof(sptRankMatrix *));
    SPTOL_CHECK_OS(!mats, "CPU HiCOO SpTns CPD-ALS");
    int r = 0;
    for (uint32_t m = 0; m <= N && !r; ++m) {
        mats[m] = (sptRankMatrix *)calloc(1, sizeof(sptRankMatrix));
        if (!mats[m]) { r = SPTERR_OS_ERROR + ENOMEM; break; }
        r = sptNewRankMatrix(mats[m], m < N ? hitsr->ndims[m] : max_dim, (sptElementIndex)rank);
        if (!r) r = m < N ? sptRandomizeRankMatrix(mats[m], hitsr->ndims[m], (sptElementIndex)rank)
                          : sptConstantRankMatrix(mats[m], 0);
    }
    sptolDevHicoo H = nullptr;
    sptolDevMatrix dm[kMaxModes];
    std::memset(dm, 0, sizeof(dm));
    // a sptMatrix view of a rank matrix (same row stride rule)
    auto view = [](sptRankMatrix *a) {
        sptMatrix v;
        v.nrows = a->nrows; v.ncols = a->ncols; v.cap = a->cap; v.stride = a->stride; v.values = a->values;
        return v;
    };
    if (!r) r = sptolDevHicooCreate(&H, hitsr);
    for (uint32_t m = 0; m < N && !r; ++m) {
        r = sptolDevMatrixCreate(&dm[m], hitsr->ndims[m], rank, SPTOL_F32);
        sptMatrix v = view(mats[m]);
        if (!r) r = sptolDevMatrixUpload(&dm[m], &v);
    }
    sptolCpdStats st;
    std::memset(&st, 0, sizeof(st));
    st.factor_dtype = SPTOL_F32;
    st.exact_fit = 1;   // the returned fit: <X,K> by an FP64 pass over the nonzeros
    auto t0 = std::chrono::steady_clock::now();
    if (!r) r = sptolCpdAlsHicooDevice(H, rank, niters, tol, dm, ktensor->lambda, &st, nullptr);
    if (!r) report_singular(st.singular_solves);
    for (uint32_t m = 0; m < N && !r; ++m) {
        sptMatrix v = view(mats[m]);
        r = sptolDevMatrixDownload(&v, &dm[m]);
    }
    for (uint32_t m = 0; m < N; ++m) sptolDevMatrixFree(&dm[m]);
    sptolDevHicooFree(H);
    if (r) {
        for (uint32_t m = 0; m <= N; ++m)
            if (mats[m]) { sptFreeRankMatrix(mats[m]); free(mats[m]); }
        free(mats);
        return r;
    }
    if (verbose())
        fprintf(stdout, "[GPU  HiCOO SpTns CPD-ALS]: %.9lf s\n",
                std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    sptFreeRankMatrix(mats[N]);
    ktensor->fit = st.fit;
    if (ktensor->factors) {
        for (uint32_t m = 0; m <= ktensor->nmodes; ++m)
            if (ktensor->factors[m]) { sptFreeRankMatrix(ktensor->factors[m]); free(ktensor->factors[m]); }
        free(ktensor->factors);
    }
    ktensor->factors = mats;   // owned by the Kruskal tensor (freed by sptFreeRankKruskalTensor)
    return 0;
}

int sptOmpCpdAlsHiCOO(sptSparseTensorHiCOO const * const hitsr, sptIndex const rank, sptIndex const niters,
                      double const tol, const int tk, int balanced, sptRankKruskalTensor * ktensor)
{
    (void)tk;
    (void)balanced;
    return sptCpdAlsHiCOO(hitsr, rank, niters, tol, ktensor);
}

}  // extern "C"
