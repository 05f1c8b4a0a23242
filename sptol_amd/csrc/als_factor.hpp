// The ALS factor step shared by the CPD kernels (cpd.hip) and the 4-mode
// MTTKRP launches that form it beside their ranges (mttkrp_seg.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace sptol {

// V = Hadamard of ata[p] for p = (m+k) % N, k = 1..N-1, and its inverse in
// FP64: by Gauss-Jordan elimination when V is positive definite (the
// Cholesky case of matrix_solver.c:50-62), else from an LU factorisation
// with partial pivoting (the gesv fallback of matrix_solver.c:66-79).
// Output: fac = V^{-1} (R*R, row-major), piv (R), kind[0] = 0 (positive
// definite) / 1 (LU) / 2 (exactly singular: no inverse, the solve leaves its
// right-hand sides), and kind[1] incremented for each singular case.
//
// One wave (64 threads, R <= RM <= 64), lane i holding row i in registers;
// the wave needs no barriers (LDS operations of a wave complete in order).
__device__ __forceinline__ void chol_wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Column c of V^{-1} from P V = L U (unit L and U in LDS, leading dimension
// RM + 1; pv: the row swaps in order): L U x = P e_c, in FP64.
template <int RM>
__device__ __forceinline__ void inv_column_lu(const double *__restrict__ LU, const int *__restrict__ pv, uint32_t R,
                                              uint32_t c, double *__restrict__ fac)
{
#pragma clang fp contract(off)
    constexpr uint32_t LD = RM + 1;
    uint32_t pos = c;   // where the swaps move the 1 of e_c
    for (uint32_t j = 0; j < R; ++j) {
        const uint32_t p = (uint32_t)pv[j];
        if (pos == j) pos = p;
        else if (pos == p) pos = j;
    }
    double y[RM];
#pragma unroll
    for (int a = 0; a < RM; ++a) {
        y[a] = 0.0;
        if (a < (int)R) {
            double s = a == (int)pos ? 1.0 : 0.0;
#pragma unroll
            for (int k = 0; k < a; ++k) s -= LU[a * LD + k] * y[k];
            y[a] = s;
        }
    }
#pragma unroll
    for (int a = RM - 1; a >= 0; --a) {
        if (a < (int)R) {
            double s = y[a];
#pragma unroll
            for (int k = a + 1; k < RM; ++k)
                if (k < (int)R) s -= LU[a * LD + k] * y[k];
            y[a] = s / LU[a * LD + a];
        }
    }
#pragma unroll
    for (int a = 0; a < RM; ++a)
        if (a < (int)R) fac[a * R + c] = y[a];
}

template <int RM>
__device__ __forceinline__ void factor_body(const double *__restrict__ ata, uint32_t N, uint32_t m, uint32_t R,
                                            double *__restrict__ fac, int *__restrict__ piv, int *__restrict__ kind)
{
#pragma clang fp contract(off)
    __shared__ double col[2][RM];         // the pivot rows of successive steps, alternately
    __shared__ double V[RM * (RM + 1)];   // the LU fallback; rows padded against bank conflicts
    __shared__ int pv[RM];
    const uint32_t RR = R * R, t = threadIdx.x;
    const bool mine = t < R;
    // row t of the Hadamard of the other modes' Grams, multiplied in mode
    // order; each Gram's row is loaded as one batch of independent loads
    double row[RM];
#pragma unroll
    for (int k = 0; k < RM; ++k) row[k] = (mine && k < (int)R) ? 1.0 : 0.0;
    for (uint32_t q = 1; q < N; ++q) {
        const double *__restrict__ g = ata + (size_t)((m + q) % N) * RR + (size_t)(mine ? t : 0) * R;
        double gv[RM];
#pragma unroll
        for (int k = 0; k < RM; ++k) gv[k] = (mine && k < (int)R) ? g[k] : 1.0;
#pragma unroll
        for (int k = 0; k < RM; ++k) row[k] *= gv[k];
    }
    // V^{-1} by Gauss-Jordan elimination in place, without pivoting: at step
    // j lane j publishes its row and every lane updates its own: lane j
    // scales its row by 1 / v_jj (its diagonal becoming 1 / v_jj), every
    // other lane i subtracts (v_ij / v_jj) times the pivot row by fused
    // multiply-adds (its column-j entry becoming -v_ij / v_jj).  The pivots are the diagonals of the
    // successive Schur complements, all positive exactly when V is SPD (the
    // condition of the Cholesky factorisation the reference tries first); a
    // pivot <= 0 sends V to the LU fallback below.  Entries k >= R are 0 in
    // every row and stay 0, so the updates run over all RM columns without a
    // test per entry; the pivot rows alternate between two LDS rows, so one
    // wave-level sync per step orders the write before the reads and the
    // reads before the rewrite two steps later.
    bool spd = true;                                    // wave-uniform
#pragma clang loop unroll(full)
    for (int j = 0; j < RM; ++j) {
        if (j < (int)R && spd) {
            double *__restrict__ pr = col[j & 1];
            if (t == (uint32_t)j) {
#pragma unroll
                for (int k = 0; k < RM; ++k) pr[k] = row[k];
            }
            chol_wave_sync();
            // the whole pivot row is read at once, its latency overlapping
            // the reciprocal of the pivot
            double prow[RM];
#pragma unroll
            for (int k = 0; k < RM; ++k) prow[k] = pr[k];
            const double p = prow[j];
            spd = __builtin_amdgcn_readfirstlane((int)(p > 0.0)) != 0;
            if (spd) {
                const double rp = 1.0 / p;
                if (t == (uint32_t)j) {
#pragma unroll
                    for (int k = 0; k < RM; ++k) row[k] = k == j ? rp : row[k] * rp;
                } else {
                    // one fused multiply-add per entry (one rounding): the
                    // entries are independent, so they issue back to back
                    const double f = row[j] * rp;
#pragma unroll
                    for (int k = 0; k < RM; ++k) row[k] = k == j ? -f : __builtin_fma(-f, prow[k], row[k]);
                }
            }
        }
    }
    constexpr uint32_t LD = RM + 1;
    if (spd) {
        if (mine) {
#pragma unroll
            for (int k = 0; k < RM; ++k)
                if (k < (int)R) fac[t * R + k] = row[k];
        }
        if (t == 0) kind[0] = 0;
        return;
    }
    // LU with partial pivoting of V (recomputed; in place): at step j swap row
    // j with the pivot row, then the lanes eliminate the rows below j.
    for (uint32_t e = t; e < RR; e += 64) {
        const uint32_t i = e / R, k = e % R;
        double v = 1.0;
        for (uint32_t q = 1; q < N; ++q) v *= ata[(size_t)((m + q) % N) * RR + e];
        V[i * LD + k] = v;
    }
    chol_wave_sync();
    for (uint32_t j = 0; j < R; ++j) {
        uint32_t best = j;
        for (uint32_t i = j + 1; i < R; ++i)
            if (fabs(V[i * LD + j]) > fabs(V[best * LD + j])) best = i;
        if (V[best * LD + j] == 0.0) {
            // exactly singular (every lane read the same LDS value): like
            // LAPACK's gesv with info > 0 the right-hand sides stay as they
            // are (kind 2); kind[1] counts such solves for the host
            if (t == 0) { kind[0] = 2; kind[1] += 1; }
            return;
        }
        if (t == 0) {
            piv[j] = (int)best;
            pv[j] = (int)best;
        }
        if (best != j)
            for (uint32_t k = t; k < R; k += 64) {
                const double tmp = V[j * LD + k];
                V[j * LD + k] = V[best * LD + k];
                V[best * LD + k] = tmp;
            }
        chol_wave_sync();
        for (uint32_t i = j + 1 + t; i < R; i += 64) {
            const double l = V[i * LD + j] / V[j * LD + j];
            V[i * LD + j] = l;
            for (uint32_t k = j + 1; k < R; ++k) V[i * LD + k] -= l * V[j * LD + k];
        }
        chol_wave_sync();
    }
    if (mine) inv_column_lu<RM>(V, pv, R, t, fac);
    if (t == 0) kind[0] = 1;
}

}  // namespace sptol
