// Segmented MTTKRP over nonzeros sorted by the output mode (3- and 4-mode
// tensors, row stride <= 64).
//
// Semantics are those of the reference (src/sptensor/mttkrp.c:39-177):
//   out[i_mode][r] = sum_x v_x * prod_{k=1..N-1} A_{order[k]}[i_{order[k]}][r]
// with every output row written, untouched rows as zeros and padding
// columns as zeros.  Each product is formed in the reference's association
// ((v*A_o1)*A_o2)*A_o3 and, with contraction off, a row whose nonzeros all
// fall in one lane group's range is summed in the reference's order: the
// sort is stable, so such rows match the sequential reference bit for bit.
//
// Layout: one 16-byte record per nonzero ({i0,i1,i2,bits(v)} for 3 modes,
// {i0,i1,i2,i3} plus a value array for 4), sorted by the output mode, and a
// bitmap of the output rows that hold at least one nonzero.
//
// Kernel (one pass, no atomics, no separate zero fill):
//  * A wave owns a contiguous range of G*chg nonzeros; each of its G lane
//    groups (LPN = stride/4 lanes, 4 columns per lane) owns chg of them.
//  * A group loads U records with one coalesced access (lane l loads record
//    z+l), passes them through LDS and reads them back as broadcasts, then
//    issues all U*(N-1) factor-row gathers before using any.
//  * Products are summed in registers while the output row repeats.  A row
//    that starts and ends inside the group is stored with one plain 16-byte
//    store per lane.  The group's first row, when it began in the previous
//    group, and its last row stay in registers.
//  * At the end the wave merges its groups' boundary rows through LDS in a
//    fixed order: the group holding a row's first nonzero stores the row.
//    A row that began in an earlier wave is written as this wave's carry.
//  * The wave then zero-fills the untouched rows of its slice of the output
//    (rows_per_wave rows), so every output row is written exactly once.
//  * k_seg_carry adds each run of carries, summed in a fixed order, to the
//    row's stored value.  The result is deterministic.
#include "dev_common.hpp"
#include <cstdlib>
#include <cstring>

namespace sptol {

void SegView::release()
{
    dfree(rec); rec = nullptr;
    if (own_vals) dfree(vals);
    vals = nullptr;
    own_vals = false;
    dfree(touched); touched = nullptr;
    dfree(carry); carry = nullptr;
    dfree(carry_row); carry_row = nullptr;
    carry_waves = 0;
    ready = false;
}

struct SegArgs {
    const uint4 *rec;
    const float *vals;
    uint64_t nnz;
    const void *A1, *A2, *A3;
    uint32_t o1, o2, o3, mode;
    uint32_t R, S, nrows, chg;
    uint32_t rows_per_wave, pad;
    uint64_t nwaves;
    float *out;
    const uint32_t *touched;
    float *carry;            // nwaves x 64 floats
    uint32_t *carry_row;     // nwaves
};

constexpr uint32_t kCarryStride = 64;

__device__ __forceinline__ uint32_t seg_comp(const uint4 r, uint32_t k)
{
    return k == 0 ? r.x : (k == 1 ? r.y : (k == 2 ? r.z : r.w));
}

__device__ __forceinline__ float4 seg_ld(const float *p) { return *reinterpret_cast<const float4 *>(p); }

__device__ __forceinline__ float4 seg_ld(const uint16_t *p)
{
    const uint2 u = *reinterpret_cast<const uint2 *>(p);
    return make_float4(__uint_as_float(u.x << 16), __uint_as_float(u.x & 0xFFFF0000u),
                       __uint_as_float(u.y << 16), __uint_as_float(u.y & 0xFFFF0000u));
}

__device__ __forceinline__ void seg_st(float *p, const float4 v) { *reinterpret_cast<float4 *>(p) = v; }

__device__ __forceinline__ void seg_add(float4 &a, const float4 b)
{
#pragma clang fp contract(off)
    a.x += b.x; a.y += b.y; a.z += b.z; a.w += b.w;
}

// LDS hand-off between lanes of one wave: LDS operations of a wave complete
// in order; the fences keep the compiler from moving accesses across.
__device__ __forceinline__ void seg_wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int LPN, int NM, int U, typename FT>
__global__ __launch_bounds__(256) void k_mttkrp_seg(const SegArgs a)
{
#pragma clang fp contract(off)
    constexpr uint32_t G = 64 / LPN;
    constexpr uint32_t SL = U + 1;            // staging slots per group (+1 spreads banks)
    __shared__ uint4 stage[4][G][SL];
    __shared__ float stagev[4][G][NM == 4 ? SL : 1];
    __shared__ float4 s_first[4][64];
    __shared__ float4 s_last[4][64];
    __shared__ uint32_t s_info[4][G][2];

    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    const uint64_t wave = (uint64_t)blockIdx.x * 4u + wv;
    if (wave >= a.nwaves) return;             // whole wave
    const uint32_t g = lane / LPN, l = lane % LPN;
    const uint32_t S = a.S, R = a.R, c0 = l * 4u;
    const bool col_ok = c0 < S;
    const uint32_t cc = col_ok ? c0 : 0u;     // idle lanes gather column 0
    const bool m0 = c0 < R, m1 = c0 + 1 < R, m2 = c0 + 2 < R, m3 = c0 + 3 < R;
    const FT *__restrict__ A1 = (const FT *)a.A1;
    const FT *__restrict__ A2 = (const FT *)a.A2;
    const FT *__restrict__ A3 = (const FT *)a.A3;
    const uint32_t mode = a.mode, o1 = a.o1, o2 = a.o2, o3 = a.o3;
    const uint4 *__restrict__ rec = a.rec;
    float *__restrict__ out = a.out;

    const uint64_t zs = (wave * G + g) * (uint64_t)a.chg;
    const uint64_t ze = zs < a.nnz ? min<uint64_t>(zs + a.chg, a.nnz) : zs;
    const bool nonempty = zs < ze;

    const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
    float4 acc = zero, facc = zero;
    uint32_t cur = kNoRow, frow = kNoRow;
    bool first_cont = false, in_first = true;
    if (nonempty) {
        frow = seg_comp(rec[zs], mode);
        first_cont = zs > 0 && seg_comp(rec[zs - 1], mode) == frow;
    }

    for (uint64_t z = zs; z < ze; z += U) {
#pragma unroll
        for (uint32_t k = 0; k < (U + LPN - 1) / LPN; ++k) {
            const uint32_t idx = l + k * LPN;
            if (idx < U) {
                // slots past the range repeat the last nonzero (in bounds, unused)
                const uint64_t zz = min<uint64_t>(z + idx, ze - 1);
                stage[wv][g][idx] = rec[zz];
                if (NM == 4) stagev[wv][g][idx] = a.vals[zz];
            }
        }
        seg_wave_sync();
        uint4 rr[U];
        float vv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            rr[u] = stage[wv][g][u];
            vv[u] = NM == 4 ? stagev[wv][g][u] : __uint_as_float(rr[u].w);
        }
        seg_wave_sync();
        float4 p[U], q[U];
#pragma unroll
        for (int u = 0; u < U; ++u) p[u] = seg_ld(A1 + (size_t)seg_comp(rr[u], o1) * S + cc);
#pragma unroll
        for (int u = 0; u < U; ++u) q[u] = seg_ld(A2 + (size_t)seg_comp(rr[u], o2) * S + cc);
        if (NM == 4) {
            float4 t[U];
#pragma unroll
            for (int u = 0; u < U; ++u) t[u] = seg_ld(A3 + (size_t)seg_comp(rr[u], o3) * S + cc);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                p[u].x = ((vv[u] * p[u].x) * q[u].x) * t[u].x;
                p[u].y = ((vv[u] * p[u].y) * q[u].y) * t[u].y;
                p[u].z = ((vv[u] * p[u].z) * q[u].z) * t[u].z;
                p[u].w = ((vv[u] * p[u].w) * q[u].w) * t[u].w;
            }
        } else {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                p[u].x = (vv[u] * p[u].x) * q[u].x;
                p[u].y = (vv[u] * p[u].y) * q[u].y;
                p[u].z = (vv[u] * p[u].z) * q[u].z;
                p[u].w = (vv[u] * p[u].w) * q[u].w;
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            // padding columns (>= R) hold unspecified factor values: keep them 0
            p[u].x = m0 ? p[u].x : 0.f;
            p[u].y = m1 ? p[u].y : 0.f;
            p[u].z = m2 ? p[u].z : 0.f;
            p[u].w = m3 ? p[u].w : 0.f;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (z + u < ze) {
                const uint32_t row = seg_comp(rr[u], mode);
                if (row != cur) {
                    if (cur != kNoRow) {
                        if (in_first && first_cont) facc = acc;
                        else if (col_ok) seg_st(out + (size_t)cur * S + c0, acc);
                        in_first = false;
                    }
                    cur = row;
                    acc = p[u];
                } else {
                    seg_add(acc, p[u]);
                }
            }
        }
    }

    // Merge the groups' boundary rows.  flags: 1 nonempty, 2 first row
    // continues from the previous nonzero, 4 single row.
    const bool single = in_first;
    const uint32_t flags = (nonempty ? 1u : 0u) | (first_cont ? 2u : 0u) | (single ? 4u : 0u);
    s_last[wv][lane] = acc;
    s_first[wv][lane] = facc;
    if (l == 0) {
        s_info[wv][g][0] = frow;
        s_info[wv][g][1] = flags;
    }
    seg_wave_sync();
    if (nonempty && !(single && first_cont)) {
        // this group holds the first nonzero of its last row
        float4 tot = acc;
        for (uint32_t k = g + 1; k < G; ++k) {
            const uint32_t f = s_info[wv][k][1];
            if ((f & 3u) != 3u) break;
            if (f & 4u) {
                seg_add(tot, s_last[wv][k * LPN + l]);
            } else {
                seg_add(tot, s_first[wv][k * LPN + l]);
                break;
            }
        }
        if (col_ok) seg_st(out + (size_t)cur * S + c0, tot);
    }
    if (g == 0) {
        uint32_t crow = kNoRow;
        if (nonempty && first_cont) {
            float4 tot = single ? acc : facc;
            if (single) {
                for (uint32_t k = 1; k < G; ++k) {
                    const uint32_t f = s_info[wv][k][1];
                    if ((f & 3u) != 3u) break;
                    if (f & 4u) {
                        seg_add(tot, s_last[wv][k * LPN + l]);
                    } else {
                        seg_add(tot, s_first[wv][k * LPN + l]);
                        break;
                    }
                }
            }
            crow = frow;
            if (col_ok) seg_st(a.carry + wave * kCarryStride + c0, tot);
        }
        if (l == 0) a.carry_row[wave] = crow;
    }

    // zero the untouched rows of this wave's slice
    const uint64_t r0 = wave * (uint64_t)a.rows_per_wave;
    const uint64_t r1 = min<uint64_t>(r0 + a.rows_per_wave, a.nrows);
    for (uint64_t r = r0 + g; r < r1; r += G) {
        const uint32_t w = a.touched[r >> 5];
        if (!((w >> (r & 31u)) & 1u) && col_ok) seg_st(out + r * S + c0, zero);
    }
}

// Add each run of carries (waves whose first row began in an earlier wave)
// to the row.  One block per wave index; only the first wave of a run works.
template <int LPN>
__global__ __launch_bounds__(256) void k_seg_carry(const float *__restrict__ carry, const uint32_t *__restrict__ carry_row,
                                                   uint64_t nwaves, float *__restrict__ out, uint32_t S)
{
#pragma clang fp contract(off)
    constexpr uint32_t G = 256 / LPN;
    __shared__ float4 part[256];
    const uint64_t k = blockIdx.x;
    const uint32_t r = carry_row[k];
    if (r == kNoRow || (k > 0 && carry_row[k - 1] == r)) return;
    // the run is contiguous: find its end 256 slots at a time
    uint64_t e = k + 1;
    for (uint64_t base = k + 1;; base += 256) {
        const uint64_t j = base + threadIdx.x;
        const bool in = j < nwaves && carry_row[j] == r;
        const int cnt = __syncthreads_count(in);
        e = base + (uint64_t)cnt;
        if (cnt < 256) break;
    }
    const uint32_t t = threadIdx.x, g = t / LPN, l = t % LPN, c0 = l * 4u;
    float4 tot = make_float4(0.f, 0.f, 0.f, 0.f);
    if (c0 < S) {
        for (uint64_t j = k + g; j < e; j += G) seg_add(tot, seg_ld(carry + j * kCarryStride + c0));
    }
    part[t] = tot;
    __syncthreads();
    if (g == 0 && c0 < S) {
        float4 s = part[l];
        for (uint32_t gg = 1; gg < G; ++gg) seg_add(s, part[gg * LPN + l]);
        float4 o = seg_ld(out + (size_t)r * S + c0);
        seg_add(o, s);
        seg_st(out + (size_t)r * S + c0, o);
    }
}

__global__ void k_seg_pack(uint4 *__restrict__ rec, float *__restrict__ vout, const uint32_t *__restrict__ i0,
                           const uint32_t *__restrict__ i1, const uint32_t *__restrict__ i2,
                           const uint32_t *__restrict__ i3, const float *__restrict__ vals, uint64_t nnz, uint32_t nm)
{
    for (uint64_t z = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; z < nnz; z += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t w = nm == 3 ? __float_as_uint(vals[z]) : i3[z];
        rec[z] = make_uint4(i0[z], i1[z], i2[z], w);
        if (vout) vout[z] = vals[z];
    }
}

__global__ void k_seg_touch(uint32_t *__restrict__ touched, const uint32_t *__restrict__ rows, uint64_t nnz)
{
    for (uint64_t z = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; z < nnz; z += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t r = rows[z];
        if (z == 0 || rows[z - 1] != r) atomicOr(&touched[r >> 5], 1u << (r & 31u));
    }
}

static unsigned seg_grid(uint64_t n)
{
    uint64_t b = (n + 255) / 256;
    if (b > 65536) b = 65536;
    return (unsigned)(b ? b : 1);
}

bool seg_ok(uint32_t nmodes, uint32_t stride)
{
    return (nmodes == 3 || nmodes == 4) && stride >= 4 && stride <= 64;
}

int seg_build(SegView &v, uint32_t nmodes, uint32_t mode, uint32_t nrows, uint64_t nnz, uint32_t *const *inds,
              const float *vals, bool copy_vals, hipStream_t s)
{
    v.release();
    if (nmodes != 3 && nmodes != 4) SPTOL_CHECK(SPTERR_VALUE_ERROR, "Seg view", "nmodes must be 3 or 4");
    v.nmodes = nmodes;
    v.mode = mode;
    v.nrows = nrows;
    v.nnz = nnz;
    int r = 0;
    const uint64_t words = ((uint64_t)nrows + 31) / 32;
    if ((r = dmalloc((void **)&v.touched, (words ? words : 1) * 4, "Seg view"))) { v.release(); return r; }
    hipError_t e = hipMemsetAsync(v.touched, 0, (words ? words : 1) * 4, s);
    if (e != hipSuccess) { v.release(); SPTOL_HIP(e, "Seg view"); }
    if (nnz) {
        if ((r = dmalloc((void **)&v.rec, nnz * sizeof(uint4), "Seg view"))) { v.release(); return r; }
        float *vout = nullptr;
        if (nmodes == 4) {
            if (copy_vals) {
                if ((r = dmalloc((void **)&v.vals, nnz * 4, "Seg view"))) { v.release(); return r; }
                v.own_vals = true;
                vout = v.vals;
            } else {
                v.vals = const_cast<float *>(vals);
            }
        }
        hipLaunchKernelGGL(k_seg_pack, dim3(seg_grid(nnz)), dim3(256), 0, s, v.rec, vout, inds[0], inds[1], inds[2],
                           nmodes == 4 ? inds[3] : inds[0], vals, nnz, nmodes);
        hipLaunchKernelGGL(k_seg_touch, dim3(seg_grid(nnz)), dim3(256), 0, s, v.touched, inds[mode], nnz);
    }
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) { v.release(); SPTOL_HIP(e, "Seg view"); }
    v.ready = true;
    return 0;
}

#define SEG_LAUNCH(LPN, NM, U, FT)                                                                   \
    hipLaunchKernelGGL((k_mttkrp_seg<LPN, NM, U, FT>), grid, dim3(256), 0, s, a)

#define SEG_NM(LPN, FT)                                                                              \
    do {                                                                                             \
        if (v.nmodes == 3) {                                                                         \
            if (useg == 8) SEG_LAUNCH(LPN, 3, 8, FT); else SEG_LAUNCH(LPN, 3, 4, FT);                \
        } else {                                                                                     \
            if (useg == 8) SEG_LAUNCH(LPN, 4, 8, FT); else SEG_LAUNCH(LPN, 4, 4, FT);                \
        }                                                                                            \
    } while (0)

#define SEG_LPN(FT)                                                                                  \
    do {                                                                                             \
        switch (lpn) {                                                                               \
        case 2: SEG_NM(2, FT); break;                                                                \
        case 4: SEG_NM(4, FT); break;                                                                \
        case 8: SEG_NM(8, FT); break;                                                                \
        default: SEG_NM(16, FT); break;                                                              \
        }                                                                                            \
    } while (0)

static int env_int(const char *name)
{
    const char *e = std::getenv(name);
    return e ? std::atoi(e) : 0;
}

// Nonzeros per lane group: about 32 waves per compute unit, between 64 and
// 512 (SPTOL_SEG_CHG overrides, for tuning).
static uint32_t seg_chg(uint64_t nnz, uint32_t G)
{
    static const int env = env_int("SPTOL_SEG_CHG");
    uint64_t c;
    if (env > 0) {
        c = (uint64_t)env;
    } else {
        const uint64_t target = (uint64_t)num_cus() * 32;
        c = (nnz + G * target - 1) / (G * target);
        if (c < 64) c = 64;
        if (c > 512) c = 512;
    }
    c = (c + 7) / 8 * 8;
    return (uint32_t)c;
}

int seg_mttkrp(SegView &v, const void *const *mats, const uint32_t *order, uint32_t R, uint32_t stride, int dtype,
               float *out, hipStream_t s)
{
    if (!v.ready) SPTOL_CHECK(SPTERR_VALUE_ERROR, "Seg MTTKRP", "view not built");
    if (!seg_ok(v.nmodes, stride)) SPTOL_CHECK(SPTERR_VALUE_ERROR, "Seg MTTKRP", "unsupported shape");
    uint32_t lpn = 2;
    while (lpn * 4 < stride) lpn <<= 1;
    const uint32_t G = 64 / lpn;
    const uint32_t chg = seg_chg(v.nnz, G);
    const uint64_t per_wave = (uint64_t)G * chg;
    uint64_t nwaves = (v.nnz + per_wave - 1) / per_wave;
    const uint64_t by_rows = ((uint64_t)v.nrows + 4095) / 4096;
    if (nwaves < by_rows) nwaves = by_rows;
    if (nwaves == 0) nwaves = 1;
    if (nwaves > 0x7FFFFFFFull) SPTOL_CHECK(SPTERR_VALUE_ERROR, "Seg MTTKRP", "tensor too large");
    if (v.carry_waves < nwaves) {
        dfree(v.carry); v.carry = nullptr;
        dfree(v.carry_row); v.carry_row = nullptr;
        v.carry_waves = 0;
        SPTOL_TRY(dmalloc((void **)&v.carry, nwaves * kCarryStride * 4, "Seg MTTKRP"));
        SPTOL_TRY(dmalloc((void **)&v.carry_row, nwaves * 4, "Seg MTTKRP"));
        v.carry_waves = nwaves;
    }
    SegArgs a;
    std::memset(&a, 0, sizeof(a));
    a.rec = v.rec;
    a.vals = v.vals;
    a.nnz = v.nnz;
    a.A1 = mats[order[1]];
    a.A2 = mats[order[2]];
    a.A3 = v.nmodes == 4 ? mats[order[3]] : mats[order[2]];
    a.o1 = order[1];
    a.o2 = order[2];
    a.o3 = v.nmodes == 4 ? order[3] : order[2];
    a.mode = v.mode;
    a.R = R;
    a.S = stride;
    a.nrows = v.nrows;
    a.chg = chg;
    a.rows_per_wave = (uint32_t)(((uint64_t)v.nrows + nwaves - 1) / nwaves);
    a.nwaves = nwaves;
    a.out = out;
    a.touched = v.touched;
    a.carry = v.carry;
    a.carry_row = v.carry_row;
    // records per step: 4 (SPTOL_SEG_U=8 for the deeper unroll)
    static const int useg = env_int("SPTOL_SEG_U") == 8 ? 8 : 4;
    dim3 grid((unsigned)((nwaves + 3) / 4));
    int tok = prof_start("mttkrp_seg", s);
    if (dtype == SPTOL_BF16) SEG_LPN(uint16_t);
    else SEG_LPN(float);
    prof_stop(tok, s);
    SPTOL_HIP(hipGetLastError(), "Seg MTTKRP launch");
    if (v.nnz) {
        int tok2 = prof_start("mttkrp_seg_carry", s);
        switch (lpn) {
        case 2: hipLaunchKernelGGL(k_seg_carry<2>, dim3((unsigned)nwaves), dim3(256), 0, s, v.carry, v.carry_row, nwaves, out, stride); break;
        case 4: hipLaunchKernelGGL(k_seg_carry<4>, dim3((unsigned)nwaves), dim3(256), 0, s, v.carry, v.carry_row, nwaves, out, stride); break;
        case 8: hipLaunchKernelGGL(k_seg_carry<8>, dim3((unsigned)nwaves), dim3(256), 0, s, v.carry, v.carry_row, nwaves, out, stride); break;
        default: hipLaunchKernelGGL(k_seg_carry<16>, dim3((unsigned)nwaves), dim3(256), 0, s, v.carry, v.carry_row, nwaves, out, stride); break;
        }
        prof_stop(tok2, s);
        SPTOL_HIP(hipGetLastError(), "Seg carry launch");
    }
    return 0;
}

}  // namespace sptol
