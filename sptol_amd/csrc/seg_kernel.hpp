// Device code of the segmented MTTKRP, shared by the single-tensor launch
// (mttkrp_seg.hip) and the batched launch (batch.hip).
//
// Semantics are those of the reference (src/sptensor/mttkrp.c:39-177):
//   out[i_mode][r] = sum_x v_x * prod_{k=1..N-1} A_{order[k]}[i_{order[k]}][r]
// with every output row written, untouched rows as zeros and padding
// columns as zeros.  Each product is formed in the reference's association
// ((v*A_o1)*A_o2)*A_o3 and, with contraction off, a row whose nonzeros all
// fall in one lane group's range is summed in the reference's order: the
// sort is stable, so such rows match the sequential reference bit for bit.
//
// Input: one 16-byte record per nonzero ({i0,i1,i2,bits(v)} for 3 modes,
// {i0,i1,i2,i3} plus a value array for 4), sorted by the output mode, and a
// bitmap of the output rows that hold at least one nonzero.
//
// One "range" (seg_range) is the unit of work of a wave:
//  * The wave owns G*chg consecutive nonzeros; each of its G lane groups
//    owns chg of them.  A group has LPN lanes of C columns each: 4 FP32
//    columns (one 16-byte gather) or 8 BF16 columns (also one 16-byte
//    gather, so BF16 factors halve the gathers per nonzero).
//  * A group loads U records with one coalesced access (lane l loads record
//    z+l), passes them through LDS and reads them back as broadcasts, then
//    issues all U*(N-1) factor-row gathers before using any; the next
//    step's records are loaded behind the gathers.
//  * Products are summed in FP32 registers while the output row repeats.  A
//    row that starts and ends inside the group is stored with plain 16-byte
//    stores.  The group's first row, when it began in the previous group,
//    and its last row stay in registers.
//  * The wave then merges its groups' boundary rows through LDS in a fixed
//    order: the group holding a row's first nonzero stores the row.  A row
//    that began in an earlier range is written as this range's carry.
//  * Finally it zero-fills the untouched rows of its slice of the output
//    (rows_per_wave rows), so every output row is written exactly once.
// seg_carry_slots then adds each run of carries, summed in a fixed order, to
// the row's stored value.  The result is deterministic.
//
// When rows are short and of even length (a uniform tensor), ranges may
// instead be runs of whole rows (SegArgs::zstart, the first nonzero of each
// range): no row then continues across ranges, so no carry exists and the
// carry pass is not launched.
#pragma once

#include "dev_common.hpp"

namespace sptol {

struct SegArgs {
    const uint4 *rec;
    const float *vals;       // N=4 values
    uint64_t nnz;
    const void *A1, *A2, *A3;
    uint32_t o1, o2, o3, mode;
    uint32_t R, S, nrows, chg;
    uint32_t rows_per_wave, lpn;
    uint64_t nwaves;         // ranges
    float *out;
    const uint32_t *touched;
    float *carry;            // nwaves x kSegCarryStride floats
    uint32_t *carry_row;     // nwaves
    uint32_t nmodes;
    uint32_t glpn;           // lanes per group of the range kernel (lpn: of the carry kernel, 4 columns a lane)
    const uint64_t *zstart;  // null: range w is nonzeros [w*G*chg, (w+1)*G*chg); else [zstart[w], zstart[w+1])
    uint32_t lk;             // 1..3: factor A1/A2/A3 is read from a copy in LDS (lrows rows); 0: none
    uint32_t lrows;
    AlsFactorReq fj;         // with fj_on (4-mode launches): the first wave of block 0 also runs this
    uint32_t fj_on;          //   ALS factor step before taking ranges
};

// Largest factor the range kernels copy to LDS (blocks of 8 waves).
constexpr uint32_t kSegLdsFactorBytes = 64u << 10;

// SegArgs of one MTTKRP on view v with factors of type dtype (carry space
// left unset; mttkrp_seg.hip).
int seg_fill_args(const SegView &v, const void *const *mats, const uint32_t *order, uint32_t R, uint32_t stride,
                  int dtype, float *out, uint32_t chg, SegArgs &a);

constexpr uint32_t kSegCarryStride = 64;
constexpr int kSegU = 4;                 // records per lane-group step (default)
constexpr uint32_t kSegCarrySlots = 64;  // ranges per carry block

__device__ __forceinline__ uint32_t seg_comp(const uint4 r, uint32_t k)
{
    return k == 0 ? r.x : (k == 1 ? r.y : (k == 2 ? r.z : r.w));
}

__device__ __forceinline__ float4 seg_ld(const float *p) { return *reinterpret_cast<const float4 *>(p); }

__device__ __forceinline__ void seg_st(float *p, const float4 v) { *reinterpret_cast<float4 *>(p) = v; }

// With NT, output stores and record loads are non-temporal: neither is
// read again by the launch, so their lines leave L2 first and factor rows
// that are gathered again stay.  The classes whose factors live in HBM use
// it (the 100M-nonzero Zipf headline: 6.17 ms per 3-mode step against
// 6.35 ms with plain accesses); classes whose factors fit in L2 do not
// (the 4-D CPD at R=32 ran 5% slower in FP32 and 13% in BF16 with it, its
// outputs being read again at once by the solve).  Gathers stay plain:
// non-temporal gathers of rarely used rows made the headline 25% slower.
typedef float seg_f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t seg_u32x4 __attribute__((ext_vector_type(4)));

template <bool NT>
__device__ __forceinline__ void seg_st_hint(float *p, const float4 v)
{
    if (NT) {
        const seg_f32x4 w = {v.x, v.y, v.z, v.w};
        __builtin_nontemporal_store(w, reinterpret_cast<seg_f32x4 *>(p));
    } else {
        seg_st(p, v);
    }
}

template <bool NT>
__device__ __forceinline__ uint4 seg_ld_rec(const uint4 *p)
{
    if (NT) {
        const seg_u32x4 v = __builtin_nontemporal_load(reinterpret_cast<const seg_u32x4 *>(p));
        return make_uint4(v.x, v.y, v.z, v.w);
    }
    return *p;
}

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

// Columns of one factor row held by a lane: C values of type FT, loaded
// with one 16-byte access and widened to FP32 one column at a time where
// they are used (a BF16 value is the upper half of its FP32 value).
template <typename FT, int C>
struct SegCols;

template <>
struct SegCols<float, 4> {
    using Raw = float4;
    static __device__ __forceinline__ Raw load(const float *p) { return *reinterpret_cast<const float4 *>(p); }
    static __device__ __forceinline__ float at(const Raw &r, int k)
    {
        return k == 0 ? r.x : (k == 1 ? r.y : (k == 2 ? r.z : r.w));
    }
};

template <>
struct SegCols<uint16_t, 8> {
    using Raw = uint4;
    static __device__ __forceinline__ Raw load(const uint16_t *p) { return *reinterpret_cast<const uint4 *>(p); }
    static __device__ __forceinline__ float at(const Raw &r, int k)
    {
        const int h = k >> 1;
        const uint32_t w = h == 0 ? r.x : (h == 1 ? r.y : (h == 2 ? r.z : r.w));
        return __uint_as_float((k & 1) ? (w & 0xFFFF0000u) : (w << 16));
    }
};

// Columns per lane of the range kernel for factor type FT.
template <typename FT>
constexpr int seg_cols() { return sizeof(FT) == 2 ? 8 : 4; }

// FP32 sums of C columns (registers once the loops are unrolled).
template <int C>
struct SegAcc {
    float v[C];
    __device__ __forceinline__ void zero()
    {
#pragma unroll
        for (int k = 0; k < C; ++k) v[k] = 0.f;
    }
    __device__ __forceinline__ void add(const SegAcc &b)
    {
#pragma clang fp contract(off)
#pragma unroll
        for (int k = 0; k < C; ++k) v[k] += b.v[k];
    }
    template <bool NT = false>
    __device__ __forceinline__ void store(float *p) const
    {
#pragma unroll
        for (int j = 0; j < C / 4; ++j)
            seg_st_hint<NT>(p + 4 * j, make_float4(v[4 * j], v[4 * j + 1], v[4 * j + 2], v[4 * j + 3]));
    }
    __device__ __forceinline__ void to_lds(float4 *p) const
    {
#pragma unroll
        for (int j = 0; j < C / 4; ++j) p[j] = make_float4(v[4 * j], v[4 * j + 1], v[4 * j + 2], v[4 * j + 3]);
    }
    __device__ __forceinline__ void from_lds(const float4 *p)
    {
#pragma unroll
        for (int j = 0; j < C / 4; ++j) {
            const float4 t = p[j];
            v[4 * j] = t.x; v[4 * j + 1] = t.y; v[4 * j + 2] = t.z; v[4 * j + 3] = t.w;
        }
    }
    __device__ __forceinline__ void add_lds(const float4 *p)
    {
#pragma clang fp contract(off)
#pragma unroll
        for (int j = 0; j < C / 4; ++j) {
            const float4 t = p[j];
            v[4 * j] += t.x; v[4 * j + 1] += t.y; v[4 * j + 2] += t.z; v[4 * j + 3] += t.w;
        }
    }
};

// Per-wave LDS of the range body, for WPB waves of C columns a lane and U
// records a lane-group step.
template <int LPN, int NM, int WPB, int C, int U = kSegU>
struct SegShared {
    static constexpr uint32_t G = 64 / LPN;
    static constexpr uint32_t SL = U + 1;   // staging slots per group (+1 spreads banks)
    uint4 stage[WPB][G][SL];
    float stagev[WPB][G][NM == 4 ? SL : 1];
    float4 piece[WPB][64][C / 4];
    uint32_t info[WPB][G][2];
};

template <int LPN, int NM, typename FT, int WPB, int U = kSegU>
using SegSharedOf = SegShared<LPN, NM, WPB, seg_cols<FT>(), U>;

// LK = k (1..3): factor A_ok is read from sfac, its copy in LDS, instead of
// global memory (a range kernel's gathers are bound by the vector-memory
// return path; an LDS read of the same shape costs less there).
template <int LPN, int NM, typename FT, int WPB, int U = kSegU, bool NT = false, int LK = 0>
__device__ __forceinline__ void seg_range(const SegArgs &a, uint64_t wave, SegSharedOf<LPN, NM, FT, WPB, U> &sh,
                                          uint32_t wv, uint32_t lane, const FT *__restrict__ sfac = nullptr)
{
#pragma clang fp contract(off)
    constexpr int C = seg_cols<FT>();
    using Cols = SegCols<FT, C>;
    using Raw = typename Cols::Raw;
    constexpr uint32_t KL = (U + LPN - 1) / LPN;   // records each lane stages per step
    constexpr uint32_t G = 64 / LPN;
    const uint32_t g = lane / LPN, l = lane % LPN;
    const uint32_t S = a.S, R = a.R, c0 = l * (uint32_t)C;
    const bool col_ok = c0 < S;                  // S is a multiple of 8: a lane's columns are all in the row
    const uint32_t cc = col_ok ? c0 : 0u;        // idle lanes gather column 0
    const uint32_t ncol = R > c0 ? R - c0 : 0u;  // this lane's columns below R
    const FT *__restrict__ A1 = LK == 1 ? sfac : (const FT *)a.A1;
    const FT *__restrict__ A2 = LK == 2 ? sfac : (const FT *)a.A2;
    const FT *__restrict__ A3 = LK == 3 ? sfac : (const FT *)a.A3;
    const uint32_t mode = a.mode, o1 = a.o1, o2 = a.o2, o3 = a.o3;
    const uint4 *__restrict__ rec = a.rec;
    float *__restrict__ out = a.out;

    uint64_t zs, ze;
    if (a.zstart) {
        // a range of whole rows, split evenly over the groups
        const uint64_t rb = a.zstart[wave], re = a.zstart[wave + 1];
        const uint64_t cg = (re - rb + G - 1) / G;
        zs = min<uint64_t>(rb + g * cg, re);
        ze = min<uint64_t>(zs + cg, re);
    } else {
        zs = (wave * G + g) * (uint64_t)a.chg;
        ze = zs < a.nnz ? min<uint64_t>(zs + a.chg, a.nnz) : zs;
    }
    const bool nonempty = zs < ze;
    // A group's "piece" (what it adds to a row begun before it) goes to its
    // lanes' LDS slots as soon as it is known, instead of being held in
    // registers across the loop: the first row's sum when that row ends
    // inside the group, else (a single row) the whole sum at the end.
    SegAcc<C> acc;
    acc.zero();
    uint32_t cur = kNoRow, frow = kNoRow;
    bool first_cont = false, in_first = true;
    if (nonempty) {
        frow = seg_comp(rec[zs], mode);
        first_cont = zs > 0 && seg_comp(rec[zs - 1], mode) == frow;
    }

    // this lane's share of a step's records; slots past the range repeat
    // the last nonzero (in bounds, unused)
    uint4 nx[KL];
    float nv[KL];
    auto load_recs = [&](uint64_t zb) {
#pragma unroll
        for (uint32_t k = 0; k < KL; ++k) {
            const uint32_t idx = l + k * LPN;
            if (idx < U) {
                const uint64_t zz = min<uint64_t>(zb + idx, ze - 1);
                nx[k] = seg_ld_rec<NT>(rec + zz);
                if (NM == 4) nv[k] = a.vals[zz];
            }
        }
    };
    if (nonempty) load_recs(zs);
    // this group's staged records, word by word
    const uint32_t *stg = reinterpret_cast<const uint32_t *>(&sh.stage[wv][g][0]);
    const uint32_t *stg_row = stg + mode, *stg_o1 = stg + o1, *stg_o2 = stg + o2;
    const uint32_t *stg_o3 = stg + (NM == 4 ? o3 : 0u), *stg_w = stg + 3;

    for (uint64_t z = zs; z < ze; z += U) {
#pragma unroll
        for (uint32_t k = 0; k < KL; ++k) {
            const uint32_t idx = l + k * LPN;
            if (idx < U) {
                sh.stage[wv][g][idx] = nx[k];
                if (NM == 4) sh.stagev[wv][g][idx] = nv[k];
            }
        }
        seg_wave_sync();
        // each word is read from LDS at its own (wave-uniform) position in
        // the record: no per-record select over the mode order
        uint32_t rw[U], j1[U], j2[U], j3[U];
        float vv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            rw[u] = stg_row[4 * u];
            j1[u] = stg_o1[4 * u];
            j2[u] = stg_o2[4 * u];
            j3[u] = NM == 4 ? stg_o3[4 * u] : 0u;
            vv[u] = NM == 4 ? sh.stagev[wv][g][u] : __uint_as_float(stg_w[4 * u]);
        }
        seg_wave_sync();
        Raw p[U], q[U], t[NM == 4 ? U : 1];
#pragma unroll
        for (int u = 0; u < U; ++u) p[u] = Cols::load(A1 + (size_t)j1[u] * S + cc);
#pragma unroll
        for (int u = 0; u < U; ++u) q[u] = Cols::load(A2 + (size_t)j2[u] * S + cc);
        if (NM == 4) {
#pragma unroll
            for (int u = 0; u < U; ++u) t[u] = Cols::load(A3 + (size_t)j3[u] * S + cc);
        }
        if (z + U < ze) load_recs(z + U);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            // a repeated slot carries the last nonzero's row, so it only
            // adds its 0 to the open run
            const uint32_t row = rw[u];
            if (row != cur) {
                if (cur != kNoRow) {
                    if (in_first && first_cont) acc.to_lds(sh.piece[wv][lane]);
                    else if (col_ok) acc.template store<NT>(out + (size_t)cur * S + c0);
                    in_first = false;
                }
                cur = row;
                acc.zero();   // the row's sum starts from 0, as the reference's output does
            }
            // ((v * A_o1) * A_o2) * A_o3 per column; padding columns (>= R)
            // hold unspecified factor values, and slots past the range
            // repeat its last nonzero: both contribute 0
            const bool ok = z + u < ze;
#pragma unroll
            for (int k = 0; k < C; ++k) {
                float x = (vv[u] * Cols::at(p[u], k)) * Cols::at(q[u], k);
                if (NM == 4) x *= Cols::at(t[u], k);
                acc.v[k] += (ok && (uint32_t)k < ncol) ? x : 0.f;
            }
        }
    }

    // Merge the groups' boundary rows.  flags: 1 nonempty, 2 first row
    // continues from the previous nonzero, 4 single row.  A piece is read
    // only from groups whose first row continues (flags 1 and 2).
    const bool single = in_first;
    const uint32_t flags = (nonempty ? 1u : 0u) | (first_cont ? 2u : 0u) | (single ? 4u : 0u);
    if (single) acc.to_lds(sh.piece[wv][lane]);
    if (l == 0) {
        sh.info[wv][g][0] = frow;
        sh.info[wv][g][1] = flags;
    }
    seg_wave_sync();
    // the chain of pieces that extends a row past group `from - 1`
    auto chain = [&](SegAcc<C> tot, uint32_t from) {
        for (uint32_t k = from; k < G; ++k) {
            const uint32_t f = sh.info[wv][k][1];
            if ((f & 3u) != 3u) break;
            tot.add_lds(sh.piece[wv][k * LPN + l]);
            if (!(f & 4u)) break;
        }
        return tot;
    };
    if (nonempty && !(single && first_cont)) {
        // this group holds the first nonzero of its last row
        const SegAcc<C> tot = chain(acc, g + 1);
        if (col_ok) tot.template store<NT>(out + (size_t)cur * S + c0);
    }
    if (g == 0) {
        uint32_t crow = kNoRow;
        if (nonempty && first_cont) {
            SegAcc<C> tot;
            if (single) tot = chain(acc, 1);
            else tot.from_lds(sh.piece[wv][lane]);
            crow = frow;
            if (col_ok) tot.store(a.carry + wave * kSegCarryStride + c0);
        }
        if (l == 0) a.carry_row[wave] = crow;
    }
    seg_wave_sync();   // piece / info are rewritten by the next range

    // zero the untouched rows of this range's slice: the slice's bitmap
    // words are loaded once, one a lane (2048 rows at a time), and each row
    // reads its word by shuffle, so the stores do not wait on loads
    SegAcc<C> zero;
    zero.zero();
    const uint64_t r0 = wave * (uint64_t)a.rows_per_wave;
    const uint64_t r1 = min<uint64_t>(r0 + a.rows_per_wave, a.nrows);
    for (uint64_t cb = r0 & ~31ull; cb < r1; cb += 2048) {
        const uint64_t w0 = cb >> 5;
        const uint64_t nw = min<uint64_t>(64, (r1 - cb + 31) >> 5);   // words w0 .. w0+nw-1 < ceil(nrows/32)
        const uint32_t myw = lane < nw ? a.touched[w0 + lane] : 0u;
        const uint64_t lo = max<uint64_t>(cb, r0), hi = min<uint64_t>(cb + 2048, r1);
        for (uint64_t rb = lo; rb < hi; rb += G) {   // wave-uniform trip count
            const uint64_t r = rb + g;
            const bool in = r < hi;
            const int k = (int)(((in ? r : lo) >> 5) - w0);
            const uint32_t w = (uint32_t)__shfl((int)myw, k);
            if (in && !((w >> (r & 31u)) & 1u) && col_ok) zero.template store<NT>(out + r * S + c0);
        }
    }
}

// One batch of a group's ranges j, j + G, ..., j + (NB-1)G of the run of
// row r (r != kNoRow): loads at indices clamped into [0, nwaves), so that
// none waits on a branch, then the ranges of the run, a prefix of the batch,
// added to tot in ascending order (bitwise tests and selects, so no load is
// moved into a branch).  False once the run has ended.
template <int NB>
__device__ __forceinline__ bool seg_carry_batch(const float *__restrict__ carry, const uint32_t *__restrict__ carry_row,
                                                uint64_t nwaves, uint32_t r, uint64_t j, uint32_t G, uint32_t c0,
                                                float4 &tot)
{
#pragma clang fp contract(off)
    float4 v[NB];
    uint32_t rr[NB];
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const uint64_t jj = min<uint64_t>(j + (uint64_t)q * G, nwaves - 1);
        rr[q] = carry_row[jj];
        v[q] = seg_ld(carry + jj * kSegCarryStride + c0);
    }
    __builtin_amdgcn_sched_barrier(0);   // all loads issue before the first add
    uint32_t m = 0;
#pragma unroll
    for (int q = 0; q < NB; ++q) m |= ((uint32_t)(j + (uint64_t)q * G < nwaves) & (uint32_t)(rr[q] == r)) << q;
    const uint32_t n1 = (uint32_t)__builtin_ctz(~m);   // <= NB
    // ranges past the run add -0.0f, which leaves every sum's bits unchanged
    const float4 nz = make_float4(-0.f, -0.f, -0.f, -0.f);
#pragma unroll
    for (int q = 0; q < NB; ++q) seg_add(tot, (uint32_t)q < n1 ? v[q] : nz);
    return n1 == (uint32_t)NB;
}

// Carries of ranges [k0, k1) of one MTTKRP (k1 - k0 <= kSegCarrySlots), by a
// 256-thread block.  A run is a maximal sequence of consecutive ranges
// carrying the same row; the block whose window holds a run's first range
// sums the run in a fixed order and adds it to the row.  Runs that end
// inside the window are summed in parallel, one lane group each; the one
// run per window that may continue past it (long runs of heavy rows) is
// summed by the whole block (groups stride the run, then an ordered sum
// over groups).  `rows` is block LDS of kSegCarrySlots + 2 entries.
__device__ __forceinline__ void seg_carry_slots(const float *__restrict__ carry, const uint32_t *__restrict__ carry_row,
                                                uint64_t nwaves, float *__restrict__ out, uint32_t S, uint32_t lpn,
                                                uint64_t k0, uint64_t k1, uint32_t *rows, float4 *part)
{
#pragma clang fp contract(off)
    const uint32_t t = threadIdx.x;
    const uint32_t n = (uint32_t)(k1 - k0);
    if (t <= n + 1) {
        // rows[i] = carry_row[k0 + i - 1]: the range before the window and the one after it
        const uint64_t k = k0 + t;   // = index + 1
        rows[t] = (t == 0 && k0 == 0) || k - 1 >= nwaves ? kNoRow : carry_row[k - 1];
    }
    __syncthreads();
    const uint32_t G = 256 / lpn, g = t / lpn, l = t % lpn, c0 = l * 4u;
    // runs that start and end in the window
    for (uint32_t i = g; i < n; i += G) {
        const uint32_t r = rows[i + 1];
        if (r == kNoRow || rows[i] == r) continue;
        uint32_t j = i + 1;
        while (j < n && rows[j + 1] == r) ++j;
        if (j == n && rows[n + 1] == r) continue;   // continues past the window
        if (c0 < S) {
            float4 o = seg_ld(out + (size_t)r * S + c0);   // in flight with the carries
            float4 s = seg_ld(carry + (k0 + i) * kSegCarryStride + c0);
            // the run's other carries, 8 loads in flight at a time (indices
            // clamped into the run), added in ascending order; slots past
            // the run add -0.0f, which leaves every sum's bits unchanged
            const float4 nz = make_float4(-0.f, -0.f, -0.f, -0.f);
            for (uint32_t q0 = i + 1; q0 < j; q0 += 8) {
                float4 v[8];
#pragma unroll
                for (uint32_t t = 0; t < 8; ++t) {
                    const uint32_t q = min(q0 + t, j - 1);
                    v[t] = seg_ld(carry + (k0 + q) * kSegCarryStride + c0);
                }
#pragma unroll
                for (uint32_t t = 0; t < 8; ++t) seg_add(s, q0 + t < j ? v[t] : nz);
            }
            seg_add(o, s);
            seg_st(out + (size_t)r * S + c0, o);
        }
    }
    // the run that continues past the window, if it starts here
    if (n == 0 || rows[n + 1] == kNoRow || rows[n + 1] != rows[n]) return;   // block-uniform
    const uint32_t r = rows[n];
    uint32_t i = n - 1;
    while (i > 0 && rows[i] == r) --i;
    if (i == 0 && rows[0] == r) return;   // began before the window: its head block sums it
    const uint64_t k = k0 + i;
    // The run's ranges are contiguous (row r's nonzeros are), so from k on
    // the ranges that carry r form a prefix (a heavy row of a skewed tensor
    // spans thousands of ranges).  Each group takes ranges k + g, k + g + G,
    // ... (one, then 16 at a time), each range's row and carry loaded
    // together, and adds them in ascending order up to the first range of
    // another row.
    float4 tot = make_float4(0.f, 0.f, 0.f, 0.f);
    float4 o = make_float4(0.f, 0.f, 0.f, 0.f);
    if (c0 < S) {
        // every group reads the row (group 0 adds to it), in flight with the
        // first batch; one range a group covers a short run whole
        o = seg_ld(out + (size_t)r * S + c0);
        uint64_t j = k + g;
        if (seg_carry_batch<1>(carry, carry_row, nwaves, r, j, G, c0, tot))
            for (j += G; seg_carry_batch<16>(carry, carry_row, nwaves, r, j, G, c0, tot);) j += 16 * (uint64_t)G;
    }
    part[t] = tot;
    __syncthreads();
    if (g == 0 && c0 < S) {
        float4 s = part[l];
        for (uint32_t gg = 1; gg < G; ++gg) seg_add(s, part[gg * lpn + l]);
        seg_add(o, s);
        seg_st(out + (size_t)r * S + c0, o);
    }
}

}  // namespace sptol
