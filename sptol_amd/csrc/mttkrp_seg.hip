// Segmented MTTKRP of one tensor: the mode-sorted record view, its build,
// and the launches (range body and carry pass in seg_kernel.hpp).
//
// The main kernel is persistent: each wave takes its first range of
// nonzeros by its own index and the following ones from a work counter
// until they run out, so ranges of uneven cost balance without a burst of
// atomics at the start; the carry kernel that follows resets the counter
// for the next launch.  A launch over ranges of whole rows takes every
// range by index (they fit in the resident waves), leaves the counter at 0
// and has no carry kernel.
// Calls on one view must be ordered (one stream): they share the view's
// counter and carry buffers.
#include "seg_kernel.hpp"
#include "als_factor.hpp"
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>
#include <algorithm>

namespace sptol {

void SegView::release()
{
    dfree(rec); rec = nullptr;
    if (own_vals) dfree(vals);
    vals = nullptr;
    own_vals = false;
    dfree(touched); touched = nullptr;   // the counter lives at its end
    dfree(carry); carry = nullptr;
    dfree(carry_row); carry_row = nullptr;
    carry_waves = 0;
    counter = nullptr;
    dfree(rr_start); rr_start = nullptr;
    rr_key = rr_nwaves = 0;
    rr_rows = 0;
    rr_ok = false;
    ready = false;
}

constexpr int kSegWPB = 4;
// register budget of the 3-mode kernels, in waves per SIMD
#ifndef SPTOL_SEG_WAVES3
#define SPTOL_SEG_WAVES3 4
#endif
#ifndef SPTOL_SEG_WAVES4
#define SPTOL_SEG_WAVES4 4
#endif

// Records a lane-group step and the register budget (waves per SIMD) of a
// class.  More records a step keep more gathers in flight:
//  * 3-mode FP32 classes of 32 or more columns take 8 records at 3 waves
//    per SIMD (the 100M-nonzero Zipf headline: 6.34 ms per 3-mode step
//    against 6.74 ms at 4 records and 4 waves);
//  * 4-mode classes take 8 records at 2 waves per SIMD (the 4-D CPD of 50M
//    nonzeros, R = 32: MTTKRP 856 -> 783 us on the mode whose factors all
//    come from global memory, 700 -> 605 us FP32 and 540 -> 480 us BF16 on
//    the modes that read one factor from LDS);
//  * the others take 4 records at the default budget.
// The sums do not depend on it: records are added one at a time, in order.
template <int LPN, int NM, typename FT, int LK = 0>
constexpr int seg_u()
{
    return (NM == 4 || (sizeof(FT) == 4 && LPN >= 8)) ? 8 : kSegU;
}
template <int LPN, int NM, typename FT, int LK = 0>
constexpr int seg_waves()
{
    return seg_u<LPN, NM, FT, LK>() == 8 ? (NM == 4 ? 2 : 3) : (NM == 3 ? SPTOL_SEG_WAVES3 : SPTOL_SEG_WAVES4);
}
// Classes with a non-temporal variant (records and outputs) for factors in HBM.
template <int LPN, int NM, typename FT>
constexpr bool seg_has_nt() { return NM == 3 && sizeof(FT) == 4 && LPN >= 8; }

// Gathered factors above this many bytes (the L2 of all XCDs) live in HBM:
// the 8-record classes then load records and store outputs non-temporally.
constexpr uint64_t kSegNtFactorBytes = (uint64_t)32 << 20;

// Waves per block: 4, or 8 with a factor copied to LDS (one copy per block,
// one block per compute unit: 2 waves per SIMD, the budget of the 4-mode
// classes and all that a 3-mode copy's launch fills).
template <int NM, int LK>
constexpr int seg_wpb() { return LK ? 8 : kSegWPB; }
// Classes whose waves' staging fits in LDS beside the factor copy (all but
// 4-mode BF16 with one lane a group).
template <int LPN, int NM, typename FT>
constexpr bool seg_lds_class() { return NM == 3 || !(sizeof(FT) == 2 && LPN == 1); }
// Register budget in waves per SIMD: a 3-mode block with a factor copy holds
// a compute unit's LDS alone, so two waves per SIMD are all it can have.
template <int LPN, int NM, typename FT, int LK>
constexpr int seg_waves_lk() { return (NM == 3 && LK) ? 2 : seg_waves<LPN, NM, FT, LK>(); }

// 4-mode classes of at most 32 columns can run the ALS factor step of their
// mode (wider ones leave it to k_factor: its 64-column form would weigh on
// their registers and build time).
template <int LPN, int NM, typename FT>
constexpr bool seg_fused_factor() { return NM == 4 && seg_cols<FT>() * LPN <= 32; }

template <int LPN, int NM, typename FT, bool NT = false, int LK = 0>
__global__ __launch_bounds__((seg_wpb<NM, LK>() * 64)) __attribute__((amdgpu_waves_per_eu(seg_waves_lk<LPN, NM, FT, LK>(), 8)))
void k_mttkrp_seg(const SegArgs a, uint32_t *__restrict__ counter)
{
    constexpr int U = seg_u<LPN, NM, FT, LK>();
    constexpr int WPB = seg_wpb<NM, LK>();
    __shared__ SegSharedOf<LPN, NM, FT, WPB, U> sh;
    __shared__ __align__(16) FT sfac[LK ? kSegLdsFactorBytes / sizeof(FT) : 1];
    if constexpr (LK != 0) {
        // the block's copy of factor A_oLK (rows x S values, a multiple of 16 bytes)
        const FT *src = (const FT *)(LK == 1 ? a.A1 : (LK == 2 ? a.A2 : a.A3));
        const uint32_t n16 = (uint32_t)((uint64_t)a.lrows * a.S * sizeof(FT) / 16);
        for (uint32_t i = threadIdx.x; i < n16; i += WPB * 64)
            reinterpret_cast<uint4 *>(sfac)[i] = reinterpret_cast<const uint4 *>(src)[i];
        __syncthreads();
    }
    if constexpr (seg_fused_factor<LPN, NM, FT>()) {
        // the ALS factor step of this mode (it needs only the other modes'
        // Grams), in the first wave of block 0 before its ranges: the
        // persistent grid's other waves take the ranges meanwhile
        if (a.fj_on && blockIdx.x == 0 && threadIdx.x < 64)
            factor_body<seg_cols<FT>() * LPN>(a.fj.ata, a.fj.N, a.fj.m, a.fj.R, a.fj.fac, a.fj.piv, a.fj.kind);
    }
    // the wave index is uniform across the wave: read from lane 0, the
    // range's indices, bounds and arguments live in scalar registers
    const uint32_t lane = threadIdx.x & 63u, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // Wave w of the W in the grid takes range w first, then pulls ranges
    // W, W+1, ... from the counter: no atomics while every wave starts, none
    // at all in a launch of at most W ranges, balancing after that.
    const uint64_t W = (uint64_t)gridDim.x * WPB;
    uint64_t wave = (uint64_t)blockIdx.x * WPB + wv;
    while (wave < a.nwaves) {
        seg_range<LPN, NM, FT, WPB, U, NT, LK>(a, wave, sh, wv, lane, sfac);
        if (W >= a.nwaves) break;   // every range was taken by index
        uint32_t wq = 0;
        if (lane == 0) wq = atomicAdd(counter, 1u);
        wave = W + __builtin_amdgcn_readfirstlane(wq);
    }
}

// kSegCarrySlots ranges per block; block 0 also resets the work counter.
__global__ __launch_bounds__(256) void k_seg_carry(const SegArgs a, uint32_t *__restrict__ counter)
{
    __shared__ uint32_t rows[kSegCarrySlots + 2];
    __shared__ float4 part[256];
    if (blockIdx.x == 0 && threadIdx.x == 0) *counter = 0u;
    const uint64_t k0 = (uint64_t)blockIdx.x * kSegCarrySlots;
    const uint64_t k1 = min<uint64_t>(k0 + kSegCarrySlots, a.nwaves);
    seg_carry_slots(a.carry, a.carry_row, a.nwaves, a.out, a.S, a.lpn, k0, k1, rows, part);
}

// rec[z] (and vout[z]) from nonzero p = perm[z] (p = z without perm).
__global__ void k_seg_pack(uint4 *__restrict__ rec, float *__restrict__ vout, const uint32_t *__restrict__ i0,
                           const uint32_t *__restrict__ i1, const uint32_t *__restrict__ i2,
                           const uint32_t *__restrict__ i3, const float *__restrict__ vals,
                           const uint32_t *__restrict__ perm, uint64_t nnz, uint32_t nm)
{
    for (uint64_t z = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; z < nnz; z += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t p = perm ? (uint64_t)perm[z] : z;
        const uint32_t w = nm == 3 ? __float_as_uint(vals[p]) : i3[p];
        rec[z] = make_uint4(i0[p], i1[p], i2[p], w);
        if (vout) vout[z] = vals[p];
    }
}

// Records in the tensor's own order, and 16-bit copies of the indices of
// the modes with at most 2^16 rows (k16.k[m], null for the others: the sort
// keys of those modes).
struct Keys16 {
    uint16_t *k[4];
};

__global__ void k_seg_pack_keys(uint4 *__restrict__ rec, const uint32_t *__restrict__ i0,
                                const uint32_t *__restrict__ i1, const uint32_t *__restrict__ i2,
                                const uint32_t *__restrict__ i3, const float *__restrict__ vals, uint64_t nnz,
                                uint32_t nm, Keys16 k16)
{
    for (uint64_t z = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; z < nnz; z += (uint64_t)gridDim.x * blockDim.x) {
        const uint4 r = make_uint4(i0[z], i1[z], i2[z], nm == 3 ? __float_as_uint(vals[z]) : i3[z]);
        rec[z] = r;
        if (k16.k[0]) k16.k[0][z] = (uint16_t)r.x;
        if (k16.k[1]) k16.k[1][z] = (uint16_t)r.y;
        if (k16.k[2]) k16.k[2][z] = (uint16_t)r.z;
        if (nm == 4 && k16.k[3]) k16.k[3][z] = (uint16_t)r.w;
    }
}

// rec[z] = rec0[perm[z]] (and vout[z] = vals[perm[z]]): one 16-byte gather
// per nonzero from records packed in the tensor's own order.  Also the
// touched-row bitmap from the sorted rows (a row's first nonzero sets its
// bit); K is the type of the sorted row keys.
template <typename K>
__global__ void k_seg_gather_rec(uint4 *__restrict__ rec, float *__restrict__ vout, const uint4 *__restrict__ rec0,
                                 const float *__restrict__ vals, const uint32_t *__restrict__ perm,
                                 const K *__restrict__ rows, uint32_t *__restrict__ touched, uint64_t nnz)
{
    for (uint64_t z = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; z < nnz; z += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t p = perm[z];
        rec[z] = rec0[p];
        if (vout) vout[z] = vals[p];
        const uint32_t r = rows[z];
        if (z == 0 || (uint32_t)rows[z - 1] != r) atomicOr(&touched[r >> 5], 1u << (r & 31u));
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

static int env_int(const char *name)
{
    const char *e = std::getenv(name);
    return e ? std::atoi(e) : 0;
}

bool seg_ok(uint32_t nmodes, uint32_t stride)
{
    return (nmodes == 3 || nmodes == 4) && stride >= 8 && stride <= 64 && stride % 8 == 0;
}

uint32_t seg_lpn(uint32_t stride)
{
    uint32_t lpn = 2;
    while (lpn * 4 < stride) lpn <<= 1;
    return lpn;
}

uint32_t seg_glpn(uint32_t stride, int dtype)
{
    const uint32_t cols = dtype == SPTOL_BF16 ? 8u : 4u;
    uint32_t lpn = dtype == SPTOL_BF16 ? 1u : 2u;
    while (lpn * cols < stride) lpn <<= 1;
    return lpn;
}

int seg_lane_class(uint32_t glpn, int dtype)
{
    int i = 0;
    for (uint32_t l = dtype == SPTOL_BF16 ? 1u : 2u; l < glpn; l <<= 1) ++i;
    return i;
}

// Records of every nonzero in the tensor's own order (the layout of
// SegView::rec), for seg_build's gather through a sort permutation.
int seg_pack_unsorted(uint4 *rec0, uint32_t nmodes, uint64_t nnz, uint32_t *const *inds, const float *vals,
                      hipStream_t s, uint16_t *const *keys16)
{
    if (!nnz) return 0;
    Keys16 k16 = {};
    for (uint32_t m = 0; m < nmodes && keys16; ++m) k16.k[m] = keys16[m];
    hipLaunchKernelGGL(k_seg_pack_keys, dim3(seg_grid(nnz)), dim3(256), 0, s, rec0, inds[0], inds[1], inds[2],
                       nmodes == 4 ? inds[3] : inds[0], vals, nnz, nmodes, k16);
    SPTOL_HIP(hipGetLastError(), "Seg view");
    return 0;
}

static int seg_row_ranges(SegView &v, hipStream_t s);

int seg_build(SegView &v, uint32_t nmodes, uint32_t mode, uint32_t nrows, uint64_t nnz, uint32_t *const *inds,
              const float *vals, bool copy_vals, hipStream_t s, const uint32_t *ndims, const uint32_t *perm,
              const uint32_t *sorted_rows, const uint4 *rec0, const uint16_t *sorted_rows16, bool row_ranges)
{
    v.release();
    if (perm && ((!sorted_rows && !sorted_rows16) || !copy_vals))
        SPTOL_CHECK(SPTERR_VALUE_ERROR, "Seg view", "perm needs sorted rows and copied values");
    if (sorted_rows16 && (!perm || !rec0 || nrows > 65536u))
        SPTOL_CHECK(SPTERR_VALUE_ERROR, "Seg view", "16-bit rows need a permutation, packed records and <= 2^16 rows");
    if (nmodes != 3 && nmodes != 4) SPTOL_CHECK(SPTERR_VALUE_ERROR, "Seg view", "nmodes must be 3 or 4");
    v.nmodes = nmodes;
    v.mode = mode;
    v.nrows = nrows;
    v.nnz = nnz;
    for (uint32_t m = 0; m < nmodes; ++m) v.ndims[m] = ndims[m];
    int r = 0;
    // the row bitmap and, after it, the work counter: one allocation, one clear
    const uint64_t words = ((uint64_t)nrows + 31) / 32;
    if ((r = dmalloc((void **)&v.touched, (words + 1) * 4, "Seg view"))) {
        v.release();
        return r;
    }
    v.counter = v.touched + words;
    hipError_t e = hipMemsetAsync(v.touched, 0, (words + 1) * 4, s);
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
        if (perm && rec0 && sorted_rows16) {
            hipLaunchKernelGGL(k_seg_gather_rec<uint16_t>, dim3(seg_grid(nnz)), dim3(256), 0, s, v.rec, vout, rec0,
                               vals, perm, sorted_rows16, v.touched, nnz);
        } else if (perm && rec0) {
            hipLaunchKernelGGL(k_seg_gather_rec<uint32_t>, dim3(seg_grid(nnz)), dim3(256), 0, s, v.rec, vout, rec0,
                               vals, perm, sorted_rows, v.touched, nnz);
        } else {
            hipLaunchKernelGGL(k_seg_pack, dim3(seg_grid(nnz)), dim3(256), 0, s, v.rec, vout, inds[0], inds[1],
                               inds[2], nmodes == 4 ? inds[3] : inds[0], vals, perm, nnz, nmodes);
            hipLaunchKernelGGL(k_seg_touch, dim3(seg_grid(nnz)), dim3(256), 0, s, v.touched,
                               perm ? sorted_rows : inds[mode], nnz);
        }
    }
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) { v.release(); SPTOL_HIP(e, "Seg view"); }
    v.ready = true;
    return row_ranges ? seg_row_ranges(v, s) : 0;
}

// Nonzeros per lane group for a tensor of nnz nonzeros: about 32 ranges per
// compute unit, between 16 and 256; when the gathered factors fit in one
// XCD's L2, about 8 ranges per compute unit (half the resident waves) and
// up to 768 (SPTOL_SEG_CHG overrides, for tuning).  With factors in HBM,
// shorter ranges balance the waves better (the 12M x 9M x 28M HiCOO
// headline: 42.1 Gnnz/s at 256, 41.0 at 512); with factors in L2 the
// per-range work (merge, carry, zero fill) counts more (the 4-D CPD at R=32,
// per iteration: FP32 4.24 ms at 256 and 4.12 at 768, BF16 3.28 and 3.22;
// the 1M-nonzero 2000 x 1500 x 1000 MTTKRP at R=16: 19.6 us per call at 16
// nonzeros a group, 15.5 at 32, 16.2 at 48).  Ranges beyond the resident
// waves are handed out by one atomic counter, whose serialised updates cost
// more than short ranges save: that MTTKRP took 144 us at 8 a group.
uint32_t seg_chg(uint64_t nnz, uint32_t G, bool l2_factors)
{
    static const int env = env_int("SPTOL_SEG_CHG");
    uint64_t c;
    if (env > 0) {
        c = (uint64_t)env;
    } else {
        const uint64_t target = (uint64_t)num_cus() * (l2_factors ? 8 : 32);
        c = (nnz + G * target - 1) / (G * target);
        const uint64_t cap = l2_factors ? 768 : 256;
        if (c < 16) c = 16;
        if (c > cap) c = cap;
    }
    c = (c + 7) / 8 * 8;
    return (uint32_t)c;
}

int seg_fill_args(const SegView &v, const void *const *mats, const uint32_t *order, uint32_t R, uint32_t stride,
                  int dtype, float *out, uint32_t chg, SegArgs &a)
{
    if (!v.ready) SPTOL_CHECK(SPTERR_VALUE_ERROR, "Seg MTTKRP", "view not built");
    if (!seg_ok(v.nmodes, stride)) SPTOL_CHECK(SPTERR_VALUE_ERROR, "Seg MTTKRP", "unsupported shape");
    const uint32_t glpn = seg_glpn(stride, dtype);
    const uint32_t G = 64 / glpn;
    const uint64_t per_wave = (uint64_t)G * chg;
    uint64_t nwaves = (v.nnz + per_wave - 1) / per_wave;
    const uint64_t by_rows = ((uint64_t)v.nrows + 4095) / 4096;
    if (nwaves < by_rows) nwaves = by_rows;
    if (nwaves == 0) nwaves = 1;
    if (nwaves > 0x7FFFFFFFull) SPTOL_CHECK(SPTERR_VALUE_ERROR, "Seg MTTKRP", "tensor too large");
    std::memset(&a, 0, sizeof(a));
    a.rec = v.rec;
    a.vals = v.vals;
    a.nnz = v.nnz;
    a.nmodes = v.nmodes;
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
    a.lpn = seg_lpn(stride);
    a.glpn = glpn;
    a.rows_per_wave = (uint32_t)(((uint64_t)v.nrows + nwaves - 1) / nwaves);
    a.nwaves = nwaves;
    a.out = out;
    a.touched = v.touched;
    return 0;
}

// First nonzero of every range of whole rows: start[k] = the first nonzero
// whose row is >= k * rows (binary search over the row-sorted records),
// for k = 0 .. nranges.
__global__ __launch_bounds__(256) void k_seg_row_starts(const uint4 *__restrict__ rec, uint64_t nnz, uint32_t mode,
                                                        uint32_t rows, uint64_t nranges, uint64_t *__restrict__ start)
{
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k > nranges) return;
    const uint64_t target = k * rows;
    uint64_t lo = 0, hi = nnz;   // first z in [lo, hi] with row >= target
    while (lo < hi) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if ((uint64_t)seg_comp(rec[mid], mode) < target) lo = mid + 1;
        else hi = mid;
    }
    start[k] = k == nranges ? nnz : lo;
}

// Ranges of whole rows for the view: about 8 ranges per compute unit (the
// range count of a tensor whose factors fit in L2), each a run of rows
// averaging nnz / ranges nonzeros.  Kept (rr_ok) when no range holds more
// than 1.5 times the average (rows of even length, as in a uniform tensor)
// and every range fits in the resident waves of one launch of any range
// kernel class (at least 3 waves per SIMD).  The range kernels split each
// range evenly over their lane groups, so one table serves every rank and
// factor type.  Synchronises s.
static int seg_row_ranges(SegView &v, hipStream_t s)
{
    dfree(v.rr_start);
    v.rr_start = nullptr;
    v.rr_ok = false;
    v.rr_nwaves = 0;
    v.rr_rows = 0;
    if (!v.nnz || !v.nrows) return 0;
    const uint64_t target = (uint64_t)num_cus() * 8, max_ranges = (uint64_t)num_cus() * 12;
    const double per_range = (double)v.nnz / (double)target;
    uint64_t rows = (uint64_t)((double)v.nrows * per_range / (double)v.nnz);
    if (rows < 1) rows = 1;
    if (rows > v.nrows) rows = v.nrows;
    const uint64_t n = (v.nrows + rows - 1) / rows;
    if (n > max_ranges || rows > 0xFFFFFFFFull) return 0;
    SPTOL_TRY(dmalloc((void **)&v.rr_start, (n + 1) * 8, "Seg row ranges"));
    hipLaunchKernelGGL(k_seg_row_starts, dim3((unsigned)((n + 256) / 256)), dim3(256), 0, s, v.rec, v.nnz, v.mode,
                       (uint32_t)rows, n, v.rr_start);
    std::vector<uint64_t> h(n + 1);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), v.rr_start, (n + 1) * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        dfree(v.rr_start);
        v.rr_start = nullptr;
        SPTOL_HIP(e, "Seg row ranges");
    }
    uint64_t most = 0;
    for (uint64_t k = 0; k < n; ++k) most = std::max(most, h[k + 1] - h[k]);
    v.rr_rows = (uint32_t)rows;
    v.rr_nwaves = n;
    v.rr_ok = (double)most <= 1.5 * (double)v.nnz / (double)n;
    if (!v.rr_ok) {
        dfree(v.rr_start);
        v.rr_start = nullptr;
    }
    return 0;
}

// The view's carry space for a launch of a.nwaves ranges.
static int seg_carry_space(SegView &v, SegArgs &a)
{
    if (v.carry_waves < a.nwaves) {
        dfree(v.carry); v.carry = nullptr;
        dfree(v.carry_row); v.carry_row = nullptr;
        v.carry_waves = 0;
        SPTOL_TRY(dmalloc((void **)&v.carry, a.nwaves * kSegCarryStride * 4, "Seg MTTKRP"));
        SPTOL_TRY(dmalloc((void **)&v.carry_row, a.nwaves * 4, "Seg MTTKRP"));
        v.carry_waves = a.nwaves;
    }
    a.carry = v.carry;
    a.carry_row = v.carry_row;
    return 0;
}

// Waves of one launch of the class's range kernel that are resident at once.
template <int LPN, int NM, typename FT, bool NT, int LK = 0>
static uint64_t seg_resident_waves()
{
    constexpr int WPB = seg_wpb<NM, LK>();
    const void *fn = reinterpret_cast<const void *>(&k_mttkrp_seg<LPN, NM, FT, NT, LK>);
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, WPB * 64, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    return (uint64_t)num_cus() * per_cu * WPB;
}

template <int LPN, int NM, typename FT, bool NT, int LK = 0>
static void seg_launch_kernel(const SegArgs &a, uint32_t *counter, hipStream_t s)
{
    constexpr int WPB = seg_wpb<NM, LK>();
    uint64_t blocks = seg_resident_waves<LPN, NM, FT, NT, LK>() / WPB;
    const uint64_t need = (a.nwaves + WPB - 1) / WPB;
    if (blocks > need) blocks = need;
    hipLaunchKernelGGL((k_mttkrp_seg<LPN, NM, FT, NT, LK>), dim3((unsigned)blocks), dim3(WPB * 64), 0, s, a,
                       counter);
}

// nt: the gathered factors live in HBM (only the 8-record classes have a
// non-temporal variant).
template <int LPN, int NM, typename FT>
static void seg_launch_main(const SegArgs &a, uint32_t *counter, hipStream_t s, bool nt)
{
    if constexpr (seg_has_nt<LPN, NM, FT>()) {
        if (nt) {
            seg_launch_kernel<LPN, NM, FT, true>(a, counter, s);
            return;
        }
    }
    if constexpr (seg_lds_class<LPN, NM, FT>()) {
        switch (a.lk) {
        case 1: seg_launch_kernel<LPN, NM, FT, false, 1>(a, counter, s); return;
        case 2: seg_launch_kernel<LPN, NM, FT, false, 2>(a, counter, s); return;
        case 3:
            if constexpr (NM == 4) {
                seg_launch_kernel<LPN, NM, FT, false, 3>(a, counter, s);
                return;
            }
            break;
        default: break;
        }
    }
    seg_launch_kernel<LPN, NM, FT, false>(a, counter, s);
}

// lanes per group of the range kernels: FP32 {2, 4, 8, 16}, BF16 {1, 2, 4, 8}
template <typename FT, int I>
constexpr int seg_lanes() { return (sizeof(FT) == 2 ? 1 : 2) << I; }

template <int LPN, int NM, typename FT>
static uint64_t seg_resident_main(bool nt, uint32_t lk)
{
    if constexpr (seg_has_nt<LPN, NM, FT>()) {
        if (nt) return seg_resident_waves<LPN, NM, FT, true>();
    }
    if constexpr (seg_lds_class<LPN, NM, FT>()) {
        if (lk == 1) return seg_resident_waves<LPN, NM, FT, false, 1>();
        if (lk == 2) return seg_resident_waves<LPN, NM, FT, false, 2>();
        if constexpr (NM == 4)
            if (lk == 3) return seg_resident_waves<LPN, NM, FT, false, 3>();
    }
    return seg_resident_waves<LPN, NM, FT, false>();
}

// Resident waves of the range kernel that seg_launch_ft runs for a.
template <typename FT>
static uint64_t seg_resident_ft(const SegArgs &a, bool nt)
{
    const bool n3 = a.nmodes == 3;
    const uint32_t lk = a.lk;
    switch (seg_lane_class(a.glpn, sizeof(FT) == 2 ? SPTOL_BF16 : SPTOL_F32)) {
    case 0: return n3 ? seg_resident_main<seg_lanes<FT, 0>(), 3, FT>(nt, lk) : seg_resident_main<seg_lanes<FT, 0>(), 4, FT>(nt, lk);
    case 1: return n3 ? seg_resident_main<seg_lanes<FT, 1>(), 3, FT>(nt, lk) : seg_resident_main<seg_lanes<FT, 1>(), 4, FT>(nt, lk);
    case 2: return n3 ? seg_resident_main<seg_lanes<FT, 2>(), 3, FT>(nt, lk) : seg_resident_main<seg_lanes<FT, 2>(), 4, FT>(nt, lk);
    default: return n3 ? seg_resident_main<seg_lanes<FT, 3>(), 3, FT>(nt, lk) : seg_resident_main<seg_lanes<FT, 3>(), 4, FT>(nt, lk);
    }
}

template <typename FT>
static void seg_launch_ft(const SegArgs &a, uint32_t *counter, hipStream_t s, bool nt)
{
    const bool n3 = a.nmodes == 3;
    switch (seg_lane_class(a.glpn, sizeof(FT) == 2 ? SPTOL_BF16 : SPTOL_F32)) {
    case 0:
        n3 ? seg_launch_main<seg_lanes<FT, 0>(), 3, FT>(a, counter, s, nt) : seg_launch_main<seg_lanes<FT, 0>(), 4, FT>(a, counter, s, nt);
        break;
    case 1:
        n3 ? seg_launch_main<seg_lanes<FT, 1>(), 3, FT>(a, counter, s, nt) : seg_launch_main<seg_lanes<FT, 1>(), 4, FT>(a, counter, s, nt);
        break;
    case 2:
        n3 ? seg_launch_main<seg_lanes<FT, 2>(), 3, FT>(a, counter, s, nt) : seg_launch_main<seg_lanes<FT, 2>(), 4, FT>(a, counter, s, nt);
        break;
    default:
        n3 ? seg_launch_main<seg_lanes<FT, 3>(), 3, FT>(a, counter, s, nt) : seg_launch_main<seg_lanes<FT, 3>(), 4, FT>(a, counter, s, nt);
        break;
    }
}

// SPTOL_SEG_ROWS=0 turns ranges of whole rows off (for measurements).
static bool seg_whole_rows()
{
    static const bool on = [] {
        const char *e = std::getenv("SPTOL_SEG_ROWS");
        return e ? std::atoi(e) != 0 : true;
    }();
    return on;
}

// The smallest gathered factor is read from a copy in LDS when it fits
// kSegLdsFactorBytes (SPTOL_SEG_LDS=0 turns this off, for measurements):
//  * 4-mode views of at least 2M nonzeros (persistent blocks of 8 waves);
//  * 3-mode views whose ranges of whole rows fill at most 8 waves per
//    compute unit (blocks of 8 waves, one a compute unit: the launch is
//    bound by each unit's gather issue, half of which the copy takes over),
//    with at least twice as many nonzeros per block as the copy has rows.
static void seg_pick_lds(const SegView &v, SegArgs &a, int dtype)
{
    static const bool on = [] {
        const char *e = std::getenv("SPTOL_SEG_LDS");
        return e ? std::atoi(e) != 0 : true;
    }();
    a.lk = 0;
    a.lrows = 0;
    if (!on) return;
    if (v.nmodes == 4) {
        if (v.nnz < ((uint64_t)2 << 20) || (dtype == SPTOL_BF16 && a.glpn == 1)) return;
    } else if (v.nmodes == 3) {
        if (!seg_whole_rows() || !v.rr_ok || !v.rr_start || v.rr_nwaves > (uint64_t)num_cus() * 8) return;
    } else {
        return;
    }
    const uint32_t o[3] = {a.o1, a.o2, a.o3};
    uint64_t best = ~0ull;
    for (uint32_t k = 0; k + 1 < v.nmodes; ++k) {
        const uint64_t bytes = (uint64_t)v.ndims[o[k]] * a.S * (dtype == SPTOL_BF16 ? 2u : 4u);
        if (bytes <= kSegLdsFactorBytes && bytes < best) {
            best = bytes;
            a.lk = k + 1;
            a.lrows = v.ndims[o[k]];
        }
    }
    if (v.nmodes == 3 && a.lk && v.nnz < 2 * (uint64_t)a.lrows * ((v.rr_nwaves + 7) / 8)) {
        a.lk = 0;
        a.lrows = 0;
    }
}

// Whether a launch on v takes its ranges of whole rows (built with the view,
// and fitting in the class's resident waves).
static bool seg_use_row_ranges(const SegView &v, const SegArgs &a, int dtype, bool nt)
{
    if (!v.rr_ok || !v.rr_start) return false;
    const uint64_t W = dtype == SPTOL_BF16 ? seg_resident_ft<uint16_t>(a, nt) : seg_resident_ft<float>(a, nt);
    return v.rr_nwaves <= W;
}

// ---------------------------------------------------------------------------
// Allocation-free launches: the carry space and the work counter of the
// segmented kernel come from `scratch`, a fixed device array of this file,
// instead of the view's own buffers (which the view allocates on its first
// MTTKRP).  Every GPU the library runs on holds its own copy until the
// process ends.  Layout: kNoAllocRanges carries of kSegCarryStride floats,
// then kNoAllocRanges row ids, then the work counter.  The counter is zero
// when the module loads and the carry kernel resets it at the end of every
// call; carries and row ids are written by every range before the carry
// kernel reads them, so nothing else needs clearing.  Calls that overlap on
// one GPU share the array and race.
// ---------------------------------------------------------------------------
constexpr uint64_t kNoAllocRanges = 65536;
constexpr size_t kNoAllocBytes = kNoAllocRanges * (kSegCarryStride + 1) * 4 + 256;
static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

constexpr int kNoAllocMaxDevices = 64;
static uint8_t *g_scratch_ptr[kNoAllocMaxDevices];
static std::mutex g_scratch_mu;

// The device address of this device's `scratch` (the array's name is a host
// stand-in in host code): looked up once per device with that device current.
static int noalloc_scratch(uint8_t **p)
{
    int dev = 0;
    SPTOL_HIP(hipGetDevice(&dev), "MTTKRP no-alloc");
    if (dev < 0 || dev >= kNoAllocMaxDevices) SPTOL_CHECK(SPTERR_VALUE_ERROR, "MTTKRP no-alloc", "device index");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *q = nullptr;
        SPTOL_HIP(hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)), "MTTKRP no-alloc");
        if (!q) SPTOL_CHECK(SPTERR_UNKNOWN, "MTTKRP no-alloc", "scratch address");
        g_scratch_ptr[dev] = (uint8_t *)q;
    }
    *p = g_scratch_ptr[dev];
    return 0;
}

int seg_mttkrp_noalloc(const SegView &v, const void *const *mats, const uint32_t *order, uint32_t R,
                       uint32_t stride, int dtype, float *out, hipStream_t s)
{
    SegArgs a;
    uint64_t fbytes = 0;
    for (uint32_t k = 1; k < v.nmodes; ++k)
        fbytes += (uint64_t)v.ndims[order[k]] * stride * (dtype == SPTOL_BF16 ? 2u : 4u);
    const bool l2 = fbytes <= ((uint64_t)4 << 20), nt = fbytes > kSegNtFactorBytes;
    SPTOL_TRY(seg_fill_args(v, mats, order, R, stride, dtype, out, seg_chg(v.nnz, 64 / seg_glpn(stride, dtype), l2), a));
    seg_pick_lds(v, a, dtype);
    const bool rows = seg_whole_rows() && seg_use_row_ranges(v, a, dtype, nt);
    if (rows) {   // the same ranges as seg_mttkrp: results bit for bit alike
        a.zstart = v.rr_start;
        a.nwaves = v.rr_nwaves;
        a.rows_per_wave = v.rr_rows;
    }
    if (a.nwaves > kNoAllocRanges)
        SPTOL_CHECK(SPTERR_VALUE_ERROR, "MTTKRP no-alloc",
                    "the tensor needs more ranges than the fixed scratch holds (at most 65536)");
    uint8_t *base = nullptr;
    SPTOL_TRY(noalloc_scratch(&base));
    a.carry = (float *)base;
    a.carry_row = (uint32_t *)(base + kNoAllocRanges * kSegCarryStride * 4);
    uint32_t *counter = (uint32_t *)(base + kNoAllocRanges * (kSegCarryStride + 1) * 4);
    if (dtype == SPTOL_BF16) seg_launch_ft<uint16_t>(a, counter, s, nt);
    else seg_launch_ft<float>(a, counter, s, nt);
    SPTOL_HIP(hipGetLastError(), "MTTKRP no-alloc launch");
    if (rows) return 0;   // no row continues across ranges: nothing to carry
    const uint64_t cblocks = (a.nwaves + kSegCarrySlots - 1) / kSegCarrySlots;
    hipLaunchKernelGGL(k_seg_carry, dim3((unsigned)cblocks), dim3(256), 0, s, a, counter);
    SPTOL_HIP(hipGetLastError(), "MTTKRP no-alloc carry launch");
    return 0;
}

int seg_mttkrp(SegView &v, const void *const *mats, const uint32_t *order, uint32_t R, uint32_t stride, int dtype,
               float *out, hipStream_t s, const AlsFactorReq *fr, bool *factored)
{
    if (factored) *factored = false;
    SegArgs a;
    // bytes of the factor rows the launch gathers from
    uint64_t fbytes = 0;
    for (uint32_t k = 1; k < v.nmodes; ++k)
        fbytes += (uint64_t)v.ndims[order[k]] * stride * (dtype == SPTOL_BF16 ? 2u : 4u);
    const bool l2 = fbytes <= ((uint64_t)4 << 20), nt = fbytes > kSegNtFactorBytes;
    SPTOL_TRY(seg_fill_args(v, mats, order, R, stride, dtype, out, seg_chg(v.nnz, 64 / seg_glpn(stride, dtype), l2), a));
    seg_pick_lds(v, a, dtype);
    // ranges of whole rows where rows are short and even: no carries, and
    // every range is taken by wave index (so the work counter stays 0)
    const bool rows = seg_whole_rows() && seg_use_row_ranges(v, a, dtype, nt);
    if (rows) {
        a.zstart = v.rr_start;
        a.nwaves = v.rr_nwaves;
        a.rows_per_wave = v.rr_rows;
    }
    SPTOL_TRY(seg_carry_space(v, a));
    if (fr && v.nmodes == 4 && fr->R <= (dtype == SPTOL_BF16 ? 8u : 4u) * a.glpn &&
        (dtype == SPTOL_BF16 ? 8u : 4u) * a.glpn <= 32u) {   // the classes of seg_fused_factor
        a.fj = *fr;
        a.fj_on = 1;
        if (factored) *factored = true;
    }
    int tok = prof_start("mttkrp_seg", s);
    if (dtype == SPTOL_BF16) seg_launch_ft<uint16_t>(a, v.counter, s, nt);
    else seg_launch_ft<float>(a, v.counter, s, nt);
    prof_stop(tok, s);
    SPTOL_HIP(hipGetLastError(), "Seg MTTKRP launch");
    if (rows) return 0;   // no row continues across ranges: nothing to carry
    // always launched: it also resets the work counter
    int tok2 = prof_start("mttkrp_seg_carry", s);
    const uint64_t cblocks = (a.nwaves + kSegCarrySlots - 1) / kSegCarrySlots;
    hipLaunchKernelGGL(k_seg_carry, dim3((unsigned)cblocks), dim3(256), 0, s, a, v.counter);
    prof_stop(tok2, s);
    SPTOL_HIP(hipGetLastError(), "Seg carry launch");
    return 0;
}

}  // namespace sptol
