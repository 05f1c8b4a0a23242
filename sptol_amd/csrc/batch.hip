// Batched MTTKRP: many independent COO tensors in one persistent launch.
//
// The reference runs one tensor per call and copies it to the device each
// time (src/sptensor/mttkrp_cuda_onekernel.cu:100-149).  Here a batch of
// device-resident tensors (each with its own factors, rank, mode and output)
// runs as persistent kernels sized to the chip that pull work from global
// counters, so small and large tensors and different ranks share the launch
// without idle CUs, and the launch can be captured once into a hipGraph and
// replayed.
//
//  * Tensors with a mode-sorted view (sptolDevCooPrepareMode; 3- or 4-mode,
//    stride <= 64) run the segmented range body of seg_kernel.hpp: one
//    persistent kernel per class of (lanes per nonzero, modes) takes ranges
//    of any tensor of its class from the class counter; one carry kernel over
//    every tensor then finishes rows that cross ranges and resets the
//    counters.  Every output row is written exactly once (no zero fill, no
//    atomics) and results are deterministic.
//  * Other tensors take the atomic path: a fill kernel zeroes their outputs
//    and resets its queue, then a persistent kernel pulls tiles of
//    consecutive nonzeros and flushes row runs with float atomics.
//
// The range length (nonzeros per lane group) and the number of resident
// blocks per CU come from an autotuner that times candidate configurations
// on the batch itself and caches the winner per batch signature.
#include "seg_kernel.hpp"
#include <vector>
#include <map>
#include <algorithm>
#include <cstring>
#include <cmath>
#include <chrono>

namespace sptol {


// Zero every output and reset the queue counter.  Work is split in chunks
// of 64K floats; chunk c of the batch belongs to tensor t with
// fill_prefix[t] <= c < fill_prefix[t+1].
__global__ __launch_bounds__(256) void k_batch_fill(const BatchDesc *__restrict__ d, const uint64_t *__restrict__ fill_prefix,
                                                    uint32_t ntensors, uint64_t nchunks, uint32_t *__restrict__ counter)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *counter = 0u;
    for (uint64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        uint32_t lo = 0, hi = ntensors;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (fill_prefix[mid] <= c) lo = mid; else hi = mid;
        }
        const uint64_t base = (c - fill_prefix[lo]) * 65536ull;
        const uint64_t n = d[lo].out_len;
        float4 *o = reinterpret_cast<float4 *>(d[lo].out + base);
        const uint64_t n4 = (min<uint64_t>(base + 65536ull, n) - base) / 4;
        for (uint64_t i = threadIdx.x; i < n4; i += 256) o[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

__device__ __forceinline__ float4 bld4(const float *p) { return *reinterpret_cast<const float4 *>(p); }
__device__ __forceinline__ float4 bld4(const uint16_t *p)
{
    const uint2 u = *reinterpret_cast<const uint2 *>(p);
    return make_float4(__uint_as_float(u.x << 16), __uint_as_float(u.x & 0xFFFF0000u),
                       __uint_as_float(u.y << 16), __uint_as_float(u.y & 0xFFFF0000u));
}

__device__ __forceinline__ void bflush(float *out, uint32_t row, uint32_t S, uint32_t c0, uint32_t R, const float4 a)
{
    float *o = out + (size_t)row * S + c0;
    if (c0 + 0 < R) unsafeAtomicAdd(o + 0, a.x);
    if (c0 + 1 < R) unsafeAtomicAdd(o + 1, a.y);
    if (c0 + 2 < R) unsafeAtomicAdd(o + 2, a.z);
    if (c0 + 3 < R) unsafeAtomicAdd(o + 3, a.w);
}

// Persistent kernel.  A wave takes one tile at a time from the queue; its
// 64/lpn lane groups each walk a contiguous slice of the tile, summing
// products of equal output rows in registers (as the single-tensor kernel).
//
// Every loop below has a wave-uniform trip count and per-lane work is
// predicated, never skipped with a divergent continue or early loop exit:
// the queue pull at the top relies on all 64 lanes arriving there together
// (lane 0 takes the tile, readfirstlane broadcasts it).
template <typename FT>
__global__ __launch_bounds__(256) void k_batch_mttkrp(const BatchDesc *__restrict__ d,
                                                      const uint64_t *__restrict__ tile_prefix, uint32_t ntensors,
                                                      uint64_t ntiles, uint32_t *__restrict__ counter)
{
    const uint32_t lane = threadIdx.x & 63u;
    for (;;) {
        uint32_t tile = 0;
        if (lane == 0) tile = atomicAdd(counter, 1u);
        tile = __builtin_amdgcn_readfirstlane(tile);
        if (tile >= ntiles) return;
        uint32_t lo = 0, hi = ntensors;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (tile_prefix[mid] <= tile) lo = mid; else hi = mid;
        }
        const BatchDesc &t = d[lo];
        const uint32_t lpn = t.lpn, S = t.stride, R = t.R;
        const uint32_t groups = 64u / lpn;
        const uint32_t g = lane / lpn, l = lane % lpn;
        const uint64_t tz0 = (tile - tile_prefix[lo]) * t.tile_nnz;
        const uint64_t tz1 = min<uint64_t>(tz0 + t.tile_nnz, t.nnz);
        const uint64_t slice = (t.tile_nnz + groups - 1) / groups;   // uniform
        const uint64_t z0 = tz0 + (uint64_t)g * slice;
        const uint64_t z1 = min<uint64_t>(z0 + slice, tz1);
        const uint32_t c0 = l * 4;                                    // < S always
        const bool lane_on = c0 < R;
        const uint32_t *__restrict__ rowi = t.inds[t.mode];
        const uint32_t o1 = t.order[1], o2 = t.order[2];
        const uint32_t *__restrict__ i1 = t.inds[o1];
        const uint32_t *__restrict__ i2 = t.inds[o2];
        const FT *__restrict__ A1 = (const FT *)t.mats[o1];
        const FT *__restrict__ A2 = (const FT *)t.mats[o2];
        const bool four = t.nmodes == 4;
        const uint32_t o3 = four ? t.order[3] : o2;
        const uint32_t *__restrict__ i3 = t.inds[o3];
        const FT *__restrict__ A3 = (const FT *)t.mats[o3];
        float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
        uint32_t cur = 0xFFFFFFFFu;
        constexpr int U = 4;
        for (uint64_t k = 0; k < slice; k += U) {
            uint32_t row[U];
            float4 p[U];
            bool valid[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint64_t z = z0 + k + u;
                valid[u] = lane_on && z < z1;
                const uint64_t zz = valid[u] ? z : tz0;   // tz0 < nnz: a safe index
                row[u] = rowi[zz];
                const float v = t.vals[zz];
                float4 a = bld4(A1 + (size_t)i1[zz] * S + c0);
                const float4 b = bld4(A2 + (size_t)i2[zz] * S + c0);
                a.x = v * a.x; a.y = v * a.y; a.z = v * a.z; a.w = v * a.w;
                a.x *= b.x; a.y *= b.y; a.z *= b.z; a.w *= b.w;
                if (four) {
                    const float4 c = bld4(A3 + (size_t)i3[zz] * S + c0);
                    a.x *= c.x; a.y *= c.y; a.z *= c.z; a.w *= c.w;
                }
                p[u] = a;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (valid[u]) {
                    if (row[u] != cur) {
                        if (cur != 0xFFFFFFFFu) bflush(t.out, cur, S, c0, R, acc);
                        cur = row[u];
                        acc = p[u];
                    } else {
                        acc.x += p[u].x; acc.y += p[u].y; acc.z += p[u].z; acc.w += p[u].w;
                    }
                }
            }
        }
        if (cur != 0xFFFFFFFFu) bflush(t.out, cur, S, c0, R, acc);
    }
}

// ---------------------------------------------------------------------------
// Segmented part
// ---------------------------------------------------------------------------
constexpr int kBatchWPB = 4;

// register budget of the batched range kernels, in waves per SIMD: 4 (128
// VGPRs) measured faster than the 3 the compiler would choose for 4-mode
// tensors, spills included
#ifndef SPTOL_BATCH_WAVES
#define SPTOL_BATCH_WAVES 4
#endif

// Records a lane-group step: FP32 classes keep 8 records' gathers in flight,
// 3-mode ones at 3 waves per SIMD (the 256-tensor batch of 390M nonzeros:
// 11.05 ms per step against 11.20 with 8 only at 32+ columns and 11.5 with
// 4 everywhere), 4-mode ones at 2 (served CPD stream of 400 requests: 189
// solves/s against 180 with 4 records at 4 waves); BF16 classes take 4 at
// the budget above.  The sums are the same either way (records are added
// one at a time, in order).
template <int LPN, int NM, typename FT>
constexpr int batch_u() { return sizeof(FT) == 4 ? 8 : kSegU; }
template <int LPN, int NM, typename FT>
constexpr int batch_waves() { return batch_u<LPN, NM, FT>() == 8 ? (NM == 4 ? 2 : 3) : SPTOL_BATCH_WAVES; }

template <int LPN, int NM, typename FT>
__global__ __launch_bounds__(kBatchWPB * 64) __attribute__((amdgpu_waves_per_eu(batch_waves<LPN, NM, FT>(), 8))) void k_batch_seg(const SegArgs *__restrict__ items,
                                                              const uint64_t *__restrict__ prefix, uint32_t n,
                                                              uint64_t total, uint32_t *__restrict__ counter)
{
    constexpr int U = batch_u<LPN, NM, FT>();
    __shared__ SegSharedOf<LPN, NM, FT, kBatchWPB, U> sh;
    // the wave index is uniform across the wave: read from lane 0, the
    // range's indices, bounds and arguments live in scalar registers
    const uint32_t lane = threadIdx.x & 63u, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // first range by wave index, then W, W+1, ... from the class counter
    const uint64_t W = (uint64_t)gridDim.x * kBatchWPB;
    for (uint64_t t = (uint64_t)blockIdx.x * kBatchWPB + wv; t < total;) {
        uint32_t lo = 0, hi = n;   // prefix[lo] <= t < prefix[hi]
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (prefix[mid] <= t) lo = mid; else hi = mid;
        }
        const SegArgs a = items[lo];
        seg_range<LPN, NM, FT, kBatchWPB, U>(a, t - prefix[lo], sh, wv, lane);
        if (W >= total) break;   // every range was taken by index
        uint32_t wq = 0;
        if (lane == 0) wq = atomicAdd(counter, 1u);
        t = W + __builtin_amdgcn_readfirstlane(wq);
    }
}

// Carry blocks of every segmented item (class order); block 0 also resets
// the class counters for the next launch.
__global__ __launch_bounds__(256) void k_batch_seg_carry(const SegArgs *__restrict__ items,
                                                         const uint64_t *__restrict__ cprefix, uint32_t n,
                                                         uint32_t *__restrict__ counters)
{
    __shared__ uint32_t rows[kSegCarrySlots + 2];
    __shared__ float4 part[256];
    if (blockIdx.x == 0 && threadIdx.x < 8) counters[threadIdx.x] = 0u;
    const uint64_t b = blockIdx.x;
    uint32_t lo = 0, hi = n;   // cprefix[lo] <= b < cprefix[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (cprefix[mid] <= b) lo = mid; else hi = mid;
    }
    const SegArgs a = items[lo];
    const uint64_t k0 = (b - cprefix[lo]) * kSegCarrySlots;
    const uint64_t k1 = min<uint64_t>(k0 + kSegCarrySlots, a.nwaves);
    seg_carry_slots(a.carry, a.carry_row, a.nwaves, a.out, a.S, a.lpn, k0, k1, rows, part);
}

SegBatchDev::~SegBatchDev()
{
    dfree(arena); dfree(carry); dfree(carry_row); dfree(counters);
}

static int seg_class(uint32_t glpn, uint32_t nmodes, int dtype)
{
    return seg_lane_class(glpn, dtype) * 2 + (nmodes == 4 ? 1 : 0);
}

// Nonzeros per lane group of a batch: 256 scaled by the tuned factor.
static uint32_t seg_tile_chg(double scale)
{
    double c = 256.0 * scale;
    if (c < 64) c = 64;
    if (c > 2048) c = 2048;
    return ((uint32_t)c + 7) / 8 * 8;
}

// Plan and upload the segmented items (a stream-ordered copy from pinned
// staging).
// scale[c]: the tile factor of class c.
static int seg_batch_prepare(const std::vector<const BatchDesc *> &items, int dtype, const double *scale,
                             SegBatchDev &dev, hipStream_t s)
{
    std::vector<SegArgs> cls[8];
    uint64_t ranges = 0;
    for (const BatchDesc *d : items) {
        const int c = seg_class(seg_glpn(d->stride, dtype), d->nmodes, dtype);
        SegArgs a;
        SPTOL_TRY(seg_fill_args(*d->seg, d->mats, d->order, d->R, d->stride, dtype, d->out, seg_tile_chg(scale[c]), a));
        ranges += a.nwaves;
        cls[c].push_back(a);
    }
    if (dev.carry_cap < ranges) {
        dfree(dev.carry); dev.carry = nullptr;
        dfree(dev.carry_row); dev.carry_row = nullptr;
        dev.carry_cap = 0;
        SPTOL_TRY(dmalloc((void **)&dev.carry, ranges * kSegCarryStride * 4, "Batch"));
        SPTOL_TRY(dmalloc((void **)&dev.carry_row, ranges * 4, "Batch"));
        dev.carry_cap = ranges;
    }
    if (!dev.counters) {
        SPTOL_TRY(dmalloc((void **)&dev.counters, 8 * 4, "Batch"));
        SPTOL_HIP(hipMemsetAsync(dev.counters, 0, 8 * 4, s), "Batch");
    }
    // carry space, in class order
    std::vector<SegArgs> all;
    std::vector<uint64_t> cprefix(1, 0);
    uint64_t off = 0;
    for (int c = 0; c < 8; ++c)
        for (SegArgs &a : cls[c]) {
            a.carry = dev.carry + off * kSegCarryStride;
            a.carry_row = dev.carry_row + off;
            off += a.nwaves;
            all.push_back(a);
            cprefix.push_back(cprefix.back() + (a.nwaves + kSegCarrySlots - 1) / kSegCarrySlots);
        }
    // staging: per class items + prefix, then all items + carry prefix
    std::vector<uint8_t> host;
    auto put = [&](const void *p, size_t bytes) {
        const size_t at = (host.size() + 255) / 256 * 256;
        host.resize(at + bytes);
        if (bytes) std::memcpy(host.data() + at, p, bytes);
        return at;
    };
    for (int c = 0; c < 8; ++c) {
        std::vector<uint64_t> pre(1, 0);
        for (const SegArgs &a : cls[c]) pre.push_back(pre.back() + a.nwaves);
        dev.n[c] = (uint32_t)cls[c].size();
        dev.total[c] = pre.back();
        if (dev.total[c] > 0xFFFFFFFFull) SPTOL_CHECK(SPTERR_VALUE_ERROR, "Batch", "too many ranges in one class");
        dev.off_items[c] = put(cls[c].data(), cls[c].size() * sizeof(SegArgs));
        dev.off_prefix[c] = put(pre.data(), pre.size() * 8);
    }
    dev.n_all = (uint32_t)all.size();
    dev.carry_blocks = cprefix.back();
    dev.off_all = put(all.data(), all.size() * sizeof(SegArgs));
    dev.off_cprefix = put(cprefix.data(), cprefix.size() * 8);
    if (dev.arena_cap < host.size()) {
        dfree(dev.arena);
        dev.arena = nullptr;
        dev.arena_cap = 0;
        const size_t cap = host.size() * 2;
        SPTOL_TRY(dmalloc((void **)&dev.arena, cap, "Batch"));
        dev.arena_cap = cap;
    }
    SPTOL_TRY(dev.stage.upload(dev.arena, host.data(), host.size(), s));
    return 0;
}

template <int LPN, int NM, typename FT>
static void seg_batch_class(const SegBatchDev &dev, int c, int blocks_per_cu, hipStream_t s)
{
    const void *fn = reinterpret_cast<const void *>(&k_batch_seg<LPN, NM, FT>);
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, kBatchWPB * 64, 0) != hipSuccess || occ < 1) occ = 1;
    const int bpc = blocks_per_cu < occ ? blocks_per_cu : occ;
    uint64_t blocks = (uint64_t)num_cus() * (uint64_t)bpc;
    const uint64_t need = (dev.total[c] + kBatchWPB - 1) / kBatchWPB;
    if (blocks > need) blocks = need;
    hipLaunchKernelGGL((k_batch_seg<LPN, NM, FT>), dim3((unsigned)blocks), dim3(kBatchWPB * 64), 0, s,
                       (const SegArgs *)(dev.arena + dev.off_items[c]), (const uint64_t *)(dev.arena + dev.off_prefix[c]),
                       dev.n[c], dev.total[c], dev.counters + c);
}

template <typename FT>
static void seg_batch_kernels_ft(const SegBatchDev &dev, const int *bpc_of, hipStream_t s)
{
    for (int c = 0; c < 8; ++c) {
        if (!dev.n[c]) continue;
        const int bpc = bpc_of[c];
        // class = lane class * 2 + (4 modes)
        constexpr int L1 = sizeof(FT) == 2 ? 1 : 2;
        switch (c) {
        case 0: seg_batch_class<L1, 3, FT>(dev, c, bpc, s); break;
        case 1: seg_batch_class<L1, 4, FT>(dev, c, bpc, s); break;
        case 2: seg_batch_class<L1 * 2, 3, FT>(dev, c, bpc, s); break;
        case 3: seg_batch_class<L1 * 2, 4, FT>(dev, c, bpc, s); break;
        case 4: seg_batch_class<L1 * 4, 3, FT>(dev, c, bpc, s); break;
        case 5: seg_batch_class<L1 * 4, 4, FT>(dev, c, bpc, s); break;
        case 6: seg_batch_class<L1 * 8, 3, FT>(dev, c, bpc, s); break;
        default: seg_batch_class<L1 * 8, 4, FT>(dev, c, bpc, s); break;
        }
    }
}

// Class kernels (blocks_per_cu[c] for class c) then the carry kernel (which
// resets the class counters).
static int seg_batch_kernels(const SegBatchDev &dev, int dtype, const int *blocks_per_cu, hipStream_t s)
{
    if (!dev.n_all) return 0;
    int tok = prof_start("batch_seg", s);
    if (dtype == SPTOL_BF16) seg_batch_kernels_ft<uint16_t>(dev, blocks_per_cu, s);
    else seg_batch_kernels_ft<float>(dev, blocks_per_cu, s);
    prof_stop(tok, s);
    SPTOL_HIP(hipGetLastError(), "Batch launch");
    tok = prof_start("batch_seg_carry", s);
    hipLaunchKernelGGL(k_batch_seg_carry, dim3((unsigned)dev.carry_blocks), dim3(256), 0, s,
                       (const SegArgs *)(dev.arena + dev.off_all), (const uint64_t *)(dev.arena + dev.off_cprefix),
                       dev.n_all, dev.counters);
    prof_stop(tok, s);
    SPTOL_HIP(hipGetLastError(), "Batch carry launch");
    return 0;
}

// Launch configuration of a batch: tile_scale multiplies the base tile
// length and blocks_per_cu caps the persistent grid.  The pair applies to
// the atomic path and, unless per_class is set, to every segmented class;
// the autotuner may then refine each class (cls_*), since the classes are
// separate launches.
struct TuneCfg {
    double tile_scale = 1.0;
    int blocks_per_cu = 8;
    bool per_class = false;
    double cls_scale[8] = {};
    int cls_bpc[8] = {};
    void seg_scales(double *sc) const
    {
        for (int c = 0; c < 8; ++c) sc[c] = per_class ? cls_scale[c] : tile_scale;
    }
    void seg_bpcs(int *b) const
    {
        for (int c = 0; c < 8; ++c) b[c] = per_class ? cls_bpc[c] : blocks_per_cu;
    }
    void make_per_class()
    {
        if (per_class) return;
        for (int c = 0; c < 8; ++c) { cls_scale[c] = tile_scale; cls_bpc[c] = blocks_per_cu; }
        per_class = true;
    }
};

static std::map<std::string, TuneCfg> g_tune_cache;

static int cu_count()
{
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) return 256;
    return n;
}

}  // namespace sptol

using namespace sptol;

struct sptolBatch_s {
    std::vector<BatchDesc> descs;
    int dtype = -1;
    TuneCfg cfg;
    bool planned = false;
    // atomic path (tensors without a usable mode-sorted view)
    uint32_t nlegacy = 0;
    BatchDesc *d_desc = nullptr;
    uint64_t *d_tile_prefix = nullptr, *d_fill_prefix = nullptr;
    uint32_t *d_counter = nullptr;
    uint64_t ntiles = 0, nchunks = 0;
    // segmented path
    SegBatchDev seg;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t graph_stream = nullptr;
    ~sptolBatch_s()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        dfree(d_desc); dfree(d_tile_prefix); dfree(d_fill_prefix); dfree(d_counter);
    }
};

namespace sptol {

// Base tile: about 2^20 * 4 / stride rank-columns of work, i.e. fewer
// nonzeros per tile for wide rows, scaled by the tuned factor.
static uint64_t tile_len(uint32_t stride, double scale)
{
    double base = 8192.0 * 16.0 / (double)stride * scale;
    uint64_t t = (uint64_t)base;
    t = (t + 63) / 64 * 64;
    return t < 64 ? 64 : t;
}

static void split_descs(const std::vector<BatchDesc> &descs, std::vector<const BatchDesc *> &seg,
                        std::vector<BatchDesc> &legacy)
{
    seg.clear();
    legacy.clear();
    for (const BatchDesc &d : descs) {
        if (d.seg) seg.push_back(&d);
        else legacy.push_back(d);
    }
}

static int build_plan(sptolBatch b)
{
    std::vector<const BatchDesc *> segs;
    std::vector<BatchDesc> leg;
    split_descs(b->descs, segs, leg);
    if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
    if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
    double sc[8];
    b->cfg.seg_scales(sc);
    SPTOL_TRY(seg_batch_prepare(segs, b->dtype, sc, b->seg, nullptr));
    const size_t n = leg.size();
    b->nlegacy = (uint32_t)n;
    dfree(b->d_desc); dfree(b->d_tile_prefix); dfree(b->d_fill_prefix);
    b->d_desc = nullptr; b->d_tile_prefix = nullptr; b->d_fill_prefix = nullptr;
    b->ntiles = b->nchunks = 0;
    if (n) {
        std::vector<uint64_t> tp(n + 1, 0), fp(n + 1, 0);
        for (size_t i = 0; i < n; ++i) {
            BatchDesc &d = leg[i];
            d.tile_nnz = tile_len(d.stride, b->cfg.tile_scale);
            tp[i + 1] = tp[i] + (d.nnz + d.tile_nnz - 1) / d.tile_nnz;
            fp[i + 1] = fp[i] + (d.out_len + 65535) / 65536;
        }
        if (tp[n] >= 0xFFFFFFFFull) SPTOL_CHECK(SPTERR_VALUE_ERROR, "Batch Plan", "too many tiles");
        b->ntiles = tp[n];
        b->nchunks = fp[n];
        SPTOL_TRY(sptolHipDuplicateMemory((void **)&b->d_desc, leg.data(), n * sizeof(BatchDesc)));
        SPTOL_TRY(sptolHipDuplicateMemory((void **)&b->d_tile_prefix, tp.data(), (n + 1) * 8));
        SPTOL_TRY(sptolHipDuplicateMemory((void **)&b->d_fill_prefix, fp.data(), (n + 1) * 8));
        if (!b->d_counter) SPTOL_TRY(dmalloc((void **)&b->d_counter, 4, "Batch Plan"));
    }
    SPTOL_HIP(hipStreamSynchronize(nullptr), "Batch Plan");
    b->planned = true;
    return 0;
}

static int launch_legacy(const BatchDesc *d_desc, const uint64_t *d_tile_prefix, const uint64_t *d_fill_prefix,
                         uint32_t *d_counter, uint32_t n, uint64_t ntiles, uint64_t nchunks, int dtype,
                         int blocks_per_cu, hipStream_t s)
{
    if (!n) return 0;
    const int cus = cu_count();
    unsigned fill_grid = (unsigned)std::min<uint64_t>(std::max<uint64_t>(nchunks, 1), (uint64_t)cus * 8);
    int tok = prof_start("batch_fill", s);
    hipLaunchKernelGGL(k_batch_fill, dim3(fill_grid), dim3(256), 0, s, d_desc, d_fill_prefix, n, nchunks, d_counter);
    prof_stop(tok, s);
    const unsigned grid = (unsigned)(cus * blocks_per_cu);
    tok = prof_start("batch_mttkrp", s);
    if (dtype == SPTOL_BF16)
        hipLaunchKernelGGL(k_batch_mttkrp<uint16_t>, dim3(grid), dim3(256), 0, s, d_desc, d_tile_prefix, n, ntiles,
                           d_counter);
    else
        hipLaunchKernelGGL(k_batch_mttkrp<float>, dim3(grid), dim3(256), 0, s, d_desc, d_tile_prefix, n, ntiles,
                           d_counter);
    prof_stop(tok, s);
    SPTOL_HIP(hipGetLastError(), "Batch launch");
    return 0;
}

static int launch_batch(sptolBatch b, hipStream_t s)
{
    SPTOL_TRY(launch_legacy(b->d_desc, b->d_tile_prefix, b->d_fill_prefix, b->d_counter, b->nlegacy, b->ntiles,
                            b->nchunks, b->dtype, b->cfg.blocks_per_cu, s));
    int bpc[8];
    b->cfg.seg_bpcs(bpc);
    return seg_batch_kernels(b->seg, b->dtype, bpc, s);
}

void make_desc(sptolDevCoo X, const sptolDevMatrix *mats, const uint32_t *mats_order, uint32_t mode,
               const sptolDevMatrix *out, BatchDesc *d)
{
    std::memset(d, 0, sizeof(*d));
    d->nnz = X->nnz;
    const SortedCopy &sc = X->sorted[mode];
    for (uint32_t m = 0; m < X->nmodes; ++m) {
        d->inds[m] = sc.vals ? sc.inds[m] : X->inds[m];
        d->mats[m] = mats[m].values;
        d->order[m] = mats_order[m];
    }
    d->vals = sc.vals ? sc.vals : X->vals;
    d->seg = (sc.ready && sc.seg.ready && seg_ok(X->nmodes, out->stride)) ? const_cast<SegView *>(&sc.seg) : nullptr;
    d->out = (float *)out->values;
    d->out_len = (uint64_t)X->ndims[mode] * out->stride;
    d->nmodes = X->nmodes;
    d->mode = mode;
    d->R = out->ncols;
    d->stride = out->stride;
    d->lpn = out->stride / 4;
}

BatchBuffers::~BatchBuffers()
{
    dfree(desc); dfree(tile_prefix); dfree(fill_prefix); dfree(counter);
}

int batch_launch(std::vector<BatchDesc> &descs, int dtype, double tile_scale, int blocks_per_cu, BatchBuffers &bb,
                 hipStream_t s)
{
    if (descs.empty()) return 0;
    std::vector<const BatchDesc *> segs;
    std::vector<BatchDesc> leg;
    split_descs(descs, segs, leg);
    if (!segs.empty()) {
        double sc[8];
        int bpc[8];
        for (int c = 0; c < 8; ++c) { sc[c] = tile_scale; bpc[c] = blocks_per_cu; }
        SPTOL_TRY(seg_batch_prepare(segs, dtype, sc, bb.seg, s));
        SPTOL_TRY(seg_batch_kernels(bb.seg, dtype, bpc, s));
    }
    const size_t n = leg.size();
    if (n == 0) return 0;
    std::vector<uint64_t> tp(n + 1, 0), fp(n + 1, 0);
    for (size_t i = 0; i < n; ++i) {
        BatchDesc &d = leg[i];
        d.tile_nnz = tile_len(d.stride, tile_scale);
        tp[i + 1] = tp[i] + (d.nnz + d.tile_nnz - 1) / d.tile_nnz;
        fp[i + 1] = fp[i] + (d.out_len + 65535) / 65536;
    }
    if (n > bb.cap) {
        dfree(bb.desc); dfree(bb.tile_prefix); dfree(bb.fill_prefix);
        bb.desc = nullptr; bb.tile_prefix = nullptr; bb.fill_prefix = nullptr; bb.cap = 0;
        const size_t cap = n * 2;
        SPTOL_TRY(dmalloc((void **)&bb.desc, cap * sizeof(BatchDesc), "Batch"));
        SPTOL_TRY(dmalloc((void **)&bb.tile_prefix, (cap + 1) * 8, "Batch"));
        SPTOL_TRY(dmalloc((void **)&bb.fill_prefix, (cap + 1) * 8, "Batch"));
        bb.cap = cap;
    }
    if (!bb.counter) SPTOL_TRY(dmalloc((void **)&bb.counter, 4, "Batch"));
    // pageable sources: each copy is staged before the call returns, so the
    // host vectors may be reused immediately; the copies stay stream-ordered
    SPTOL_HIP(hipMemcpyAsync(bb.desc, leg.data(), n * sizeof(BatchDesc), hipMemcpyHostToDevice, s), "Batch");
    SPTOL_HIP(hipMemcpyAsync(bb.tile_prefix, tp.data(), (n + 1) * 8, hipMemcpyHostToDevice, s), "Batch");
    SPTOL_HIP(hipMemcpyAsync(bb.fill_prefix, fp.data(), (n + 1) * 8, hipMemcpyHostToDevice, s), "Batch");
    return launch_legacy(bb.desc, bb.tile_prefix, bb.fill_prefix, bb.counter, (uint32_t)n, tp[n], fp[n], dtype,
                         blocks_per_cu, s);
}

static std::string signature(sptolBatch b)
{
    // rank mix and log2 total nonzeros per rank class
    std::map<uint32_t, uint64_t> by;
    for (auto &d : b->descs) by[d.stride] += d.nnz;
    std::string s = std::to_string(b->dtype);
    for (auto &kv : by) s += "/" + std::to_string(kv.first) + ":" + std::to_string((int)std::log2((double)kv.second + 1));
    return s;
}

}  // namespace sptol

extern "C" {

int sptolBatchCreate(sptolBatch *b)
{
    *b = new sptolBatch_s();
    return 0;
}

int sptolBatchAdd(sptolBatch b, sptolDevCoo X, const sptolDevMatrix *mats, const sptIndex mats_order[], sptIndex mode,
                  sptolDevMatrix *out)
{
    if (X->nmodes != 3 && X->nmodes != 4) SPTOL_CHECK(SPTERR_VALUE_ERROR, "Batch Add", "batched MTTKRP takes 3-D or 4-D tensors");
    if (mode >= X->nmodes) SPTOL_CHECK(SPTERR_VALUE_ERROR, "Batch Add", "mode out of range");
    const uint32_t S = out->stride;
    if (S > 256 || out->dtype != SPTOL_F32) SPTOL_CHECK(SPTERR_VALUE_ERROR, "Batch Add", "stride must be <= 256, output FP32");
    for (uint32_t m = 0; m < X->nmodes; ++m) {
        if (mats[m].nrows != X->ndims[m] || mats[m].stride != S || mats[m].ncols != out->ncols)
            SPTOL_CHECK(SPTERR_SHAPE_MISMATCH, "Batch Add", "factor shape");
        if (b->dtype >= 0 && mats[m].dtype != b->dtype) SPTOL_CHECK(SPTERR_VALUE_ERROR, "Batch Add", "mixed factor dtypes");
    }
    if (out->nrows < X->ndims[mode]) SPTOL_CHECK(SPTERR_SHAPE_MISMATCH, "Batch Add", "output rows");
    uint32_t seen = 0;
    for (uint32_t k = 1; k < X->nmodes; ++k) {
        if (mats_order[k] >= X->nmodes || mats_order[k] == mode || ((seen >> mats_order[k]) & 1u))
            SPTOL_CHECK(SPTERR_VALUE_ERROR, "Batch Add", "mats_order[1..N-1] must list the other modes");
        seen |= 1u << mats_order[k];
    }
    b->dtype = mats[0].dtype;
    BatchDesc d;
    make_desc(X, mats, mats_order, mode, out, &d);
    b->descs.push_back(d);
    b->planned = false;
    return 0;
}

int sptolBatchSetConfig(sptolBatch b, double tile_scale, int blocks_per_cu)
{
    if (!(tile_scale > 0) || blocks_per_cu < 1 || blocks_per_cu > 32)
        SPTOL_CHECK(SPTERR_VALUE_ERROR, "Batch Config", "tile_scale > 0, blocks_per_cu in [1, 32]");
    b->cfg.tile_scale = tile_scale;
    b->cfg.blocks_per_cu = blocks_per_cu;
    b->cfg.per_class = false;
    return build_plan(b);
}

int sptolBatchGetConfig(sptolBatch b, double *tile_scale, int *blocks_per_cu)
{
    *tile_scale = b->cfg.tile_scale;
    *blocks_per_cu = b->cfg.blocks_per_cu;
    return 0;
}

int sptolBatchRun(sptolBatch b, void *stream)
{
    if (b->descs.empty()) return 0;
    if (!b->planned) SPTOL_TRY(build_plan(b));
    hipStream_t s = stream_of(stream);
    if (b->exec && b->graph_stream == s) {
        SPTOL_HIP(hipGraphLaunch(b->exec, s), "Batch Graph Launch");
        return 0;
    }
    return launch_batch(b, s);
}

// Capture the fill + persistent launch into a hipGraph on `stream`; later
// sptolBatchRun calls on the same stream replay it.
int sptolBatchCapture(sptolBatch b, void *stream)
{
    if (b->descs.empty()) return 0;
    if (!b->planned) SPTOL_TRY(build_plan(b));
    hipStream_t s = stream_of(stream);
    if (s == nullptr) SPTOL_CHECK(SPTERR_VALUE_ERROR, "Batch Capture", "capture needs a created (non-default) stream");
    if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
    if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
    SPTOL_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "Batch Capture");
    int r = launch_batch(b, s);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (r) { if (g) (void)hipGraphDestroy(g); return r; }
    SPTOL_HIP(e, "Batch Capture");
    b->graph = g;
    SPTOL_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0), "Batch Capture");
    b->graph_stream = s;
    return 0;
}

// Autotuner: time every (tile scale, blocks per CU) candidate on the batch
// (median of `reps` runs after one warm-up) and keep the fastest; then, for
// each segmented class in turn, every candidate for that class alone with
// the others held, keeping any that shortens the whole launch.  The result
// is cached under the batch's signature (factor dtype, rank classes, log2
// nonzeros per class); a cached signature is reused without timing.
int sptolBatchAutotune(sptolBatch b, int reps, void *stream)
{
    if (b->descs.empty()) return 0;
    const std::string key = signature(b);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
        b->cfg = it->second;
        return build_plan(b);
    }
    hipStream_t s = stream_of(stream);
    const double scales[] = {0.25, 0.5, 1.0, 2.0, 4.0};
    const int occs[] = {2, 4, 8, 16};
    hipEvent_t e0, e1;
    SPTOL_HIP(hipEventCreate(&e0), "Autotune");
    SPTOL_HIP(hipEventCreate(&e1), "Autotune");
    if (reps < 1) reps = 1;
    // median time of one batch launch under b->cfg (plan rebuilt for it)
    auto time_cfg = [&](double &ms_out) {
        int r = build_plan(b);
        if (!r) r = launch_batch(b, s);   // warm-up
        std::vector<float> t;
        for (int k = 0; k < reps && !r; ++k) {
            (void)hipEventRecord(e0, s);
            r = launch_batch(b, s);
            (void)hipEventRecord(e1, s);
            (void)hipEventSynchronize(e1);
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            t.push_back(ms);
        }
        if (r) return r;
        std::sort(t.begin(), t.end());
        ms_out = t[t.size() / 2];
        return 0;
    };
    int r = 0;
    double best = 1e300;
    TuneCfg bestc = b->cfg;
    // one pair for every class (and the atomic path)
    for (double sc : scales)
        for (int oc : occs) {
            if (r) break;
            b->cfg = TuneCfg();
            b->cfg.tile_scale = sc;
            b->cfg.blocks_per_cu = oc;
            double ms = 0;
            r = time_cfg(ms);
            if (!r && ms < best) { best = ms; bestc = b->cfg; }
        }
    // then each segmented class on its own, the others held at their best
    for (int c = 0; c < 8 && !r; ++c) {
        if (!b->seg.n[c]) continue;
        b->cfg = bestc;
        double ms = 0;
        r = time_cfg(ms);   // the current best, timed again beside the candidates
        if (r) break;
        best = ms;
        for (double sc : scales)
            for (int oc : occs) {
                if (r) break;
                b->cfg = bestc;
                b->cfg.make_per_class();
                b->cfg.cls_scale[c] = sc;
                b->cfg.cls_bpc[c] = oc;
                r = time_cfg(ms);
                if (!r && ms < best) { best = ms; bestc = b->cfg; }
            }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (r) return r;
    b->cfg = bestc;
    g_tune_cache[key] = bestc;
    return build_plan(b);
}

int sptolBatchSize(sptolBatch b) { return (int)b->descs.size(); }

void sptolBatchFree(sptolBatch b) { delete b; }

}  // extern "C"
