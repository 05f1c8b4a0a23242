// HiCOO MTTKRP with output tiles staged in LDS.
//
// In the HiCOO order (superblocks row-major, Z-Morton inside) consecutive
// nonzeros of a hypersparse tensor land on unrelated output rows, so a
// nonzero-parallel kernel pays one global float atomic per nonzero and
// column.  The reference's remedy is its superblock-row schedule
// (hicoo/mttkrp_omp.c:555-820: a thread owns a band of output rows).  The
// GPU form of that idea, per output mode n:
//   * a "mode-n view" of the tensor: the HiCOO blocks stably reordered by
//     their mode-n block coordinate binds[n] (each block keeps its entries,
//     in their Morton order), built once on the device;
//   * work items: runs of consecutive nonzeros whose blocks share binds[n],
//     cut into pieces of at most kItemNnz nonzeros;
//   * one workgroup per item accumulates its 2^sb x R output tile in LDS with
//     LDS float atomics, then adds the rows it touched to the output with one
//     global atomic per element (items of one hot tile meet only there).
// Within an item each lane group walks a contiguous slice of nonzeros with
// the same gather pipeline as the plain HiCOO kernel.
#include "dev_common.hpp"
#include <atomic>
#include <hipcub/hipcub.hpp>
#include <vector>
#include <cstring>

namespace sptol {

constexpr uint32_t kItemNnz = 2048;

struct HicooItem {
    uint64_t z0, z1;     // nonzero range
    uint64_t b0;         // block containing z0
    uint32_t tile;       // mode-n block coordinate
    uint32_t pad;
};

// One mode-n view: blocks sorted by binds[n], entries moved with them.
struct HicooView {
    SegView seg;         // segmented kind (when use_seg)
    bool use_seg = false;
    uint64_t *bptr = nullptr;
    uint32_t *binds[kMaxModes] = {};
    uint8_t *einds[kMaxModes] = {};
    float *vals = nullptr;
    HicooItem *items = nullptr;
    uint64_t nitems = 0;
    bool ready = false;
    void release()
    {
        dfree(bptr); bptr = nullptr;
        for (auto &p : binds) { dfree(p); p = nullptr; }
        for (auto &p : einds) { dfree(p); p = nullptr; }
        dfree(vals); vals = nullptr;
        dfree(items); items = nullptr;
        nitems = 0;
        seg.release();
        use_seg = false;
        ready = false;
    }
};

}  // namespace sptol

// The views live beside the HiCOO handle (declared opaque in dev_common).
struct sptolHicooViews {
    sptol::HicooView v[sptol::kMaxModes];
};

namespace sptol {

__global__ void hv_iota(uint32_t *p, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = (uint32_t)i;
}

__global__ void hv_block_sizes(const uint64_t *__restrict__ bptr, const uint32_t *__restrict__ perm, uint64_t nb,
                               uint64_t *__restrict__ sizes)
{
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nb; j += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t b = perm[j];
        sizes[j] = bptr[b + 1] - bptr[b];
    }
}

// One thread per new block: move its block coordinates and entries.
__global__ void hv_move(const uint64_t *__restrict__ bptr, const uint32_t *__restrict__ perm, uint64_t nb,
                        const uint64_t *__restrict__ nbptr, uint32_t nmodes, const MttkrpArgs src, uint32_t **dbinds,
                        uint8_t **deinds, float *__restrict__ dvals)
{
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nb; j += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t b = perm[j];
        const uint64_t s0 = bptr[b], s1 = bptr[b + 1], d0 = nbptr[j];
        for (uint32_t m = 0; m < nmodes; ++m) {
            dbinds[m][j] = src.binds[m][b];
            for (uint64_t z = s0; z < s1; ++z) deinds[m][d0 + (z - s0)] = src.einds[m][z];
        }
        for (uint64_t z = s0; z < s1; ++z) dvals[d0 + (z - s0)] = src.vals[z];
    }
}

// Tile boundaries: flag[j] = 1 where the block's mode-n coordinate changes.
__global__ void hv_tile_flags(const uint32_t *__restrict__ key, uint64_t nb, uint8_t *__restrict__ flag)
{
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nb; j += (uint64_t)gridDim.x * blockDim.x)
        flag[j] = (j == 0 || key[j] != key[j - 1]) ? 1 : 0;
}

__global__ void hv_tile_info(const uint32_t *__restrict__ start_block, uint64_t ntiles, uint64_t nb,
                             const uint64_t *__restrict__ nbptr, const uint32_t *__restrict__ key,
                             uint64_t *__restrict__ tz0, uint64_t *__restrict__ tz1, uint32_t *__restrict__ tcoord)
{
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < ntiles; t += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t b0 = start_block[t];
        const uint64_t b1 = t + 1 < ntiles ? start_block[t + 1] : nb;
        tz0[t] = nbptr[b0];
        tz1[t] = nbptr[b1];
        tcoord[t] = key[b0];
    }
}

// items[i].b0 = the block containing items[i].z0 (search over all blocks).
__global__ void hv_item_blocks(HicooItem *__restrict__ items, uint64_t nitems, const uint64_t *__restrict__ bptr,
                               uint64_t nb)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nitems; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t z = items[i].z0;
        uint64_t lo = 0, hi = nb;   // bptr[lo] <= z < bptr[hi]
        while (hi - lo > 1) {
            const uint64_t mid = (lo + hi) >> 1;
            if (bptr[mid] <= z) lo = mid; else hi = mid;
        }
        items[i].b0 = lo;
    }
}

static unsigned hv_grid(uint64_t n)
{
    uint64_t g = (n + 255) / 256;
    if (g > 65536) g = 65536;
    return (unsigned)(g ? g : 1);
}

__device__ __forceinline__ uint64_t hv_find_block(const uint64_t *__restrict__ bptr, uint64_t lo, uint64_t hi, uint64_t z)
{
    // invariant bptr[lo] <= z < bptr[hi]
    while (hi - lo > 1) {
        const uint64_t mid = (lo + hi) >> 1;
        if (bptr[mid] <= z) lo = mid; else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ float4 hv_ld4(const float *p) { return *reinterpret_cast<const float4 *>(p); }
__device__ __forceinline__ float4 hv_ld4(const uint16_t *p)
{
    const uint2 u = *reinterpret_cast<const uint2 *>(p);
    return make_float4(__uint_as_float(u.x << 16), __uint_as_float(u.x & 0xFFFF0000u),
                       __uint_as_float(u.y << 16), __uint_as_float(u.y & 0xFFFF0000u));
}

// One workgroup per item; tile in dynamic LDS: (1 << sb) rows x S floats.
template <int LPN, int NM, int U, typename FT>
__global__ __launch_bounds__(256) void k_mttkrp_hicoo_tiled(const MttkrpArgs a, const HicooItem *__restrict__ items,
                                                            uint64_t nb)
{
    extern __shared__ float tile[];
    __shared__ uint32_t touched[8];
    const HicooItem it = items[blockIdx.x];
    const uint32_t S = a.stride, R = a.R, sb = a.sb_bits;
    const uint32_t rows = 1u << sb;
    for (uint32_t i = threadIdx.x; i < rows * S; i += 256) tile[i] = 0.0f;
    if (threadIdx.x < 8) touched[threadIdx.x] = 0u;
    __syncthreads();

    constexpr uint32_t groups = 256 / LPN;
    const uint32_t g = threadIdx.x / LPN, l = threadIdx.x % LPN;
    const uint64_t n = it.z1 - it.z0;
    const uint64_t slice = (n + groups - 1) / groups;
    const uint64_t z0 = it.z0 + (uint64_t)g * slice;
    const uint64_t z1 = min<uint64_t>(z0 + slice, it.z1);
    const uint32_t nm = NM > 0 ? NM : a.nmodes;
    const uint32_t mode = a.mode;
    const uint64_t *__restrict__ bptr = a.bptr;
    if (z0 < z1) {
        // b0 holds the item's first nonzero; every block holds at least one
        // nonzero, so the item ends before block b0 + kItemNnz + 1
        uint64_t b = hv_find_block(bptr, it.b0, min<uint64_t>(it.b0 + kItemNnz + 1, nb), z0);
        for (uint32_t c0 = l * 4; c0 < S && c0 < R; c0 += 4 * LPN) {
            uint64_t bb_start = b;
            for (uint64_t z = z0; z < z1; z += U) {
                uint64_t bnext[U];
#pragma unroll
                for (int t = 0; t < U; ++t) bnext[t] = bptr[min<uint64_t>(bb_start + 1 + t, nb)];
                uint64_t blk[U];
                bool valid[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    uint64_t c = 0;
#pragma unroll
                    for (int t = 0; t < U; ++t) c += (bnext[t] <= z + u) ? 1u : 0u;
                    blk[u] = bb_start + c;
                    valid[u] = (z + u) < z1;
                }
                uint32_t coord[U][NM > 0 ? NM : kMaxModes];
                uint32_t erow[U];
                float v[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    // padding lanes repeat the chunk's first nonzero z (block blk[0]):
                    // a block coordinate and an element offset of the same nonzero
                    const uint64_t zz = valid[u] ? z + u : z;
                    const uint64_t bq = valid[u] ? blk[u] : blk[0];
                    v[u] = a.vals[zz];
                    erow[u] = a.einds[mode][zz];
#pragma unroll
                    for (int m = 0; m < (NM > 0 ? NM : kMaxModes); ++m) {
                        if (NM == 0 && m >= (int)nm) break;
                        coord[u][m] = (a.binds[m][bq] << sb) + (uint32_t)a.einds[m][zz];
                    }
                }
                const uint32_t o1 = a.order[1];
                const FT *__restrict__ A1 = (const FT *)a.mats[o1];
                float4 p[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    p[u] = hv_ld4(A1 + (size_t)coord[u][o1] * S + c0);
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
                    for (int u = 0; u < U; ++u) q[u] = hv_ld4(Ak + (size_t)coord[u][ok] * S + c0);
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        p[u].x *= q[u].x; p[u].y *= q[u].y; p[u].z *= q[u].z; p[u].w *= q[u].w;
                    }
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    if (!valid[u]) continue;
                    float *tr = tile + erow[u] * S + c0;
                    atomicAdd(tr + 0, p[u].x);
                    if (c0 + 1 < R) atomicAdd(tr + 1, p[u].y);
                    if (c0 + 2 < R) atomicAdd(tr + 2, p[u].z);
                    if (c0 + 3 < R) atomicAdd(tr + 3, p[u].w);
                    if (l == 0) atomicOr(&touched[erow[u] >> 5], 1u << (erow[u] & 31));
                }
                bb_start = blk[U - 1];
            }
        }
    }
    __syncthreads();
    const uint64_t base_row = (uint64_t)it.tile << sb;
    for (uint32_t i = threadIdx.x; i < rows * S; i += 256) {
        const uint32_t r = i / S, c = i % S;
        if (c < R && ((touched[r >> 5] >> (r & 31)) & 1u))
            unsafeAtomicAdd(a.out + (base_row + r) * S + c, tile[i]);
    }
}

#define HV_NM(LPN, U, FT, nm, grid, shm, s, a, items, nb)                                                   \
    do {                                                                                                     \
        if ((nm) == 3) hipLaunchKernelGGL((k_mttkrp_hicoo_tiled<LPN, 3, U, FT>), grid, dim3(256), shm, s, a, items, nb); \
        else if ((nm) == 4) hipLaunchKernelGGL((k_mttkrp_hicoo_tiled<LPN, 4, U, FT>), grid, dim3(256), shm, s, a, items, nb); \
        else hipLaunchKernelGGL((k_mttkrp_hicoo_tiled<LPN, 0, U, FT>), grid, dim3(256), shm, s, a, items, nb); \
    } while (0)

#define HV_LPN(FT, lpn, nm, grid, shm, s, a, items, nb)                                 \
    do {                                                                                \
        switch (lpn) {                                                                  \
        case 1: case 2: HV_NM(2, 4, FT, nm, grid, shm, s, a, items, nb); break;         \
        case 4: HV_NM(4, 4, FT, nm, grid, shm, s, a, items, nb); break;                 \
        case 8: HV_NM(8, 4, FT, nm, grid, shm, s, a, items, nb); break;                 \
        default: HV_NM(16, 4, FT, nm, grid, shm, s, a, items, nb); break;               \
        }                                                                               \
    } while (0)

// Whether the tiled kernel applies: tile fits in 64 KB of LDS, <= 256 rows.
bool hicoo_tiled_ok(uint32_t sb_bits, uint32_t stride)
{
    return sb_bits <= 8 && stride <= 64 && ((size_t)(1u << sb_bits) * stride * 4) <= 65536;
}

int build_hicoo_view(sptolDevHicoo X, uint32_t mode, HicooView &v, hipStream_t s)
{
    const uint64_t nb = X->nb, nnz = X->nnz;
    if (nb > kMaxSortItems)
        SPTOL_CHECK(SPTERR_VALUE_ERROR, "HiCOO view", "nb > 2^31 - 1 (the device sort counts items in 32-bit ints)");
    uint32_t *perm = nullptr, *perm2 = nullptr, *key2 = nullptr, *start_block = nullptr, *tcoord = nullptr;
    uint64_t *sizes = nullptr, *tz0 = nullptr, *tz1 = nullptr;
    uint8_t *flags = nullptr;
    void *tmp = nullptr;
    uint64_t *d_nsel = nullptr;
    uint32_t **d_binds = nullptr;
    uint8_t **d_einds = nullptr;
    auto cleanup = [&]() {
        dfree(perm); dfree(perm2); dfree(key2); dfree(start_block); dfree(tcoord);
        dfree(sizes); dfree(tz0); dfree(tz1); dfree(flags); dfree(tmp); dfree(d_nsel);
        dfree(d_binds); dfree(d_einds);
    };
    int r = 0;
#define HV_TRY(x) do { r = (x); if (r) { cleanup(); v.release(); return r; } } while (0)
#define HV_HIP(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { cleanup(); v.release(); SPTOL_HIP(e_, "HiCOO view"); } } while (0)
    HV_TRY(dmalloc((void **)&perm, nb * 4, "HiCOO view"));
    HV_TRY(dmalloc((void **)&perm2, nb * 4, "HiCOO view"));
    HV_TRY(dmalloc((void **)&key2, nb * 4, "HiCOO view"));
    hipLaunchKernelGGL(hv_iota, dim3(hv_grid(nb)), dim3(256), 0, s, perm, nb);
    int endbit = 1;
    const uint64_t maxkey = ((uint64_t)X->ndims[mode] >> X->sb_bits) + 1;
    while (endbit < 32 && (1ull << endbit) < maxkey) ++endbit;
    size_t tb = 0;
    HV_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, X->binds[mode], key2, perm, perm2, (int)nb, 0, endbit, s));
    HV_TRY(dmalloc(&tmp, tb ? tb : 4, "HiCOO view"));
    HV_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tb, X->binds[mode], key2, perm, perm2, (int)nb, 0, endbit, s));
    dfree(tmp); tmp = nullptr;
    // new block pointers
    HV_TRY(dmalloc((void **)&sizes, (nb + 1) * 8, "HiCOO view"));
    HV_TRY(dmalloc((void **)&v.bptr, (nb + 1) * 8, "HiCOO view"));
    hipLaunchKernelGGL(hv_block_sizes, dim3(hv_grid(nb)), dim3(256), 0, s, X->bptr, perm2, nb, sizes);
    HV_HIP(hipMemsetAsync(sizes + nb, 0, 8, s));
    tb = 0;
    HV_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, sizes, v.bptr, (int)(nb + 1), s));
    HV_TRY(dmalloc(&tmp, tb ? tb : 4, "HiCOO view"));
    HV_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tb, sizes, v.bptr, (int)(nb + 1), s));
    dfree(tmp); tmp = nullptr;
    // move blocks and entries
    for (uint32_t m = 0; m < X->nmodes; ++m) {
        HV_TRY(dmalloc((void **)&v.binds[m], nb * 4, "HiCOO view"));
        HV_TRY(dmalloc((void **)&v.einds[m], nnz, "HiCOO view"));
    }
    HV_TRY(dmalloc((void **)&v.vals, nnz * 4, "HiCOO view"));
    HV_TRY(dmalloc((void **)&d_binds, kMaxModes * sizeof(void *), "HiCOO view"));
    HV_TRY(dmalloc((void **)&d_einds, kMaxModes * sizeof(void *), "HiCOO view"));
    HV_HIP(hipMemcpyAsync(d_binds, v.binds, kMaxModes * sizeof(void *), hipMemcpyHostToDevice, s));
    HV_HIP(hipMemcpyAsync(d_einds, v.einds, kMaxModes * sizeof(void *), hipMemcpyHostToDevice, s));
    MttkrpArgs src;
    std::memset(&src, 0, sizeof(src));
    for (uint32_t m = 0; m < X->nmodes; ++m) { src.binds[m] = X->binds[m]; src.einds[m] = X->einds[m]; }
    src.vals = X->vals;
    hipLaunchKernelGGL(hv_move, dim3(hv_grid(nb)), dim3(256), 0, s, X->bptr, perm2, nb, v.bptr, X->nmodes, src,
                       d_binds, d_einds, v.vals);
    // tiles: runs of equal mode-n block coordinate
    HV_TRY(dmalloc((void **)&flags, nb, "HiCOO view"));
    hipLaunchKernelGGL(hv_tile_flags, dim3(hv_grid(nb)), dim3(256), 0, s, v.binds[mode], nb, flags);
    HV_TRY(dmalloc((void **)&start_block, nb * 4, "HiCOO view"));
    HV_TRY(dmalloc((void **)&d_nsel, 8, "HiCOO view"));
    tb = 0;
    HV_HIP(hipcub::DeviceSelect::Flagged(nullptr, tb, perm /* iota input */, flags, start_block, d_nsel, (int)nb, s));
    HV_TRY(dmalloc(&tmp, tb ? tb : 4, "HiCOO view"));
    // perm still holds 0..nb-1 (SortPairs wrote perm2)
    HV_HIP(hipcub::DeviceSelect::Flagged(tmp, tb, perm, flags, start_block, d_nsel, (int)nb, s));
    uint64_t ntiles = 0;
    HV_HIP(hipMemcpyAsync(&ntiles, d_nsel, 8, hipMemcpyDeviceToHost, s));
    HV_HIP(hipStreamSynchronize(s));
    HV_TRY(dmalloc((void **)&tz0, ntiles * 8, "HiCOO view"));
    HV_TRY(dmalloc((void **)&tz1, ntiles * 8, "HiCOO view"));
    HV_TRY(dmalloc((void **)&tcoord, ntiles * 4, "HiCOO view"));
    hipLaunchKernelGGL(hv_tile_info, dim3(hv_grid(ntiles)), dim3(256), 0, s, start_block, ntiles, nb, v.bptr,
                       v.binds[mode], tz0, tz1, tcoord);
    std::vector<uint64_t> h0(ntiles), h1(ntiles);
    std::vector<uint32_t> hc(ntiles), hsb(ntiles);
    HV_HIP(hipMemcpyAsync(h0.data(), tz0, ntiles * 8, hipMemcpyDeviceToHost, s));
    HV_HIP(hipMemcpyAsync(h1.data(), tz1, ntiles * 8, hipMemcpyDeviceToHost, s));
    HV_HIP(hipMemcpyAsync(hc.data(), tcoord, ntiles * 4, hipMemcpyDeviceToHost, s));
    HV_HIP(hipMemcpyAsync(hsb.data(), start_block, ntiles * 4, hipMemcpyDeviceToHost, s));
    HV_HIP(hipStreamSynchronize(s));
    // items: each tile cut into pieces of at most kItemNnz nonzeros; the block
    // containing each piece's first nonzero is then found on the device.
    std::vector<HicooItem> items;
    items.reserve(ntiles + nnz / kItemNnz + 1);
    for (uint64_t t = 0; t < ntiles; ++t) {
        for (uint64_t z = h0[t]; z < h1[t]; z += kItemNnz) {
            HicooItem itm;
            itm.z0 = z;
            itm.z1 = std::min<uint64_t>(z + kItemNnz, h1[t]);
            itm.b0 = hsb[t];
            itm.tile = hc[t];
            itm.pad = 0;
            items.push_back(itm);
        }
    }
    v.nitems = items.size();
    HV_TRY(dmalloc((void **)&v.items, std::max<size_t>(items.size(), 1) * sizeof(HicooItem), "HiCOO view"));
    HV_HIP(hipMemcpyAsync(v.items, items.data(), items.size() * sizeof(HicooItem), hipMemcpyHostToDevice, s));
    if (!items.empty())
        hipLaunchKernelGGL(hv_item_blocks, dim3(hv_grid(items.size())), dim3(256), 0, s, v.items, (uint64_t)items.size(),
                           v.bptr, nb);
    HV_HIP(hipGetLastError());
    HV_HIP(hipStreamSynchronize(s));
    cleanup();
    v.ready = true;
    return 0;
#undef HV_TRY
#undef HV_HIP
}

}  // namespace sptol

namespace sptol {

// Segmented view: one thread per block rebuilds its entries' coordinates.
__global__ void hv_decompress(const uint64_t *__restrict__ bptr, uint64_t nb, uint32_t nmodes, uint32_t sb,
                              const MttkrpArgs src, uint32_t **coords)
{
    for (uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t z0 = bptr[b], z1 = bptr[b + 1];
        for (uint32_t m = 0; m < nmodes; ++m) {
            const uint32_t base = src.binds[m][b] << sb;
            for (uint64_t z = z0; z < z1; ++z) coords[m][z] = base + (uint32_t)src.einds[m][z];
        }
    }
}

template <typename T>
__global__ void hv_gather(T *__restrict__ dst, const T *__restrict__ src, const uint32_t *__restrict__ perm, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[perm[i]];
}

int build_hicoo_seg(sptolDevHicoo X, uint32_t mode, HicooView &v, hipStream_t s)
{
    const uint64_t nb = X->nb, nnz = X->nnz;
    const uint32_t N = X->nmodes;
    if (nnz > kMaxSortItems)
        SPTOL_CHECK(SPTERR_VALUE_ERROR, "HiCOO view", "nnz > 2^31 - 1 (the device sort counts items in 32-bit ints)");
    uint32_t *coords[kMaxModes] = {}, *sorted[kMaxModes] = {};
    uint32_t *perm = nullptr, *perm2 = nullptr, *key2 = nullptr, **d_coords = nullptr;
    float *svals = nullptr;
    void *tmp = nullptr;
    auto cleanup = [&]() {
        for (uint32_t m = 0; m < kMaxModes; ++m) { dfree(coords[m]); dfree(sorted[m]); }
        dfree(perm); dfree(perm2); dfree(key2); dfree(d_coords); dfree(svals); dfree(tmp);
    };
    int r = 0;
#define HS_TRY(x) do { r = (x); if (r) { cleanup(); v.release(); return r; } } while (0)
#define HS_HIP(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { cleanup(); v.release(); SPTOL_HIP(e_, "HiCOO view"); } } while (0)
    for (uint32_t m = 0; m < N; ++m) HS_TRY(dmalloc((void **)&coords[m], nnz * 4, "HiCOO view"));
    HS_TRY(dmalloc((void **)&d_coords, kMaxModes * sizeof(void *), "HiCOO view"));
    HS_HIP(hipMemcpyAsync(d_coords, coords, kMaxModes * sizeof(void *), hipMemcpyHostToDevice, s));
    MttkrpArgs src;
    std::memset(&src, 0, sizeof(src));
    for (uint32_t m = 0; m < N; ++m) { src.binds[m] = X->binds[m]; src.einds[m] = X->einds[m]; }
    hipLaunchKernelGGL(hv_decompress, dim3(hv_grid(nb)), dim3(256), 0, s, X->bptr, nb, N, X->sb_bits, src, d_coords);
    HS_TRY(dmalloc((void **)&perm, nnz * 4, "HiCOO view"));
    HS_TRY(dmalloc((void **)&perm2, nnz * 4, "HiCOO view"));
    HS_TRY(dmalloc((void **)&key2, nnz * 4, "HiCOO view"));
    hipLaunchKernelGGL(hv_iota, dim3(hv_grid(nnz)), dim3(256), 0, s, perm, nnz);
    int endbit = 1;
    while (endbit < 32 && (1ull << endbit) < X->ndims[mode]) ++endbit;
    size_t tb = 0;
    HS_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, coords[mode], key2, perm, perm2, (int)nnz, 0, endbit, s));
    HS_TRY(dmalloc(&tmp, tb ? tb : 4, "HiCOO view"));
    HS_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tb, coords[mode], key2, perm, perm2, (int)nnz, 0, endbit, s));
    dfree(tmp); tmp = nullptr;
    dfree(key2); key2 = nullptr;
    dfree(perm); perm = nullptr;
    for (uint32_t m = 0; m < N; ++m) {
        HS_TRY(dmalloc((void **)&sorted[m], nnz * 4, "HiCOO view"));
        hipLaunchKernelGGL(hv_gather<uint32_t>, dim3(hv_grid(nnz)), dim3(256), 0, s, sorted[m], coords[m], perm2, nnz);
    }
    HS_TRY(dmalloc((void **)&svals, nnz * 4, "HiCOO view"));
    hipLaunchKernelGGL(hv_gather<float>, dim3(hv_grid(nnz)), dim3(256), 0, s, svals, X->vals, perm2, nnz);
    HS_HIP(hipGetLastError());
    HS_TRY(seg_build(v.seg, N, mode, X->ndims[mode], nnz, sorted, svals, true, s, X->ndims));
    cleanup();
    v.use_seg = true;
    v.ready = true;
    return 0;
#undef HS_TRY
#undef HS_HIP
}

static std::atomic<int> g_view_kind{SPTOL_HICOO_VIEW_AUTO};   // process-wide, set from any thread

}  // namespace sptol

using namespace sptol;

extern "C" {

int sptolSetHicooViewKind(int kind)
{
    if (kind < SPTOL_HICOO_VIEW_AUTO || kind > SPTOL_HICOO_VIEW_SEG)
        SPTOL_CHECK(SPTERR_VALUE_ERROR, "HiCOO view", "unknown view kind");
    g_view_kind = kind;
    return 0;
}

int sptolDevHicooPrepareMode(sptolDevHicoo h, sptIndex mode) { return hicoo_prepare(h, mode, nullptr); }

}  // extern "C"

namespace sptol {

int hicoo_prepare(sptolDevHicoo h, uint32_t mode, hipStream_t s)
{
    if (mode >= h->nmodes) SPTOL_CHECK(SPTERR_VALUE_ERROR, "HiCOO view", "mode out of range");
    if (!h->views) h->views = new sptolHicooViews();
    HicooView &v = h->views->v[mode];
    if (v.ready || h->nnz == 0) return 0;
    const bool seg_shape = h->nmodes == 3 || h->nmodes == 4;
    bool use_seg = false;
    const int kind = g_view_kind.load();
    // The segmented view wins at every block density measured, dense
    // blocks included (651 nonzeros a block: 0.013 against 0.094 ms per
    // MTTKRP; 15 and 41 a block: about 10x), so the default takes it
    // wherever its kernel runs and its sort takes the nonzeros; the tiled
    // view (a sort of the blocks) stays for explicit requests and beyond.
    if (kind == SPTOL_HICOO_VIEW_SEG) use_seg = seg_shape;
    else if (kind == SPTOL_HICOO_VIEW_AUTO) use_seg = seg_shape && h->nnz <= kMaxSortItems;
    if (use_seg) return build_hicoo_seg(h, mode, v, s);
    return build_hicoo_view(h, mode, v, s);
}

const SegView *hicoo_seg_records(sptolDevHicoo X)
{
    if (!X->views) return nullptr;
    for (uint32_t q = 0; q < X->nmodes; ++q) {
        const HicooView &v = X->views->v[q];
        if (v.ready && v.use_seg && v.seg.ready && v.seg.rec && v.seg.nnz == X->nnz) return &v.seg;
    }
    return nullptr;
}

void free_hicoo_views(sptolDevHicoo h)
{
    if (!h->views) return;
    for (auto &v : h->views->v) v.release();
    delete h->views;
    h->views = nullptr;
}

// Run the mode's view (segmented, or zero fill + tiled kernel); *used = 0
// when no view applies (the caller runs the block-order kernel).
int launch_hicoo_view(sptolDevHicoo X, const MttkrpArgs &in, int dtype, hipStream_t s, int *used,
                      const AlsFactorReq *fr, bool *factored)
{
    *used = 0;
    if (!X->views || !X->views->v[in.mode].ready) return 0;
    HicooView &v = X->views->v[in.mode];
    if (v.use_seg) {
        if (!seg_ok(in.nmodes, in.stride)) return 0;
        *used = 1;
        return seg_mttkrp(v.seg, in.mats, in.order, in.R, in.stride, dtype, in.out, s, fr, factored);
    }
    if (!hicoo_tiled_ok(in.sb_bits, in.stride)) return 0;
    *used = 1;
    SPTOL_TRY(launch_fill_zero(in.out, (uint64_t)X->ndims[in.mode] * in.stride, s));
    if (v.nitems == 0) return 0;
    MttkrpArgs a = in;
    a.bptr = v.bptr;
    for (uint32_t m = 0; m < a.nmodes; ++m) { a.binds[m] = v.binds[m]; a.einds[m] = v.einds[m]; }
    a.vals = v.vals;
    uint32_t lpn = a.stride / 4;
    if (lpn > 16) lpn = 16;
    uint32_t p = 1;
    while (p < lpn) p <<= 1;
    const size_t shm = (size_t)(1u << a.sb_bits) * a.stride * 4;
    if (v.nitems > 0x7FFFFFFFull) SPTOL_CHECK(SPTERR_VALUE_ERROR, "HiCOO tiled", "too many items");
    dim3 grid((unsigned)v.nitems);
    int tok = prof_start("mttkrp_hicoo", s);
    if (dtype == SPTOL_BF16) HV_LPN(uint16_t, p, a.nmodes, grid, shm, s, a, v.items, X->nb);
    else HV_LPN(float, p, a.nmodes, grid, shm, s, a, v.items, X->nb);
    prof_stop(tok, s);
    SPTOL_HIP(hipGetLastError(), "HiCOO tiled launch");
    return 0;
}

}  // namespace sptol
