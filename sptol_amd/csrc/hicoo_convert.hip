// COO -> HiCOO on the GPU (sptolSparseTensorToHiCOODevice).
//
// Produces the same arrays, bit for bit, as the host converter
// sptSparseTensorToHiCOO (the reference's convert.c:300-497) and leaves the
// caller's COO tensor permuted into HiCOO order, as the reference does.
//   1. Order: stable LSD radix sorts (hipCUB) over the same key chunks as
//      the host: the Z-Morton key of the low min(sk_bits, 24|32) index bits
//      (3-D and 4-D tensors only, mode 0 the most significant bit of each
//      group), then the superblock coordinates (i_m >> sk_bits) packed
//      lexicographically into 64-bit chunks, least significant chunk first.
//   2. Blocks and superblocks start where some i_m >> sb_bits (resp.
//      sk_bits) changes: flags, inclusive scans, scatters give bptr, binds,
//      einds, kptr (as block offsets).
//   3. Chunk pointers (cptr) follow the reference's rule, run serially per
//      superblock (one thread each): a superblock opens a chunk; a later
//      block opens one when the running chunk size plus the previous
//      block's count reaches 2^14, the previous block's count carrying into
//      the first block of a superblock (convert.c:386-456).
//   4. Per-mode superblock schedules kschr[m][r]: superblock ids with
//      i_m >> sk_bits == r, ascending (a stable sort by r); nkiters[m] is the
//      longest list (convert.c:103-146).
#include "dev_common.hpp"
#include <hipcub/hipcub.hpp>
#include <vector>
#include <cstring>
#include <cstdlib>

namespace sptol {

namespace {

enum KeyKind { kMortonLo = 0, kMortonHi = 1, kRowBlock = 2 };

struct KeyChunk {
    int kind;
    uint32_t m_lo, m_hi;     // kRowBlock: modes [m_lo, m_hi) packed, most significant first
    uint32_t shift;          // kRowBlock: right shift of each index
    uint32_t mbits;          // Morton: low bits of each index interleaved
    uint32_t bits[8];        // kRowBlock: per-mode widths
    uint32_t total;          // key bits
};

struct Coords {
    const uint32_t *c[8];
};

unsigned bits_for(uint64_t n)
{
    unsigned b = 0;
    while (b < 64 && n > (1ull << b)) ++b;
    return b;
}

// The row-block (superblock) key split into 64-bit chunks, least
// significant first (the host converter's lex_chunks).
int rowblock_chunks(uint32_t N, const uint32_t *ndims, uint32_t shift, KeyChunk *out)
{
    int nc = 0;
    uint32_t hi = N;
    while (hi > 0) {
        KeyChunk c;
        std::memset(&c, 0, sizeof(c));
        c.kind = kRowBlock;
        c.shift = shift;
        c.m_hi = hi;
        uint32_t lo = hi, tot = 0;
        while (lo > 0) {
            unsigned w = bits_for(((uint64_t)ndims[lo - 1] >> shift) + 1);
            if (w == 0) w = 1;
            if (tot + w > 64) break;
            tot += w;
            --lo;
        }
        c.m_lo = lo;
        for (uint32_t m = lo; m < hi; ++m) {
            unsigned w = bits_for(((uint64_t)ndims[m] >> shift) + 1);
            c.bits[m - lo] = w ? w : 1;
        }
        c.total = tot;
        out[nc++] = c;
        hi = lo;
    }
    return nc;
}

}  // namespace

__global__ void hc_iota(uint32_t *p, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = (uint32_t)i;
}

__global__ void hc_check(const Coords c, uint32_t N, const uint32_t *ndims_d, uint64_t nnz, uint32_t *bad)
{
    for (uint64_t z = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; z < nnz; z += (uint64_t)gridDim.x * blockDim.x)
        for (uint32_t m = 0; m < N; ++m)
            if (c.c[m][z] >= ndims_d[m]) atomicOr(bad, 1u);
}

// key of element perm[z] under chunk k
__global__ void hc_key(uint64_t *__restrict__ keys, const uint32_t *__restrict__ perm, const Coords c, uint32_t N,
                       uint64_t nnz, const KeyChunk k)
{
    for (uint64_t z = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; z < nnz; z += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t x = perm[z];
        uint64_t key = 0;
        if (k.kind == kRowBlock) {
            for (uint32_t m = k.m_lo; m < k.m_hi; ++m)
                key = (key << k.bits[m - k.m_lo]) | (uint64_t)(c.c[m][x] >> k.shift);
        } else {
            uint32_t v[4];
            for (uint32_t m = 0; m < N; ++m) v[m] = c.c[m][x];
            uint64_t lo = 0, hi = 0;
            for (int b = (int)k.mbits - 1; b >= 0; --b)
                for (uint32_t m = 0; m < N; ++m) {
                    hi = (hi << 1) | (lo >> 63);
                    lo = (lo << 1) | ((v[m] >> b) & 1u);
                }
            key = k.kind == kMortonLo ? lo : hi;
        }
        keys[z] = key;
    }
}

template <typename T>
__global__ void hc_gather(T *__restrict__ dst, const T *__restrict__ src, const uint32_t *__restrict__ perm, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[perm[i]];
}

// newb / newk flags (1 at z = 0)
__global__ void hc_flags(const Coords s, uint32_t N, uint64_t nnz, uint32_t sb, uint32_t sk, uint32_t *__restrict__ fb,
                         uint32_t *__restrict__ fk)
{
    for (uint64_t z = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; z < nnz; z += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t nb = z == 0, nk = z == 0;
        if (z) {
            for (uint32_t m = 0; m < N; ++m) {
                const uint32_t a = s.c[m][z], b = s.c[m][z - 1];
                if ((a >> sb) != (b >> sb)) nb = 1;
                if ((a >> sk) != (b >> sk)) { nk = 1; nb = 1; break; }
            }
        }
        fb[z] = nb;
        fk[z] = nk;
    }
}

// bid / kid are inclusive scans of the flags (1-based ids)
__global__ void hc_scatter(const Coords s, uint32_t N, uint64_t nnz, uint32_t sb, uint32_t sk,
                           const uint32_t *__restrict__ fb, const uint32_t *__restrict__ fk,
                           const uint32_t *__restrict__ bid, const uint32_t *__restrict__ kid,
                           uint64_t *__restrict__ bptr, uint32_t **binds, uint8_t **einds, uint64_t *__restrict__ kptr,
                           uint64_t *__restrict__ kstart)
{
    const uint32_t mask = (1u << sb) - 1u;
    for (uint64_t z = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; z < nnz; z += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t b = (uint64_t)bid[z] - 1;
        for (uint32_t m = 0; m < N; ++m) einds[m][z] = (uint8_t)(s.c[m][z] & mask);
        if (fb[z]) {
            bptr[b] = z;
            for (uint32_t m = 0; m < N; ++m) binds[m][b] = s.c[m][z] >> sb;
        }
        if (fk[z]) {
            const uint64_t k = (uint64_t)kid[z] - 1;
            kptr[k] = b;
            kstart[k] = z;
        }
    }
}

// Chunk pointers of superblock k: count (out == null) or write at off[k].
__global__ void hc_chunks(const uint64_t *__restrict__ bptr, const uint64_t *__restrict__ kptr, uint64_t nk,
                          uint64_t sc, uint64_t *__restrict__ cnt, const uint64_t *__restrict__ off,
                          uint64_t *__restrict__ out)
{
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nk; k += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t b0 = kptr[k], b1 = kptr[k + 1];
        uint64_t n = 0, chunk = 0;
        uint64_t w = out ? off[k] : 0;
        if (out) out[w++] = b0;
        ++n;
        for (uint64_t b = b0; b < b1; ++b) {
            if (b == 0) continue;   // the tensor's first block: no previous count
            const uint64_t ne = bptr[b] - bptr[b - 1];
            if (chunk + ne >= sc) {
                if (out) out[w++] = b;
                ++n;
                chunk = 0;
            } else {
                chunk += ne;
            }
        }
        if (cnt) cnt[k] = n;
    }
}

__global__ void hc_sched_keys(const uint32_t *__restrict__ coord, const uint64_t *__restrict__ kstart, uint64_t nk,
                              uint32_t sk, uint32_t *__restrict__ rk, uint32_t *__restrict__ q)
{
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nk; k += (uint64_t)gridDim.x * blockDim.x) {
        rk[k] = coord[kstart[k]] >> sk;
        q[k] = (uint32_t)k;
    }
}

static unsigned hc_grid(uint64_t n)
{
    uint64_t g = (n + 255) / 256;
    if (g > 65536) g = 65536;
    return (unsigned)(g ? g : 1);
}

// Device scratch freed on every exit path.
struct HcBufs {
    std::vector<void *> p;
    ~HcBufs() { for (void *x : p) dfree(x); }
    template <typename T>
    int get(T **out, uint64_t n, const char *what)
    {
        void *q = nullptr;
        int r = dmalloc(&q, (n ? n : 1) * sizeof(T), what);
        if (r) return r;
        p.push_back(q);
        *out = (T *)q;
        return 0;
    }
};

template <typename T>
static int replace_vec(T **data, sptNnzIndex *len, sptNnzIndex *cap, uint64_t n)
{
    free(*data);
    *data = (T *)malloc((n ? n : 1) * sizeof(T));
    SPTOL_CHECK_OS(!*data, "HiSpTns Convert (GPU)");
    *len = n;
    *cap = n ? n : 1;
    return 0;
}

}  // namespace sptol

using namespace sptol;

extern "C" {

int sptolSparseTensorToHiCOODevice(sptSparseTensorHiCOO *hitsr, sptNnzIndex *max_nnzb, sptSparseTensor *tsr,
                                   sptElementIndex sb_bits, sptElementIndex sk_bits)
{
    const char *mod = "HiSpTns Convert (GPU)";
    const sptElementIndex sc_bits = 14;
    if (sk_bits < sb_bits) SPTOL_CHECK(SPTERR_VALUE_ERROR, mod, "sk_bits < sb_bits");
    if (sb_bits > 8) SPTOL_CHECK(SPTERR_VALUE_ERROR, mod, "sb_bits > 8 does not fit 8-bit element offsets");
    if (sk_bits > 31) SPTOL_CHECK(SPTERR_VALUE_ERROR, mod, "sk_bits > 31");
    if (tsr->nnz == 0) SPTOL_CHECK(SPTERR_VALUE_ERROR, mod, "empty tensor");
    const uint32_t N = tsr->nmodes;
    const uint64_t nnz = tsr->nnz;
    if (N < 2 || N > 8) SPTOL_CHECK(SPTERR_VALUE_ERROR, mod, "nmodes must be in [2, 8]");
    if (nnz > kMaxSortItems) SPTOL_CHECK(SPTERR_VALUE_ERROR, mod, "nnz > 2^31 - 1 (the device sort counts items in 32-bit ints)");
    hipStream_t s = nullptr;
    HcBufs bufs;
    int r = 0;
#define HC_TRY(x) do { r = (x); if (r) return r; } while (0)
#define HC_HIP(x) SPTOL_HIP((x), mod)

    // upload and validate
    uint32_t *in[8] = {}, *sc[8] = {}, *ndims_d = nullptr, *bad = nullptr;
    float *vin = nullptr, *vsc = nullptr;
    for (uint32_t m = 0; m < N; ++m) {
        HC_TRY(bufs.get(&in[m], nnz, mod));
        if ((r = upload_pageable(in[m], tsr->inds[m].data, nnz * 4, s))) SPTOL_CHECK(r, mod, "upload failed");
    }
    HC_TRY(bufs.get(&vin, nnz, mod));
    if ((r = upload_pageable(vin, tsr->values.data, nnz * 4, s))) SPTOL_CHECK(r, mod, "upload failed");
    HC_TRY(bufs.get(&ndims_d, N, mod));
    HC_HIP(hipMemcpyAsync(ndims_d, tsr->ndims, N * 4, hipMemcpyHostToDevice, s));
    HC_TRY(bufs.get(&bad, 1, mod));
    HC_HIP(hipMemsetAsync(bad, 0, 4, s));
    Coords ci;
    std::memset(&ci, 0, sizeof(ci));
    for (uint32_t m = 0; m < N; ++m) ci.c[m] = in[m];
    hipLaunchKernelGGL(hc_check, dim3(hc_grid(nnz)), dim3(256), 0, s, ci, N, ndims_d, nnz, bad);
    uint32_t hbad = 0;
    HC_HIP(hipMemcpyAsync(&hbad, bad, 4, hipMemcpyDeviceToHost, s));
    HC_HIP(hipStreamSynchronize(s));
    if (hbad) SPTOL_CHECK(SPTERR_VALUE_ERROR, mod, "index >= ndims");

    // 1. order
    KeyChunk ch[24];
    int nc = 0;
    if (N == 3 || N == 4) {
        unsigned mb = N == 3 ? 24 : 32;
        if (sk_bits < mb) mb = sk_bits;
        const unsigned tot = N * mb;
        KeyChunk c;
        std::memset(&c, 0, sizeof(c));
        c.mbits = mb;
        c.kind = kMortonLo;
        c.total = tot > 64 ? 64 : tot;
        if (tot > 0) ch[nc++] = c;
        if (tot > 64) {
            c.kind = kMortonHi;
            c.total = tot - 64;
            ch[nc++] = c;
        }
    }
    nc += rowblock_chunks(N, tsr->ndims, sk_bits, ch + nc);
    uint32_t *perm = nullptr, *perm2 = nullptr;
    uint64_t *keys = nullptr, *keys2 = nullptr;
    HC_TRY(bufs.get(&perm, nnz, mod));
    HC_TRY(bufs.get(&perm2, nnz, mod));
    HC_TRY(bufs.get(&keys, nnz, mod));
    HC_TRY(bufs.get(&keys2, nnz, mod));
    hipLaunchKernelGGL(hc_iota, dim3(hc_grid(nnz)), dim3(256), 0, s, perm, nnz);
    size_t tb = 0;
    HC_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, keys, keys2, perm, perm2, (int)nnz, 0, 64, s));
    void *tmp = nullptr;
    HC_TRY(bufs.get((uint8_t **)&tmp, tb, mod));
    for (int c = 0; c < nc; ++c) {
        if (ch[c].total == 0) continue;
        hipLaunchKernelGGL(hc_key, dim3(hc_grid(nnz)), dim3(256), 0, s, keys, perm, ci, N, nnz, ch[c]);
        size_t tb2 = tb;
        HC_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tb2, keys, keys2, perm, perm2, (int)nnz, 0, (int)ch[c].total, s));
        std::swap(perm, perm2);
    }
    for (uint32_t m = 0; m < N; ++m) {
        HC_TRY(bufs.get(&sc[m], nnz, mod));
        hipLaunchKernelGGL(hc_gather<uint32_t>, dim3(hc_grid(nnz)), dim3(256), 0, s, sc[m], in[m], perm, nnz);
    }
    HC_TRY(bufs.get(&vsc, nnz, mod));
    hipLaunchKernelGGL(hc_gather<float>, dim3(hc_grid(nnz)), dim3(256), 0, s, vsc, vin, perm, nnz);
    Coords cs;
    std::memset(&cs, 0, sizeof(cs));
    for (uint32_t m = 0; m < N; ++m) cs.c[m] = sc[m];

    // 2. blocks and superblocks
    uint32_t *fb = nullptr, *fk = nullptr, *bid = nullptr, *kid = nullptr;
    HC_TRY(bufs.get(&fb, nnz, mod));
    HC_TRY(bufs.get(&fk, nnz, mod));
    HC_TRY(bufs.get(&bid, nnz, mod));
    HC_TRY(bufs.get(&kid, nnz, mod));
    hipLaunchKernelGGL(hc_flags, dim3(hc_grid(nnz)), dim3(256), 0, s, cs, N, nnz, (uint32_t)sb_bits, (uint32_t)sk_bits,
                       fb, fk);
    tb = 0;
    HC_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tb, fb, bid, (int)nnz, s));
    void *tmp2 = nullptr;
    HC_TRY(bufs.get((uint8_t **)&tmp2, tb, mod));
    size_t tb2 = tb;
    HC_HIP(hipcub::DeviceScan::InclusiveSum(tmp2, tb2, fb, bid, (int)nnz, s));
    tb2 = tb;
    HC_HIP(hipcub::DeviceScan::InclusiveSum(tmp2, tb2, fk, kid, (int)nnz, s));
    uint32_t last[2] = {0, 0};
    HC_HIP(hipMemcpyAsync(&last[0], bid + nnz - 1, 4, hipMemcpyDeviceToHost, s));
    HC_HIP(hipMemcpyAsync(&last[1], kid + nnz - 1, 4, hipMemcpyDeviceToHost, s));
    HC_HIP(hipStreamSynchronize(s));
    const uint64_t nb = last[0], nk = last[1];
    uint64_t *bptr = nullptr, *kptr = nullptr, *kstart = nullptr;
    uint32_t *binds[8] = {};
    uint8_t *einds[8] = {};
    uint32_t **d_binds = nullptr;
    uint8_t **d_einds = nullptr;
    HC_TRY(bufs.get(&bptr, nb + 1, mod));
    HC_TRY(bufs.get(&kptr, nk + 1, mod));
    HC_TRY(bufs.get(&kstart, nk, mod));
    for (uint32_t m = 0; m < N; ++m) {
        HC_TRY(bufs.get(&binds[m], nb, mod));
        HC_TRY(bufs.get(&einds[m], nnz, mod));
    }
    HC_TRY(bufs.get(&d_binds, 8, mod));
    HC_TRY(bufs.get(&d_einds, 8, mod));
    HC_HIP(hipMemcpyAsync(d_binds, binds, sizeof(binds), hipMemcpyHostToDevice, s));
    HC_HIP(hipMemcpyAsync(d_einds, einds, sizeof(einds), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(hc_scatter, dim3(hc_grid(nnz)), dim3(256), 0, s, cs, N, nnz, (uint32_t)sb_bits, (uint32_t)sk_bits,
                       fb, fk, bid, kid, bptr, d_binds, d_einds, kptr, kstart);
    const uint64_t tail_b = nnz, tail_k = nb;
    HC_HIP(hipMemcpyAsync(bptr + nb, &tail_b, 8, hipMemcpyHostToDevice, s));
    HC_HIP(hipMemcpyAsync(kptr + nk, &tail_k, 8, hipMemcpyHostToDevice, s));

    // 3. chunk pointers
    uint64_t *ccnt = nullptr, *coff = nullptr, *cptr = nullptr;
    HC_TRY(bufs.get(&ccnt, nk, mod));
    HC_TRY(bufs.get(&coff, nk + 1, mod));
    hipLaunchKernelGGL(hc_chunks, dim3(hc_grid(nk)), dim3(256), 0, s, bptr, kptr, nk, 1ull << sc_bits, ccnt,
                       (const uint64_t *)nullptr, (uint64_t *)nullptr);
    tb = 0;
    HC_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, ccnt, coff, (int)nk, s));
    void *tmp3 = nullptr;
    HC_TRY(bufs.get((uint8_t **)&tmp3, tb, mod));
    HC_HIP(hipcub::DeviceScan::ExclusiveSum(tmp3, tb, ccnt, coff, (int)nk, s));
    uint64_t lastc[2] = {0, 0};
    HC_HIP(hipMemcpyAsync(&lastc[0], coff + nk - 1, 8, hipMemcpyDeviceToHost, s));
    HC_HIP(hipMemcpyAsync(&lastc[1], ccnt + nk - 1, 8, hipMemcpyDeviceToHost, s));
    HC_HIP(hipStreamSynchronize(s));
    const uint64_t ncp = lastc[0] + lastc[1];   // + the closing entry nb
    HC_TRY(bufs.get(&cptr, ncp + 1, mod));
    hipLaunchKernelGGL(hc_chunks, dim3(hc_grid(nk)), dim3(256), 0, s, bptr, kptr, nk, 1ull << sc_bits,
                       (uint64_t *)nullptr, (const uint64_t *)coff, cptr);
    HC_HIP(hipMemcpyAsync(cptr + ncp, &tail_k, 8, hipMemcpyHostToDevice, s));

    // 4. per-mode superblock schedules
    std::vector<std::vector<uint32_t>> sched_q(N), sched_r(N);
    {
        uint32_t *rk = nullptr, *q = nullptr, *rk2 = nullptr, *q2 = nullptr;
        HC_TRY(bufs.get(&rk, nk, mod));
        HC_TRY(bufs.get(&q, nk, mod));
        HC_TRY(bufs.get(&rk2, nk, mod));
        HC_TRY(bufs.get(&q2, nk, mod));
        tb = 0;
        HC_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, rk, rk2, q, q2, (int)nk, 0, 32, s));
        void *tmp4 = nullptr;
        HC_TRY(bufs.get((uint8_t **)&tmp4, tb, mod));
        for (uint32_t m = 0; m < N; ++m) {
            hipLaunchKernelGGL(hc_sched_keys, dim3(hc_grid(nk)), dim3(256), 0, s, sc[m], kstart, nk, (uint32_t)sk_bits,
                               rk, q);
            size_t t4 = tb;
            HC_HIP(hipcub::DeviceRadixSort::SortPairs(tmp4, t4, rk, rk2, q, q2, (int)nk, 0, 32, s));
            sched_q[m].resize(nk);
            sched_r[m].resize(nk);
            HC_HIP(hipMemcpyAsync(sched_q[m].data(), q2, nk * 4, hipMemcpyDeviceToHost, s));
            HC_HIP(hipMemcpyAsync(sched_r[m].data(), rk2, nk * 4, hipMemcpyDeviceToHost, s));
            HC_HIP(hipStreamSynchronize(s));
        }
    }
    HC_HIP(hipGetLastError());

    // 5. host arrays
    HC_TRY(sptNewSparseTensorHiCOO(hitsr, N, tsr->ndims, nnz, sb_bits, sk_bits, sc_bits));
    auto fail = [&](int code) {
        sptFreeSparseTensorHiCOO(hitsr);
        return code;
    };
    if ((r = replace_vec(&hitsr->kptr.data, &hitsr->kptr.len, &hitsr->kptr.cap, nk + 1))) return fail(r);
    if ((r = replace_vec(&hitsr->bptr.data, &hitsr->bptr.len, &hitsr->bptr.cap, nb + 1))) return fail(r);
    if ((r = replace_vec(&hitsr->cptr.data, &hitsr->cptr.len, &hitsr->cptr.cap, ncp + 1))) return fail(r);
    if ((r = replace_vec(&hitsr->values.data, &hitsr->values.len, &hitsr->values.cap, nnz))) return fail(r);
    for (uint32_t m = 0; m < N; ++m) {
        if ((r = replace_vec(&hitsr->binds[m].data, &hitsr->binds[m].len, &hitsr->binds[m].cap, nb))) return fail(r);
        if ((r = replace_vec(&hitsr->einds[m].data, &hitsr->einds[m].len, &hitsr->einds[m].cap, nnz))) return fail(r);
    }
    // the host arrays are fresh (their pages not yet touched): each comes
    // down through pinned chunks, copied out on the host's threads
    r = download_parallel(hitsr->kptr.data, kptr, (nk + 1) * 8, s);
    if (!r) r = download_parallel(hitsr->bptr.data, bptr, (nb + 1) * 8, s);
    if (!r) r = download_parallel(hitsr->cptr.data, cptr, (ncp + 1) * 8, s);
    if (!r) r = download_parallel(hitsr->values.data, vsc, nnz * 4, s);
    for (uint32_t m = 0; m < N && !r; ++m) {
        r = download_parallel(hitsr->binds[m].data, binds[m], nb * 4, s);
        if (!r) r = download_parallel(hitsr->einds[m].data, einds[m], nnz, s);
        // the caller's tensor, permuted into HiCOO order
        if (!r) r = download_parallel(tsr->inds[m].data, sc[m], nnz * 4, s);
    }
    if (!r) r = download_parallel(tsr->values.data, vsc, nnz * 4, s);
    if (r) return fail(r);
    for (uint32_t m = 0; m < N; ++m) tsr->sortorder[m] = m;
    for (uint32_t m = 0; m < N; ++m) {
        const uint64_t kd = ((uint64_t)tsr->ndims[m] + (1ull << sk_bits) - 1) >> sk_bits;
        std::vector<uint64_t> cnt(kd ? kd : 1, 0);
        for (uint64_t k = 0; k < nk; ++k) cnt[sched_r[m][k]]++;
        sptIndex mx = 0;
        uint64_t pos = 0;
        for (uint64_t i = 0; i < kd; ++i) {
            sptIndexVector *v = &hitsr->kschr[m][i];
            free(v->data);
            v->cap = cnt[i] < 2 ? 2 : cnt[i];
            v->data = (sptIndex *)malloc(v->cap * sizeof(sptIndex));
            if (!v->data) return fail(SPTERR_OS_ERROR + 12);
            v->len = cnt[i];
            if (cnt[i]) std::memcpy(v->data, sched_q[m].data() + pos, cnt[i] * sizeof(sptIndex));
            pos += cnt[i];
            if (cnt[i] > mx) mx = (sptIndex)cnt[i];
        }
        hitsr->nkiters[m] = mx;
    }
    if (max_nnzb) {
        sptNnzIndex mx = 0;
        for (uint64_t q = 0; q < nb; ++q) {
            const sptNnzIndex c = hitsr->bptr.data[q + 1] - hitsr->bptr.data[q];
            if (c > mx) mx = c;
        }
        *max_nnzb = mx;
    }
    return 0;
#undef HC_TRY
#undef HC_HIP
}

}  // extern "C"
