// hipwrap: the C-ABI device layer that replaces src/cudawrap.{h,cu}.
//
// The reference duplicates host arrays with cudaMalloc+cudaMemcpy
// (cudawrap.cu:54-93) and packs arrays of device pointers into one
// allocation (cudawrap.h:66-350).  Here every device allocation goes through
// one budget-tracked allocator (so a serving loop can admit work by HBM
// budget, and tests can inject allocation failures), and kernels take
// pointers by value in their argument block instead of through device-side
// pointer arrays.
#include "dev_common.hpp"
#include <atomic>
#include <mutex>
#include <unordered_map>
#include <map>
#include <cstdlib>
#include <vector>
#include <cstring>
#include <cstdio>
#include <string>

namespace sptol {

static std::mutex g_mem_mu;
struct LiveBlock { size_t bytes, rounded; };
static std::unordered_map<void *, LiveBlock> g_live;  // device pointer -> sizes
static size_t g_in_use = 0;
static size_t g_budget = 0;                          // 0 = unlimited
static int g_fail_countdown = 0;
static std::atomic<int> g_verbose{0};

int verbose() { return g_verbose.load(); }

// ---------------------------------------------------------------------------
// Block cache.  hipFree waits for every stream on the device, so a serving
// loop that frees a finished request's arrays while other requests run
// would stall them all.  Blocks freed inside a StreamScope are kept instead,
// tagged with an event recorded on the scope's stream: a later allocation
// on that same stream may take the block at once (stream order covers it),
// an allocation elsewhere once the event has completed.  Blocks freed
// outside a scope go back to HIP as before.  Sizes are rounded up to one of
// eight steps per power of two so that blocks of similar size are shared.
// ---------------------------------------------------------------------------
struct CachedBlock {
    void *p;
    hipStream_t s;
    hipEvent_t ev;   // nullptr: no pending work
};
static std::multimap<size_t, CachedBlock> g_cache;   // rounded size -> block
static size_t g_cached = 0;
static size_t g_cache_cap = (size_t)64 << 30;
static bool g_cache_cap_set = false;
static std::vector<hipEvent_t> g_spare_events;
static thread_local hipStream_t tl_stream = nullptr;
static thread_local int tl_scoped = 0;

StreamScope::StreamScope(hipStream_t s) : prev_(tl_stream), prev_on_(tl_scoped)
{
    tl_stream = s;
    tl_scoped = 1;
}

StreamScope::~StreamScope()
{
    tl_stream = prev_;
    tl_scoped = prev_on_;
}

static size_t round_block(size_t n)
{
    if (n <= 512) return 512;
    int lg = 63 - __builtin_clzll((unsigned long long)n);
    const size_t step = (size_t)1 << (lg - 3);
    return (n + step - 1) / step * step;
}

// Take a cached block of `rounded` bytes usable on the caller's stream
// (caller holds g_mem_mu).
static void *cache_take(size_t rounded)
{
    auto range = g_cache.equal_range(rounded);
    int scanned = 0;
    for (auto it = range.first; it != range.second && scanned < 16; ++it, ++scanned) {
        CachedBlock &b = it->second;
        bool ok = b.ev == nullptr || (tl_scoped && b.s == tl_stream);
        if (!ok && hip_quiet([&] { return hipEventQuery(b.ev); }) == hipSuccess) ok = true;
        if (!ok) continue;
        void *p = b.p;
        if (b.ev) g_spare_events.push_back(b.ev);
        g_cache.erase(it);
        g_cached -= rounded;
        return p;
    }
    return nullptr;
}

// Return every cached block to HIP after the device has drained.
static void cache_release_all()
{
    std::vector<void *> ps;
    std::vector<hipEvent_t> evs;
    {
        std::lock_guard<std::mutex> lk(g_mem_mu);
        for (auto &kv : g_cache) {
            ps.push_back(kv.second.p);
            if (kv.second.ev) evs.push_back(kv.second.ev);
        }
        g_cache.clear();
        g_cached = 0;
        for (auto e : g_spare_events) evs.push_back(e);
        g_spare_events.clear();
    }
    if (ps.empty() && evs.empty()) return;
    (void)hipDeviceSynchronize();
    for (void *p : ps) (void)hipFree(p);
    for (auto e : evs) (void)hipEventDestroy(e);
}

static size_t cache_cap()
{
    if (!g_cache_cap_set) {
        g_cache_cap_set = true;
        if (const char *e = getenv("SPTOL_CACHE_BYTES")) g_cache_cap = (size_t)strtoull(e, nullptr, 10);
    }
    return g_cache_cap;
}

// ---------------------------------------------------------------------------
// Pinned host slots: small page-locked buffers (fits, weights) that async
// device-to-host copies land in.  Carved from 1 MiB pinned slabs that live
// until the process ends, so taking and returning a slot never calls the
// runtime's pinned allocator (whose free waits for the device).
// ---------------------------------------------------------------------------
static std::mutex g_pin_mu;
static std::vector<void *> g_pin_free;

int pinned_slot(void **p)
{
    std::lock_guard<std::mutex> lk(g_pin_mu);
    if (g_pin_free.empty()) {
        void *slab = nullptr;
        SPTOL_HIP(hipHostMalloc(&slab, (size_t)1 << 20, hipHostMallocDefault), "HipWrap pinned slot");
        for (size_t k = 0; k < ((size_t)1 << 20) / kPinnedSlotBytes; ++k)
            g_pin_free.push_back((char *)slab + k * kPinnedSlotBytes);
    }
    *p = g_pin_free.back();
    g_pin_free.pop_back();
    return 0;
}

void pinned_slot_free(void *p)
{
    if (!p) return;
    std::lock_guard<std::mutex> lk(g_pin_mu);
    g_pin_free.push_back(p);
}

// Pinned chunks for streaming file data to the device (kPinnedChunkBytes
// each), kept for the life of the process and shared by all threads.
static std::mutex g_chunk_mu;
static std::vector<void *> g_chunk_free;

int pinned_chunk(void **p)
{
    {
        std::lock_guard<std::mutex> lk(g_chunk_mu);
        if (!g_chunk_free.empty()) {
            *p = g_chunk_free.back();
            g_chunk_free.pop_back();
            return 0;
        }
    }
    SPTOL_HIP(hipHostMalloc(p, kPinnedChunkBytes, hipHostMallocDefault), "HipWrap pinned chunk");
    return 0;
}

void pinned_chunk_free(void *p)
{
    if (!p) return;
    std::lock_guard<std::mutex> lk(g_chunk_mu);
    g_chunk_free.push_back(p);
}

int download_parallel(void *dst, const void *src, size_t n, hipStream_t s)
{
    if (!n) return 0;
    if (n <= 2 * kPinnedChunkBytes) {   // small: one runtime copy
        SPTOL_HIP(hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, s), "Download");
        SPTOL_HIP(hipStreamSynchronize(s), "Download");
        return 0;
    }
    void *pin[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    int r = pinned_chunk(&pin[0]);
    if (!r) r = pinned_chunk(&pin[1]);
    for (int k = 0; k < 2 && !r; ++k)
        if (hipEventCreateWithFlags(&ev[k], hipEventDisableTiming) != hipSuccess) r = SPTERR_CUDA_ERROR;
    const size_t C = kPinnedChunkBytes, nc = (n + C - 1) / C;
    auto issue = [&](size_t i) -> int {
        const size_t off = i * C, len = n - off < C ? n - off : C;
        SPTOL_HIP(hipMemcpyAsync(pin[i & 1], (const uint8_t *)src + off, len, hipMemcpyDeviceToHost, s), "Download");
        SPTOL_HIP(hipEventRecord(ev[i & 1], s), "Download");
        return 0;
    };
    if (!r) r = issue(0);
    for (size_t i = 0; i < nc && !r; ++i) {
        // chunk i + 1 goes to the other buffer, whose copy-out (chunk i - 1)
        // finished in the previous iteration
        if (i + 1 < nc) r = issue(i + 1);
        if (!r && hipEventSynchronize(ev[i & 1]) != hipSuccess) r = SPTERR_CUDA_ERROR;
        if (!r) {
            const size_t off = i * C, len = n - off < C ? n - off : C;
            sptol_parallel_memcpy((uint8_t *)dst + off, pin[i & 1], len);
        }
    }
    if (r) (void)hip_quiet([&] { return hipStreamSynchronize(s); });
    for (int k = 0; k < 2; ++k) {
        if (ev[k]) (void)hipEventDestroy(ev[k]);
        pinned_chunk_free(pin[k]);
    }
    if (r == SPTERR_CUDA_ERROR) SPTOL_CHECK(SPTERR_CUDA_ERROR, "Download", "device-to-host copy failed");
    return r;
}

int upload_pageable(void *dst, const void *src, size_t bytes, hipStream_t s)
{
    static const int chunked = [] {
        const char *e = std::getenv("SPTOL_UPLOAD_CHUNKED");
        return e ? std::atoi(e) : 1;
    }();
    hipError_t e = hipSuccess;
    if (!chunked || bytes < ((size_t)4 << 20)) {
        e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        return e == hipSuccess ? 0 : SPTERR_CUDA_ERROR + (int)e;
    }
    void *buf[2] = {};
    hipEvent_t ev[2] = {};
    bool used[2] = {false, false};
    int r = 0;
    for (int k = 0; k < 2 && !r; ++k) {
        r = pinned_chunk(&buf[k]);
        if (!r && (e = hipEventCreateWithFlags(&ev[k], hipEventDisableTiming)) != hipSuccess)
            r = SPTERR_CUDA_ERROR + (int)e;
    }
    int k = 0;
    for (size_t off = 0; off < bytes && !r; off += kPinnedChunkBytes) {
        const size_t n = std::min(kPinnedChunkBytes, bytes - off);
        if (used[k] && (e = hipEventSynchronize(ev[k])) != hipSuccess) { r = SPTERR_CUDA_ERROR + (int)e; break; }
        sptol_parallel_memcpy(buf[k], (const char *)src + off, n);
        e = hipMemcpyAsync((char *)dst + off, buf[k], n, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipEventRecord(ev[k], s);
        if (e != hipSuccess) { r = SPTERR_CUDA_ERROR + (int)e; break; }
        used[k] = true;
        k ^= 1;
    }
    e = hipStreamSynchronize(s);
    if (!r && e != hipSuccess) r = SPTERR_CUDA_ERROR + (int)e;
    for (int q = 0; q < 2; ++q) {
        pinned_chunk_free(buf[q]);
        if (ev[q]) (void)hipEventDestroy(ev[q]);
    }
    return r;
}

int PinnedStage::upload(void *dev, const void *host, size_t n, hipStream_t s)
{
    if (!n) return 0;
    const int k = cur;
    cur ^= 1;
    if (used[k]) SPTOL_HIP(hipEventSynchronize(ev[k]), "Pinned stage");
    used[k] = false;
    if (cap[k] < n) {
        if (buf[k]) (void)hipHostFree(buf[k]);
        buf[k] = nullptr;
        cap[k] = 0;
        const size_t c = n < 65536 ? 65536 : n * 2;
        SPTOL_HIP(hipHostMalloc((void **)&buf[k], c, hipHostMallocDefault), "Pinned stage");
        cap[k] = c;
    }
    if (!ev[k]) SPTOL_HIP(hipEventCreateWithFlags(&ev[k], hipEventDisableTiming), "Pinned stage");
    std::memcpy(buf[k], host, n);
    SPTOL_HIP(hipMemcpyAsync(dev, buf[k], n, hipMemcpyHostToDevice, s), "Pinned stage");
    SPTOL_HIP(hipEventRecord(ev[k], s), "Pinned stage");
    used[k] = true;
    return 0;
}

PinnedStage::~PinnedStage()
{
    for (int k = 0; k < 2; ++k) {
        if (used[k]) (void)hipEventSynchronize(ev[k]);
        if (buf[k]) (void)hipHostFree(buf[k]);
        if (ev[k]) (void)hipEventDestroy(ev[k]);
    }
}

int dmalloc(void **ptr, size_t bytes, const char *module)
{
    *ptr = nullptr;
    if (bytes == 0) return 0;
    const size_t rounded = round_block(bytes);
    {
        std::lock_guard<std::mutex> lk(g_mem_mu);
        if (g_fail_countdown > 0 && --g_fail_countdown == 0) {
            sptol_complain(module, SPTERR_CUDA_ERROR + (int)hipErrorOutOfMemory, __FILE__, __LINE__,
                           "injected allocation failure");
            return SPTERR_CUDA_ERROR + (int)hipErrorOutOfMemory;
        }
        if (g_budget && g_in_use + bytes > g_budget) {
            sptol_complain(module, SPTERR_CUDA_ERROR + (int)hipErrorOutOfMemory, __FILE__, __LINE__,
                           "device memory budget exceeded");
            return SPTERR_CUDA_ERROR + (int)hipErrorOutOfMemory;
        }
        g_in_use += bytes;
        if (void *p = cache_take(rounded)) {
            g_live[p] = LiveBlock{bytes, rounded};
            *ptr = p;
            return 0;
        }
    }
    hipError_t e = hipMalloc(ptr, rounded);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        cache_release_all();
        e = hipMalloc(ptr, rounded);
    }
    if (e != hipSuccess) {
        std::lock_guard<std::mutex> lk(g_mem_mu);
        g_in_use -= bytes;
        *ptr = nullptr;
        sptol_complain(module, SPTERR_CUDA_ERROR + (int)e, __FILE__, __LINE__, hipGetErrorString(e));
        return SPTERR_CUDA_ERROR + (int)e;
    }
    std::lock_guard<std::mutex> lk(g_mem_mu);
    g_live[*ptr] = LiveBlock{bytes, rounded};
    return 0;
}

void dfree(void *ptr)
{
    if (!ptr) return;
    size_t rounded = 0;
    hipEvent_t ev = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mem_mu);
        auto it = g_live.find(ptr);
        if (it != g_live.end()) {
            g_in_use -= it->second.bytes;
            rounded = it->second.rounded;
            g_live.erase(it);
        }
        if (rounded && tl_scoped && g_cached + rounded <= cache_cap()) {
            if (!g_spare_events.empty()) { ev = g_spare_events.back(); g_spare_events.pop_back(); }
        } else {
            rounded = 0;
        }
    }
    if (rounded) {
        if (!ev && hip_quiet([&] { return hipEventCreateWithFlags(&ev, hipEventDisableTiming); }) != hipSuccess)
            ev = nullptr;
        if (ev && hip_quiet([&] { return hipEventRecord(ev, tl_stream); }) == hipSuccess) {
            std::lock_guard<std::mutex> lk(g_mem_mu);
            g_cache.emplace(rounded, CachedBlock{ptr, tl_stream, ev});
            g_cached += rounded;
            return;
        }
        if (ev) (void)hip_quiet([&] { return hipEventDestroy(ev); });
    }
    (void)hip_quiet([&] { return hipFree(ptr); });
}

hipStream_t stream_of(void *s) { return (hipStream_t)s; }

// ---------------------------------------------------------------------------
// Kernel timing
// ---------------------------------------------------------------------------
struct ProfRec {
    std::string name;
    hipEvent_t a = nullptr, b = nullptr;
    bool closed = false;
};
static std::mutex g_prof_mu;
static std::atomic<int> g_prof_on{0};
// open and closed records by token; a drain takes only the closed ones, so a
// launch still between prof_start and prof_stop on another thread keeps its
// record (and its token cannot name another launch's record)
static std::unordered_map<int, ProfRec> g_prof;
static int g_prof_next = 0;
struct ProfTotal { double ms = 0; uint64_t n = 0; };
static std::unordered_map<std::string, ProfTotal> g_prof_tot;

int prof_start(const char *name, hipStream_t s)
{
    if (!g_prof_on.load()) return -1;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return -1;
    ProfRec r;
    r.name = name;
    if (hipEventCreate(&r.a) != hipSuccess || hipEventCreate(&r.b) != hipSuccess) return -1;
    if (hipEventRecord(r.a, s) != hipSuccess) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
        return -1;
    }
    std::lock_guard<std::mutex> lk(g_prof_mu);
    const int token = g_prof_next;
    g_prof_next = g_prof_next == 0x7FFFFFFF ? 0 : g_prof_next + 1;
    g_prof[token] = r;
    return token;
}

void prof_stop(int token, hipStream_t s)
{
    if (token < 0) return;
    std::lock_guard<std::mutex> lk(g_prof_mu);
    auto it = g_prof.find(token);
    if (it != g_prof.end()) {
        (void)hipEventRecord(it->second.b, s);
        it->second.closed = true;
    }
}

static void prof_drain()
{
    std::lock_guard<std::mutex> lk(g_prof_mu);
    for (auto it = g_prof.begin(); it != g_prof.end();) {
        ProfRec &r = it->second;
        if (!r.closed) {
            ++it;
            continue;
        }
        float ms = 0;
        if (hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            auto &t = g_prof_tot[r.name];
            t.ms += ms;
            t.n += 1;
        }
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
        it = g_prof.erase(it);
    }
}

}  // namespace sptol

using namespace sptol;

extern "C" {

int sptolHipSetDevice(int device)
{
    SPTOL_HIP(hipSetDevice(device), "HipWrap SetDevice");
    return 0;
}

int sptolHipGetLastError(void)
{
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : SPTERR_CUDA_ERROR + (int)e;
}

int sptolHipDeviceCount(int *count)
{
    SPTOL_HIP(hipGetDeviceCount(count), "HipWrap DeviceCount");
    return 0;
}

int sptolHipMalloc(void **ptr, size_t bytes) { return dmalloc(ptr, bytes, "HipWrap Malloc"); }

int sptolHipFree(void *ptr)
{
    dfree(ptr);
    return 0;
}

int sptolHipMemcpyH2D(void *dst, const void *src, size_t bytes, void *stream)
{
    if (!bytes) return 0;
    SPTOL_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream_of(stream)), "HipWrap H2D");
    SPTOL_HIP(hipStreamSynchronize(stream_of(stream)), "HipWrap H2D");
    return 0;
}

int sptolHipMemcpyD2H(void *dst, const void *src, size_t bytes, void *stream)
{
    if (!bytes) return 0;
    SPTOL_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream_of(stream)), "HipWrap D2H");
    SPTOL_HIP(hipStreamSynchronize(stream_of(stream)), "HipWrap D2H");
    return 0;
}

int sptolHipDuplicateMemory(void **dst, const void *src, size_t bytes)
{
    SPTOL_TRY(dmalloc(dst, bytes, "HipWrap Duplicate"));
    int r = sptolHipMemcpyH2D(*dst, src, bytes, nullptr);
    if (r) { dfree(*dst); *dst = nullptr; }
    return r;
}

int sptolHipStreamCreate(void **stream)
{
    hipStream_t s = nullptr;
    SPTOL_HIP(hipStreamCreate(&s), "HipWrap StreamCreate");
    *stream = (void *)s;
    return 0;
}

int sptolHipStreamDestroy(void *stream)
{
    // drain first: cached blocks tagged with this stream must not meet a
    // new stream at the same address while the old work is still pending
    if (stream) SPTOL_HIP(hipStreamSynchronize((hipStream_t)stream), "HipWrap StreamDestroy");
    if (stream) SPTOL_HIP(hipStreamDestroy((hipStream_t)stream), "HipWrap StreamDestroy");
    return 0;
}

int sptolHipStreamSynchronize(void *stream)
{
    SPTOL_HIP(hipStreamSynchronize(stream_of(stream)), "HipWrap Sync");
    return 0;
}

int sptolHipSetMemoryBudget(size_t bytes)
{
    std::lock_guard<std::mutex> lk(g_mem_mu);
    g_budget = bytes;
    return 0;
}

size_t sptolHipMemoryInUse(void)
{
    std::lock_guard<std::mutex> lk(g_mem_mu);
    return g_in_use;
}

size_t sptolHipMemoryCached(void)
{
    std::lock_guard<std::mutex> lk(g_mem_mu);
    return g_cached;
}

int sptolHipReleaseCache(void)
{
    cache_release_all();
    return 0;
}

int sptolHipInjectMallocFailure(int n)
{
    std::lock_guard<std::mutex> lk(g_mem_mu);
    g_fail_countdown = n > 0 ? n : 0;
    return 0;
}

void sptolSetVerbose(int v) { g_verbose.store(v); }

void sptolProfileEnable(int on) { g_prof_on.store(on ? 1 : 0); }

void sptolProfileReset(void)
{
    prof_drain();
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof_tot.clear();
}

int sptolProfileQuery(const char *name, double *total_ms, uint64_t *count)
{
    prof_drain();
    std::lock_guard<std::mutex> lk(g_prof_mu);
    auto it = g_prof_tot.find(name);
    *total_ms = it == g_prof_tot.end() ? 0.0 : it->second.ms;
    *count = it == g_prof_tot.end() ? 0 : it->second.n;
    return 0;
}

// ---------------------------------------------------------------------------
// Device matrices
// ---------------------------------------------------------------------------
static inline uint16_t f32_to_bf16_rne(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) return (uint16_t)((u >> 16) | 0x40);  // quiet NaN
    uint32_t lsb = (u >> 16) & 1u;
    u += 0x7FFFu + lsb;
    return (uint16_t)(u >> 16);
}

static inline float bf16_to_f32(uint16_t h)
{
    uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

int sptolDevMatrixCreate(sptolDevMatrix *m, sptIndex nrows, sptIndex ncols, int dtype)
{
    if (dtype != SPTOL_F32 && dtype != SPTOL_BF16)
        SPTOL_CHECK(SPTERR_VALUE_ERROR, "DevMtx New", "dtype must be SPTOL_F32 or SPTOL_BF16");
    m->nrows = nrows;
    m->ncols = ncols;
    m->stride = ncols ? ((ncols - 1) / 8 + 1) * 8 : 8;
    m->dtype = dtype;
    m->values = nullptr;
    size_t esz = dtype == SPTOL_BF16 ? 2 : 4;
    size_t rows = nrows ? nrows : 1;
    return dmalloc(&m->values, rows * m->stride * esz, "DevMtx New");
}

int sptolDevMatrixUpload(sptolDevMatrix *m, const sptMatrix *src)
{
    if (!m->values && m->nrows) SPTOL_CHECK(SPTERR_VALUE_ERROR, "DevMtx Upload", "matrix not created");
    if (src->stride != m->stride || src->nrows < m->nrows)
        SPTOL_CHECK(SPTERR_SHAPE_MISMATCH, "DevMtx Upload", "stride or rows differ");
    size_t n = (size_t)m->nrows * m->stride;
    if (m->dtype == SPTOL_F32) return sptolHipMemcpyH2D(m->values, src->values, n * 4, nullptr);
    std::vector<uint16_t> tmp(n);
    for (size_t i = 0; i < n; ++i) tmp[i] = f32_to_bf16_rne(src->values[i]);
    return sptolHipMemcpyH2D(m->values, tmp.data(), n * 2, nullptr);
}

int sptolDevMatrixDownload(sptMatrix *dst, const sptolDevMatrix *m)
{
    if (!m->values && m->nrows) SPTOL_CHECK(SPTERR_VALUE_ERROR, "DevMtx Download", "matrix not created");
    if (dst->stride != m->stride || dst->cap < m->nrows)
        SPTOL_CHECK(SPTERR_SHAPE_MISMATCH, "DevMtx Download", "stride or capacity differ");
    size_t n = (size_t)m->nrows * m->stride;
    if (m->dtype == SPTOL_F32) return sptolHipMemcpyD2H(dst->values, m->values, n * 4, nullptr);
    std::vector<uint16_t> tmp(n);
    SPTOL_TRY(sptolHipMemcpyD2H(tmp.data(), m->values, n * 2, nullptr));
    for (size_t i = 0; i < n; ++i) dst->values[i] = bf16_to_f32(tmp[i]);
    return 0;
}

void sptolDevMatrixFree(sptolDevMatrix *m)
{
    dfree(m->values);
    m->values = nullptr;
}

}  // extern "C"
