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
#include <vector>
#include <cstring>
#include <cstdio>
#include <string>

namespace sptol {

static std::mutex g_mem_mu;
static std::unordered_map<void *, size_t> g_live;   // device pointer -> bytes
static size_t g_in_use = 0;
static size_t g_budget = 0;                          // 0 = unlimited
static int g_fail_countdown = 0;
static std::atomic<int> g_verbose{0};

int verbose() { return g_verbose.load(); }

int dmalloc(void **ptr, size_t bytes, const char *module)
{
    *ptr = nullptr;
    if (bytes == 0) return 0;
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
    }
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) {
        std::lock_guard<std::mutex> lk(g_mem_mu);
        g_in_use -= bytes;
        *ptr = nullptr;
        sptol_complain(module, SPTERR_CUDA_ERROR + (int)e, __FILE__, __LINE__, hipGetErrorString(e));
        return SPTERR_CUDA_ERROR + (int)e;
    }
    std::lock_guard<std::mutex> lk(g_mem_mu);
    g_live[*ptr] = bytes;
    return 0;
}

void dfree(void *ptr)
{
    if (!ptr) return;
    {
        std::lock_guard<std::mutex> lk(g_mem_mu);
        auto it = g_live.find(ptr);
        if (it != g_live.end()) {
            g_in_use -= it->second;
            g_live.erase(it);
        }
    }
    (void)hipFree(ptr);
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
static std::vector<ProfRec> g_prof;
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
    if (hipEventRecord(r.a, s) != hipSuccess) return -1;
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof.push_back(r);
    return (int)g_prof.size() - 1;
}

void prof_stop(int token, hipStream_t s)
{
    if (token < 0) return;
    std::lock_guard<std::mutex> lk(g_prof_mu);
    if ((size_t)token < g_prof.size()) {
        (void)hipEventRecord(g_prof[token].b, s);
        g_prof[token].closed = true;
    }
}

static void prof_drain()
{
    std::lock_guard<std::mutex> lk(g_prof_mu);
    for (auto &r : g_prof) {
        if (r.closed) {
            float ms = 0;
            if (hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
                auto &t = g_prof_tot[r.name];
                t.ms += ms;
                t.n += 1;
            }
        }
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    g_prof.clear();
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
