// CPU stand-in for the host API of the HIP runtime: the functions the
// library's host code calls, so that this code (the block cache, pinned
// pools, CPD jobs, batched launches, the server's admission / compute /
// release threads) can run under ThreadSanitizer and AddressSanitizer on a
// machine without a GPU.  Test infrastructure only.
//
//  * A stream is a worker thread that runs its queued operations in order;
//    the null stream is one more such worker.
//  * Kernels are no-ops that take a few microseconds (so that host threads
//    interleave with "device" work); copies and fills move real bytes.
//  * Device memory is zeroed host memory; hipFree waits for every stream
//    first, as the runtime's does.
//  * An event record is an operation on its stream; a query reports
//    hipErrorNotReady until the last record has run.
//  * Using a destroyed stream or event, or freeing memory not allocated
//    here, is reported on stderr ("HIPSTUB misuse: ...") and counted, and the
//    call fails with hipErrorInvalidHandle / hipErrorInvalidValue.
#include <hip/hip_runtime_api.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <random>
#include <set>
#include <thread>
#include <unordered_map>
#include <unordered_set>

struct ihipStream_t {
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::function<void()>> q;
    uint64_t queued = 0, done = 0;
    bool stop = false;
    std::thread worker;
    void run()
    {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            cv.wait(lk, [&] { return stop || !q.empty(); });
            if (q.empty()) return;
            auto f = std::move(q.front());
            q.pop_front();
            lk.unlock();
            f();
            lk.lock();
            ++done;
            cv.notify_all();
        }
    }
    void push(std::function<void()> f)
    {
        std::lock_guard<std::mutex> lk(mu);
        q.push_back(std::move(f));
        ++queued;
        cv.notify_all();
    }
    void drain()
    {
        std::unique_lock<std::mutex> lk(mu);
        const uint64_t target = queued;
        cv.wait(lk, [&] { return done >= target; });
    }
};

struct EventState {
    std::mutex mu;
    std::condition_variable cv;
    uint64_t recorded = 0, completed = 0;
};

struct ihipEvent_t {
    std::shared_ptr<EventState> st = std::make_shared<EventState>();
};

namespace {

std::mutex g_mu;                                  // guards the registries below
std::unordered_set<ihipStream_t *> g_streams;
std::unordered_set<ihipEvent_t *> g_events;
std::unordered_map<void *, size_t> g_dev, g_host;
std::unordered_map<const void *, void *> g_symbols;
std::atomic<int> g_misuse{0};
thread_local hipError_t t_last = hipSuccess;

struct CallConfig {
    dim3 grid, block;
    size_t shmem;
    hipStream_t stream;
};
thread_local std::vector<CallConfig> t_cfg;

ihipStream_t *null_stream()
{
    static ihipStream_t *s = [] {
        auto *n = new ihipStream_t();
        n->worker = std::thread([n] { n->run(); });
        n->worker.detach();
        return n;
    }();
    return s;
}

hipError_t fail(hipError_t e)
{
    t_last = e;
    return e;
}

void misuse(const char *what, const void *h)
{
    g_misuse++;
    fprintf(stderr, "HIPSTUB misuse: %s %p\n", what, h);
}

ihipStream_t *resolve(hipStream_t s, bool *ok)
{
    *ok = true;
    if (!s) return null_stream();
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_streams.count(s)) {
        *ok = false;
        misuse("unknown or destroyed stream", s);
        return nullptr;
    }
    return s;
}

bool live_event(hipEvent_t e)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_events.count(e)) {
        misuse("unknown or destroyed event", e);
        return false;
    }
    return true;
}

void jitter()
{
    thread_local std::minstd_rand rng(std::hash<std::thread::id>()(std::this_thread::get_id()));
    std::this_thread::sleep_for(std::chrono::microseconds(rng() % 40));
}

void drain_all()
{
    std::vector<ihipStream_t *> ss;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        ss.assign(g_streams.begin(), g_streams.end());
    }
    null_stream()->drain();
    for (auto *s : ss) s->drain();
}

}  // namespace

extern "C" int hipstub_misuse_count() { return g_misuse.load(); }

// ---- errors and devices ----------------------------------------------------
hipError_t hipGetLastError()
{
    hipError_t e = t_last;
    t_last = hipSuccess;
    return e;
}
hipError_t hipPeekAtLastError() { return t_last; }
const char *hipGetErrorString(hipError_t e)
{
    switch (e) {
    case hipSuccess: return "no error";
    case hipErrorNotReady: return "device not ready";
    case hipErrorInvalidHandle: return "invalid resource handle";
    case hipErrorInvalidValue: return "invalid argument";
    case hipErrorOutOfMemory: return "out of memory";
    case hipErrorNotSupported: return "operation not supported";
    default: return "hip stub error";
    }
}
hipError_t hipSetDevice(int d) { return d == 0 ? hipSuccess : fail(hipErrorInvalidDevice); }
hipError_t hipGetDevice(int *d) { *d = 0; return hipSuccess; }
hipError_t hipGetDeviceCount(int *n) { *n = 1; return hipSuccess; }
int hipGetStreamDeviceId(hipStream_t) { return 0; }
hipError_t hipGetDevicePropertiesR0600(hipDeviceProp_tR0600 *p, int)
{
    std::memset(p, 0, sizeof(*p));
    std::snprintf(p->name, sizeof(p->name), "hip stub");
    std::snprintf(p->gcnArchName, sizeof(p->gcnArchName), "gfx950");
    p->multiProcessorCount = 256;
    p->maxThreadsPerBlock = 1024;
    p->warpSize = 64;
    p->sharedMemPerBlock = 160 * 1024;
    p->totalGlobalMem = (size_t)288 << 30;
    p->l2CacheSize = 4 << 20;
    p->maxThreadsPerMultiProcessor = 2048;
    p->regsPerBlock = 65536;
    for (int k = 0; k < 3; ++k) { p->maxThreadsDim[k] = 1024; p->maxGridSize[k] = 0x7FFFFFFF; }
    return hipSuccess;
}
hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t a, int)
{
    switch (a) {
    case hipDeviceAttributeWarpSize: *v = 64; break;
    case hipDeviceAttributeMultiprocessorCount: *v = 256; break;
    case hipDeviceAttributeMaxThreadsPerBlock: *v = 1024; break;
    case hipDeviceAttributeMaxSharedMemoryPerBlock: *v = 160 * 1024; break;
    case hipDeviceAttributeMaxThreadsPerMultiProcessor: *v = 2048; break;
    default: *v = 1; break;
    }
    return hipSuccess;
}
hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int *n, const void *, int, size_t)
{
    *n = 2;
    return hipSuccess;
}

// ---- memory -----------------------------------------------------------------
static void *zeroed(size_t n)
{
    void *p = nullptr;
    if (posix_memalign(&p, 256, n ? n : 1) != 0) return nullptr;
    std::memset(p, 0, n ? n : 1);
    return p;
}
hipError_t hipMalloc(void **p, size_t n)
{
    *p = zeroed(n);
    if (!*p) return fail(hipErrorOutOfMemory);
    std::lock_guard<std::mutex> lk(g_mu);
    g_dev[*p] = n;
    return hipSuccess;
}
hipError_t hipFree(void *p)
{
    if (!p) return hipSuccess;
    drain_all();   // the runtime's hipFree waits for the device
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (!g_dev.erase(p)) {
            misuse("hipFree of memory not from hipMalloc", p);
            return fail(hipErrorInvalidValue);
        }
    }
    std::free(p);
    return hipSuccess;
}
hipError_t hipHostMalloc(void **p, size_t n, unsigned int)
{
    *p = zeroed(n);
    if (!*p) return fail(hipErrorOutOfMemory);
    std::lock_guard<std::mutex> lk(g_mu);
    g_host[*p] = n;
    return hipSuccess;
}
hipError_t hipHostFree(void *p)
{
    if (!p) return hipSuccess;
    drain_all();
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (!g_host.erase(p)) {
            misuse("hipHostFree of memory not from hipHostMalloc", p);
            return fail(hipErrorInvalidValue);
        }
    }
    std::free(p);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t n, hipMemcpyKind, hipStream_t s)
{
    bool ok;
    ihipStream_t *st = resolve(s, &ok);
    if (!ok) return fail(hipErrorInvalidHandle);
    if (n) st->push([dst, src, n] { std::memcpy(dst, src, n); });
    return hipSuccess;
}
hipError_t hipMemsetAsync(void *dst, int v, size_t n, hipStream_t s)
{
    bool ok;
    ihipStream_t *st = resolve(s, &ok);
    if (!ok) return fail(hipErrorInvalidHandle);
    if (n) st->push([dst, v, n] { std::memset(dst, v, n); });
    return hipSuccess;
}
hipError_t hipGetSymbolAddress(void **p, const void *sym)
{
    std::lock_guard<std::mutex> lk(g_mu);
    void *&q = g_symbols[sym];
    if (!q) q = zeroed((size_t)17 << 20);
    *p = q;
    return hipSuccess;
}

// ---- streams ------------------------------------------------------------------
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned int)
{
    auto *n = new ihipStream_t();
    n->worker = std::thread([n] { n->run(); });
    std::lock_guard<std::mutex> lk(g_mu);
    g_streams.insert(n);
    *s = n;
    return hipSuccess;
}
hipError_t hipStreamCreate(hipStream_t *s) { return hipStreamCreateWithFlags(s, 0); }
hipError_t hipStreamDestroy(hipStream_t s)
{
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (!g_streams.erase(s)) {
            misuse("destroy of unknown or destroyed stream", s);
            return fail(hipErrorInvalidHandle);
        }
    }
    s->drain();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv.notify_all();
    }
    s->worker.join();
    delete s;
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t s)
{
    bool ok;
    ihipStream_t *st = resolve(s, &ok);
    if (!ok) return fail(hipErrorInvalidHandle);
    st->drain();
    return hipSuccess;
}
hipError_t hipDeviceSynchronize()
{
    drain_all();
    return hipSuccess;
}
hipError_t hipStreamIsCapturing(hipStream_t, hipStreamCaptureStatus *c)
{
    *c = hipStreamCaptureStatusNone;
    return hipSuccess;
}
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return fail(hipErrorNotSupported); }
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *) { return fail(hipErrorNotSupported); }
hipError_t hipGraphInstantiate(hipGraphExec_t *, hipGraph_t, hipGraphNode_t *, char *, size_t)
{
    return fail(hipErrorNotSupported);
}
hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return fail(hipErrorNotSupported); }
hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }
hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }

// ---- events -------------------------------------------------------------------
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned int)
{
    auto *n = new ihipEvent_t();
    std::lock_guard<std::mutex> lk(g_mu);
    g_events.insert(n);
    *e = n;
    return hipSuccess;
}
hipError_t hipEventCreate(hipEvent_t *e) { return hipEventCreateWithFlags(e, 0); }
hipError_t hipEventDestroy(hipEvent_t e)
{
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (!g_events.erase(e)) {
            misuse("destroy of unknown or destroyed event", e);
            return fail(hipErrorInvalidHandle);
        }
    }
    delete e;   // a record still queued keeps the state alive
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s)
{
    if (!live_event(e)) return fail(hipErrorInvalidHandle);
    bool ok;
    ihipStream_t *st = resolve(s, &ok);
    if (!ok) return fail(hipErrorInvalidHandle);
    std::shared_ptr<EventState> es = e->st;
    uint64_t gen;
    {
        std::lock_guard<std::mutex> lk(es->mu);
        gen = ++es->recorded;
    }
    st->push([es, gen] {
        std::lock_guard<std::mutex> lk(es->mu);
        if (gen > es->completed) es->completed = gen;
        es->cv.notify_all();
    });
    return hipSuccess;
}
hipError_t hipEventQuery(hipEvent_t e)
{
    if (!live_event(e)) return fail(hipErrorInvalidHandle);
    std::lock_guard<std::mutex> lk(e->st->mu);
    return e->st->completed >= e->st->recorded ? hipSuccess : fail(hipErrorNotReady);
}
hipError_t hipEventSynchronize(hipEvent_t e)
{
    if (!live_event(e)) return fail(hipErrorInvalidHandle);
    std::shared_ptr<EventState> es = e->st;
    std::unique_lock<std::mutex> lk(es->mu);
    const uint64_t target = es->recorded;
    es->cv.wait(lk, [&] { return es->completed >= target; });
    return hipSuccess;
}
hipError_t hipEventElapsedTime(float *ms, hipEvent_t, hipEvent_t)
{
    *ms = 0.f;
    return hipSuccess;
}

// ---- kernels ------------------------------------------------------------------
hipError_t __hipPushCallConfiguration(dim3 g, dim3 b, size_t shm, hipStream_t s)
{
    t_cfg.push_back(CallConfig{g, b, shm, s});
    return hipSuccess;
}
hipError_t __hipPopCallConfiguration(dim3 *g, dim3 *b, size_t *shm, hipStream_t *s)
{
    if (t_cfg.empty()) return fail(hipErrorInvalidValue);
    *g = t_cfg.back().grid;
    *b = t_cfg.back().block;
    *shm = t_cfg.back().shmem;
    *s = t_cfg.back().stream;
    t_cfg.pop_back();
    return hipSuccess;
}
hipError_t hipLaunchKernel(const void *, dim3 g, dim3 b, void **, size_t, hipStream_t s)
{
    if ((uint64_t)g.x * g.y * g.z == 0 || (uint64_t)b.x * b.y * b.z == 0 || (uint64_t)b.x * b.y * b.z > 1024)
        return fail(hipErrorInvalidConfiguration);
    bool ok;
    ihipStream_t *st = resolve(s, &ok);
    if (!ok) return fail(hipErrorInvalidHandle);
    st->push([] { jitter(); });
    return hipSuccess;
}

extern "C" {
void **__hipRegisterFatBinary(const void *)
{
    static void *module = nullptr;
    return &module;
}
void __hipRegisterFunction(void **, const void *, char *, const char *, unsigned int, uint3 *, uint3 *, dim3 *,
                           dim3 *, int *)
{
}
void __hipRegisterVar(void **, void *, char *, char *, int, size_t, int, int) {}
void __hipUnregisterFatBinary(void **) {}
}
