// Continuous-batching CPD-ALS server.
//
// The reference solves one tensor at a time on the host (src/sptensor/cpd.c).
// Here requests (tensor, rank, iteration cap, tolerance, seed) arrive at any
// time and are served concurrently on one GPU:
//   * admission threads (admit_threads, each with its own stream) take
//     queued requests in arrival order, reserve their device footprint
//     against an HBM budget (waiting while admitted work holds the budget or
//     all slots are full, rejecting requests that could never fit), upload
//     and prepare them, and hand them to the compute thread -- so uploads
//     and preparation overlap with each other and with compute;
//   * a compute thread advances the active requests by one ALS mode-step
//     per tick.  The requests form two cohorts (balanced by nonzeros), each
//     on its own compute stream, and ticks alternate between them.  Within
//     a tick the MTTKRPs of a cohort run as one persistent batched launch,
//     and their FP64 solve / normalise / Gram (and fit) steps as one launch
//     per phase.  At the end of an iteration the fit kernel stores the
//     request's fit in its pinned slot, and the request sits out ticks until
//     the event behind it has completed, so the host never blocks the
//     device to read a fit;
//   * a finished request's final kernels are queued behind an event and the
//     request goes to a completion thread, which copies its factors to host
//     memory, returns its device memory to the block cache (no device-wide
//     hipFree wait while other requests run), releases its budget and wakes
//     the admission threads and waiters.
// A request that fails (malformed input, allocation failure, budget) is
// reported with its error code and never affects the others: each request
// owns its tensor, factors and workspace, and the batched kernel writes only
// each request's own output.
#include "cpd_job.hpp"
#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>
#include <cmath>
#include <algorithm>
#include <cstring>
#include <chrono>
#include <string>
#include <cerrno>
#include <cstdio>

namespace sptol {

using clk = std::chrono::steady_clock;

enum { REQ_QUEUED = 0, REQ_PREPARING = 1, REQ_RUNNING = 2, REQ_DONE = 3, REQ_FAILED = 4 };

struct Request {
    uint64_t id = 0;
    // input (host copy, released after upload; or the caller's arrays when
    // the server borrows its inputs)
    uint32_t nmodes = 0;
    uint32_t ndims[kMaxModes] = {};
    uint64_t nnz = 0;
    std::vector<uint32_t> inds[kMaxModes];
    std::vector<float> vals;
    const uint32_t *src_inds[kMaxModes] = {};
    const float *src_vals = nullptr;
    std::string path;              // or a binary tensor file, read at admission
    uint32_t R = 0, niters = 0;
    double tol = 0.0;
    uint64_t seed = 0;
    int fdt = SPTOL_F32;
    // state
    int status = REQ_QUEUED;
    int error = 0;
    size_t reserved = 0;
    bool admitted_once = false;
    clk::time_point t_submit, t_admit, t_done;
    // device
    sptolDevCoo X = nullptr;
    sptolDevMatrix mats[kMaxModes] = {};
    CpdJob *job = nullptr;
    hipEvent_t fit_event = nullptr;
    int cohort = 0;                // compute stream that runs its iterations
    // results
    std::vector<float> lambda;
    std::vector<std::vector<float>> factors;   // per mode, ndims[m] x stride
    double fit = 0.0;
    int iterations = 0;

    // Callers hold a StreamScope of a stream ordered after every use.
    void free_device()
    {
        delete job;
        job = nullptr;
        for (auto &m : mats) { dfree(m.values); m.values = nullptr; }
        if (X) { sptolDevCooFree(X); X = nullptr; }
        if (fit_event) { (void)hipEventDestroy(fit_event); fit_event = nullptr; }
    }
};

// Compute streams (cohorts of active requests; 3 or 4 measured slower than 2)
// and the most admission workers a server runs.
constexpr int kCohorts = 2;
#ifndef SPTOL_MAX_ADMIT_THREADS
#define SPTOL_MAX_ADMIT_THREADS 4
#endif
constexpr int kMaxAdmitThreads = SPTOL_MAX_ADMIT_THREADS;

static size_t estimate_bytes(const Request &r)
{
    const uint32_t N = r.nmodes;
    const uint32_t S = ((r.R - 1) / 8 + 1) * 8;
    // a count whose byte total could wrap is more than any budget
    if (r.nnz > (SIZE_MAX >> 12)) return SIZE_MAX;
    size_t b = (size_t)r.nnz * (4 * N + 4);                    // tensor
    if (N == 3 || N == 4)
        b += (size_t)r.nnz * (N == 4 ? 20 : 16) * N;           // one packed mode-sorted view per mode
    else
        b += (size_t)r.nnz * (4 * N + 4) * N;                  // one sorted copy per mode
    for (uint32_t m = 0; m < N; ++m) b += ((size_t)r.ndims[m] + 31) / 32 * 4;   // touched-row bitmaps
    b += (size_t)r.nnz * 32 + (1u << 20);                      // packed records and sort temporaries
    for (uint32_t m = 0; m < N; ++m) b += (size_t)r.ndims[m] * S * 4;
    b += CpdJob::workspace_bytes(r.ndims, N, r.R, r.fdt);
    return b;
}

// Seeded U[0,1) factor initialisation local to a request (splitmix64): the
// request's stream of draws fills the rows of every mode in mode order, R
// draws a row, so entry (i, c < R) of mode m is draw n = off_m + i * R + c
// with state seed + (n + 1) * golden; padding columns are zero.  Each entry
// is computed from its own index, on the device.
__global__ void k_seeded_uniform(float *__restrict__ A, uint32_t I, uint32_t R, uint32_t S, uint64_t seed,
                                 uint64_t off)
{
    const uint64_t n = (uint64_t)I * S;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t i = e / S, c = e % S;
        float v = 0.0f;
        if (c < R) {
            uint64_t z = seed + (off + i * R + c + 1) * 0x9E3779B97F4A7C15ull;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            z ^= z >> 31;
            v = (float)(z >> 40) * (1.0f / 16777216.0f);
        }
        A[e] = v;
    }
}

}  // namespace sptol

using namespace sptol;

struct sptolServer_s {
    sptolServerConfig cfg;
    std::mutex mu;
    std::condition_variable cv_admit, cv_compute, cv_done, cv_release;
    std::deque<Request *> queued;
    std::vector<Request *> ready;
    std::deque<Request *> completing;   // final kernels queued behind fit_event
    std::unordered_map<uint64_t, Request *> all;
    size_t reserved = 0;
    int admitted = 0;           // preparing + ready + active
    uint64_t next_id = 1;
    bool stop = false;
    std::vector<std::thread> admit_thr;
    std::thread compute_thr, release_thr;
    bool stop_release = false;
    std::vector<hipStream_t> admit_stream;
    hipStream_t compute_stream[kCohorts] = {}, release_stream = nullptr;
    int device = 0;
    std::atomic<uint64_t> ticks{0};
};

namespace sptol {

static void finish_request(sptolServer srv, Request *r, int err)
{
    std::lock_guard<std::mutex> lk(srv->mu);
    r->status = err ? REQ_FAILED : REQ_DONE;
    r->error = err;
    r->t_done = clk::now();
    srv->reserved -= r->reserved;
    r->reserved = 0;
    srv->admitted -= 1;
    srv->cv_admit.notify_all();
    srv->cv_done.notify_all();
}

static int prepare_loaded(sptolServer srv, Request *r, hipStream_t s);

static int prepare_request(sptolServer srv, Request *r, hipStream_t s)
{
    if (!r->path.empty()) {
        SPTOL_TRY(coo_load_binary(&r->X, r->path.c_str(), s, true));
        bool same = r->X->nmodes == r->nmodes && r->X->nnz == r->nnz;
        for (uint32_t m = 0; m < r->nmodes && same; ++m) same = r->X->ndims[m] == r->ndims[m];
        if (!same) SPTOL_CHECK(SPTERR_VALUE_ERROR, "Serve", "tensor file changed after submission");
        return prepare_loaded(srv, r, s);
    }
    // host view of the copied tensor
    sptIndexVector iv[kMaxModes];
    sptIndex nd[kMaxModes], so[kMaxModes];
    for (uint32_t m = 0; m < r->nmodes; ++m) {
        iv[m].len = iv[m].cap = r->nnz;
        iv[m].data = const_cast<uint32_t *>(r->src_inds[m]);
        nd[m] = r->ndims[m];
        so[m] = m;
    }
    sptSparseTensor X;
    X.nmodes = r->nmodes;
    X.sortorder = so;
    X.ndims = nd;
    X.nnz = r->nnz;
    X.inds = iv;
    X.values.len = X.values.cap = r->nnz;
    X.values.data = const_cast<float *>(r->src_vals);
    SPTOL_TRY(coo_create(&r->X, &X, s, true));
    return prepare_loaded(srv, r, s);
}

// The request's tensor is on the device (indices and values checked):
// prepare every mode, create the factors and set up the job.
static int prepare_loaded(sptolServer srv, Request *r, hipStream_t s)
{
    // mode views need device sorts (at most 2^31 - 1 items); a larger tensor
    // runs the unsorted (atomic) batched MTTKRP
    r->X->batch_only = true;
    if (r->nnz <= kMaxSortItems) SPTOL_TRY(coo_prepare_all(r->X, s));
    uint64_t off = 0;
    for (uint32_t m = 0; m < r->nmodes; ++m) {
        SPTOL_TRY(sptolDevMatrixCreate(&r->mats[m], r->ndims[m], r->R, SPTOL_F32));
        const uint32_t S = r->mats[m].stride;
        const uint64_t n = (uint64_t)r->ndims[m] * S;
        const unsigned grid = (unsigned)std::min<uint64_t>((n + 255) / 256, 4096);
        hipLaunchKernelGGL(k_seeded_uniform, dim3(grid ? grid : 1), dim3(256), 0, s, (float *)r->mats[m].values,
                           r->ndims[m], r->R, S, r->seed, off);
        SPTOL_HIP(hipGetLastError(), "Serve init");
        off += (uint64_t)r->ndims[m] * r->R;
    }
    r->job = new CpdJob();
    r->job->exact_fit = srv->cfg.exact_fit != 0;
    SPTOL_TRY(r->job->setup(r->X, r->R, r->niters, r->tol, r->mats, r->fdt, s));
    SPTOL_HIP(hipEventCreateWithFlags(&r->fit_event, hipEventDisableTiming), "Serve");
    SPTOL_HIP(hipStreamSynchronize(s), "Serve prepare");
    return 0;
}

static void admit_loop(sptolServer srv, int worker)
{
    (void)hip_quiet([&] { return hipSetDevice(srv->device); });
    hipStream_t s = srv->admit_stream[worker];
    for (;;) {
        Request *r = nullptr;
        {
            std::unique_lock<std::mutex> lk(srv->mu);
            // the first queued request, once it fits the budget and a slot
            // (every worker waits on the same front, so arrival order holds)
            for (;;) {
                srv->cv_admit.wait(lk, [&] { return srv->stop || !srv->queued.empty(); });
                if (srv->stop) return;
                Request *f = srv->queued.front();
                const size_t need = estimate_bytes(*f);
                if (need > srv->cfg.hbm_budget) {
                    srv->queued.pop_front();
                    srv->admitted += 1;   // balanced by finish_request
                    lk.unlock();
                    sptol_complain("Serve", SPTERR_CUDA_ERROR + (int)hipErrorOutOfMemory, __FILE__, __LINE__,
                                   "request exceeds the HBM budget");
                    finish_request(srv, f, SPTERR_CUDA_ERROR + (int)hipErrorOutOfMemory);
                    lk.lock();
                    continue;
                }
                if (srv->reserved + need <= srv->cfg.hbm_budget && srv->admitted < srv->cfg.max_active) {
                    srv->queued.pop_front();
                    srv->reserved += need;
                    f->reserved = need;
                    srv->admitted += 1;
                    f->status = REQ_PREPARING;
                    f->admitted_once = true;
                    f->t_admit = clk::now();
                    r = f;
                    break;
                }
                srv->cv_admit.wait(lk);   // budget or slots free up, or the queue changes
            }
        }
        StreamScope scope(s);   // the request's arrays are used on s until it is handed over
        int err = prepare_request(srv, r, s);
        for (auto &v : r->inds) std::vector<uint32_t>().swap(v);
        std::vector<float>().swap(r->vals);
        for (auto &p : r->src_inds) p = nullptr;
        r->src_vals = nullptr;
        if (err) {
            (void)hipStreamSynchronize(s);
            r->free_device();
            finish_request(srv, r, err);
            continue;
        }
        std::lock_guard<std::mutex> lk(srv->mu);
        r->status = REQ_RUNNING;
        srv->ready.push_back(r);
        srv->cv_compute.notify_all();
        srv->cv_admit.notify_all();
    }
}

// Queue a converged request's final kernels on the compute stream and hand
// it to the completion thread.
static void complete(sptolServer srv, Request *r)
{
    hipStream_t s = srv->compute_stream[r->cohort];
    r->lambda.resize(r->R);
    int err = r->job->finish_enqueue(s);
    if (!err) {
        hipError_t e = hipEventRecord(r->fit_event, s);
        if (e != hipSuccess) err = SPTERR_CUDA_ERROR + (int)e;
    }
    if (err) {
        (void)hipStreamSynchronize(s);
        StreamScope scope(s);
        r->free_device();
        finish_request(srv, r, err);
        return;
    }
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->completing.push_back(r);
    srv->cv_release.notify_one();
}

// Completion thread: wait for a request's final kernels, copy its results
// to host memory and release its device memory.
static void release_loop(sptolServer srv)
{
    (void)hip_quiet([&] { return hipSetDevice(srv->device); });
    hipStream_t s = srv->release_stream;
    for (;;) {
        Request *r = nullptr;
        {
            std::unique_lock<std::mutex> lk(srv->mu);
            srv->cv_release.wait(lk, [&] { return srv->stop_release || !srv->completing.empty(); });
            if (srv->completing.empty()) return;
            r = srv->completing.front();
            srv->completing.pop_front();
        }
        int err = 0;
        hipError_t e = hipEventSynchronize(r->fit_event);
        if (e != hipSuccess) err = SPTERR_CUDA_ERROR + (int)e;
        if (!err) {
            r->job->finish_read(r->lambda.data());
            r->fit = r->job->fit;
            r->iterations = r->job->iters;
            r->factors.resize(r->nmodes);
            for (uint32_t m = 0; m < r->nmodes && !err; ++m) {
                r->factors[m].resize((size_t)r->ndims[m] * r->mats[m].stride);
                e = hipMemcpyAsync(r->factors[m].data(), r->mats[m].values, r->factors[m].size() * 4,
                                   hipMemcpyDeviceToHost, s);
                if (e == hipSuccess) e = hipStreamSynchronize(s);
                if (e != hipSuccess) err = SPTERR_CUDA_ERROR + (int)e;
            }
        }
        {
            // every use of the request's memory has completed (event and
            // copies waited for above)
            StreamScope scope(s);
            r->free_device();
        }
        finish_request(srv, r, err);
    }
}

// The compute thread.  Active requests are split into two cohorts, each with
// its own compute stream and batch buffers, and ticks alternate between
// them: one cohort's solve and Gram kernels and the tail of its MTTKRP
// launches overlap the other cohort's MTTKRP.  A request's arithmetic does
// not depend on which requests share its batch, so results do not depend
// on the cohorts.
static void compute_loop(sptolServer srv)
{
    (void)hip_quiet([&] { return hipSetDevice(srv->device); });
    std::vector<Request *> active;
    BatchBuffers bb[kCohorts];
    AlsBatchBuffers ab[kCohorts];
    std::vector<BatchDesc> descs[2];
    std::vector<Request *> stepped;
    std::vector<CpdJob *> jobs;
    const int nc = kCohorts;
    uint64_t load[kCohorts] = {};   // nonzeros of each cohort's active requests
    int turn = 0;
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(srv->mu);
            if (active.empty())
                srv->cv_compute.wait(lk, [&] { return srv->stop || !srv->ready.empty(); });
            if (srv->stop && srv->ready.empty() && active.empty()) return;
            for (auto *r : srv->ready) {
                int best = 0;
                for (int k = 1; k < nc; ++k)
                    if (load[k] < load[best]) best = k;
                r->cohort = best;
                load[r->cohort] += r->nnz;
                active.push_back(r);
            }
            srv->ready.clear();
        }
        // collect fits that have arrived; finish converged requests
        bool ready_in[kCohorts] = {};
        for (size_t i = 0; i < active.size();) {
            Request *r = active[i];
            if (r->job->fit_pending) {
                if (hip_quiet([&] { return hipEventQuery(r->fit_event); }) != hipSuccess) { ++i; continue; }
                r->job->fit_ready();
            }
            if (r->job->done) {
                load[r->cohort] -= r->nnz;
                complete(srv, r);
                active[i] = active.back();
                active.pop_back();
                continue;
            }
            ready_in[r->cohort] = true;
            ++i;
        }
        if (active.empty()) continue;
        if (std::none_of(ready_in, ready_in + kCohorts, [](bool b) { return b; })) {
            // every active request waits on its fit: block on the oldest
            for (auto *r : active)
                if (r->job->fit_pending) { (void)hipEventSynchronize(r->fit_event); break; }
            continue;
        }
        // one tick of one cohort: the current MTTKRP of its ready requests
        int c = turn;
        while (!ready_in[c]) c = (c + 1) % nc;
        turn = (c + 1) % nc;
        hipStream_t s = srv->compute_stream[c];
        descs[0].clear();
        descs[1].clear();
        stepped.clear();
        for (auto *r : active) {
            if (r->cohort != c || r->job->fit_pending || r->job->done) continue;
            CpdJob &j = *r->job;
            uint32_t o[kMaxModes];
            j.order(o);
            j.Mout.nrows = r->ndims[j.m];
            if (srv->cfg.batch_mttkrp) {
                BatchDesc d;
                make_desc(r->X, j.mt, o, j.m, &j.Mout, &d);
                descs[j.fdt == SPTOL_BF16 ? 1 : 0].push_back(d);
            }
            stepped.push_back(r);
        }
        int err = 0;
        if (srv->cfg.batch_mttkrp) {
            // every stepped request's MTTKRP in one persistent launch per
            // factor type, then their solve / normalise / Gram / fit steps
            // as one set of launches
            StreamScope scope(s);   // batch buffers that grow were used on s only
            for (int k = 0; k < 2 && !err; ++k)
                err = batch_launch(descs[k], k ? SPTOL_BF16 : SPTOL_F32, srv->cfg.tile_scale, srv->cfg.blocks_per_cu,
                                   bb[c], s);
            jobs.clear();
            for (auto *r : stepped) jobs.push_back(r->job);
            if (!err) err = als_step_batch(jobs, ab[c], s);
        }
        for (auto *r : stepped) {
            int e = err;
            if (!e && !srv->cfg.batch_mttkrp) e = r->job->mttkrp(s);
            if (!e && !srv->cfg.batch_mttkrp) e = r->job->after_mttkrp(s);
            if (!e && r->job->fit_pending) {
                hipError_t he = hipEventRecord(r->fit_event, s);
                if (he != hipSuccess) e = SPTERR_CUDA_ERROR + (int)he;
            }
            if (e) {
                (void)hipStreamSynchronize(s);
                r->job->done = true;
                load[r->cohort] -= r->nnz;
                StreamScope scope(s);
                r->free_device();
                finish_request(srv, r, e);
                active.erase(std::find(active.begin(), active.end(), r));
            }
        }
        srv->ticks++;
    }
}

}  // namespace sptol

extern "C" {

int sptolServerCreate(sptolServer *out, const sptolServerConfig *cfg)
{
    *out = nullptr;
    if (!cfg || cfg->hbm_budget == 0 || cfg->max_active < 1)
        SPTOL_CHECK(SPTERR_VALUE_ERROR, "Serve", "hbm_budget > 0 and max_active >= 1 required");
    // More than 4 admission workers is refused: a stream with 8 workers and
    // 96 slots failed view preparation with runtime errors and, once,
    // faulted the device, for a reason not yet found.
    if (cfg->admit_threads > kMaxAdmitThreads)
        SPTOL_CHECK(SPTERR_VALUE_ERROR, "Serve", "admit_threads must be at most 4");
    sptolServer s = new sptolServer_s();
    s->cfg = *cfg;
    if (!(s->cfg.tile_scale > 0)) s->cfg.tile_scale = 1.0;
    if (s->cfg.blocks_per_cu < 1) s->cfg.blocks_per_cu = 8;
    if (s->cfg.admit_threads < 1) s->cfg.admit_threads = kMaxAdmitThreads;
    (void)hipGetDevice(&s->device);
    hipError_t e = hipSuccess;
    for (int k = 0; k < kCohorts && e == hipSuccess; ++k)
        e = hipStreamCreateWithFlags(&s->compute_stream[k], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->release_stream, hipStreamNonBlocking);
    for (int w = 0; w < s->cfg.admit_threads && e == hipSuccess; ++w) {
        hipStream_t as = nullptr;
        e = hipStreamCreateWithFlags(&as, hipStreamNonBlocking);
        if (e == hipSuccess) s->admit_stream.push_back(as);
    }
    if (e != hipSuccess) {
        for (auto as : s->admit_stream) (void)hipStreamDestroy(as);
        for (auto cs : s->compute_stream)
            if (cs) (void)hipStreamDestroy(cs);
        if (s->release_stream) (void)hipStreamDestroy(s->release_stream);
        delete s;
        SPTOL_HIP(e, "Serve");
    }
    for (int w = 0; w < s->cfg.admit_threads; ++w) s->admit_thr.emplace_back(admit_loop, s, w);
    s->compute_thr = std::thread(compute_loop, s);
    s->release_thr = std::thread(release_loop, s);
    *out = s;
    return 0;
}

int sptolServerSubmit(sptolServer srv, const sptSparseTensor *X, sptIndex rank, sptIndex niters, double tol,
                      uint64_t seed, int factor_dtype, uint64_t *id)
{
    if (X->nmodes < 2 || X->nmodes > (uint32_t)kMaxModes) SPTOL_CHECK(SPTERR_VALUE_ERROR, "Serve", "nmodes must be in [2, 8]");
    if (rank < 1 || rank > (uint32_t)kMaxRank) SPTOL_CHECK(SPTERR_VALUE_ERROR, "Serve", "rank must be in [1, 64]");
    if (factor_dtype != SPTOL_F32 && factor_dtype != SPTOL_BF16) SPTOL_CHECK(SPTERR_VALUE_ERROR, "Serve", "bad dtype");
    Request *r = new Request();
    r->nmodes = X->nmodes;
    r->nnz = X->nnz;
    for (uint32_t m = 0; m < X->nmodes; ++m) {
        r->ndims[m] = X->ndims[m];
        if (srv->cfg.borrow_inputs) {
            r->src_inds[m] = X->inds[m].data;
        } else {
            r->inds[m].assign(X->inds[m].data, X->inds[m].data + X->nnz);
            r->src_inds[m] = r->inds[m].data();
        }
    }
    if (srv->cfg.borrow_inputs) {
        r->src_vals = X->values.data;   // non-finite values fail the request at admission
    } else {
        for (uint64_t z = 0; z < X->nnz; ++z)
            if (!std::isfinite(X->values.data[z])) {
                delete r;
                SPTOL_CHECK(SPTERR_VALUE_ERROR, "Serve", "non-finite value");
            }
        r->vals.assign(X->values.data, X->values.data + X->nnz);
        r->src_vals = r->vals.data();
    }
    r->R = rank;
    r->niters = niters;
    r->tol = tol;
    r->seed = seed;
    r->fdt = factor_dtype;
    r->t_submit = clk::now();
    std::lock_guard<std::mutex> lk(srv->mu);
    r->id = srv->next_id++;
    srv->all[r->id] = r;
    srv->queued.push_back(r);
    srv->cv_admit.notify_all();
    *id = r->id;
    return 0;
}

int sptolServerSubmitFile(sptolServer srv, const char *path, sptIndex rank, sptIndex niters, double tol,
                          uint64_t seed, int factor_dtype, uint64_t *id)
{
    if (rank < 1 || rank > (uint32_t)kMaxRank) SPTOL_CHECK(SPTERR_VALUE_ERROR, "Serve", "rank must be in [1, 64]");
    if (factor_dtype != SPTOL_F32 && factor_dtype != SPTOL_BF16) SPTOL_CHECK(SPTERR_VALUE_ERROR, "Serve", "bad dtype");
    // the header sizes the request's reservation; the arrays are read at admission
    FILE *fp = fopen(path, "rb");
    if (!fp) {
        const int en = errno ? errno : ENOENT;
        sptol_complain("Serve", SPTERR_OS_ERROR + en, __FILE__, __LINE__, strerror(en));
        return SPTERR_OS_ERROR + en;
    }
    uint32_t nm = 0, nd[8] = {};
    uint64_t nnz = 0;
    int rc = sptol_read_binary_header(fp, &nm, &nnz, nd);
    fclose(fp);
    if (rc) return rc;
    if (nm < 2) SPTOL_CHECK(SPTERR_VALUE_ERROR, "Serve", "nmodes must be in [2, 8]");
    Request *r = new Request();
    r->nmodes = nm;
    r->nnz = nnz;
    for (uint32_t m = 0; m < nm; ++m) r->ndims[m] = nd[m];
    r->path = path;
    r->R = rank;
    r->niters = niters;
    r->tol = tol;
    r->seed = seed;
    r->fdt = factor_dtype;
    r->t_submit = clk::now();
    std::lock_guard<std::mutex> lk(srv->mu);
    r->id = srv->next_id++;
    srv->all[r->id] = r;
    srv->queued.push_back(r);
    srv->cv_admit.notify_all();
    *id = r->id;
    return 0;
}

int sptolServerWait(sptolServer srv, uint64_t id, double timeout_s, sptolRequestInfo *info)
{
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->all.find(id);
    if (it == srv->all.end()) SPTOL_CHECK(SPTERR_VALUE_ERROR, "Serve", "unknown request id");
    Request *r = it->second;
    auto pred = [&] { return r->status == REQ_DONE || r->status == REQ_FAILED; };
    if (timeout_s < 0) srv->cv_done.wait(lk, pred);
    else if (!srv->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred)) {
        if (info) info->status = r->status;
        return SPTERR_NO_MORE;
    }
    if (info) {
        info->status = r->status;
        info->error = r->error;
        info->iterations = r->iterations;
        info->rank = (int)r->R;
        info->fit = r->fit;
        info->latency_s = std::chrono::duration<double>(r->t_done - r->t_submit).count();
        info->queue_s = r->admitted_once ? std::chrono::duration<double>(r->t_admit - r->t_submit).count()
                                         : info->latency_s;
    }
    return 0;
}

int sptolServerFetch(sptolServer srv, uint64_t id, sptMatrix **mats, sptValue *lambda)
{
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->all.find(id);
    if (it == srv->all.end()) SPTOL_CHECK(SPTERR_VALUE_ERROR, "Serve", "unknown request id");
    Request *r = it->second;
    if (r->status != REQ_DONE) SPTOL_CHECK(SPTERR_VALUE_ERROR, "Serve", "request not completed");
    for (uint32_t m = 0; m < r->nmodes && mats; ++m) {
        const uint32_t S = ((r->R - 1) / 8 + 1) * 8;
        if (mats[m]->stride != S || mats[m]->cap < r->ndims[m] || mats[m]->ncols != r->R)
            SPTOL_CHECK(SPTERR_SHAPE_MISMATCH, "Serve", "factor matrix shape");
        std::memcpy(mats[m]->values, r->factors[m].data(), r->factors[m].size() * 4);
        mats[m]->nrows = r->ndims[m];
    }
    if (lambda) std::memcpy(lambda, r->lambda.data(), r->R * sizeof(float));
    return 0;
}

int sptolServerRelease(sptolServer srv, uint64_t id)
{
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->all.find(id);
    if (it == srv->all.end()) return 0;
    Request *r = it->second;
    if (r->status != REQ_DONE && r->status != REQ_FAILED) SPTOL_CHECK(SPTERR_VALUE_ERROR, "Serve", "request still running");
    delete r;
    srv->all.erase(it);
    return 0;
}

int sptolServerDrain(sptolServer srv)
{
    std::unique_lock<std::mutex> lk(srv->mu);
    srv->cv_done.wait(lk, [&] {
        for (auto &kv : srv->all)
            if (kv.second->status != REQ_DONE && kv.second->status != REQ_FAILED) return false;
        return true;
    });
    return 0;
}

uint64_t sptolServerTicks(sptolServer srv) { return srv->ticks.load(); }

void sptolServerDestroy(sptolServer srv)
{
    if (!srv) return;
    sptolServerDrain(srv);
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
        srv->cv_admit.notify_all();
        srv->cv_compute.notify_all();
    }
    for (auto &t : srv->admit_thr) t.join();
    srv->compute_thr.join();
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop_release = true;
        srv->cv_release.notify_all();
    }
    srv->release_thr.join();
    (void)hipDeviceSynchronize();
    for (auto &kv : srv->all) {
        {
            StreamScope scope(srv->release_stream);   // the device has drained
            kv.second->free_device();
        }
        delete kv.second;
    }
    for (auto as : srv->admit_stream) (void)hipStreamDestroy(as);
    for (auto cs : srv->compute_stream)
        if (cs) (void)hipStreamDestroy(cs);
    (void)hipStreamDestroy(srv->release_stream);
    delete srv;
}

}  // extern "C"
