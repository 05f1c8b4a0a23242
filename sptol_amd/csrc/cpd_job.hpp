// One CPD-ALS solve as a resumable job: the standalone loop and the server
// drive the same steps.  Follows src/sptensor/cpd.c:32-137.
#pragma once
#include "dev_common.hpp"
#include <chrono>
#include <vector>

namespace sptol {

constexpr int kRedBlocks = 240;   // fixed block count of the two-stage reductions
constexpr int kMaxRank = 64;

struct CpdWork {
    float *M = nullptr;            // MTTKRP output, maxI x S
    uint16_t *bf[kMaxModes] = {};  // BF16 factor copies
    double *ata = nullptr;         // N x R x R Grams
    double *partial = nullptr;     // kRedBlocks x R x R (Gram partials), then kRedBlocks x R (column partials)
    double *fac = nullptr;         // R x R Cholesky or LU factor
    int *piv = nullptr;            // R
    int *kind = nullptr;           // [0]: 0 Cholesky, 1 LU, 2 singular; [1]: singular solves so far
    double *lambda = nullptr;      // R
    double *norms = nullptr;       // R
    double *normx = nullptr;       // ||X||^2
    double *fit = nullptr;         // last fit
    ~CpdWork();
};

struct AlsStepDesc;

// Device table of a batched post-MTTKRP step (reused from step to step).
struct AlsBatchBuffers {
    AlsStepDesc *dev = nullptr;
    size_t cap = 0;               // descriptors
    PinnedStage stage;
    ~AlsBatchBuffers();
};

struct CpdJob {
    sptolDevCoo X = nullptr;              // the tensor: COO ...
    sptolDevHicoo H = nullptr;            // ... or HiCOO
    uint32_t ndims[kMaxModes] = {};
    uint64_t nnz = 0;
    const float *vals = nullptr;          // the tensor's values (device)
    uint32_t N = 0, R = 0, S = 0, maxI = 0;
    int fdt = SPTOL_F32;
    uint32_t niters = 0;
    double tol = 0.0;
    sptolDevMatrix *mats = nullptr;       // FP32 factors, owned by the caller
    CpdWork w;
    sptolDevMatrix mt[kMaxModes] = {};    // what the MTTKRP reads (FP32 or the BF16 copies)
    sptolDevMatrix Mout = {};
    uint32_t it = 0, m = 0;               // current iteration and mode
    double fit = 0.0, oldfit = 0.0;
    int iters = 0;
    int singular_solves = 0;              // mode updates whose normal equations were exactly singular
    bool done = false, fit_pending = false;
    bool exact_fit = false;               // final fit with <X,K> by an FP64 pass over the nonzeros
    bool factored = false;                // the current mode's MTTKRP launch also ran its factor step
    double *h_fit = nullptr;              // pinned slot: [0] fit, [1 .. R] final lambda,
                                          // [1 + kMaxRank] singular-solve count (an int)
    std::chrono::steady_clock::time_point t_iter;

    ~CpdJob();
    // Device bytes setup() allocates beyond the tensor and the factors.
    static size_t workspace_bytes(const uint32_t *ndims, uint32_t N, uint32_t R, int fdt);
    // Allocate, ||X||^2, initial Grams (and BF16 copies).  Every mode of X
    // must already have its sorted copy (coo_prepare).
    int setup(sptolDevCoo X, uint32_t R, uint32_t niters, double tol, sptolDevMatrix *mats, int fdt, hipStream_t s);
    // The same over a device HiCOO tensor whose every mode has its view.
    int setup_hicoo(sptolDevHicoo H, uint32_t R, uint32_t niters, double tol, sptolDevMatrix *mats, int fdt,
                    hipStream_t s);
    void order(uint32_t *o) const;        // mats_order of the current mode
    int init(uint32_t R, uint32_t niters, double tol, sptolDevMatrix *mats, int fdt, hipStream_t s);
    int mttkrp(hipStream_t s);            // the current mode's MTTKRP into Mout
    // Solve, normalise and Gram for the current mode, then advance; at the
    // end of an iteration queue the fit and its copy to the host and set
    // fit_pending (call fit_ready() after the stream synchronises).
    int after_mttkrp(hipStream_t s);
    // The tail of after_mttkrp once the step's kernels are queued (by
    // als_step_batch, whose fit kernel also stores the fit in the pinned
    // slot): advance the mode and, at the end of an iteration, set
    // fit_pending.
    void advance();
    void fit_ready();
    // End of an iteration whose fit is not read (no positive tolerance):
    // advance on the iteration count alone.
    void fit_skipped();
    // Final weights (and the exact fit when exact_fit): finish_enqueue
    // queues the kernels and the copies to the pinned slot; finish_read
    // takes the results once the stream has passed them.  finish does both.
    int finish_enqueue(hipStream_t s);
    void finish_read(sptValue *lambda_out);
    int finish(hipStream_t s, sptValue *lambda_out);
};

// The post-MTTKRP step of every job (solve, normalise, Gram and, for jobs
// at the end of an iteration, the fit) as one set of launches over all of
// them, then advance() of each.  Each job's arithmetic, block partition and
// summation order equal its after_mttkrp, so results are bitwise the same.
int als_step_batch(const std::vector<CpdJob *> &jobs, AlsBatchBuffers &bb, hipStream_t s);

}  // namespace sptol
