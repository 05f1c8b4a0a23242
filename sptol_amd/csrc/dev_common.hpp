// Internal helpers shared by the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <cstdint>
#include <cstddef>
#include <vector>
#include "../../include/sptol_hip.h"
#include "host_internal.h"

// Map a hipError_t to the reference's device-error slot and record it.
#define SPTOL_HIP(call, module)                                                        \
    do {                                                                               \
        hipError_t sptol_he_ = (call);                                                 \
        if (sptol_he_ != hipSuccess) {                                                 \
            sptol_complain((module), SPTERR_CUDA_ERROR + (int)sptol_he_, __FILE__,     \
                           __LINE__, hipGetErrorString(sptol_he_));                    \
            return SPTERR_CUDA_ERROR + (int)sptol_he_;                                 \
        }                                                                              \
    } while (0)

// Propagate an SptError returned by an internal call.
#define SPTOL_TRY(expr)                 \
    do {                                \
        int sptol_r_ = (expr);          \
        if (sptol_r_) return sptol_r_;  \
    } while (0)

namespace sptol {

// Runs a HIP call whose failure the caller handles or ignores, and drops the
// error that call leaves in the thread's last-error slot: the next
// hipGetLastError (after a kernel launch) would otherwise report it as the
// launch's.  An error already pending before the call is kept.
template <class F>
inline hipError_t hip_quiet(F &&f)
{
    const hipError_t pending = hipPeekAtLastError();
    const hipError_t e = f();
    if (e != hipSuccess && pending == hipSuccess) (void)hipGetLastError();
    return e;
}

// Budget-tracked device allocation (zero-size requests return nullptr).
// dmalloc may hand out a cached block (see StreamScope); dfree returns the
// block to HIP, or, inside a StreamScope, to the cache.
int dmalloc(void **ptr, size_t bytes, const char *module);
void dfree(void *ptr);

// While a StreamScope(s) is alive on a thread, the thread's dfree calls keep
// their blocks for reuse instead of returning them to HIP.  The caller
// promises that every use of a block freed in the scope is ordered on s
// before the dfree call (work enqueued on s, or on streams s was made to
// wait for, or already complete).  dmalloc calls in the scope may reuse such
// a block before that work has run, as their own work is queued behind it.
// Page-locked chunks of kPinnedChunkBytes for streaming uploads (a
// process-lifetime pool shared by all threads).
constexpr size_t kPinnedChunkBytes = (size_t)16 << 20;
int pinned_chunk(void **p);
void pinned_chunk_free(void *p);

// dst (pageable host memory) = src (device) [0, n), through two pinned chunks
// in turn: the copy out of each chunk runs on the host's OpenMP threads
// while the next chunk is in flight.  Synchronises s.
int download_parallel(void *dst, const void *src, size_t n, hipStream_t s);
// dst (device) = src (pageable host memory) [0, n), the other way round:
// each pinned chunk is filled on the host's OpenMP threads while the other
// is in flight, and each calling thread streams on its own (the runtime's
// pageable path stages through shared buffers).  Synchronises s; returns 0
// or SPTERR_CUDA_ERROR + the HIP error.
int upload_pageable(void *dst, const void *src, size_t bytes, hipStream_t s);

// Page-locked host slot of kPinnedSlotBytes (process-lifetime slabs).
constexpr size_t kPinnedSlotBytes = 1024;
int pinned_slot(void **p);
void pinned_slot_free(void *p);

// Uploads of small per-launch tables through two pinned host buffers used
// in turn: each is reused once the copy that read it has run, so the host
// can queue the next launch without waiting for the device to drain.
struct PinnedStage {
    uint8_t *buf[2] = {};
    size_t cap[2] = {};
    hipEvent_t ev[2] = {};
    bool used[2] = {};
    int cur = 0;
    // dev[0..n) = host[0..n), stream-ordered on s
    int upload(void *dev, const void *host, size_t n, hipStream_t s);
    PinnedStage() = default;
    PinnedStage(const PinnedStage &) = delete;
    PinnedStage &operator=(const PinnedStage &) = delete;
    ~PinnedStage();
};

struct StreamScope {
    explicit StreamScope(hipStream_t s);
    ~StreamScope();
    StreamScope(const StreamScope &) = delete;
    StreamScope &operator=(const StreamScope &) = delete;
private:
    hipStream_t prev_;
    int prev_on_;
};
hipStream_t stream_of(void *s);
int verbose();

constexpr int kMaxModes = 8;
// Most items one device sort, scan or select takes (their counts are int).
constexpr uint64_t kMaxSortItems = 0x7FFFFFFFull;
constexpr uint32_t kNoRow = 0xFFFFFFFFu;

// Compute units of the current device (cached).
int num_cus();

// Nonzeros sorted by one mode and packed for the segmented MTTKRP kernel
// (mttkrp_seg.hip), for 3- and 4-mode tensors.
struct SegView {
    uint32_t nmodes = 0, mode = 0, nrows = 0;
    uint32_t ndims[4] = {};        // rows of every mode's factor
    uint64_t nnz = 0;
    uint4 *rec = nullptr;          // N=3: {i0, i1, i2, bits(v)}; N=4: {i0, i1, i2, i3}
    float *vals = nullptr;         // N=4: values (owned when own_vals)
    bool own_vals = false;
    uint32_t *touched = nullptr;   // bitmap: row r of the mode holds a nonzero
    float *carry = nullptr;        // per wave: partial row carried into an earlier wave's row
    uint32_t *carry_row = nullptr; // per wave: that row, or kNoRow
    uint64_t carry_waves = 0;      // capacity in waves (64 floats each)
    uint32_t *counter = nullptr;   // work counter of the persistent kernel (0 between launches), after touched
    // Ranges of whole rows (built on demand, for one range size): range k
    // holds the nonzeros of rows [k*rr_rows, (k+1)*rr_rows), from
    // rr_start[k] to rr_start[k+1], so no row continues across ranges
    uint64_t *rr_start = nullptr;  // device, rr_nwaves + 1 offsets
    uint64_t rr_key = 0;           // the nonzeros-per-range target it was built for (0: not built)
    uint64_t rr_nwaves = 0;
    uint32_t rr_rows = 0;
    bool rr_ok = false;            // balanced enough to use
    bool ready = false;
    void release();
};

// The ALS factor step of one mode (k_factor's arguments: the Grams, the
// inverse, pivots and kind), which a 4-mode MTTKRP launch can run in one
// wave of its first block, beside its ranges.
struct AlsFactorReq {
    const double *ata;
    double *fac;
    int *piv;
    int *kind;
    uint32_t N, m, R, pad;
};

// Pack a mode-sorted SoA copy (inds[0..N-1], vals) into v (allocates).
// With perm, inds and vals are in any order and nonzero perm[z] becomes
// record z; sorted_rows[z] is then its mode index, and copy_vals must be set.
// With rec0 as well (records of the tensor in its own order, from
// seg_pack_unsorted), each record is one 16-byte gather from rec0, and the
// sorted rows may come as 16-bit keys (sorted_rows16, at most 2^16 rows)
// instead of sorted_rows.
int seg_build(SegView &v, uint32_t nmodes, uint32_t mode, uint32_t nrows, uint64_t nnz,
              uint32_t *const *inds, const float *vals, bool copy_vals, hipStream_t s, const uint32_t *ndims,
              const uint32_t *perm = nullptr, const uint32_t *sorted_rows = nullptr, const uint4 *rec0 = nullptr,
              const uint16_t *sorted_rows16 = nullptr, bool row_ranges = true);
// Records in the tensor's own order into rec0 and, for each mode m with
// keys16[m] non-null (keys16 itself may be null), the mode's indices as
// 16-bit sort keys.
int seg_pack_unsorted(uint4 *rec0, uint32_t nmodes, uint64_t nnz, uint32_t *const *inds, const float *vals,
                      hipStream_t s, uint16_t *const *keys16 = nullptr);
// Whether the segmented kernel handles this shape.
bool seg_ok(uint32_t nmodes, uint32_t stride);
// Lanes per nonzero (one float4 of columns each) for a row stride.
uint32_t seg_lpn(uint32_t stride);
// Lanes per lane group of the range kernels for factors of type dtype (4
// FP32 or 8 BF16 columns a lane), and its class 0..3 among that type's
// kernels ({2, 4, 8, 16} lanes for FP32, {1, 2, 4, 8} for BF16).
uint32_t seg_glpn(uint32_t stride, int dtype);
int seg_lane_class(uint32_t glpn, int dtype);
// Nonzeros per lane group for a tensor of nnz nonzeros and G groups per
// wave; l2_factors: the gathered factor rows total at most 4 MiB.
uint32_t seg_chg(uint64_t nnz, uint32_t G, bool l2_factors);
// out[0 .. nrows*stride) = MTTKRP of the view (no fill needed beforehand).
int seg_mttkrp(SegView &v, const void *const *mats, const uint32_t *order, uint32_t R, uint32_t stride,
               int dtype, float *out, hipStream_t s,
               const AlsFactorReq *fr = nullptr, bool *factored = nullptr);
// The same kernels with the carry space and work counter in a fixed device
// array (no allocation; at most 65536 ranges; overlapping calls on one GPU race).
int seg_mttkrp_noalloc(const SegView &v, const void *const *mats, const uint32_t *order, uint32_t R,
                       uint32_t stride, int dtype, float *out, hipStream_t s);

// Device COO tensor: structure-of-arrays copy of the host tensor, plus
// optional copies sorted by one mode.
struct SortedCopy {
    uint32_t *inds[kMaxModes] = {};
    float *vals = nullptr;
    SegView seg;
    bool ready = false;
};

}  // namespace sptol

struct sptolDevCoo_s {
    uint32_t nmodes = 0;
    uint32_t ndims[sptol::kMaxModes] = {};
    uint64_t nnz = 0;
    uint32_t *inds[sptol::kMaxModes] = {};
    float *vals = nullptr;
    sptol::SortedCopy sorted[sptol::kMaxModes];
    // mode m's indices are already non-decreasing in the tensor's own order
    // (found by the upload's check pass): its view needs no sort
    bool in_order[sptol::kMaxModes] = {};
    // views serve batched launches only (the server): no table of ranges of
    // whole rows, which only single-tensor launches use
    bool batch_only = false;
};

struct sptolHicooViews;

struct sptolDevHicoo_s {
    uint32_t nmodes = 0;
    uint32_t ndims[sptol::kMaxModes] = {};
    uint64_t nnz = 0, nb = 0, nk = 0;
    uint32_t sb_bits = 0, sk_bits = 0;
    uint64_t *bptr = nullptr;                       // nb+1
    uint32_t *binds[sptol::kMaxModes] = {};         // nb each
    uint8_t *einds[sptol::kMaxModes] = {};          // nnz each
    float *vals = nullptr;                          // nnz
    sptolHicooViews *views = nullptr;               // per-mode reordered copies (hicoo_view.hip)
};

namespace sptol {

// Launch parameters for the MTTKRP kernels (passed by value).
struct MttkrpArgs {
    uint64_t nnz;
    uint32_t nmodes;
    uint32_t mode;
    uint32_t R;            // columns computed
    uint32_t stride;       // row stride of factors and output
    uint32_t order[kMaxModes];      // order[1..nmodes-1]: Khatri-Rao order
    const uint32_t *inds[kMaxModes];
    const float *vals;
    const void *mats[kMaxModes];    // device factor pointers (dtype per launch)
    float *out;
    uint64_t per_group;    // nonzeros per lane group
    // HiCOO
    const uint64_t *bptr;
    const uint32_t *binds[kMaxModes];
    const uint8_t *einds[kMaxModes];
    uint64_t nb;
    uint32_t sb_bits;
};

// Optional per-kernel timing with hipEvents recorded on the launch stream.
// prof_start returns a token (or -1 when profiling is off); prof_stop closes it.
int prof_start(const char *name, hipStream_t s);
void prof_stop(int token, hipStream_t s);

// Allocate + copy on stream s (synchronises s).
int dup_on(void **dst, const void *src, size_t bytes, hipStream_t s, const char *module);
// Stream-aware forms of sptolDevCooCreate / sptolDevCooPrepareMode.
int coo_create(sptolDevCoo *h, const sptSparseTensor *X, hipStream_t s, bool check_vals = false);
int coo_prepare(sptolDevCoo h, uint32_t mode, hipStream_t s);
// Every mode (3- and 4-mode tensors pack their records once and gather the
// views from them).
int coo_prepare_all(sptolDevCoo h, hipStream_t s);
// Device HiCOO MTTKRP on stream s (sptolDevMTTKRPHiCOO without the checks)
// and the stream-aware form of sptolDevHicooPrepareMode.
int hicoo_mttkrp(sptolDevHicoo X, const sptolDevMatrix *mats, const uint32_t *order, uint32_t mode,
                 sptolDevMatrix *out, hipStream_t s, const AlsFactorReq *fr = nullptr, bool *factored = nullptr);
int hicoo_prepare(sptolDevHicoo h, uint32_t mode, hipStream_t s);
// Device COO tensor straight from a binary tensor file (the layout of
// sptolSaveSparseTensorBinary), streamed through pinned chunks; indices (and
// with check_vals, values) are checked on the device.  Synchronises s.
int coo_load_binary(sptolDevCoo *h, const char *path, hipStream_t s, bool check_vals = false);
// Device MTTKRP on stream s (sptolDevMTTKRP without the argument checks).
// With fr, a 4-mode launch of the segmented kernel also runs that factor
// step (*factored says whether it did).
int coo_mttkrp(sptolDevCoo X, const sptolDevMatrix *mats, const uint32_t *order, uint32_t mode,
               sptolDevMatrix *out, hipStream_t s, const AlsFactorReq *fr = nullptr, bool *factored = nullptr);

// One MTTKRP of a batch (device pointers; shared by the batch API and the server).
struct BatchDesc {
    uint64_t nnz;
    uint64_t tile_nnz;         // nonzeros per tile (multiple of 64)
    const uint32_t *inds[4];
    const float *vals;
    const void *mats[4];
    float *out;
    uint64_t out_len;          // floats in the output (rows * stride)
    uint32_t nmodes, mode, R, stride;
    uint32_t order[4];
    uint32_t lpn;              // lanes per nonzero = stride / 4
    uint32_t pad;
    SegView *seg;              // host-side: mode-sorted view for the segmented launch, or null
};

void make_desc(sptolDevCoo X, const sptolDevMatrix *mats, const uint32_t *mats_order, uint32_t mode,
               const sptolDevMatrix *out, BatchDesc *d);

// Device side of the segmented part of a batch (batch.hip): per class of
// (lanes per group, modes) the items and their range prefix; all items in
// class order with their carry-block prefix; per-batch carry space; one
// work counter per class (zero between launches).
struct SegBatchDev {
    uint8_t *arena = nullptr;
    size_t arena_cap = 0;
    float *carry = nullptr;
    uint32_t *carry_row = nullptr;
    uint64_t carry_cap = 0;              // ranges
    uint32_t *counters = nullptr;        // 8
    size_t off_items[8] = {}, off_prefix[8] = {}, off_all = 0, off_cprefix = 0;
    uint32_t n[8] = {}, n_all = 0;
    uint64_t total[8] = {}, carry_blocks = 0;
    PinnedStage stage;
    ~SegBatchDev();
};

// Reusable device buffers of a per-step batched launch.
struct BatchBuffers {
    BatchDesc *desc = nullptr;
    uint64_t *tile_prefix = nullptr, *fill_prefix = nullptr;
    uint32_t *counter = nullptr;
    size_t cap = 0;
    SegBatchDev seg;
    ~BatchBuffers();
};

// Fill + persistent batched MTTKRP of `descs` (all with factor dtype `dtype`) on s.
int batch_launch(std::vector<BatchDesc> &descs, int dtype, double tile_scale, int blocks_per_cu, BatchBuffers &bb,
                 hipStream_t s);

// HiCOO per-mode views (hicoo_view.hip).
void free_hicoo_views(sptolDevHicoo h);
int launch_hicoo_view(sptolDevHicoo X, const MttkrpArgs &a, int dtype, hipStream_t s, int *used,
                      const AlsFactorReq *fr = nullptr, bool *factored = nullptr);
// A built segmented view of X holding all its nonzeros as packed records
// (every mode's index in mode order), or nullptr.
const SegView *hicoo_seg_records(sptolDevHicoo X);

int launch_mttkrp_coo(const MttkrpArgs &a, int dtype, hipStream_t s);
int launch_mttkrp_hicoo(const MttkrpArgs &a, int dtype, hipStream_t s);
int launch_fill_zero(float *p, uint64_t n, hipStream_t s);

}  // namespace sptol
