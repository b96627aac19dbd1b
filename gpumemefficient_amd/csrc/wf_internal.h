// wf_internal.h -- host-side planner structures shared by the library's
// translation units (not part of the C ABI).
#ifndef WF_INTERNAL_H
#define WF_INTERNAL_H
#include <stdint.h>

#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "../../include/wfdp.h"
#include "wf_device.h"
#include "wf_variants.h"

// Grow-only device buffer.
struct DevBuf {
    void *ptr = nullptr;
    size_t cap = 0;
    // zero_fill: clear new memory (boundary rows must not hold a stale tag).
    int ensure(size_t bytes, bool zero_fill = false)
    {
        if (bytes <= cap) return WF_OK;
        if (ptr) {
            (void)hipFree(ptr);
            ptr = nullptr;
            cap = 0;
        }
        size_t want = std::max<size_t>(bytes, 256);
        want = (want + 255) & ~size_t(255);
        if (hipMalloc(&ptr, want) != hipSuccess) {
            (void)hipGetLastError();
            ptr = nullptr;
            return WF_ERR_NOMEM;
        }
        if (zero_fill && hipMemset(ptr, 0, want) != hipSuccess) {
            (void)hipFree(ptr);
            ptr = nullptr;
            return WF_ERR_DEVICE;
        }
        cap = want;
        return WF_OK;
    }
    void release()
    {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        cap = 0;
    }
};

struct WfArenas {
    DevBuf codes, pairs, tasks, bnd, res, counters, epoch, splits, colw;
    void release()
    {
        splits.release();
        colw.release();
        epoch.release();
        codes.release();
        pairs.release();
        tasks.release();
        bnd.release();
        res.release();
        counters.release();
    }
};

struct wf_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipDeviceProp_t prop{};
    double timeout_s = 10.0;
    WfArenas scratch;    // used by wf_align / wf_align_batch
    // Pinned host block of wf_align / wf_align_batch: stages a plan's codes
    // and descriptors for upload and receives the results (hipHostMalloc,
    // grown on demand, freed with the context).
    uint8_t *pinned = nullptr;
    size_t pinned_cap = 0;
    bool pinned_busy = false;  // copies from or to the block may be in flight on `stream`
};

struct WfPlanHost {
    int64_t njobs = 0;   // result slots 0..njobs-1
    int64_t npairs = 0;  // njobs plus the reverse halves of split pairs
    std::vector<WfSplitDev> splits;
    int64_t colw_entries = 0;
    const WfVariant *variant = nullptr;
    std::vector<WfPairDev> pairs;
    std::vector<WfTaskDev> tasks;
    std::vector<uint8_t> codes;
    std::vector<int64_t> task_begin, task_end, order;
    std::vector<int32_t> host_result;  // scores of pairs with an empty sequence
    std::vector<uint8_t> host_only;
    std::vector<uint64_t> cells;
    int64_t bnd_ints = 0;
    int grid = 0;
};

// True if every job can run on the bit-parallel engine (LCS or ED, at most
// 32 distinct symbols).
bool wf_bitpar_eligible(const wf_job *jobs, int64_t njobs);
int wf_encode_pair(int problem, const int32_t *a, int64_t n1, const int32_t *b, int64_t n2,
                   uint8_t *out_a, uint8_t *out_b, int32_t *mode);
// Codes and WFK_MODE_* bits of one job for a cell or bit-parallel engine:
// symbol encoding, the ED row-0 flag (job or tile flags) and, with `perm4`,
// SW's byte row tables for pairs of at most 4 symbols.
int wf_encode_job(const wf_job &j, int32_t tile_flags, bool perm4, uint8_t *a_out, uint8_t *b_out, int32_t *mode);
// Score of a pair with an empty sequence (the reference boundaries).
int32_t wf_empty_score(int problem, int64_t n1, int64_t n2, int flags);
// Byte row tables for SW on cell engines with two rows per lane (WFDP_PERM=0: off).
bool wf_perm4_enabled();
const WfVariant *wf_pick_variant(struct wf_ctx *ctx, const wf_tile *tile, const wf_job *jobs, int64_t njobs,
                                 int64_t max_rows, int64_t total_cells);
void wf_tune_key(const wf_job *jobs, int64_t njobs, int *problem, int *batch, int *bucket);
// Load the autotune table shipped with the library (once per process).
void wf_tune_autoload();
bool wf_tune_lookup(const char *arch, int cus, int problem, int batch, int bucket, int *engine, int *R, int *W,
                    int *K);
int wf_default_waves(int64_t max_rows, int64_t total_cells, int npairs);
int wf_default_rows_per_lane(int waves, int64_t max_rows, int64_t total_cells, int npairs);
// max_groups > 0: the launch runs at most that many workgroups (a server
// slot's share of the chip); an SW column split is planned only if all its
// strips fit in them.
int wf_build_plan(struct wf_ctx *ctx, const wf_job *jobs, int64_t njobs, const wf_tile *tile,
                  WfPlanHost *ph, int64_t max_groups = 0);
// Device bytes of every buffer a plan needs.
struct WfPlanBytes {
    size_t codes, pairs, tasks, bnd, res, counters, epoch, colw, splits;
};
void wf_plan_bytes(const WfPlanHost *ph, WfPlanBytes *b);
// Copy the plan's codes and descriptors into buffers that already hold
// wf_plan_bytes() (asynchronous on `stream`; the host vectors must outlive it).
int wf_copy_plan(WfPlanHost *ph, WfArenas *ar, hipStream_t stream);
int wf_upload_plan(wf_ctx *ctx, WfPlanHost *ph, WfArenas *ar);
int wf_upload_plan_async(wf_ctx *ctx, WfPlanHost *ph, WfArenas *ar, hipStream_t stream, bool sync);
int wf_max_resident_groups(wf_ctx *ctx, const WfVariant *v);
int wf_enqueue_plan(wf_ctx *ctx, WfPlanHost *ph, WfArenas *ar, int mode, int maxg, hipStream_t stream);
// Reset + engine launch(es) (or a captured graph), timed with HIP events on
// the context stream; returns after the stream has drained.
int wf_run_plan(wf_ctx *ctx, WfPlanHost *ph, WfArenas *ar, int mode, double *kernel_ms, hipGraphExec_t exec);
// wf_noalloc.hip: device address of this device's static scratch array
// (looked up and cleared on the first call per device).
int wf_noalloc_setup(int device, hipStream_t stream, uint8_t **out);
int wf_download_results(wf_ctx *ctx, WfPlanHost *ph, WfArenas *ar, double kernel_ms, wf_result *out, int64_t n);

#endif
