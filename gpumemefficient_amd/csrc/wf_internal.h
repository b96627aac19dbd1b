// wf_internal.h -- host-side planner structures shared by the library's
// translation units (not part of the C ABI).
#ifndef WF_INTERNAL_H
#define WF_INTERNAL_H
#include <stdint.h>

#include <vector>

#include "../../include/wfdp.h"
#include "wf_device.h"
#include "wf_variants.h"

struct WfPlanHost {
    int64_t npairs = 0;
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

int wf_encode_pair(int problem, const int32_t *a, int64_t n1, const int32_t *b, int64_t n2,
                   uint8_t *out_a, uint8_t *out_b, int32_t *mode);
const WfVariant *wf_pick_variant(const wf_tile *tile, int64_t max_rows, int64_t total_cells, int npairs);
int wf_default_waves(int64_t max_rows, int64_t total_cells, int npairs);
int wf_default_rows_per_lane(int waves, int64_t max_rows, int64_t total_cells, int npairs);
int wf_build_plan(struct wf_ctx *ctx, const wf_job *jobs, int64_t njobs, const wf_tile *tile,
                  WfPlanHost *ph);

#endif
