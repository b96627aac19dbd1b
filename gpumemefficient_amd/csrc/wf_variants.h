// wf_variants.h -- compiled kernel variants (rows per lane, waves per
// workgroup, chunk width) and the launch entry points the host planner uses.
#ifndef WF_VARIANTS_H
#define WF_VARIANTS_H
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wf_device.h"

struct WfVariant {
    int rows_per_lane;  // R
    int waves;          // W: wave64s per workgroup (strip height = 64*W*R rows)
    int chunk;          // K: steps between workgroup barriers / boundary publishes
    const void *fn;     // __global__ wf_engine<R, W, K>
};

extern "C" const WfVariant wf_variants[];
extern "C" const int wf_num_variants;

extern "C" hipError_t wf_launch_reset(WfResultDev *res, int64_t nres, int32_t *counters, int64_t ncounters,
                                      hipStream_t stream);
extern "C" hipError_t wf_launch_engine(const WfVariant *v, int grid, const WfKernelArgs *ka,
                                       hipStream_t stream);

#endif
