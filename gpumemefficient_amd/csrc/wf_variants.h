// wf_variants.h -- compiled kernel variants (rows per lane, waves per
// workgroup, chunk width) and the launch entry points the host planner uses.
#ifndef WF_VARIANTS_H
#define WF_VARIANTS_H
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wf_device.h"

// WF_ENGINE_BITPAR: LCS and ED only, 32 rows per lane as one bit-vector word
// (wf_bitpar.hip).
enum { WF_ENGINE_LOCKSTEP = 0, WF_ENGINE_ASYNC = 1, WF_ENGINE_BITPAR = 2, WF_ENGINE_BITPAR_STAGED = 3 };

struct WfVariant {
    int engine;         // WF_ENGINE_*
    int rows_per_lane;  // R
    int waves;          // W: compute wave64s per workgroup (strip height = 64*W*R rows)
    int chunk;          // lock-step: steps between barriers; async: steps between progress updates
    int block;          // threads per workgroup (async adds one I/O wave)
    const void *fn;     // the __global__ kernel
};

extern "C" const WfVariant wf_variants_lockstep[];
extern "C" const int wf_num_variants_lockstep;
extern "C" const WfVariant wf_variants_async[];
extern "C" const int wf_num_variants_async;
extern "C" const WfVariant wf_variants_bitpar[];
extern "C" const int wf_num_variants_bitpar;

extern "C" hipError_t wf_launch_reset(WfResultDev *res, int64_t nres, int32_t *counters, int64_t ncounters,
                                      uint32_t *epoch, hipStream_t stream);
// Lock-step <2,4,32> sweep that also writes every cell of pair 0 to ka->table.
extern "C" hipError_t wf_launch_table(int grid, const WfKernelArgs *ka, hipStream_t stream);
// Naive baselines (wf_naive.hip): method 0 one launch per cell
// anti-diagonal, 1 one launch per tile anti-diagonal; the full table t.
extern "C" hipError_t wf_launch_naive(int method, int prob, int textbook, int *t, const int32_t *a,
                                      const int32_t *b, int n1, int n2, int *smax, hipStream_t stream);
// Combine the two halves of each split pair (WfSplitDev) into res[job].
extern "C" hipError_t wf_launch_bitpar_combine(const WfSplitDev *splits, int nsplits, const uint32_t *colw,
                                               WfResultDev *res, hipStream_t stream);
extern "C" hipError_t wf_launch_engine(const WfVariant *v, int grid, const WfKernelArgs *ka,
                                       hipStream_t stream);
// Resident server engine (lock-step <2, 4, 64>): the kernel (for occupancy
// queries, with its block size) and its launch.
extern "C" const void *wf_serve_kernel(int *block);
extern "C" hipError_t wf_launch_serve(int grid, const WfServeArgs *sa, hipStream_t stream);

// Stencil sweep (wf_stencil.hip): op WFS_SAT / WFS_SOR, waves 1, 2 or 4 per
// workgroup; grid <= 0 picks one workgroup per strip up to the resident limit.
extern "C" hipError_t wf_launch_stencil(int op, int waves, int grid, const WfStencilArgs *a, hipStream_t stream);
extern "C" int wf_stencil_resident(int op, int waves, int cus);

#endif
