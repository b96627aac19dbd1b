// wf_noalloc.hip -- wf_align_noalloc(): the single-pair path of wf_align()
// with every device buffer taken from one static __device__ array per GPU,
// for callers that cannot allocate device memory (code that must not call
// hipMalloc: tight serving loops, harnesses that forbid it, work recorded
// into a graph).
//
// The array `scratch` has a fixed layout:
//   [0, 256)              launch epoch (the tag of strip-boundary words)
//   [256, kBndOff)        counters, results, pair/task/split descriptors,
//                         bit-parallel last-column words, symbol codes,
//                         carved afresh for each call
//   [kBndOff, end)        strip-boundary rows
// Boundary words are only ever written as (epoch << 32 | value) and the
// epoch only grows, so the boundary region never has to be cleared between
// calls: a stale word carries an older tag.  The whole array is cleared once
// per device, at set-up.  Every call on one GPU shares that GPU's array and
// leaves its contents behind; two calls that overlap on one GPU (two threads,
// two contexts) race and must be serialised by the caller.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cstring>
#include <mutex>

#include "wf_internal.h"

static_assert(WF_NOALLOC_SCRATCH_BYTES % 256 == 0, "scratch size");

static __device__ __align__(256) uint8_t scratch[WF_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr size_t kBndOff = (size_t)8 << 20;  // 8 MiB for everything but the boundary rows
constexpr size_t kBndBytes = (size_t)WF_NOALLOC_SCRATCH_BYTES - kBndOff;
constexpr int kMaxDevices = 64;

std::mutex g_mu;
uint8_t *g_ptr[kMaxDevices];  // device address of `scratch` on each device, once set up

__global__ void wf_noalloc_clear(uint4 *p, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = make_uint4(0u, 0u, 0u, 0u);
}

}  // namespace

// Look up this device's copy of `scratch` (the device must be current) and
// clear it, once per device.  Host code uses only the looked-up address: the
// array's name on the host is a stand-in, not a device pointer.
int wf_noalloc_setup(int device, hipStream_t stream, uint8_t **out)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (device < 0 || device >= kMaxDevices) return WF_ERR_INVALID;
    if (!g_ptr[device]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            (void)hipGetLastError();
            return WF_ERR_DEVICE;
        }
        const size_t n = WF_NOALLOC_SCRATCH_BYTES / sizeof(uint4);
        hipLaunchKernelGGL(wf_noalloc_clear, dim3(1024), dim3(256), 0, stream, (uint4 *)p, n);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(stream) != hipSuccess) return WF_ERR_DEVICE;
        g_ptr[device] = (uint8_t *)p;
    }
    if (out) *out = g_ptr[device];
    return WF_OK;
}

extern "C" int wf_align_noalloc(wf_ctx *ctx, int problem, const int32_t *a, int64_t n1, const int32_t *b,
                                int64_t n2, const wf_tile *tile, wf_result *out)
{
    if (!ctx || !out) return WF_ERR_INVALID;
    if (hipSetDevice(ctx->device) != hipSuccess) return WF_ERR_DEVICE;
    uint8_t *base = nullptr;
    int st = wf_noalloc_setup(ctx->device, ctx->stream, &base);
    if (st != WF_OK) return st;
    const wf_job j{problem, tile ? (tile->flags & WF_TILE_ED_TEXTBOOK) : 0, a, n1, b, n2};
    WfPlanHost ph;
    if ((st = wf_build_plan(ctx, &j, 1, tile, &ph)) != WF_OK) return st;
    WfPlanBytes nb;
    wf_plan_bytes(&ph, &nb);
    auto fits_scratch = [](const WfPlanBytes &b) {
        size_t small = 256;
        for (size_t x : {b.counters, b.res, b.pairs, b.tasks, b.splits, b.colw, b.codes})
            small += (std::max<size_t>(x, 256) + 255) & ~size_t(255);
        return b.bnd <= kBndBytes && small <= kBndOff;
    };
    // A column-split SW pair needs 1.5 to 2 times the boundary rows of the
    // pair: where the split does not fit and the pair does, it runs unsplit
    // (a launch of one workgroup leaves no room for a split).
    if (!fits_scratch(nb) && ph.variant->engine == WF_ENGINE_ASYNC && !ph.splits.empty()) {
        ph = WfPlanHost();
        if ((st = wf_build_plan(ctx, &j, 1, tile, &ph, 1)) != WF_OK) return st;
        wf_plan_bytes(&ph, &nb);
    }
    if (!fits_scratch(nb)) return WF_ERR_SCRATCH;
    // Carve the small buffers out of [256, kBndOff), 256-byte aligned.
    WfArenas ar;
    size_t off = 256;
    bool fits = true;
    auto carve = [&](DevBuf &d, size_t bytes) {
        bytes = (std::max<size_t>(bytes, 256) + 255) & ~size_t(255);
        if (off + bytes > kBndOff) {
            fits = false;
            return;
        }
        d.ptr = base + off;
        d.cap = bytes;
        off += bytes;
    };
    carve(ar.counters, nb.counters);
    carve(ar.res, nb.res);
    carve(ar.pairs, nb.pairs);
    carve(ar.tasks, nb.tasks);
    carve(ar.splits, nb.splits);
    carve(ar.colw, nb.colw);
    carve(ar.codes, nb.codes);
    if (!fits) return WF_ERR_SCRATCH;
    ar.epoch.ptr = base;
    ar.epoch.cap = 256;
    ar.bnd.ptr = base + kBndOff;
    ar.bnd.cap = kBndBytes;
    // The arenas point into `scratch`: nothing here may release or grow them.
    if ((st = wf_copy_plan(&ph, &ar, ctx->stream)) != WF_OK) return st;
    double ms = 0.0;
    if ((st = wf_run_plan(ctx, &ph, &ar, WF_RUN_PERSISTENT, &ms, nullptr)) != WF_OK) return st;
    return wf_download_results(ctx, &ph, &ar, ms, out, 1);
}
