// wf_stencil_api.cpp -- C ABI of the in-place table stencils (wf_stencil.hip):
// the reference's summed-area table and 2D-SOR sweeps behind
// void SOR(int n1, int n2, int *table[, int tX, int tY])
// (SAT/GPU_Hyperlane_Share.h:5, 2D-SOR/nested_loop/GPU_Hyperlane_Share.h:5 and
// the Cache headers).
#include <cstdint>
#include <cstring>

#include "wf_internal.h"

namespace {

int pick_waves(wf_ctx *ctx, const wf_tile *tile, int64_t n2, int *waves)
{
    if (tile && tile->rows_per_lane > 1) return WF_ERR_NOKERNEL;
    int w = 0;
    if (tile && tile->strip_rows > 0) {
        if (tile->strip_rows % 64) return WF_ERR_NOKERNEL;
        w = tile->strip_rows / 64;
        if (w != 1 && w != 2 && w != 4) return WF_ERR_NOKERNEL;
    } else {
        // Four waves per strip: the sweep is bound by the chain of waves down
        // the table, and a hop between waves of one strip (three 32-step
        // chunks through LDS) is shorter than one between strips (through
        // HBM), so taller strips win even with CUs left idle (2^15: 7.15 ms
        // SAT / 7.69 ms SOR against 7.22 / 7.95 ms with 128-row strips;
        // 2^14: 3.41 / 3.77 against 3.57 / 3.93).  Shorter strips only for
        // tables of fewer than four strips.
        (void)ctx;
        w = 4;
        while (w > 1 && n2 < 4 * 64 * w) w /= 2;
    }
    *waves = w;
    return WF_OK;
}

// The sweep on a device table (pitch ld, ld % 4 == 0, 16-byte aligned).
int run_device(wf_ctx *ctx, int op, int64_t n1, int64_t n2, int32_t *dt, int64_t ld, int waves, double *ms)
{
    *ms = 0.0;
    if (n1 == 0 || n2 == 0) return WF_OK;
    WfArenas *ar = &ctx->scratch;
    const int64_t strips = (n2 + 64 * waves - 1) / (64 * waves);
    int st;
    if ((st = ar->bnd.ensure((size_t)std::max<int64_t>(1, (strips - 1) * (n1 + 1)) * sizeof(uint64_t), true)) !=
        WF_OK)
        return st;
    if ((st = ar->res.ensure(sizeof(WfResultDev))) != WF_OK) return st;
    if ((st = ar->counters.ensure(2 * sizeof(int32_t))) != WF_OK) return st;
    if ((st = ar->epoch.ensure(sizeof(uint32_t), true)) != WF_OK) return st;
    WfStencilArgs a{};
    a.t = dt;
    a.ld = ld;
    a.n1 = (int32_t)n1;
    a.n2 = (int32_t)n2;
    a.nstrips = (int32_t)strips;
    a.op = op;
    a.bnd = (uint64_t *)ar->bnd.ptr;
    a.counter = (int32_t *)ar->counters.ptr;
    a.res = (WfResultDev *)ar->res.ptr;
    a.epoch = (const uint32_t *)ar->epoch.ptr;
    a.timeout_ticks = (int64_t)(ctx->timeout_s * 1.0e8);
    const int lim = wf_stencil_resident(op, waves, ctx->prop.multiProcessorCount);
    const int grid = (int)std::min<int64_t>(strips, std::max(1, lim));
    if (hipEventRecord(ctx->ev0, ctx->stream) != hipSuccess) return WF_ERR_DEVICE;
    if (wf_launch_reset(a.res, 1, a.counter, 2, (uint32_t *)ar->epoch.ptr, ctx->stream) != hipSuccess)
        return WF_ERR_DEVICE;
    if (wf_launch_stencil(op, waves, grid, &a, ctx->stream) != hipSuccess) return WF_ERR_DEVICE;
    if (hipEventRecord(ctx->ev1, ctx->stream) != hipSuccess) return WF_ERR_DEVICE;
    if (hipEventSynchronize(ctx->ev1) != hipSuccess) return WF_ERR_DEVICE;
    float f = 0.f;
    if (hipEventElapsedTime(&f, ctx->ev0, ctx->ev1) != hipSuccess) return WF_ERR_DEVICE;
    *ms = f;
    WfResultDev r{};
    if (hipMemcpy(&r, ar->res.ptr, sizeof(r), hipMemcpyDeviceToHost) != hipSuccess) return WF_ERR_DEVICE;
    return r.status ? WF_ERR_TIMEOUT : WF_OK;
}

// Host table -> device copy of pitch ldd -> sweep -> host table (or only the
// result cell with no_copyback).
int upload_run_download(wf_ctx *ctx, int op, int64_t n1, int64_t n2, int32_t *table, int64_t ld, int32_t *d,
                        int64_t ldd, int waves, bool no_copyback, wf_result *out)
{
    const int64_t rows = n2 + 2, cols = n1 + 3;
    if (hipMemcpy2D(d, (size_t)ldd * sizeof(int32_t), table, (size_t)ld * sizeof(int32_t),
                    (size_t)cols * sizeof(int32_t), (size_t)rows, hipMemcpyHostToDevice) != hipSuccess)
        return WF_ERR_DEVICE;
    double ms = 0.0;
    const int st = run_device(ctx, op, n1, n2, d, ldd, waves, &ms);
    out->kernel_ms = ms;
    if (st != WF_OK) return st;
    if (no_copyback)
        return hipMemcpy(&out->last, d + n2 * ldd + n1, sizeof(int32_t), hipMemcpyDeviceToHost) == hipSuccess
                   ? WF_OK
                   : WF_ERR_DEVICE;
    if (hipMemcpy2D(table, (size_t)ld * sizeof(int32_t), d, (size_t)ldd * sizeof(int32_t),
                    (size_t)cols * sizeof(int32_t), (size_t)rows, hipMemcpyDeviceToHost) != hipSuccess)
        return WF_ERR_DEVICE;
    out->last = table[n2 * ld + n1];
    return WF_OK;
}

bool valid_dims(int op, int64_t n1, int64_t n2, int64_t ld)
{
    // The kernel addresses a wave's 65-row window (and 8 rows of overrun)
    // with 32-bit byte offsets.
    return (op == WF_SAT || op == WF_SOR) && n1 >= 0 && n2 >= 0 && n1 + 3 <= ld && ld <= (int64_t)WF_STENCIL_MAX_PITCH &&
           n2 <= INT32_MAX - 256;
}

}  // namespace

extern "C" {

int wf_stencil_device(wf_ctx *ctx, int op, int64_t n1, int64_t n2, int32_t *dev_table, int64_t ld,
                      const wf_tile *tile, wf_result *out)
{
    if (!ctx || !out || !dev_table || !valid_dims(op, n1, n2, ld) || ld % 4 != 0 ||
        (reinterpret_cast<uintptr_t>(dev_table) & 15) != 0)
        return WF_ERR_INVALID;
    std::memset(out, 0, sizeof(*out));
    out->cells = (uint64_t)n1 * (uint64_t)n2;
    if (hipSetDevice(ctx->device) != hipSuccess) return WF_ERR_DEVICE;
    int waves = 0;
    int st = pick_waves(ctx, tile, n2, &waves);
    if (st != WF_OK) return st;
    double ms = 0.0;
    st = run_device(ctx, op, n1, n2, dev_table, ld, waves, &ms);
    out->kernel_ms = ms;
    if (st != WF_OK) return st;
    if (hipMemcpy(&out->last, dev_table + n2 * ld + n1, sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess)
        return WF_ERR_DEVICE;
    return WF_OK;
}

int wf_stencil(wf_ctx *ctx, int op, int64_t n1, int64_t n2, int32_t *table, int64_t ld, const wf_tile *tile,
               wf_result *out)
{
    if (!ctx || !out || !table || !valid_dims(op, n1, n2, ld)) return WF_ERR_INVALID;
    std::memset(out, 0, sizeof(*out));
    out->cells = (uint64_t)n1 * (uint64_t)n2;
    if (hipSetDevice(ctx->device) != hipSuccess) return WF_ERR_DEVICE;
    int waves = 0;
    int st = pick_waves(ctx, tile, n2, &waves);
    if (st != WF_OK) return st;
    const int64_t rows = n2 + 2, cols = n1 + 3;
    const int64_t ldd = (cols + 3) & ~int64_t(3);
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return WF_ERR_DEVICE;
    const size_t bytes = (size_t)(rows * ldd) * sizeof(int32_t);
    if (bytes + (64u << 20) > fr) return WF_ERR_NOMEM;
    DevBuf dt;
    if ((st = dt.ensure(bytes)) != WF_OK) return st;
    st = upload_run_download(ctx, op, n1, n2, table, ld, (int32_t *)dt.ptr, ldd, waves,
                             tile && (tile->flags & WF_TILE_NO_COPYBACK), out);
    dt.release();
    return st;
}

}  // extern "C"
