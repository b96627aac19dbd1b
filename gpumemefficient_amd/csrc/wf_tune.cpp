// wf_tune.cpp -- autotuner for the tile choice (kernel variant).
//
// For a workload the tuner times every compiled variant (lock-step and
// asynchronous engines; rows per lane, waves per strip, chunk width) with
// the workload's own pairs resident in HBM, and keeps the fastest in a table
// keyed by (problem, workload class, size bucket, GPU arch, CU count):
//   class "single": one pair, bucket = floor(log2(n2 rows))
//   class "batch":  several pairs, bucket = floor(log2(mean cells per pair))
// The table is a small text file (one entry per line, '#' comments):
//   <arch> <cus> <problem> <class> <bucket> <engine> <rows_per_lane> <waves> <chunk> <ms>
// wf_pick_variant() consults it before the built-in defaults.
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/wfdp.h"
#include "wf_internal.h"
#include "wf_variants.h"

namespace {

struct TuneEntry {
    std::string arch;
    int cus;
    int problem;
    int batch;   // 0 single, 1 batch
    int bucket;
    int engine, rows_per_lane, waves, chunk;
    double ms;
};

std::mutex g_mu;
std::vector<TuneEntry> g_table;

int ilog2(uint64_t v)
{
    int r = -1;
    while (v) {
        v >>= 1;
        ++r;
    }
    return r < 0 ? 0 : r;
}

}  // namespace

// Workload key of a job list (problem of the first job; mixed-problem
// batches use the majority problem).
void wf_tune_key(const wf_job *jobs, int64_t njobs, int *problem, int *batch, int *bucket)
{
    int cnt[3] = {0, 0, 0};
    uint64_t cells = 0;
    int64_t rows = 0;
    for (int64_t i = 0; i < njobs; ++i) {
        if (jobs[i].problem >= 0 && jobs[i].problem <= 2) cnt[jobs[i].problem]++;
        cells += (uint64_t)(jobs[i].n1 * jobs[i].n2);
        rows = std::max<int64_t>(rows, jobs[i].n2);
    }
    *problem = (int)(std::max_element(cnt, cnt + 3) - cnt);
    *batch = njobs > 1 ? 1 : 0;
    *bucket = *batch ? ilog2(njobs ? cells / (uint64_t)njobs : 0) : ilog2((uint64_t)rows);
}

// Tuned variant for a workload key on this device, or nullptr.
bool wf_tune_lookup(const char *arch, int cus, int problem, int batch, int bucket, int *engine, int *R,
                    int *W, int *K)
{
    std::lock_guard<std::mutex> lk(g_mu);
    const TuneEntry *best = nullptr;
    for (const TuneEntry &e : g_table) {
        if (e.arch != arch || e.cus != cus || e.problem != problem || e.batch != batch) continue;
        // Exact bucket first, else the nearest one within 2 octaves.
        if (!best || std::abs(e.bucket - bucket) < std::abs(best->bucket - bucket)) best = &e;
    }
    if (!best || std::abs(best->bucket - bucket) > 2) return false;
    *engine = best->engine;
    *R = best->rows_per_lane;
    *W = best->waves;
    *K = best->chunk;
    return true;
}

static void tune_insert(const TuneEntry &n)
{
    std::lock_guard<std::mutex> lk(g_mu);
    for (TuneEntry &e : g_table) {
        if (e.arch == n.arch && e.cus == n.cus && e.problem == n.problem && e.batch == n.batch &&
            e.bucket == n.bucket) {
            e = n;
            return;
        }
    }
    g_table.push_back(n);
}

// Load the shipped table once per process: $WFDP_TUNE, else
// <directory of libwfdp.so>/../tune/wfdp_tune.txt.
void wf_tune_autoload()
{
    static std::once_flag once;
    std::call_once(once, [] {
        const char *env = std::getenv("WFDP_TUNE");
        if (env && std::strcmp(env, "0") == 0) return;
        std::string path;
        if (env && *env) {
            path = env;
        } else {
            Dl_info info{};
            if (!dladdr(reinterpret_cast<const void *>(&wf_tune_autoload), &info) || !info.dli_fname) return;
            path = info.dli_fname;
            const size_t slash = path.rfind('/');
            path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) + "/../tune/wfdp_tune.txt";
        }
        (void)wf_tune_load(path.c_str());
    });
}

extern "C" {

int wf_tune_load(const char *path)
{
    if (!path) return WF_ERR_INVALID;
    FILE *f = std::fopen(path, "r");
    if (!f) return WF_ERR_INVALID;
    char line[512];
    int n = 0;
    while (std::fgets(line, sizeof(line), f)) {
        if (line[0] == '#' || line[0] == '\n') continue;
        char arch[128], cls[16];
        TuneEntry e;
        if (std::sscanf(line, "%127s %d %d %15s %d %d %d %d %d %lf", arch, &e.cus, &e.problem, cls, &e.bucket,
                        &e.engine, &e.rows_per_lane, &e.waves, &e.chunk, &e.ms) != 10)
            continue;
        e.arch = arch;
        e.batch = std::strcmp(cls, "batch") == 0;
        tune_insert(e);
        ++n;
    }
    std::fclose(f);
    return n > 0 ? WF_OK : WF_ERR_INVALID;
}

int wf_tune_save(const char *path)
{
    if (!path) return WF_ERR_INVALID;
    FILE *f = std::fopen(path, "w");
    if (!f) return WF_ERR_INVALID;
    std::fprintf(f, "# wfdp autotune table: arch cus problem(0 LCS,1 ED,2 SW) class bucket "
                    "engine(0 lockstep,1 async,2 bitpar,3 bitpar staged) rows_per_lane waves chunk ms\n");
    std::lock_guard<std::mutex> lk(g_mu);
    for (const TuneEntry &e : g_table)
        std::fprintf(f, "%s %d %d %s %d %d %d %d %d %.4f\n", e.arch.c_str(), e.cus, e.problem,
                     e.batch ? "batch" : "single", e.bucket, e.engine, e.rows_per_lane, e.waves, e.chunk, e.ms);
    std::fclose(f);
    return WF_OK;
}

int wf_tune_clear(void)
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_table.clear();
    return WF_OK;
}

static int32_t engine_flag(int engine)
{
    return engine == WF_ENGINE_LOCKSTEP ? WF_TILE_LOCKSTEP
           : engine == WF_ENGINE_ASYNC  ? WF_TILE_ASYNC
           : engine == WF_ENGINE_BITPAR ? WF_TILE_BITPAR
                                        : WF_TILE_BITPAR_STAGED;
}

int wf_tune_lookup_tile(wf_ctx *ctx, const wf_job *jobs, int64_t njobs, wf_tile *out)
{
    if (!ctx || !jobs || njobs <= 0 || !out) return WF_ERR_INVALID;
    int p, b, bk, e, R, W, K;
    wf_tune_key(jobs, njobs, &p, &b, &bk);
    if (!wf_tune_lookup(ctx->prop.gcnArchName, ctx->prop.multiProcessorCount, p, b, bk, &e, &R, &W, &K))
        return WF_ERR_NOKERNEL;
    *out = wf_tile{K, 64 * W * R, R, engine_flag(e)};
    return WF_OK;
}

int wf_autotune(wf_ctx *ctx, const wf_job *jobs, int64_t njobs, int reps, wf_tile *best, double *best_ms)
{
    if (!ctx || !jobs || njobs <= 0 || !best) return WF_ERR_INVALID;
    if (reps < 1) reps = 1;
    wf_device_info info;
    int st = wf_device_query(ctx, &info);
    if (st != WF_OK) return st;
    const int nv = wf_variant_count();
    double bms = 1e30;
    int bi = -1;
    for (int i = 0; i < nv; ++i) {
        int32_t e, R, W, K;
        wf_variant_info(i, &e, &R, &W, &K);
        wf_tile t{K, 64 * W * R, R, engine_flag(e)};
        wf_plan *p = nullptr;
        if (wf_plan_create(ctx, jobs, njobs, &t, &p) != WF_OK) continue;
        double ms = 1e30;
        bool ok = wf_plan_run(p, WF_RUN_PERSISTENT) == WF_OK;  // warm-up
        for (int r = 0; ok && r < reps; ++r) {
            ok = wf_plan_run(p, WF_RUN_PERSISTENT) == WF_OK;
            if (ok) ms = std::min(ms, wf_plan_kernel_ms(p));
        }
        wf_plan_destroy(p);
        if (ok && ms < bms) {
            bms = ms;
            bi = i;
        }
    }
    if (bi < 0) return WF_ERR_NOKERNEL;
    int32_t e, R, W, K;
    wf_variant_info(bi, &e, &R, &W, &K);
    *best = wf_tile{K, 64 * W * R, R, engine_flag(e)};
    if (best_ms) *best_ms = bms;
    TuneEntry n;
    n.arch = info.arch;
    n.cus = info.compute_units;
    wf_tune_key(jobs, njobs, &n.problem, &n.batch, &n.bucket);
    n.engine = e;
    n.rows_per_lane = R;
    n.waves = W;
    n.chunk = K;
    n.ms = bms;
    tune_insert(n);
    return WF_OK;
}

}  // extern "C"
