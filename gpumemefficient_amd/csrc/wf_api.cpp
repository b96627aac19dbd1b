// wf_api.cpp -- host side of libwfdp.so: contexts, symbol encoding, strip
// planning, kernel selection and launches behind the C ABI in
// include/wfdp.h.
//
// The reference's host functions (LCS/GPU_Hyperlane_Share.cu:696-792,
// SmithWaterman/GPU_Hyperlane_Share.cu:683-781,
// EditDistance/GPU_Hyperlane_Share.cu:689-789) cudaMalloc and cudaMemset the
// whole (n1+3) x (n2+1) table, create multiProcessorCount streams and launch
// one kernel per strip on every call.  Here a context owns one stream and
// grow-only device arenas; a call encodes the two sequences to one-byte
// codes, uploads them, and runs one persistent launch whose workgroups claim
// strips in dependency order.  Device memory per pair is the codes plus one
// boundary row per strip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/wfdp.h"
#include "wf_device.h"
#include "wf_internal.h"
#include "wf_variants.h"

struct wf_plan {
    wf_ctx *ctx = nullptr;
    WfArenas own;
    WfPlanHost host;
    double kernel_ms = 0.0;
    bool ran = false;
    hipGraph_t graph[2] = {nullptr, nullptr};      // per run mode, when captured
    hipGraphExec_t exec[2] = {nullptr, nullptr};
};

// ----------------------------------------------------------------------------
// Encoding
// ----------------------------------------------------------------------------

namespace {

// The encoder's two scans over the symbols, also built for AVX2 and chosen
// at run time: the baseline x86-64 target has no packed 32-bit min/max, and
// scalar scans took 40 us of a 65536-symbol pair's encoding (10 us vectorised).
__attribute__((always_inline)) inline void minmax_body(const int32_t *v, int64_t n, int32_t &lo, int32_t &hi)
{
    int32_t l = lo, h = hi;
    for (int64_t i = 0; i < n; ++i) {
        l = std::min(l, v[i]);
        h = std::max(h, v[i]);
    }
    lo = l;
    hi = h;
}
__attribute__((always_inline)) inline void scale_body(const int32_t *v, int64_t n, int scale, uint8_t *out)
{
    for (int64_t i = 0; i < n; ++i) out[i] = (uint8_t)(v[i] * scale);
}
__attribute__((target("avx2"))) void minmax_avx2(const int32_t *v, int64_t n, int32_t &lo, int32_t &hi)
{
    minmax_body(v, n, lo, hi);
}
__attribute__((target("avx2"))) void scale_avx2(const int32_t *v, int64_t n, int scale, uint8_t *out)
{
    scale_body(v, n, scale, out);
}
__attribute__((always_inline)) inline uint32_t umax_body(const int32_t *v, int64_t n)
{
    uint32_t m = 0;
    for (int64_t i = 0; i < n; ++i) m = std::max(m, (uint32_t)v[i]);
    return m;
}
__attribute__((target("avx2"))) uint32_t umax_avx2(const int32_t *v, int64_t n) { return umax_body(v, n); }
__attribute__((always_inline)) inline uint8_t bmax_body(const uint8_t *v, int64_t n)
{
    uint8_t m = 0;
    for (int64_t i = 0; i < n; ++i) m = std::max(m, v[i]);
    return m;
}
__attribute__((target("avx2"))) uint8_t bmax_avx2(const uint8_t *v, int64_t n) { return bmax_body(v, n); }
__attribute__((always_inline)) inline void shr2_body(uint8_t *v, int64_t n)
{
    for (int64_t i = 0; i < n; ++i) v[i] = (uint8_t)(v[i] >> 2);
}
__attribute__((target("avx2"))) void shr2_avx2(uint8_t *v, int64_t n) { shr2_body(v, n); }
__attribute__((always_inline)) inline void reverse_body(const uint8_t *src, int64_t n, uint8_t *dst)
{
    for (int64_t k = 0; k < n; ++k) dst[k] = src[n - 1 - k];
}
__attribute__((target("avx2"))) void reverse_avx2(const uint8_t *src, int64_t n, uint8_t *dst)
{
    reverse_body(src, n, dst);
}
bool host_avx2()
{
    static const bool yes = __builtin_cpu_supports("avx2");
    return yes;
}
void minmax(const int32_t *v, int64_t n, int32_t &lo, int32_t &hi)
{
    if (host_avx2())
        minmax_avx2(v, n, lo, hi);
    else
        minmax_body(v, n, lo, hi);
}
void scale_codes(const int32_t *v, int64_t n, int scale, uint8_t *out)
{
    if (host_avx2())
        scale_avx2(v, n, scale, out);
    else
        scale_body(v, n, scale, out);
}
// The largest symbol read as unsigned (negative symbols count as large).
uint32_t umax(const int32_t *v, int64_t n)
{
    return host_avx2() ? umax_avx2(v, n) : umax_body(v, n);
}
// The largest code, and codes divided by 4 (the byte tables' plain codes).
uint8_t bmax(const uint8_t *v, int64_t n)
{
    return host_avx2() ? bmax_avx2(v, n) : bmax_body(v, n);
}
void shr2(uint8_t *v, int64_t n)
{
    if (host_avx2())
        shr2_avx2(v, n);
    else
        shr2_body(v, n);
}
// dst[k] = src[n - 1 - k] (the reverse halves of column-split pairs).
void reverse_codes(const uint8_t *src, int64_t n, uint8_t *dst)
{
    if (host_avx2())
        reverse_avx2(src, n, dst);
    else
        reverse_body(src, n, dst);
}

}  // namespace

// Maps the symbols of one pair to one-byte codes.  Equality is all the three
// recurrences look at, so any injective renaming gives identical tables.
// Row tables need codes < 8 (SW, 4-bit fields) or < 32 (LCS, ED, 1-bit
// fields); larger alphabets fall back to comparing codes.
int wf_encode_pair(int problem, const int32_t *a, int64_t n1, const int32_t *b, int64_t n2,
                   uint8_t *out_a, uint8_t *out_b, int32_t *mode)
{
    const int table_limit = (problem == WF_SW) ? 8 : 32;
    int32_t lo = INT32_MAX, hi = INT32_MIN;
    minmax(a, n1, lo, hi);
    minmax(b, n2, lo, hi);
    const int scale = (problem == WF_SW) ? 4 : 1;  // SW codes are bit offsets of 4-bit fields
    *mode = 0;
    if (n1 + n2 == 0) return WF_OK;
    // Unmasked steps left of column 0 (WFK_MODE_FREE_HEAD): always for LCS
    // and ED; for SW only with a code that matches no row -- the unused last
    // 4-bit field of the row tables, or any value above the codes in CMP mode.
    auto free_head = [&](int max_symbol, bool cmp) {
        if (problem != WF_SW) {
            *mode |= WFK_MODE_FREE_HEAD;
        } else if (cmp) {
            *mode |= WFK_MODE_FREE_HEAD | (0x1FF << WFK_MODE_NEVER_SHIFT);
        } else if (max_symbol < table_limit - 1) {
            *mode |= WFK_MODE_FREE_HEAD | (((table_limit - 1) * scale) << WFK_MODE_NEVER_SHIFT);
        }
    };
    if (lo >= 0 && hi < table_limit) {
        scale_codes(a, n1, scale, out_a);
        scale_codes(b, n2, scale, out_b);
        free_head(hi, false);
        return WF_OK;
    }
    // Dense renaming in order of first appearance.
    std::unordered_map<int32_t, int> ids;
    ids.reserve(64);
    auto id_of = [&](int32_t v) -> int {
        auto it = ids.find(v);
        if (it != ids.end()) return it->second;
        const int id = (int)ids.size();
        ids.emplace(v, id);
        return id;
    };
    std::vector<int> ia(n1), ib(n2);
    for (int64_t i = 0; i < n1; ++i) ia[i] = id_of(a[i]);
    for (int64_t i = 0; i < n2; ++i) ib[i] = id_of(b[i]);
    const int nsym = (int)ids.size();
    if (nsym > 256) return WF_ERR_ALPHABET;
    if (nsym <= table_limit) {
        for (int64_t i = 0; i < n1; ++i) out_a[i] = (uint8_t)(ia[i] * scale);
        for (int64_t i = 0; i < n2; ++i) out_b[i] = (uint8_t)(ib[i] * scale);
        free_head(nsym - 1, false);
    } else {
        *mode |= WFK_MODE_CMP;
        for (int64_t i = 0; i < n1; ++i) out_a[i] = (uint8_t)ia[i];
        for (int64_t i = 0; i < n2; ++i) out_b[i] = (uint8_t)ib[i];
        free_head(nsym - 1, true);
    }
    return WF_OK;
}

int wf_encode_job(const wf_job &j, int32_t tile_flags, bool perm4, uint8_t *a_out, uint8_t *b_out, int32_t *mode_out)
{
    int32_t mode = 0;
    const int st = wf_encode_pair(j.problem, j.a, j.n1, j.b, j.n2, a_out, b_out, &mode);
    if (st != WF_OK) return st;
    if (j.problem == WF_ED && ((j.flags | tile_flags) & WF_TILE_ED_TEXTBOOK)) mode |= WFK_MODE_TEXTBOOK;
    // SW with at most 4 symbols on a cell engine with two rows per lane:
    // byte row tables and plain codes (WFK_MODE_PERM4).
    if (perm4 && j.problem == WF_SW && !(mode & WFK_MODE_CMP)) {
        if (std::max(bmax(a_out, j.n1), bmax(b_out, j.n2)) <= 12) {
            shr2(a_out, j.n1);
            shr2(b_out, j.n2);
            mode = (mode & WFK_MODE_TEXTBOOK) | WFK_MODE_PERM4 | WFK_MODE_FREE_HEAD;
        }
    }
    *mode_out = mode;
    return WF_OK;
}

// Scores of degenerate pairs (an empty sequence), per the reference
// boundaries: LCS/SW 0; ED D(n2,0) = n2, D(0,n1) = n1-1 (n1 >= 1, ED.cpp
// :48-54) or n1 (textbook).
int32_t wf_empty_score(int problem, int64_t n1, int64_t n2, int flags)
{
    if (problem != WF_ED) return 0;
    if (n2 == 0) {
        if (n1 == 0) return 0;
        return (flags & WF_TILE_ED_TEXTBOOK) ? (int32_t)n1 : (int32_t)(n1 - 1);
    }
    return (int32_t)n2;
}

bool wf_perm4_enabled()
{
    const char *e = std::getenv("WFDP_PERM");
    return !(e && e[0] == '0');
}

// ----------------------------------------------------------------------------
// Variant choice
// ----------------------------------------------------------------------------

static int variant_count() { return wf_num_variants_async + wf_num_variants_lockstep + wf_num_variants_bitpar; }

static const WfVariant *variant_at(int i)
{
    if (i < 0 || i >= variant_count()) return nullptr;
    if (i < wf_num_variants_async) return &wf_variants_async[i];
    i -= wf_num_variants_async;
    if (i < wf_num_variants_lockstep) return &wf_variants_lockstep[i];
    return &wf_variants_bitpar[i - wf_num_variants_lockstep];
}

// The bit-parallel engine (wf_bitpar.hip) takes LCS and ED pairs whose
// symbols map to codes below 32 (its match-mask table has 32 rows).
bool wf_bitpar_eligible(const wf_job *jobs, int64_t njobs)
{
    for (int64_t i = 0; i < njobs; ++i) {
        const wf_job &j = jobs[i];
        if (j.problem != WF_LCS && j.problem != WF_ED) return false;
        if (std::max(umax(j.a, j.n1), umax(j.b, j.n2)) < 32u) continue;
        std::unordered_map<int32_t, int> seen;
        for (int64_t k = 0; k < j.n1 && seen.size() <= 32; ++k) seen.emplace(j.a[k], 0);
        for (int64_t k = 0; k < j.n2 && seen.size() <= 32; ++k) seen.emplace(j.b[k], 0);
        if (seen.size() > 32) return false;
    }
    return true;
}

static const WfVariant *find_variant(int E, int R, int W, int K)
{
    for (int i = 0; i < variant_count(); ++i) {
        const WfVariant *v = variant_at(i);
        if (v->engine == E && v->rows_per_lane == R && v->waves == W && v->chunk == K) return v;
    }
    return nullptr;
}

// Kernel variant for a plan, in order of precedence: explicit tile fields,
// the WFDP_VARIANT environment override ("engine,R,W,K"), the autotuned
// table (wf_tune.cpp) for this workload and device, built-in defaults.
// strip_rows = 64 * waves * rows_per_lane; tile_w maps to the chunk (the
// lock-step barrier interval, or the async progress interval).
const WfVariant *wf_pick_variant(wf_ctx *ctx, const wf_tile *tile, const wf_job *jobs, int64_t njobs,
                                 int64_t max_rows, int64_t total_cells)
{
    int E = WF_ENGINE_LOCKSTEP, R = 0, W = 0, K = 0;
    if (const char *env = std::getenv("WFDP_VARIANT")) {
        if (std::sscanf(env, "%d,%d,%d,%d", &E, &R, &W, &K) == 4) {
            if (const WfVariant *v = find_variant(E, R, W, K)) return v;
        }
        E = WF_ENGINE_LOCKSTEP;
        R = W = K = 0;
    }
    // A hint tile (the reference's tX, tY) only keeps its flags: the
    // variant is the planner's own choice.
    wf_tile hinted{};
    if (tile && (tile->flags & WF_TILE_HINT)) {
        hinted.flags = tile->flags & ~WF_TILE_HINT;
        tile = &hinted;
    }
    const bool cache = tile && (tile->flags & WF_TILE_CACHE);
    // The tuned table applies only when the caller chose neither a size nor
    // an engine.
    const int32_t engine_flags = WF_TILE_LOCKSTEP | WF_TILE_ASYNC | WF_TILE_BITPAR | WF_TILE_BITPAR_STAGED;
    const bool explicit_tile = tile && (tile->tile_w > 0 || tile->strip_rows > 0 || tile->rows_per_lane > 0 ||
                                        (tile->flags & engine_flags));
    if (!explicit_tile && !cache && ctx) {
        int p, b, bk, te, tr, tw, tk;
        wf_tune_key(jobs, njobs, &p, &b, &bk);
        if (wf_tune_lookup(ctx->prop.gcnArchName, ctx->prop.multiProcessorCount, p, b, bk, &te, &tr, &tw, &tk)) {
            if (const WfVariant *v = find_variant(te, tr, tw, tk)) return v;
        }
    }
    // LCS and ED pairs with at most 32 symbols go to the bit-parallel engine
    // (32 rows per lane word) unless a cell engine is asked for: its chain of
    // waves down a pair is 32 times shorter and a column step costs about
    // 20 VALU operations for 32 cells.
    const bool cell_flags = tile && (tile->flags & (WF_TILE_LOCKSTEP | WF_TILE_ASYNC | WF_TILE_CACHE));
    const int32_t bp_flags = WF_TILE_BITPAR | WF_TILE_BITPAR_STAGED;
    const bool want_bitpar = (tile && ((tile->flags & bp_flags) || tile->rows_per_lane == 32)) ||
                             (!cell_flags && !(tile && tile->rows_per_lane > 0));
    if (want_bitpar && wf_bitpar_eligible(jobs, njobs)) {
        int Wb = 0, Kb = 64;
        if (tile && tile->strip_rows > 0) Wb = std::max(1, tile->strip_rows / 2048);
        if (tile && tile->tile_w > 0) Kb = tile->tile_w;
        // The staged variant (a mask-staging wave per workgroup) when every
        // 2048-row strip gets a CU of its own; the plain one, which packs
        // more compute waves per CU, for larger launches.
        bool staged;
        if (tile && (tile->flags & WF_TILE_BITPAR_STAGED))
            staged = true;
        else if (tile && (tile->flags & WF_TILE_BITPAR))
            staged = false;
        else {
            int64_t strips = 0;
            for (int64_t i = 0; i < njobs; ++i)
                if (jobs[i].n1 > 0 && jobs[i].n2 > 0) strips += (jobs[i].n2 + 2047) / 2048;
            staged = strips <= (ctx ? ctx->prop.multiProcessorCount : 256);
        }
        // Strip height, measured on ED 65536^2: staged, one compute wave
        // (2.75 ms; W = 2 3.14); plain, two waves (3.23 ms; W = 1 3.68,
        // W = 4 3.32; K = 128 3.39).  Likewise for LCS 8192^2.
        if (Wb == 0) Wb = staged ? 1 : (max_rows > 2048 ? 2 : 1);
        const int Eb = staged ? WF_ENGINE_BITPAR_STAGED : WF_ENGINE_BITPAR;
        const WfVariant *best = nullptr;
        int best_d = 1 << 30;
        for (int i = 0; i < wf_num_variants_bitpar; ++i) {
            const WfVariant &v = wf_variants_bitpar[i];
            if (v.engine != Eb) continue;
            const int d = std::abs(v.waves - Wb) * 1000 + std::abs(v.chunk - Kb);
            if (d < best_d) {
                best_d = d;
                best = &v;
            }
        }
        return best;
    }
    if (tile && (tile->flags & bp_flags)) return nullptr;  // asked for, not possible
    // Default engine: a launch whose strips all fit on the chip at once is
    // bound by the wavefront's critical path, where the asynchronous engine's
    // shorter lag between waves and strips wins; a launch with many more
    // strips than CUs is bound by throughput, where the lock-step engine's
    // higher occupancy wins (measured on MI355X: SW 131072^2 and ED 65536^2
    // against the 4096-pair batch).
    // The async engine's counter group G trades per-step cost (a check, a
    // branch that ends the compiler's scheduling region, per group) against
    // the lag a wave adds to the chain (64 + G steps): G = 32 measured best
    // on both single pairs (SW 131072^2: G = 8 / 16 / 32 8.43 / 7.88 /
    // 7.80 ms; ED 65536^2 on this engine: 3.77 / 3.39 / 3.26 ms).
    const int async_k = 32;
    {
        int64_t strips = 0;
        for (int64_t i = 0; i < njobs; ++i)
            if (jobs[i].n1 > 0 && jobs[i].n2 > 0) strips += (jobs[i].n2 + 511) / 512;
        const int cus = ctx ? ctx->prop.multiProcessorCount : 256;
        if (strips <= 2LL * cus) E = WF_ENGINE_ASYNC;
    }
    if (tile && (tile->flags & WF_TILE_ASYNC)) E = WF_ENGINE_ASYNC;
    if (tile && (tile->flags & WF_TILE_LOCKSTEP)) E = WF_ENGINE_LOCKSTEP;
    if (cache) E = WF_ENGINE_ASYNC;
    if (tile && tile->rows_per_lane > 0) R = tile->rows_per_lane;
    if (tile && tile->tile_w > 0) K = tile->tile_w;
    if (tile && tile->strip_rows > 0) {
        const int rows = tile->strip_rows;
        if (cache) W = 1;
        if (R == 0 && W == 0) {
            // Prefer 4 waves (one per SIMD), then fewer rows per lane.
            if (rows >= 256) {
                W = 4;
                R = std::max(1, std::min(4, rows / 256));
            } else if (rows >= 128) {
                W = 2;
                R = 1;
            } else {
                W = 1;
                R = 1;
            }
        } else if (W == 0) {
            W = std::max(1, rows / (64 * R));
        } else if (R == 0) {
            R = std::max(1, rows / (64 * W));
        }
    }
    if (W == 0) W = cache ? 1 : wf_default_waves(max_rows, total_cells, (int)njobs);
    if (R == 0) R = wf_default_rows_per_lane(W, max_rows, total_cells, (int)njobs);
    if (K == 0) K = (E == WF_ENGINE_LOCKSTEP) ? 64 : async_k;
    // Snap to the nearest compiled variant of the requested engine.
    const WfVariant *best = nullptr;
    int best_d = 1 << 30;
    for (int i = 0; i < variant_count(); ++i) {
        const WfVariant &v = *variant_at(i);
        if (v.engine != E) continue;
        const int d = std::abs(v.waves - W) * 1000 + std::abs(v.rows_per_lane - R) * 10 +
                      std::abs(v.chunk - K) / 8;
        if (d < best_d) {
            best_d = d;
            best = &v;
        }
    }
    return best;
}

// Built-in defaults (measured on MI355X: SW 2^17 single pair, ED 2^16, the
// 4096-pair batch): strips of 512 rows = 4 waves x 2 rows per lane; smaller
// strips for pairs too short to fill four waves.
int wf_default_waves(int64_t max_rows, int64_t total_cells, int npairs)
{
    (void)total_cells;
    (void)npairs;
    if (max_rows >= 512) return 4;
    if (max_rows >= 256) return 2;
    return 1;
}

int wf_default_rows_per_lane(int waves, int64_t max_rows, int64_t total_cells, int npairs)
{
    (void)total_cells;
    (void)npairs;
    if (max_rows >= 64LL * waves * 2) return 2;
    return 1;
}

// ----------------------------------------------------------------------------
// Planning
// ----------------------------------------------------------------------------

// Column split of bit-parallel pairs (WfSplitDev): on unless WFDP_SPLIT=0.
static bool split_enabled()
{
    const char *e = std::getenv("WFDP_SPLIT");
    return !(e && e[0] == '0');
}
// Byte row tables for SW pairs of at most 4 symbols (WFK_MODE_PERM4): on
// unless WFDP_PERM=0 (wf_perm4_enabled).
// Pairs narrower than this are not split.
constexpr int64_t kSplitMinCols = 1024;
// SW pairs narrower than this are not split (measured gains of the split:
// 13% at 1024 x 1024, 17% at 2048 x 2048, 30% at 3000 x 600).
constexpr int64_t kSwSplitMinCols = 1024;

int wf_build_plan(wf_ctx *ctx, const wf_job *jobs, int64_t njobs, const wf_tile *tile,
                  WfPlanHost *ph, int64_t max_groups)
{
    int64_t max_rows = 0, total_cells = 0;
    for (int64_t i = 0; i < njobs; ++i) {
        const wf_job &j = jobs[i];
        if (j.problem < WF_LCS || j.problem > WF_SW) return WF_ERR_INVALID;
        if (j.n1 < 0 || j.n2 < 0 || j.n1 > (1LL << 30) || j.n2 > (1LL << 30)) return WF_ERR_INVALID;
        if ((j.n1 > 0 && !j.a) || (j.n2 > 0 && !j.b)) return WF_ERR_INVALID;
        if (j.n1 > 0 && j.n2 > 0) max_rows = std::max(max_rows, j.n2);
        total_cells += j.n1 * j.n2;
    }
    const WfVariant *v = wf_pick_variant(ctx, tile, jobs, njobs, max_rows, total_cells);
    if (!v) return WF_ERR_NOKERNEL;
    ph->variant = v;
    const int64_t strip = 64LL * v->waves * v->rows_per_lane;
    const bool bitpar = v->engine == WF_ENGINE_BITPAR || v->engine == WF_ENGINE_BITPAR_STAGED;
    const bool perm4 = (v->engine == WF_ENGINE_ASYNC || v->engine == WF_ENGINE_LOCKSTEP) && v->rows_per_lane == 2 &&
                       wf_perm4_enabled();

    // Bit-parallel pairs are split into two column halves computed at once
    // (WfSplitDev) when the launch has room for twice its strips: a single
    // pair's time is its column count times the step time, so the halves
    // finish in about half of it.
    std::vector<int64_t> mid(njobs, 0);
    std::vector<uint8_t> swsplit(njobs, 0);
    std::vector<std::vector<uint8_t>> trial;  // SW split candidates' codes (a then b) and modes
    std::vector<int32_t> trial_mode;
    int64_t nsplit = 0, nextra = 0;
    if (bitpar && split_enabled()) {
        int64_t strips = 0;
        for (int64_t i = 0; i < njobs; ++i)
            if (jobs[i].n1 > 0 && jobs[i].n2 > 0) strips += (jobs[i].n2 + strip - 1) / strip;
        const int64_t cus = ctx ? ctx->prop.multiProcessorCount : 256;
        if (2 * strips <= cus)
            for (int64_t i = 0; i < njobs; ++i)
                if (jobs[i].n1 >= kSplitMinCols && jobs[i].n2 > 0) {
                    mid[i] = jobs[i].n1 / 2;
                    ++nsplit;
                }
        nextra = nsplit;
    } else if (v->engine == WF_ENGINE_ASYNC && perm4 && split_enabled()) {
        // SW pairs on the byte tables split in three (bottom-right cell
        // only) or four (with the running maximum) pairs of half the columns
        // (WfSplitDev) when all of the launch's strips fit on the chip at
        // once: the halves' strip chains run side by side.
        int64_t strips = 0, extra = 0;
        std::vector<uint8_t> cand(njobs, 0);
        trial.resize(njobs);
        trial_mode.assign(njobs, 0);
        for (int64_t i = 0; i < njobs; ++i) {
            const wf_job &j = jobs[i];
            if (j.n1 <= 0 || j.n2 <= 0) continue;
            const int64_t si = (j.n2 + strip - 1) / strip;
            strips += si;
            if (j.problem != WF_SW || j.n1 < kSwSplitMinCols) continue;
            const bool nomax = ((j.flags | (tile ? tile->flags : 0)) & WF_TILE_NO_MAX) != 0;
            // The trial encoding is kept for the plan's codes below.
            std::vector<uint8_t> t((size_t)(j.n1 + j.n2));
            int32_t mode = 0;
            if (wf_encode_job(j, tile ? tile->flags : 0, perm4, t.data(), t.data() + j.n1, &mode) != WF_OK ||
                !(mode & WFK_MODE_PERM4))
                continue;
            trial[i] = std::move(t);
            trial_mode[i] = mode;
            cand[i] = nomax ? 2 : 3;  // extra pairs
            extra += cand[i] * si;
        }
        // A split pays while every strip has a CU of its own.  Beyond that
        // only three pairs of a launch of one pair still do, whose strips
        // start staggered along one long chain (SW 65536², 384 strips on
        // 256 CUs: 3.18 -> 3.03 ms); four pairs do not (with the maximum,
        // 512 strips: 3.62 -> 4.27 ms), nor a batch, whose strips all run at
        // once (16 pairs of 4096², 384 strips: 0.27 ms against 0.21).
        const bool four = std::find(cand.begin(), cand.end(), 3) != cand.end();
        const int64_t cus = ctx ? ctx->prop.multiProcessorCount : 256;
        const int64_t room = std::min({ctx ? (int64_t)wf_max_resident_groups(ctx, v) : 256,
                                       (four || njobs > 1) ? cus : INT64_MAX,
                                       max_groups > 0 ? max_groups : INT64_MAX});
        if (extra > 0 && strips + extra <= room)
            for (int64_t i = 0; i < njobs; ++i)
                if (cand[i]) {
                    mid[i] = jobs[i].n1 / 2;
                    swsplit[i] = cand[i];
                    ++nsplit;
                    nextra += cand[i];
                }
    }
    const int64_t np = njobs + nextra;
    ph->njobs = njobs;
    ph->npairs = np;
    ph->pairs.assign(np, WfPairDev{});
    ph->host_result.assign(np, 0);
    ph->host_only.assign(np, 0);
    ph->cells.assign(np, 0);
    ph->splits.clear();
    ph->colw_entries = 0;
    for (int64_t i = 0; i < njobs; ++i) ph->cells[i] = (uint64_t)(jobs[i].n1 * jobs[i].n2);

    // Codes: a then b per pair, 16-byte aligned; a split pair's reverse half
    // gets its own reversed copies after all jobs.
    int64_t code_bytes = 0, bnd_ints = 0;
    auto place = [&](WfPairDev &p, int prob, int64_t n1, int64_t n2, int64_t a_bytes) {
        p.prob = prob;
        p.n1 = (int32_t)n1;
        p.n2 = (int32_t)n2;
        p.a_off = code_bytes;
        code_bytes += (a_bytes + 15) & ~15LL;
        p.b_off = code_bytes;
        code_bytes += (n2 + 15) & ~15LL;
        p.ngroups = (int32_t)((n2 + strip - 1) / strip);
        p.bnd_off = bnd_ints;
        bnd_ints += (int64_t)(p.ngroups - 1) * n1;
    };
    for (int64_t i = 0; i < njobs; ++i) {
        const wf_job &j = jobs[i];
        WfPairDev &p = ph->pairs[i];
        if (j.n1 == 0 || j.n2 == 0) {
            p.prob = j.problem;
            p.n1 = (int32_t)j.n1;
            p.n2 = (int32_t)j.n2;
            p.a_off = code_bytes;
            code_bytes += (j.n1 + 15) & ~15LL;
            p.b_off = code_bytes;
            code_bytes += (j.n2 + 15) & ~15LL;
            ph->host_only[i] = 1;
            ph->host_result[i] = wf_empty_score(j.problem, j.n1, j.n2, j.flags);
            p.ngroups = 0;
            continue;
        }
        place(p, j.problem, mid[i] ? mid[i] : j.n1, j.n2, j.n1);
    }
    {
        int64_t r = njobs;
        for (int64_t i = 0; i < njobs; ++i) {
            if (!mid[i]) continue;
            const wf_job &j = jobs[i];
            if (swsplit[i]) {
                // Right half from a zero boundary (pair r), reversed right
                // half (pair r + 1) and, with the running maximum, the
                // clamped reversed right half (pair r + 2); the left and
                // the reversed halves keep their last column.
                const int64_t nr = j.n1 - mid[i];
                for (int k = 0; k < swsplit[i]; ++k) place(ph->pairs[r + k], j.problem, nr, j.n2, nr);
                WfSplitDev sp{};
                sp.job = (int32_t)i;
                sp.rev = (int32_t)(r + 1);
                sp.prob = WFK_SW;
                sp.n2 = (int32_t)j.n2;
                sp.f0 = (int32_t)r;
                sp.r0 = (int32_t)nr;
                const int64_t entries = (j.n2 + 1) / 2;  // two words per entry, one per row
                ph->pairs[i].col_off = ph->colw_entries;
                sp.fwd_off = ph->colw_entries;
                ph->colw_entries += entries;
                ph->pairs[r + 1].col_off = ph->colw_entries;
                sp.rev_off = ph->colw_entries;
                ph->colw_entries += entries;
                if (swsplit[i] == 3) {
                    sp.crev = (int32_t)(r + 2);
                    ph->pairs[r + 2].col_off = ph->colw_entries;
                    sp.crev_off = ph->colw_entries;
                    ph->colw_entries += entries;
                }
                ph->splits.push_back(sp);
                r += swsplit[i];
                continue;
            }
            place(ph->pairs[r], j.problem, j.n1 - mid[i], j.n2, j.n1 - mid[i]);
            WfSplitDev sp{};
            sp.job = (int32_t)i;
            sp.rev = (int32_t)r;
            sp.prob = j.problem;
            sp.n2 = (int32_t)j.n2;
            ph->pairs[i].col_off = ph->colw_entries;
            sp.fwd_off = ph->colw_entries;
            ph->colw_entries += (int64_t)ph->pairs[i].ngroups * v->waves * 64;
            ph->pairs[r].col_off = ph->colw_entries;
            sp.rev_off = ph->colw_entries;
            ph->colw_entries += (int64_t)ph->pairs[r].ngroups * v->waves * 64;
            ph->splits.push_back(sp);
            ++r;
        }
    }
    ph->codes.assign((size_t)std::max<int64_t>(code_bytes, 16), 0);
    for (int64_t i = 0; i < njobs; ++i) {
        const wf_job &j = jobs[i];
        WfPairDev &p = ph->pairs[i];
        int32_t mode = 0;
        if (i < (int64_t)trial.size() && !trial[i].empty()) {
            std::memcpy(ph->codes.data() + p.a_off, trial[i].data(), (size_t)j.n1);
            std::memcpy(ph->codes.data() + p.b_off, trial[i].data() + j.n1, (size_t)j.n2);
            mode = trial_mode[i];
        } else {
            const int st = wf_encode_job(j, tile ? tile->flags : 0, perm4, ph->codes.data() + p.a_off,
                                         ph->codes.data() + p.b_off, &mode);
            if (st != WF_OK) return st;
        }
        // SW pairs on the row tables may skip the running maximum (the cell
        // engines have instances without it).
        if (j.problem == WF_SW && !(mode & WFK_MODE_CMP) && ((j.flags | (tile ? tile->flags : 0)) & WF_TILE_NO_MAX))
            mode |= WFK_MODE_NOMAX;
        p.mode = mode;
        // The bit-parallel kernel indexes a 32-row table with the codes.
        if (bitpar && (j.problem == WF_SW || (mode & WFK_MODE_CMP))) return WF_ERR_NOKERNEL;
    }
    for (WfSplitDev &sp : ph->splits) {
        if (sp.prob == WFK_SW) {
            // Right half: a[mid..n1) and b as they are, the pair's modes;
            // reverse half: both reversed, the unclamped recurrence.
            const wf_job &j = jobs[sp.job];
            const WfPairDev &f = ph->pairs[sp.job];
            WfPairDev &rp = ph->pairs[sp.f0];
            WfPairDev &rv = ph->pairs[sp.rev];
            const int64_t m = mid[sp.job], nr = sp.r0;
            const uint8_t *fa = ph->codes.data() + f.a_off;
            const uint8_t *fb = ph->codes.data() + f.b_off;
            std::memcpy(ph->codes.data() + rp.a_off, fa + m, (size_t)nr);
            std::memcpy(ph->codes.data() + rp.b_off, fb, (size_t)j.n2);
            uint8_t *ra = ph->codes.data() + rv.a_off;
            uint8_t *rb = ph->codes.data() + rv.b_off;
            reverse_codes(fa + m, nr, ra);
            reverse_codes(fb, j.n2, rb);
            rp.mode = f.mode;
            rv.mode = (f.mode & (WFK_MODE_PERM4 | WFK_MODE_FREE_HEAD)) | WFK_MODE_NOMAX | WFK_MODE_SWE;
            if (!(f.mode & WFK_MODE_PERM4) || !(f.mode & WFK_MODE_NOMAX) != (sp.crev >= 0)) return WF_ERR_INVALID;
            if (sp.crev >= 0) {
                WfPairDev &cv = ph->pairs[sp.crev];
                std::memcpy(ph->codes.data() + cv.a_off, ra, (size_t)nr);
                std::memcpy(ph->codes.data() + cv.b_off, rb, (size_t)j.n2);
                cv.mode = (f.mode & (WFK_MODE_PERM4 | WFK_MODE_FREE_HEAD)) | WFK_MODE_NOMAX;
            }
            continue;
        }
        // Reverse half: a[mid..n1) and b reversed, textbook ED boundary (its
        // row 0 is a plain distance from the empty prefix).
        const wf_job &j = jobs[sp.job];
        const WfPairDev &f = ph->pairs[sp.job];
        WfPairDev &r = ph->pairs[sp.rev];
        const int64_t m = mid[sp.job], nr = j.n1 - m;
        const uint8_t *fa = ph->codes.data() + f.a_off;
        const uint8_t *fb = ph->codes.data() + f.b_off;
        uint8_t *ra = ph->codes.data() + r.a_off;
        uint8_t *rb = ph->codes.data() + r.b_off;
        reverse_codes(fa + m, nr, ra);
        reverse_codes(fb, j.n2, rb);
        r.mode = (f.mode & ~WFK_MODE_TEXTBOOK) | (j.problem == WF_ED ? WFK_MODE_TEXTBOOK : 0);
        if (j.problem == WF_ED) {
            sp.f0 = (int32_t)((f.mode & WFK_MODE_TEXTBOOK) ? m : m - 1);  // row 0 at column mid
            sp.r0 = (int32_t)nr;
        }
    }

    // Tasks: pairs with the longest dependency chain first (columns plus the
    // start-up lag of every strip), then by size; the strips of a pair stay
    // contiguous and in order, which is what makes the persistent claim loop
    // deadlock-free.  The longest chains also get the highest wave issue
    // priority, so that co-resident strips of short pairs fill the issue
    // slots they leave idle instead of stretching the critical path.
    std::vector<int64_t> chain(np, 0);
    int64_t max_chain = 1;
    const int64_t strip_lag = (int64_t)v->waves * (64 + 2 * v->chunk);
    for (int64_t i = 0; i < np; ++i) {
        const WfPairDev &p = ph->pairs[i];
        if (p.ngroups == 0) continue;
        chain[i] = p.n1 + 63 + (int64_t)p.ngroups * strip_lag;
        max_chain = std::max(max_chain, chain[i]);
    }
    static const bool use_prio = [] {
        const char *e = std::getenv("WFDP_PRIORITY");
        return !(e && e[0] == '0');
    }();
    for (int64_t i = 0; i < np; ++i)
        ph->pairs[i].prio = use_prio ? (int32_t)std::min<int64_t>(3, 4 * chain[i] / (max_chain + 1)) : 0;
    std::vector<int64_t> order(np);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int64_t x, int64_t y) {
        if (chain[x] != chain[y]) return chain[x] > chain[y];
        return ph->cells[x] > ph->cells[y];
    });
    ph->tasks.clear();
    ph->task_begin.assign(np, 0);
    ph->task_end.assign(np, 0);
    for (int64_t k = 0; k < np; ++k) {
        const int64_t i = order[k];
        ph->task_begin[i] = (int64_t)ph->tasks.size();
        for (int32_t g = 0; g < ph->pairs[i].ngroups; ++g) ph->tasks.push_back(WfTaskDev{(int32_t)i, g});
        ph->task_end[i] = (int64_t)ph->tasks.size();
    }
    ph->bnd_ints = bnd_ints;
    ph->order = order;
    return WF_OK;
}

int wf_upload_plan(wf_ctx *ctx, WfPlanHost *ph, WfArenas *ar)
{
    return wf_upload_plan_async(ctx, ph, ar, ctx->stream, true);
}

void wf_plan_bytes(const WfPlanHost *ph, WfPlanBytes *b)
{
    b->codes = ph->codes.size();
    b->pairs = std::max<size_t>(1, ph->pairs.size()) * sizeof(WfPairDev);
    b->tasks = std::max<size_t>(1, ph->tasks.size()) * sizeof(WfTaskDev);
    b->bnd = (size_t)std::max<int64_t>(1, ph->bnd_ints) * sizeof(uint64_t);
    b->res = (size_t)std::max<int64_t>(1, ph->npairs) * sizeof(WfResultDev);
    b->counters = (size_t)(ph->npairs + 1) * sizeof(int32_t);
    b->epoch = sizeof(uint32_t);
    b->colw = (size_t)std::max<int64_t>(1, ph->colw_entries) * 2 * sizeof(uint32_t);
    b->splits = std::max<size_t>(1, ph->splits.size()) * sizeof(WfSplitDev);
}

int wf_copy_plan(WfPlanHost *ph, WfArenas *ar, hipStream_t stream)
{
    const size_t ntasks = ph->tasks.size();
    if (hipMemcpyAsync(ar->codes.ptr, ph->codes.data(), ph->codes.size(), hipMemcpyHostToDevice,
                       stream) != hipSuccess)
        return WF_ERR_DEVICE;
    if (!ph->splits.empty() &&
        hipMemcpyAsync(ar->splits.ptr, ph->splits.data(), ph->splits.size() * sizeof(WfSplitDev),
                       hipMemcpyHostToDevice, stream) != hipSuccess)
        return WF_ERR_DEVICE;
    if (!ph->pairs.empty() &&
        hipMemcpyAsync(ar->pairs.ptr, ph->pairs.data(), ph->pairs.size() * sizeof(WfPairDev),
                       hipMemcpyHostToDevice, stream) != hipSuccess)
        return WF_ERR_DEVICE;
    if (ntasks &&
        hipMemcpyAsync(ar->tasks.ptr, ph->tasks.data(), ntasks * sizeof(WfTaskDev),
                       hipMemcpyHostToDevice, stream) != hipSuccess)
        return WF_ERR_DEVICE;
    return WF_OK;
}

// Sizes every device buffer of the plan (growing them as needed).
static int ensure_arenas(WfPlanHost *ph, WfArenas *ar)
{
    WfPlanBytes nb;
    wf_plan_bytes(ph, &nb);
    int st;
    if ((st = ar->codes.ensure(nb.codes)) != WF_OK) return st;
    if ((st = ar->pairs.ensure(nb.pairs)) != WF_OK) return st;
    if ((st = ar->tasks.ensure(nb.tasks)) != WF_OK) return st;
    if ((st = ar->bnd.ensure(nb.bnd, true)) != WF_OK) return st;
    if ((st = ar->res.ensure(nb.res)) != WF_OK) return st;
    if ((st = ar->counters.ensure(nb.counters)) != WF_OK) return st;
    // Launch epoch: zero on allocation, so it never matches a zeroed boundary word.
    if ((st = ar->epoch.ensure(nb.epoch, true)) != WF_OK) return st;
    if ((st = ar->colw.ensure(nb.colw)) != WF_OK) return st;
    return ar->splits.ensure(nb.splits);
}

int wf_upload_plan_async(wf_ctx *ctx, WfPlanHost *ph, WfArenas *ar, hipStream_t stream, bool sync)
{
    (void)ctx;
    int st;
    if ((st = ensure_arenas(ph, ar)) != WF_OK) return st;
    if ((st = wf_copy_plan(ph, ar, stream)) != WF_OK) return st;
    if (sync && hipStreamSynchronize(stream) != hipSuccess) return WF_ERR_DEVICE;
    return WF_OK;
}

namespace {

size_t round256(size_t n) { return (n + 255) & ~size_t(255); }

// The context's pinned block, at least `bytes` long and idle (a call that
// failed after queueing its copies may have left them in flight).
int ensure_pinned(wf_ctx *ctx, size_t bytes)
{
    if (ctx->pinned_busy) {
        if (hipStreamSynchronize(ctx->stream) != hipSuccess) return WF_ERR_DEVICE;
        ctx->pinned_busy = false;
    }
    if (ctx->pinned_cap >= bytes) return WF_OK;
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    ctx->pinned = nullptr;
    ctx->pinned_cap = 0;
    const size_t cap = std::max<size_t>(round256(bytes), (size_t)1 << 20);
    if (hipHostMalloc((void **)&ctx->pinned, cap, hipHostMallocDefault) != hipSuccess) {
        ctx->pinned = nullptr;
        return WF_ERR_NOMEM;
    }
    ctx->pinned_cap = cap;
    return WF_OK;
}

// Byte offsets of a plan's buffers in the pinned block: codes, splits,
// pairs, tasks, then the results.
struct PinnedLayout {
    size_t codes = 0, splits, pairs, tasks, res, end;
    explicit PinnedLayout(const WfPlanHost *ph)
    {
        splits = round256(ph->codes.size());
        pairs = splits + round256(ph->splits.size() * sizeof(WfSplitDev));
        tasks = pairs + round256(ph->pairs.size() * sizeof(WfPairDev));
        res = tasks + round256(ph->tasks.size() * sizeof(WfTaskDev));
        end = res + round256((size_t)std::max<int64_t>(1, ph->npairs) * sizeof(WfResultDev));
    }
};

// wf_upload_plan through the context's pinned block: the copies are
// asynchronous DMA from pinned memory, and the stream orders them before
// the launch, so the host does not wait for them.  The block must not be
// rewritten before the stream has drained (wf_align_batch synchronises
// before it returns).
int upload_pinned(wf_ctx *ctx, WfPlanHost *ph, WfArenas *ar)
{
    int st;
    if ((st = ensure_arenas(ph, ar)) != WF_OK) return st;
    const PinnedLayout L(ph);
    if ((st = ensure_pinned(ctx, L.end)) != WF_OK) return st;
    ctx->pinned_busy = true;
    auto put = [&](size_t off, void *dst, const void *src, size_t n) -> bool {
        if (!n) return true;
        std::memcpy(ctx->pinned + off, src, n);
        return hipMemcpyAsync(dst, ctx->pinned + off, n, hipMemcpyHostToDevice, ctx->stream) == hipSuccess;
    };
    if (!put(L.codes, ar->codes.ptr, ph->codes.data(), ph->codes.size()) ||
        !put(L.splits, ar->splits.ptr, ph->splits.data(), ph->splits.size() * sizeof(WfSplitDev)) ||
        !put(L.pairs, ar->pairs.ptr, ph->pairs.data(), ph->pairs.size() * sizeof(WfPairDev)) ||
        !put(L.tasks, ar->tasks.ptr, ph->tasks.data(), ph->tasks.size() * sizeof(WfTaskDev)))
        return WF_ERR_DEVICE;
    return WF_OK;
}

}  // namespace

int wf_max_resident_groups(wf_ctx *ctx, const WfVariant *v)
{
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, v->fn, v->block, 0) != hipSuccess ||
        per_cu < 1) {
        (void)hipGetLastError();
        per_cu = 1;
    }
    if (const char *env = std::getenv("WFDP_WG_PER_CU")) {
        const int e = std::atoi(env);
        if (e > 0) per_cu = std::min(per_cu, e);
    }
    return per_cu * ctx->prop.multiProcessorCount;
}

// The engine launch(es) of a plan: one persistent launch, or one per pair.
static int enqueue_engine(WfPlanHost *ph, const WfVariant *v, int lim, int mode, WfKernelArgs ka, WfArenas *ar,
                          hipStream_t stream)
{
    if (mode == WF_RUN_PER_PAIR) {
        for (int64_t k = 0; k < ph->npairs; ++k) {
            const int64_t i = ph->order[k];
            const int64_t tb = ph->task_begin[i], te = ph->task_end[i];
            if (te <= tb) continue;
            ka.counter = (int32_t *)ar->counters.ptr + 1 + i;
            ka.task_begin = (int32_t)tb;
            ka.task_end = (int32_t)te;
            const int grid = (int)std::min<int64_t>(te - tb, lim);
            ph->grid = std::max(ph->grid, grid);
            if (wf_launch_engine(v, grid, &ka, stream) != hipSuccess) return WF_ERR_DEVICE;
        }
        return WF_OK;
    }
    ka.counter = (int32_t *)ar->counters.ptr;
    ka.task_begin = 0;
    ka.task_end = (int32_t)ph->tasks.size();
    ph->grid = (int)std::min<int64_t>((int64_t)ph->tasks.size(), lim);
    if (wf_launch_engine(v, ph->grid, &ka, stream) != hipSuccess) return WF_ERR_DEVICE;
    return WF_OK;
}

// Enqueue the reset kernel and the engine launch(es) of a plan on the
// context stream (also used inside hipGraph capture: no memset nodes, and
// the epoch that tags boundary values is advanced on the device).
int wf_enqueue_plan(wf_ctx *ctx, WfPlanHost *ph, WfArenas *ar, int mode, int maxg, hipStream_t stream)
{
    const int64_t ncounters = ph->npairs + 1;
    if (wf_launch_reset((WfResultDev *)ar->res.ptr, ph->npairs, (int32_t *)ar->counters.ptr, ncounters,
                        (uint32_t *)ar->epoch.ptr, stream) != hipSuccess)
        return WF_ERR_DEVICE;
    WfKernelArgs ka{};
    ka.pairs = (const WfPairDev *)ar->pairs.ptr;
    ka.tasks = (const WfTaskDev *)ar->tasks.ptr;
    ka.codes = (const uint8_t *)ar->codes.ptr;
    ka.bnd = (uint64_t *)ar->bnd.ptr;
    ka.res = (WfResultDev *)ar->res.ptr;
    ka.epoch = (const uint32_t *)ar->epoch.ptr;
    ka.timeout_ticks = (int64_t)(ctx->timeout_s * 1.0e8);
    ka.colw = (uint32_t *)ar->colw.ptr;
    ph->grid = 0;
    if (ph->tasks.empty()) return WF_OK;
    const WfVariant *v = ph->variant;
    const int lim = maxg > 0 ? maxg : wf_max_resident_groups(ctx, v);
    const int st = enqueue_engine(ph, v, lim, mode, ka, ar, stream);
    if (st != WF_OK) return st;
    if (!ph->splits.empty() &&
        wf_launch_bitpar_combine((const WfSplitDev *)ar->splits.ptr, (int)ph->splits.size(),
                                 (const uint32_t *)ar->colw.ptr, (WfResultDev *)ar->res.ptr, stream) != hipSuccess)
        return WF_ERR_DEVICE;
    return WF_OK;
}

// Run a plan (directly, or by replaying its captured graph) and time the
// device work with HIP events on the context stream.
int wf_run_plan(wf_ctx *ctx, WfPlanHost *ph, WfArenas *ar, int mode, double *kernel_ms, hipGraphExec_t exec)
{
    const int maxg = wf_max_resident_groups(ctx, ph->variant);
    if (hipEventRecord(ctx->ev0, ctx->stream) != hipSuccess) return WF_ERR_DEVICE;
    if (exec) {
        if (hipGraphLaunch(exec, ctx->stream) != hipSuccess) return WF_ERR_DEVICE;
    } else {
        const int st = wf_enqueue_plan(ctx, ph, ar, mode, maxg, ctx->stream);
        if (st != WF_OK) return st;
    }
    if (hipEventRecord(ctx->ev1, ctx->stream) != hipSuccess) return WF_ERR_DEVICE;
    if (hipEventSynchronize(ctx->ev1) != hipSuccess) return WF_ERR_DEVICE;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) != hipSuccess) return WF_ERR_DEVICE;
    *kernel_ms = ms;
    return WF_OK;
}

// The caller's results from the device results r of the plan's pairs.
static int fill_results(const WfPlanHost *ph, const WfResultDev *r, double kernel_ms, wf_result *out, int64_t n)
{
    int worst = WF_OK;
    for (int64_t i = 0; i < std::min(n, ph->njobs); ++i) {
        wf_result &o = out[i];
        std::memset(&o, 0, sizeof(o));
        o.cells = ph->cells[i];
        o.kernel_ms = kernel_ms;
        if (ph->host_only[i]) {
            o.last = ph->host_result[i];
            o.max_score = 0;
            continue;
        }
        o.last = r[i].last;
        o.max_score = (ph->pairs[i].mode & WFK_MODE_NOMAX) ? -1 : r[i].maxs;
        o.status = r[i].status ? WF_ERR_TIMEOUT : WF_OK;
        if (o.status) worst = o.status;
    }
    return worst;
}

int wf_download_results(wf_ctx *ctx, WfPlanHost *ph, WfArenas *ar, double kernel_ms,
                            wf_result *out, int64_t n)
{
    std::vector<WfResultDev> r((size_t)std::max<int64_t>(1, ph->npairs));
    if (ph->npairs > 0 &&
        hipMemcpy(r.data(), ar->res.ptr, ph->npairs * sizeof(WfResultDev), hipMemcpyDeviceToHost) !=
            hipSuccess)
        return WF_ERR_DEVICE;
    (void)ctx;
    return fill_results(ph, r.data(), kernel_ms, out, n);
}

// ----------------------------------------------------------------------------
// C ABI
// ----------------------------------------------------------------------------

extern "C" {

int wf_abi_version(void) { return WFDP_ABI_VERSION; }

int wf_variant_count(void) { return variant_count(); }

int wf_variant_info(int index, int32_t *engine, int32_t *rows_per_lane, int32_t *waves, int32_t *chunk)
{
    const WfVariant *v = variant_at(index);
    if (!v) return WF_ERR_INVALID;
    if (engine) *engine = v->engine;
    if (rows_per_lane) *rows_per_lane = v->rows_per_lane;
    if (waves) *waves = v->waves;
    if (chunk) *chunk = v->chunk;
    return WF_OK;
}

const char *wf_strerror(int status)
{
    switch (status) {
    case WF_OK: return "ok";
    case WF_ERR_INVALID: return "invalid argument";
    case WF_ERR_DEVICE: return "HIP runtime error";
    case WF_ERR_NOMEM: return "device memory exhausted";
    case WF_ERR_ALPHABET: return "more than 256 distinct symbols in one pair";
    case WF_ERR_TIMEOUT: return "device-side dependency wait timed out";
    case WF_ERR_NOKERNEL: return "no compiled kernel for the requested tile";
    case WF_ERR_AGAIN: return "queue full, retry";
    case WF_ERR_SCRATCH: return "the pair does not fit the fixed scratch area of wf_align_noalloc (64 MiB)";
    default: return "unknown error";
    }
}

int wf_ctx_create(int device, wf_ctx **out)
{
    if (!out) return WF_ERR_INVALID;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        return WF_ERR_DEVICE;
    }
    if (device < 0 || device >= ndev) return WF_ERR_INVALID;
    if (hipSetDevice(device) != hipSuccess) return WF_ERR_DEVICE;
    wf_ctx *c = new wf_ctx();
    c->device = device;
    if (hipGetDeviceProperties(&c->prop, device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        delete c;
        return WF_ERR_DEVICE;
    }
    // Set up this device's static scratch area of wf_align_noalloc (once per
    // device); a failure here is reported by that entry point's own call.
    (void)wf_noalloc_setup(device, c->stream, nullptr);
    wf_tune_autoload();
    *out = c;
    return WF_OK;
}

void wf_ctx_destroy(wf_ctx *ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    ctx->scratch.release();
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);  // after the stream's synchronisation above
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int wf_ctx_set_timeout(wf_ctx *ctx, double seconds)
{
    if (!ctx || !(seconds > 0.0)) return WF_ERR_INVALID;
    ctx->timeout_s = seconds;
    return WF_OK;
}

int wf_device_query(wf_ctx *ctx, wf_device_info *out)
{
    if (!ctx || !out) return WF_ERR_INVALID;
    std::memset(out, 0, sizeof(*out));
    out->compute_units = ctx->prop.multiProcessorCount;
    out->max_clock_mhz = ctx->prop.clockRate / 1000;
    out->simds_per_cu = 4;
    out->wave_size = ctx->prop.warpSize;
    (void)hipSetDevice(ctx->device);
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) == hipSuccess) {
        out->free_mem = (int64_t)fr;
        out->total_mem = (int64_t)tot;
    }
    std::snprintf(out->arch, sizeof(out->arch), "%s", ctx->prop.gcnArchName);
    return WF_OK;
}

int wf_align_batch(wf_ctx *ctx, const wf_job *jobs, int64_t njobs, const wf_tile *tile, wf_result *out)
{
    if (!ctx || (njobs > 0 && (!jobs || !out)) || njobs < 0) return WF_ERR_INVALID;
    if (njobs == 0) return WF_OK;
    if (hipSetDevice(ctx->device) != hipSuccess) return WF_ERR_DEVICE;
    WfPlanHost ph;
    int st = wf_build_plan(ctx, jobs, njobs, tile, &ph);
    if (st != WF_OK) return st;
    if ((st = upload_pinned(ctx, &ph, &ctx->scratch)) != WF_OK) return st;
    double ms = 0.0;
    if ((st = wf_run_plan(ctx, &ph, &ctx->scratch, WF_RUN_PERSISTENT, &ms, nullptr)) != WF_OK) return st;
    // Results into the pinned block (its upload part is no longer read).
    const PinnedLayout L(&ph);
    WfResultDev *r = reinterpret_cast<WfResultDev *>(ctx->pinned + L.res);
    if (ph.npairs > 0 &&
        (hipMemcpyAsync(r, ctx->scratch.res.ptr, ph.npairs * sizeof(WfResultDev), hipMemcpyDeviceToHost,
                        ctx->stream) != hipSuccess ||
         hipStreamSynchronize(ctx->stream) != hipSuccess))
        return WF_ERR_DEVICE;
    ctx->pinned_busy = false;
    return fill_results(&ph, r, ms, out, njobs);
}

int wf_align(wf_ctx *ctx, int problem, const int32_t *a, int64_t n1, const int32_t *b, int64_t n2,
             const wf_tile *tile, wf_result *out)
{
    if (!ctx || !out) return WF_ERR_INVALID;
    wf_job j{problem, tile ? (tile->flags & WF_TILE_ED_TEXTBOOK) : 0, a, n1, b, n2};
    return wf_align_batch(ctx, &j, 1, tile, out);
}

int wf_align_table(wf_ctx *ctx, int problem, const int32_t *a, int64_t n1, const int32_t *b, int64_t n2,
                   const wf_tile *tile, int32_t *table, wf_result *out)
{
    if (!ctx || !out || !table || n1 <= 0 || n2 <= 0 || n1 * n2 > WF_TABLE_MAX_CELLS) return WF_ERR_INVALID;
    if (hipSetDevice(ctx->device) != hipSuccess) return WF_ERR_DEVICE;
    const int flags = tile ? (tile->flags & WF_TILE_ED_TEXTBOOK) : 0;
    const wf_job j{problem, flags, a, n1, b, n2};
    // The table kernel is the lock-step sweep <R=2, W=4, K=32>; the plan's
    // strips must have its geometry.
    const wf_tile lock{32, 512, 2, WF_TILE_LOCKSTEP | flags};
    WfPlanHost ph;
    int st = wf_build_plan(ctx, &j, 1, &lock, &ph);
    if (st != WF_OK) return st;
    const WfVariant *v = ph.variant;
    if (v->engine != WF_ENGINE_LOCKSTEP || v->rows_per_lane != 2 || v->waves != 4 || v->chunk != 32)
        return WF_ERR_NOKERNEL;
    if ((st = wf_upload_plan(ctx, &ph, &ctx->scratch)) != WF_OK) return st;
    DevBuf dt;
    if ((st = dt.ensure((size_t)(n1 * n2) * sizeof(int32_t))) != WF_OK) return st;
    WfArenas *ar = &ctx->scratch;
    if (wf_launch_reset((WfResultDev *)ar->res.ptr, 1, (int32_t *)ar->counters.ptr, 2, (uint32_t *)ar->epoch.ptr,
                        ctx->stream) != hipSuccess)
        return WF_ERR_DEVICE;
    WfKernelArgs ka{};
    ka.pairs = (const WfPairDev *)ar->pairs.ptr;
    ka.tasks = (const WfTaskDev *)ar->tasks.ptr;
    ka.codes = (const uint8_t *)ar->codes.ptr;
    ka.bnd = (uint64_t *)ar->bnd.ptr;
    ka.res = (WfResultDev *)ar->res.ptr;
    ka.epoch = (const uint32_t *)ar->epoch.ptr;
    ka.timeout_ticks = (int64_t)(ctx->timeout_s * 1.0e8);
    ka.counter = (int32_t *)ar->counters.ptr;
    ka.task_begin = 0;
    ka.task_end = (int32_t)ph.tasks.size();
    ka.table = (int32_t *)dt.ptr;
    const int grid = (int)std::min<int64_t>((int64_t)ph.tasks.size(), ctx->prop.multiProcessorCount);
    if (hipEventRecord(ctx->ev0, ctx->stream) != hipSuccess) return WF_ERR_DEVICE;
    if (grid > 0 && wf_launch_table(grid, &ka, ctx->stream) != hipSuccess) return WF_ERR_DEVICE;
    if (hipEventRecord(ctx->ev1, ctx->stream) != hipSuccess) return WF_ERR_DEVICE;
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return WF_ERR_DEVICE;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    if (hipMemcpy(table, dt.ptr, (size_t)(n1 * n2) * sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess)
        return WF_ERR_DEVICE;
    dt.release();
    return wf_download_results(ctx, &ph, ar, ms, out, 1);
}

int wf_align_naive(wf_ctx *ctx, int problem, const int32_t *a, int64_t n1, const int32_t *b, int64_t n2,
                   const wf_tile *tile, int method, int32_t *table, wf_result *out)
{
    if (!ctx || !out || n1 < 0 || n2 < 0 || (n1 > 0 && !a) || (n2 > 0 && !b)) return WF_ERR_INVALID;
    if (problem != WF_LCS && problem != WF_ED && problem != WF_SW) return WF_ERR_INVALID;
    if (method != WF_NAIVE_DIAGONAL && method != WF_NAIVE_TILE) return WF_ERR_INVALID;
    if (n1 >= INT32_MAX || n2 >= INT32_MAX) return WF_ERR_INVALID;
    const bool textbook = tile && (tile->flags & WF_TILE_ED_TEXTBOOK);
    *out = wf_result{};
    out->cells = (uint64_t)n1 * (uint64_t)n2;
    if (n1 == 0 || n2 == 0) {
        // One side empty: the table is its boundary.
        out->last = (problem == WF_ED) ? (int32_t)(n2 > 0 ? n2 : (textbook ? n1 : std::max<int64_t>(0, n1 - 1))) : 0;
        return WF_OK;
    }
    if (hipSetDevice(ctx->device) != hipSuccess) return WF_ERR_DEVICE;
    const int64_t cells = (n1 + 1) * (n2 + 1);
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return WF_ERR_DEVICE;
    const size_t need = (size_t)cells * sizeof(int32_t) + (size_t)(n1 + n2 + 64) * sizeof(int32_t);
    if (need > fr) return WF_ERR_NOMEM;
    DevBuf dt, dab;
    int st;
    if ((st = dt.ensure((size_t)cells * sizeof(int32_t))) != WF_OK) return st;
    if ((st = dab.ensure((size_t)(n1 + n2 + 64) * sizeof(int32_t))) != WF_OK) return st;
    int32_t *da = (int32_t *)dab.ptr, *db = da + n1, *smax = db + n2;
    if (hipMemcpyAsync(da, a, (size_t)n1 * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
        hipMemcpyAsync(db, b, (size_t)n2 * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
        hipMemsetAsync(smax, 0, sizeof(int32_t), ctx->stream) != hipSuccess)
        return WF_ERR_DEVICE;
    if (hipEventRecord(ctx->ev0, ctx->stream) != hipSuccess) return WF_ERR_DEVICE;
    if (wf_launch_naive(method, problem, textbook ? 1 : 0, (int *)dt.ptr, da, db, (int)n1, (int)n2,
                        problem == WF_SW ? smax : nullptr, ctx->stream) != hipSuccess)
        return WF_ERR_DEVICE;
    if (hipEventRecord(ctx->ev1, ctx->stream) != hipSuccess) return WF_ERR_DEVICE;
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return WF_ERR_DEVICE;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    const int32_t *dlast = (const int32_t *)dt.ptr + cells - 1;
    int32_t last = 0, mx = 0;
    if (hipMemcpy(&last, dlast, sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(&mx, smax, sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess)
        return WF_ERR_DEVICE;
    if (table && hipMemcpy2D(table, (size_t)n1 * sizeof(int32_t), (const int32_t *)dt.ptr + n1 + 2,
                             (size_t)(n1 + 1) * sizeof(int32_t), (size_t)n1 * sizeof(int32_t), (size_t)n2,
                             hipMemcpyDeviceToHost) != hipSuccess)
        return WF_ERR_DEVICE;
    out->last = last;
    out->max_score = problem == WF_SW ? mx : 0;
    out->kernel_ms = ms;
    return WF_OK;
}

int wf_plan_create(wf_ctx *ctx, const wf_job *jobs, int64_t njobs, const wf_tile *tile, wf_plan **out)
{
    if (!ctx || !out || njobs <= 0 || !jobs) return WF_ERR_INVALID;
    *out = nullptr;
    if (hipSetDevice(ctx->device) != hipSuccess) return WF_ERR_DEVICE;
    wf_plan *p = new wf_plan();
    p->ctx = ctx;
    int st = wf_build_plan(ctx, jobs, njobs, tile, &p->host);
    if (st == WF_OK) st = wf_upload_plan(ctx, &p->host, &p->own);
    if (st != WF_OK) {
        p->own.release();
        delete p;
        return st;
    }
    *out = p;
    return WF_OK;
}

int wf_plan_run(wf_plan *plan, int mode)
{
    if (!plan || (mode != WF_RUN_PERSISTENT && mode != WF_RUN_PER_PAIR)) return WF_ERR_INVALID;
    if (hipSetDevice(plan->ctx->device) != hipSuccess) return WF_ERR_DEVICE;
    const int st = wf_run_plan(plan->ctx, &plan->host, &plan->own, mode, &plan->kernel_ms, plan->exec[mode]);
    plan->ran = (st == WF_OK);
    return st;
}

int wf_plan_capture(wf_plan *plan, int mode)
{
    if (!plan || (mode != WF_RUN_PERSISTENT && mode != WF_RUN_PER_PAIR)) return WF_ERR_INVALID;
    wf_ctx *ctx = plan->ctx;
    if (hipSetDevice(ctx->device) != hipSuccess) return WF_ERR_DEVICE;
    if (plan->exec[mode]) return WF_OK;
    const int maxg = wf_max_resident_groups(ctx, plan->host.variant);
    if (hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return WF_ERR_DEVICE;
    const int st = wf_enqueue_plan(ctx, &plan->host, &plan->own, mode, maxg, ctx->stream);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (st != WF_OK || e != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        return st != WF_OK ? st : WF_ERR_DEVICE;
    }
    hipGraphExec_t x = nullptr;
    if (hipGraphInstantiate(&x, g, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(g);
        return WF_ERR_DEVICE;
    }
    plan->graph[mode] = g;
    plan->exec[mode] = x;
    return WF_OK;
}

int wf_plan_results(wf_plan *plan, wf_result *out, int64_t n)
{
    if (!plan || !out || n < 0) return WF_ERR_INVALID;
    if (!plan->ran) return WF_ERR_INVALID;
    if (hipSetDevice(plan->ctx->device) != hipSuccess) return WF_ERR_DEVICE;
    return wf_download_results(plan->ctx, &plan->host, &plan->own, plan->kernel_ms, out, n);
}

double wf_plan_kernel_ms(const wf_plan *plan) { return plan ? plan->kernel_ms : -1.0; }

int wf_plan_describe(const wf_plan *plan, int32_t *engine, int32_t *rows_per_lane, int32_t *waves,
                     int32_t *chunk, int64_t *ntasks, int32_t *grid)
{
    if (!plan) return WF_ERR_INVALID;
    const WfVariant *v = plan->host.variant;
    if (engine) *engine = v->engine;
    if (rows_per_lane) *rows_per_lane = v->rows_per_lane;
    if (waves) *waves = v->waves;
    if (chunk) *chunk = v->chunk;
    if (ntasks) *ntasks = (int64_t)plan->host.tasks.size();
    if (grid) *grid = plan->host.grid;
    return WF_OK;
}

void wf_plan_destroy(wf_plan *plan)
{
    if (!plan) return;
    (void)hipSetDevice(plan->ctx->device);
    (void)hipStreamSynchronize(plan->ctx->stream);
    for (int m = 0; m < 2; ++m) {
        if (plan->exec[m]) (void)hipGraphExecDestroy(plan->exec[m]);
        if (plan->graph[m]) (void)hipGraphDestroy(plan->graph[m]);
    }
    plan->own.release();
    delete plan;
}

}  // extern "C"
