// wf_compat.cpp -- the reference's per-problem C++ entry points over the C
// ABI.  Built six times: -DWF_COMPAT_POLICY_SHARE and _CACHE (the hyperlane
// drivers' functions), _NAIVE and _TILE (the naive baselines' functions),
// _SAT and _SOR (the SAT and 2D-SOR programs' SOR()).
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "wfdp.h"
#include "wfdp_compat.h"

static wf_ctx *g_ctx = nullptr;

static void fail(int st)
{
    std::printf("GPU error in %s: %s\n", __FILE__, wf_strerror(st));
    std::exit(EXIT_FAILURE);
}

static wf_ctx *ctx()
{
    if (!g_ctx) {
        const int st = wf_ctx_create(0, &g_ctx);
        if (st != WF_OK) fail(st);
    }
    return g_ctx;
}

// The reference copies its device table back to the host only in DEBUG
// builds (LCS/GPU_Hyperlane_Share.cu:770-772; the SAT and 2D-SOR hosts not
// even there, SAT/GPU_Hyperlane_Share.cu:317).  The environment variable
// WFDP_DEBUG_TABLE=1 plays that role here.
static bool debug_table()
{
    static const int on = [] {
        const char *e = std::getenv("WFDP_DEBUG_TABLE");
        return (e && *e && *e != '0') ? 1 : 0;
    }();
    return on != 0;
}

#if defined(WF_COMPAT_POLICY_SAT) || defined(WF_COMPAT_POLICY_SOR)
#if defined(WF_COMPAT_POLICY_SAT)
static const int kStencil = WF_SAT;
#else
static const int kStencil = WF_SOR;
#endif

// One sweep of the (n2 + 2) x (n1 + 3) table the driver read (it holds n2 + 3
// rows; the reference uploads n2 + 2, SAT/GPU_Hyperlane_Share.cu:261-278).
// The reference discards the result; with WFDP_DEBUG_TABLE=1 it is written
// back into `table`.  tX and tY only shape the reference's tiles.
static void run_stencil(int n1, int n2, int *table)
{
    wf_tile t{0, 0, 0, debug_table() ? 0 : WF_TILE_NO_COPYBACK};
    wf_result r{};
    const int st = wf_stencil(ctx(), kStencil, n1, n2, reinterpret_cast<int32_t *>(table), (int64_t)n1 + 3, &t, &r);
    if (st != WF_OK) fail(st);
}

// SAT/GPU_Hyperlane_Share.h:5, 2D-SOR/nested_loop/GPU_Hyperlane_Share.h:5
void SOR(int n1, int n2, int *table, int tX, int tY)
{
    (void)tX;
    (void)tY;
    run_stencil(n1, n2, table);
}

// SAT/GPU_Hyperlane_Cache.h:5, 2D-SOR/nested_loop/GPU_Hyperlane_Cache.h:5
void SOR(int n1, int n2, int *table)
{
    run_stencil(n1, n2, table);
}

#elif defined(WF_COMPAT_POLICY_NAIVE) || defined(WF_COMPAT_POLICY_TILE)
#if defined(WF_COMPAT_POLICY_NAIVE)
static const int kMethod = WF_NAIVE_DIAGONAL;
#else
static const int kMethod = WF_NAIVE_TILE;
#endif

// The memory line the reference's naive functions print before they
// allocate the table (LCS/GPU.cu:51-56): free and total device memory, the
// padded table's dimensions and its size in bytes.
static int run_naive(int problem, int n1, int n2, const int *arr1, const int *arr2, int flags)
{
    wf_device_info d{};
    int st = wf_device_query(ctx(), &d);
    if (st != WF_OK) fail(st);
    const long long colsize = 1LL + n1, rowsize = 1LL + n2;
    std::printf("current GPU memory info FREE: %lld Bytes, Total: %lld Bytes.colsize: %lld, rowsize: %lld, "
                "allocates: %lld Bytes.\n",
                (long long)d.free_mem, (long long)d.total_mem, colsize, rowsize,
                colsize * rowsize * (long long)sizeof(int));
    std::fflush(stdout);
    wf_tile t{0, 0, 0, flags};
    wf_result r{};
    st = wf_align_naive(ctx(), problem, reinterpret_cast<const int32_t *>(arr1), n1,
                        reinterpret_cast<const int32_t *>(arr2), n2, &t, kMethod, nullptr, &r);
    if (st != WF_OK) fail(st);
    return r.last;
}

// LCS/GPU.h:5 and LCS/GPU_TILE_SHARE.h:5 (whose SmithWaterman and
// EditDistance copies declare, and compute, LCS as well).
int LCS(int n1, int n2, int *arr1, int *arr2)
{
    return run_naive(WF_LCS, n1, n2, arr1, arr2, 0);
}

#if defined(WF_COMPAT_POLICY_NAIVE)
// SmithWaterman/GPU.h:5
int SW(int n1, int n2, int *arr1, int *arr2)
{
    return run_naive(WF_SW, n1, n2, arr1, arr2, 0);
}

// EditDistance/GPU.h:5; row 0 is the textbook x (EditDistance/GPU.cu:12-18).
int ED(int n1, int n2, int *arr1, int *arr2)
{
    return run_naive(WF_ED, n1, n2, arr1, arr2, WF_TILE_ED_TEXTBOOK);
}
#endif

#else  // hyperlane drivers

// The reference sizes tX x tY tiles for its shared-memory staging on a
// Pascal GPU; here they are hints (WF_TILE_HINT): every variant gives the
// same score, and the planner runs its fastest one for the pair.  The Cache
// library keeps its policy (strip boundaries through L2 only).
#if defined(WF_COMPAT_POLICY_CACHE)
static const int kDefaultTX = 256, kDefaultTY = 1024, kPolicyFlags = WF_TILE_CACHE;
#else
static const int kDefaultTX = 32, kDefaultTY = 256, kPolicyFlags = WF_TILE_HINT;
#endif

// With WFDP_DEBUG_TABLE=1 a non-null table of a square pair receives the
// whole DP table in the reference's padded layout,
// table[(i + paddY) * (n2 + paddX) + (j + paddX)] = H(i + 1, j + 1), which its
// drivers' DEBUG dump prints (LCS/GPU_Hyperlane_Share.cpp:109-131).  Without
// it the score-only path runs and `table` is left alone.

static int run(int problem, int n1, int n2, const int *arr1, const int *arr2, int tX, int tY, int paddX = 0,
               int paddY = 0, int *table = nullptr)
{
    wf_ctx *c = ctx();
    // The reference's SW() returns the bottom-right cell only, so the running
    // maximum is not computed (WF_TILE_NO_MAX).
    wf_tile t{tX, tY, 0, kPolicyFlags | (problem == WF_SW ? WF_TILE_NO_MAX : 0)};
    wf_result r{};
    const int32_t *a = reinterpret_cast<const int32_t *>(arr1), *b = reinterpret_cast<const int32_t *>(arr2);
    if (table && debug_table() && n1 == n2 && n1 > 0 && (int64_t)n1 * n2 <= WF_TABLE_MAX_CELLS) {
        std::vector<int32_t> h((size_t)n1 * n2);
        const int st = wf_align_table(c, problem, a, n1, b, n2, &t, h.data(), &r);
        if (st != WF_OK) fail(st);
        for (int i = 0; i < n2; i++)
            for (int j = 0; j < n1; j++)
                table[(int64_t)(i + paddY) * (n2 + paddX) + (j + paddX)] = h[(size_t)i * n1 + j];
        return r.last;
    }
    const int st = wf_align(c, problem, a, n1, b, n2, &t, &r);
    if (st != WF_OK) fail(st);
    return r.last;
}

int LCS(int n1, int n2, int *arr1, int *arr2, int paddX, int paddY, int *table)
{
    return run(WF_LCS, n1, n2, arr1, arr2, kDefaultTX, kDefaultTY, paddX, paddY, table);
}

int LCS(int n1, int n2, int *arr1, int *arr2, int paddX, int paddY, int *table, int tX, int tY)
{
    return run(WF_LCS, n1, n2, arr1, arr2, tX, tY, paddX, paddY, table);
}

int LCS(int n1, int n2, int *arr1, int *arr2)
{
    return run(WF_LCS, n1, n2, arr1, arr2, kDefaultTX, kDefaultTY);
}

int SW(int n1, int n2, int *arr1, int *arr2, int paddX, int paddY, int *table, int tX, int tY)
{
    return run(WF_SW, n1, n2, arr1, arr2, tX, tY, paddX, paddY, table);
}

int SW(int n1, int n2, int *arr1, int *arr2, int paddX, int paddY, int *table)
{
    return run(WF_SW, n1, n2, arr1, arr2, kDefaultTX, kDefaultTY, paddX, paddY, table);
}

int ED(int n1, int n2, int *arr1, int *arr2, int paddX, int paddY, int *table, int tX, int tY)
{
    return run(WF_ED, n1, n2, arr1, arr2, tX, tY, paddX, paddY, table);
}

int ED(int n1, int n2, int *arr1, int *arr2, int paddX, int paddY, int *table)
{
    return run(WF_ED, n1, n2, arr1, arr2, kDefaultTX, kDefaultTY, paddX, paddY, table);
}

#endif
