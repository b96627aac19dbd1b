// wf_compat.cpp -- the reference's per-problem C++ entry points over the C
// ABI.  Built twice: -DWF_COMPAT_POLICY_SHARE and -DWF_COMPAT_POLICY_CACHE.
#include <cstdio>
#include <cstdlib>

#include "wfdp.h"
#include "wfdp_compat.h"

#if defined(WF_COMPAT_POLICY_CACHE)
static const int kDefaultTX = 256, kDefaultTY = 1024, kPolicyFlags = WF_TILE_CACHE;
#else
static const int kDefaultTX = 32, kDefaultTY = 256, kPolicyFlags = 0;
#endif

static wf_ctx *g_ctx = nullptr;

static void fail(int st)
{
    std::printf("GPU error in %s: %s\n", __FILE__, wf_strerror(st));
    std::exit(EXIT_FAILURE);
}

static int run(int problem, int n1, int n2, const int *arr1, const int *arr2, int tX, int tY)
{
    if (!g_ctx) {
        const int st = wf_ctx_create(0, &g_ctx);
        if (st != WF_OK) fail(st);
    }
    wf_tile t{tX, tY, 0, kPolicyFlags};
    wf_result r{};
    const int st = wf_align(g_ctx, problem, reinterpret_cast<const int32_t *>(arr1), n1,
                            reinterpret_cast<const int32_t *>(arr2), n2, &t, &r);
    if (st != WF_OK) fail(st);
    return r.last;
}

int LCS(int n1, int n2, int *arr1, int *arr2, int, int, int *)
{
    return run(WF_LCS, n1, n2, arr1, arr2, kDefaultTX, kDefaultTY);
}

int LCS(int n1, int n2, int *arr1, int *arr2, int, int, int *, int tX, int tY)
{
    return run(WF_LCS, n1, n2, arr1, arr2, tX, tY);
}

int LCS(int n1, int n2, int *arr1, int *arr2)
{
    return run(WF_LCS, n1, n2, arr1, arr2, kDefaultTX, kDefaultTY);
}

int SW(int n1, int n2, int *arr1, int *arr2, int, int, int *, int tX, int tY)
{
    return run(WF_SW, n1, n2, arr1, arr2, tX, tY);
}

int SW(int n1, int n2, int *arr1, int *arr2, int, int, int *)
{
    return run(WF_SW, n1, n2, arr1, arr2, kDefaultTX, kDefaultTY);
}

int ED(int n1, int n2, int *arr1, int *arr2, int, int, int *, int tX, int tY)
{
    return run(WF_ED, n1, n2, arr1, arr2, tX, tY);
}

int ED(int n1, int n2, int *arr1, int *arr2, int, int, int *)
{
    return run(WF_ED, n1, n2, arr1, arr2, kDefaultTX, kDefaultTY);
}
