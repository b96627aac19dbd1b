// wf_naive.hip -- the reference's naive GPU baselines, rebuilt for gfx950.
//
//   diagonal: one launch per anti-diagonal of cells, one thread per cell,
//             every operand read from and every cell written to the full
//             (n2+1) x (n1+1) int32 table in HBM (LCS/GPU.cu:11-27, 66-102;
//             SmithWaterman/GPU.cu:16-29; EditDistance/GPU.cu:12-31).
//   tile:     one launch per anti-diagonal of 64 x 128 tiles, one workgroup
//             per tile; the tile's top row and left column come from the
//             HBM table, the tile is swept in LDS one cell anti-diagonal at a
//             time with a barrier per level, and written back whole
//             (LCS/GPU_TILE_SHARE.cu:12-80, 140-182).  The reference issues
//             one launch per tile over 32 streams; the tiles of one
//             anti-diagonal are independent, so here they share a launch.
//
// Both keep the whole table in HBM.  They are the baselines the strip
// engines are reported against and a second, independent route to every
// cell of the table, not a product path.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wf_device.h"

namespace wfn {

// The three recurrences on plain int32 cells (the oracle's representation:
// ED distances D, not the engines' shifted form).
template <int P>
__device__ __forceinline__ int cell(int up, int left, int diag, bool eq)
{
    if (P == WFK_LCS) return eq ? diag + 1 : max(up, left);
    if (P == WFK_SW) return max(max(up - 2, left - 2), max(diag + (eq ? 3 : -3), 0));
    return eq ? diag : 1 + min(min(up, left), diag);
}

// Row 0 and column 0.  ED: D(y, 0) = y; D(0, x) = x with the textbook
// boundary (the reference's GPU_setInit), else ED.cpp's x - 1 for x >= 1.
__global__ __launch_bounds__(256) void naive_init(int *t, int n1, int n2, int prob, int textbook)
{
    const int64_t st = (int64_t)n1 + 1;
    const int stride = gridDim.x * blockDim.x;
    for (int x = blockIdx.x * blockDim.x + threadIdx.x; x <= n1; x += stride)
        t[x] = (prob == WFK_ED) ? (textbook || x == 0 ? x : x - 1) : 0;
    for (int y = blockIdx.x * blockDim.x + threadIdx.x; y <= n2; y += stride)
        if (y > 0) t[(int64_t)y * st] = (prob == WFK_ED) ? y : 0;
}

// Cells (y, x) with x + y = d, rows ylo .. ylo + cnt - 1 (1-based table
// coordinates, all inside the table).
template <int P>
__global__ __launch_bounds__(256) void naive_diag(int *t, const int32_t *a, const int32_t *b, int n1, int d,
                                                  int ylo, int cnt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt) return;
    const int y = ylo + i, x = d - y;
    const int64_t st = (int64_t)n1 + 1;
    int *c = t + (int64_t)y * st + x;
    c[0] = cell<P>(c[-st], c[-1], c[-st - 1], a[x - 1] == b[y - 1]);
}

constexpr int kTX = 64, kTY = 128;

// Tiles (tx, ty) with tx + ty = k, ty = tylo + blockIdx.x.
template <int P>
__global__ __launch_bounds__(256) void naive_tile(int *t, const int32_t *a, const int32_t *b, int n1, int n2, int k,
                                                  int tylo)
{
    constexpr int SX = kTX + 1;
    __shared__ int s[(kTY + 1) * SX];
    __shared__ int32_t sa[kTX], sb[kTY];
    const int ty = tylo + blockIdx.x, tx = k - ty;
    const int x0 = tx * kTX, y0 = ty * kTY;
    const int w = min(kTX, n1 - x0), h = min(kTY, n2 - y0);
    const int64_t st = (int64_t)n1 + 1;
    const int *tt = t + (int64_t)y0 * st + x0;
    for (int i = threadIdx.x; i <= w; i += blockDim.x) s[i] = tt[i];
    for (int j = threadIdx.x + 1; j <= h; j += blockDim.x) s[j * SX] = tt[(int64_t)j * st];
    for (int i = threadIdx.x; i < w; i += blockDim.x) sa[i] = a[x0 + i];
    for (int j = threadIdx.x; j < h; j += blockDim.x) sb[j] = b[y0 + j];
    __syncthreads();
    for (int lev = 2; lev <= w + h; ++lev) {
        const int jlo = max(1, lev - w), jhi = min(h, lev - 1);
        for (int j = jlo + (int)threadIdx.x; j <= jhi; j += blockDim.x) {
            const int i = lev - j;
            int *c = s + j * SX + i;
            c[0] = cell<P>(c[-SX], c[-1], c[-SX - 1], sa[i - 1] == sb[j - 1]);
        }
        __syncthreads();
    }
    int *to = t + (int64_t)y0 * st + x0;
    for (int idx = threadIdx.x; idx < w * h; idx += blockDim.x) {
        const int j = idx / w + 1, i = idx % w + 1;
        to[(int64_t)j * st + i] = s[j * SX + i];
    }
}

// Maximum cell of the table (Smith-Waterman's local score); cells are >= 0.
__global__ __launch_bounds__(256) void naive_max(const int *t, int64_t n, int *out)
{
    int m = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        m = max(m, t[i]);
    for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

template <int P>
hipError_t run_diag(int *t, const int32_t *a, const int32_t *b, int n1, int n2, hipStream_t stream)
{
    for (int d = 2; d <= n1 + n2; ++d) {
        const int ylo = max(1, d - n1), yhi = min(n2, d - 1);
        const int cnt = yhi - ylo + 1;
        hipLaunchKernelGGL(naive_diag<P>, dim3((cnt + 255) / 256), dim3(256), 0, stream, t, a, b, n1, d, ylo, cnt);
    }
    return hipGetLastError();
}

template <int P>
hipError_t run_tile(int *t, const int32_t *a, const int32_t *b, int n1, int n2, hipStream_t stream)
{
    const int nx = (n1 + kTX - 1) / kTX, ny = (n2 + kTY - 1) / kTY;
    for (int k = 0; k <= nx + ny - 2; ++k) {
        const int tylo = max(0, k - (nx - 1)), tyhi = min(ny - 1, k);
        hipLaunchKernelGGL(naive_tile<P>, dim3(tyhi - tylo + 1), dim3(256), 0, stream, t, a, b, n1, n2, k, tylo);
    }
    return hipGetLastError();
}

}  // namespace wfn

// Computes the whole table t ((n2+1) x (n1+1) int32, row stride n1+1) of
// one pair with the naive method (0 diagonal, 1 tile) on `stream`; `smax`
// (one device int, or null) receives the maximum cell.  n1, n2 >= 1.
extern "C" hipError_t wf_launch_naive(int method, int prob, int textbook, int *t, const int32_t *a,
                                      const int32_t *b, int n1, int n2, int *smax, hipStream_t stream)
{
    hipLaunchKernelGGL(wfn::naive_init, dim3(64), dim3(256), 0, stream, t, n1, n2, prob, textbook);
    hipError_t e;
    if (method == 0) {
        e = prob == WFK_LCS ? wfn::run_diag<WFK_LCS>(t, a, b, n1, n2, stream)
            : prob == WFK_SW ? wfn::run_diag<WFK_SW>(t, a, b, n1, n2, stream)
                             : wfn::run_diag<WFK_ED>(t, a, b, n1, n2, stream);
    } else {
        e = prob == WFK_LCS ? wfn::run_tile<WFK_LCS>(t, a, b, n1, n2, stream)
            : prob == WFK_SW ? wfn::run_tile<WFK_SW>(t, a, b, n1, n2, stream)
                             : wfn::run_tile<WFK_ED>(t, a, b, n1, n2, stream);
    }
    if (e != hipSuccess) return e;
    if (smax) {
        const int64_t n = ((int64_t)n1 + 1) * ((int64_t)n2 + 1);
        hipLaunchKernelGGL(wfn::naive_max, dim3(1024), dim3(256), 0, stream, t, n, smax);
    }
    return hipGetLastError();
}
