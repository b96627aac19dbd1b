// wf_stencil.hip -- gfx950 wavefront engine for the reference's other users of
// the hyperlane schedule: in-place sweeps over a whole int32 table whose cell
// depends on its left and upper neighbours after they were updated.
//
//   SAT  summed-area table:
//          T(y,x) = T(y,x-1) + T(y-1,x) + T(y,x) - T(y-1,x-1)
//        (SAT/GPU_Hyperlane_Share.cu:169, :196, :233), int32 wrap-around.
//   SOR  one Gauss-Seidel sweep of the 5-point stencil:
//          T(y,x) = (T(y,x-1) + T(y-1,x) + T(y,x) + T(y,x+1) + T(y+1,x)) / 5
//        with left and up already updated and right and down not
//        (2D-SOR/nested_loop/GPU_Hyperlane_Share.cu:169, :197, :233); C
//        division truncating toward zero.
//
// The table has the reference's layout (SAT/GPU_Hyperlane_Share.cu:255-263):
// rows 0 .. n2+1 of pitch ld >= n1+3; row 0, row n2+1 and columns 0, n1+1,
// n1+2 are boundary values that are read, never written; the interior
// 1 <= y <= n2, 1 <= x <= n1 is updated.  Any order that visits each cell
// after its left and upper neighbours and before its right and lower ones
// gives the row-major sequential sweep's result; the reference's hyperlane
// tiles are one such order and so is this engine's.
//
// Geometry.  A strip of 64*W rows is swept left to right by one workgroup of
// W waves; lane l of a wave owns one row and computes column x = s + 1 - l at
// step s (the systolic front of the alignment engines).  The updated left
// value stays in a VGPR, the updated upper value comes from lane l-1 through
// one DPP wave_shr:1 (lane 0: from an LDS ring that the wave above, or for the
// strip's first wave the strip above through HBM, fills).  Old values (self,
// right, and the row below) come from an LDS window of the wave's 64 rows plus
// the row below, staged from HBM in 32-column blocks with coalesced 16-byte
// loads issued two chunks ahead of use; each new value is written back into the
// window in place, and finished blocks go to HBM with coalesced 16-byte
// stores.  Every cell is read once from and written once to HBM: the sweep
// is bound by HBM bandwidth, not by the VALU.
//
// Waves of a workgroup advance in lock-step chunks of K = 32 steps between
// s_barriers, wave w D = 3 chunks behind wave w-1.  A strip hands its bottom
// row (columns 0..n1) to the next strip through HBM as (launch epoch << 32 |
// value) words written and polled with relaxed device-scope atomics, as the
// alignment engines do; strips are claimed in order from an atomic counter,
// so a strip only waits on one that a running workgroup holds.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wf_cells.h"
#include "wf_device.h"

namespace wfs {

using wfk::kDppWaveShr1;
using wfk::now_ticks;

constexpr int kK = 32;                  // steps per chunk
constexpr int kD = (kK + 62) / kK + 1;  // chunks between consecutive waves
constexpr int kTW = 128;                // window columns per row (4 blocks of 32)
constexpr int kTS = kTW + 8;            // row stride in dwords: slot 128 mirrors slot 0
constexpr int kTR = 65;                 // window rows: the wave's 64 plus the row below
constexpr int kRS = 256;                // ring slots; column x lives in slot (x + 63) mod kRS
constexpr int kVec = kTR * 8;           // 16-byte vectors in one block of the window

template <int W>
struct StShared {
    // Window rows r = 0..64 of each wave; column x in slot x mod 128.  The
    // stride of 136 dwords makes the lanes' skewed reads (column s + 1 - l in
    // row l) fall in distinct banks.
    __attribute__((aligned(16))) int win[W][kTR * kTS];
    // ring[0]: the strip's top row (wave 0's lane 0 input); ring[w+1]: lane
    // 63 of wave w, read by wave w+1 (w = W-1: copied to HBM).
    __attribute__((aligned(16))) int ring[W + 1][kRS];
    // Lanes 0..62 store their value here so the ring capture is one unmasked
    // ds_write per step.
    int dummy[W][64 + kK];
    int task;
    int abort_at;
};

// Barrier between the waves of the workgroup that orders their LDS accesses
// only: global loads issued before it stay in flight (a __syncthreads would
// wait for them, and for the flushes' stores, at every chunk).
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// One block of the window (columns [32b, 32b+32) of the wave's 65 rows) in
// flight between its global loads and its LDS stores: vector v = lane + 64k
// is row v / 8, columns 32b + 4 (v % 8) .. + 3.
struct Stage {
    int4 v[(kVec + 63) / 64];
};

// The wave's 65 window rows as a buffer resource: loads past its last row
// (rows below the table) read zeros, so every block load is issued
// unconditionally, with no per-lane branch (see the chunk loop).  Offsets
// are 32-bit: 73 rows of the pitch must fit (checked by the host).
// The descriptor's fields pass through readfirstlane: a descriptor the
// compiler cannot prove uniform lives in VGPRs, and every buffer load then
// becomes a waterfall loop (readfirstlane, compare, exec mask, load, repeat).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t window_rsrc(const WfStencilArgs &a, int yw)
{
    const int rows = max(0, min(kTR, a.n2 + 2 - yw));
    const uint64_t base = reinterpret_cast<uint64_t>(a.t + (int64_t)yw * a.ld);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)base);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(base >> 32));
    int *p = reinterpret_cast<int *>(((uint64_t)hi << 32) | lo);
    const int bytes = __builtin_amdgcn_readfirstlane((int)(rows * a.ld * 4));
    // Third descriptor dword of an untyped buffer on the gfx9 family.
    return __builtin_amdgcn_make_buffer_rsrc(p, 0, bytes, 0x00020000);
}

__device__ __forceinline__ void stage_load(Stage &sg, __amdgpu_buffer_rsrc_t rs, int64_t ld, int b, int lane)
{
#pragma unroll
    for (int k = 0; k < (kVec + 63) / 64; ++k) {
        const int v = lane + 64 * k;
        // Row v / 8 of the window; a vector past the pitch reads the next
        // row's first columns (never used: the table's columns end at n1 + 2).
        const uint32_t off = (uint32_t)(((v >> 3) * ld + 32 * b + 4 * (v & 7)) * 4);
        sg.v[k] = __builtin_bit_cast(int4, __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0));
    }
}

__device__ __forceinline__ void stage_store(const Stage &sg, int *win, int b, int lane)
{
#pragma unroll
    for (int k = 0; k < (kVec + 63) / 64; ++k) {
        const int v = lane + 64 * k;
        if (k < kVec / 64 || v < kVec) {
            const int slot = (32 * b + 4 * (v & 7)) & (kTW - 1);
            *reinterpret_cast<int4 *>(win + (v >> 3) * kTS + slot) = sg.v[k];
        }
    }
    // A block at slot 0 is mirrored at slots 128..131 (the vectors of lanes
    // with v % 8 == 0); the other lanes rewrite their own vector, so the
    // stores need no per-lane branch.
    if ((b & 3) == 0) {
#pragma unroll
        for (int k = 0; k < (kVec + 63) / 64; ++k) {
            const int v = lane + 64 * k;
            if (k < kVec / 64 || v < kVec) {
                const int slot = (v & 7) == 0 ? kTW : 4 * (v & 7);
                *reinterpret_cast<int4 *>(win + (v >> 3) * kTS + slot) = sg.v[k];
            }
        }
    }
}

// Store block b of the wave's interior rows (r < nrows) back to the table
// through a buffer descriptor over those rows: a vector outside them, or
// holding no interior column, gets an offset past the descriptor's end and
// the hardware drops it (boundary columns inside a stored vector are written
// back unchanged).  With no branch around the stores the eight LDS reads are
// issued together, not one round trip per vector.
__device__ __forceinline__ void flush_block(const WfStencilArgs &a, const int *win, int b, int nrows, int lane,
                                            __amdgpu_buffer_rsrc_t out)
{
    if (b < 0 || 32 * b > a.n1) return;
    int4 val[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int v = lane + 64 * k;
        val[k] = *reinterpret_cast<const int4 *>(win + (v >> 3) * kTS + ((32 * b + 4 * (v & 7)) & (kTW - 1)));
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int v = lane + 64 * k;
        const int rr = v >> 3;
        const int c = 32 * b + 4 * (v & 7);
        const uint32_t off = (rr < nrows && c + 3 >= 1 && c <= a.n1) ? (uint32_t)((rr * a.ld + c) * 4) : 0xfffffff0u;
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(__attribute__((ext_vector_type(4))) uint32_t, val[k]),
                                               out, off, 0, 0);
    }
}

// The wave's interior rows (nrows of them from yw) as the descriptor of
// flush_block, its fields through readfirstlane as in window_rsrc.  The 4
// bytes past the last row cover the one dword a stored vector can reach
// beyond its row when the pitch is not a multiple of 4: the next row's
// column 0, a boundary value written back unchanged.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t interior_rsrc(const WfStencilArgs &a, int yw, int nrows)
{
    const uint64_t base = reinterpret_cast<uint64_t>(a.t + (int64_t)yw * a.ld);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)base);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(base >> 32));
    int *p = reinterpret_cast<int *>(((uint64_t)hi << 32) | lo);
    const int bytes = __builtin_amdgcn_readfirstlane(nrows > 0 ? (int)(nrows * a.ld * 4 + 4) : 0);
    return __builtin_amdgcn_make_buffer_rsrc(p, 0, bytes, 0x00020000);
}

// One chunk of kK steps from s0 for one wave.  `h` is the lane's last updated
// value (its left neighbour), `dg` the upper neighbour of the previous step
// (SAT's diagonal).  MASKED: some lane is outside its columns 1..n1 or holds
// no interior row; those lanes keep their state and write nothing.
//
// The old values a group of 4 steps needs (own row: columns x..x+4; row
// below: x..x+3), and lane 0's upper neighbours from the ring, are read one
// group ahead, before the current group writes
// its new values: the reads are of columns no lane has updated yet, and the
// LDS latency overlaps a group of arithmetic.  Reads may run past slot 127
// into the mirror of slots 0..3 (old values, which is what they need);
// writes always go to the slot itself.
template <int OP, bool MASKED>
__device__ __forceinline__ void st_chunk(int s0, int lane, int n1, bool rowok, int &h, int &dg, int *wrow,
                                         const int *ring_prev, int *ring_cur, int *dummy)
{
    constexpr int NO = (OP == WFS_SOR) ? 5 : 4;  // own-row values per group
    constexpr int ND = (OP == WFS_SOR) ? 4 : 1;  // row-below values per group (SAT: none)
    const int *__restrict__ inj_base = ring_prev + ((s0 + 64) & (kRS - 1));
    int *cap = (lane == 63) ? ring_cur + ((s0 + 1) & (kRS - 1)) : dummy + lane;
    int *cap_last = (lane == 63) ? ring_cur + ((s0 + kK) & (kRS - 1)) : dummy + lane + (kK - 1);
    const int xb = s0 + 1 - lane;  // the lane's column at step s0
    int own[NO], dn[ND];
    int4 inj4;
    auto load_ops = [&](int q, int (&o)[NO], int (&d)[ND], int4 &in4) {
        in4 = *reinterpret_cast<const int4 *>(inj_base + q);
        const int p0 = (xb + q) & (kTW - 1);
#pragma unroll
        for (int u = 0; u < NO; ++u) o[u] = wrow[p0 + u];
        if (OP == WFS_SOR) {
#pragma unroll
            for (int u = 0; u < ND; ++u) d[u] = wrow[kTS + p0 + u];
        } else {
            d[0] = 0;
        }
    };
    load_ops(0, own, dn, inj4);
#pragma unroll
    for (int q = 0; q < kK; q += 4) {
        const int injv[4] = {inj4.x, inj4.y, inj4.z, inj4.w};
        int nown[NO], ndn[ND];
        int4 ninj4 = inj4;
        int capv[4];
        if (q + 4 < kK) load_ops(q + 4, nown, ndn, ninj4);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int j = q + u;
            const int x = xb + j;
            const int self = own[u];
            const int up = __builtin_amdgcn_update_dpp(injv[u], h, kDppWaveShr1, 0xf, 0xf, false);
            int nv;
            if (OP == WFS_SAT) {
                nv = (int)((uint32_t)h + (uint32_t)up + (uint32_t)self - (uint32_t)dg);
            } else {
                const int right = own[u + 1];
                const int down = dn[u];
                const int sum = (int)((uint32_t)h + (uint32_t)up + (uint32_t)self + (uint32_t)right +
                                      (uint32_t)down);
                nv = sum / 5;
            }
            bool valid = true;
            if (MASKED) valid = rowok && (unsigned)(x - 1) < (unsigned)n1;
            if (valid) wrow[x & (kTW - 1)] = nv;
            h = valid ? nv : h;
            dg = up;
            capv[u] = h;
        }
        // The group's ring captures together (adjacent slots: paired writes).
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int j = q + u;
            if (j == kK - 1)
                *cap_last = capv[u];
            else
                cap[j] = capv[u];
        }
        if (q + 4 < kK) {
            inj4 = ninj4;
#pragma unroll
            for (int u = 0; u < NO; ++u) own[u] = nown[u];
#pragma unroll
            for (int u = 0; u < ND; ++u) dn[u] = ndn[u];
        }
    }
}

template <int OP, int W>
__device__ void stencil_strip(const WfStencilArgs &a, int g, StShared<W> &sh, int lane, int wave, uint32_t epoch)
{
    const int n1 = a.n1;
    const int yw = 1 + g * 64 * W + wave * 64;  // the wave's first row
    const int rows_left = a.n2 - g * 64 * W;     // interior rows from the strip's first
    const int wa = min(W, (rows_left + 63) / 64);
    const bool active = wave < wa;
    const int nrows = active ? min(64, a.n2 - yw + 1) : 0;  // interior rows of the wave
    const bool rowok = lane < nrows;
    const int nsteps = n1 + 63;                 // lane 63 updates column n1 at step n1 + 62
    const int nch = (nsteps + kK - 1) / kK;
    const int total = nch + (wa - 1) * kD;
    const int nfill = (n1 + 1 + kK - 1) / kK;   // ring chunks covering columns 0..n1
    const bool has_prev = g > 0;
    const bool has_next = g + 1 < a.nstrips;
    const uint64_t *bprev = a.bnd + (int64_t)(g - 1) * (n1 + 1);
    uint64_t *bnext = a.bnd + (int64_t)g * (n1 + 1);
    const uint64_t tag = (uint64_t)epoch << 32;
    int *win = sh.win[wave];
    int *wrow = win + lane * kTS;
    const __amdgpu_buffer_rsrc_t rs = window_rsrc(a, yw);
    const __amdgpu_buffer_rsrc_t outr = interior_rsrc(a, yw, nrows);

    // Top row (columns [32c, 32c+32), wave 0, lanes < 32): row 0 of the table
    // for the first strip, else the previous strip's bottom row.
    auto load_top = [&](int c) -> uint64_t {
        const int x = c * kK + lane;
        if (lane >= kK || x > n1) return tag;
        if (!has_prev)  // row 0 is never written: any cache level may serve it
            return tag | (uint32_t)__hip_atomic_load(a.t + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        return __hip_atomic_load(bprev + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    };
    auto fill = [&](int c, uint64_t top) -> bool {
        const int x = c * kK + lane;
        const bool mine = lane < kK && x <= n1;
        if (__builtin_amdgcn_ballot_w64(mine && (top >> 32) != (uint64_t)epoch) != 0) {
            const int64_t t0 = now_ticks();
            for (;;) {
                __builtin_amdgcn_s_sleep(1);
                if (mine && (top >> 32) != (uint64_t)epoch)
                    top = __hip_atomic_load(bprev + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (__builtin_amdgcn_ballot_w64(mine && (top >> 32) != (uint64_t)epoch) == 0) break;
                if (now_ticks() - t0 > a.timeout_ticks) return false;
            }
        }
        if (mine) sh.ring[0][(x + 63) & (kRS - 1)] = (int)(uint32_t)top;
        return true;
    };

    // Prologue: block 0 of the window in LDS, blocks 1 and 2 in flight (block
    // b travels in sg[b & 1], two chunks ahead of its LDS store); the top
    // ring's first two chunks.
    Stage sg0, sg1;
    if (active) {
        stage_load(sg0, rs, a.ld, 0, lane);
        stage_store(sg0, win, 0, lane);
        stage_load(sg1, rs, a.ld, 1, lane);
        stage_load(sg0, rs, a.ld, 2, lane);
    }
    // Top-row loads run two chunks ahead of their ring fill (top0 / top1
    // alternate like the blocks), so an HBM round trip hides behind two
    // chunks of arithmetic.
    uint64_t top0 = tag, top1 = tag;
    if (wave == 0) {
        bool ok = fill(0, load_top(0));
        if (ok && nfill > 1) ok = fill(1, load_top(1));
        if (!ok && lane == 0) sh.abort_at = -1;
        top1 = load_top(2);
    }
    __syncthreads();

    int h = active ? wrow[0] : 0;  // column 0: the left boundary
    int dg = 0;
    int copied = 0;  // wave W-1: bottom-row columns stored to bnext
    auto copy_out = [&](int upto, bool from_ring) {
        while (copied < upto) {
            const int x = copied + lane;
            if (x < upto) {
                const int v = from_ring ? sh.ring[W][(x + 63) & (kRS - 1)] : 0;
                __hip_atomic_store(bnext + x, tag | (uint32_t)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            copied = min(upto, copied + 64);
        }
    };

    // Every wave passes the same `total` barriers: kD idle chunks per wave
    // above it, its nch chunks, then idle chunks until the last wave is done.
    // The active chunks run in a loop of their own so that their nine block
    // loads are issued on every iteration: the compiler then waits for a
    // block's loads (two chunks old) without waiting for the newer ones.
    bool aborted = sh.abort_at <= -1;
    int j = 0;
    const int jstart = active ? wave * kD : total;
    for (; j < jstart && !aborted; ++j) {
        lds_barrier();
        aborted = sh.abort_at <= j;
    }
    // One active chunk; `sg` holds block cj+1 (loaded two chunks ago) and
    // takes block cj+3; `top_old` holds the top row's chunk cj+2 (loaded
    // during the previous chunk), `top_new` takes chunk cj+3.
    auto chunk = [&](int cj, Stage &sg, uint64_t &top_new, uint64_t &top_old) {
        const int s0 = cj * kK;
        if (wave == 0 && cj + 3 < nfill) top_new = load_top(cj + 3);
        // Keep the top-row load ahead of the block loads (the compiler would
        // otherwise sink it toward its use, where waiting for it would also
        // wait for the block loads issued after it).
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        // Window: block cj+1 takes the slot of block cj-3, flushed at the end
        // of the previous chunk.
        stage_store(sg, win, cj + 1, lane);
        stage_load(sg, rs, a.ld, cj + 3, lane);
        if (cj == 0 && lane == 0) dg = sh.ring[wave][63];  // column 0 of the row above
        const bool steady = nrows == 64 && s0 >= 63 && s0 + kK - 1 <= n1 - 1;
        if (steady)
            st_chunk<OP, false>(s0, lane, n1, rowok, h, dg, wrow, sh.ring[wave], sh.ring[wave + 1], sh.dummy[wave]);
        else
            st_chunk<OP, true>(s0, lane, n1, rowok, h, dg, wrow, sh.ring[wave], sh.ring[wave + 1], sh.dummy[wave]);
        // Steps < s0 + kK are done: lane 63 has finished columns < s0 + kK - 62,
        // so every lane has finished block cj-2.
        if (wave == W - 1 && has_next) copy_out(min(n1 + 1, s0 + kK - 62), true);
        flush_block(a, win, cj - 2, nrows, lane, outr);
        if (wave == 0 && cj + 2 < nfill) {
            if (!fill(cj + 2, top_old) && lane == 0) sh.abort_at = j;
        }
        lds_barrier();
        aborted = sh.abort_at <= j;
        ++j;
    };
    // Even chunks stage from sg1, odd ones from sg0; the loop takes them in
    // pairs so that the alternation is visible to the compiler.
    for (int cj = 0; cj < nch && !aborted; cj += 2) {
        chunk(cj, sg1, top0, top1);
        if (cj + 1 >= nch || aborted) break;
        chunk(cj + 1, sg0, top1, top0);
    }
    for (; j < total && !aborted; ++j) {
        lds_barrier();
        aborted = sh.abort_at <= j;
    }

    if (aborted) {
        // A dependency wait timed out: mark the launch and release the strip
        // below with placeholder values so it does not wait out its own timeout.
        if (threadIdx.x == 0) atomicOr(&a.res->status, 1);
        if (wave == W - 1 && has_next) copy_out(n1 + 1, false);
        return;
    }
    if (wave == W - 1 && has_next) copy_out(n1 + 1, true);
    // Blocks not yet flushed: nch-2 .. the one holding column n1 (all of
    // them are still in the window).
    if (active)
        for (int b = max(0, nch - 2); 32 * b <= n1; ++b) flush_block(a, win, b, nrows, lane, outr);
}

template <int OP, int W>
__global__ __launch_bounds__(64 * W) void wf_stencil_kernel(WfStencilArgs a)
{
    __shared__ StShared<W> sh;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t epoch = __builtin_amdgcn_readfirstlane(*a.epoch);
    for (;;) {
        if (threadIdx.x == 0) {
            sh.task = atomicAdd(a.counter, 1);
            sh.abort_at = 0x7fffffff;
        }
        __syncthreads();
        const int g = __builtin_amdgcn_readfirstlane(sh.task);
        __syncthreads();
        if (g >= a.nstrips) break;
        stencil_strip<OP, W>(a, g, sh, lane, wave, epoch);
        __syncthreads();
    }
}

}  // namespace wfs

static const void *stencil_fn(int op, int waves)
{
    if (op == WFS_SAT) {
        if (waves == 1) return reinterpret_cast<const void *>(&wfs::wf_stencil_kernel<WFS_SAT, 1>);
        if (waves == 2) return reinterpret_cast<const void *>(&wfs::wf_stencil_kernel<WFS_SAT, 2>);
        if (waves == 4) return reinterpret_cast<const void *>(&wfs::wf_stencil_kernel<WFS_SAT, 4>);
    } else if (op == WFS_SOR) {
        if (waves == 1) return reinterpret_cast<const void *>(&wfs::wf_stencil_kernel<WFS_SOR, 1>);
        if (waves == 2) return reinterpret_cast<const void *>(&wfs::wf_stencil_kernel<WFS_SOR, 2>);
        if (waves == 4) return reinterpret_cast<const void *>(&wfs::wf_stencil_kernel<WFS_SOR, 4>);
    }
    return nullptr;
}

extern "C" int wf_stencil_resident(int op, int waves, int cus)
{
    const void *fn = stencil_fn(op, waves);
    if (!fn) return 0;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 64 * waves, 0) != hipSuccess || per_cu < 1) {
        (void)hipGetLastError();
        per_cu = 1;
    }
    return per_cu * cus;
}

extern "C" hipError_t wf_launch_stencil(int op, int waves, int grid, const WfStencilArgs *a, hipStream_t stream)
{
    const void *fn = stencil_fn(op, waves);
    if (!fn || grid <= 0) return hipErrorInvalidValue;
    void *args[] = {const_cast<WfStencilArgs *>(a)};
    return hipLaunchKernel(fn, dim3(grid), dim3(64 * waves), args, 0, stream);
}
