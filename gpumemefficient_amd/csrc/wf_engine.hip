// wf_engine.hip -- asynchronous strip engine for gfx950 (the default engine).
//
// Same lane geometry as the lock-step engine in wf_kernels.hip (lane l owns
// R rows and computes column s - l at step s; DPP wave_shr:1 passes the
// bottom cell to the next lane), but the waves of a workgroup no longer meet
// at barriers:
//
//   * W compute waves stack vertically (wave w owns rows w*64*R ... of the
//     strip).  Each advances in groups of G steps and publishes its progress
//     in an LDS counter; the wave below starts a group once the counter says
//     the lane-63 values it needs are in the LDS ring.  The check uses the
//     counter read one group earlier, so the read latency overlaps compute,
//     and the steady-state lag between two waves is 64 + 2G steps instead of
//     the two chunks of the lock-step engine.
//   * One extra I/O wave per workgroup owns all global-memory traffic.  It
//     stages the strip's top boundary (the previous strip's bottom row) and
//     the column codes into LDS as soon as the producing workgroup has
//     stored them, two 64-column batches in flight, and stores this strip's
//     bottom row as compute wave W-1 finishes columns.  Compute waves never
//     wait on HBM or L2 latency.
//
// Strip-boundary words in HBM are (launch epoch << 32 | value), written and
// read with relaxed device-scope atomics, so readiness is the tag itself and
// no cache maintenance is needed (the reference's volatile flag + spin,
// LCS/GPU_Hyperlane_Share.cu:102-121, without its missing fences).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wf_cells.h"
#include "wf_variants.h"

namespace wfk {

template <int W, int G, int AW, int RS>
struct AsyncShared {
    // ring[0]: the strip's top input, staged by the I/O wave; ring[w+1]:
    // lane 63 of compute wave w, read by wave w+1 (w = W-1: by the I/O wave).
    // Column x lives in slot (x + 64) mod RS.
    __attribute__((aligned(16))) int ring[W + 1][RS];
    // Lanes 0..62 store their bottom cell here so the capture is one
    // unmasked ds_write per step.
    int dummy[W][64 + G];
    // Column codes: column x at x mod AW, first G slots mirrored at AW.
    uint8_t aw[AW + G];
    // Progress counters, 64 copies each so a wave stores one value with an
    // unmasked ds_write:
    //   cnt[0]     columns staged in ring[0] and aw
    //   cnt[1+w]   steps finished by compute wave w
    //   cnt[W+1]   bottom-row columns stored to HBM by the I/O wave
    int cnt[W + 2][64];
    int task;
    int abort;
};

__device__ __forceinline__ int cnt_load(const int *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void cnt_store(int *p, int v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Ordering of LDS progress counters.  LDS instructions of one wave are
// executed in issue order, so a counter stored after a wave's ring stores
// can only be seen once those stores are visible, and ring loads issued
// after a counter load returned see them; only the compiler has to be kept
// from moving the accesses across the counter (a signal fence).  Building
// with -DWFK_LDS_HW_FENCE=1 uses s_waitcnt-based workgroup fences instead.
#ifndef WFK_LDS_HW_FENCE
#define WFK_LDS_HW_FENCE 1
#endif
__device__ __forceinline__ void lds_release()
{
#if WFK_LDS_HW_FENCE
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
#else
    __atomic_signal_fence(__ATOMIC_RELEASE);
#endif
}
__device__ __forceinline__ void lds_acquire()
{
#if WFK_LDS_HW_FENCE
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
#else
    __atomic_signal_fence(__ATOMIC_ACQUIRE);
#endif
}

// One group of G steps starting at s0 (a multiple of G) for one compute wave.
template <int P, bool CMP, int R, int G, int AW, int RS, bool MASKED, bool TRACK>
__device__ __forceinline__ void run_steps(LaneState<R> &st, int s0, int lane, int n1, const int *ring_prev,
                                          int *ring_cur, int *dummy, const uint8_t *aw)
{
    const int *__restrict__ inj_base = ring_prev + ((s0 + 64) & (RS - 1));
    const uint8_t *__restrict__ cp = aw + ((s0 - lane) & (AW - 1));
    int *cap = (lane == 63) ? ring_cur + ((s0 + 1) & (RS - 1)) : dummy + lane;
    int *cap_last = (lane == 63) ? ring_cur + ((s0 + G) & (RS - 1)) : dummy + lane + (G - 1);
    int4 inj_n = *reinterpret_cast<const int4 *>(inj_base);
    uint32_t ca_n[4] = {cp[0], cp[1], cp[2], cp[3]};
#pragma unroll
    for (int q = 0; q < G; q += 4) {
        const int injv[4] = {inj_n.x, inj_n.y, inj_n.z, inj_n.w};
        const uint32_t cav[4] = {ca_n[0], ca_n[1], ca_n[2], ca_n[3]};
        if (q + 4 < G) {
            inj_n = *reinterpret_cast<const int4 *>(inj_base + q + 4);
#pragma unroll
            for (int u = 0; u < 4; ++u) ca_n[u] = cp[q + 4 + u];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int j = q + u;
            bool valid = true;
            if (MASKED) valid = (unsigned)(s0 + j - lane) < (unsigned)n1;
            step<P, CMP, R, MASKED, TRACK>(st, injv[u], cav[u], valid);
            if (j == G - 1)
                *cap_last = st.h[R - 1];
            else
                cap[j] = st.h[R - 1];
        }
    }
}

template <int P, bool CMP, int R, int W, int G, int AW, int RS>
__device__ void async_strip(const WfKernelArgs &ka, const WfPairDev &pd, int pair, int g,
                            AsyncShared<W, G, AW, RS> &sh, int lane, int wave, uint32_t epoch)
{
    constexpr bool TRACK = (P == WFK_SW);
    const int n1 = pd.n1, n2 = pd.n2;
    const int nsteps = n1 + 63;
    const bool has_prev = g > 0;
    const bool has_next = g + 1 < pd.ngroups;
    const uint64_t tag = (uint64_t)epoch << 32;

    if (wave < W) {
        // ------------------------------------------------------ compute wave
        const int w = wave;
        const int64_t row0 = (int64_t)g * (W * 64 * R) + w * (64 * R) + lane * R;
        const uint8_t *bcodes = ka.codes + pd.b_off;
        LaneState<R> st;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int64_t y = row0 + r;
            const bool v = y < n2;
            st.t[r] = Cell<P, CMP>::row_entry(v ? (uint32_t)bcodes[y] : 0u, v);
            st.h[r] = 0;
        }
        st.dg = 0;
        st.best = 0;

        const int *cin = sh.cnt[w] + lane;        // producer of my input
        const int *cout = sh.cnt[w + 2] + lane;   // consumer of my output ring
        int *cmine = sh.cnt[w + 1] + lane;
        const int ngr = (nsteps + G - 1) / G;
        // Group q reads input columns < (q+1)G: staged (w = 0) or written by
        // lane 63 of wave w-1, which finishes column c at its step c + 63.
        auto need_in = [&](int q) { return w == 0 ? min(n1, (q + 1) * G) : min(nsteps, (q + 1) * G + 63); };
        // Group q overwrites ring slots of columns up to (q+1)G - 64 - RS,
        // which the consumer must have read.
        auto need_out = [&](int q) { return (q + 1) * G - 63 - RS; };

        int vin = cnt_load(cin), vout = cnt_load(cout);
        bool aborted = false;
        for (int q = 0; q < ngr; ++q) {
            if (vin < need_in(q) || vout < need_out(q)) {
                const int64_t t0 = now_ticks();
                for (;;) {
                    __builtin_amdgcn_s_sleep(1);
                    vin = cnt_load(cin);
                    vout = cnt_load(cout);
                    if (vin >= need_in(q) && vout >= need_out(q)) break;
                    if (__hip_atomic_load(&sh.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) ||
                        now_ticks() - t0 > ka.timeout_ticks) {
                        aborted = true;
                        break;
                    }
                }
                if (aborted) break;
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            // Counters for the next group's check, read now so the latency
            // hides behind this group's arithmetic.
            const int vin_n = cnt_load(cin), vout_n = cnt_load(cout);
            const int s0 = q * G;
            const bool steady = (s0 >= 63) && (s0 + G - 1 <= n1 - 1);
            if (steady)
                run_steps<P, CMP, R, G, AW, RS, false, TRACK>(st, s0, lane, n1, sh.ring[w], sh.ring[w + 1],
                                                             sh.dummy[w], sh.aw);
            else
                run_steps<P, CMP, R, G, AW, RS, true, TRACK>(st, s0, lane, n1, sh.ring[w], sh.ring[w + 1],
                                                            sh.dummy[w], sh.aw);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            cnt_store(cmine, min(nsteps, s0 + G));
            vin = vin_n;
            vout = vout_n;
        }
        if (aborted) {
            if (lane == 0) __hip_atomic_store(&sh.abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        } else {
            const int64_t ylast = (int64_t)n2 - 1;
            if (ylast >= row0 && ylast < row0 + R) {
                const int rr = (int)(ylast - row0);
                int v = 0;
#pragma unroll
                for (int r = 0; r < R; ++r)
                    if (r == rr) v = st.h[r];
                if (P == WFK_ED) v += n1 + n2;
                ka.res[pair].last = v;
            }
            if (TRACK) {
                int m = st.best;
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off));
                if (lane == 0 && m > 0) atomicMax(&ka.res[pair].maxs, m);
            }
        }
        return;
    }

    // -------------------------------------------------------------- I/O wave
    const uint8_t *acodes = ka.codes + pd.a_off;
    const uint64_t *bprev = ka.bnd + pd.bnd_off + (int64_t)(g - 1) * n1;
    uint64_t *bnext = ka.bnd + pd.bnd_off + (int64_t)g * n1;
    // Row 0 of the table in the kernels' value representation.
    const int topv = (P == WFK_ED && !(pd.mode & WFK_MODE_TEXTBOOK)) ? -1 : 0;
    const int *c0 = sh.cnt[1] + lane;       // compute wave 0: reads ring[0] and aw
    const int *clast = sh.cnt[W] + lane;    // compute wave W-1: slowest reader of aw, producer of ring[W]

    int in_next = 0;    // columns staged and published in cnt[0]
    int out_done = 0;   // bottom-row columns stored and published in cnt[W+1]
    uint64_t top[2];
    uint32_t code[2];
    int have = 0;       // batches loaded: columns [in_next, in_next + 64*have)
    int staged = 0;     // leading columns of the current batch already staged

    auto load_batch = [&](int base, uint64_t &t, uint32_t &c) {
        const int x = base + lane;
        const bool in = x < n1;
        c = in ? (uint32_t)acodes[x] : 0u;
        t = (has_prev && in) ? __hip_atomic_load(bprev + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : tag;
    };

    // The last strip has no consumer of its bottom row: never block wave W-1.
    if (!has_next) cnt_store(sh.cnt[W + 1] + lane, 1 << 30);

    // Idle time only counts while the compute waves are not moving either,
    // i.e. while everything waits on the producing strip.  The wave sleeps
    // whenever it moved no data, so it takes few issue slots from the compute
    // wave that shares its SIMD.
    int64_t t_idle = now_ticks();
    int seen = -1;
    bool aborted = false;
    for (;;) {
        bool did = false;
        const int p0 = cnt_load(c0), pl = cnt_load(clast);
        const bool moving = (p0 + pl != seen);
        seen = p0 + pl;
        if (in_next < n1) {
            // Room for column in_next + 64*k + 63: its ring[0] slot must have
            // been read by wave 0 (steps < p0 done) and its code-window slot
            // by wave W-1 (reads columns >= pl - 63).
            auto room = [&](int k) {
                const int hi = in_next + 64 * k + 63;
                return hi - RS < p0 && hi - AW < pl - 63 - G;
            };
            while (have < 2 && in_next + 64 * have < n1 && room(have)) {
                load_batch(in_next + 64 * have, top[have], code[have]);
                ++have;
            }
            if (have > 0) {
                const bool stale = (top[0] >> 32) != (uint64_t)epoch;
                const uint64_t smask = __builtin_amdgcn_ballot_w64(stale);
                // Columns [in_next, in_next + ready) of the batch are valid.
                const int ready = smask ? (int)__builtin_ctzll(smask) : 64;
                if (ready > staged) {
                    const int x = in_next + lane;
                    if (lane >= staged && lane < ready && x < n1) {
                        sh.ring[0][(x + 64) & (RS - 1)] = has_prev ? (int)(uint32_t)top[0] : topv;
                        const int slot = x & (AW - 1);
                        sh.aw[slot] = (uint8_t)code[0];
                        if (slot < G) sh.aw[AW + slot] = (uint8_t)code[0];
                    }
                    staged = ready;
                    lds_release();
                    cnt_store(sh.cnt[0] + lane, min(n1, in_next + ready));
                    did = true;
                }
                if (staged == 64 || in_next + staged >= n1) {
                    in_next = min(n1, in_next + 64);
                    staged = 0;
                    top[0] = top[1];
                    code[0] = code[1];
                    --have;
                } else if (stale) {
                    // The producing strip has not stored these yet: re-poll.
                    top[0] = __hip_atomic_load(bprev + in_next + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        }
        if (has_next && out_done < n1) {
            const int avail = min(n1, pl - 63);
            if (avail > out_done && (avail - out_done >= 8 || avail == n1)) {
                const int end = min(avail, out_done + 64);
                const int x = out_done + lane;
                if (x < end) {
                    const int v = sh.ring[W][(x + 64) & (RS - 1)];
                    __hip_atomic_store(bnext + x, tag | (uint32_t)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
                out_done = end;
                lds_release();
                cnt_store(sh.cnt[W + 1] + lane, out_done);
                did = true;
            }
        }
        if (in_next >= n1 && (!has_next || out_done >= n1)) break;
        if (__hip_atomic_load(&sh.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) {
            aborted = true;
            break;
        }
        if (did || moving) t_idle = now_ticks();
        if (!did) {
            __builtin_amdgcn_s_sleep(2);
            if (now_ticks() - t_idle > ka.timeout_ticks) {
                aborted = true;
                if (lane == 0) __hip_atomic_store(&sh.abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                break;
            }
        }
    }
    if (aborted) {
        // Mark the pair and release the strip below with placeholder values
        // so that it does not wait out its own timeout in turn.
        if (lane == 0) atomicOr(&ka.res[pair].status, 1);
        if (has_next)
            for (int x = out_done + lane; x < n1; x += 64)
                __hip_atomic_store(bnext + x, tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

template <int R, int W, int G>
struct AsyncCfg {
    static constexpr int RS = 256;
    static constexpr int AW = pow2_at_least(W * (64 + 2 * G + 16) + 256);
    static_assert(G == 8 || G == 16 || G == 32, "G must divide 64");
};

template <int R, int W, int G>
__global__ __launch_bounds__(64 * (W + 1)) void wf_engine_async(WfKernelArgs ka)
{
    constexpr int AW = AsyncCfg<R, W, G>::AW;
    constexpr int RS = AsyncCfg<R, W, G>::RS;
    __shared__ AsyncShared<W, G, AW, RS> sh;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const uint32_t epoch = __builtin_amdgcn_readfirstlane(*ka.epoch);
    for (;;) {
        if (threadIdx.x == 0) sh.task = ka.task_begin + atomicAdd(ka.counter, 1);
        // Fresh counters for the strip.
        for (int i = threadIdx.x; i < (W + 2) * 64; i += blockDim.x) (&sh.cnt[0][0])[i] = 0;
        if (threadIdx.x == 0) sh.abort = 0;
        __syncthreads();
        const int t = sh.task;
        if (t >= ka.task_end) break;
        const WfTaskDev td = ka.tasks[t];
        const WfPairDev pd = ka.pairs[td.pair];
        const bool cmp = (pd.mode & WFK_MODE_CMP) != 0;
        switch (pd.prob * 2 + (cmp ? 1 : 0)) {
        case WFK_LCS * 2 + 0: async_strip<WFK_LCS, false, R, W, G, AW, RS>(ka, pd, td.pair, td.group, sh, lane, wave, epoch); break;
        case WFK_LCS * 2 + 1: async_strip<WFK_LCS, true, R, W, G, AW, RS>(ka, pd, td.pair, td.group, sh, lane, wave, epoch); break;
        case WFK_ED * 2 + 0: async_strip<WFK_ED, false, R, W, G, AW, RS>(ka, pd, td.pair, td.group, sh, lane, wave, epoch); break;
        case WFK_ED * 2 + 1: async_strip<WFK_ED, true, R, W, G, AW, RS>(ka, pd, td.pair, td.group, sh, lane, wave, epoch); break;
        case WFK_SW * 2 + 0: async_strip<WFK_SW, false, R, W, G, AW, RS>(ka, pd, td.pair, td.group, sh, lane, wave, epoch); break;
        default: async_strip<WFK_SW, true, R, W, G, AW, RS>(ka, pd, td.pair, td.group, sh, lane, wave, epoch); break;
        }
        __syncthreads();
    }
}

}  // namespace wfk

#define WF_ASYNC(R_, W_, G_) \
    {WF_ENGINE_ASYNC, R_, W_, G_, 64 * ((W_) + 1), reinterpret_cast<const void *>(&wfk::wf_engine_async<R_, W_, G_>)}

extern "C" const WfVariant wf_variants_async[] = {
    WF_ASYNC(1, 4, 16), WF_ASYNC(2, 4, 16), WF_ASYNC(4, 4, 16),
    WF_ASYNC(1, 4, 8), WF_ASYNC(2, 4, 8), WF_ASYNC(2, 4, 32),
    WF_ASYNC(1, 2, 16), WF_ASYNC(2, 2, 16), WF_ASYNC(4, 2, 16),
    WF_ASYNC(1, 1, 16), WF_ASYNC(2, 1, 16), WF_ASYNC(4, 1, 16), WF_ASYNC(8, 1, 16),
    WF_ASYNC(1, 8, 16), WF_ASYNC(2, 8, 16), WF_ASYNC(2, 3, 16), WF_ASYNC(2, 7, 16),
};
extern "C" const int wf_num_variants_async = sizeof(wf_variants_async) / sizeof(wf_variants_async[0]);
