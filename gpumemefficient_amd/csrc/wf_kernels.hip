// wf_kernels.hip -- gfx950 wavefront kernels for LCS, edit distance and
// Smith-Waterman (the reference's hyperlane engine, LCS/GPU_Hyperlane_Share.cu
// :123-687, SmithWaterman/GPU_Hyperlane_Share.cu:127-674,
// EditDistance/GPU_Hyperlane_Share.cu:131-680, redesigned for CDNA4).
//
// Geometry.  A pair's DP table has n2 rows (b) and n1 columns (a).  Rows are
// cut into strips of W*64*R rows; one workgroup of W wave64s sweeps a strip
// left to right.  Inside a wave, lane l owns R consecutive rows and at step s
// computes column x = s - l of all of them: the lanes form a systolic array
// whose front is a parallelogram (the reference's "hyperlane" tile), so every
// lane is busy except in the first and last 63 steps of a strip.  A step
// passes the lane's bottom cell to the next lane with one DPP wave_shr:1
// (no LDS, no shuffle); lane 0 takes its upper neighbour from an LDS ring
// written by the wave above.  The tile body never leaves VGPRs.
//
// Waves of a workgroup advance in lock-step chunks of K steps separated by
// one s_barrier; wave w runs D chunks behind wave w-1, which is exactly the
// time lane 63 of wave w-1 needs to produce the columns lane 0 of wave w
// reads.  The strip's bottom row (lane 63 of the last wave) is the only data
// that goes to HBM: after every chunk its finished columns are stored as
// 64-bit (launch epoch, value) words with relaxed device-scope stores, and
// the next strip's first wave prefetches them one chunk ahead and re-polls
// any whose tag is not yet this launch's epoch.  This replaces the
// reference's one-launch-per-strip + volatile flag scheme
// (LCS/GPU_Hyperlane_Share.cu:102-121, 741-767) with one persistent launch
// that claims strips from an atomic counter in dependency order.
//
// Cell functions are the reference recurrences rewritten so that each costs
// 3-4 VALU ops with a per-row lookup table indexed by the column symbol:
//   LCS  H = max3(up, left, diag + eq)          (LCS/LCS.cpp:58-63)
//   ED   E = min3(E_up, E_left, E_diag + ne - 2), E = D - x - y
//                                               (EditDistance/ED.cpp:61-63)
//   SW   H = usat(max3(up, left, diag + s + 2) - 2)
//                                               (SmithWaterman/SW.cpp:63-65)
// The ED and LCS rewrites rely on adjacent cells differing by at most one;
// the oracle tests check them against the original recurrences.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wf_device.h"

#include "wf_cells.h"

namespace wfk {

// LDS workspace of one workgroup.
template <int W, int K, int AW>
struct WgShared {
    static constexpr int RS = ChunkGeom<K>::RS;
    // ring[0]: this strip's top input (lane 0 of wave 0); ring[w+1]: bottom
    // lane of wave w, read by wave w+1 (and by the copy-out for w = W-1).
    __attribute__((aligned(16))) int ring[W + 1][RS];
    // Lanes 0..62 store their bottom cells here (lane l at 4l plus the
    // group's offset) so that a 4-step group's capture is one unmasked
    // ds_write_b128; only lane 63's store lands in a ring.
    __attribute__((aligned(16))) int dummy[W][64 * 4 + K];
    // Window of column codes, one per dword (aligned dword reads are the
    // cheapest LDS reads for the lane-skewed access): column x lives at
    // x mod AW, the first K slots are mirrored after the end so a chunk never
    // wraps.
    uint32_t aw[AW + K];
    int task;
    // Chunk index at which wave 0 hit a dependency timeout (INT_MAX: none).
    // Waves test `abort_at <= j` after barrier j; wave 0 can only store a
    // value > j before the next barrier, so every wave leaves the chunk loop
    // at the same barrier.
    int abort_at;
};

// Table-dump destination of one lane (DUMP kernels only): the lane's first
// row and the pair's row count; values are written in the reference's
// representation (ED: D = E + x + y with 1-based table indices).
struct DumpRows {
    int32_t *table;
    int64_t row0;
    int n2;
};

// One chunk of K steps starting at s0 (a multiple of K) for one wave.
template <int P, bool CMP, int R, int K, int AW, bool MASKED, bool TRACK, bool DUMP = false>
__device__ __forceinline__ void run_chunk(LaneState<R> &st, int s0, int lane, int n1,
                                          const int *ring_prev, int *ring_cur, int *dummy,
                                          const uint32_t *aw, int &pend, DumpRows dr = DumpRows{})
{
    constexpr int RS = ChunkGeom<K>::RS;
    const int *__restrict__ inj_base = ring_prev + ((s0 + 64) & (RS - 1));
    const uint32_t *__restrict__ cp = aw + ((s0 - lane) & (AW - 1));
    // Lane 63's bottom cell of step s belongs in ring slot s + 1.  After the
    // 4-step group at S the values of steps S-1..S+2 go to slots S..S+3 in
    // one 16-byte store; step S+3's value is carried in `pend` and stored
    // alone at the chunk's end, before the barrier that hands the chunk to
    // the wave below.
    int *capb = (lane == 63) ? ring_cur + (s0 & (RS - 1)) : dummy + lane * 4;
    // Software pipeline: the LDS reads of group q+1 (4 column codes and the
    // 4 top-neighbour values of lane 0) are issued before group q's capture
    // stores, so their latency hides behind group q's arithmetic.
    int4 inj_n = *reinterpret_cast<const int4 *>(inj_base);
    uint32_t ca_n[4] = {cp[0], cp[1], cp[2], cp[3]};
#pragma unroll
    for (int q = 0; q < K; q += 4) {
        const int injv[4] = {inj_n.x, inj_n.y, inj_n.z, inj_n.w};
        const uint32_t cav[4] = {ca_n[0], ca_n[1], ca_n[2], ca_n[3]};
        if (q + 4 < K) {
            inj_n = *reinterpret_cast<const int4 *>(inj_base + q + 4);
#pragma unroll
            for (int u = 0; u < 4; ++u) ca_n[u] = cp[q + 4 + u];
        }
        // The next group's reads stay here, a group ahead of their use.
        __builtin_amdgcn_sched_barrier(0x207);
        int capv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int j = q + u;
            bool valid = true;
            if (MASKED) valid = (unsigned)(s0 + j - lane) < (unsigned)n1;
            step<P, CMP, R, MASKED, TRACK>(st, injv[u], cav[u], valid);
            capv[u] = st.h[R - 1];
            if (DUMP) {
                const int x = s0 + j - lane;
                if ((unsigned)x < (unsigned)n1) {
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const int64_t y = dr.row0 + r;
                        if (y < dr.n2)
                            dr.table[y * n1 + x] = (P == WFK_ED) ? st.h[r] + x + (int)y + 2 : st.h[r];
                    }
                }
            }
        }
        *reinterpret_cast<int4 *>(capb + q) = make_int4(pend, capv[0], capv[1], capv[2]);
        pend = capv[3];
    }
    *((lane == 63) ? ring_cur + ((s0 + K) & (RS - 1)) : dummy + lane * 4) = pend;
}

// One strip (`g`) of one pair, swept by the whole workgroup.
template <int P, bool CMP, int R, int W, int K, int AW, bool DUMP = false, bool NOMAX = false>
__device__ void run_group(const WfKernelArgs &ka, const WfPairDev &pd, int pair, int g,
                          WgShared<W, K, AW> &sh, int lane, int wave, uint32_t epoch)
{
    constexpr int D = ChunkGeom<K>::D;
    constexpr int RS = ChunkGeom<K>::RS;
    constexpr bool TRACK = (P == WFK_SW || P == WFK_SWP) && !NOMAX;
    const int n1 = pd.n1, n2 = pd.n2;
    const int64_t row0 = (int64_t)g * (W * 64 * R) + wave * (64 * R) + lane * R;
    const uint8_t *acodes = ka.codes + pd.a_off;
    const uint8_t *bcodes = ka.codes + pd.b_off;
    // Column value staged for the step: the code, or (WFK_SWP) its selector.
    auto colv = [](uint32_t c) -> uint32_t {
        if constexpr (P == WFK_SWP) return Cell<WFK_SWP, false>::selector(c & 3u);
        return c;
    };

    LaneState<R> st;
    if constexpr (P == WFK_SWP) {
        const uint32_t c0 = row0 < n2 ? (uint32_t)bcodes[row0] : 4u;
        const uint32_t c1 = row0 + 1 < n2 ? (uint32_t)bcodes[row0 + 1] : 4u;
        st.t[0] = Cell<P, CMP>::tables(c0, c1, 0);
        st.t[1] = Cell<P, CMP>::tables(c0, c1, 1);
#pragma unroll
        for (int r = 0; r < R; ++r) st.h[r] = 0;
    } else {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int64_t y = row0 + r;
            const bool v = y < n2;
            st.t[r] = Cell<P, CMP>::row_entry(v ? (uint32_t)bcodes[y] : 0u, v);
            st.h[r] = 0;
        }
    }
    st.dg = 0;
    st.best = 0;

    const int nsteps = n1 + 63;
    const int nch = (nsteps + K - 1) / K;
    // Waves holding at least one row of the pair (fewer than W only in a
    // pair's last strip): the strip ends when the last of them is done.
    const int64_t rows_left = (int64_t)n2 - (int64_t)g * (W * 64 * R);
    const int wa = (int)min<int64_t>(W, (rows_left + 64 * R - 1) / (64 * R));
    const int total = nch + (wa - 1) * D;
    const bool has_prev = g > 0;
    const bool has_next = g + 1 < pd.ngroups;
    // Strip-boundary rows hold (launch epoch << 32 | value): a reader knows
    // a column is ready when its tag equals this launch's epoch, so the rows
    // need no separate progress flag, no release/acquire cache maintenance
    // and no clearing between launches.
    const uint64_t *bprev = ka.bnd + pd.bnd_off + (int64_t)(g - 1) * n1;
    uint64_t *bnext = ka.bnd + pd.bnd_off + (int64_t)g * n1;
    const uint64_t tag = (uint64_t)epoch << 32;
    // Row 0 of the table in this kernel's value representation.
    const int topv = (P == WFK_ED && !(pd.mode & WFK_MODE_TEXTBOOK)) ? -1 : 0;

    // Top-boundary loads for columns [c*K, c*K+K) (wave 0, lanes < K).
    auto load_top = [&](int c) -> uint64_t {
        const int x = c * K + lane;
        if (!has_prev || lane >= K || x >= n1) return tag;  // nothing to wait for
        return __hip_atomic_load(bprev + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    };
    // Make sure every lane's value carries this epoch (re-polling the missing
    // ones), then store columns [c*K, c*K+K) into ring[0] and the code window.
    auto fill = [&](int c, uint64_t top, uint32_t code_pre) -> bool {
        const int x = c * K + lane;
        if (__builtin_amdgcn_ballot_w64((top >> 32) != (uint64_t)epoch) != 0) {
            const int64_t t0 = now_ticks();
            for (;;) {
                __builtin_amdgcn_s_sleep(1);
                if ((top >> 32) != (uint64_t)epoch)
                    top = __hip_atomic_load(bprev + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (__builtin_amdgcn_ballot_w64((top >> 32) != (uint64_t)epoch) == 0) break;
                if (now_ticks() - t0 > ka.timeout_ticks) return false;
            }
        }
        if (lane < K && x < n1) {
            const int v = has_prev ? (int)(uint32_t)top : topv;
            sh.ring[0][(x + 64) & (RS - 1)] = v;
            const int slot = x & (AW - 1);
            sh.aw[slot] = code_pre;
            if (slot < K) sh.aw[AW + slot] = code_pre;
        }
        return true;
    };

    if (wave == 0) {
        const uint32_t c0 = colv((lane < K && lane < n1) ? acodes[lane] : 0u);
        if (!fill(0, load_top(0), c0) && lane == 0) sh.abort_at = -1;
    }
    __syncthreads();

    int copied = 0;  // wave W-1: columns already stored to bnext
    // Store finished bottom-row columns [copied, upto) from ring[W].
    auto copy_out = [&](int upto, bool from_ring) {
        while (copied < upto) {
            const int x = copied + lane;
            if (x < upto) {
                const int v = from_ring ? sh.ring[W][(x + 64) & (RS - 1)] : 0;
                __hip_atomic_store(bnext + x, tag | (uint32_t)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            copied = min(upto, copied + 64);
        }
    };

    bool aborted = sh.abort_at <= -1;
    int pend = 0;  // lane 63's last capture not yet in the ring (see run_chunk)
    for (int j = 0; j < total && !aborted; ++j) {
        const int cj = j - wave * D;
        // Wave 0 prefetches the next chunk's column codes and top boundary.
        uint32_t code_pre = 0;
        uint64_t top_pre = tag;
        if (wave == 0 && j + 1 < nch) {
            const int x = (j + 1) * K + lane;
            code_pre = colv((lane < K && x < n1) ? acodes[x] : 0u);
            top_pre = load_top(j + 1);
        }
        if (cj >= 0 && cj < nch && wave < wa) {
            const int s0 = cj * K;
            const bool steady = (s0 >= 63) && (s0 + K - 1 <= n1 - 1);
            const DumpRows dr{ka.table, row0, n2};
            if (steady)
                run_chunk<P, CMP, R, K, AW, false, TRACK, DUMP>(st, s0, lane, n1, sh.ring[wave],
                                                                sh.ring[wave + 1], sh.dummy[wave], sh.aw, pend, dr);
            else
                run_chunk<P, CMP, R, K, AW, true, TRACK, DUMP>(st, s0, lane, n1, sh.ring[wave],
                                                               sh.ring[wave + 1], sh.dummy[wave], sh.aw, pend, dr);
            if (wave == W - 1 && has_next) {
                // Steps < s0 + K are done: lane 63 has finished columns < s0 + K - 63.
                copy_out(min(n1, s0 + K - 63), true);
            }
        }
        if (wave == 0 && j + 1 < nch) {
            if (!fill(j + 1, top_pre, code_pre) && lane == 0) sh.abort_at = j;
        }
        __syncthreads();
        aborted = sh.abort_at <= j;
    }

    if (aborted) {
        // A dependency wait timed out.  Mark the pair and release the strip
        // below (tagged placeholder values) so that it does not wait out its
        // own timeout in turn.
        if (threadIdx.x == 0) atomicOr(&ka.res[pair].status, 1);
        if (wave == W - 1 && has_next) copy_out(n1, false);
        return;
    }
    if (wave == W - 1 && has_next) copy_out(n1, true);

    // Bottom-right cell.
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

template <int R, int W, int K>
struct EngineCfg {
    static constexpr int D = ChunkGeom<K>::D;
    static constexpr int AW = pow2_at_least((2 + (W - 1) * D) * K + 64 + K);
};

template <int R, int W, int K, bool DUMP>
__device__ __forceinline__ void engine_body(const WfKernelArgs &ka)
{
    constexpr int AW = EngineCfg<R, W, K>::AW;
    __shared__ WgShared<W, K, AW> sh;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t epoch = __builtin_amdgcn_readfirstlane(*ka.epoch);
    for (;;) {
        if (threadIdx.x == 0) {
            sh.task = ka.task_begin + atomicAdd(ka.counter, 1);
            sh.abort_at = 0x7fffffff;
        }
        __syncthreads();
        const int t = __builtin_amdgcn_readfirstlane(sh.task);
        __syncthreads();
        if (t >= ka.task_end) break;
        const WfTaskDev td = ka.tasks[t];
        const WfPairDev pd = ka.pairs[td.pair];
        set_wave_priority(pd.prio);
        const bool cmp = (pd.mode & WFK_MODE_CMP) != 0;
        switch (pd.prob * 2 + (cmp ? 1 : 0)) {
        case WFK_LCS * 2 + 0: run_group<WFK_LCS, false, R, W, K, AW, DUMP>(ka, pd, td.pair, td.group, sh, lane, wave, epoch); break;
        case WFK_LCS * 2 + 1: run_group<WFK_LCS, true, R, W, K, AW, DUMP>(ka, pd, td.pair, td.group, sh, lane, wave, epoch); break;
        case WFK_ED * 2 + 0: run_group<WFK_ED, false, R, W, K, AW, DUMP>(ka, pd, td.pair, td.group, sh, lane, wave, epoch); break;
        case WFK_ED * 2 + 1: run_group<WFK_ED, true, R, W, K, AW, DUMP>(ka, pd, td.pair, td.group, sh, lane, wave, epoch); break;
        case WFK_SW * 2 + 0: {
            // Instances without the running maximum for WFK_MODE_NOMAX (the
            // planner never sets it for the table dump).
            const bool nomax = !DUMP && (pd.mode & WFK_MODE_NOMAX) != 0;
            if constexpr (R == 2) {
                if (pd.mode & WFK_MODE_PERM4) {
                    if constexpr (!DUMP) {
                        if (nomax) {
                            run_group<WFK_SWP, false, R, W, K, AW, false, true>(ka, pd, td.pair, td.group, sh, lane,
                                                                               wave, epoch);
                            break;
                        }
                    }
                    run_group<WFK_SWP, false, R, W, K, AW, DUMP>(ka, pd, td.pair, td.group, sh, lane, wave, epoch);
                    break;
                }
            }
            if constexpr (!DUMP) {
                if (nomax) {
                    run_group<WFK_SW, false, R, W, K, AW, false, true>(ka, pd, td.pair, td.group, sh, lane, wave, epoch);
                    break;
                }
            }
            run_group<WFK_SW, false, R, W, K, AW, DUMP>(ka, pd, td.pair, td.group, sh, lane, wave, epoch);
            break;
        }
        default: run_group<WFK_SW, true, R, W, K, AW, DUMP>(ka, pd, td.pair, td.group, sh, lane, wave, epoch); break;
        }
        __syncthreads();
    }
}

// Residency bound of the lock-step engine, the throughput engine (batches,
// served mixes), where residency wins: four-wave workgroups of one or two
// rows per lane at 8 per CU (8 waves per SIMD, at most 64 VGPRs; the
// 4096-pair SW batch 0.99 ms against 1.03 ms at 6 per CU / 80 VGPRs, 1.09 ms
// at 5 / 96 and 1.19 ms at 4 / 104, spill-free, so residency beats the
// spills); other shapes at 6 per CU.
template <int R, int W>
constexpr int lockstep_min_blocks()
{
    return (W == 4 && R <= 2) ? 8 : 6;
}

template <int R, int W, int K>
__global__ __launch_bounds__(64 * W, (lockstep_min_blocks<R, W>())) void wf_engine(WfKernelArgs ka)
{
    engine_body<R, W, K, false>(ka);
}

// Same sweep, also writing every cell of pair 0 to ka.table (the reference
// drivers' DEBUG table dump; small tables only).
__global__ __launch_bounds__(256) void wf_engine_table(WfKernelArgs ka)
{
    engine_body<2, 4, 32, true>(ka);
}

// ----------------------------------------------------------------------------
// Resident server engine: one persistent launch of the lock-step engine
// serves every request of a server (structures in wf_device.h).
//
// Thread 0 of a free workgroup claims the next task index and waits for it
// to become valid.  While it waits it may take the poll lease (one
// workgroup at a time reads the host's submission ring, so the waiting grid
// does not flood the host link); a new submission is then taken in by that
// whole workgroup: the pair's codes are copied from pinned host memory to
// HBM, its descriptor, result and done counter are set up, and its strips
// are appended to the task ring in order, each made valid after a release
// fence.  Strips are claimed in task order, so a strip only waits on strips
// already claimed by running workgroups, as in a plan launch.  A strip runs
// the ordinary lock-step sweep with the pair's own boundary tag; the last
// strip of a pair to finish writes the completion record to pinned host
// memory.  The launch ends when the host sets its stop flag (only when no
// request is in flight).
// ----------------------------------------------------------------------------
enum { kServeRun = 0, kServeIntake = 1, kServeExit = 2 };

__device__ __forceinline__ uint64_t ld_agent(const uint64_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t ld_system(const uint64_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Take submission `s` in (whole workgroup).
__device__ void serve_intake(const WfServeArgs &sa, int64_t s)
{
    // (Raw storage: the descriptor type has default member initializers,
    // which __shared__ variables cannot take.)
    __shared__ __attribute__((aligned(16))) unsigned char se_raw[sizeof(WfSubmitDev)];
    WfSubmitDev &se = *reinterpret_cast<WfSubmitDev *>(se_raw);
    const WfSubmitDev *e = sa.subq + (s % sa.subq_cap);
    if (threadIdx.x == 0) {
        // The host writes the entry before it publishes its tail; the check
        // of the entry's own sequence number is bounded all the same.
        const int64_t t0 = now_ticks();
        while (ld_system(&e->seq) != (uint64_t)s + 1 && now_ticks() - t0 < sa.ka.timeout_ticks)
            __builtin_amdgcn_s_sleep(1);
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");
        se = *e;
    }
    __syncthreads();
    const int slot = se.slot;
    // Codes: one contiguous, 16-byte aligned run from a_off.
    const uint4 *src = reinterpret_cast<const uint4 *>(sa.staging + se.stage_off);
    uint4 *dst = reinterpret_cast<uint4 *>(const_cast<uint8_t *>(sa.ka.codes) + se.pd.a_off);
    const int64_t n16 = se.code_bytes / 16;
    for (int64_t i = threadIdx.x; i < n16; i += blockDim.x) dst[i] = src[i];
    if (threadIdx.x == 0) {
        const_cast<WfPairDev *>(sa.ka.pairs)[slot] = se.pd;
        sa.ka.res[slot] = WfResultDev{0, 0, 0, 0};
        sa.pair_done[slot] = 0;
    }
    for (int g = threadIdx.x; g < se.pd.ngroups; g += blockDim.x) {
        WfTaskQDev *t = sa.taskq + ((se.task_base + g) % sa.taskq_cap);
        t->slot = slot;
        t->group = g;
    }
    // One release for the whole workgroup's stores (the barrier orders them
    // before it), then the strips become valid.
    __syncthreads();
    if (threadIdx.x == 0) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    for (int g = threadIdx.x; g < se.pd.ngroups; g += blockDim.x) {
        const int64_t idx = se.task_base + g;
        __hip_atomic_store(&sa.taskq[idx % sa.taskq_cap].seq, (uint64_t)idx + 1, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
}

// After the last strip of the pair in `slot` (thread 0).
__device__ void serve_complete(const WfServeArgs &sa, int slot, int ngroups)
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const int done = atomicAdd(sa.pair_done + slot, 1);
    if (done != ngroups - 1) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    const WfResultDev r = sa.ka.res[slot];
    const uint64_t k = atomicAdd(reinterpret_cast<unsigned long long *>(&sa.ctl->comp_tail), 1ULL);
    WfCompDev *c = sa.compq + (k % sa.compq_cap);
    __hip_atomic_store(&c->last, r.last, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&c->maxs, r.maxs, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&c->status, r.status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&c->slot, slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
    __hip_atomic_store(&c->seq, k + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

template <int R, int W, int K>
__global__ __launch_bounds__(64 * W, (lockstep_min_blocks<R, W>())) void wf_serve_engine(WfServeArgs sa)
{
    constexpr int AW = EngineCfg<R, W, K>::AW;
    __shared__ WgShared<W, K, AW> sh;
    __shared__ int64_t act[2];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    int64_t idx = -1;     // thread 0: claimed task index not yet run
    bool lease = false;   // thread 0: holds the poll lease
    for (;;) {
        if (threadIdx.x == 0) {
            if (idx < 0) idx = (int64_t)atomicAdd(reinterpret_cast<unsigned long long *>(&sa.ctl->head), 1ULL);
            // Waiting workgroups back off exponentially (up to about 3.4 us
            // between polls) and look at the stop flag and the lease only
            // every 8th poll: the grid's idle polling must not load the
            // memory system the running strips use.
            int backoff = 1;
            for (int it = 0;; ++it) {
                const WfTaskQDev *t = sa.taskq + (idx % sa.taskq_cap);
                if (ld_agent(&t->seq) == (uint64_t)idx + 1) {
                    act[0] = kServeRun;
                    act[1] = idx;
                    idx = -1;
                    break;
                }
                const bool rare = (it & 7) == 0;
                if (rare && __hip_atomic_load(&sa.ctl->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    act[0] = kServeExit;
                    break;
                }
                if (!lease && rare &&
                    __hip_atomic_load(&sa.ctl->poller, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
                    lease = atomicCAS(&sa.ctl->poller, 0, 1) == 0;
                if (lease) {
                    backoff = 1;  // the lease holder polls the host at full rate
                    if (__hip_atomic_load(&sa.hctl->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM)) {
                        __hip_atomic_store(&sa.ctl->stop, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        continue;
                    }
                    const uint64_t tail = ld_system(&sa.hctl->sub_tail);
                    const uint64_t sub = ld_agent(&sa.ctl->sub_head);
                    if (sub < tail) {
                        // Only the lease holder advances sub_head.
                        __hip_atomic_store(&sa.ctl->sub_head, sub + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        __hip_atomic_store(&sa.ctl->poller, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        lease = false;
                        act[0] = kServeIntake;
                        act[1] = (int64_t)sub;
                        break;
                    }
                }
                // s_sleep takes an immediate: the backoff picks one.
                switch (backoff) {
                case 1: __builtin_amdgcn_s_sleep(1); break;
                case 2: __builtin_amdgcn_s_sleep(2); break;
                case 4: __builtin_amdgcn_s_sleep(4); break;
                case 8: __builtin_amdgcn_s_sleep(8); break;
                case 16: __builtin_amdgcn_s_sleep(16); break;
                case 32: __builtin_amdgcn_s_sleep(32); break;
                case 64: __builtin_amdgcn_s_sleep(64); break;
                default: __builtin_amdgcn_s_sleep(127); break;
                }
                if (backoff < 128) backoff *= 2;
            }
            if (lease && act[0] != kServeIntake) {
                __hip_atomic_store(&sa.ctl->poller, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                lease = false;
            }
            // Everything the strip reads was written before its task became
            // valid (intake's release fence).
            if (act[0] == kServeRun) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        __syncthreads();
        const int a = __builtin_amdgcn_readfirstlane((int)act[0]);
        const int64_t arg = act[1];
        __syncthreads();
        if (a == kServeExit) break;
        if (a == kServeIntake) {
            serve_intake(sa, arg);
            continue;
        }
        const WfTaskQDev td = sa.taskq[arg % sa.taskq_cap];
        const WfPairDev pd = sa.ka.pairs[td.slot];
        if (threadIdx.x == 0) sh.abort_at = 0x7fffffff;
        __syncthreads();
        set_wave_priority(pd.prio);
        const bool cmp = (pd.mode & WFK_MODE_CMP) != 0;
        const uint32_t tag = pd.tag;
        switch (pd.prob * 2 + (cmp ? 1 : 0)) {
        case WFK_LCS * 2 + 0: run_group<WFK_LCS, false, R, W, K, AW>(sa.ka, pd, td.slot, td.group, sh, lane, wave, tag); break;
        case WFK_LCS * 2 + 1: run_group<WFK_LCS, true, R, W, K, AW>(sa.ka, pd, td.slot, td.group, sh, lane, wave, tag); break;
        case WFK_ED * 2 + 0: run_group<WFK_ED, false, R, W, K, AW>(sa.ka, pd, td.slot, td.group, sh, lane, wave, tag); break;
        case WFK_ED * 2 + 1: run_group<WFK_ED, true, R, W, K, AW>(sa.ka, pd, td.slot, td.group, sh, lane, wave, tag); break;
        case WFK_SW * 2 + 0: {
            const bool nomax = (pd.mode & WFK_MODE_NOMAX) != 0;
            if constexpr (R == 2) {
                if (pd.mode & WFK_MODE_PERM4) {
                    if (nomax)
                        run_group<WFK_SWP, false, R, W, K, AW, false, true>(sa.ka, pd, td.slot, td.group, sh, lane, wave,
                                                                           tag);
                    else
                        run_group<WFK_SWP, false, R, W, K, AW>(sa.ka, pd, td.slot, td.group, sh, lane, wave, tag);
                    break;
                }
            }
            if (nomax)
                run_group<WFK_SW, false, R, W, K, AW, false, true>(sa.ka, pd, td.slot, td.group, sh, lane, wave, tag);
            else
                run_group<WFK_SW, false, R, W, K, AW>(sa.ka, pd, td.slot, td.group, sh, lane, wave, tag);
            break;
        }
        default: run_group<WFK_SW, true, R, W, K, AW>(sa.ka, pd, td.slot, td.group, sh, lane, wave, tag); break;
        }
        __syncthreads();
        if (threadIdx.x == 0) serve_complete(sa, td.slot, pd.ngroups);
    }
}

// Zero progress counters, results and claim counters before a launch (a
// kernel rather than hipMemsetAsync so the sequence can be graph-captured).
// Also advances the launch epoch that tags strip-boundary values.
__global__ void wf_reset(WfResultDev *res, int64_t nres, int32_t *counters, int64_t ncounters, uint32_t *epoch)
{
    const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    if (i0 == 0) *epoch = *epoch + 1u;
    for (int64_t i = i0; i < nres; i += stride) res[i] = WfResultDev{0, 0, 0, 0};
    for (int64_t i = i0; i < ncounters; i += stride) counters[i] = 0;
}

}  // namespace wfk

// ----------------------------------------------------------------------------
// Variant table used by the host planner.
// ----------------------------------------------------------------------------
#include "wf_variants.h"

#define WF_VARIANT(R_, W_, K_) \
    {WF_ENGINE_LOCKSTEP, R_, W_, K_, 64 * (W_), reinterpret_cast<const void *>(&wfk::wf_engine<R_, W_, K_>)}

extern "C" const WfVariant wf_variants_lockstep[] = {
    WF_VARIANT(1, 4, 16), WF_VARIANT(2, 4, 16), WF_VARIANT(4, 4, 16),
    WF_VARIANT(1, 4, 32), WF_VARIANT(2, 4, 32), WF_VARIANT(4, 4, 32),
    WF_VARIANT(1, 4, 64), WF_VARIANT(2, 4, 64), WF_VARIANT(4, 4, 64),
    WF_VARIANT(1, 1, 32), WF_VARIANT(2, 1, 32), WF_VARIANT(4, 1, 32), WF_VARIANT(8, 1, 32),
    WF_VARIANT(1, 8, 16), WF_VARIANT(2, 8, 16),
    WF_VARIANT(1, 2, 32), WF_VARIANT(2, 2, 32), WF_VARIANT(4, 2, 32),
};
extern "C" const int wf_num_variants_lockstep = sizeof(wf_variants_lockstep) / sizeof(wf_variants_lockstep[0]);

extern "C" hipError_t wf_launch_reset(WfResultDev *res, int64_t nres, int32_t *counters, int64_t ncounters,
                                      uint32_t *epoch, hipStream_t stream)
{
    int64_t n = nres;
    if (ncounters > n) n = ncounters;
    int blocks = (int)((n + 255) / 256);
    if (blocks < 1) blocks = 1;
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(wfk::wf_reset, dim3(blocks), dim3(256), 0, stream, res, nres, counters, ncounters, epoch);
    return hipGetLastError();
}

extern "C" hipError_t wf_launch_table(int grid, const WfKernelArgs *ka, hipStream_t stream)
{
    void *args[] = {const_cast<WfKernelArgs *>(ka)};
    return hipLaunchKernel(reinterpret_cast<const void *>(&wfk::wf_engine_table), dim3(grid), dim3(256), args, 0,
                           stream);
}

// The resident server engine: lock-step <2, 4, 64> (four-wave workgroups of
// 512-row strips, byte row tables for SW of at most 4 symbols).
extern "C" const void *wf_serve_kernel(int *block)
{
    *block = 64 * 4;
    return reinterpret_cast<const void *>(&wfk::wf_serve_engine<2, 4, 64>);
}

extern "C" hipError_t wf_launch_serve(int grid, const WfServeArgs *sa, hipStream_t stream)
{
    void *args[] = {const_cast<WfServeArgs *>(sa)};
    return hipLaunchKernel(reinterpret_cast<const void *>(&wfk::wf_serve_engine<2, 4, 64>), dim3(grid), dim3(64 * 4),
                           args, 0, stream);
}

extern "C" hipError_t wf_launch_engine(const WfVariant *v, int grid, const WfKernelArgs *ka,
                                       hipStream_t stream)
{
    void *args[] = {const_cast<WfKernelArgs *>(ka)};
    return hipLaunchKernel(v->fn, dim3(grid), dim3(v->block), args, 0, stream);
}
