// wf_engine.hip -- asynchronous strip engine for gfx950.
//
// Same lane geometry as the lock-step engine in wf_kernels.hip (lane l owns
// R rows and computes column s - l at step s; DPP wave_shr:1 passes the
// bottom cell to the next lane), but the waves of a workgroup no longer meet
// at barriers:
//
//   * W compute waves stack vertically (wave w owns rows w*64*R ... of the
//     strip).  Each advances in groups of G steps and publishes its progress
//     in an LDS counter; the wave below starts a group once the counter says
//     the lane-63 values it needs are in the LDS ring.  The counter is read
//     just before the group's first operands and checked when the group
//     starts (check on use), so the steady-state lag between two waves is
//     64 + G steps plus the LDS latency.
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

#ifdef WFK_TRACE
// Diagnostic build only (-DWFK_TRACE): per (task, wave) timestamps of the
// strip schedule, read back with wf_trace_read_async().
struct TraceRec {
    int64_t t_begin, t_start, t_end;   // s_memrealtime (100 MHz)
    int64_t wait_cyc, tot_cyc;         // s_memtime cycles
    int32_t nwait, hwid, xcc, nwait_out;
    int64_t t_ck[16];                  // when steps 16384 * (k + 1) were done
    int64_t c_ck[16];                  // the same in s_memtime cycles
    int64_t w_ck[16];                  // wait cycles so far
    int32_t n_ck[16];                  // waits so far
};
constexpr int kTraceTasks = 8192, kTraceWaves = 9;
__device__ TraceRec g_trace[kTraceTasks * kTraceWaves];
#endif

// Steps per iteration of the steady compute loop (a multiple of the check
// granularity G): one loop back-edge per kU steps, the counter checks in
// between are not-taken branches.  kB: steps per block of LDS operands.
constexpr int kU = 32;
// The SW byte-table strips of the four-wave, G = 32 kernels (the single-pair
// SW path) run 256 steps per iteration instead: one lap of the 256-slot
// rings, so every ring address of an iteration is a constant offset and the
// per-iteration scalar address arithmetic is amortised over 8 times the
// steps (SW 131072^2 7.07 -> 6.75 ms).  Other strips keep kU, which keeps
// the build time of the many variants in check.
constexpr int kUMax = 256;
template <int P, int W, int G>
constexpr int steady_unroll()
{
    return ((P == WFK_SWP || P == WFK_SWE) && W == 4 && G == 32) ? kUMax : kU;
}
constexpr int kB = 8;  // a 4-step block (prefetch 4 steps ahead) measured slower: 7.76 against 7.71 ms

template <int W, int G, int AW, int RS>
struct AsyncShared {
    // Column codes, one per dword (a lane's codes of a block are then aligned
    // dword reads, which LDS serves faster than unaligned or byte reads):
    // column x at slot (x + 64) mod AW, the first kUMax + kB slots mirrored
    // at AW so that a lane's reads of one loop iteration never wrap.  First in
    // the workgroup's LDS so that a lane's reads fit the 8-bit dword offsets
    // of ds_read2 without extra address arithmetic.
    __attribute__((aligned(16))) uint32_t aw[AW + kUMax + kB];
    // ring[0]: the strip's top input, staged by the input wave; ring[w+1]:
    // lane 63 of compute wave w, read by wave w+1 (w = W-1: by the output
    // wave).  Column x lives in slot (x + 64) mod RS.
    __attribute__((aligned(16))) int ring[W + 1][RS];
    // Lanes 0..62 store their bottom cells here (lane l at 4l plus the
    // block's offset: each 16-byte write of the wave covers consecutive
    // banks) so that a block's capture is two unmasked ds_write_b128.
    __attribute__((aligned(16))) int dummy[W][64 * 4 + kUMax];
    // Progress counters, 64 copies each so a wave stores one value with an
    // unmasked ds_write:
    //   cnt[0]     columns staged in ring[0] and aw (input wave)
    //   cnt[1+w]   steps finished by compute wave w
    //   cnt[W+1]   bottom-row columns stored to HBM (output wave)
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
// with -DWFK_LDS_HW_FENCE=1 uses s_waitcnt-based workgroup fences instead
// (which drain the read pipeline at every group boundary).
// Sleep (s_sleep units of 64 clocks) of the output wave between polls of
// the last compute wave's progress, and of the input wave while it has no
// room to stage into.
#ifndef WFK_IO_SLEEP
#define WFK_IO_SLEEP 3
#endif
#ifndef WFK_IN_SLEEP
#define WFK_IN_SLEEP 16
#endif
// Polls of a waiting compute wave before it sleeps between polls.
#ifndef WFK_SPIN_FREE
#define WFK_SPIN_FREE 64
#endif

#ifndef WFK_LDS_HW_FENCE
#define WFK_LDS_HW_FENCE 0
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

// Operands of one block of kB steps: lane 0's top neighbours of kB columns
// (broadcast ds_read_b128s) and the lane's kB column codes.  Block k+1 is
// read while block k is computed, kB steps ahead of its use, which covers
// the LDS latency with six waves of a workgroup sharing the LDS.
struct Block {
    int4 inj[kB / 4];
    uint32_t ca[kB];
    __device__ int top(int u) const
    {
        const int4 v = inj[u / 4];
        return (u & 3) == 0 ? v.x : (u & 3) == 1 ? v.y : (u & 3) == 2 ? v.z : v.w;
    }
};

// Read the block of steps [s, s + kB) (s a multiple of kB; RS and AW are
// multiples of kB and the code window is mirrored past AW, so nothing wraps
// inside a block).
template <int AW, int RS>
__device__ __forceinline__ void load_block(Block &b, const int *ring_prev, const uint32_t *aw, int s, int lane)
{
    const int *ip = ring_prev + ((s + 64) & (RS - 1));
#pragma unroll
    for (int i = 0; i < kB / 4; ++i) b.inj[i] = *reinterpret_cast<const int4 *>(ip + 4 * i);
    const uint32_t *cp = aw + ((s - lane + 64) & (AW - 1));  // column x in slot (x + 64) mod AW
#pragma unroll
    for (int u = 0; u < kB; ++u) b.ca[u] = cp[u];
}

// U steps of one compute wave, starting at step q0*G (U a multiple of G, G
// of kB; the progress counters work in groups of G steps).  `blk` holds the
// first block's operands on entry and the block after the last on return.
// Check on use: the counters that vouch for group g are read just before the
// operands of g's first block (LDS executes one wave's accesses in order, so
// operands read after a counter that shows the producer's stores see those
// stores), and checked when the group starts.  A short counter sends the
// wave to a poll loop that re-reads the counters and the block's operands
// together (one LDS round trip per poll; the operands are valid once the
// counters read with them suffice), with `wait_slow` between polls.  A wave
// thus trails its producer by 64 + G steps plus the LDS latency, instead of
// 64 + 2G with a check one group ahead.  The group's
// progress is published after its last block (publishing after every block
// measured slower: the step is bound by the wave's instruction issue).  There is one load site per
// block on the fast path, so the loop-carried block stays in the same VGPRs.
// Returns false if a wait was abandoned.
template <int P, bool CMP, int R, int G, int U, int AW, int RS, bool MASKED, bool TRACK, class NeedIn,
          class NeedOut, class WaitFn, class PubFn>
__device__ __forceinline__ bool run_steps(LaneState<R> &st, int q0, int ngr, int lane, int n1, int xlo,
                                          const int *ring_prev,
                                          int *ring_cur, int *dummy, const uint32_t *aw, Block &blk,
                                          int &pend, const int *cin, const int *cout, int &vin, int &vout,
                                          NeedIn &&need_in, NeedOut &&need_out, WaitFn &&wait_slow,
                                          PubFn &&publish)
{
    static_assert(U % G == 0 && G % kB == 0 && U <= RS, "bad unroll");
    (void)ngr;
    const int s0 = q0 * G;
    // s0 is a multiple of U, so neither the ring nor the (mirrored) code
    // window wraps inside the iteration; only the next iteration's first
    // block is addressed afresh.
    const int *__restrict__ ib = ring_prev + ((s0 + 64) & (RS - 1));
    // Lane 0's ring operands of the block at step s0 + qq (qq a multiple of
    // kB): contiguous from ib unless the iteration is a whole lap of the
    // ring (U == RS, s0 a multiple of RS), where each block sits at a
    // constant offset.
    auto ring_at = [&](int qq) __attribute__((always_inline)) -> const int * {
        return (U == RS) ? ring_prev + ((64 + qq) & (RS - 1)) : ib + qq;
    };
    const uint32_t *__restrict__ cp = aw + ((s0 - lane + 64) & (AW - 1));
    // Capture of lane 63's bottom cell (column s - 63 at step s, ring slot
    // s + 1), a block at a time: after the block of steps S..S+7 the values
    // of steps S-1..S+6 go to slots S..S+7 (16-byte aligned) with two
    // ds_write_b128, and the last step's value is carried in `pend`.  Each
    // group ends with `pend` stored alone, so a published count covers every
    // column it names.  Against a ds_write2 per two steps: SW 131072^2
    // 6.76-6.78 -> 6.56-6.58 ms.
    int *capb = (lane == 63) ? ring_cur + (s0 & (RS - 1)) : dummy + lane * 4;
    bool ok = true;
    // The consumer of this wave's ring trails it by about 64 + G steps and
    // the ring holds RS = 256 columns, so its counter (vout) is read and
    // checked only once per kOutEvery groups, for all of them.
    constexpr int kOutEvery = (G >= 32) ? 1 : 32 / G;
#pragma unroll
    for (int q = 0; q < U; q += kB) {
        const int gi = q0 + q / G;  // group of this block
        if (q % G == 0) {
            // Every lane holds its own copy of the counters, so the compare
            // is wave-uniform: one VALU compare and a branch on its mask.
            bool short_ = vin < need_in(gi);
            // Global group index (the steady loop starts at multiples of
            // kU / G groups, so there it is known at compile time).
            const bool out_group = MASKED ? (gi % kOutEvery == 0) : ((q / G) % kOutEvery == 0);
            // Output room is checked for every group up to the next checked
            // one (the groups in between do not look at vout).
            const int gout = out_group ? gi + kOutEvery - 1 : gi;
            if (out_group) short_ = short_ || vout < need_out(gout);
            if (__builtin_expect(__builtin_amdgcn_ballot_w64(short_) != 0, 0)) {
                for (int spins = 0;; ++spins) {
                    vin = cnt_load(cin);
                    vout = cnt_load(cout);
                    lds_acquire();
#pragma unroll
                    for (int i = 0; i < kB / 4; ++i)
                        blk.inj[i] = *reinterpret_cast<const int4 *>(ring_at(q) + 4 * i);
#pragma unroll
                    for (int u = 0; u < kB; ++u) blk.ca[u] = cp[q + u];
                    if (__builtin_amdgcn_ballot_w64(vin < need_in(gi) || vout < need_out(gout)) == 0) break;
                    if (!wait_slow(spins)) {
                        ok = false;
                        break;
                    }
                }
            }
            lds_acquire();
        }
        int injv[kB];
        uint32_t cav[kB];
#pragma unroll
        for (int u = 0; u < kB; ++u) {
            injv[u] = blk.top(u);
            cav[u] = blk.ca[u];
        }
        // The counters for the next group's check, read before that group's
        // first operands (the scheduling barriers keep the compiler from
        // moving either across the other).
        if ((q + kB) % G == 0) {
            // Only LDS accesses are kept from crossing (arithmetic may).
            __builtin_amdgcn_sched_barrier(0x7);
            vin = cnt_load(cin);
            if (MASKED ? ((gi + 1) % kOutEvery == 0) : (((q + kB) / G) % kOutEvery == 0)) vout = cnt_load(cout);
            lds_acquire();
            __builtin_amdgcn_sched_barrier(0x7);
        }
        // Past the last group the next block reads stale slots (in bounds)
        // that are never used.
        const int *ip = (q + kB < U) ? ring_at(q + kB) : ring_prev + ((s0 + U + 64) & (RS - 1));
#pragma unroll
        for (int i = 0; i < kB / 4; ++i) blk.inj[i] = *reinterpret_cast<const int4 *>(ip + 4 * i);
#pragma unroll
        for (int u = 0; u < kB; ++u) blk.ca[u] = cp[q + kB + u];
        // Keep the next block's LDS reads here, a block ahead of their use
        // (the scheduler would otherwise sink them to the block's end); every
        // other kind of instruction may still move across (mask: ALU, VALU,
        // SALU, DS write).
        __builtin_amdgcn_sched_barrier(0x207);
        int capv[kB];
#pragma unroll
        for (int u = 0; u < kB; ++u) {
            const int j = q + u;
            bool valid = true;
            // Columns [xlo, n1) are computed (xlo = -64 with a free head).
            if (MASKED) valid = (unsigned)(s0 + j - lane - xlo) < (unsigned)(n1 - xlo);
            step<P, CMP, R, MASKED, TRACK>(st, injv[u], cav[u], valid);
            capv[u] = st.h[R - 1];
        }
        {
            int4 *d = reinterpret_cast<int4 *>(capb + q);
            d[0] = make_int4(pend, capv[0], capv[1], capv[2]);
            d[1] = make_int4(capv[3], capv[4], capv[5], capv[6]);
            pend = capv[kB - 1];
        }
        if (q % G == G - kB) {
            *((lane == 63) ? ring_cur + ((s0 + q + kB) & (RS - 1)) : dummy + lane * 4) = pend;
            publish(gi);
        }
    }
    return ok;
}

// Forced inline into the kernel: left as a call it takes the kernel
// arguments by address and copies them (and the pair) to scratch memory,
// which the steady loop then reloads.
template <int P, bool CMP, int R, int W, int G, int AW, int RS, bool NOMAX = false>
__device__ __forceinline__ void async_strip(const WfKernelArgs &ka, const WfPairDev &pd, int pair, int g,
                            AsyncShared<W, G, AW, RS> &sh, int lane, int wave, uint32_t epoch, int task)
{
#ifdef WFK_TRACE
    const int64_t tr_begin = __builtin_amdgcn_s_memrealtime();
    const int64_t tr_c0 = clock64();
    int64_t tr_start = 0, tr_wait = 0;
    int tr_nwait = 0, tr_nwait_out = 0;
    bool tr_out_short = false;
    auto tr_store = [&]() {
        if (lane == 0 && task < kTraceTasks) {
            TraceRec &r = g_trace[task * kTraceWaves + wave];
            r.t_begin = tr_begin;
            r.t_start = tr_start;
            r.t_end = __builtin_amdgcn_s_memrealtime();
            r.wait_cyc = tr_wait;
            r.tot_cyc = clock64() - tr_c0;
            r.nwait = tr_nwait;
            r.nwait_out = tr_nwait_out;
            r.hwid = __builtin_amdgcn_s_getreg((31 << 11) | 4);
            r.xcc = __builtin_amdgcn_s_getreg((31 << 11) | 20);
        }
    };
#endif
    constexpr bool TRACK = (P == WFK_SW || P == WFK_SWP) && !NOMAX;
    const int n1 = pd.n1, n2 = pd.n2;
    const int nsteps = n1 + 63;
    const bool has_prev = g > 0;
    const bool has_next = g + 1 < pd.ngroups;
    const uint64_t tag = (uint64_t)epoch << 32;
    // Compute waves holding at least one row of the pair (fewer than W only
    // in a pair's last strip, which has no consumer below it).
    const int64_t rows_left = (int64_t)n2 - (int64_t)g * (W * 64 * R);
    const int wa = (int)min<int64_t>(W, (rows_left + 64 * R - 1) / (64 * R));

    if (wave < W && wave >= wa) {
        // Only padding rows: never block the wave above (its consumer).
        cnt_store(sh.cnt[wave + 1] + lane, 1 << 30);
        return;
    }
    if (wave < W) {
        // ------------------------------------------------------ compute wave
        const int w = wave;
        const int64_t row0 = (int64_t)g * (W * 64 * R) + w * (64 * R) + lane * R;
        const uint8_t *bcodes = ka.codes + pd.b_off;
        LaneState<R> st;
        if constexpr (P == WFK_SWP || P == WFK_SWE) {
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
        // Lanes left of column 0 run unmasked steps (see Cell::idle): lane 0
        // starts at column 0, lane l >= 1 at column -l from the idle state.
        const bool free_head = (pd.mode & WFK_MODE_FREE_HEAD) != 0;
        if (free_head && lane > 0) {
#pragma unroll
            for (int r = 0; r < R; ++r) st.h[r] = Cell<P, CMP>::idle;
            st.dg = Cell<P, CMP>::idle;
        }

        const int *cin = sh.cnt[w] + lane;        // producer of my input
        const int *cout = sh.cnt[w + 2] + lane;   // consumer of my output ring
        int *cmine = sh.cnt[w + 1] + lane;
        const int ngr = (nsteps + G - 1) / G;
        // Group q reads input columns < (q+1)G: staged (w = 0) or written by
        // lane 63 of wave w-1, which finishes column c at its step c + 63.
        auto need_in = [&](int q) __attribute__((always_inline)) { return w == 0 ? min(n1, (q + 1) * G) : min(nsteps, (q + 1) * G + 63); };
        // Group q overwrites ring slots of columns up to (q+1)G - 64 - RS,
        // which the consumer must have read.
        auto need_out = [&](int q) __attribute__((always_inline)) { return (q + 1) * G - 63 - RS; };

        // Wait until group q's inputs are there and its output slots free.
        // Every lane holds its own copy of a counter, so the values are read
        // through readfirstlane and the whole loop stays wave-uniform
        // (scalar branches, no exec-mask bookkeeping per group).  These
        // helpers are forced inline: a helper left out of line takes its
        // captures by address and moves the strip's state to scratch memory.
        auto ready = [&](int q) __attribute__((always_inline)) -> bool {
            const int vi = __builtin_amdgcn_readfirstlane(cnt_load(cin));
            const int vo = __builtin_amdgcn_readfirstlane(cnt_load(cout));
#ifdef WFK_TRACE
            tr_out_short = vi >= need_in(q) && vo < need_out(q);
#endif
            return vi >= need_in(q) && vo >= need_out(q);
        };
        auto wait_ready = [&](int q) __attribute__((always_inline)) -> bool {
            if (!ready(q)) {
#ifdef WFK_TRACE
                const int64_t c0 = clock64();
                ++tr_nwait;
                tr_nwait_out += tr_out_short ? 1 : 0;
                struct Acc {
                    int64_t c0, *acc;
                    __device__ ~Acc() { *acc += clock64() - c0; }
                } acc{c0, &tr_wait};
#endif
                // Spin first (the producer is usually a few steps away; an
                // LDS round trip is the natural poll period), then sleep.
                const int64_t t0 = now_ticks();
                for (int spins = 0;; ++spins) {
                    if (spins >= WFK_SPIN_FREE) __builtin_amdgcn_s_sleep(1);
                    if (ready(q)) break;
                    if (__builtin_amdgcn_readfirstlane(
                            __hip_atomic_load(&sh.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) ||
                        now_ticks() - t0 > ka.timeout_ticks)
                        return false;
                }
            }
            lds_acquire();
            return true;
        };
        Block blk;
        bool aborted = !wait_ready(0);
#ifdef WFK_TRACE
        tr_start = __builtin_amdgcn_s_memrealtime();
        tr_nwait = 0;
        tr_nwait_out = 0;
        tr_wait = 0;
#endif
        // Counters first, then the first block's operands (check on use).
        int vin = cnt_load(cin), vout = cnt_load(cout);
        lds_acquire();
        load_block<AW, RS>(blk, sh.ring[w], sh.aw, 0, lane);
        // Published progress, kept in a VGPR of its own (it is loaded, so the
        // compiler treats it as per-lane): materialising it from an SGPR right
        // before the store would reuse a register of a capture store still in
        // flight and make the wave wait for its block reads.
        int prog = cnt_load(cmine);  // 0: counters are cleared before each strip
        auto publish = [&](int gi) __attribute__((always_inline)) {
            (void)gi;
#ifdef WFK_TRACE
            if ((((gi + 1) * G) & 16383) == 0 && (gi + 1) * G <= 16 * 16384 && lane == 0 && task < kTraceTasks)
            {
                g_trace[task * kTraceWaves + wave].t_ck[(gi + 1) * G / 16384 - 1] = __builtin_amdgcn_s_memrealtime();
                g_trace[task * kTraceWaves + wave].c_ck[(gi + 1) * G / 16384 - 1] = clock64();
                g_trace[task * kTraceWaves + wave].w_ck[(gi + 1) * G / 16384 - 1] = tr_wait;
                g_trace[task * kTraceWaves + wave].n_ck[(gi + 1) * G / 16384 - 1] = tr_nwait;
            }
#endif
            prog = min(nsteps, prog + G);
            lds_release();
            cnt_store(cmine, prog);
        };
        // Between two polls of a short group: sleep once the wave has
        // polled WFK_SPIN_FREE times, and look at the abort flag and the
        // clock only every 16th poll (both cost a round trip of their own).
        int64_t t_wait0 = 0;
        auto wait_slow = [&](int spins) __attribute__((always_inline)) -> bool {
#ifdef WFK_TRACE
            if (spins == 0) ++tr_nwait;
#endif
            if (spins >= WFK_SPIN_FREE) __builtin_amdgcn_s_sleep(1);
            if ((spins & 15) == 15) {
                if (__builtin_amdgcn_readfirstlane(
                        __hip_atomic_load(&sh.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)))
                    return false;
                const int64_t t = now_ticks();
                if (spins == 15)
                    t_wait0 = t;
                else if (t - t_wait0 > ka.timeout_ticks)
                    return false;
            }
            return true;
        };
        // Steps whose columns are valid for every lane run the unmasked body,
        // kU steps per loop iteration; the masked ones (lanes before column
        // 0 at the start, past column n1 - 1 at the end) run one group per
        // iteration.  One body per loop keeps `blk` in fixed registers.
        constexpr int KU = steady_unroll<P, W, G>();
        constexpr int UG = KU / G;
        const int xlo = free_head ? -64 : 0;
        auto steady_from = [&](int q) __attribute__((always_inline)) {
            return (free_head || q * G >= 63) && q * G + KU - 1 <= n1 - 1 && q + UG <= ngr;
        };
        int q = 0;
        int pend = 0;  // lane 63's last capture not yet in the ring (see run_steps)
        for (; !aborted && q < ngr && !(steady_from(q) && (q * G) % KU == 0); ++q)
            aborted = !run_steps<P, CMP, R, G, G, AW, RS, true, TRACK>(
                st, q, ngr, lane, n1, xlo, sh.ring[w], sh.ring[w + 1], sh.dummy[w], sh.aw, blk, pend, cin, cout, vin, vout,
                need_in, need_out, wait_slow, publish);
        for (; !aborted && steady_from(q); q += UG)
            aborted = !run_steps<P, CMP, R, G, KU, AW, RS, false, TRACK>(
                st, q, ngr, lane, n1, xlo, sh.ring[w], sh.ring[w + 1], sh.dummy[w], sh.aw, blk, pend, cin, cout, vin, vout,
                need_in, need_out, wait_slow, publish);
        for (; !aborted && q < ngr; ++q)
            aborted = !run_steps<P, CMP, R, G, G, AW, RS, true, TRACK>(
                st, q, ngr, lane, n1, xlo, sh.ring[w], sh.ring[w + 1], sh.dummy[w], sh.aw, blk, pend, cin, cout, vin, vout,
                need_in, need_out, wait_slow, publish);
        if (aborted) {
            // Mark the pair failed: in a pair's last strip no output wave
            // exists, and an input wave that has staged every column is gone,
            // so nobody else would.
            if (lane == 0) {
                atomicOr(&ka.res[pair].status, 1);
                __hip_atomic_store(&sh.abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
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
            // The halves of a column-split SW pair keep their last column
            // (row y at word 2 * col_off + y) for the combine kernel.
            if ((P == WFK_SWP || P == WFK_SWE) && pd.col_off >= 0) {
#pragma unroll
                for (int r = 0; r < R; ++r)
                    if (row0 + r < n2) ka.colw[2 * pd.col_off + row0 + r] = (uint32_t)st.h[r];
            }
        }
#ifdef WFK_TRACE
        tr_store();
#endif
        return;
    }

    // The two I/O waves share SIMDs with compute waves 0 and 1 and run below
    // them: they only take issue slots those waves leave idle.
    __builtin_amdgcn_s_setprio(0);
    const uint8_t *acodes = ka.codes + pd.a_off;
    const uint64_t *bprev = ka.bnd + pd.bnd_off + (int64_t)(g - 1) * n1;
    uint64_t *bnext = ka.bnd + pd.bnd_off + (int64_t)g * n1;
    const int *clast = sh.cnt[wa] + lane;   // last compute wave: slowest reader of aw, producer of ring[W]
    bool aborted = false;

    if (wave == W + 1) {
        // ------------------------------------------------------ output wave
        // Stores the strip's bottom row (ring[W], written by lane 63 of the
        // last compute wave) to HBM as soon as 8 columns are finished, and
        // publishes the stored count in cnt[W+1] (which frees ring slots).
        if (!has_next) {
            // No consumer of the bottom row: never block the last wave.
            cnt_store(sh.cnt[W + 1] + lane, 1 << 30);
        } else {
            int out_done = 0;
            int64_t t_idle = now_ticks();
            while (out_done < n1) {
                const int pl = __builtin_amdgcn_readfirstlane(cnt_load(clast));
                const int avail = min(n1, pl - 63);
                if (avail - out_done >= 8 || (avail == n1 && avail > out_done)) {
                    const int end = min(avail, out_done + 64);
                    const int x = out_done + lane;
                    lds_acquire();
                    if (x < end) {
                        const int v = sh.ring[W][(x + 64) & (RS - 1)];
                        __hip_atomic_store(bnext + x, tag | (uint32_t)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                    out_done = end;
                    lds_release();
                    cnt_store(sh.cnt[W + 1] + lane, out_done);
                    t_idle = now_ticks();
                    continue;
                }
                if (__builtin_amdgcn_readfirstlane(
                        __hip_atomic_load(&sh.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) ||
                    now_ticks() - t_idle > ka.timeout_ticks) {
                    aborted = true;
                    break;
                }
                __builtin_amdgcn_s_sleep(WFK_IO_SLEEP);
            }
            if (aborted) {
                // Mark the pair and release the strip below with placeholder
                // values so that it does not wait out its own timeout in turn.
                if (lane == 0) {
                    atomicOr(&ka.res[pair].status, 1);
                    __hip_atomic_store(&sh.abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                }
                for (int x = out_done + lane; x < n1; x += 64)
                    __hip_atomic_store(bnext + x, tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
#ifdef WFK_TRACE
        tr_store();
#endif
        return;
    }

    // ---------------------------------------------------------- input wave
    // Stages the strip's top boundary (the previous strip's bottom row, as
    // it appears in HBM) and the column codes into LDS and publishes the
    // staged count in cnt[0].  Two polls of the next 64 unstaged columns are
    // kept in flight, half a poll latency apart, so a column is seen about a
    // quarter latency after it lands instead of half a latency; each poll
    // stages its ready prefix (the lanes up to the first stale tag).
    // Row 0 of the table in the kernels' value representation.
    const int topv = (P == WFK_ED && !(pd.mode & WFK_MODE_TEXTBOOK)) ? -1 : 0;
    const int *c0 = sh.cnt[1] + lane;       // compute wave 0: reads ring[0] and aw
    // Column x lives in code slot (x + 64) mod AW; the slots of columns
    // -64..-1, read by lanes left of column 0, hold a code that matches no
    // row (see WFK_MODE_FREE_HEAD).  Written before the first staging
    // counter store, so every compute wave reads them after they are set.
    {
        const uint32_t never =
            (P == WFK_SWP || P == WFK_SWE) ? Cell<WFK_SWP, false>::kNever
                                           : ((uint32_t)pd.mode >> WFK_MODE_NEVER_SHIFT) & 0x1FFu;
        sh.aw[lane] = never;
        if (lane < kUMax + kB) sh.aw[AW + lane] = never;
    }
    int in_next = 0;  // columns staged and published in cnt[0]
    struct Poll {
        uint64_t t;
        uint32_t c;
        int base;
        int width;  // columns polled: lanes past it count as not ready
#ifdef WFK_TRACE
        int64_t t_issue;
#endif
    };
#ifdef WFK_TRACE
    int64_t tr_lat[8] = {0, 0, 0, 0, 0, 0, 0, 0}, tr_cols[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int tr_polls[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#endif
    // Poll width: 32 columns while the polls keep up with the producing
    // strip (a poll then finds about 20 new columns), 64 after a poll that
    // came back ready in full (catching up); narrower polls fetch fewer
    // boundary words that are not written yet.
    int width = 64;
    auto issue = [&](Poll &p) __attribute__((always_inline)) {
        p.base = in_next;
        p.width = width;
#ifdef WFK_TRACE
        p.t_issue = __builtin_amdgcn_s_memrealtime();
#endif
        const int x = in_next + lane;
        const bool in = x < n1;
        p.c = in ? (uint32_t)acodes[x] : 0u;
        if (P == WFK_SWP || P == WFK_SWE) p.c = Cell<WFK_SWP, false>::selector(p.c & 3u);
        p.t = (has_prev && in)
                  ? (lane < p.width ? __hip_atomic_load(bprev + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0)
                  : tag;
    };
    // Stage what poll p shows beyond in_next, within the room the compute
    // waves have freed: ring[0] slot of column c is free once wave 0 has
    // read it (steps < p0 done), its code slot once the last wave has
    // (it reads columns >= pl - 63).  Then poll again from in_next.
    auto examine = [&](Poll &p, int p0, int pl) __attribute__((always_inline)) -> bool {
        const bool stale = (p.t >> 32) != (uint64_t)epoch;
        const uint64_t smask = __builtin_amdgcn_ballot_w64(stale);
        const int ready = smask ? (int)__builtin_ctzll(smask) : 64;
        width = (ready >= p.width) ? 64 : 32;
        const int room = min(p0 + RS, pl - 63 - G + AW);
        const int limit = min(min(n1, p.base + ready), room);
#ifdef WFK_TRACE
        {
            const int bk = min(7, p.base >> 14);
            tr_lat[bk] += __builtin_amdgcn_s_memrealtime() - p.t_issue;
            tr_polls[bk] += 1;
            tr_cols[bk] += max(0, limit - in_next);
        }
#endif
        bool did = false;
        if (limit > in_next) {
            const int x = p.base + lane;
            if (x >= in_next && x < limit) {
                sh.ring[0][(x + 64) & (RS - 1)] = has_prev ? (int)(uint32_t)p.t : topv;
                const int slot = (x + 64) & (AW - 1);
                sh.aw[slot] = p.c;
                if (slot < kUMax + kB) sh.aw[AW + slot] = p.c;
            }
            in_next = limit;
            lds_release();
            cnt_store(sh.cnt[0] + lane, in_next);
            did = true;
        }
        if (in_next < n1) issue(p);
        return did;
    };
    constexpr int NP = 2;  // four in flight measured slower (more, smaller bursts)
    Poll pp[NP];
#pragma unroll
    for (int i = 0; i < NP; ++i) {
        if (i > 0 && has_prev) __builtin_amdgcn_s_sleep(32 / NP);  // spread over a poll latency
        issue(pp[i]);
    }
    // Idle time only counts while the compute waves are not moving either,
    // i.e. while everything waits on the producing strip.
    int64_t t_idle = now_ticks();
    int seen = -1;
    while (in_next < n1) {
        int p0 = __builtin_amdgcn_readfirstlane(cnt_load(c0));
        int pl = __builtin_amdgcn_readfirstlane(cnt_load(clast));
        bool did = false;
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            if (in_next >= n1) break;
            did = examine(pp[i], p0, pl) || did;
        }
        if (in_next >= n1) break;
        const bool moving = (p0 + pl != seen);
        seen = p0 + pl;
        if (did || moving) t_idle = now_ticks();
        if (__builtin_amdgcn_readfirstlane(
                __hip_atomic_load(&sh.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) ||
            now_ticks() - t_idle > ka.timeout_ticks) {
            aborted = true;
            break;
        }
        // Only waiting for room: the compute waves free it at one column per
        // step, and the wave shares a SIMD with compute wave 0.
        if (!did && !has_prev) __builtin_amdgcn_s_sleep(WFK_IN_SLEEP);
    }
    if (aborted && lane == 0) {
        atomicOr(&ka.res[pair].status, 1);
        __hip_atomic_store(&sh.abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
#ifdef WFK_TRACE
    tr_store();
    if (lane == 0 && task < kTraceTasks)
        for (int k = 0; k < 8; ++k) {
            g_trace[task * kTraceWaves + wave].w_ck[k] = tr_lat[k];
            g_trace[task * kTraceWaves + wave].n_ck[k] = tr_polls[k];
            g_trace[task * kTraceWaves + wave].c_ck[k] = tr_cols[k];
        }
#endif
}

template <int R, int W, int G>
struct AsyncCfg {
    static constexpr int RS = 256;
    static constexpr int AW = pow2_at_least(W * (64 + 2 * G + 16) + 256);
    static_assert(G == 8 || G == 16 || G == 32, "G must divide 64");
};

template <int R, int W, int G>
__global__ __launch_bounds__(64 * (W + 2)) void wf_engine_async(WfKernelArgs ka)
{
    constexpr int AW = AsyncCfg<R, W, G>::AW;
    constexpr int RS = AsyncCfg<R, W, G>::RS;
    __shared__ AsyncShared<W, G, AW, RS> sh;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t epoch = __builtin_amdgcn_readfirstlane(*ka.epoch);
    for (;;) {
        if (threadIdx.x == 0) sh.task = ka.task_begin + atomicAdd(ka.counter, 1);
        // Fresh counters for the strip.
        for (int i = threadIdx.x; i < (W + 2) * 64; i += blockDim.x) (&sh.cnt[0][0])[i] = 0;
        if (threadIdx.x == 0) sh.abort = 0;
        __syncthreads();
        const int t = __builtin_amdgcn_readfirstlane(sh.task);
        if (t >= ka.task_end) break;
        const WfTaskDev td = ka.tasks[t];
        const WfPairDev pd = ka.pairs[td.pair];
        set_wave_priority(pd.prio);
        const bool cmp = (pd.mode & WFK_MODE_CMP) != 0;
        switch (pd.prob * 2 + (cmp ? 1 : 0)) {
        case WFK_LCS * 2 + 0: async_strip<WFK_LCS, false, R, W, G, AW, RS>(ka, pd, td.pair, td.group, sh, lane, wave, epoch, t); break;
        case WFK_LCS * 2 + 1: async_strip<WFK_LCS, true, R, W, G, AW, RS>(ka, pd, td.pair, td.group, sh, lane, wave, epoch, t); break;
        case WFK_ED * 2 + 0: async_strip<WFK_ED, false, R, W, G, AW, RS>(ka, pd, td.pair, td.group, sh, lane, wave, epoch, t); break;
        case WFK_ED * 2 + 1: async_strip<WFK_ED, true, R, W, G, AW, RS>(ka, pd, td.pair, td.group, sh, lane, wave, epoch, t); break;
        case WFK_SW * 2 + 0: {
            // Instances without the running maximum for WFK_MODE_NOMAX.
            const bool nomax = (pd.mode & WFK_MODE_NOMAX) != 0;
            if constexpr (R == 2) {
                if (pd.mode & WFK_MODE_SWE) {
                    async_strip<WFK_SWE, false, R, W, G, AW, RS, true>(ka, pd, td.pair, td.group, sh, lane, wave,
                                                                      epoch, t);
                    break;
                }
                if (pd.mode & WFK_MODE_PERM4) {
                    if (nomax)
                        async_strip<WFK_SWP, false, R, W, G, AW, RS, true>(ka, pd, td.pair, td.group, sh, lane, wave,
                                                                          epoch, t);
                    else
                        async_strip<WFK_SWP, false, R, W, G, AW, RS>(ka, pd, td.pair, td.group, sh, lane, wave, epoch, t);
                    break;
                }
            }
            if (nomax)
                async_strip<WFK_SW, false, R, W, G, AW, RS, true>(ka, pd, td.pair, td.group, sh, lane, wave, epoch, t);
            else
                async_strip<WFK_SW, false, R, W, G, AW, RS>(ka, pd, td.pair, td.group, sh, lane, wave, epoch, t);
            break;
        }
        default: async_strip<WFK_SW, true, R, W, G, AW, RS>(ka, pd, td.pair, td.group, sh, lane, wave, epoch, t); break;
        }
        __syncthreads();
    }
}

}  // namespace wfk

#define WF_ASYNC(R_, W_, G_) \
    {WF_ENGINE_ASYNC, R_, W_, G_, 64 * ((W_) + 2), reinterpret_cast<const void *>(&wfk::wf_engine_async<R_, W_, G_>)}

extern "C" const WfVariant wf_variants_async[] = {
    WF_ASYNC(1, 4, 16), WF_ASYNC(2, 4, 16), WF_ASYNC(4, 4, 16),
    WF_ASYNC(1, 4, 8), WF_ASYNC(2, 4, 8), WF_ASYNC(2, 4, 32),
    WF_ASYNC(1, 2, 16), WF_ASYNC(2, 2, 16), WF_ASYNC(4, 2, 16),
    WF_ASYNC(1, 1, 16), WF_ASYNC(2, 1, 16), WF_ASYNC(4, 1, 16), WF_ASYNC(8, 1, 16),
    WF_ASYNC(1, 8, 16), WF_ASYNC(2, 8, 16), WF_ASYNC(2, 3, 16), WF_ASYNC(2, 7, 16),
    WF_ASYNC(1, 4, 32), WF_ASYNC(2, 2, 32), WF_ASYNC(1, 2, 32), WF_ASYNC(2, 1, 32), WF_ASYNC(1, 1, 32),
};
extern "C" const int wf_num_variants_async = sizeof(wf_variants_async) / sizeof(wf_variants_async[0]);

#ifdef WFK_TRACE
extern "C" hipError_t wf_trace_read_async(void *host, size_t bytes)
{
    return hipMemcpyFromSymbol(host, HIP_SYMBOL(wfk::g_trace), bytes, 0, hipMemcpyDeviceToHost);
}
#endif
