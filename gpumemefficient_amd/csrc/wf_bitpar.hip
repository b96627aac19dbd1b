// wf_bitpar.hip -- bit-parallel wavefront engine for LCS and edit distance on
// gfx950.
//
// LCS (LCS/LCS.cpp:45-82) and edit distance (EditDistance/ED.cpp:45-81) have
// unit costs, so neighbouring cells differ by -1, 0 or +1 and a column of the
// table is fully described by bit vectors of its vertical differences.  A
// lane holds 32 consecutive rows as one 32-bit word and advances its word by
// one column per step with a fixed sequence of word operations:
//
//   ED   Myers' bit-vector algorithm in Hyyro's formulation, blocked as in
//        Myers (1999) "A fast bit-vector algorithm for approximate string
//        matching based on dynamic programming", J. ACM 46(3), section 4:
//        state Pv/Mv (vertical +1/-1), horizontal deltas Ph/Mh; the word
//        above passes its last row's horizontal delta (hin) down, and a -1
//        enters the carry chain of the addition as Eq | 1.
//   LCS  the bit-vector LCS of Crochemore, Iliopoulos, Pinzon and Reid (2001)
//        "A fast and practical bit-vector algorithm for the longest common
//        subsequence problem", IPL 80(6): V' = (V + (V & Eq)) | (V & ~Eq);
//        the word above passes its addition's carry down.
//
// Eq is the lane's 32-row match mask for the column's symbol, looked up in a
// per-wave LDS table (symbol code x lane) built from the row symbols at strip
// start.  The words of a wave form the same 64-lane systolic array as the
// cell engines (lane l handles column s - l at step s; the word above's
// output reaches lane l through one DPP wave_shr:1), so a wave covers 2048
// rows and a column step costs about 20 VALU operations for 32 cells instead
// of 3-4 per cell.  The chain of waves down a pair is 32 times shorter than
// the cell engines' (one wave per 2048 rows instead of per 128), which is
// where a single large pair spends half its time there.
//
// Scheduling is the lock-step engine's (wf_kernels.hip): W waves per strip
// advance in chunks of K = 64 or 128 steps between barriers, D chunks apart;
// wave 0 stages the column codes and the strip's top boundary into LDS one
// chunk ahead; the last wave stores the bottom row's boundary (ED: hin bits, LCS:
// carry) to HBM as epoch-tagged words for the next strip.  The score comes
// from the last column's vertical differences: ED D(n2,n1) = D(0,n1) + sum of
// Pv - Mv over the rows, LCS = number of zero bits of V, summed over the
// lanes with atomics.
//
// Also here: the staged variant (a staging wave per workgroup looks up the
// match masks for the compute waves) and the combine kernel of pairs split
// into a forward and a reversed column half (WfSplitDev).
#include <hip/hip_runtime.h>
#include <climits>
#include <stdint.h>

#include "wf_cells.h"
#include "wf_device.h"
#include "wf_variants.h"

namespace wfb {

using wfk::kDppWaveShr1;
using wfk::now_ticks;
using wfk::pow2_at_least;

constexpr int kRowsPerWave = 64 * 32;

// Chunk geometry for K steps per chunk: D chunks between consecutive waves,
// ring slots (column c in slot (c + 64) mod RS), code window AW.  A chunk
// must outlast a global-memory round trip (wave 0 loads the next chunk's
// codes and top boundary during it): K = 64 or 128.
template <int W, int K>
struct BpCfg {
    static constexpr int D = (K + 62) / K + 1;
    static constexpr int RS = pow2_at_least((D + 2) * K + 64);
    static constexpr int AW = pow2_at_least((2 + (W - 1) * D) * K + 64 + K);
};

template <int W, int K>
struct BpShared {
    static constexpr int AW = BpCfg<W, K>::AW;
    static constexpr int RS = BpCfg<W, K>::RS;
    // Match masks: peq[wave][code * 64 + lane] = bit r set iff row r of the
    // lane's word holds symbol `code` (codes < 32).
    int peq[W][32 * 64];
    // ring[0]: the strip's top boundary; ring[w+1]: lane 63 of wave w, read
    // by wave w+1 (w = W-1: stored to HBM).  ED: (Ph, Mh) words, whose top
    // bits are the horizontal delta (+1, -1) of the word's last row; LCS:
    // (a word whose bit 31 is the carry, 0).
    __attribute__((aligned(16))) int2 ring[W + 1][RS];
    // Lanes 0..62 store here so that the ring capture is one unmasked write.
    int2 dummy[W][64 + K];
    // Column codes, column c at slot c mod AW, the first K slots mirrored.
    uint32_t aw[AW + K];
    int task;
    int abort_at;
};

struct BpState {
    uint32_t pv, mv;   // ED: vertical +1 / -1 bits; LCS: pv = V
    int ph, mh;        // ED: this column's horizontal deltas (pre-shift); LCS: ph = carry out in bit 31
};

template <int P>
__device__ __forceinline__ void bp_step(BpState &st, int2 inj, uint32_t eq)
{
    if (P == WFK_ED) {
        const int nph = __builtin_amdgcn_update_dpp(inj.x, st.ph, kDppWaveShr1, 0xf, 0xf, false);
        const int nmh = __builtin_amdgcn_update_dpp(inj.y, st.mh, kDppWaveShr1, 0xf, 0xf, false);
        const uint32_t hm = (uint32_t)nmh >> 31;  // hin == -1
        const uint32_t xv = eq | st.mv;
        const uint32_t eq2 = eq | hm;
        const uint32_t xh = (((eq2 & st.pv) + st.pv) ^ st.pv) | eq2;
        const uint32_t ph = st.mv | ~(xh | st.pv);
        const uint32_t mh = st.pv & xh;
        // (h << 1) | the word above's last-row delta.
        const uint32_t phs = __builtin_amdgcn_alignbit(ph, (uint32_t)nph, 31);
        const uint32_t mhs = __builtin_amdgcn_alignbit(mh, (uint32_t)nmh, 31);
        st.pv = mhs | ~(xv | phs);
        st.mv = phs & xv;
        st.ph = (int)ph;
        st.mh = (int)mh;
    } else {
        // st.ph carries the word's carry out in bit 31 (the lane below
        // shifts it down); the sum and its carry are three-input operations.
        const int cw = __builtin_amdgcn_update_dpp(inj.x, st.ph, kDppWaveShr1, 0xf, 0xf, false);
        const uint32_t u = st.pv & eq;
        const uint32_t sum = st.pv + u + ((uint32_t)cw >> 31);
        // Carry out of bit 31: majority of the operands' top bits and the
        // carry into it, i.e. (V & U) | ((V | U) & ~sum).
        st.ph = (int)((st.pv & u) | ((st.pv | u) & ~sum));
        st.pv = sum | (st.pv & ~u);
    }
}

// One chunk of K steps from s0 for one wave.  MASKED: some lane's column is
// outside 0..n1-1; such lanes keep their state (so the lane below reads the
// delta or carry of the lane's last real column).
template <int P, int K, int AW, int RS, bool MASKED>
__device__ __forceinline__ void bp_chunk(BpState &st, int s0, int lane, int n1, const int2 *ring_prev, int2 *ring_cur,
                                         int2 *dummy, const uint32_t *aw, const int *peq)
{
    // Lane 0's column c = s0 + q lives in slot (c + 64) mod RS; a run of 64
    // slots from a multiple of 64 never wraps (K and RS are multiples of 64).
    const uint32_t *__restrict__ cp = aw + ((s0 - lane) & (AW - 1));
    int2 *cap = (lane == 63) ? ring_cur + ((s0 + 1) & (RS - 1)) : dummy + lane;
    int2 *cap_last = (lane == 63) ? ring_cur + ((s0 + K) & (RS - 1)) : dummy + lane + (K - 1);
    // Operands are read ahead of their group of 4 steps: the column codes
    // two groups ahead, the match masks they select and lane 0's upper
    // words one group ahead, so no LDS latency (nor the code -> mask
    // dependency) sits in front of a group's arithmetic.
    auto load_codes = [&](int q, uint32_t (&c)[4]) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            c[u] = cp[q + u];
            if (MASKED) c[u] &= 31u;  // window slots of columns outside the pair hold stale codes
        }
    };
    auto load_eq = [&](const uint32_t (&c)[4], uint32_t (&e)[4]) {
#pragma unroll
        for (int u = 0; u < 4; ++u) e[u] = (uint32_t)peq[c[u] * 64 + lane];
    };
    auto load_inj = [&](int q, int2 (&v)[4]) {
        const int2 *ib = ring_prev + ((s0 + 64 + (q & ~63)) & (RS - 1)) + (q & 63);
        const int4 i01 = *reinterpret_cast<const int4 *>(ib);
        const int4 i23 = *reinterpret_cast<const int4 *>(ib + 2);
        v[0] = make_int2(i01.x, i01.y);
        v[1] = make_int2(i01.z, i01.w);
        v[2] = make_int2(i23.x, i23.y);
        v[3] = make_int2(i23.z, i23.w);
    };
    uint32_t code_n[4], eqv[4];
    int2 injv[4];
    {
        uint32_t c0[4];
        load_codes(0, c0);
        load_codes(4, code_n);
        load_eq(c0, eqv);
        load_inj(0, injv);
    }
#pragma unroll
    for (int q = 0; q < K; q += 4) {
        uint32_t eqn[4] = {0u, 0u, 0u, 0u}, code_nn[4] = {0u, 0u, 0u, 0u};
        int2 injn[4] = {};
        if (q + 4 < K) {
            load_eq(code_n, eqn);
            load_inj(q + 4, injn);
        }
        if (q + 8 < K) load_codes(q + 8, code_nn);
        int2 capv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (MASKED) {
                const bool valid = (unsigned)(s0 + q + u - lane) < (unsigned)n1;
                BpState nx = st;
                bp_step<P>(nx, injv[u], eqv[u]);
                if (valid) st = nx;
            } else {
                bp_step<P>(st, injv[u], eqv[u]);
            }
            capv[u] = make_int2(st.ph, st.mh);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int j = q + u;
            if (j == K - 1)
                *cap_last = capv[u];
            else
                cap[j] = capv[u];
        }
        if (q + 4 < K) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                eqv[u] = eqn[u];
                injv[u] = injn[u];
                code_n[u] = code_nn[u];
            }
        }
    }
}

// One strip (`g`) of one pair, swept by the whole workgroup.
template <int P, int W, int K>
__device__ void bp_group(const WfKernelArgs &ka, const WfPairDev &pd, int pair, int g, BpShared<W, K> &sh, int lane,
                         int wave, uint32_t epoch)
{
    constexpr int AW = BpCfg<W, K>::AW;
    constexpr int RS = BpCfg<W, K>::RS;
    constexpr int kD = BpCfg<W, K>::D;
    constexpr int kK = K;
    const int n1 = pd.n1, n2 = pd.n2;
    const int64_t row0 = (int64_t)g * (W * kRowsPerWave) + wave * kRowsPerWave + lane * 32;
    const uint8_t *acodes = ka.codes + pd.a_off;
    const uint8_t *bcodes = ka.codes + pd.b_off;
    const int64_t rows_left = (int64_t)n2 - (int64_t)g * (W * kRowsPerWave);
    const int wa = (int)min<int64_t>(W, (rows_left + kRowsPerWave - 1) / kRowsPerWave);
    const bool active = wave < wa;
    const int nvalid = (int)max<int64_t>(0, min<int64_t>(32, (int64_t)n2 - row0));
    const uint32_t rowmask = nvalid >= 32 ? 0xffffffffu : ((1u << nvalid) - 1u);
    const bool textbook = (pd.mode & WFK_MODE_TEXTBOOK) != 0;

    // Match masks of this wave's rows.
    int *peq = sh.peq[wave];
    if (active) {
        for (int i = lane; i < 32 * 64; i += 64) peq[i] = 0;
        for (int r = 0; r < nvalid; ++r) {
            const uint32_t code = bcodes[row0 + r] & 31u;
            atomicOr(&peq[code * 64 + lane], (int)(1u << r));
        }
    }

    const int nsteps = n1 + 63;
    const int nch = (nsteps + kK - 1) / kK;
    const int total = nch + (wa - 1) * kD;
    const bool has_prev = g > 0;
    const bool has_next = g + 1 < pd.ngroups;
    const uint64_t *bprev = ka.bnd + pd.bnd_off + (int64_t)(g - 1) * n1;
    uint64_t *bnext = ka.bnd + pd.bnd_off + (int64_t)g * n1;
    const uint64_t tag = (uint64_t)epoch << 32;

    // Top boundary and column codes of columns [c*K, c*K+K), staged by wave
    // 0 in K/64 batches of 64 columns (one per lane): the previous strip's
    // bottom-row words, or row 0 of the table (ED: delta 0 at column 1 and +1
    // after it, ED.cpp:48-54; +1 throughout for the textbook row; LCS: no
    // carry).
    constexpr int NB = K / 64;
    static_assert(K % 64 == 0, "K must be a multiple of 64");
    struct Pre {
        uint64_t top[NB];
        uint32_t code[NB];
    };
    auto load_pre = [&](int c, Pre &pr) {
#pragma unroll
        for (int h = 0; h < NB; ++h) {
            const int x = c * kK + h * 64 + lane;
            pr.code[h] = x < n1 ? (uint32_t)acodes[x] : 0u;
            pr.top[h] = (!has_prev || x >= n1) ? tag
                                               : __hip_atomic_load(bprev + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    };
    auto fill = [&](int c, Pre &pr) -> bool {
#pragma unroll
        for (int h = 0; h < NB; ++h) {
            const int x = c * kK + h * 64 + lane;
            uint64_t top = pr.top[h];
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
            if (x < n1) {
                uint32_t v;
                if (has_prev)
                    v = (uint32_t)top;
                else
                    v = (P == WFK_ED && (x >= 1 || textbook)) ? 1u : 0u;
                sh.ring[0][(x + 64) & (RS - 1)] = (P == WFK_ED)
                                                      ? make_int2((int)((v & 1u) << 31), (int)(((v >> 1) & 1u) << 31))
                                                      : make_int2((int)((v & 1u) << 31), 0);
                const int slot = x & (AW - 1);
                sh.aw[slot] = pr.code[h];
                if (slot < kK) sh.aw[AW + slot] = pr.code[h];
            }
        }
        return true;
    };

    if (wave == 0) {
        Pre p0;
        load_pre(0, p0);
        if (!fill(0, p0) && lane == 0) sh.abort_at = -1;
    }
    __syncthreads();

    int copied = 0;  // wave W-1: columns already stored to bnext
    auto copy_out = [&](int upto, bool from_ring) {
        while (copied < upto) {
            const int x = copied + lane;
            if (x < upto) {
                uint32_t v = 0;
                if (from_ring) {
                    const int2 w = sh.ring[W][(x + 64) & (RS - 1)];
                    v = (P == WFK_ED) ? (((uint32_t)w.x >> 31) | (((uint32_t)w.y >> 31) << 1)) : ((uint32_t)w.x >> 31);
                }
                __hip_atomic_store(bnext + x, tag | v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            copied = min(upto, copied + 64);
        }
    };

    BpState st;
    st.pv = (P == WFK_ED) ? 0xffffffffu : 0xffffffffu;  // ED: column 0 rises by 1 per row; LCS: V = all ones
    st.mv = 0u;
    st.ph = 0;
    st.mh = 0;

    bool aborted = sh.abort_at <= -1;
    for (int j = 0; j < total && !aborted; ++j) {
        const int cj = j - wave * kD;
        Pre pre;
        if (wave == 0 && j + 1 < nch) load_pre(j + 1, pre);
        if (cj >= 0 && cj < nch && active) {
            const int s0 = cj * kK;
            const bool steady = (s0 >= 63) && (s0 + kK - 1 <= n1 - 1);
            if (steady)
                bp_chunk<P, K, AW, RS, false>(st, s0, lane, n1, sh.ring[wave], sh.ring[wave + 1], sh.dummy[wave], sh.aw, peq);
            else
                bp_chunk<P, K, AW, RS, true>(st, s0, lane, n1, sh.ring[wave], sh.ring[wave + 1], sh.dummy[wave], sh.aw, peq);
            // Steps < s0 + K are done: lane 63 has finished columns < s0 + K - 63.
            if (wave == W - 1 && has_next) copy_out(min(n1, s0 + kK - 63), true);
        }
        if (wave == 0 && j + 1 < nch) {
            if (!fill(j + 1, pre) && lane == 0) sh.abort_at = j;
        }
        __syncthreads();
        aborted = sh.abort_at <= j;
    }

    if (aborted) {
        if (threadIdx.x == 0) atomicOr(&ka.res[pair].status, 1);
        if (wave == W - 1 && has_next) copy_out(n1, false);
        return;
    }
    if (wave == W - 1 && has_next) copy_out(n1, true);

    // Score: the last column's vertical differences over the rows.
    if (active) {
        int v;
        if (P == WFK_ED)
            v = __builtin_popcount(st.pv & rowmask) - __builtin_popcount(st.mv & rowmask);
        else
            v = __builtin_popcount(~st.pv & rowmask);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (P == WFK_ED && g == 0 && wave == 0) v += textbook ? n1 : n1 - 1;  // D(0, n1)
        if (lane == 0) atomicAdd(&ka.res[pair].last, v);
        if (pd.col_off >= 0) {
            uint32_t *cw = ka.colw + 2 * (pd.col_off + ((int64_t)g * W + wave) * 64 + lane);
            cw[0] = st.pv;
            cw[1] = st.mv;
        }
    }
}

template <int W, int K>
__global__ __launch_bounds__(64 * W) void wf_engine_bitpar(WfKernelArgs ka)
{
    __shared__ BpShared<W, K> sh;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t epoch = __builtin_amdgcn_readfirstlane(*ka.epoch);
    // Codes in the window are always < 32, even in slots no strip wrote yet.
    for (int i = threadIdx.x; i < BpShared<W, K>::AW + K; i += blockDim.x) sh.aw[i] = 0;
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
        wfk::set_wave_priority(pd.prio);
        if (pd.prob == WFK_ED)
            bp_group<WFK_ED, W, K>(ka, pd, td.pair, td.group, sh, lane, wave, epoch);
        else
            bp_group<WFK_LCS, W, K>(ka, pd, td.pair, td.group, sh, lane, wave, epoch);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Staged variant: one more wave per workgroup looks up the match masks.
//
// In the engine above a compute wave spends about a sixth of its
// instructions per step on finding Eq (the column code's LDS read, the table
// address, the table read).  Here a staging wave does those lookups for the
// compute waves one chunk ahead and leaves each lane's masks in LDS in step
// order, four steps per 16-byte word, so a compute wave fetches the masks of
// four steps with one ds_read_b128 and no address arithmetic (ED: 16.8
// instructions per step instead of 19.6).  The staging wave also loads the
// column codes and the strip's top boundary and stores the bottom row, so
// the compute waves issue no global memory operation at all.  The mask
// buffers take 16 KB of LDS per compute wave and chunk parity (W = 1: 51 KB
// per workgroup): the planner takes this variant for launches whose strips
// fit on the chip at once (single large pairs), and the engine above, which
// packs more compute waves per CU, for throughput-bound batches.  Measured
// on ED 65536^2: W = 1 2.75 ms (plain engine 3.23 ms at its best W = 2);
// W = 2 3.14 ms, W = 3 3.93 ms (one staging wave cannot feed more).
// ---------------------------------------------------------------------------

template <int W>
struct BsCfg {
    static constexpr int K = 64;
    static constexpr int D = (K + 62) / K + 1;
    static constexpr int RS = pow2_at_least((D + 2) * K + 64);
    static constexpr int AW = pow2_at_least((2 + (W - 1) * D) * K + 64 + K);
};

template <int W>
struct BsShared {
    static constexpr int K = BsCfg<W>::K;
    static constexpr int AW = BsCfg<W>::AW;
    static constexpr int RS = BsCfg<W>::RS;
    // eqb[i & 1][w][q / 4][lane]: lane's match masks for steps q..q+3 of
    // compute wave w's chunk in iteration i (written during iteration i - 1).
    int4 eqb[2][W][K / 4][64];
    // As in BpShared.
    __attribute__((aligned(16))) int2 ring[W + 1][RS];
    int peq[W][32 * 64];
    int2 dummy[W][64 + K];
    uint32_t aw[AW + K];
    int task;
    int abort_at;
};

// One chunk of K steps of a compute wave; masks from `eqw` (= eqb[i & 1][w]).
template <int P, int K, int RS, bool MASKED>
__device__ __forceinline__ void bs_chunk(BpState &st, int s0, int lane, int n1, const int2 *ring_prev, int2 *ring_cur,
                                         int2 *dummy, const int4 *eqw)
{
    int2 *cap = (lane == 63) ? ring_cur + ((s0 + 1) & (RS - 1)) : dummy + lane;
    int2 *cap_last = (lane == 63) ? ring_cur + ((s0 + K) & (RS - 1)) : dummy + lane + (K - 1);
    auto load_inj = [&](int q, int2 (&v)[4]) {
        const int2 *ib = ring_prev + ((s0 + 64 + (q & ~63)) & (RS - 1)) + (q & 63);
        const int4 i01 = *reinterpret_cast<const int4 *>(ib);
        const int4 i23 = *reinterpret_cast<const int4 *>(ib + 2);
        v[0] = make_int2(i01.x, i01.y);
        v[1] = make_int2(i01.z, i01.w);
        v[2] = make_int2(i23.x, i23.y);
        v[3] = make_int2(i23.z, i23.w);
    };
    const int4 *ep = eqw + lane;
    // Masks and lane 0's upper words are read one group of 4 steps ahead.
    int4 eqv = ep[0];
    int2 injv[4];
    load_inj(0, injv);
#pragma unroll
    for (int q = 0; q < K; q += 4) {
        int4 eqn = make_int4(0, 0, 0, 0);
        int2 injn[4] = {};
        if (q + 4 < K) {
            eqn = ep[(q + 4) * 16];
            load_inj(q + 4, injn);
        }
        const uint32_t e[4] = {(uint32_t)eqv.x, (uint32_t)eqv.y, (uint32_t)eqv.z, (uint32_t)eqv.w};
        int2 capv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (MASKED) {
                const bool valid = (unsigned)(s0 + q + u - lane) < (unsigned)n1;
                BpState nx = st;
                bp_step<P>(nx, injv[u], e[u]);
                if (valid) st = nx;
            } else {
                bp_step<P>(st, injv[u], e[u]);
            }
            capv[u] = make_int2(st.ph, st.mh);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int j = q + u;
            if (j == K - 1)
                *cap_last = capv[u];
            else
                cap[j] = capv[u];
        }
        if (q + 4 < K) {
            eqv = eqn;
#pragma unroll
            for (int u = 0; u < 4; ++u) injv[u] = injn[u];
        }
    }
}

// One strip (`g`) of one pair: waves 0..W-1 compute, wave W stages.
template <int P, int W>
__device__ void bs_group(const WfKernelArgs &ka, const WfPairDev &pd, int pair, int g, BsShared<W> &sh, int lane,
                         int wave, uint32_t epoch)
{
    constexpr int K = BsCfg<W>::K;
    constexpr int AW = BsCfg<W>::AW;
    constexpr int RS = BsCfg<W>::RS;
    constexpr int kD = BsCfg<W>::D;
    const bool stager = wave == W;
    const int n1 = pd.n1, n2 = pd.n2;
    const int64_t row0 = (int64_t)g * (W * kRowsPerWave) + wave * kRowsPerWave + lane * 32;
    const uint8_t *acodes = ka.codes + pd.a_off;
    const uint8_t *bcodes = ka.codes + pd.b_off;
    const int64_t rows_left = (int64_t)n2 - (int64_t)g * (W * kRowsPerWave);
    const int wa = (int)min<int64_t>(W, (rows_left + kRowsPerWave - 1) / kRowsPerWave);
    const bool active = !stager && wave < wa;
    const int nvalid = active ? (int)max<int64_t>(0, min<int64_t>(32, (int64_t)n2 - row0)) : 0;
    const uint32_t rowmask = nvalid >= 32 ? 0xffffffffu : ((1u << nvalid) - 1u);
    const bool textbook = (pd.mode & WFK_MODE_TEXTBOOK) != 0;

    if (active) {
        int *peq = sh.peq[wave];
        for (int i = lane; i < 32 * 64; i += 64) peq[i] = 0;
        for (int r = 0; r < nvalid; ++r) {
            const uint32_t code = bcodes[row0 + r] & 31u;
            atomicOr(&peq[code * 64 + lane], (int)(1u << r));
        }
    }

    const int nsteps = n1 + 63;
    const int nch = (nsteps + K - 1) / K;
    const int total = nch + (wa - 1) * kD;
    const bool has_prev = g > 0;
    const bool has_next = g + 1 < pd.ngroups;
    const uint64_t *bprev = ka.bnd + pd.bnd_off + (int64_t)(g - 1) * n1;
    uint64_t *bnext = ka.bnd + pd.bnd_off + (int64_t)g * n1;
    const uint64_t tag = (uint64_t)epoch << 32;

    // Staging wave: column codes of chunk c (column c*K + lane per lane) into
    // the code window, the strip's top boundary of chunk c into ring[0]
    // (values as in bp_group), and the masks of an iteration's chunks.
    auto load_code = [&](int c) -> uint32_t {
        const int x = c * K + lane;
        return x < n1 ? (uint32_t)acodes[x] : 0u;
    };
    auto put_code = [&](int c, uint32_t code) {
        const int x = c * K + lane;
        if (x < n1) {
            const int slot = x & (AW - 1);
            sh.aw[slot] = code;
            if (slot < K) sh.aw[AW + slot] = code;
        }
    };
    auto load_top = [&](int c) -> uint64_t {
        const int x = c * K + lane;
        return (!has_prev || x >= n1) ? tag : __hip_atomic_load(bprev + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    };
    auto fill_top = [&](int c, uint64_t top) -> bool {
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
        if (x < n1) {
            uint32_t v;
            if (has_prev)
                v = (uint32_t)top;
            else
                v = (P == WFK_ED && (x >= 1 || textbook)) ? 1u : 0u;
            sh.ring[0][(x + 64) & (RS - 1)] = (P == WFK_ED)
                                                  ? make_int2((int)((v & 1u) << 31), (int)(((v >> 1) & 1u) << 31))
                                                  : make_int2((int)((v & 1u) << 31), 0);
        }
        return true;
    };
    // Masks of iteration `it`: compute wave w runs chunk it - w*D; its lane
    // l at step q reads column (it - w*D)*K + q - l (codes of columns
    // outside the pair are stale, masked to a table row; those steps keep
    // their state).
    auto produce = [&](int it) {
        for (int w = 0; w < wa; ++w) {
            const int c = it - w * kD;
            if (c < 0 || c >= nch) continue;
            const uint32_t *cp = sh.aw + ((c * K - lane) & (AW - 1));
            const int *pq = sh.peq[w] + lane;
            int4 *d = &sh.eqb[it & 1][w][0][lane];
            // All codes, then all table reads, then the stores: the table
            // reads issue back to back instead of each group waiting out two
            // dependent LDS round trips.
            uint32_t code[K];
#pragma unroll
            for (int q = 0; q < K; ++q) code[q] = cp[q];
            int e[K];
#pragma unroll
            for (int q = 0; q < K; ++q) e[q] = pq[(code[q] & 31u) * 64];
#pragma unroll
            for (int gq = 0; gq < K / 4; ++gq) d[gq * 64] = make_int4(e[4 * gq], e[4 * gq + 1], e[4 * gq + 2], e[4 * gq + 3]);
        }
    };

    __syncthreads();  // match-mask tables built
    uint32_t code_next = 0u;  // stager: codes of chunk j + 2, loaded in iteration j
    if (stager) {
        put_code(0, load_code(0));
        if (1 < nch) code_next = load_code(1);
        produce(0);
        if (!fill_top(0, load_top(0)) && lane == 0) sh.abort_at = -1;
    }
    __syncthreads();

    int copied = 0;  // stager: columns already stored to bnext
    auto copy_out = [&](int upto, bool from_ring) {
        while (copied < upto) {
            const int x = copied + lane;
            if (x < upto) {
                uint32_t v = 0;
                if (from_ring) {
                    const int2 w = sh.ring[W][(x + 64) & (RS - 1)];
                    v = (P == WFK_ED) ? (((uint32_t)w.x >> 31) | (((uint32_t)w.y >> 31) << 1)) : ((uint32_t)w.x >> 31);
                }
                __hip_atomic_store(bnext + x, tag | v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            copied = min(upto, copied + 64);
        }
    };

    BpState st;
    st.pv = 0xffffffffu;  // ED: column 0 rises by 1 per row; LCS: V = all ones
    st.mv = 0u;
    st.ph = 0;
    st.mh = 0;

    bool aborted = sh.abort_at <= -1;
    for (int j = 0; j < total && !aborted; ++j) {
        if (stager) {
            // The bottom row's columns the last compute wave finished in
            // iteration j - 1, to HBM for the strip below (here rather than
            // in that wave, whose next chunk would otherwise wait for the
            // stores to complete).
            const int cl = j - 1 - (wa - 1) * kD;
            if (has_next && cl >= 0 && cl < nch) copy_out(min(n1, cl * K + K - 63), true);
            if (j + 1 < nch) put_code(j + 1, code_next);
            if (j + 2 < nch) code_next = load_code(j + 2);
            const uint64_t top = (j + 1 < nch) ? load_top(j + 1) : tag;
            if (j + 1 < total) produce(j + 1);
            if (j + 1 < nch && !fill_top(j + 1, top) && lane == 0) sh.abort_at = j;
        } else {
            const int cj = j - wave * kD;
            if (cj >= 0 && cj < nch && active) {
                const int s0 = cj * K;
                const bool steady = (s0 >= 63) && (s0 + K - 1 <= n1 - 1);
                const int4 *eqw = &sh.eqb[j & 1][wave][0][0];
                if (steady)
                    bs_chunk<P, K, RS, false>(st, s0, lane, n1, sh.ring[wave], sh.ring[wave + 1], sh.dummy[wave], eqw);
                else
                    bs_chunk<P, K, RS, true>(st, s0, lane, n1, sh.ring[wave], sh.ring[wave + 1], sh.dummy[wave], eqw);
            }
        }
        __syncthreads();
        aborted = sh.abort_at <= j;
    }

    if (aborted) {
        if (threadIdx.x == 0) atomicOr(&ka.res[pair].status, 1);
        if (stager && has_next) copy_out(n1, false);
        return;
    }
    if (stager && has_next) copy_out(n1, true);

    if (active) {
        int v;
        if (P == WFK_ED)
            v = __builtin_popcount(st.pv & rowmask) - __builtin_popcount(st.mv & rowmask);
        else
            v = __builtin_popcount(~st.pv & rowmask);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (P == WFK_ED && g == 0 && wave == 0) v += textbook ? n1 : n1 - 1;  // D(0, n1)
        if (lane == 0) atomicAdd(&ka.res[pair].last, v);
        if (pd.col_off >= 0) {
            uint32_t *cw = ka.colw + 2 * (pd.col_off + ((int64_t)g * W + wave) * 64 + lane);
            cw[0] = st.pv;
            cw[1] = st.mv;
        }
    }
}

// ---------------------------------------------------------------------------
// Split pairs: the score from the two halves' last-column words.
//
// Row r (0-based, r < n2) of the forward half's column mid changes by
// d_r = Pv_r - Mv_r (ED) or z_r = 1 - V_r (LCS, the LCS grows by one where V
// has a zero); F(y) = f0 + sum_{r<y} d_r.  The reverse half's last column
// at reverse row n2 - y is R(y) = r0 + T' - sum_{r<y} d'_{n2-1-r}, T' the sum
// of all its deltas.  So F(y) + R(y) = f0 + r0 + T' + sum_{r<y} e_r with
// e_r = d_r - d'_{n2-1-r}, and the score is that constant plus the smallest
// (ED) or largest (LCS) prefix sum of e, the empty prefix included.  One
// workgroup per split; the reverse words are read bit-reversed so that both
// halves line up row by row.
// ---------------------------------------------------------------------------

// Column-split SW pair (prob WFK_SW): job holds columns 1..mid, f0 the
// columns mid+1..n1 from a zero boundary (its bottom-right cell R0: the best
// local alignment starting right of the cut), rev the reversed right half,
// whose last column E(i) gives the best path weight from (n2 - i, mid) to
// (n2, n1) as E(i) - 2 (i + r0).  The bottom-right cell is
// max(R0, max_y H(y, mid) + that weight), H(0, mid) = 0, E(0) = 0.
__device__ void sw_split_combine(const WfSplitDev &sp, const uint32_t *colw, WfResultDev *res)
{
    // Bottom-right cell: an alignment into (n2, n1) starts right of column
    // mid (pair f0's cell) or leaves column mid last at some row y after a
    // best local alignment into (y, mid), H(y, mid), and then takes the best
    // path from (y, mid) to the corner, E(n2 - y) - 2 (n2 - y + r0) from the
    // unclamped reverse pair (E(0) = H(0, mid) = 0).  The running maximum
    // likewise: the halves' own maxima, or H(y, mid) plus the best path from
    // (y, mid) to any cell right of it, C(n2 - y) from the clamped reverse
    // pair (C(0) = 0).
    const int n2 = sp.n2, nr = sp.r0;
    const bool mx = sp.crev >= 0;
    const int *h = reinterpret_cast<const int *>(colw) + 2 * sp.fwd_off;
    const int *e = reinterpret_cast<const int *>(colw) + 2 * sp.rev_off;
    const int *c = reinterpret_cast<const int *>(colw) + 2 * sp.crev_off;
    int best = INT_MIN, bestm = 0;
    for (int y = threadIdx.x; y <= n2; y += blockDim.x) {
        const int i = n2 - y;
        const int hv = y > 0 ? h[y - 1] : 0;
        const int ev = i > 0 ? e[i - 1] : 0;
        best = max(best, hv + ev - 2 * (i + nr));
        if (mx) bestm = max(bestm, hv + (i > 0 ? c[i - 1] : 0));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        best = max(best, __shfl_xor(best, off));
        bestm = max(bestm, __shfl_xor(bestm, off));
    }
    __shared__ int wb[16], wm[16];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        wb[wv] = best;
        wm[wv] = bestm;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int b = wb[0], m = wm[0];
        for (int k = 1; k < (int)(blockDim.x >> 6); ++k) {
            b = max(b, wb[k]);
            m = max(m, wm[k]);
        }
        WfResultDev &r = res[sp.job];
        r.last = max(b, res[sp.f0].last);
        r.status |= res[sp.rev].status | res[sp.f0].status;
        if (mx) {
            r.maxs = max(max(m, r.maxs), res[sp.f0].maxs);
            r.status |= res[sp.crev].status;
        }
    }
}

__global__ __launch_bounds__(1024) void wf_bitpar_combine(const WfSplitDev *splits, const uint32_t *colw,
                                                          WfResultDev *res)
{
    const WfSplitDev sp = splits[blockIdx.x];
    if (sp.prob == WFK_SW) {
        sw_split_combine(sp, colw, res);
        return;
    }
    const bool ed = sp.prob == WFK_ED;
    const int n2 = sp.n2;
    const int nw = (n2 + 31) >> 5;
    const uint32_t *fw = colw + 2 * sp.fwd_off;
    const uint32_t *rw = colw + 2 * sp.rev_off;
    // Bits of reverse rows [base, base + 32) (base > -32), bit k = row
    // base + k; rows below 0 read as 0.  h: 0 Pv/V, 1 Mv.
    auto rev_bits = [&](int base, int h) -> uint32_t {
        const int lo = base >= 0 ? (base >> 5) : -1;
        const int sh = base - 32 * lo;
        const uint32_t wl = lo >= 0 ? rw[2 * lo + h] : 0u;
        const uint32_t wh = (lo + 1 < nw) ? rw[2 * (lo + 1) + h] : 0u;
        return (uint32_t)((((uint64_t)wh << 32) | wl) >> sh);
    };
    const int tid = threadIdx.x;
    const int per = (nw + 1023) >> 10;
    const int w0 = min(nw, tid * per), w1 = min(nw, w0 + per);
    int run = 0, best = 0, rtot = 0;
    for (int w = w0; w < w1; ++w) {
        const int nk = min(32, n2 - 32 * w);
        const uint32_t m = nk >= 32 ? 0xffffffffu : ((1u << nk) - 1u);
        const int base = n2 - 32 * w - 32;  // reverse rows of original rows 32w+31 .. 32w
        uint32_t fa, fb, ra, rb;
        if (ed) {
            fa = fw[2 * w] & m;
            fb = fw[2 * w + 1] & m;
            ra = __builtin_bitreverse32(rev_bits(base, 0)) & m;
            rb = __builtin_bitreverse32(rev_bits(base, 1)) & m;
            rtot += __builtin_popcount(ra) - __builtin_popcount(rb);
        } else {
            fa = ~fw[2 * w] & m;  // zeros of V: the LCS grows at that row
            fb = 0u;
            ra = __builtin_bitreverse32(~rev_bits(base, 0)) & m;
            rb = 0u;
            rtot += __builtin_popcount(ra);
        }
        for (int k = 0; k < nk; ++k) {
            run += (int)((fa >> k) & 1u) - (int)((fb >> k) & 1u) - (int)((ra >> k) & 1u) + (int)((rb >> k) & 1u);
            best = ed ? min(best, run) : max(best, run);
        }
    }
    // Exclusive scan of the threads' sums (in word order) and the reductions.
    __shared__ int wsum[16], wbest[16], wrt[16];
    const int lane = tid & 63, wv = tid >> 6;
    int incl = run;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int t = __shfl_up(incl, off);
        if (lane >= off) incl += t;
    }
    if (lane == 63) wsum[wv] = incl;
    int rt = rtot;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) rt += __shfl_xor(rt, off);
    if (lane == 0) wrt[wv] = rt;
    __syncthreads();
    int before = 0, rall = 0;
    for (int i = 0; i < 16; ++i) {
        if (i < wv) before += wsum[i];
        rall += wrt[i];
    }
    int cand = before + incl - run + best;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const int o = __shfl_xor(cand, off);
        cand = ed ? min(cand, o) : max(cand, o);
    }
    if (lane == 0) wbest[wv] = cand;
    __syncthreads();
    if (tid == 0) {
        int b = wbest[0];
        for (int i = 1; i < 16; ++i) b = ed ? min(b, wbest[i]) : max(b, wbest[i]);
        res[sp.job].last = sp.f0 + sp.r0 + rall + b;
        res[sp.job].status |= res[sp.rev].status;
    }
}

template <int W>
__global__ __launch_bounds__(64 * (W + 1)) void wf_engine_bitpar_staged(WfKernelArgs ka)
{
    __shared__ BsShared<W> sh;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t epoch = __builtin_amdgcn_readfirstlane(*ka.epoch);
    for (int i = threadIdx.x; i < BsShared<W>::AW + BsShared<W>::K; i += blockDim.x) sh.aw[i] = 0;
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
        wfk::set_wave_priority(pd.prio);
        if (pd.prob == WFK_ED)
            bs_group<WFK_ED, W>(ka, pd, td.pair, td.group, sh, lane, wave, epoch);
        else
            bs_group<WFK_LCS, W>(ka, pd, td.pair, td.group, sh, lane, wave, epoch);
        __syncthreads();
    }
}

}  // namespace wfb

extern "C" hipError_t wf_launch_bitpar_combine(const WfSplitDev *splits, int nsplits, const uint32_t *colw,
                                               WfResultDev *res, hipStream_t stream)
{
    if (nsplits <= 0) return hipSuccess;
    hipLaunchKernelGGL(wfb::wf_bitpar_combine, dim3(nsplits), dim3(1024), 0, stream, splits, colw, res);
    return hipGetLastError();
}

#define WF_BITPAR_STAGED(W_) \
    {WF_ENGINE_BITPAR_STAGED, 32, W_, 64, 64 * ((W_) + 1), reinterpret_cast<const void *>(&wfb::wf_engine_bitpar_staged<W_>)}

#define WF_BITPAR(W_, K_) \
    {WF_ENGINE_BITPAR, 32, W_, K_, 64 * (W_), reinterpret_cast<const void *>(&wfb::wf_engine_bitpar<W_, K_>)}

extern "C" const WfVariant wf_variants_bitpar[] = {WF_BITPAR(1, 64),        WF_BITPAR(2, 64),        WF_BITPAR(4, 64),
                                                   WF_BITPAR(1, 128),       WF_BITPAR(2, 128),       WF_BITPAR(4, 128),
                                                   WF_BITPAR_STAGED(1),     WF_BITPAR_STAGED(2)};
extern "C" const int wf_num_variants_bitpar = sizeof(wf_variants_bitpar) / sizeof(wf_variants_bitpar[0]);
