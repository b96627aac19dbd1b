// wf_cells.h -- device building blocks shared by the wavefront engines:
// cell functions (the three reference recurrences), per-lane state and the
// one-column step of a lane.  Included by wf_kernels.hip and wf_engine.hip.
#ifndef WF_CELLS_H
#define WF_CELLS_H
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wf_device.h"

namespace wfk {

constexpr int kDppWaveShr1 = 0x138;  // DPP_CTRL wave_shr:1 (GFX9 family)

constexpr int pow2_at_least(int v)
{
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

// Chunk geometry for chunk width K (steps between barriers).
template <int K>
struct ChunkGeom {
    static_assert(K == 16 || K == 32 || K == 64, "K must divide 64");
    // Lag in chunks between consecutive waves: lane 0 of wave w reads, at
    // chunk-aligned groups of 4 columns, values lane 63 of wave w-1 wrote
    // 63..66 steps later in its own step count.
    static constexpr int D = (K + 62) / K + 1;
    // LDS ring of bottom-lane values, indexed by (column + 64) mod RS.
    static constexpr int RS = pow2_at_least((D + 2) * K);
};

// ----------------------------------------------------------------------------
// Cell functions.  `ca` is the column symbol code, `t` the lane's row entry.
// BFE mode: t is a per-row table, ca a bit offset into it.
// CMP mode: t is the row symbol code (0x100 for padding rows), ca the code.
// ----------------------------------------------------------------------------
template <int P, bool CMP>
struct Cell;

// Cell::idle: the starting state of a lane that runs unmasked steps left of
// column 0.  Every such step maps values at columns <= -2 to values that
// stay far from the table's (min for ED, max for LCS), and column -1 to the
// left boundary 0 exactly, because the lane above delivers 0 there.  SW
// starts at 0 and relies on a never-matching column code instead (its
// saturation at 0 keeps no sentinel).
constexpr int kIdleBig = 1 << 29;

template <bool CMP>
struct Cell<WFK_LCS, CMP> {
    static constexpr int idle = -kIdleBig;
    static __device__ __forceinline__ uint32_t row_entry(uint32_t bc, bool valid)
    {
        if (CMP) return valid ? bc : 0x100u;
        return valid ? (1u << bc) : 0u;
    }
    static __device__ __forceinline__ int f(int up, int left, int diag, uint32_t ca, uint32_t t)
    {
        const int eq = CMP ? (int)(ca == t) : (int)__builtin_amdgcn_ubfe(t, ca, 1);
        return max(max(up, left), diag + eq);
    }
};

template <bool CMP>
struct Cell<WFK_ED, CMP> {
    static constexpr int idle = kIdleBig;
    static __device__ __forceinline__ uint32_t row_entry(uint32_t bc, bool valid)
    {
        if (CMP) return valid ? bc : 0x100u;
        return valid ? ~(1u << bc) : 0xffffffffu;
    }
    static __device__ __forceinline__ int f(int up, int left, int diag, uint32_t ca, uint32_t t)
    {
        const int ne = CMP ? (int)(ca != t) : (int)__builtin_amdgcn_ubfe(t, ca, 1);
        return min(min(up, left), diag + ne - 2);
    }
};

template <bool CMP>
struct Cell<WFK_SW, CMP> {
    static constexpr int idle = 0;
    // BFE: 4-bit signed field at bit 4*code holds s+2: 5 on a match, -1 else.
    static __device__ __forceinline__ uint32_t row_entry(uint32_t bc, bool valid)
    {
        if (CMP) return valid ? bc : 0x100u;
        return valid ? ((~(0xfu << bc)) | (5u << bc)) : 0xffffffffu;
    }
    static __device__ __forceinline__ int f(int up, int left, int diag, uint32_t ca, uint32_t t)
    {
        const int s2 = CMP ? ((ca == t) ? 5 : -1) : __builtin_amdgcn_sbfe((int)t, ca, 4);
        const int m = max(max(up, left), diag + s2);
        // max(m - 2, 0) for m >= 0: unsigned saturating subtract.
        return (int)__builtin_elementwise_sub_sat((uint32_t)m, 2u);
    }
};

// SW, at most 4 symbols, two rows per lane (WFK_MODE_PERM4): the lane's two
// row tables hold one byte per (symbol, row), s+2 as a signed byte (5 on a
// match, -1 else), t[0] = symbols 0 and 1, t[1] = symbols 2 and 3, each as
// (row 0, row 1); the column value is a v_perm_b32 selector that moves the
// column symbol's two bytes to bytes 0 and 1.  One perm per step replaces
// the two per-row field extracts, and each row's add reads its byte
// directly.
template <bool CMP>
struct Cell<WFK_SWP, CMP> {
    static constexpr int idle = 0;
    static __device__ __forceinline__ uint32_t tables(uint32_t c0, uint32_t c1, int half)
    {
        // c0, c1: the rows' symbols (>= 4 for rows past the table: no match).
        uint32_t t = 0;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const uint32_t sym = 2u * (uint32_t)half + (uint32_t)k;
            t |= (c0 == sym ? 0x05u : 0xffu) << (16 * k);
            t |= (c1 == sym ? 0x05u : 0xffu) << (16 * k + 8);
        }
        return t;
    }
    // Selector of symbol c: bytes 2c and 2c+1 of {t[1], t[0]} to bytes 0, 1.
    static __device__ __forceinline__ uint32_t selector(uint32_t c)
    {
        return 0x0c0c0000u | ((2u * c + 1u) << 8) | (2u * c);
    }
    // A column symbol that matches no row (selector byte 0x0d reads 0xff).
    static constexpr uint32_t kNever = 0x0c0c0d0du;
};

// The reverse half of a column-split SW pair (WFK_MODE_SWE): the same byte
// tables and selectors with s + 4 (7 on a match, 1 else) and no clamp,
// E = max3(up, left, diag + s + 4).  A never-matching column reads -1, so
// lanes left of column 0 keep the zero boundary.
template <bool CMP>
struct Cell<WFK_SWE, CMP> {
    static constexpr int idle = 0;
    static __device__ __forceinline__ uint32_t tables(uint32_t c0, uint32_t c1, int half)
    {
        uint32_t t = 0;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const uint32_t sym = 2u * (uint32_t)half + (uint32_t)k;
            t |= (c0 == sym ? 0x07u : 0x01u) << (16 * k);
            t |= (c1 == sym ? 0x07u : 0x01u) << (16 * k + 8);
        }
        return t;
    }
    static __device__ __forceinline__ uint32_t selector(uint32_t c) { return Cell<WFK_SWP, CMP>::selector(c); }
    static constexpr uint32_t kNever = Cell<WFK_SWP, CMP>::kNever;
};

template <int R>
struct LaneState {
    int h[R];        // H(row r, last column) -- the next step's "left"
    int dg;          // H(row above row 0, last column) -- the next step's "diag"
    int best;        // SW running maximum
    uint32_t t[R];   // per-row table / code
};

template <int P, bool CMP, int R, bool MASKED, bool TRACK>
__device__ __forceinline__ void step(LaneState<R> &st, int inj, uint32_t ca, bool valid)
{
    const int up0 = __builtin_amdgcn_update_dpp(inj, st.h[R - 1], kDppWaveShr1, 0xf, 0xf, false);
    int up = up0;
    int diag = st.dg;
    uint32_t pp = 0;
    if (P == WFK_SWP || P == WFK_SWE) pp = __builtin_amdgcn_perm(st.t[1], st.t[0], ca);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int left = st.h[r];
        int h;
        if constexpr (P == WFK_SWP || P == WFK_SWE) {
            static_assert(R == 2, "the byte tables hold two rows per lane");
            const int s2 = (int)(int8_t)(uint8_t)(pp >> (8 * r));
            const int m = max(max(up, left), diag + s2);
            h = (P == WFK_SWE) ? m : (int)__builtin_elementwise_sub_sat((uint32_t)m, 2u);
        } else {
            h = Cell<P, CMP>::f(up, left, diag, ca, st.t[r]);
        }
        if (MASKED) {
            h = valid ? h : left;
            if (TRACK) st.best = max(st.best, valid ? h : 0);
        } else if (TRACK) {
            st.best = max(st.best, h);
        }
        diag = left;
        st.h[r] = h;
        up = h;
    }
    st.dg = MASKED ? (valid ? up0 : st.dg) : up0;
}

// Issue priority of the calling wave (s_setprio takes an immediate).
__device__ __forceinline__ void set_wave_priority(int p)
{
    switch (p) {
    case 3: __builtin_amdgcn_s_setprio(3); break;
    case 2: __builtin_amdgcn_s_setprio(2); break;
    case 1: __builtin_amdgcn_s_setprio(1); break;
    default: __builtin_amdgcn_s_setprio(0); break;
    }
}

__device__ __forceinline__ int64_t now_ticks()
{
    return (int64_t)__builtin_amdgcn_s_memrealtime();
}

}  // namespace wfk

#endif
