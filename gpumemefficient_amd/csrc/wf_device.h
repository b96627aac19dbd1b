// wf_device.h -- descriptors shared by the host planner (wf_api.cpp) and the
// gfx950 wavefront kernels (wf_kernels.hip).  Plain structs, no HIP types.
#ifndef WF_DEVICE_H
#define WF_DEVICE_H
#include <stdint.h>

// Problem ids match include/wfdp.h (WF_LCS, WF_ED, WF_SW).
enum { WFK_LCS = 0, WFK_ED = 1, WFK_SW = 2 };
// Kernel-internal problem ids (never stored in a pair): SW in
// WFK_MODE_PERM4, and the reverse half of a column-split SW pair
// (WFK_MODE_SWE).
enum { WFK_SWP = 3, WFK_SWE = 4 };

// Pair mode bits.
enum {
    WFK_MODE_CMP = 1,       // symbol codes are compared (alphabet too large for a row table)
    WFK_MODE_TEXTBOOK = 2,  // ED: D(0,x) = x instead of ED.cpp's x-1
    // Lanes still left of column 0 may run the unmasked step: their state
    // starts at a value the recurrence maps back to the left boundary (LCS,
    // ED), or (SW) the column codes left of column 0 are a code that matches
    // no row, bits WFK_MODE_NEVER_SHIFT.. of the mode.
    WFK_MODE_FREE_HEAD = 4,
    // SW with at most 4 symbols on a cell engine (asynchronous or
    // lock-step), two rows per lane: the codes are plain 0..3 and a lane's
    // row tables are bytes (see WFK_SWP in wf_cells.h).
    WFK_MODE_PERM4 = 16,
    // SW without the running maximum (only the bottom-right cell).
    WFK_MODE_NOMAX = 32,
    // The reverse half of a column-split SW pair (with WFK_MODE_PERM4): the
    // unclamped best path weight into each cell of the reversed sequences,
    // kept as E = F + 2 (x + y), so E = max3(up, left, diag + s + 4) with
    // zero boundaries.
    WFK_MODE_SWE = 64,
};
enum { WFK_MODE_NEVER_SHIFT = 8 };

// One sequence pair as the kernels see it.  x runs over a (n1 columns),
// y over b (n2 rows).  Codes are one byte per symbol in the code arena.
struct WfPairDev {
    int64_t a_off;     // byte offset of a's codes
    int64_t b_off;     // byte offset of b's codes
    int64_t bnd_off;   // element offset of this pair's strip-boundary rows
    int32_t n1, n2;
    int32_t prob;      // WFK_LCS / WFK_ED / WFK_SW
    int32_t mode;      // WFK_MODE_* bits
    int32_t prio;      // wave issue priority (s_setprio 0..3) of this pair's strips
    int32_t ngroups;   // strips of (waves * 64 * rows_per_lane) rows
    // Bit-parallel engine: entry offset in WfKernelArgs::colw where each
    // 32-row word stores its last column's vertical deltas (two words per
    // entry, ED: Pv, Mv; LCS: V, 0), at entry (strip * waves + wave) * 64 +
    // lane; -1: not stored.
    int64_t col_off = -1;
    // Tag of this pair's strip-boundary words; 0: the launch epoch.  The
    // resident server, whose one launch outlives many pairs in recycled
    // memory, gives every request a tag of its own.
    uint32_t tag = 0;
    int32_t pad = 0;
};

// A pair computed as two column halves (bit-parallel engine, LCS and ED):
// pair `job` sweeps columns 1..mid forward, pair `rev` sweeps the reversed
// columns mid+1..n1 against the reversed rows.  With F(y) the forward
// half's column mid and R(y) the reverse half's last column at row n2 - y,
// the score is min (ED) / max (LCS) over y of F(y) + R(y) (Hirschberg's
// split, J. ACM 1975; CACM 18(6) 1975 for LCS).
//
// An SW pair (prob WFK_SW, cell engine with byte tables) splits into three:
// `job` (columns 1..mid), `f0` (columns mid+1..n1 from a zero boundary) and
// `rev` (the reversed right half, WFK_MODE_SWE); r0 is the right half's
// column count.  A pair that keeps the running maximum adds a fourth, `crev`
// (the reversed right half with the clamped recurrence, no maximum), whose
// last column starts at colw entry crev_off (see sw_split_combine in
// wf_bitpar.hip).
struct WfSplitDev {
    int32_t job;       // result slot (the forward half's pair)
    int32_t rev;       // the reverse half's pair
    int32_t prob;      // WFK_ED / WFK_LCS / WFK_SW
    int32_t n2;        // rows
    int32_t f0;        // ED: F(0), row 0 of the table at column mid; LCS: 0; SW: the right half's pair
    int32_t r0;        // ED: the reverse half's row 0 at its last column (n1 - mid); LCS: 0; SW: n1 - mid
    int64_t fwd_off;   // colw entry offsets of the two halves
    int64_t rev_off;
    int32_t crev = -1; // SW with the running maximum: the clamped reverse pair; -1: none
    int32_t pad = 0;
    int64_t crev_off = 0;
};

// One unit of persistent work: strip `group` of pair `pair`.  Tasks of a
// pair appear in strip order, so a strip only ever waits on a strip that was
// claimed before it.
struct WfTaskDev {
    int32_t pair;
    int32_t group;
};

struct WfResultDev {
    int32_t last;      // bottom-right cell (the score every reference program prints)
    int32_t maxs;      // SW: maximum cell; 0 otherwise
    int32_t status;    // 0 ok, else WF_ERR_* bits set on the device
    int32_t pad;
};

// Kernel argument block (passed by value).
struct WfKernelArgs {
    const WfPairDev *pairs;
    const WfTaskDev *tasks;
    const uint8_t *codes;
    uint64_t *bnd;           // strip-boundary rows: (epoch << 32) | value per column
    WfResultDev *res;
    int32_t *counter;        // task claim counter for this launch
    int64_t timeout_ticks;   // s_memrealtime ticks (100 MHz) a dependency wait may take
    int32_t task_begin;
    int32_t task_end;
    const uint32_t *epoch;   // launch number (device memory, bumped by the reset kernel before
                             // every launch, so graph replays get fresh tags)
    int32_t *table;          // table-dump kernel only: H(y+1, x+1) of pair 0 at [y * n1 + x]
    uint32_t *colw;          // bit-parallel engine: last-column entries, two words each (WfPairDev::col_off)
};

// ----------------------------------------------------------------------------
// Resident server: one persistent launch of the lock-step engine serves every
// request for the server's lifetime (wf_serve_engine in wf_kernels.hip).
// Requests enter through a ring in pinned host memory; the kernel's waiting
// workgroups take them in (copy the codes to HBM, write the pair
// descriptor, append the pair's strips to a task ring in HBM), run strips in
// task order, and the last strip of a pair writes a completion record to a
// second pinned ring.
// ----------------------------------------------------------------------------

// Submission ring entry (pinned host memory, written by the host).
struct WfSubmitDev {
    WfPairDev pd;        // a_off/b_off/bnd_off in the server's device arena, tag, ngroups
    int64_t task_base;   // index of the pair's first strip in the task ring
    int64_t stage_off;   // byte offset of the pair's codes in the staging buffer (same layout as from a_off)
    int64_t code_bytes;  // bytes from a_off to the end of b's codes, a multiple of 16
    int32_t slot;        // pair slot: descriptor, result and done counter
    int32_t pad;
    uint64_t seq;        // submission index + 1 once the entry is complete (written last)
};

// Task ring entry (HBM): strip `group` of the pair in `slot`; seq = task
// index + 1 once valid.
struct WfTaskQDev {
    int32_t slot;
    int32_t group;
    uint64_t seq;
};

// Completion ring entry (pinned host memory, written by the device).
struct WfCompDev {
    int32_t last, maxs, status, slot;
    uint64_t seq;        // completion index + 1 once the entry is complete (written last)
};

// Device-side control block (HBM).
struct WfServeCtl {
    uint64_t head;       // next task index to claim
    uint64_t sub_head;   // next submission to take in
    uint64_t comp_tail;  // next completion record
    int32_t poller;      // 1 while a workgroup holds the lease to poll the host ring
    int32_t stop;        // copy of the host's stop flag
};

// Host-side control block (pinned host memory).
struct WfHostCtl {
    uint64_t sub_tail;   // submissions published
    int32_t stop;        // set by the host to end the launch once idle
    int32_t pad;
};

struct WfServeArgs {
    WfKernelArgs ka;            // pairs, codes, bnd, res: the server's tables; timeout
    WfTaskQDev *taskq;
    int64_t taskq_cap;
    const WfSubmitDev *subq;    // pinned host
    int64_t subq_cap;
    const uint8_t *staging;     // pinned host
    WfCompDev *compq;           // pinned host
    int64_t compq_cap;
    WfServeCtl *ctl;
    const WfHostCtl *hctl;      // pinned host
    int32_t *pair_done;         // strips finished, per slot
};

// In-place table stencils swept with the same strip schedule (wf_stencil.hip):
// the reference's summed-area table and one Gauss-Seidel SOR sweep.
enum { WFS_SAT = 0, WFS_SOR = 1 };

// Kernel argument block of a stencil sweep.  The table is row-major with
// pitch `ld` int32 (ld >= n1 + 3, ld % 4 == 0, 16-byte aligned base); rows
// 0 and n2 + 1 and columns 0, n1 + 1, n1 + 2 are fixed boundary values and
// the interior 1 <= y <= n2, 1 <= x <= n1 is updated in place.
struct WfStencilArgs {
    int32_t *t;
    int64_t ld;
    int32_t n1, n2;
    int32_t nstrips;         // strips of 64 * waves rows
    int32_t op;              // WFS_SAT / WFS_SOR
    uint64_t *bnd;           // (nstrips - 1) rows of n1 + 1 words: (epoch << 32) | value
    int32_t *counter;        // strip claim counter (zeroed by the reset kernel)
    WfResultDev *res;        // res[0].status: timeout flag
    const uint32_t *epoch;
    int64_t timeout_ticks;
};

#endif
