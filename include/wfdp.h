/*
 * wfdp.h -- C ABI of libwfdp.so, the gfx950 wavefront DP engine for LCS,
 * edit distance and Smith-Waterman.
 *
 * The reference exposes one C++ function per problem and per driver, called
 * from its driver main()s (all paths relative to the reference tree):
 *
 *   int LCS(int n1, int n2, int *arr1, int *arr2, int paddX, int paddY, int *table);
 *       LCS/GPU_Hyperlane_Share.h:5, LCS/GPU_Hyperlane_Cache.h:5,
 *       LCS-SHFL/GPU_Hyperlane_Share.h:5, LCS-SHFL/GPU_Hyperlane_Cache.h:5
 *   int LCS(int n1, int n2, int *arr1, int *arr2);
 *       LCS/GPU.h:5, LCS-SHFL/GPU.h:5 (the naive programs)
 *   int SW(int n1, int n2, int *arr1, int *arr2, int paddX, int paddY, int *table, int tX, int tY);
 *       SmithWaterman/GPU_Hyperlane_Share.h:5
 *   int SW(int n1, int n2, int *arr1, int *arr2, int paddX, int paddY, int *table);
 *       SmithWaterman/GPU_Hyperlane_Cache.h:5
 *   int ED(int n1, int n2, int *arr1, int *arr2, int paddX, int paddY, int *table, int tX, int tY);
 *       EditDistance/GPU_Hyperlane_Share.h:5
 *   int ED(int n1, int n2, int *arr1, int *arr2, int paddX, int paddY, int *table);
 *       EditDistance/GPU_Hyperlane_Cache.h:5
 *
 * Each of those allocates the full (n1+3) x (n2+1) table on the device per
 * call and returns the bottom-right cell.  wf_align() below is the single
 * entry point those six signatures map onto (wfdp_compat.h gives them back
 * with their exact C++ signatures); it keeps only strip boundaries in HBM.
 * Everything else here (contexts, plans, batches, the server) has no
 * counterpart in the reference and exists for the batched and served
 * workloads.
 *
 * Conventions: x runs over `a` (the reference's arr1, n1 symbols), y over `b`
 * (arr2, n2 symbols); the returned score is the table cell (n2, n1), exactly
 * what LCS.cpp / SW.cpp / ED.cpp print for n1 == n2.  Symbols are arbitrary
 * int32 values compared for equality; at most 256 distinct symbols per pair.
 * All lengths are 64-bit and may be 0.  Functions return WF_OK or a WF_ERR_*
 * code instead of exiting (the reference exits: LCS/GPU_Hyperlane_Share.cu
 * :689-694).  A context is used by one host thread at a time.
 */
#ifndef WFDP_H
#define WFDP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WFDP_ABI_VERSION 1

typedef enum { WF_LCS = 0, WF_ED = 1, WF_SW = 2 } wf_problem;

enum {
    WF_OK = 0,
    WF_ERR_INVALID = 1,   /* bad argument */
    WF_ERR_DEVICE = 2,    /* HIP runtime error */
    WF_ERR_NOMEM = 3,     /* device allocation failed or over the HBM budget */
    WF_ERR_ALPHABET = 4,  /* more than 256 distinct symbols in one pair */
    WF_ERR_TIMEOUT = 5,   /* a device-side dependency wait exceeded its limit */
    WF_ERR_NOKERNEL = 6,  /* no compiled kernel matches the requested tile */
    WF_ERR_AGAIN = 7,     /* server: queue full, retry later */
    WF_ERR_SCRATCH = 8    /* wf_align_noalloc: the pair needs more than its fixed scratch area */
};

/* Tile selection.  strip_rows = 64 * waves * rows_per_lane rows are swept by
 * one workgroup; tile_w = steps between boundary exchanges (16, 32 or 64).
 * Zero fields mean "choose" (autotuned table, then built-in defaults). */
typedef struct {
    int32_t tile_w;
    int32_t strip_rows;
    int32_t rows_per_lane;
    int32_t flags;
} wf_tile;

#define WF_TILE_ED_TEXTBOOK 0x1 /* ED row 0 = x (reference GPU_setInit), not ED.cpp's x-1 */
#define WF_TILE_CACHE 0x2       /* hyperlane-Cache policy: one wave per workgroup, all
                                   strip boundaries through L2 (GPU_Hyperlane_Cache.cu) */
/* Engine choice; without a flag the planner takes the bit-parallel engine
 * for LCS and ED pairs of at most 32 symbols, else the asynchronous engine
 * when all strips of the launch fit on the chip at once (at most two 512-row
 * strips per CU) and the lock-step engine otherwise. */
#define WF_TILE_LOCKSTEP 0x4    /* barrier-synchronised engine */
#define WF_TILE_ASYNC 0x8       /* asynchronous engine: LDS progress counters + I/O wave */
#define WF_TILE_BITPAR 0x20     /* bit-parallel engine: LCS and ED with at most 32 symbols,
                                   32 rows per lane as one bit-vector word; the default for
                                   such pairs unless a cell engine flag is given */
#define WF_TILE_BITPAR_STAGED 0x40 /* bit-parallel engine with a mask-staging wave per
                                      workgroup (one workgroup per CU); without either
                                      bit-parallel flag the planner takes it when every
                                      strip of the launch gets a CU of its own */
#define WF_TILE_NO_COPYBACK 0x10 /* wf_stencil: leave the host table unchanged (the
                                    reference never copies its result back) */
#define WF_TILE_HINT 0x80       /* tile_w, strip_rows and rows_per_lane are hints from a
                                   caller that sizes tiles for another machine (the
                                   reference's tX, tY): the planner picks its own variant
                                   (autotune table, then defaults); every variant gives
                                   the same results */
#define WF_TILE_NO_MAX 0x100    /* SW: compute only the bottom-right cell, which is all the
                                   reference's SW() returns (SmithWaterman/SW.cpp:80,
                                   GPU_Hyperlane_Share.cu:781), not the running maximum
                                   (about 8% of the cell engines' instructions); such pairs
                                   report max_score = -1.  Pairs with more symbols than the
                                   kernels' row tables hold keep the maximum.  Also a job
                                   flag (wf_job.flags), which the server honours. */

typedef struct {
    int32_t last;      /* bottom-right cell: the score the reference prints */
    int32_t max_score; /* SW: maximum cell (local alignment score); 0 for LCS/ED */
    int32_t status;    /* WF_OK or WF_ERR_* for this pair */
    int32_t reserved;
    double kernel_ms;  /* device time of the launch(es) that computed this result */
    uint64_t cells;    /* n1 * n2 */
} wf_result;

typedef struct {
    int32_t problem; /* wf_problem */
    int32_t flags;   /* WF_TILE_ED_TEXTBOOK */
    const int32_t *a;
    int64_t n1;
    const int32_t *b;
    int64_t n2;
} wf_job;

typedef struct {
    int32_t compute_units;
    int32_t max_clock_mhz;
    int32_t simds_per_cu;
    int32_t wave_size;
    int64_t total_mem;
    int64_t free_mem;
    char arch[32];
} wf_device_info;

typedef struct wf_ctx wf_ctx;
typedef struct wf_plan wf_plan;

/* Context: owns a HIP stream, device buffer pools and events on one device. */
int wf_ctx_create(int device, wf_ctx **out);
void wf_ctx_destroy(wf_ctx *ctx);
int wf_device_query(wf_ctx *ctx, wf_device_info *out);
const char *wf_strerror(int status);
int wf_abi_version(void);
/* Seconds a device-side dependency wait may take before the pair fails with
 * WF_ERR_TIMEOUT (default 10). */
int wf_ctx_set_timeout(wf_ctx *ctx, double seconds);

/* One pair, synchronous: encode + upload + run + download (the drop-in path
 * behind the reference's LCS()/SW()/ED()). */
int wf_align(wf_ctx *ctx, int problem, const int32_t *a, int64_t n1, const int32_t *b,
             int64_t n2, const wf_tile *tile, wf_result *out);

/* wf_align() without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code that
 * must not call hipMalloc, such as tight serving loops, harnesses that forbid
 * it, work recorded into a graph).  Every device buffer the call needs
 * (symbol codes, descriptors, strip boundaries, results, counters) comes from
 * one static device array of WF_NOALLOC_SCRATCH_BYTES per GPU, held until the
 * process ends; the caller passes the same host inputs and output as to
 * wf_align().  Pairs whose strip boundaries or codes do not fit return
 * WF_ERR_SCRATCH (roughly: (n2 / strip_rows - 1) * n1 * 8 bytes of boundary
 * rows within 56 MiB, and n1 + n2 bytes of codes within 8 MiB; 16384 x 16384
 * SW fits; an SW pair whose column split would not fit runs unsplit).  All calls on one GPU share that GPU's array, and calls that
 * overlap on one GPU (several threads or contexts) race: the caller
 * serialises them. */
#define WF_NOALLOC_SCRATCH_BYTES (64LL << 20)
int wf_align_noalloc(wf_ctx *ctx, int problem, const int32_t *a, int64_t n1, const int32_t *b,
                     int64_t n2, const wf_tile *tile, wf_result *out);

/* Full DP table of one pair (the reference drivers' DEBUG dump,
 * LCS/GPU_Hyperlane_Share.cpp:109-131): fills table[(y-1)*n1 + (x-1)] = H(y, x)
 * for 1 <= y <= n2, 1 <= x <= n1 in the reference's values (LCS and SW cell
 * scores, ED distances D with ED.cpp's or, with WF_TILE_ED_TEXTBOOK in tile,
 * the textbook row 0).  Runs the lock-step engine with a per-cell store; for
 * cell-level parity checks on small pairs (n1 * n2 <= WF_TABLE_MAX_CELLS).
 * `table` is a caller-owned host buffer of n1 * n2 int32. */
#define WF_TABLE_MAX_CELLS (1LL << 28)
int wf_align_table(wf_ctx *ctx, int problem, const int32_t *a, int64_t n1, const int32_t *b,
                   int64_t n2, const wf_tile *tile, int32_t *table, wf_result *out);

/* The reference's naive GPU baselines, for "vs naive" comparisons: the whole
 * (n2+1) x (n1+1) int32 table lives in HBM and every cell is written there.
 *   WF_NAIVE_DIAGONAL: one launch per cell anti-diagonal, one thread per cell
 *     (LCS/GPU.cu:66-102, SmithWaterman/GPU.cu, EditDistance/GPU.cu).
 *   WF_NAIVE_TILE: one launch per anti-diagonal of 64 x 128 tiles, each tile
 *     swept in LDS (LCS/GPU_TILE_SHARE.cu:140-182; the reference launches
 *     each tile separately over 32 streams).
 * Symbols are compared as int32 (no alphabet limit).  tile: only its
 * WF_TILE_ED_TEXTBOOK flag is read (the reference's naive ED uses the
 * textbook row 0, EditDistance/GPU.cu:12-18).  table, if not NULL, receives
 * table[(y-1)*n1 + (x-1)] = H(y, x) as wf_align_table does.  out->kernel_ms
 * spans all launches.  WF_ERR_NOMEM when the table does not fit in free HBM. */
#define WF_NAIVE_DIAGONAL 0
#define WF_NAIVE_TILE 1
int wf_align_naive(wf_ctx *ctx, int problem, const int32_t *a, int64_t n1, const int32_t *b,
                   int64_t n2, const wf_tile *tile, int method, int32_t *table, wf_result *out);

/* In-place table stencils swept with the same wavefront schedule: the
 * reference's other users of the hyperlane engine, whose host function is
 *   void SOR(int n1, int n2, int *table, int tX, int tY);
 *       SAT/GPU_Hyperlane_Share.h:5, 2D-SOR/nested_loop/GPU_Hyperlane_Share.h:5
 *   void SOR(int n1, int n2, int *table);
 *       SAT/GPU_Hyperlane_Cache.h:5, 2D-SOR/nested_loop/GPU_Hyperlane_Cache.h:5
 *   WF_SAT  summed-area table: T(y,x) = T(y,x-1) + T(y-1,x) + T(y,x) - T(y-1,x-1)
 *           (SAT/GPU_Hyperlane_Share.cu:169), int32 arithmetic wrapping mod 2^32.
 *   WF_SOR  one Gauss-Seidel sweep: T(y,x) = (T(y,x-1) + T(y-1,x) + T(y,x) +
 *           T(y,x+1) + T(y+1,x)) / 5 with left and up already updated
 *           (2D-SOR/nested_loop/GPU_Hyperlane_Share.cu:169); the sum wraps mod
 *           2^32, the division truncates toward zero.
 * The table is the reference's: n2 + 2 rows of pitch ld >= n1 + 3 int32
 * (rowsize = n1 + 3, colsize = n2 + 2, SAT/GPU_Hyperlane_Share.cu:261-263).
 * Row 0, row n2 + 1 and columns 0, n1 + 1, n1 + 2 are boundary values; the
 * interior 1 <= y <= n2, 1 <= x <= n1 is updated in place, as the row-major
 * sequential sweep would update it.  The reference computes on a device copy
 * that it never copies back (SAT/GPU_Hyperlane_Share.cu:317); these functions
 * return the result (wf_stencil: unless WF_TILE_NO_COPYBACK).  tile:
 * strip_rows 64, 128 or 256 (0 = choose); rows_per_lane 0 or 1.  out->last =
 * the updated cell (n2, n1). */
typedef enum { WF_SAT = 0, WF_SOR = 1 } wf_stencil_op;
#define WF_STENCIL_MAX_PITCH 14000000 /* largest ld (and n1 + 3) accepted */
/* Host table, updated in place (uploaded, swept, downloaded).  WF_ERR_NOMEM if
 * the table does not fit in free HBM. */
int wf_stencil(wf_ctx *ctx, int op, int64_t n1, int64_t n2, int32_t *table, int64_t ld,
               const wf_tile *tile, wf_result *out);
/* Device-resident table (caller-owned device memory): ld % 4 == 0 and a
 * 16-byte aligned base, else WF_ERR_INVALID.  Synchronous on the context's
 * stream; out->kernel_ms is the sweep's device time. */
int wf_stencil_device(wf_ctx *ctx, int op, int64_t n1, int64_t n2, int32_t *dev_table, int64_t ld,
                      const wf_tile *tile, wf_result *out);

int wf_align_batch(wf_ctx *ctx, const wf_job *jobs, int64_t njobs, const wf_tile *tile,
                   wf_result *out);

/* Plans: inputs encoded and resident in HBM, run any number of times. */
enum { WF_RUN_PERSISTENT = 0, WF_RUN_PER_PAIR = 1 };
int wf_plan_create(wf_ctx *ctx, const wf_job *jobs, int64_t njobs, const wf_tile *tile,
                   wf_plan **out);
int wf_plan_run(wf_plan *plan, int mode);                 /* enqueue + wait */
/* Capture the plan's launches for `mode` into a hipGraph; later
 * wf_plan_run(plan, mode) calls replay the graph. */
int wf_plan_capture(wf_plan *plan, int mode);
int wf_plan_results(wf_plan *plan, wf_result *out, int64_t n);
double wf_plan_kernel_ms(const wf_plan *plan);            /* last run, HIP events */
/* engine: 0 lock-step, 1 asynchronous, 2 bit-parallel, 3 bit-parallel staged.
 * ntasks: strips of the launch.  A bit-parallel pair of at least 1024
 * columns in a launch with room for twice its strips runs as two column
 * halves (the left one forward, the right one reversed) whose scores a
 * small kernel combines; ntasks then counts the strips of both halves.
 * WFDP_SPLIT=0 in the environment turns the split off. */
int wf_plan_describe(const wf_plan *plan, int32_t *engine, int32_t *rows_per_lane,
                     int32_t *waves, int32_t *chunk, int64_t *ntasks, int32_t *grid);

/* Autotuner: time every compiled variant on `jobs` (resident in HBM, `reps`
 * timed runs each), return the fastest as a tile and remember it for
 * workloads of the same class on this device; plans created without an
 * explicit tile use the remembered choice.  The table persists through
 * wf_tune_save / wf_tune_load (text, one entry per line). */
int wf_autotune(wf_ctx *ctx, const wf_job *jobs, int64_t njobs, int reps, wf_tile *best,
                double *best_ms);
int wf_tune_load(const char *path);
int wf_tune_save(const char *path);
int wf_tune_clear(void);
/* The tile the autotune table holds for `jobs` on this context's device
 * (what a plan without an explicit tile would use), or WF_ERR_NOKERNEL if
 * it holds none.  The first context of a process loads the table shipped
 * with the library (tune/wfdp_tune.txt beside lib/), or the file named by
 * WFDP_TUNE in the environment; WFDP_TUNE=0 loads none. */
int wf_tune_lookup_tile(wf_ctx *ctx, const wf_job *jobs, int64_t njobs, wf_tile *out);

/* Compiled kernel variants (engine 0 = lock-step, 1 = asynchronous, 2 =
 * bit-parallel, whose rows_per_lane is the 32 bits of a lane's word). */
int wf_variant_count(void);
int wf_variant_info(int index, int32_t *engine, int32_t *rows_per_lane, int32_t *waves,
                    int32_t *chunk);
void wf_plan_destroy(wf_plan *plan);

/* ------------------------------------------------------------------------
 * Server: continuous batching of requests that arrive over time.  A
 * scheduler thread admits queued requests in arrival order while their
 * device footprint fits the HBM budget and launches every admitted request
 * not yet on the device as one persistent launch on the next free stream,
 * so requests that arrive while others run overlap with them on the device.
 * In resident mode one persistent launch lives as long as the server and
 * takes admitted requests in as they come (device-resident task queue).
 * No counterpart in the reference.
 * ---------------------------------------------------------------------- */
typedef struct wf_server wf_server;

typedef struct {
    int32_t streams;        /* concurrent launches (HIP streams); 0 = 4 */
    int32_t max_batch;      /* requests per launch; 0 = 512 */
    int64_t hbm_budget;     /* device bytes the server may hold; 0 = half of free memory */
    int32_t queue_capacity; /* queued requests before submit returns WF_ERR_AGAIN; 0 = 2^20 */
    int32_t per_request;    /* 1: one launch per request on one stream (baseline) */
    int32_t resident;       /* 1: one persistent launch for the server's lifetime: requests
                               are handed to it through pinned host memory, their strips
                               queued in HBM and run as workgroups free up (no launch per
                               batch); `streams` and `max_batch` do not apply */
} wf_server_opts;

typedef struct {
    uint64_t ticket;
    int32_t last;
    int32_t max_score;
    int32_t status;
    int32_t reserved;
    double t_submit; /* seconds on the server clock (replay: scheduled arrival) */
    double t_start;  /* launch */
    double t_done;   /* completion observed by the host */
} wf_completion;

typedef struct {
    uint64_t launches;
    uint64_t completed;
    uint64_t cells;
    int64_t hbm_budget;
    int64_t hbm_used;
    int64_t pending;
} wf_server_stats_t;

int wf_server_create(int device, const wf_server_opts *opts, wf_server **out);
/* Copies the job's sequences; returns at once with a ticket.  A request that
 * no launch could run is refused here (WF_ERR_INVALID, WF_ERR_ALPHABET for
 * more than 256 distinct symbols) and never joins a batch. */
int wf_server_submit(wf_server *s, const wf_job *job, uint64_t *ticket);
/* Completed requests since the last poll (up to max); *n is set and the
 * completions are returned even when the status is an error.  Per-request
 * failures are in each completion's status; the function's status is
 * WF_ERR_DEVICE only after a HIP runtime error, which stops the server. */
int wf_server_poll(wf_server *s, wf_completion *out, int max, int *n);
int wf_server_wait_idle(wf_server *s, double timeout_s);
int wf_server_stats(wf_server *s, wf_server_stats_t *st);
/* Submit jobs[i] at arrival_s[i] seconds after the call (real time, from
 * the calling thread) and wait for all of them; out[i] gets job i's
 * completion with times relative to the start. */
int wf_server_replay(wf_server *s, const wf_job *jobs, const double *arrival_s, int64_t n,
                     wf_completion *out);
void wf_server_destroy(wf_server *s);

#ifdef __cplusplus
}
#endif
#endif /* WFDP_H */
