/*
 * wf_oracle.c -- CPU restatement of the reference's three wavefront DPs.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product path (the HIP library,
 * the drivers, the Python host package) links, loads or calls this file.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
 * it, and only as the checker or the timed CPU baseline.
 *
 * Semantics (x runs over a = arr1, n1 symbols; y runs over b = arr2, n2):
 *
 *   LCS  LCS/LCS.cpp:45-82
 *        H(0,.) = H(.,0) = 0                                   (:48-52)
 *        H(y,x) = a[x-1]==b[y-1] ? H(y-1,x-1)+1
 *                                : max(H(y,x-1), H(y-1,x))     (:58-63)
 *        result H(n2,n1)                                       (:79)
 *   SW   SmithWaterman/SW.cpp:45-83, s(a,b) = a==b ? 3 : -3    (:45-47)
 *        H(0,.) = H(.,0) = 0                                   (:52-56)
 *        H(y,x) = max(H(y,x-1)-2, H(y-1,x)-2,
 *                     H(y-1,x-1)+s(a[x-1],b[y-1]), 0)          (:63-65)
 *        result "last" = H(n2,n1) (bottom-right, :80); the maximum cell is
 *        reported separately (the reference does not print it).
 *   ED   EditDistance/ED.cpp:45-81
 *        D(0,0) = 0, D(0,x) = x-1 for x >= 1, D(y,0) = y       (:48-54)
 *        D(y,x) = a[x-1]==b[y-1] ? D(y-1,x-1)
 *                                : 1+min(D(y-1,x), D(y,x-1), D(y-1,x-1))
 *                                                              (:61-63)
 *        result D(n2,n1)                                       (:78)
 *        The "textbook" flag selects D(0,x) = x instead, which is the
 *        boundary the reference GPU code writes
 *        (EditDistance/GPU_Hyperlane_Share.cu:16-22).
 *
 * The reference CPU programs index the table as table[y*(n2+1)+x] with the
 * outer loop over arr2 and the inner over arr1, and are only defined for
 * n1 == n2 (LCS/LCS.cpp:54-59).  For n1 == n2 the functions here compute the
 * identical table; for n1 != n2 they follow the GPU reference geometry
 * (rowsize = 3+n1, x over arr1: LCS/GPU_Hyperlane_Share.cu:705,417-420).
 *
 * All indices are 64-bit and only two table rows are kept (rolling rows), so
 * 2^17 x 2^17 runs in O(n) memory.  The *_table functions keep the whole
 * (n2+1) x (n1+1) table, in the reference's loop order, for cell-level
 * comparison and for the full-table timing mode.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "wf_oracle.h"

static inline int32_t imax(int32_t a, int32_t b) { return a > b ? a : b; }
static inline int32_t imin(int32_t a, int32_t b) { return a < b ? a : b; }

/* LCS/LCS.cpp:45-82 with rolling rows. */
int32_t oracle_lcs(const int32_t *a, int64_t n1, const int32_t *b, int64_t n2)
{
    if (n1 <= 0 || n2 <= 0) return 0;
    int32_t *prev = (int32_t *)calloc((size_t)n1 + 1, sizeof(int32_t));
    int32_t *cur = (int32_t *)calloc((size_t)n1 + 1, sizeof(int32_t));
    if (!prev || !cur) { free(prev); free(cur); return -1; }
    for (int64_t y = 1; y <= n2; ++y) {
        const int32_t by = b[y - 1];
        cur[0] = 0;
        for (int64_t x = 1; x <= n1; ++x) {
            if (a[x - 1] == by)
                cur[x] = prev[x - 1] + 1;
            else
                cur[x] = imax(cur[x - 1], prev[x]);
        }
        int32_t *t = prev; prev = cur; cur = t;
    }
    int32_t r = prev[n1];
    free(prev); free(cur);
    return r;
}

/* SmithWaterman/SW.cpp:45-83 with rolling rows; *max_out gets max H. */
int32_t oracle_sw(const int32_t *a, int64_t n1, const int32_t *b, int64_t n2,
                  int32_t *max_out)
{
    if (max_out) *max_out = 0;
    if (n1 <= 0 || n2 <= 0) return 0;
    int32_t *prev = (int32_t *)calloc((size_t)n1 + 1, sizeof(int32_t));
    int32_t *cur = (int32_t *)calloc((size_t)n1 + 1, sizeof(int32_t));
    if (!prev || !cur) { free(prev); free(cur); return -1; }
    int32_t best = 0;
    for (int64_t y = 1; y <= n2; ++y) {
        const int32_t by = b[y - 1];
        cur[0] = 0;
        for (int64_t x = 1; x <= n1; ++x) {
            int32_t s = (a[x - 1] == by) ? 3 : -3;
            int32_t h = imax(cur[x - 1] - 2,
                        imax(prev[x] - 2, imax(prev[x - 1] + s, 0)));
            cur[x] = h;
            best = imax(best, h);
        }
        int32_t *t = prev; prev = cur; cur = t;
    }
    int32_t r = prev[n1];
    free(prev); free(cur);
    if (max_out) *max_out = best;
    return r;
}

/* Row-0 boundary of EditDistance/ED.cpp:48-54 (textbook=0) or of the GPU
 * reference's GPU_setInit (textbook=1). */
static inline int32_t ed_top(int64_t x, int textbook)
{
    if (x == 0) return 0;
    return textbook ? (int32_t)x : (int32_t)(x - 1);
}

/* EditDistance/ED.cpp:45-81 with rolling rows. */
int32_t oracle_ed(const int32_t *a, int64_t n1, const int32_t *b, int64_t n2,
                  int textbook)
{
    if (n1 < 0 || n2 < 0) return -1;
    if (n2 == 0) return ed_top(n1, textbook);
    if (n1 == 0) return (int32_t)n2;
    int32_t *prev = (int32_t *)malloc(((size_t)n1 + 1) * sizeof(int32_t));
    int32_t *cur = (int32_t *)malloc(((size_t)n1 + 1) * sizeof(int32_t));
    if (!prev || !cur) { free(prev); free(cur); return -1; }
    for (int64_t x = 0; x <= n1; ++x) prev[x] = ed_top(x, textbook);
    for (int64_t y = 1; y <= n2; ++y) {
        const int32_t by = b[y - 1];
        cur[0] = (int32_t)y;
        for (int64_t x = 1; x <= n1; ++x) {
            int32_t d = imin(prev[x] + 1, imin(cur[x - 1] + 1, prev[x - 1] + 1));
            if (a[x - 1] == by) d = prev[x - 1];
            cur[x] = d;
        }
        int32_t *t = prev; prev = cur; cur = t;
    }
    int32_t r = prev[n1];
    free(prev); free(cur);
    return r;
}

/* Full-table forms: table has (n2+1)*(n1+1) cells, row-major with stride
 * n1+1, filled in the reference's loop order.  Returns the bottom-right
 * cell.  The caller owns `table`. */
int32_t oracle_table(int problem, const int32_t *a, int64_t n1,
                     const int32_t *b, int64_t n2, int textbook,
                     int32_t *table)
{
    const int64_t st = n1 + 1;
    for (int64_t x = 0; x <= n1; ++x)
        table[x] = (problem == ORACLE_ED) ? ed_top(x, textbook) : 0;
    for (int64_t y = 0; y <= n2; ++y)
        table[y * st] = (problem == ORACLE_ED) ? (int32_t)y : 0;
    for (int64_t y = 1; y <= n2; ++y) {
        for (int64_t x = 1; x <= n1; ++x) {
            const int64_t i = y * st + x;
            const int32_t up = table[i - st], left = table[i - 1],
                          diag = table[i - st - 1];
            const int eq = a[x - 1] == b[y - 1];
            int32_t v;
            if (problem == ORACLE_LCS) {
                v = eq ? diag + 1 : imax(left, up);
            } else if (problem == ORACLE_SW) {
                v = imax(left - 2, imax(up - 2, imax(diag + (eq ? 3 : -3), 0)));
            } else {
                v = imin(up + 1, imin(left + 1, diag + 1));
                if (eq) v = diag;
            }
            table[i] = v;
        }
    }
    return table[n2 * st + n1];
}

/* Last row H(n2, 0..n1) written to out[0..n1]; used to check boundary rows
 * exchanged between strips.  Returns 0 on success. */
int oracle_last_row(int problem, const int32_t *a, int64_t n1,
                    const int32_t *b, int64_t n2, int textbook, int32_t *out)
{
    int32_t *prev = (int32_t *)malloc(((size_t)n1 + 1) * sizeof(int32_t));
    int32_t *cur = (int32_t *)malloc(((size_t)n1 + 1) * sizeof(int32_t));
    if (!prev || !cur) { free(prev); free(cur); return -1; }
    for (int64_t x = 0; x <= n1; ++x)
        prev[x] = (problem == ORACLE_ED) ? ed_top(x, textbook) : 0;
    for (int64_t y = 1; y <= n2; ++y) {
        const int32_t by = b[y - 1];
        cur[0] = (problem == ORACLE_ED) ? (int32_t)y : 0;
        for (int64_t x = 1; x <= n1; ++x) {
            const int32_t up = prev[x], left = cur[x - 1], diag = prev[x - 1];
            const int eq = a[x - 1] == by;
            int32_t v;
            if (problem == ORACLE_LCS) {
                v = eq ? diag + 1 : imax(left, up);
            } else if (problem == ORACLE_SW) {
                v = imax(left - 2, imax(up - 2, imax(diag + (eq ? 3 : -3), 0)));
            } else {
                v = imin(up + 1, imin(left + 1, diag + 1));
                if (eq) v = diag;
            }
            cur[x] = v;
        }
        int32_t *t = prev; prev = cur; cur = t;
    }
    memcpy(out, prev, ((size_t)n1 + 1) * sizeof(int32_t));
    free(prev); free(cur);
    return 0;
}

/* Batch helper: results[i] = last score of pair i (ED uses ED.cpp's
 * boundary), maxs[i] = SW maximum (0 for LCS/ED). */
void oracle_batch(const int32_t *problems, const int32_t *const *as,
                  const int64_t *n1s, const int32_t *const *bs,
                  const int64_t *n2s, int64_t npairs, int32_t *results,
                  int32_t *maxs)
{
    for (int64_t i = 0; i < npairs; ++i) {
        int32_t m = 0;
        switch (problems[i]) {
        case ORACLE_LCS: results[i] = oracle_lcs(as[i], n1s[i], bs[i], n2s[i]); break;
        case ORACLE_SW: results[i] = oracle_sw(as[i], n1s[i], bs[i], n2s[i], &m); break;
        default: results[i] = oracle_ed(as[i], n1s[i], bs[i], n2s[i], 0); break;
        }
        if (maxs) maxs[i] = m;
    }
}

/*
 * In-place table stencils: the reference's other users of the hyperlane
 * schedule.  The table has n2 + 2 rows of pitch ld >= n1 + 3 (rowsize = n1+3,
 * colsize = n2+2: SAT/GPU_Hyperlane_Share.cu:261-263); row 0, row n2+1 and
 * columns 0, n1+1, n1+2 are read-only boundary values and the interior
 * 1 <= y <= n2, 1 <= x <= n1 is updated in row-major order.
 *
 *   SAT  T(y,x) = T(y,x-1) + T(y-1,x) + T(y,x) - T(y-1,x-1)
 *        (SAT/GPU_Hyperlane_Share.cu:169, :196, :233)
 *   SOR  T(y,x) = (T(y,x-1) + T(y-1,x) + T(y,x) + T(y,x+1) + T(y+1,x)) / 5
 *        (2D-SOR/nested_loop/GPU_Hyperlane_Share.cu:169, :197, :233)
 *
 * The reference's hyperlane tiles visit every cell after its left and upper
 * neighbours and before its right and lower ones (the strip below lags by
 * YoverX + 1 tiles, SAT/GPU_Hyperlane_Share.cu:106-112), which is the
 * row-major order's dependence pattern, so this sequential sweep is its
 * result.  The reference computes in C int; here sums wrap mod 2^32 (the
 * reference's overflow is undefined behaviour, two's complement wrap on the
 * GPU) and SOR's division truncates toward zero as C's does.
 */
void oracle_stencil(int op, int64_t n1, int64_t n2, int32_t *t, int64_t ld)
{
    for (int64_t y = 1; y <= n2; ++y) {
        int32_t *row = t + y * ld;
        const int32_t *up = row - ld;
        const int32_t *dn = row + ld;
        for (int64_t x = 1; x <= n1; ++x) {
            if (op == ORACLE_SAT) {
                row[x] = (int32_t)((uint32_t)row[x - 1] + (uint32_t)up[x] + (uint32_t)row[x] - (uint32_t)up[x - 1]);
            } else {
                const int32_t s = (int32_t)((uint32_t)row[x - 1] + (uint32_t)up[x] + (uint32_t)row[x] +
                                            (uint32_t)row[x + 1] + (uint32_t)dn[x]);
                row[x] = s / 5;
            }
        }
    }
}

/*
 * All-cores variant of oracle_lcs / oracle_sw / oracle_ed: the table is cut
 * into tiles of `tile` x `tile` cells and the tiles of one anti-diagonal of
 * tiles are computed in parallel by `threads` POSIX threads, with a barrier
 * between anti-diagonals (the CPU form of the reference's wavefront).  Each
 * tile reads its top row from top[] (the bottom row of the tile above), its
 * left column from left[] (the right column of the tile to its left) and
 * its top-left corner from corner[], and writes its own bottom row, right
 * column and bottom-right corner back.  Same recurrences and boundaries as
 * the rolling-row functions above, so the results are identical; *max_out
 * gets the SW maximum.  Returns H(n2, n1).
 */
#include <pthread.h>

typedef struct {
    int problem, textbook, threads;
    const int32_t *a, *b;
    int64_t n1, n2, tile, nI, nJ;
    int32_t *top, *left, *corner; /* corner[(I) * (nJ + 1) + J] = H(I*tile, J*tile) */
    int32_t *best;                /* per thread */
    pthread_barrier_t bar;
    int64_t next;                 /* tile index within the current anti-diagonal */
    pthread_mutex_t mu;
} Blocked;

typedef struct {
    Blocked *B;
    int id;
} BlockedArg;

static void blocked_tile(Blocked *B, int64_t I, int64_t J, int32_t *row, int32_t *best)
{
    const int64_t T = B->tile;
    const int64_t y0 = I * T + 1, y1 = (I + 1) * T < B->n2 ? (I + 1) * T : B->n2;
    const int64_t x0 = J * T + 1, x1 = (J + 1) * T < B->n1 ? (J + 1) * T : B->n1;
    const int64_t w = x1 - x0 + 1;
    /* row[0] = H(y, x0 - 1), row[1..w] = H(y, x0..x1) of the row above */
    memcpy(row + 1, B->top + x0, (size_t)w * sizeof(int32_t));
    int32_t diag0 = B->corner[I * (B->nJ + 1) + J];
    int32_t m = *best;
    for (int64_t y = y0; y <= y1; ++y) {
        const int32_t by = B->b[y - 1];
        const int32_t lf = B->left[y];
        int32_t diag = diag0, left = lf;
        for (int64_t k = 1; k <= w; ++k) {
            const int32_t up = row[k];
            const int eq = B->a[x0 + k - 2] == by;
            int32_t v;
            if (B->problem == ORACLE_LCS) {
                v = eq ? diag + 1 : imax(left, up);
            } else if (B->problem == ORACLE_SW) {
                v = imax(left - 2, imax(up - 2, imax(diag + (eq ? 3 : -3), 0)));
                m = imax(m, v);
            } else {
                v = imin(up + 1, imin(left + 1, diag + 1));
                if (eq) v = diag;
            }
            diag = up;
            row[k] = v;
            left = v;
        }
        diag0 = lf;
        B->left[y] = left; /* H(y, x1): the left column of tile (I, J + 1) */
    }
    memcpy(B->top + x0, row + 1, (size_t)w * sizeof(int32_t));
    B->corner[(I + 1) * (B->nJ + 1) + J + 1] = row[w];
    *best = m;
}

static void *blocked_worker(void *p)
{
    BlockedArg *arg = (BlockedArg *)p;
    Blocked *B = arg->B;
    int32_t *row = (int32_t *)malloc(((size_t)B->tile + 1) * sizeof(int32_t));
    int32_t best = 0;
    for (int64_t d = 0; d <= B->nI + B->nJ - 2; ++d) {
        const int64_t ilo = d - (B->nJ - 1) > 0 ? d - (B->nJ - 1) : 0;
        const int64_t ihi = d < B->nI - 1 ? d : B->nI - 1;
        for (;;) {
            pthread_mutex_lock(&B->mu);
            const int64_t k = B->next++;
            pthread_mutex_unlock(&B->mu);
            if (ilo + k > ihi) break;
            blocked_tile(B, ilo + k, d - (ilo + k), row, &best);
        }
        /* Everyone has finished diagonal d; thread 0 resets the counter
         * between the two barriers. */
        pthread_barrier_wait(&B->bar);
        if (arg->id == 0) B->next = 0;
        pthread_barrier_wait(&B->bar);
    }
    B->best[arg->id] = best;
    free(row);
    return NULL;
}

int32_t oracle_blocked(int problem, const int32_t *a, int64_t n1, const int32_t *b, int64_t n2, int textbook,
                       int threads, int64_t tile, int32_t *max_out)
{
    if (max_out) *max_out = 0;
    if (n1 <= 0 || n2 <= 0) {
        if (problem != ORACLE_ED) return 0;
        return n2 == 0 ? ed_top(n1, textbook) : (int32_t)n2;
    }
    if (threads < 1) threads = 1;
    if (tile < 16) tile = 16;
    Blocked B;
    memset(&B, 0, sizeof(B));
    B.problem = problem;
    B.textbook = textbook;
    B.threads = threads;
    B.a = a;
    B.b = b;
    B.n1 = n1;
    B.n2 = n2;
    B.tile = tile;
    B.nI = (n2 + tile - 1) / tile;
    B.nJ = (n1 + tile - 1) / tile;
    B.top = (int32_t *)malloc(((size_t)n1 + 1) * sizeof(int32_t));
    B.left = (int32_t *)malloc(((size_t)n2 + 1) * sizeof(int32_t));
    B.corner = (int32_t *)malloc((size_t)(B.nI + 1) * (size_t)(B.nJ + 1) * sizeof(int32_t));
    B.best = (int32_t *)calloc((size_t)threads, sizeof(int32_t));
    BlockedArg *args = (BlockedArg *)malloc((size_t)threads * sizeof(BlockedArg));
    pthread_t *tid = (pthread_t *)malloc((size_t)threads * sizeof(pthread_t));
    if (!B.top || !B.left || !B.corner || !B.best || !args || !tid) {
        free(B.top); free(B.left); free(B.corner); free(B.best); free(args); free(tid);
        return -1;
    }
    for (int64_t x = 0; x <= n1; ++x) B.top[x] = (problem == ORACLE_ED) ? ed_top(x, textbook) : 0;
    for (int64_t y = 0; y <= n2; ++y) B.left[y] = (problem == ORACLE_ED) ? (int32_t)y : 0;
    for (int64_t I = 0; I <= B.nI; ++I)
        for (int64_t J = 0; J <= B.nJ; ++J) {
            int32_t v = 0;
            if (problem == ORACLE_ED) {
                const int64_t y = I * tile < n2 ? I * tile : n2, x = J * tile < n1 ? J * tile : n1;
                v = (I == 0) ? ed_top(x, textbook) : (J == 0 ? (int32_t)y : 0);
            }
            B.corner[I * (B.nJ + 1) + J] = v;
        }
    pthread_barrier_init(&B.bar, NULL, (unsigned)threads);
    pthread_mutex_init(&B.mu, NULL);
    for (int t = 0; t < threads; ++t) {
        args[t].B = &B;
        args[t].id = t;
        pthread_create(&tid[t], NULL, blocked_worker, &args[t]);
    }
    for (int t = 0; t < threads; ++t) pthread_join(tid[t], NULL);
    pthread_barrier_destroy(&B.bar);
    pthread_mutex_destroy(&B.mu);
    int32_t m = 0;
    for (int t = 0; t < threads; ++t) m = imax(m, B.best[t]);
    if (max_out && problem == ORACLE_SW) *max_out = m;
    const int32_t r = B.top[n1];
    free(B.top); free(B.left); free(B.corner); free(B.best); free(args); free(tid);
    return r;
}
