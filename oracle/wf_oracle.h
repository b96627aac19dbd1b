/* wf_oracle.h -- CPU restatement of LCS/LCS.cpp, SmithWaterman/SW.cpp and
 * EditDistance/ED.cpp.  Test infrastructure only (see wf_oracle.c). */
#ifndef WF_ORACLE_H
#define WF_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORACLE_LCS = 0, ORACLE_ED = 1, ORACLE_SW = 2 };

int32_t oracle_lcs(const int32_t *a, int64_t n1, const int32_t *b, int64_t n2);
int32_t oracle_sw(const int32_t *a, int64_t n1, const int32_t *b, int64_t n2,
                  int32_t *max_out);
int32_t oracle_ed(const int32_t *a, int64_t n1, const int32_t *b, int64_t n2,
                  int textbook);
int32_t oracle_table(int problem, const int32_t *a, int64_t n1,
                     const int32_t *b, int64_t n2, int textbook,
                     int32_t *table);
int oracle_last_row(int problem, const int32_t *a, int64_t n1,
                    const int32_t *b, int64_t n2, int textbook, int32_t *out);
/* In-place table stencils (SAT/GPU_Hyperlane_Share.cu,
 * 2D-SOR/nested_loop/GPU_Hyperlane_Share.cu): op 0 = SAT, 1 = SOR. */
enum { ORACLE_SAT = 0, ORACLE_SOR = 1 };
void oracle_stencil(int op, int64_t n1, int64_t n2, int32_t *t, int64_t ld);
/* The same results computed by `threads` threads over anti-diagonals of
 * tile x tile blocks (the all-cores CPU baseline). */
int32_t oracle_blocked(int problem, const int32_t *a, int64_t n1,
                       const int32_t *b, int64_t n2, int textbook,
                       int threads, int64_t tile, int32_t *max_out);
void oracle_batch(const int32_t *problems, const int32_t *const *as,
                  const int64_t *n1s, const int32_t *const *bs,
                  const int64_t *n2s, int64_t npairs, int32_t *results,
                  int32_t *maxs);

#ifdef __cplusplus
}
#endif
#endif
