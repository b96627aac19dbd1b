/*
 * wfdp_compat.h -- the reference's per-problem host functions with their
 * exact C++ signatures, implemented over wf_align() (include/wfdp.h).
 *
 * Link libwfdp_compat_share.so or libwfdp_compat_cache.so (plus libwfdp.so)
 * in place of the reference's .cu objects and its driver .cpp files build
 * unchanged.  (The SmithWaterman and EditDistance Cache drivers include the
 * Share header, GPU_Hyperlane_Cache.cpp:7, whose 9-argument declaration does
 * not match their 7-argument call at :101; compile them with
 * -include wfdp_compat.h, which declares both.)  The two libraries differ
 * only in the tile policy of the calls that carry no tile arguments:
 *   share: LCS 32x256 (LCS/GPU_Hyperlane_Share.cu:703-704)
 *   cache: 256x1024, strip boundaries through L2 only
 *          (LCS/GPU_Hyperlane_Cache.cu:254-255 and the SW/ED Cache hosts)
 *
 * Semantics kept: return value = bottom-right DP cell; `table`, `paddX` and
 * `paddY` are ignored unless WFDP_DEBUG_TABLE=1 is set in the environment, the
 * run-time counterpart of the reference's DEBUG build, which copies the whole
 * table back (LCS/GPU_Hyperlane_Share.cu:770-772): then a square pair's
 * table[(i + paddY) * (n2 + paddX) + (j + paddX)] holds H(i + 1, j + 1) for
 * i, j < n (wf_align_table); on a device error the process prints the error and exits
 * with EXIT_FAILURE, as checkGPUError does (LCS/GPU_Hyperlane_Share.cu
 * :689-694).  The device is 0, as in the reference (cudaSetDevice(0)).
 */
#ifndef WFDP_COMPAT_H
#define WFDP_COMPAT_H

/* LCS/GPU_Hyperlane_Share.h:5, LCS/GPU_Hyperlane_Cache.h:5,
 * LCS-SHFL/GPU_Hyperlane_Share.h:5, LCS-SHFL/GPU_Hyperlane_Cache.h:5 */
int LCS(int n1, int n2, int *arr1, int *arr2, int paddX, int paddY, int *table);
/* LCS/GPU.h:5, LCS-SHFL/GPU.h:5 (the naive programs' form; in the share and
 * cache libraries it runs the hyperlane path with the library's tile) */
int LCS(int n1, int n2, int *arr1, int *arr2);
/* SmithWaterman/GPU_Hyperlane_Share.h:5 */
int SW(int n1, int n2, int *arr1, int *arr2, int paddX, int paddY, int *table, int tX, int tY);
/* SmithWaterman/GPU_Hyperlane_Cache.h:5 */
int SW(int n1, int n2, int *arr1, int *arr2, int paddX, int paddY, int *table);
/* EditDistance/GPU_Hyperlane_Share.h:5 */
int ED(int n1, int n2, int *arr1, int *arr2, int paddX, int paddY, int *table, int tX, int tY);
/* EditDistance/GPU_Hyperlane_Cache.h:5 */
int ED(int n1, int n2, int *arr1, int *arr2, int paddX, int paddY, int *table);

/* The naive baselines' functions (libwfdp_compat_naive.so: one launch per
 * cell anti-diagonal; libwfdp_compat_tile.so: one launch per tile
 * anti-diagonal, LCS only).  They print the reference's memory line first
 * (LCS/GPU.cu:51-56).  LCS(n1, n2, arr1, arr2) above is LCS/GPU.h:5 and
 * LCS/GPU_TILE_SHARE.h:5 in these libraries. */
/* SmithWaterman/GPU.h:5 */
int SW(int n1, int n2, int *arr1, int *arr2);
/* EditDistance/GPU.h:5 (textbook row 0, EditDistance/GPU.cu:12-18) */
int ED(int n1, int n2, int *arr1, int *arr2);

/* The SAT and 2D-SOR programs' host function, named SOR in both
 * (libwfdp_compat_sat.so: summed-area table; libwfdp_compat_sor.so: one
 * Gauss-Seidel sweep).  `table` holds (n2 + 2) rows of n1 + 3 values; like
 * the reference it is left unchanged unless WFDP_DEBUG_TABLE=1, which writes
 * the swept table back. */
/* SAT/GPU_Hyperlane_Share.h:5, 2D-SOR/nested_loop/GPU_Hyperlane_Share.h:5 */
void SOR(int n1, int n2, int *table, int tX, int tY);
/* SAT/GPU_Hyperlane_Cache.h:5, 2D-SOR/nested_loop/GPU_Hyperlane_Cache.h:5 */
void SOR(int n1, int n2, int *table);

/* Extension used by the drop-in LCS driver's optional 5-argument form. */
int LCS(int n1, int n2, int *arr1, int *arr2, int paddX, int paddY, int *table, int tX, int tY);

#endif /* WFDP_COMPAT_H */
