/*
 * wf_cpu_main.c -- the reference's CPU programs as drop-ins, over the C
 * restatement in wf_oracle.c.  Built three times (oracle/Makefile):
 * -DCPU_LCS -> bin/LCS (LCS/LCS.cpp:86-166), -DCPU_SW -> bin/SW
 * (SmithWaterman/SW.cpp:87-167), -DCPU_ED -> bin/ED (EditDistance/ED.cpp
 * :85-165).
 *
 * TEST INFRASTRUCTURE AND CPU BASELINE ONLY, like the rest of oracle/: run by
 * tests/ and by bench.py's cpu_baseline leg, never by the product path.
 *
 *   LCS|SW|ED e1 e2 [--faithful] [--threads N]
 *
 * Reads ./Data/str1_2_<e1>_str2_2_<e2>.txt (LCS/LCS.cpp:96-121), times only
 * the DP call with gettimeofday (:127-135) and prints what the reference
 * prints (:137-138; SW.cpp:136-139 labels its milliseconds "second", and so
 * does this program; ED.cpp:134-137).  A wrong argc prints the reference's
 * usage line and exits 1; an unreadable file prints its error line and exits
 * 1 (LCS/LCS.cpp:16-17, :88-90).
 *
 *   default      rolling rows, 64-bit indices: any size (the reference
 *                programs overflow int past n = 46340, LCS.cpp:51)
 *   --faithful   the reference's own memory behaviour: the whole
 *                (n2+1) x (n1+1) int32 table allocated before the timer and
 *                filled in its loop order inside it (LCS/LCS.cpp:123-133)
 *   --threads N  N threads over anti-diagonals of 256 x 256 tiles (0 = all
 *                online cores); prints "threads: N" after the result lines
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>
#include <unistd.h>

#include "wf_oracle.h"

#if defined(CPU_LCS)
#define PROBLEM ORACLE_LCS
#elif defined(CPU_SW)
#define PROBLEM ORACLE_SW
#elif defined(CPU_ED)
#define PROBLEM ORACLE_ED
#else
#error "build with -DCPU_LCS, -DCPU_SW or -DCPU_ED"
#endif

static void usage_exit(void)
{
    printf("Incorrect Input Parameters. Must be two string sizes.\n");
    exit(EXIT_FAILURE);
}

/* "n1 n2" then n1 ints then n2 ints, whitespace separated
 * (LCS/LCS.cpp:11-31). */
static int read_input(const char *path, int64_t *n1, int64_t *n2, int32_t **a, int32_t **b)
{
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    long long x = 0, y = 0;
    if (fscanf(f, "%lld %lld", &x, &y) != 2 || x < 0 || y < 0) {
        fclose(f);
        return -1;
    }
    *n1 = x;
    *n2 = y;
    *a = (int32_t *)malloc(((size_t)x + 1) * sizeof(int32_t));
    *b = (int32_t *)malloc(((size_t)y + 1) * sizeof(int32_t));
    if (!*a || !*b) {
        fclose(f);
        return -1;
    }
    for (long long i = 0; i < x; ++i)
        if (fscanf(f, "%d", &(*a)[i]) != 1) (*a)[i] = 0;
    for (long long i = 0; i < y; ++i)
        if (fscanf(f, "%d", &(*b)[i]) != 1) (*b)[i] = 0;
    fclose(f);
    return 0;
}

int main(int argc, char **argv)
{
    int faithful = 0, threads = -1;
    int pos[2], npos = 0;
    for (int i = 1; i < argc; ++i) {
        if (strcmp(argv[i], "--faithful") == 0) {
            faithful = 1;
        } else if (strcmp(argv[i], "--threads") == 0 && i + 1 < argc) {
            threads = atoi(argv[++i]);
        } else if (npos < 2) {
            pos[npos++] = atoi(argv[i]);
        } else {
            usage_exit();
        }
    }
    if (npos != 2) usage_exit();
    if (threads == 0) threads = (int)sysconf(_SC_NPROCESSORS_ONLN);

    char path[256];
    snprintf(path, sizeof(path), "./Data/str1_2_%d_str2_2_%d.txt", pos[0], pos[1]);
    int64_t n1 = 0, n2 = 0;
    int32_t *a = NULL, *b = NULL;
    if (read_input(path, &n1, &n2, &a, &b) != 0) {
        printf("ERROR: Input file cannot be opened!\n");
        exit(EXIT_FAILURE);
    }

    int32_t *table = NULL;
    if (faithful) {
        table = (int32_t *)malloc((size_t)(n1 + 1) * (size_t)(n2 + 1) * sizeof(int32_t));
        if (!table) {
            printf("ERROR: the full table does not fit in memory.\n");
            exit(EXIT_FAILURE);
        }
    }
    int32_t score = 0, mx = 0;
    struct timeval t0, t1;
    gettimeofday(&t0, NULL);
    if (faithful)
        score = oracle_table(PROBLEM, a, n1, b, n2, 0, table);
    else if (threads > 0)
        score = oracle_blocked(PROBLEM, a, n1, b, n2, 0, threads, 256, &mx);
    else if (PROBLEM == ORACLE_LCS)
        score = oracle_lcs(a, n1, b, n2);
    else if (PROBLEM == ORACLE_SW)
        score = oracle_sw(a, n1, b, n2, &mx);
    else
        score = oracle_ed(a, n1, b, n2, 0);
    gettimeofday(&t1, NULL);
    const double ms = (double)(t1.tv_sec - t0.tv_sec) * 1000.0 + (double)(t1.tv_usec - t0.tv_usec) / 1000.0;

#if defined(CPU_LCS)
    printf("lcs length: %d\n", score);
    printf("execution time: %g ms.\n", ms);
#elif defined(CPU_SW)
    printf("last: %d\n", score);
    printf("execution time: %g second.\n", ms);
#else
    printf("last: %d\n", score);
    printf("execution time: %g ms.\n", ms);
#endif
    if (threads > 0) printf("threads: %d\n", threads);
    free(table);
    free(a);
    free(b);
    return 0;
}
