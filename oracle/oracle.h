/*
 * oracle.h — CPU oracle for k-core peeling. TEST INFRASTRUCTURE ONLY: it is imported by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker or the CPU
 * baseline, never by the product path (libkcore_hip.so / the kcore driver).
 *
 * Functions and what pins each (details in oracle.c):
 *   kco_bz            Batagelj-Zaversnik O(n+m) bucket peel (Batagelj & Zaversnik 2003).
 *                     Pinned against kco_bruteforce on every graph with n <= 16, closed forms,
 *                     networkx core_number fixtures, and kco_levelsync.
 *   kco_levelsync     Restatement of the reference's level-synchronous peel (src/ours.cc:2-160):
 *                     full degree==level scan per level, check / fetch_sub / enqueue on
 *                     a == level+1 / undo on a <= level, stop at count == V, kmax = level-1.
 *                     OpenMP, all host threads by default. Pinned against kco_bz and brute force.
 *   kco_bruteforce    Definition-level k-core for n <= 16: core[v] = max over vertex subsets S
 *                     containing v of the minimum degree of G[S]. Independent of any peel.
 *   kco_check_certificate  Sound and complete check of core numbers given a peel order:
 *                     (A) every v has >= core[v] neighbours u with core[u] >= core[v]
 *                     (so core <= true core); (B) the order is a permutation along which core
 *                     is non-decreasing and every v has <= core[v] later neighbours
 *                     (so core >= true core).
 */
#ifndef KCORE_ORACLE_H
#define KCORE_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint32_t level;     /* peel level */
    uint32_t selected;  /* vertices found by the degree==level scan */
    uint32_t peeled;    /* all vertices peeled at this level (scan + cascade) */
    uint32_t rounds;    /* synchronous cascade sub-rounds needed to drain the level */
    uint64_t arcs;      /* adjacency entries visited while processing the level */
    uint64_t decrements;/* successful (not undone) decrements */
} kco_level_stat;

/* Returns kmax; core[V] receives the core numbers. */
uint32_t kco_bz(uint32_t V, const uint32_t* offsets, const uint32_t* nbrs, uint32_t* core);

/* Returns 0 on success, -1 on a no-progress guard trip (cannot happen on a valid CSR).
 * order (nullable, V entries) receives the enqueue order; shell_off (nullable, dmax+2
 * entries) the start of each level in order; stats (nullable, stats_cap entries) per-level
 * counters; *nlevels the number of levels run (= kmax + 1). nthreads <= 0 = all threads. */
int kco_levelsync(uint32_t V, const uint32_t* offsets, const uint32_t* nbrs,
                  const uint32_t* degrees, int nthreads, uint32_t* core, uint32_t* kmax,
                  uint32_t* order, uint32_t* shell_off, kco_level_stat* stats,
                  uint32_t stats_cap, uint32_t* nlevels);

/* n <= 16. Returns kmax, or UINT32_MAX when n is too large. */
uint32_t kco_bruteforce(uint32_t V, const uint32_t* offsets, const uint32_t* nbrs, uint32_t* core);

/* 0 = valid; otherwise a code naming the first failed condition (1 = A, 2 = order not a
 * permutation, 3 = core decreases along order, 4 = too many later neighbours). */
int kco_check_certificate(uint32_t V, const uint32_t* offsets, const uint32_t* nbrs,
                          const uint32_t* core, const uint32_t* order);

/* Condition (A) only (core numbers never exceed the true ones). 0 = holds. */
int kco_check_lower(uint32_t V, const uint32_t* offsets, const uint32_t* nbrs, const uint32_t* core);

#ifdef __cplusplus
}
#endif
#endif
