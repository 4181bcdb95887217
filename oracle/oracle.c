/*
 * oracle.c — CPU oracle for k-core peeling (see oracle.h). TEST INFRASTRUCTURE ONLY.
 * Plain C11 + OpenMP. Every function cites the reference behaviour it restates.
 */
#include "oracle.h"

#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* Batagelj & Zaversnik, "An O(m) Algorithm for Cores Decomposition of Networks", 2003:
 * vertices kept in an array sorted by current degree with bucket starts; taking vertices in
 * that order, each neighbour u with deg[u] > deg[v] is moved to the front of its bucket and
 * its degree lowered by one. The degree at removal is the core number. */
uint32_t kco_bz(uint32_t V, const uint32_t* off, const uint32_t* nb, uint32_t* core) {
    uint32_t md = 0;
    for (uint32_t v = 0; v < V; ++v) {
        core[v] = off[v + 1] - off[v];
        if (core[v] > md) md = core[v];
    }
    uint32_t* bin = (uint32_t*)calloc((size_t)md + 2, sizeof(uint32_t));
    uint32_t* pos = (uint32_t*)malloc((size_t)V * sizeof(uint32_t) + 4);
    uint32_t* vert = (uint32_t*)malloc((size_t)V * sizeof(uint32_t) + 4);
    for (uint32_t v = 0; v < V; ++v) bin[core[v]]++;
    uint32_t start = 0;
    for (uint32_t d = 0; d <= md; ++d) {
        uint32_t num = bin[d];
        bin[d] = start;
        start += num;
    }
    for (uint32_t v = 0; v < V; ++v) {
        pos[v] = bin[core[v]];
        vert[pos[v]] = v;
        bin[core[v]]++;
    }
    for (uint32_t d = md; d >= 1; --d) bin[d] = bin[d - 1];
    bin[0] = 0;
    uint32_t kmax = 0;
    for (uint32_t i = 0; i < V; ++i) {
        uint32_t v = vert[i];
        uint32_t cv = core[v];
        if (cv > kmax) kmax = cv;
        for (uint32_t j = off[v]; j < off[v + 1]; ++j) {
            uint32_t u = nb[j];
            if (core[u] > cv) {
                uint32_t du = core[u], pu = pos[u], pw = bin[du], w = vert[pw];
                if (u != w) {
                    pos[u] = pw;
                    vert[pu] = w;
                    pos[w] = pu;
                    vert[pw] = u;
                }
                bin[du]++;
                core[u]--;
            }
        }
    }
    free(bin);
    free(pos);
    free(vert);
    return kmax;
}

/* Level-synchronous peel, restating src/ours.cc:
 *   select  (src/ours.cc:2-33):   every v with deg[v] == level is appended to the frontier;
 *   process (src/ours.cc:38-114): for each frontier v and each neighbour u with deg[u] > level,
 *            a = fetch_sub(deg[u], 1); a == level+1 appends u to the same level's frontier;
 *            a <= level undoes the decrement;
 *   host loop (src/ours.cc:138-151): repeat per level until count == V; kmax = level - 1.
 * The reference drains cascades inside per-block queues; here each level is drained in
 * synchronous sub-rounds over one shared queue. Core numbers do not depend on the order. */
int kco_levelsync(uint32_t V, const uint32_t* off, const uint32_t* nb, const uint32_t* degrees,
                  int nthreads, uint32_t* deg, uint32_t* kmax_out, uint32_t* order,
                  uint32_t* shell_off, kco_level_stat* stats, uint32_t stats_cap,
                  uint32_t* nlevels) {
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
    uint32_t* q = order ? order : (uint32_t*)malloc((size_t)V * sizeof(uint32_t) + 4);
    uint32_t dmax = 0;
    for (uint32_t v = 0; v < V; ++v) {
        deg[v] = degrees[v];
        if (deg[v] > dmax) dmax = deg[v];
    }
    uint32_t level = 0;
    uint64_t tail = 0, head = 0;
    int rc = 0;
    while (tail < V) {
        if (level > dmax) { /* no-progress guard */
            rc = -1;
            break;
        }
        if (shell_off) shell_off[level] = (uint32_t)tail;
        uint64_t t0 = tail;
#pragma omp parallel for num_threads(nt) schedule(static)
        for (int64_t v = 0; v < (int64_t)V; ++v) {
            if (deg[v] == level) {
                uint64_t p = __atomic_fetch_add(&tail, 1, __ATOMIC_RELAXED);
                q[p] = (uint32_t)v;
            }
        }
        uint32_t selected = (uint32_t)(tail - t0);
        uint32_t rounds = 0;
        uint64_t arcs = 0, decs = 0;
        while (head < tail) {
            uint64_t end = tail;
            rounds++;
#pragma omp parallel for num_threads(nt) schedule(dynamic, 64) reduction(+ : arcs, decs)
            for (int64_t i = (int64_t)head; i < (int64_t)end; ++i) {
                uint32_t v = q[i];
                for (uint32_t j = off[v]; j < off[v + 1]; ++j) {
                    uint32_t u = nb[j];
                    arcs++;
                    if (__atomic_load_n(&deg[u], __ATOMIC_RELAXED) > level) {
                        uint32_t a = __atomic_fetch_sub(&deg[u], 1u, __ATOMIC_RELAXED);
                        if (a == level + 1) {
                            uint64_t p = __atomic_fetch_add(&tail, 1, __ATOMIC_RELAXED);
                            q[p] = u;
                        }
                        if (a <= level)
                            __atomic_fetch_add(&deg[u], 1u, __ATOMIC_RELAXED);
                        else
                            decs++;
                    }
                }
            }
            head = end;
        }
        if (stats && level < stats_cap) {
            stats[level].level = level;
            stats[level].selected = selected;
            stats[level].peeled = (uint32_t)(tail - t0);
            stats[level].rounds = rounds;
            stats[level].arcs = arcs;
            stats[level].decrements = decs;
        }
        level++;
    }
    if (shell_off) shell_off[level] = (uint32_t)tail;
    if (kmax_out) *kmax_out = level ? level - 1 : 0;
    if (nlevels) *nlevels = level;
    if (!order) free(q);
    return rc;
}

/* Definition-level brute force: the k-core is the largest vertex set whose induced subgraph
 * has minimum degree >= k, and the union of such sets is again one, so core[v] is the
 * maximum, over subsets S containing v, of the minimum induced degree of S. */
uint32_t kco_bruteforce(uint32_t V, const uint32_t* off, const uint32_t* nb, uint32_t* core) {
    if (V > 16) return UINT32_MAX;
    uint32_t adj[16] = {0};
    for (uint32_t v = 0; v < V; ++v)
        for (uint32_t j = off[v]; j < off[v + 1]; ++j) adj[v] |= 1u << nb[j];
    for (uint32_t v = 0; v < V; ++v) core[v] = 0;
    for (uint32_t S = 1; S < (1u << V); ++S) {
        uint32_t md = UINT32_MAX;
        for (uint32_t v = 0; v < V; ++v)
            if (S >> v & 1) {
                uint32_t d = (uint32_t)__builtin_popcount(adj[v] & S);
                if (d < md) md = d;
            }
        for (uint32_t v = 0; v < V; ++v)
            if ((S >> v & 1) && md > core[v]) core[v] = md;
    }
    uint32_t kmax = 0;
    for (uint32_t v = 0; v < V; ++v)
        if (core[v] > kmax) kmax = core[v];
    return kmax;
}

int kco_check_lower(uint32_t V, const uint32_t* off, const uint32_t* nb, const uint32_t* core) {
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(| : bad)
    for (int64_t v = 0; v < (int64_t)V; ++v) {
        uint32_t c = core[v], k = 0;
        for (uint32_t j = off[v]; j < off[v + 1]; ++j) k += core[nb[j]] >= c;
        if (k < c) bad = 1;
    }
    return bad;
}

int kco_check_certificate(uint32_t V, const uint32_t* off, const uint32_t* nb,
                          const uint32_t* core, const uint32_t* order) {
    if (kco_check_lower(V, off, nb, core)) return 1;
    uint32_t* rank = (uint32_t*)malloc((size_t)V * sizeof(uint32_t) + 4);
    for (uint32_t v = 0; v < V; ++v) rank[v] = UINT32_MAX;
    int rc = 0;
    for (uint32_t i = 0; i < V && !rc; ++i) {
        uint32_t v = order[i];
        if (v >= V || rank[v] != UINT32_MAX) rc = 2;
        else rank[v] = i;
        if (!rc && i > 0 && core[order[i - 1]] > core[v]) rc = 3;
    }
    if (!rc) {
        int bad = 0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(| : bad)
        for (int64_t v = 0; v < (int64_t)V; ++v) {
            uint32_t later = 0;
            for (uint32_t j = off[v]; j < off[v + 1]; ++j) later += rank[nb[j]] > rank[v];
            if (later > core[v]) bad = 1;
        }
        if (bad) rc = 4;
    }
    free(rank);
    return rc;
}
