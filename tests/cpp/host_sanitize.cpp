// The host library (loader, caches, writers, generators, batches) under AddressSanitizer and
// UndefinedBehaviorSanitizer: every entry point on generated graphs and on the loader fixtures
// given as arguments. Built and run by tests/test_host_logic.py.
#include <cstdio>
#include <string>
#include <vector>
#include "../../include/kcore_graph.h"
int main(int argc, char** argv) {
    int bad = 0;
    // generators
    kcg_graph* g = nullptr;
    if (kcg_gen_er(5000, 40000, 1, 4, &g)) return 1;
    kcg_write_edgelist(g, "asan_er.txt");
    kcg_graph* h = nullptr;
    if (kcg_load_edgelist("asan_er.txt", 4, &h)) return 2;
    kcg_write_serialized(h, "asan_ser");
    kcg_graph* s = nullptr;
    if (kcg_read_serialized("asan_ser", &s)) return 3;
    kcg_write_bincache(h, "asan_bin.kcg", 7);
    kcg_graph* b = nullptr;
    if (kcg_read_bincache("asan_bin.kcg", 7, &b)) return 4;
    int fc = 0; kcg_graph* c = nullptr;
    if (kcg_load_cached_bin("asan_er.txt", "./", 2, &c, &fc)) return 5;
    kcg_graph* cl = nullptr; if (kcg_gen_chung_lu(20000, 2.1, 1, 2800, 12, 3, 4, &cl)) return 6;
    kcg_graph* rm = nullptr; if (kcg_gen_rmat(14, 16 << 14, .57, .19, .19, 3, 4, &rm)) return 7;
    kcg_batch* bt = nullptr; if (kcg_gen_batch(40, 100, 5000, 4, 32, 2.1, 1, 4, &bt)) return 8;
    kcg_graph* bg = nullptr; if (kcg_batch_graph(bt, 3, &bg)) return 9;
    for (int i = 0; i < 30; ++i) { kcg_graph* r = nullptr; int k; if (kcg_gen_request(5, i, 50000, 1.0, 1000, 50000000, 4, &r, &k)) return 10; kcg_graph_free(r); }
    // fixtures of the loader
    for (int i = 1; i < argc; ++i) { kcg_graph* f = nullptr; int st = kcg_load_edgelist(argv[i], 2, &f); if (!st) kcg_graph_free(f); }
    uint32_t V; uint64_t arcs; uint32_t dmax;
    kcg_graph_info(h, &V, &arcs, &dmax);
    const uint32_t *off, *nbr, *deg; kcg_graph_arrays(h, &off, &nbr, &deg);
    kcg_write_kcore_json("asan.json", V, deg);
    kcg_graph_free(g); kcg_graph_free(h); kcg_graph_free(s); kcg_graph_free(b); kcg_graph_free(c);
    kcg_graph_free(cl); kcg_graph_free(rm); kcg_graph_free(bg); kcg_batch_free(bt);
    std::printf("ok V=%u arcs=%llu\n", V, (unsigned long long)arcs);
    return bad;
}
