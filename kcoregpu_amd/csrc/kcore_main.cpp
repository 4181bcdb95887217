// kcore_main.cpp — the `./kcore <edgelist>` driver (reference main.cu:44-91).
//
// With one argument it prints the reference's lines: loader messages, the dataset banner,
// V and E (arcs), one timing line per variant label (integer ms of the peel), Kmax after
// "Ours", and "Deallocated... " from the Graph destructor. The reference's nine labels map
// to this engine's variants: Ours/BC/EC -> wave, SM/BC + SM/EC + SM -> LDS buffer,
// VP/BC + VP/EC + VP -> prefetch (its BC/EC compaction flavours are one wave64 ballot
// compaction here).
//
// Extra options (after the edge list), none needed for the reference behaviour:
//   --core-out PATH   write the core numbers as raw little-endian uint32[V]
//   --core-text PATH  write the core numbers as text, one per line (the serialized cache's
//                     degree-block form)
//   --json            write ./output/pkc-kcore-<edgelist> like Graph::writeKCoreToDisk
//   --order-out PATH  write the peel order (uint32[V]) and PATH.shells (k-shell starts)
//   --variant NAME    run only one variant: wave | lds | prefetch | async | auto | stepwise (the
//                     reference-shaped host level loop)
//   --reps N          repeat each variant N times (the reference uses 1)
//   --stats           print a JSON stats line per run
//   --output-dir DIR  directory for the serialized cache and JSON (default ./output/)
//   --binary-cache    use the content-keyed binary CSR cache instead of the text one
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "../../include/kcore.h"
#include "graph.h"

static int run_variant(kc_ctx* ctx, kc_graph* dg, int variant, int reps, bool stats, uint32_t* kmax) {
    kc_params p{};
    p.variant = variant;
    for (int r = 0; r < reps; ++r) {
        kc_stats st{};
        kc_status s = kc_peel(ctx, dg, &p, nullptr, kmax, &st);
        if (s != KC_OK) {
            std::cout << "peel failed: " << kc_status_str(s) << std::endl;
            return -1;
        }
        std::cout << (long long)st.ms_peel << " ";
        if (stats)
            std::fprintf(stderr,
                         "{\"variant\": %d, \"ms_peel\": %.4f, \"kmax\": %u, \"levels\": %u, "
                         "\"bytes_alg\": %llu, \"GBps\": %.2f, \"grid\": %u, \"block\": %u}\n",
                         variant, st.ms_peel, st.kmax, st.levels, (unsigned long long)st.bytes_alg,
                         st.bytes_alg / (st.ms_peel * 1e6), st.grid, st.block);
    }
    std::cout << std::endl;
    return 0;
}

int main(int argc, char* argv[]) {
    if (argc < 2) {
        std::cout << "Please provide data file" << std::endl;
        return -1;
    }
    std::string ds = argv[1];
    std::string core_out, core_text, order_out, only;
    bool json = false, stats = false, binary = false;
    int reps = 1;
    for (int i = 2; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() -> std::string {
            if (i + 1 >= argc) {
                std::cerr << "missing value for " << a << std::endl;
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "--core-out") core_out = next();
        else if (a == "--core-text") core_text = next();
        else if (a == "--order-out") order_out = next();
        else if (a == "--json") json = true;
        else if (a == "--stats") stats = true;
        else if (a == "--binary-cache") binary = true;
        else if (a == "--variant") only = next();
        else if (a == "--reps") reps = std::atoi(next().c_str());
        else if (a == "--output-dir") {
            Graph::output_dir = next();
            if (!Graph::output_dir.empty() && Graph::output_dir.back() != '/') Graph::output_dir += '/';
        } else {
            std::cerr << "unknown option " << a << std::endl;
            return 2;
        }
    }
    if (reps < 1) reps = 1;

    std::cout << "Graph loading Started... " << std::endl;
    std::unique_ptr<Graph> gp;
    if (binary) {
        gp.reset(new Graph());
        gp->loadBinaryCached(ds);
    } else {
        gp.reset(new Graph(ds));
    }
    Graph& g = *gp;
    // the reference exits inside readFile on a missing file (src/graph.cpp:63-66), before the
    // Graph destructor could print "Deallocated... ": std::exit does not unwind this frame either
    if (g.V == 0) std::exit(-1);
    std::cout << "******************  " << ds << " ****************** " << std::endl;
    std::cout << "V: " << g.V << std::endl;
    std::cout << "E: " << g.E << std::endl;

    kc_ctx* ctx = nullptr;
    kc_status s = kc_open(0, &ctx);
    if (s != KC_OK) {
        std::cout << kc_status_str(s) << std::endl;
        return -1;
    }
    kc_csr h{g.V, g.E, g.neighbors_offset, g.neighbors, g.degrees};
    kc_graph* dg = nullptr;
    s = kc_graph_upload(ctx, &h, &dg);
    if (s != KC_OK) {
        std::cout << kc_status_str(s) << std::endl;
        kc_close(ctx);
        return -1;
    }

    struct Label {
        const char* name;
        int variant;
    };
    const Label labels[] = {{"Ours: ", KC_VARIANT_WAVE},     {"SM:  ", KC_VARIANT_LDS},
                            {"VP: ", KC_VARIANT_PREFETCH},    {"BC: ", KC_VARIANT_WAVE},
                            {"BC + SM: ", KC_VARIANT_LDS},     {"BC + VP: ", KC_VARIANT_PREFETCH},
                            {"EC: ", KC_VARIANT_WAVE},         {"EC + SM: ", KC_VARIANT_LDS},
                            {"EC + VP: ", KC_VARIANT_PREFETCH}};
    int rc = 0;
    uint32_t kmax = 0;
    if (!only.empty()) {
        int v = only == "wave" ? KC_VARIANT_WAVE
              : only == "lds" ? KC_VARIANT_LDS
              : only == "prefetch" ? KC_VARIANT_PREFETCH
              : only == "async" ? KC_VARIANT_ASYNC
              : only == "auto" ? KC_VARIANT_AUTO
              : only == "stepwise" ? KC_VARIANT_STEPWISE : -1;
        if (v < 0) {
            std::cerr << "unknown variant " << only << std::endl;
            rc = 2;
        } else {
            std::cout << only << ": ";
            rc = run_variant(ctx, dg, v, reps, stats, &kmax);
            if (!rc) std::cout << "Kmax: " << kmax << std::endl;
        }
    } else {
        for (size_t i = 0; i < sizeof(labels) / sizeof(labels[0]) && !rc; ++i) {
            std::cout << labels[i].name;
            rc = run_variant(ctx, dg, labels[i].variant, reps, stats, &kmax);
            if (i == 0 && !rc) std::cout << "Kmax: " << kmax << std::endl;
        }
    }
    g.kmax = kmax;
    if (!rc && (!core_out.empty() || !core_text.empty() || json)) {
        std::vector<uint32_t> core(g.V);
        s = kc_graph_download_core(ctx, dg, core.data());
        if (s != KC_OK) {
            std::cout << kc_status_str(s) << std::endl;
            rc = -1;
        } else {
            if (!core_out.empty() && kcg::write_core_u32(core_out, g.V, core.data()) != kcg::OK) {
                std::cout << core_out << " : File couldn't open" << std::endl;
                rc = -1;
            }
            if (!core_text.empty() && kcg::write_core_text(core_text, g.V, core.data()) != kcg::OK) {
                std::cout << core_text << " : File couldn't open" << std::endl;
                rc = -1;
            }
            if (json) kcg::write_kcore_json(Graph::output_dir + "pkc-kcore-" + ds, g.V, core.data());
        }
    }
    if (!rc && !order_out.empty()) {
        std::vector<uint32_t> order(g.V), shells(kmax + 2);
        s = kc_graph_download_order(ctx, dg, order.data(), shells.data());
        if (s == KC_OK) {
            kcg::write_core_u32(order_out, g.V, order.data());
            kcg::write_core_u32(order_out + ".shells", kmax + 2, shells.data());
        } else {
            rc = -1;
        }
    }
    kc_graph_free(dg);
    kc_close(ctx);
    return rc;
}
