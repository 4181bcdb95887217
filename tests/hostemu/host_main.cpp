// host_main.cpp — drives the real host code of libkcore_hip (kcore_hip.hip, serve.hip, tune.cpp,
// built against the CPU runtime stand-in in tests/hostemu/hip) through the C ABI of
// include/kcore.h, with every kernel run by the wave64 emulator. Test infrastructure:
// tests/test_hostemu.py builds it and checks every result against the oracle.
//
//   host peel <graph> <out>                 every variant and a few settings, repeated peels
//   host batch <out> <graphs...>            eager, captured-graph replays, tuner
//   host serve <out> <large_arcs> <budget> <graphs...>
//                                            a served stream: pinned and pageable outputs, a
//                                            malformed request, an oversized one, frees and a
//                                            whole single-graph peel while the kernel is resident
//   host tune <graph> <cache>                tuner, then the cached answer
// <graph>: u32 V, u64 arcs, u32 offsets[V+1], u32 neighbours[arcs]. Outputs: see each scenario.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/kcore.h"

namespace {

struct G {
    uint32_t V = 0;
    uint64_t arcs = 0;
    std::vector<uint32_t> off, nbr;
};

G read_graph(const char* path) {
    G g;
    FILE* f = std::fopen(path, "rb");
    if (!f) std::exit(3);
    if (std::fread(&g.V, 4, 1, f) != 1 || std::fread(&g.arcs, 8, 1, f) != 1) std::exit(3);
    g.off.resize((size_t)g.V + 1);
    g.nbr.resize(g.arcs);
    if (std::fread(g.off.data(), 4, g.off.size(), f) != g.off.size()) std::exit(3);
    if (g.arcs && std::fread(g.nbr.data(), 4, g.arcs, f) != g.arcs) std::exit(3);
    std::fclose(f);
    return g;
}

void put(FILE* f, uint32_t x) { std::fwrite(&x, 4, 1, f); }
void put(FILE* f, const uint32_t* p, size_t n) { std::fwrite(p, 4, n, f); }

#define CHECK(c)                                                              \
    do {                                                                      \
        if (!(c)) {                                                           \
            std::fprintf(stderr, "host: check failed %s:%d: %s\n", __FILE__, __LINE__, #c); \
            std::exit(5);                                                     \
        }                                                                     \
    } while (0)

kc_ctx* open_ctx() {
    kc_ctx* c = nullptr;
    CHECK(kc_open(0, &c) == KC_OK);
    return c;
}

// per run: u32 status, u32 kmax, u32 levels, core[V], order[V], shells[kmax + 2]
int run_peel(char** argv) {
    G g = read_graph(argv[2]);
    FILE* out = std::fopen(argv[3], "wb");
    kc_ctx* c = open_ctx();
    CHECK(kc_selftest(c) == KC_OK);
    double gbps = 0;
    CHECK(kc_copy_bandwidth(c, 1u << 20, 2, &gbps) == KC_OK && gbps > 0);
    CHECK(kc_copy_bandwidth(c, 8, 2, &gbps) == KC_ERR_ARG);
    kc_csr h{g.V, g.arcs, g.off.data(), g.nbr.data(), nullptr};
    kc_graph* dg = nullptr;
    CHECK(kc_graph_upload(c, &h, &dg) == KC_OK);
    struct Run {
        int variant;
        uint32_t block, heavy, chunk, flags, ratio;
    };
    const Run runs[] = {
        {0, 0, 0, 0, 0, 0},                                  // defaults (LDS)
        {KC_VARIANT_WAVE, 256, 16, 64, 0, 0},                // small chunks: chunked hubs
        {KC_VARIANT_PREFETCH, 1024, 0, 0, 0, 0},
        {KC_VARIANT_LDS, 512, 0, 0, KC_FLAG_NO_LEVEL_SKIP, 0},
        {KC_VARIANT_STEPWISE, 0, 0, 0, 0, 0},
        {KC_VARIANT_STEPWISE, 0, 0, 0, KC_FLAG_NO_LEVEL_SKIP, 0},
        {KC_VARIANT_ASYNC, 256, 16, 64, 0, 1},               // compaction at every chance
        {KC_VARIANT_ASYNC, 0, 0, 0, 0, 0},                   // the same graph again: queues re-armed
        {KC_VARIANT_ASYNC, 512, 0, 0, KC_FLAG_DIRECT_POLL | KC_FLAG_NO_LEVEL_SKIP, 0},
        {KC_VARIANT_ASYNC, 256, 16, 64, KC_FLAG_LEVEL_HANDOFF, 1},  // level hand-off, no grid barriers
        {KC_VARIANT_ASYNC, 0, 0, 0, KC_FLAG_LEVEL_HANDOFF | KC_FLAG_NO_LEVEL_SKIP, 0},
        {0, 0, 0, 0, 0, 0},                                  // defaults after the cascade-queue runs
    };
    for (const Run& r : runs) {
        kc_params p{};
        p.variant = r.variant;
        p.block = r.block;
        p.heavy_threshold = r.heavy;
        p.chunk_arcs = r.chunk;
        p.flags = r.flags;
        p.compact_ratio = r.ratio;
        std::vector<uint32_t> core(g.V, 0xEEEEEEEEu), order(g.V, 0xEEEEEEEEu);
        uint32_t kmax = 0;
        kc_stats st{};
        const kc_status s = kc_peel(c, dg, &p, core.data(), &kmax, &st);
        std::vector<uint32_t> shells(kmax + 2, 0xEEEEEEEEu);
        if (s == KC_OK) CHECK(kc_graph_download_order(c, dg, order.data(), shells.data()) == KC_OK);
        put(out, (uint32_t)s);
        put(out, kmax);
        put(out, st.levels);
        put(out, core.data(), g.V);
        put(out, order.data(), g.V);
        put(out, shells.data(), kmax + 2);
    }
    kc_graph_free(dg);
    kc_close(c);
    std::fclose(out);
    return 0;
}

// per run (eager, three replays, tuned): per graph u32 kmax, core[V_g]
int run_batch(int argc, char** argv) {
    std::vector<G> gs;
    for (int i = 3; i < argc; ++i) gs.push_back(read_graph(argv[i]));
    const uint32_t Gn = (uint32_t)gs.size();
    std::vector<uint64_t> vo(Gn + 1, 0), ao(Gn + 1, 0);
    std::vector<uint32_t> off, nbr;
    for (uint32_t i = 0; i < Gn; ++i) {
        vo[i + 1] = vo[i] + gs[i].V;
        ao[i + 1] = ao[i] + gs[i].arcs;
        off.insert(off.end(), gs[i].off.begin(), gs[i].off.end());
        nbr.insert(nbr.end(), gs[i].nbr.begin(), gs[i].nbr.end());
    }
    FILE* out = std::fopen(argv[2], "wb");
    kc_ctx* c = open_ctx();
    kc_batch_csr hb{Gn, vo[Gn], ao[Gn], vo.data(), ao.data(), off.data(), nbr.data(), nullptr};
    kc_batch* db = nullptr;
    CHECK(kc_batch_upload(c, &hb, &db) == KC_OK);
    uint64_t scanned = 0;
    auto emit = [&](const kc_params& p) {
        std::vector<uint32_t> core(vo[Gn], 0xEEEEEEEEu), kmax(Gn, 0xEEEEEEEEu);
        kc_stats st{};
        CHECK(kc_batch_peel(c, db, &p, core.data(), kmax.data(), &st) == KC_OK);
        CHECK(st.bytes_alg == st.bytes_scan + st.bytes_edge && st.bytes_scan == 4 * st.scanned_entries);
        if (scanned) CHECK(st.scanned_entries == scanned);  // the same levels every time
        scanned = st.scanned_entries;
        for (uint32_t i = 0; i < Gn; ++i) {
            put(out, kmax[i]);
            put(out, core.data() + vo[i], gs[i].V);
        }
    };
    kc_params p{};
    emit(p);
    p.flags = KC_FLAG_GRAPH;
    for (int r = 0; r < 3; ++r) emit(p);
    kc_tune_cfg tc{};
    tc.reps = 1;
    kc_params best{};
    double ms = 0;
    uint32_t trials = 0;
    CHECK(kc_tune_batch(c, db, &tc, &best, &ms, &trials) == KC_OK);
    CHECK(trials == 3 && (best.block == 256 || best.block == 512 || best.block == 1024));
    emit(best);
    // a malformed batch (a local id == V_g) is refused before any peel
    std::vector<uint32_t> bad = nbr;
    if (!bad.empty()) {
        bad[0] = gs[0].V;
        kc_batch_csr hbad{Gn, vo[Gn], ao[Gn], vo.data(), ao.data(), off.data(), bad.data(), nullptr};
        kc_batch* dbad = nullptr;
        CHECK(kc_batch_upload(c, &hbad, &dbad) == KC_ERR_ARG && dbad == nullptr);
    }
    kc_batch_free(db);
    kc_close(c);
    std::fclose(out);
    FILE* sf = std::fopen((std::string(argv[2]) + ".scan").c_str(), "wb");
    std::fwrite(&scanned, 8, 1, sf);
    std::fclose(sf);
    return 0;
}

// per request in submission order: u32 status, u32 kmax, core[V]; then u32 requests served by a
// lane (large), u32 completions seen
int run_serve(int argc, char** argv) {
    const uint32_t large_arcs = (uint32_t)std::strtoul(argv[3], nullptr, 10);
    const uint64_t budget = std::strtoull(argv[4], nullptr, 10);
    std::vector<G> gs;
    for (int i = 5; i < argc; ++i) gs.push_back(read_graph(argv[i]));
    const size_t N = gs.size();
    kc_ctx* c = open_ctx();
    kc_serve_cfg cfg{};
    cfg.hbm_budget_bytes = budget;
    cfg.ring_slots = 4;
    cfg.block = 256;
    cfg.workgroups = 2;
    cfg.large_arcs = large_arcs;
    cfg.large_lanes = 2;
    kc_server* srv = nullptr;
    CHECK(kc_serve_open(c, &cfg, &srv) == KC_OK);
    // outputs: even requests pinned (the kernel writes them), odd pageable (copied down)
    std::vector<uint32_t*> outp(N, nullptr);
    std::vector<std::vector<uint32_t>> pageable(N);
    std::vector<kc_completion> done(N);
    std::vector<int> seen(N, 0);
    std::vector<std::vector<uint32_t>> result(N);
    for (size_t i = 0; i < N; ++i) {
        if (i % 2 == 0) {
            void* b = nullptr;
            CHECK(kc_host_alloc(4ull * gs[i].V, &b) == KC_OK);
            outp[i] = static_cast<uint32_t*>(b);
        } else {
            pageable[i].assign(gs[i].V, 0xEEEEEEEEu);
            outp[i] = pageable[i].data();
        }
    }
    // a malformed request: neighbour id == V (answered with KC_ERR_ARG, never peeled)
    G bad = gs[0];
    if (!bad.nbr.empty()) bad.nbr[bad.nbr.size() / 2] = bad.V;
    std::vector<uint32_t> bad_out(bad.V);
    // a request larger than the whole arena is refused at submission
    {
        const uint32_t bigV = (uint32_t)std::min<uint64_t>(budget / 16 + 1024, (1u << 30) - 1024);
        std::vector<uint32_t> big_off((size_t)bigV + 1, 0);
        kc_csr hb{bigV, 0, big_off.data(), nullptr, nullptr};
        CHECK(kc_serve_submit(srv, 999999, &hb, nullptr) == KC_ERR_CAPACITY);
    }
    size_t next = 0, completed = 0;
    bool bad_sent = false, bad_done = false, side_done = false;
    std::vector<kc_completion> comp(16);
    while (completed < N || !bad_done) {
        // submissions interleave with polls, so requests join a running kernel
        for (int k = 0; k < 3 && next < N; ++k, ++next) {
            kc_csr h{gs[next].V, gs[next].arcs, gs[next].off.data(), gs[next].nbr.data(), nullptr};
            CHECK(kc_serve_submit(srv, next, &h, outp[next]) == KC_OK);
            if (next == N / 2 && !bad_sent) {
                kc_csr hb{bad.V, bad.arcs, bad.off.data(), bad.nbr.data(), nullptr};
                CHECK(kc_serve_submit(srv, 777777, &hb, bad_out.data()) == KC_OK);
                bad_sent = true;
            }
        }
        if (next >= N && !bad_sent) {
            kc_csr hb{bad.V, bad.arcs, bad.off.data(), bad.nbr.data(), nullptr};
            CHECK(kc_serve_submit(srv, 777777, &hb, bad_out.data()) == KC_OK);
            bad_sent = true;
        }
        uint32_t n = 0;
        const kc_status ps = kc_serve_poll(srv, comp.data(), (uint32_t)comp.size(), &n);
        CHECK(ps == KC_OK);
        if (n == 0) std::this_thread::sleep_for(std::chrono::microseconds(200));
        for (uint32_t j = 0; j < n; ++j) {
            const kc_completion& d = comp[j];
            if (d.req_id == 777777) {
                CHECK(d.status == KC_ERR_ARG);
                bad_done = true;
                continue;
            }
            CHECK(d.req_id < N && !seen[d.req_id]);
            CHECK(d.t_submit <= d.t_admit && d.t_admit <= d.t_done);
            seen[d.req_id] = 1;
            done[d.req_id] = d;
            result[d.req_id].assign(outp[d.req_id], outp[d.req_id] + gs[d.req_id].V);
            // a pinned result buffer freed while the serving kernel is resident
            if (d.req_id % 2 == 0) {
                CHECK(kc_host_free(outp[d.req_id]) == KC_OK);
                CHECK(kc_host_free(outp[d.req_id]) == KC_ERR_ARG);  // a second free is refused
                outp[d.req_id] = nullptr;
            }
            ++completed;
        }
        // once, mid-stream: a whole single-graph upload, peel and free beside the resident kernel
        if (!side_done && completed >= N / 3) {
            kc_csr h{gs[0].V, gs[0].arcs, gs[0].off.data(), gs[0].nbr.data(), nullptr};
            kc_graph* dg = nullptr;
            CHECK(kc_graph_upload(c, &h, &dg) == KC_OK);
            std::vector<uint32_t> core(gs[0].V);
            uint32_t km = 0;
            CHECK(kc_peel(c, dg, nullptr, core.data(), &km, nullptr) == KC_OK);
            kc_graph_free(dg);
            side_done = true;
            FILE* sf = std::fopen((std::string(argv[2]) + ".side").c_str(), "wb");
            put(sf, km);
            put(sf, core.data(), gs[0].V);
            std::fclose(sf);
        }
    }
    uint32_t q = 9, inflight = 9;
    uint64_t used = 9;
    CHECK(kc_serve_pending(srv, &q, &inflight, &used) == KC_OK);
    CHECK(q == 0 && inflight == 0 && used == 0);
    CHECK(kc_serve_close(srv) == KC_OK);
    uint32_t lanes = 0;
    FILE* out = std::fopen(argv[2], "wb");
    for (size_t i = 0; i < N; ++i) {
        put(out, (uint32_t)done[i].status);
        put(out, done[i].kmax);
        put(out, result[i].data(), gs[i].V);
        if (gs[i].arcs >= large_arcs) ++lanes;
    }
    put(out, lanes);
    put(out, (uint32_t)completed);
    std::fclose(out);
    kc_close(c);
    return 0;
}

// u32 trials (first call), u32 trials (second call), then the chosen kc_params fields, u32 status
// of a peel with them, core[V]
int run_tune(char** argv) {
    G g = read_graph(argv[2]);
    kc_ctx* c = open_ctx();
    kc_csr h{g.V, g.arcs, g.off.data(), g.nbr.data(), nullptr};
    kc_graph* dg = nullptr;
    CHECK(kc_graph_upload(c, &h, &dg) == KC_OK);
    kc_tune_cfg tc{};
    tc.reps = 1;
    tc.cache_path = argv[3];
    kc_params best{}, again{};
    double ms = 0, ms2 = 0;
    uint32_t t1 = 0, t2 = 0;
    CHECK(kc_tune_peel(c, dg, &tc, &best, &ms, &t1) == KC_OK);
    CHECK(kc_tune_peel(c, dg, &tc, &again, &ms2, &t2) == KC_OK);
    CHECK(std::memcmp(&best, &again, sizeof(best)) == 0);
    std::vector<uint32_t> core(g.V);
    uint32_t km = 0;
    const kc_status s = kc_peel(c, dg, &best, core.data(), &km, nullptr);
    FILE* out = std::fopen(argv[4], "wb");
    put(out, t1);
    put(out, t2);
    put(out, (uint32_t)best.variant);
    put(out, best.block);
    put(out, (uint32_t)s);
    put(out, km);
    put(out, core.data(), g.V);
    std::fclose(out);
    kc_graph_free(dg);
    kc_close(c);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    const std::string mode = argv[1];
    if (mode == "peel" && argc == 4) return run_peel(argv);
    if (mode == "batch" && argc >= 4) return run_batch(argc, argv);
    if (mode == "serve" && argc >= 6) return run_serve(argc, argv);
    if (mode == "tune" && argc == 5) return run_tune(argv);
    return 2;
}
