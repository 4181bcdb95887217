// kcore_serve — open-loop served k-core workload: Poisson arrivals for a fixed time, each request
// a graph from a seeded pool (R-MAT or Chung-Lu by coin flip, lognormal edge count), answered
// either by the server (kc_serve_*, continuous batching in the resident kernel) or by the plain
// per-request path (upload, single-graph peel, download, free, one request after another).
// Prints one JSON object: graphs/s, p50/p99/max latency (arrival to result on the host) and the
// check of every result (against --expect core numbers, else across repeats of a pool graph).
//
//   kcore_serve [--mode served|plain] [--rate R] [--seconds T] [--pool P] [--seed S]
//               [--median M] [--sigma S] [--mmin A] [--mmax B] [--budget-gb G] [--block B]
//               [--workgroups W] [--slots N] [--expect FILE] [--drain-limit S] [--closed N]
//               [--max-backlog N] [--large-arcs A] [--large-lanes K]
//
// --closed N replaces the Poisson stream by N requests that all arrive at time 0: the run then
// measures the path's capacity (graphs/s), and its latencies are queueing times.
// Pool graph j is request j of the seeded stream (kcg_gen_request). --expect FILE holds the core
// numbers of every pool graph, little-endian uint32, concatenated in pool order.
#include <algorithm>
#include <chrono>
#include <csignal>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/kcore.h"
#include "../../include/kcore_graph.h"

namespace {

struct Opts {
    std::string mode = "served";
    double rate = 2000, seconds = 10, median = 50000, sigma = 1.0;
    uint64_t mmin = 1000, mmax = 50000000;
    uint32_t pool = 256;
    uint64_t seed = 1;
    double budget_gb = 64;
    uint32_t block = 0, workgroups = 0, slots = 0;
    std::string expect;
    double drain_limit = 120;
    uint64_t closed = 0;
    uint64_t max_backlog = 20000;
    uint32_t large_arcs = 0, large_lanes = 0;
};

struct PoolGraph {
    uint32_t V = 0;
    uint64_t arcs = 0;
    uint32_t* off = nullptr;  // pinned: offsets, neighbours, degrees back to back
    uint32_t* nbr = nullptr;
    uint32_t* deg = nullptr;
    std::vector<uint32_t> expect;
    std::vector<uint32_t*> free_out;  // pinned output buffers of V words
    bool have_ref = false;
    int kind = 0;
};

uint64_t splitmix(uint64_t x) { return kcg_splitmix64(x); }
double unif(uint64_t key, uint64_t i) {
    return ((double)(splitmix(key + (i + 1) * 0x9E3779B97F4A7C15ull) >> 11) + 1.0) * (1.0 / 9007199254740992.0);
}

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

kc_server* g_srv = nullptr;  // closed on every exit path, so the serving kernel never outlives us
volatile std::sig_atomic_t g_quit = 0;
void on_signal(int) { g_quit = 1; }

[[noreturn]] void die(const char* what, int st) {
    std::fprintf(stderr, "kcore_serve: %s failed (%d)\n", what, st);
    if (g_srv) (void)kc_serve_close(g_srv);
    g_srv = nullptr;
    std::exit(1);
}

double pct(std::vector<double> v, double p) {
    if (v.empty()) return 0;
    std::sort(v.begin(), v.end());
    const size_t i = (size_t)std::min<double>(v.size() - 1, std::ceil(p * v.size()) - 1);
    return v[i];
}

}  // namespace

int main(int argc, char** argv) {
    Opts o;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto val = [&]() -> const char* {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "missing value for %s\n", a.c_str());
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "--mode") o.mode = val();
        else if (a == "--rate") o.rate = std::atof(val());
        else if (a == "--seconds") o.seconds = std::atof(val());
        else if (a == "--pool") o.pool = (uint32_t)std::atoi(val());
        else if (a == "--seed") o.seed = std::strtoull(val(), nullptr, 10);
        else if (a == "--median") o.median = std::atof(val());
        else if (a == "--sigma") o.sigma = std::atof(val());
        else if (a == "--mmin") o.mmin = std::strtoull(val(), nullptr, 10);
        else if (a == "--mmax") o.mmax = std::strtoull(val(), nullptr, 10);
        else if (a == "--budget-gb") o.budget_gb = std::atof(val());
        else if (a == "--block") o.block = (uint32_t)std::atoi(val());
        else if (a == "--workgroups") o.workgroups = (uint32_t)std::atoi(val());
        else if (a == "--slots") o.slots = (uint32_t)std::atoi(val());
        else if (a == "--expect") o.expect = val();
        else if (a == "--drain-limit") o.drain_limit = std::atof(val());
        else if (a == "--closed") o.closed = std::strtoull(val(), nullptr, 10);
        else if (a == "--max-backlog") o.max_backlog = std::strtoull(val(), nullptr, 10);
        else if (a == "--large-arcs") o.large_arcs = (uint32_t)std::strtoul(val(), nullptr, 10);
        else if (a == "--large-lanes") o.large_lanes = (uint32_t)std::atoi(val());
        else {
            std::fprintf(stderr, "unknown option %s\n", a.c_str());
            return 2;
        }
    }
    if (o.mode != "served" && o.mode != "plain") {
        std::fprintf(stderr, "--mode is served or plain\n");
        return 2;
    }
    if (o.pool == 0 || o.rate <= 0 || o.seconds <= 0) {
        std::fprintf(stderr, "--pool, --rate and --seconds must be positive\n");
        return 2;
    }

    // ---- the request pool (graph j of the pool is request j of the seeded stream) ----
    std::vector<PoolGraph> pool(o.pool);
    uint64_t pool_arcs = 0;
    const double tg0 = now_s();
    for (uint32_t j = 0; j < o.pool; ++j) {
        kcg_graph* g = nullptr;
        int kind = 0;
        int st = kcg_gen_request(o.seed, j, o.median, o.sigma, o.mmin, o.mmax, 0, &g, &kind);
        if (st) die("kcg_gen_request", st);
        uint32_t V, dmax;
        uint64_t arcs;
        kcg_graph_info(g, &V, &arcs, &dmax);
        const uint32_t *off, *nbr, *deg;
        kcg_graph_arrays(g, &off, &nbr, &deg);
        PoolGraph& p = pool[j];
        p.V = V;
        p.arcs = arcs;
        p.kind = kind;
        void* buf = nullptr;
        st = kc_host_alloc(((uint64_t)V + 1 + arcs + V) * 4, &buf);
        if (st) die("kc_host_alloc", st);
        p.off = (uint32_t*)buf;
        p.nbr = p.off + V + 1;
        p.deg = p.nbr + arcs;
        std::memcpy(p.off, off, ((size_t)V + 1) * 4);
        if (arcs) std::memcpy(p.nbr, nbr, arcs * 4);
        std::memcpy(p.deg, deg, (size_t)V * 4);
        kcg_graph_free(g);
        pool_arcs += arcs;
    }
    const double gen_s = now_s() - tg0;

    if (!o.expect.empty()) {
        FILE* f = std::fopen(o.expect.c_str(), "rb");
        if (!f) die("open --expect", 1);
        for (auto& p : pool) {
            p.expect.resize(p.V);
            if (std::fread(p.expect.data(), 4, p.V, f) != p.V) die("read --expect", 1);
            p.have_ref = true;
        }
        std::fclose(f);
    }

    // ---- arrivals: Poisson at `rate` for `seconds`, each picking a pool graph uniformly ----
    struct Arrival {
        double t;
        uint32_t g;
    };
    std::vector<Arrival> arr;
    {
        const uint64_t key = splitmix(o.seed ^ 0xA11CE5ull);
        double t = 0;
        for (uint64_t k = 0;; ++k) {
            if (o.closed) {
                if (k == o.closed) break;
            } else {
                t += -std::log(unif(key, 2 * k)) / o.rate;
                if (t >= o.seconds) break;
            }
            const uint32_t g = (uint32_t)std::min<double>(o.pool - 1, unif(key, 2 * k + 1) * o.pool);
            arr.push_back({t, g});
        }
    }
    const uint64_t N = arr.size();

    std::signal(SIGINT, on_signal);
    std::signal(SIGTERM, on_signal);
    kc_ctx* ctx = nullptr;
    int st = kc_open(0, &ctx);
    if (st) die("kc_open", st);

    std::vector<double> lat(N, -1.0);
    std::vector<uint8_t> ok(N, 0);
    uint64_t mismatches = 0, done_n = 0, done_arcs = 0;
    auto check = [&](uint64_t k, const uint32_t* core) {
        PoolGraph& p = pool[arr[k].g];
        if (!p.have_ref) {  // first result of a pool graph becomes the reference for its repeats
            p.expect.assign(core, core + p.V);
            p.have_ref = true;
        }
        const bool same = std::memcmp(core, p.expect.data(), (size_t)p.V * 4) == 0;
        ok[k] = same;
        if (!same) ++mismatches;
    };
    auto take_out = [&](PoolGraph& p) {
        if (!p.free_out.empty()) {
            uint32_t* b = p.free_out.back();
            p.free_out.pop_back();
            return b;
        }
        void* b = nullptr;
        int s2 = kc_host_alloc((uint64_t)std::max<uint32_t>(p.V, 1) * 4, &b);
        if (s2) die("kc_host_alloc", s2);
        return (uint32_t*)b;
    };

    double t_start = 0, t_end = 0;
    uint64_t unfinished = 0, dropped = 0;
    bool overloaded = false, interrupted = false;
    int device_error = 0;
    if (o.mode == "served") {
        kc_serve_cfg cfg{};
        cfg.hbm_budget_bytes = (uint64_t)(o.budget_gb * (1ull << 30));
        cfg.block = o.block;
        cfg.workgroups = o.workgroups;
        cfg.ring_slots = o.slots;
        cfg.large_arcs = o.large_arcs;
        cfg.large_lanes = o.large_lanes;
        st = kc_serve_open(ctx, &cfg, &g_srv);
        if (st) die("kc_serve_open", st);
        std::vector<uint32_t*> outbuf(N, nullptr);
        std::vector<kc_completion> comp(4096);
        uint64_t next = 0;
        t_start = now_s();
        const double deadline = t_start + o.seconds + o.drain_limit;
        while (done_n + dropped < N && !g_quit) {
            const double now = now_s() - t_start;
            while (next < N && arr[next].t <= now) {
                // an open-loop stream above capacity grows without bound: past max_backlog
                // outstanding requests the rest of the stream is dropped and the run is flagged
                if (overloaded || next - done_n >= o.max_backlog) {
                    overloaded = true;
                    ++dropped;
                    ++next;
                    continue;
                }
                PoolGraph& p = pool[arr[next].g];
                outbuf[next] = take_out(p);
                kc_csr c{p.V, p.arcs, p.off, p.nbr, nullptr};
                st = kc_serve_submit(g_srv, next, &c, outbuf[next]);
                if (st) die("kc_serve_submit", st);
                ++next;
            }
            uint32_t n = 0;
            st = kc_serve_poll(g_srv, comp.data(), (uint32_t)comp.size(), &n);
            if (st) {
                device_error = st;
                break;
            }
            for (uint32_t i = 0; i < n; ++i) {
                const uint64_t k = comp[i].req_id;
                if (comp[i].status) die("served request", comp[i].status);
                lat[k] = comp[i].t_done - (t_start + arr[k].t);
                check(k, outbuf[k]);
                pool[arr[k].g].free_out.push_back(outbuf[k]);
                done_arcs += pool[arr[k].g].arcs;
                ++done_n;
            }
            if (now_s() > deadline) break;
        }
        interrupted = g_quit != 0;
        t_end = now_s();
        unfinished = N - done_n - dropped;
        st = kc_serve_close(g_srv);
        g_srv = nullptr;
        if (st && !device_error) device_error = st;
    } else {
        kc_params prm{};
        t_start = now_s();
        const double deadline = t_start + o.seconds + o.drain_limit;
        for (uint64_t k = 0; k < N && !g_quit; ++k) {
            double now = now_s() - t_start;
            while (now < arr[k].t && !g_quit) now = now_s() - t_start;  // idle until the next arrival
            PoolGraph& p = pool[arr[k].g];
            uint32_t* out = take_out(p);
            kc_csr c{p.V, p.arcs, p.off, p.nbr, p.deg};
            kc_graph* dg = nullptr;
            st = kc_graph_upload(ctx, &c, &dg);
            if (st) die("kc_graph_upload", st);
            uint32_t kmax = 0;
            st = kc_peel(ctx, dg, &prm, out, &kmax, nullptr);
            if (st) die("kc_peel", st);
            kc_graph_free(dg);
            lat[k] = now_s() - (t_start + arr[k].t);
            check(k, out);
            p.free_out.push_back(out);
            done_arcs += p.arcs;
            ++done_n;
            if (now_s() > deadline) break;
        }
        interrupted = g_quit != 0;
        t_end = now_s();
        unfinished = N - done_n;
    }

    std::vector<double> L;
    for (uint64_t k = 0; k < N; ++k)
        if (lat[k] >= 0) L.push_back(lat[k]);
    // throughput over the whole run: arrivals span `seconds`, the run ends when the last finishes
    const double wall = t_end - t_start;
    uint64_t pool_kind1 = 0;
    for (auto& p : pool) pool_kind1 += p.kind;
    std::printf(
        "{\"mode\": \"%s\", \"offered_rate\": %.1f, \"seconds\": %.1f, \"closed\": %llu, \"requests\": %llu, "
        "\"completed\": %llu, "
        "\"unfinished\": %llu, \"wall_s\": %.3f, \"graphs_per_s\": %.1f, \"edges_per_s\": %.4g, "
        "\"p50_ms\": %.3f, \"p90_ms\": %.3f, \"p99_ms\": %.3f, \"max_ms\": %.3f, \"mismatches\": %llu, "
        "\"checked_against\": \"%s\", \"pool\": %u, \"pool_edges\": %llu, \"pool_chung_lu\": %llu, "
        "\"pool_gen_s\": %.2f, \"median_m\": %.0f, \"sigma\": %.2f, \"dropped\": %llu, \"overloaded\": %s, "
        "\"interrupted\": %s, \"device_error\": %d}\n",
        o.mode.c_str(), o.closed ? 0.0 : o.rate, o.closed ? 0.0 : o.seconds, (unsigned long long)o.closed,
        (unsigned long long)N, (unsigned long long)done_n,
        (unsigned long long)unfinished, wall, done_n / wall, done_arcs / 2.0 / wall, pct(L, 0.5) * 1e3,
        pct(L, 0.9) * 1e3, pct(L, 0.99) * 1e3, pct(L, 1.0) * 1e3, (unsigned long long)mismatches,
        o.expect.empty() ? "first result of each pool graph" : "expected core numbers", o.pool,
        (unsigned long long)(pool_arcs / 2), (unsigned long long)pool_kind1, gen_s, o.median, o.sigma,
        (unsigned long long)dropped, overloaded ? "true" : "false", interrupted ? "true" : "false", device_error);
    kc_close(ctx);
    return mismatches || unfinished || device_error ? 3 : (overloaded ? 4 : 0);
}
