// tune.cpp — the autotuner: picks kc_params for a resident graph or batch by timing real peels.
//
// Single graph: coordinate descent from the defaults over variant, workgroup size, the degree
// above which a vertex's arcs are split into chunks shared by many waves (the wave/block
// threshold), chunk length, then the LDS frontier depth (sub-round variants) or the alive-list
// compaction ratio and idle poll step (cascade-queue variant); each axis keeps its fastest value
// (median of `reps` timed peels, device time) before the next axis is tried. Batch: workgroup size.
// Every candidate must produce identical core numbers; the tuner enforces it: a candidate is
// adopted only if the core numbers of its last timed peel equal those of the default settings.
//
// Results can be kept in a text cache, one line per key:
//   peel <V> <arcs> <dmax> <flags> <arch> <cus> : <variant> <block> <heavy> <chunk> <lds> <compact> <poll> <ms>
//   batch <G> <vertices> <arcs> <flags> <arch> <cus> : <block> <ms>
// A cached key is answered without timing anything.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "host_internal.h"

namespace {

std::string cache_lookup(const char* path, const std::string& key) {
    if (!path) return {};
    FILE* f = std::fopen(path, "r");
    if (!f) return {};
    std::string found;
    char line[512];
    while (std::fgets(line, sizeof(line), f)) {
        std::string l(line);
        const size_t c = l.find(" : ");
        if (c != std::string::npos && l.compare(0, c, key) == 0 && c == key.size()) found = l.substr(c + 3);
    }
    std::fclose(f);
    return found;  // the last entry for a key wins
}

void cache_store(const char* path, const std::string& key, const std::string& val) {
    if (!path) return;
    FILE* f = std::fopen(path, "a");
    if (!f) return;
    std::fprintf(f, "%s : %s\n", key.c_str(), val.c_str());
    std::fclose(f);
}

template <class Run>
kc_status time_median(Run run, uint32_t reps, double* ms) {
    std::vector<double> t;
    for (uint32_t r = 0; r < reps; ++r) {
        double x = 0;
        const kc_status s = run(&x);
        if (s != KC_OK) return s;
        t.push_back(x);
    }
    std::sort(t.begin(), t.end());
    *ms = t[t.size() / 2];
    return KC_OK;
}

}  // namespace

extern "C" {

kc_status kc_tune_peel(kc_ctx* c, kc_graph* g, const kc_tune_cfg* cfg, kc_params* best, double* best_ms,
                       uint32_t* trials) {
    if (!c || !g || !best) return KC_ERR_ARG;
    kc_tune_cfg tc{};
    if (cfg) tc = *cfg;
    const uint32_t reps = tc.reps ? tc.reps : 2;
    uint32_t V, dmax;
    uint64_t arcs;
    kc_graph_info(g, &V, &arcs, &dmax);
    char key[256];
    std::snprintf(key, sizeof(key), "peel %u %llu %u %u %s %d", V, (unsigned long long)arcs, dmax, tc.flags,
                  c->arch, c->cus);
    const std::string hit = cache_lookup(tc.cache_path, key);
    if (!hit.empty()) {
        kc_params p{};
        double ms = 0;
        if (std::sscanf(hit.c_str(), "%d %u %u %u %u %u %u %lf", &p.variant, &p.block, &p.heavy_threshold,
                        &p.chunk_arcs, &p.lds_depth, &p.compact_ratio, &p.poll_max, &ms) == 8) {
            p.flags = tc.flags;
            *best = p;
            if (best_ms) *best_ms = ms;
            if (trials) *trials = 0;
            return KC_OK;
        }
    }
    kc_params cur{};
    cur.variant = KC_VARIANT_LDS;
    cur.block = 512;
    cur.heavy_threshold = 64;
    cur.chunk_arcs = 512;
    cur.lds_depth = 0;
    cur.compact_ratio = 4;
    cur.poll_max = 2;
    cur.flags = tc.flags;
    uint32_t n = 0;
    auto measure = [&](const kc_params& p, double* ms) {
        ++n;
        return time_median(
            [&](double* x) {
                kc_stats st{};
                const kc_status s = kc_peel(c, g, &p, nullptr, nullptr, &st);
                *x = st.ms_peel;
                return s;
            },
            reps, ms);
    };
    double cur_ms = 0;
    kc_status s = measure(cur, &cur_ms);
    if (s != KC_OK) return s;
    std::vector<uint32_t> ref(V), got(V);
    if ((s = kc_graph_download_core(c, g, ref.data())) != KC_OK) return s;
    auto same = [&]() {
        return kc_graph_download_core(c, g, got.data()) == KC_OK && got == ref;
    };
    struct Axis {
        uint32_t kc_params::*field;
        std::vector<uint32_t> values;
    };
    // A candidate that fails with a device-side status (a wait past its time limit, the
    // no-progress guard) is skipped; a HIP runtime failure ends the tuning.
    auto fatal = [](kc_status st) { return st != KC_OK && st != KC_ERR_DEVICE && st != KC_ERR_NO_PROGRESS; };
    // the variant axis first (an int field), then the unsigned axes
    for (int v : {KC_VARIANT_WAVE, KC_VARIANT_PREFETCH, KC_VARIANT_ASYNC}) {
        kc_params p = cur;
        p.variant = v;
        double ms = 0;
        if ((s = measure(p, &ms)) != KC_OK) {
            if (fatal(s)) return s;
            continue;
        }
        if (ms < cur_ms && same()) cur = p, cur_ms = ms;
    }
    std::vector<Axis> axes = {
        {&kc_params::block, {256, 512, 1024}},
        {&kc_params::heavy_threshold, {32, 64, 128, 256, 1024}},
        {&kc_params::chunk_arcs, {256, 512, 1024, 2048}},
    };
    if (cur.variant == KC_VARIANT_ASYNC) {
        axes.push_back({&kc_params::compact_ratio, {2, 4, 8, 16}});
        axes.push_back({&kc_params::poll_max, {1, 2, 3}});
    } else {
        axes.push_back({&kc_params::lds_depth, {0, 512, 256}});
    }
    for (const Axis& ax : axes) {
        for (uint32_t v : ax.values) {
            if (cur.*ax.field == v) continue;
            kc_params p = cur;
            p.*ax.field = v;
            double ms = 0;
            if ((s = measure(p, &ms)) != KC_OK) {
                if (fatal(s)) return s;
                continue;
            }
            if (ms < cur_ms && same()) cur = p, cur_ms = ms;
        }
    }
    *best = cur;
    if (best_ms) *best_ms = cur_ms;
    if (trials) *trials = n;
    char val[160];
    std::snprintf(val, sizeof(val), "%d %u %u %u %u %u %u %.4f", cur.variant, cur.block, cur.heavy_threshold,
                  cur.chunk_arcs, cur.lds_depth, cur.compact_ratio, cur.poll_max, cur_ms);
    cache_store(tc.cache_path, key, val);
    return KC_OK;
}

kc_status kc_tune_batch(kc_ctx* c, kc_batch* b, const kc_tune_cfg* cfg, kc_params* best, double* best_ms,
                        uint32_t* trials) {
    if (!c || !b || !best) return KC_ERR_ARG;
    kc_tune_cfg tc{};
    if (cfg) tc = *cfg;
    const uint32_t reps = tc.reps ? tc.reps : 2;
    uint32_t G;
    uint64_t nv, na;
    kc_batch_info(b, &G, &nv, &na);
    char key[256];
    std::snprintf(key, sizeof(key), "batch %u %llu %llu %u %s %d", G, (unsigned long long)nv,
                  (unsigned long long)na, tc.flags, c->arch, c->cus);
    const std::string hit = cache_lookup(tc.cache_path, key);
    if (!hit.empty()) {
        kc_params p{};
        double ms = 0;
        if (std::sscanf(hit.c_str(), "%u %lf", &p.block, &ms) == 2) {
            p.flags = tc.flags;
            *best = p;
            if (best_ms) *best_ms = ms;
            if (trials) *trials = 0;
            return KC_OK;
        }
    }
    kc_params cur{};
    double cur_ms = 1e300;
    uint32_t n = 0;
    for (uint32_t blk : {256u, 512u, 1024u}) {
        kc_params p{};
        p.block = blk;
        p.flags = tc.flags;
        double ms = 0;
        ++n;
        const kc_status s = time_median(
            [&](double* x) {
                kc_stats st{};
                const kc_status r = kc_batch_peel(c, b, &p, nullptr, nullptr, &st);
                *x = st.ms_peel;
                return r;
            },
            reps, &ms);
        if (s != KC_OK) return s;
        if (ms < cur_ms) cur = p, cur_ms = ms;
    }
    *best = cur;
    if (best_ms) *best_ms = cur_ms;
    if (trials) *trials = n;
    char val[64];
    std::snprintf(val, sizeof(val), "%u %.4f", cur.block, cur_ms);
    cache_store(tc.cache_path, key, val);
    return KC_OK;
}

}  // extern "C"
