// generators.cpp — seeded synthetic undirected graphs (Erdos-Renyi, Chung-Lu, R-MAT).
//
// Every random value is a counter-based draw rnd(key, i) = splitmix64(key + (i+1)*G): the
// output depends only on (seed, stream, index), never on the thread count or schedule, so
// a graph is reproducible from its seed on any host. Integer arithmetic decides every
// sample (Lemire multiply-high for bounded draws, fixed-point thresholds for R-MAT).
// The pair lists are turned into a CSR by build_csr, which applies exactly the loader's
// rules, so a generated graph equals the loader's result on its own 'u v' text form.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "graph.h"

namespace kcg {

static const uint64_t kGolden = 0x9E3779B97F4A7C15ull;

uint64_t splitmix64(uint64_t x) {
    x += kGolden;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

static inline uint64_t stream_key(uint64_t seed, uint64_t stream) {
    return splitmix64(splitmix64(seed) ^ (stream * 0xD1B54A32D192ED03ull));
}
static inline uint64_t rnd(uint64_t key, uint64_t i) { return splitmix64(key + (i + 1) * kGolden); }
static inline uint64_t bounded(uint64_t r, uint64_t n) {
    return (uint64_t)(((unsigned __int128)r * n) >> 64);
}
static int nthr(int n) { return n > 0 ? n : omp_get_max_threads(); }

// Fisher-Yates permutation of [0, n) driven by the seed.
static std::vector<uint32_t> permutation(uint32_t n, uint64_t key) {
    std::vector<uint32_t> p(n);
    for (uint32_t i = 0; i < n; ++i) p[i] = i;
    for (uint32_t i = n; i > 1; --i) {
        uint32_t j = (uint32_t)bounded(rnd(key, i), i);
        std::swap(p[i - 1], p[j]);
    }
    return p;
}

int gen_er(uint32_t n, uint64_t m, uint64_t seed, int nthreads, Csr& out) {
    if (n < 2) return ERR_ARG;
    const uint64_t maxpairs = (uint64_t)n * (n - 1) / 2;
    if (m > maxpairs) return ERR_ARG;
    const int nt = nthr(nthreads);
    const uint64_t key = stream_key(seed, 1);
    // candidate i = (rnd(2i) mod n, rnd(2i+1) mod n); the graph is the first m distinct
    // unordered non-loop pairs in candidate order
    uint64_t K = m + m / 64 + 1024;
    std::vector<uint32_t> S, T;
    while (true) {
        std::vector<std::pair<uint64_t, uint64_t>> c(K);  // (pair key, candidate index)
#pragma omp parallel for num_threads(nt) schedule(static)
        for (int64_t i = 0; i < (int64_t)K; ++i) {
            uint64_t u = bounded(rnd(key, 2 * (uint64_t)i), n);
            uint64_t v = bounded(rnd(key, 2 * (uint64_t)i + 1), n);
            uint64_t lo = std::min(u, v), hi = std::max(u, v);
            c[i] = {u == v ? ~0ull : (lo << 32 | hi), (uint64_t)i};
        }
        std::sort(c.begin(), c.end());
        std::vector<std::pair<uint64_t, uint64_t>> keep;  // (index, pair)
        keep.reserve(K);
        for (uint64_t i = 0; i < K; ++i) {
            if (c[i].first == ~0ull) break;
            if (i > 0 && c[i].first == c[i - 1].first) continue;
            keep.push_back({c[i].second, c[i].first});
        }
        if (keep.size() >= m) {
            std::sort(keep.begin(), keep.end());
            S.resize(m);
            T.resize(m);
            for (uint64_t i = 0; i < m; ++i) {
                S[i] = (uint32_t)(keep[i].second >> 32);
                T[i] = (uint32_t)(keep[i].second & 0xffffffffu);
            }
            break;
        }
        uint64_t missing = m - keep.size();
        K += missing + missing / 2 + 1024 + K / 8;
    }
    return build_csr(S.data(), T.data(), m, nt, out);
}

int gen_chung_lu(uint32_t n, double gamma, double wmin, double wmax, double avg, uint64_t seed,
                 int nthreads, Csr& out) {
    if (n < 2 || gamma <= 1.0 || wmin <= 0 || wmax < wmin || avg <= 0) return ERR_ARG;
    const int nt = nthr(nthreads);
    // raw power-law weights, then the scale c with mean(clip(c*r, wmin, wmax)) = avg by
    // bisection in log space (a fixed number of steps, so the result is deterministic)
    const double ex = -1.0 / (gamma - 1.0);
    std::vector<double> r(n);
    for (uint32_t i = 0; i < n; ++i) r[i] = std::pow((double)(i + 1), ex);
    auto mean_at = [&](double c) {
        double s = 0;
        for (uint32_t i = 0; i < n; ++i) s += std::min(wmax, std::max(wmin, c * r[i]));
        return s / n;
    };
    double lo = std::log(1e-6), hi = std::log(1e12);
    for (int it = 0; it < 200; ++it) {
        double mid = 0.5 * (lo + hi);
        if (mean_at(std::exp(mid)) < avg) lo = mid; else hi = mid;
    }
    const double c = std::exp(0.5 * (lo + hi));
    // fixed-point cumulative weights
    std::vector<uint64_t> cum(n);
    uint64_t tot = 0;
    for (uint32_t i = 0; i < n; ++i) {
        double w = std::min(wmax, std::max(wmin, c * r[i]));
        tot += (uint64_t)std::llround(w * 1048576.0);
        cum[i] = tot;
    }
    const uint64_t M = (uint64_t)std::llround((double)n * avg / 2.0);
    const uint64_t key = stream_key(seed, 2);
    std::vector<uint32_t> perm = permutation(n, stream_key(seed, 3));
    u32vec S(M), T(M);  // written in parallel below: no zero fill
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int64_t i = 0; i < (int64_t)M; ++i) {
        uint64_t x = bounded(rnd(key, 2 * (uint64_t)i), tot);
        uint64_t y = bounded(rnd(key, 2 * (uint64_t)i + 1), tot);
        uint32_t a = (uint32_t)(std::upper_bound(cum.begin(), cum.end(), x) - cum.begin());
        uint32_t b = (uint32_t)(std::upper_bound(cum.begin(), cum.end(), y) - cum.begin());
        S[i] = perm[a];
        T[i] = perm[b];
    }
    return build_csr(S.data(), T.data(), M, nt, out);
}

int gen_rmat(uint32_t scale, uint64_t ndraws, double a, double b, double c, uint64_t seed,
             int nthreads, Csr& out) {
    if (scale < 1 || scale > 31 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return ERR_ARG;
    const int nt = nthr(nthreads);
    const double two32 = 4294967296.0;
    const uint64_t ta = (uint64_t)(a * two32);
    const uint64_t tab = (uint64_t)((a + b) * two32);
    const uint64_t tabc = (uint64_t)((a + b + c) * two32);
    const uint64_t key = stream_key(seed, 4);
    const uint32_t nv = 1u << scale;
    std::vector<uint32_t> perm = permutation(nv, stream_key(seed, 5));
    const uint32_t words = (scale + 1) / 2;  // two 32-bit quadrant draws per 64-bit value
    u32vec S(ndraws), T(ndraws);  // written in parallel below: no zero fill
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int64_t i = 0; i < (int64_t)ndraws; ++i) {
        uint32_t u = 0, v = 0;
        uint64_t w = 0;
        for (uint32_t l = 0; l < scale; ++l) {
            if ((l & 1) == 0) w = rnd(key, (uint64_t)i * words + l / 2);
            uint64_t x = (l & 1) ? (w >> 32) : (w & 0xffffffffull);
            uint32_t bu = 0, bv = 0;
            if (x < ta) {
            } else if (x < tab) {
                bv = 1;
            } else if (x < tabc) {
                bu = 1;
            } else {
                bu = 1;
                bv = 1;
            }
            u = (u << 1) | bu;
            v = (v << 1) | bv;
        }
        S[i] = perm[u];
        T[i] = perm[v];
    }
    return build_csr(S.data(), T.data(), ndraws, nt, out);
}

}  // namespace kcg
