// graph.cpp — edge list -> CSR loader, serialized cache and core writers with the reference's
// semantics (src/graph.cpp:9-161), built for large inputs: the text is parsed in parallel
// chunks and the CSR is built by a counting scatter plus per-vertex sort instead of one
// std::set per vertex. The output arrays are identical to the reference loader's.
#include "graph.h"

#include <omp.h>

#include <algorithm>
#include <cctype>
#include <charconv>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>

namespace kcg {

static int threads_or_default(int nthreads) {
    return nthreads > 0 ? nthreads : omp_get_max_threads();
}

std::string basename_of(const std::string& path) {
    // reference src/graph.cpp:10: substr(find_last_of("/") + 1)
    size_t p = path.find_last_of('/');
    return p == std::string::npos ? path : path.substr(p + 1);
}

int build_csr(const uint32_t* s, const uint32_t* t, uint64_t np, int nthreads, Csr& out) {
    const int nt = threads_or_default(nthreads);
    // V = 1 + max id over non-loop pairs (reference src/graph.cpp:89-96)
    uint32_t vmax = 0;
    bool any = false;
#pragma omp parallel for num_threads(nt) reduction(max : vmax) reduction(|| : any) schedule(static)
    for (int64_t i = 0; i < (int64_t)np; ++i) {
        if (s[i] == t[i]) continue;
        any = true;
        vmax = std::max(vmax, std::max(s[i], t[i]));
    }
    if (any && vmax == UINT32_MAX) return ERR_RANGE;  // V would wrap to 0
    const uint32_t V = any ? vmax + 1 : 1;

    // counting scatter of both directions
    std::vector<uint64_t> start((size_t)V + 1, 0);
    {
        std::vector<uint32_t> cnt(V, 0);
        uint32_t* c = cnt.data();
#pragma omp parallel for num_threads(nt) schedule(static)
        for (int64_t i = 0; i < (int64_t)np; ++i) {
            if (s[i] == t[i]) continue;
            __atomic_fetch_add(&c[s[i]], 1u, __ATOMIC_RELAXED);
            __atomic_fetch_add(&c[t[i]], 1u, __ATOMIC_RELAXED);
        }
        for (uint32_t v = 0; v < V; ++v) start[v + 1] = start[v] + cnt[v];
    }
    const uint64_t total = start[V];
    std::vector<uint32_t> buf(total);
    {
        std::vector<uint64_t> cur(start.begin(), start.end() - 1);
        uint64_t* cp = cur.data();
        uint32_t* bp = buf.data();
#pragma omp parallel for num_threads(nt) schedule(static)
        for (int64_t i = 0; i < (int64_t)np; ++i) {
            uint32_t a = s[i], b = t[i];
            if (a == b) continue;
            bp[__atomic_fetch_add(&cp[a], 1ull, __ATOMIC_RELAXED)] = b;
            bp[__atomic_fetch_add(&cp[b], 1ull, __ATOMIC_RELAXED)] = a;
        }
    }
    // per-vertex sort + unique (the reference's std::set gives ascending unique lists)
    out.V = V;
    out.degrees.assign(V, 0);
    uint32_t* deg = out.degrees.data();
#pragma omp parallel for num_threads(nt) schedule(dynamic, 1024)
    for (int64_t v = 0; v < (int64_t)V; ++v) {
        uint32_t* b = buf.data() + start[v];
        uint32_t* e = buf.data() + start[v + 1];
        std::sort(b, e);
        deg[v] = (uint32_t)(std::unique(b, e) - b);
    }
    uint64_t arcs = 0;
    out.offsets.resize((size_t)V + 1);
    out.offsets[0] = 0;
    uint32_t dmax = 0;
    for (uint32_t v = 0; v < V; ++v) {
        arcs += deg[v];
        if (arcs > UINT32_MAX) return ERR_RANGE;  // reference offsets are uint32 (inc/graph.h:11-13)
        out.offsets[v + 1] = (uint32_t)arcs;
        dmax = std::max(dmax, deg[v]);
    }
    out.arcs = arcs;
    out.dmax = dmax;
    out.neighbors.resize(arcs);
    uint32_t* nb = out.neighbors.data();
    const uint32_t* off = out.offsets.data();
#pragma omp parallel for num_threads(nt) schedule(dynamic, 4096)
    for (int64_t v = 0; v < (int64_t)V; ++v)
        if (deg[v]) std::memcpy(nb + off[v], buf.data() + start[v], (size_t)deg[v] * 4);
    return OK;
}

// ---------------------------------------------------------------------------------------
// Text parsing. A line is handled by a fast path when its shape is `[ \t]*D+[ \t]+D+...`
// with at most 9 digits per number; any other line makes the whole file go through the
// exact sequential emulation of `std::istringstream iss(line); iss >> s >> t;` that the
// reference performs (src/graph.cpp:84-95), since such lines carry state between lines.

static inline bool is_ws(char c) {
    return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r';
}

// returns 1 = pair parsed, 0 = comment line (skipped), -1 = needs the exact path
static inline int fast_line(const char* p, const char* e, uint32_t& s, uint32_t& t) {
    if (p == e) return -1;  // empty line: extraction fails, previous values kept
    if (std::isalpha((unsigned char)*p)) return 0;
    while (p < e && (*p == ' ' || *p == '\t')) ++p;
    uint32_t a = 0;
    int nd = 0;
    while (p < e && *p >= '0' && *p <= '9') {
        a = a * 10 + (uint32_t)(*p - '0');
        ++p;
        if (++nd > 9) return -1;
    }
    if (nd == 0 || p == e || !(*p == ' ' || *p == '\t')) return -1;
    while (p < e && (*p == ' ' || *p == '\t')) ++p;
    uint32_t b = 0;
    nd = 0;
    while (p < e && *p >= '0' && *p <= '9') {
        b = b * 10 + (uint32_t)(*p - '0');
        ++p;
        if (++nd > 9) return -1;
    }
    if (nd == 0) return -1;
    // whatever follows the second number is ignored by `iss >> t`
    s = a;
    t = b;
    return 1;
}

static int read_whole(const std::string& path, std::string& data) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return ERR_IO;
    std::fseek(f, 0, SEEK_END);
    long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    data.resize(sz > 0 ? (size_t)sz : 0);
    size_t got = sz > 0 ? std::fread(&data[0], 1, (size_t)sz, f) : 0;
    std::fclose(f);
    if (got != data.size()) return ERR_IO;
    return OK;
}

int parse_edgelist(const std::string& path, int nthreads, std::vector<uint32_t>& S,
                   std::vector<uint32_t>& T) {
    std::string data;
    int st = read_whole(path, data);
    if (st) return st;
    const int nt = threads_or_default(nthreads);
    const char* base = data.data();
    const size_t n = data.size();
    // chunk boundaries at line starts
    std::vector<size_t> cut(nt + 1, n);
    cut[0] = 0;
    for (int k = 1; k < nt; ++k) {
        size_t p = std::max(cut[k - 1], n * (size_t)k / nt);
        while (p < n && p > 0 && base[p - 1] != '\n') ++p;
        cut[k] = p;
    }
    std::vector<std::vector<uint32_t>> ls(nt), lt(nt);
    int slow = 0;
#pragma omp parallel num_threads(nt) reduction(| : slow)
    {
        int k = omp_get_thread_num();
        const char* p = base + cut[k];
        const char* end = base + cut[k + 1];
        auto& vs = ls[k];
        auto& vt = lt[k];
        vs.reserve((end - p) / 12 + 16);
        vt.reserve((end - p) / 12 + 16);
        while (p < end && !slow) {
            const char* nl = (const char*)std::memchr(p, '\n', (size_t)(end - p));
            const char* le = nl ? nl : end;
            uint32_t a, b;
            int r = fast_line(p, le, a, b);
            if (r < 0) {
                slow = 1;
                break;
            }
            if (r == 1 && a != b) {
                vs.push_back(a);
                vt.push_back(b);
            }
            p = nl ? nl + 1 : end;
        }
    }
    S.clear();
    T.clear();
    if (!slow) {
        size_t tot = 0;
        for (int k = 0; k < nt; ++k) tot += ls[k].size();
        S.reserve(tot);
        T.reserve(tot);
        for (int k = 0; k < nt; ++k) {
            S.insert(S.end(), ls[k].begin(), ls[k].end());
            T.insert(T.end(), lt[k].begin(), lt[k].end());
        }
        return OK;
    }
    // exact sequential emulation, line for line. The reference leaves s and t
    // uninitialised before the first line; they start at 0 here.
    unsigned int s = 0, t = 0;
    size_t p = 0;
    std::string line;
    while (p < n) {
        size_t q = data.find('\n', p);
        size_t le = q == std::string::npos ? n : q;
        line.assign(base + p, le - p);
        p = q == std::string::npos ? n : q + 1;
        if (std::isalpha(line[0])) continue;
        std::istringstream iss(line);
        iss >> s >> t;
        if (s == t) continue;
        S.push_back(s);
        T.push_back(t);
    }
    return OK;
}

int load_edgelist(const std::string& path, int nthreads, Csr& out) {
    std::vector<uint32_t> s, t;
    int st = parse_edgelist(path, nthreads, s, t);
    if (st) return st;
    return build_csr(s.data(), t.data(), s.size(), nthreads, out);
}

// ---------------------------------------------------------------------------------------
// Serialized cache (reference src/graph.cpp:35-54 writer, :9-33 reader):
//   "<V>\n<E>\n" + "<deg>\n" * V + "<off> " * (V+1) + "<nbr> " * E

static void put_u32(std::string& buf, uint32_t v, char sep) {
    char tmp[16];
    auto r = std::to_chars(tmp, tmp + 16, v);
    buf.append(tmp, r.ptr);
    buf.push_back(sep);
}

int write_serialized(const Csr& g, const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return ERR_IO;
    std::string buf;
    buf.reserve(1 << 22);
    bool short_write = false;
    auto flush = [&]() {
        short_write |= std::fwrite(buf.data(), 1, buf.size(), f) != buf.size();
        buf.clear();
    };
    put_u32(buf, g.V, '\n');
    put_u32(buf, (uint32_t)g.arcs, '\n');
    for (uint32_t i = 0; i < g.V; ++i) {
        put_u32(buf, g.degrees[i], '\n');
        if (buf.size() > (1 << 22) - 64) flush();
    }
    for (uint64_t i = 0; i <= g.V; ++i) {
        put_u32(buf, g.offsets[i], ' ');
        if (buf.size() > (1 << 22) - 64) flush();
    }
    for (uint64_t i = 0; i < g.arcs; ++i) {
        put_u32(buf, g.neighbors[i], ' ');
        if (buf.size() > (1 << 22) - 64) flush();
    }
    flush();
    return (std::fclose(f) == 0 && !short_write) ? OK : ERR_IO;
}

// One core number per line, each followed by '\n': the form of the degree block of the
// serialized cache (src/graph.cpp:35-54), so the same tools read both.
int write_core_text(const std::string& path, uint32_t V, const uint32_t* core) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return ERR_IO;
    std::string buf;
    buf.reserve(1 << 22);
    bool short_write = false;
    for (uint32_t i = 0; i < V; ++i) {
        put_u32(buf, core[i], '\n');
        if (buf.size() > (1 << 22) - 64) {
            short_write |= std::fwrite(buf.data(), 1, buf.size(), f) != buf.size();
            buf.clear();
        }
    }
    short_write |= std::fwrite(buf.data(), 1, buf.size(), f) != buf.size();
    return (std::fclose(f) == 0 && !short_write) ? OK : ERR_IO;
}

namespace {
struct TokenReader {
    const char* p;
    const char* e;
    bool next(uint32_t& v) {
        while (p < e && is_ws(*p)) ++p;
        if (p >= e) return false;
        uint64_t a = 0;
        const char* q = p;
        while (p < e && *p >= '0' && *p <= '9') a = a * 10 + (uint64_t)(*p++ - '0');
        if (p == q || a > UINT32_MAX) return false;
        v = (uint32_t)a;
        return true;
    }
};
}  // namespace

int read_serialized(const std::string& path, Csr& g) {
    std::string data;
    int st = read_whole(path, data);
    if (st) return st;
    TokenReader r{data.data(), data.data() + data.size()};
    uint32_t V, E;
    if (!r.next(V) || !r.next(E)) return ERR_FORMAT;
    g.V = V;
    g.arcs = E;
    g.degrees.resize(V);
    g.offsets.resize((size_t)V + 1);
    g.neighbors.resize(E);
    g.dmax = 0;
    for (uint32_t i = 0; i < V; ++i) {
        if (!r.next(g.degrees[i])) return ERR_FORMAT;
        g.dmax = std::max(g.dmax, g.degrees[i]);
    }
    for (uint64_t i = 0; i <= V; ++i)
        if (!r.next(g.offsets[i])) return ERR_FORMAT;
    for (uint64_t i = 0; i < E; ++i)
        if (!r.next(g.neighbors[i])) return ERR_FORMAT;
    return OK;
}

// writeKCoreToDisk (reference src/graph.cpp:126-140): `{ ` then `"v": c, \n` for every
// non-zero core, then the last 3 bytes (", \n") are overwritten from their start by " }",
// leaving the final newline. With no non-zero core the seek fails and the file is `{ `.
int write_kcore_json(const std::string& path, uint32_t V, const uint32_t* core) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return ERR_IO;
    std::string buf = "{ ";
    bool any = false;
    char tmp[48];
    for (uint64_t i = 0; i < V; ++i) {
        if (core[i] == 0) continue;
        any = true;
        buf.push_back('"');
        auto r = std::to_chars(tmp, tmp + 24, i);
        buf.append(tmp, r.ptr);
        buf.append("\": ");
        r = std::to_chars(tmp, tmp + 24, core[i]);
        buf.append(tmp, r.ptr);
        buf.append(", \n");
        if (buf.size() > (1 << 22)) {
            // keep the last 3 bytes buffered so the final overwrite stays in memory
            std::fwrite(buf.data(), 1, buf.size() - 3, f);
            buf.erase(0, buf.size() - 3);
        }
    }
    if (any) {
        size_t k = buf.size() - 3;
        buf[k] = ' ';
        buf[k + 1] = '}';
    }
    std::fwrite(buf.data(), 1, buf.size(), f);
    return std::fclose(f) == 0 ? OK : ERR_IO;
}

int write_core_u32(const std::string& path, uint32_t V, const uint32_t* core) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return ERR_IO;
    size_t w = std::fwrite(core, 4, V, f);
    int rc = std::fclose(f);
    return (w == V && rc == 0) ? OK : ERR_IO;
}

int write_edgelist(const Csr& g, const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return ERR_IO;
    std::string buf;
    buf.reserve(1 << 22);
    for (uint32_t v = 0; v < g.V; ++v) {
        for (uint32_t j = g.offsets[v]; j < g.offsets[v + 1]; ++j) {
            uint32_t u = g.neighbors[j];
            if (u <= v) continue;
            put_u32(buf, v, ' ');
            put_u32(buf, u, '\n');
        }
        if (buf.size() > (1 << 22) - 64) {
            std::fwrite(buf.data(), 1, buf.size(), f);
            buf.clear();
        }
    }
    std::fwrite(buf.data(), 1, buf.size(), f);
    return std::fclose(f) == 0 ? OK : ERR_IO;
}

}  // namespace kcg

// ---------------------------------------------------------------------------------------
// class Graph (reference inc/graph.h:7-23, src/graph.cpp:9-161)

std::string Graph::output_dir = "./output/";

Graph::Graph() {}

void Graph::adopt() {
    V = csr_.V;
    E = (unsigned int)csr_.arcs;
    neighbors = csr_.neighbors.data();
    neighbors_offset = csr_.offsets.data();
    degrees = csr_.degrees.data();
    dmax = csr_.dmax;
    AVG_DEGREE = V ? (unsigned int)(csr_.arcs / V) : 0;
}

bool Graph::readSerialized(std::string input_file) {
    input_file = kcg::basename_of(input_file);
    std::string path = output_dir + "serialized-" + input_file;
    int st = kcg::read_serialized(path, csr_);
    if (st == kcg::OK) {
        std::cout << "Reading serialized file... " << std::endl;
        adopt();
        return true;
    }
    std::cout << std::string("serialized-") + input_file << " : File couldn't open" << std::endl;
    return false;
}

void Graph::writeSerialized(std::string input_file) {
    input_file = kcg::basename_of(input_file);
    if (kcg::write_serialized(csr_, output_dir + "serialized-" + input_file) != kcg::OK)
        std::cout << std::string("serialized-") + input_file << " : File couldn't open" << std::endl;
}

bool Graph::readFile(std::string input_file) {
    int st = kcg::load_edgelist(input_file, 0, csr_);
    if (st == kcg::ERR_IO) {
        std::cout << "load graph file failed " << std::endl;
        return false;
    }
    if (st != kcg::OK) {
        std::cout << "load graph file failed: vertex id or arc count beyond uint32" << std::endl;
        return false;
    }
    adopt();
    writeSerialized(input_file);
    return true;
}

void Graph::writeKCoreToDisk(std::string file) {
    kcg::write_kcore_json(output_dir + "pkc-kcore-" + file, V, degrees);
}

bool Graph::loadBinaryCached(std::string input_file) {
    uint64_t hs = 0;
    if (kcg::hash_file(input_file, &hs) == kcg::OK &&
        kcg::read_bincache(kcg::bincache_path(input_file, output_dir, hs), csr_, hs) == kcg::OK) {
        std::cout << "Reading binary cache... " << std::endl;
        adopt();
        return true;
    }
    std::cout << "Reading normal file... " << std::endl;
    auto start = std::chrono::steady_clock::now();
    int from_cache = 0;
    const int st = kcg::load_cached_bin(input_file, output_dir, 0, csr_, &from_cache);
    if (st != kcg::OK) {
        std::cout << (st == kcg::ERR_IO ? "load graph file failed "
                                        : "load graph file failed: vertex id or arc count beyond uint32")
                  << std::endl;
        V = 0;
        return false;
    }
    adopt();
    auto end = std::chrono::steady_clock::now();
    std::cout << "File Loaded in: "
              << std::chrono::duration_cast<std::chrono::milliseconds>(end - start).count() << std::endl;
    return true;
}

Graph::Graph(std::string input_file) {
    if (readSerialized(input_file)) return;
    std::cout << "Reading normal file... " << std::endl;
    auto start = std::chrono::steady_clock::now();
    if (!readFile(input_file)) {
        V = 0;
        return;
    }
    auto end = std::chrono::steady_clock::now();
    std::cout << "File Loaded in: "
              << std::chrono::duration_cast<std::chrono::milliseconds>(end - start).count()
              << std::endl;
}

Graph::~Graph() { std::cout << "Deallocated... " << std::endl; }
