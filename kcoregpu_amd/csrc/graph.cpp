// graph.cpp — edge list -> CSR loader, serialized cache and core writers with the reference's
// semantics (src/graph.cpp:9-161), built for large inputs: the text is read and parsed in
// parallel chunks straight into the pair arrays, and the CSR is built by a bucket partition,
// a counting sort per bucket and a sort per vertex instead of one std::set per vertex. The
// output arrays are identical to the reference loader's.
#include "graph.h"

#include <fcntl.h>
#include <omp.h>
#include <sys/stat.h>
#include <unistd.h>

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

// The pair list becomes a CSR in three passes over cache-friendly ranges, without atomics:
//  1. each thread histograms its share of the arcs (both directions of each pair) by vertex
//     bucket (V / 2^shift <= 4096 buckets of 2^shift consecutive vertices);
//  2. each thread writes its arcs into the bucket-major order (every thread owns a slice of each
//     bucket, so no two threads write the same place);
//  3. buckets in parallel: a counting sort by source inside the bucket, then per vertex a sort
//     and duplicate removal (the reference's std::set gives ascending unique lists).
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
    int shift = 0;
    while ((((uint64_t)V - 1) >> shift) >= 4096) ++shift;
    const uint32_t B = (uint32_t)(((uint64_t)V - 1) >> shift) + 1;
    const uint32_t span = 1u << shift;
    auto lo = [&](int k) { return (int64_t)(np * (uint64_t)k / (uint64_t)nt); };

    // 1. per-thread bucket histograms
    std::vector<uint64_t> hist((size_t)nt * B, 0);
#pragma omp parallel num_threads(nt)
    {
        const int k = omp_get_thread_num();
        uint64_t* h = hist.data() + (size_t)k * B;
        for (int64_t i = lo(k); i < lo(k + 1); ++i) {
            const uint32_t a = s[i], b = t[i];
            if (a == b) continue;
            ++h[a >> shift];
            ++h[b >> shift];
        }
    }
    // bucket starts, and each thread's first slot in each bucket
    std::vector<uint64_t> bstart((size_t)B + 1, 0);
    for (uint32_t b = 0; b < B; ++b) {
        uint64_t acc = bstart[b];
        for (int k = 0; k < nt; ++k) {
            const uint64_t c = hist[(size_t)k * B + b];
            hist[(size_t)k * B + b] = acc;
            acc += c;
        }
        bstart[b + 1] = acc;
    }
    const uint64_t total = bstart[B];

    // 2. arcs into bucket-major order
    std::unique_ptr<uint32_t[]> ks(new uint32_t[total + 1]), kd(new uint32_t[total + 1]);
#pragma omp parallel num_threads(nt)
    {
        const int k = omp_get_thread_num();
        uint64_t* cur = hist.data() + (size_t)k * B;
        for (int64_t i = lo(k); i < lo(k + 1); ++i) {
            const uint32_t a = s[i], b = t[i];
            if (a == b) continue;
            uint64_t p = cur[a >> shift]++;
            ks[p] = a;
            kd[p] = b;
            p = cur[b >> shift]++;
            ks[p] = b;
            kd[p] = a;
        }
    }

    // 3. per bucket: counting sort by source, then sorted unique lists per vertex
    std::unique_ptr<uint32_t[]> buf(new uint32_t[total + 1]);
    std::unique_ptr<uint64_t[]> start(new uint64_t[(size_t)V + 1]);
    out.V = V;
    out.degrees.resize(V);
    uint32_t* deg = out.degrees.data();
#pragma omp parallel num_threads(nt)
    {
        std::vector<uint64_t> pos(span);
#pragma omp for schedule(dynamic, 1)
        for (int64_t b = 0; b < (int64_t)B; ++b) {
            const uint32_t v0 = (uint32_t)b << shift;
            const uint32_t nv = (uint32_t)std::min<uint64_t>(span, (uint64_t)V - v0);
            std::fill(pos.begin(), pos.begin() + nv, 0);
            for (uint64_t p = bstart[b]; p < bstart[b + 1]; ++p) ++pos[ks[p] - v0];
            uint64_t acc = bstart[b];
            for (uint32_t j = 0; j < nv; ++j) {
                const uint64_t c = pos[j];
                start[v0 + j] = acc;
                pos[j] = acc;
                acc += c;
            }
            for (uint64_t p = bstart[b]; p < bstart[b + 1]; ++p) buf[pos[ks[p] - v0]++] = kd[p];
            for (uint32_t j = 0; j < nv; ++j) {
                uint32_t* lb = buf.get() + start[v0 + j];
                uint32_t* le = buf.get() + pos[j];
                std::sort(lb, le);
                deg[v0 + j] = (uint32_t)(std::unique(lb, le) - lb);
            }
        }
    }
    ks.reset();
    kd.reset();
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
        if (deg[v]) std::memcpy(nb + off[v], buf.get() + start[v], (size_t)deg[v] * 4);
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

// The whole file in one buffer (no zero fill), read by nt threads with pread.
static int read_whole(const std::string& path, int nt, std::unique_ptr<char[]>& data, size_t* size) {
    const int fd = ::open(path.c_str(), O_RDONLY);
    if (fd < 0) return ERR_IO;
    struct stat st;
    if (::fstat(fd, &st) != 0) {
        ::close(fd);
        return ERR_IO;
    }
    const size_t n = st.st_size > 0 ? (size_t)st.st_size : 0;
    data.reset(new char[n + 1]);
    bool bad = false;
#pragma omp parallel num_threads(nt) reduction(|| : bad)
    {
        const int k = omp_get_thread_num(), m = omp_get_num_threads();
        size_t p = n * (size_t)k / m;
        const size_t e = n * (size_t)(k + 1) / m;
        while (p < e && !bad) {
            const ssize_t r = ::pread(fd, data.get() + p, e - p, (off_t)p);
            if (r <= 0) bad = true;
            else p += (size_t)r;
        }
    }
    ::close(fd);
    data[n] = '\n';
    *size = n;
    return bad ? ERR_IO : OK;
}

int parse_edgelist(const std::string& path, int nthreads, u32vec& S, u32vec& T) {
    std::unique_ptr<char[]> buf;
    size_t n = 0;
    const int nt = threads_or_default(nthreads);
    int st = read_whole(path, nt, buf, &n);
    if (st) return st;
    const char* base = buf.get();
    // chunk boundaries at line starts
    std::vector<size_t> cut(nt + 1, n);
    cut[0] = 0;
    for (int k = 1; k < nt; ++k) {
        size_t p = std::max(cut[k - 1], n * (size_t)k / nt);
        while (p < n && p > 0 && base[p - 1] != '\n') ++p;
        cut[k] = p;
    }
    // lines per chunk (a last line without '\n' counts too): every line gets one pair slot, so
    // each chunk writes straight to its place; a comment line's slot holds the self loop (0, 0),
    // which build_csr drops like any other self loop
    std::vector<uint64_t> first(nt + 1, 0);
#pragma omp parallel for num_threads(nt) schedule(static, 1)
    for (int k = 0; k < nt; ++k) {
        const char* p = base + cut[k];
        const char* e = base + cut[k + 1];
        uint64_t c = (uint64_t)std::count(p, e, '\n');
        if (e > p && e[-1] != '\n') ++c;
        first[k + 1] = c;
    }
    for (int k = 0; k < nt; ++k) first[k + 1] += first[k];
    S.resize(first[nt]);
    T.resize(first[nt]);
    uint32_t* sp = S.data();
    uint32_t* tp = T.data();
    int slow = 0;
#pragma omp parallel for num_threads(nt) schedule(static, 1) reduction(| : slow)
    for (int k = 0; k < nt; ++k) {
        const char* p = base + cut[k];
        const char* end = base + cut[k + 1];
        uint64_t i = first[k];
        while (p < end) {
            const char* nl = (const char*)std::memchr(p, '\n', (size_t)(end - p));
            const char* le = nl ? nl : end;
            uint32_t a = 0, b = 0;
            const int r = fast_line(p, le, a, b);
            if (r < 0) {
                slow = 1;
                break;
            }
            sp[i] = a;  // (0, 0) for a comment line
            tp[i] = b;
            ++i;
            p = nl ? nl + 1 : end;
        }
    }
    if (!slow) return OK;
    // exact sequential emulation, line for line. The reference leaves s and t
    // uninitialised before the first line; they start at 0 here.
    S.clear();
    T.clear();
    unsigned int s = 0, t = 0;
    size_t p = 0;
    std::string line;
    while (p < n) {
        const char* q = (const char*)std::memchr(base + p, '\n', n - p);
        const size_t le = q ? (size_t)(q - base) : n;
        line.assign(base + p, le - p);
        p = q ? le + 1 : n;
        if (std::isalpha((unsigned char)line[0])) continue;
        std::istringstream iss(line);
        iss >> s >> t;
        if (s == t) continue;
        S.push_back(s);
        T.push_back(t);
    }
    return OK;
}

int load_edgelist(const std::string& path, int nthreads, Csr& out) {
    u32vec s, t;
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

// Items [0, n) are formatted by nt threads in rounds of nt slices (so the memory held is bounded):
// fmt(i, q) writes item i (at most `maxb` bytes) at q and returns the end. Each thread formats
// into its own buffer through a local cursor, so nothing it writes shares a cache line with
// another thread's writes; the round's buffers are then passed in order to emit(p, len).
struct FmtBuf {
    std::unique_ptr<char[]> p;
    size_t len = 0;
    char pad[64 - sizeof(std::unique_ptr<char[]>) - sizeof(size_t)];
};
constexpr uint64_t kFmtSlice = 1u << 18;  // items per thread and round

template <class Fmt, class Emit>
static void format_parallel(uint64_t n, size_t maxb, int nt, Fmt fmt, Emit emit) {
    std::vector<FmtBuf> bufs(nt);
    for (auto& b : bufs) b.p.reset(new char[kFmtSlice * maxb]);
    for (uint64_t r0 = 0; r0 < n; r0 += kFmtSlice * (uint64_t)nt) {
#pragma omp parallel for num_threads(nt) schedule(static, 1)
        for (int k = 0; k < nt; ++k) {
            char* q = bufs[k].p.get();
            const uint64_t lo = std::min(n, r0 + kFmtSlice * (uint64_t)k);
            const uint64_t hi = std::min(n, lo + kFmtSlice);
            for (uint64_t i = lo; i < hi; ++i) q = fmt(i, q);
            bufs[k].len = (size_t)(q - bufs[k].p.get());
        }
        for (int k = 0; k < nt; ++k) emit(bufs[k].p.get(), bufs[k].len);
    }
}

// n values, each followed by `sep`
static bool write_values(FILE* f, const uint32_t* a, uint64_t n, char sep, int nt) {
    bool ok = true;
    format_parallel(
        n, 11, nt,
        [&](uint64_t i, char* q) {
            q = std::to_chars(q, q + 10, a[i]).ptr;
            *q++ = sep;
            return q;
        },
        [&](const char* p, size_t len) { ok = ok && std::fwrite(p, 1, len, f) == len; });
    return ok;
}

int write_serialized(const Csr& g, const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return ERR_IO;
    const int nt = threads_or_default(0);
    std::string head;
    put_u32(head, g.V, '\n');
    put_u32(head, (uint32_t)g.arcs, '\n');
    bool ok = std::fwrite(head.data(), 1, head.size(), f) == head.size();
    ok = ok && write_values(f, g.degrees.data(), g.V, '\n', nt);
    ok = ok && write_values(f, g.offsets.data(), (uint64_t)g.V + 1, ' ', nt);
    ok = ok && write_values(f, g.neighbors.data(), g.arcs, ' ', nt);
    return (std::fclose(f) == 0 && ok) ? OK : ERR_IO;
}

// One core number per line, each followed by '\n': the form of the degree block of the
// serialized cache (src/graph.cpp:35-54), so the same tools read both.
int write_core_text(const std::string& path, uint32_t V, const uint32_t* core) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return ERR_IO;
    const bool ok = write_values(f, core, V, '\n', threads_or_default(0));
    return (std::fclose(f) == 0 && ok) ? OK : ERR_IO;
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

// The serial reader defines the format: whitespace-separated decimal tokens, V, E, V degrees,
// V + 1 offsets, E neighbours, anything after them ignored (reference src/graph.cpp:9-33).
static int read_serialized_serial(TokenReader r, uint32_t V, uint32_t E, Csr& g) {
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

int read_serialized(const std::string& path, Csr& g) {
    std::unique_ptr<char[]> data;
    size_t n = 0;
    const int nt = omp_get_max_threads();
    int st = read_whole(path, nt, data, &n);
    if (st) return st;
    TokenReader r{data.get(), data.get() + n};
    uint32_t V, E;
    if (!r.next(V) || !r.next(E)) return ERR_FORMAT;
    g.V = V;
    g.arcs = E;
    g.degrees.resize(V);
    g.offsets.resize((size_t)V + 1);
    g.neighbors.resize(E);
    // Parallel path: the rest cut at whitespace into nt chunks, tokens counted per chunk, then
    // each chunk parses its tokens into their places. Any token among the needed ones that is not
    // plain digits within uint32, or too few tokens, hands the file to the serial reader, which
    // decides (it accepts, e.g., garbage glued to the very last needed token).
    const char* b0 = r.p;
    const char* e0 = r.e;
    const uint64_t need = 2ull * V + 1 + E;
    std::vector<const char*> cut(nt + 1, e0);
    cut[0] = b0;
    for (int k = 1; k < nt; ++k) {
        const char* p = std::max(cut[k - 1], b0 + (size_t)(e0 - b0) * k / nt);
        while (p < e0 && !is_ws(*p)) ++p;
        cut[k] = p;
    }
    std::vector<uint64_t> first(nt + 1, 0);
#pragma omp parallel for num_threads(nt) schedule(static, 1)
    for (int k = 0; k < nt; ++k) {
        uint64_t c = 0;
        bool in = false;
        for (const char* p = cut[k]; p < cut[k + 1]; ++p) {
            const bool ws = is_ws(*p);
            c += (!ws && !in) ? 1 : 0;
            in = !ws;
        }
        first[k + 1] = c;
    }
    for (int k = 0; k < nt; ++k) first[k + 1] += first[k];
    bool irregular = first[nt] < need;
    uint32_t dmax = 0;
    uint32_t* deg = g.degrees.data();
    uint32_t* off = g.offsets.data();
    uint32_t* nbr = g.neighbors.data();
    if (!irregular) {
#pragma omp parallel for num_threads(nt) schedule(static, 1) reduction(|| : irregular) reduction(max : dmax)
        for (int k = 0; k < nt; ++k) {
            uint64_t idx = first[k];
            const char* p = cut[k];
            const char* e = cut[k + 1];
            while (idx < need) {
                while (p < e && is_ws(*p)) ++p;
                if (p >= e) break;
                uint64_t a = 0;
                int nd = 0;
                while (p < e && *p >= '0' && *p <= '9' && nd < 11) {
                    a = a * 10 + (uint64_t)(*p++ - '0');
                    ++nd;
                }
                if (nd == 0 || nd > 10 || a > UINT32_MAX || (p < e && !is_ws(*p))) {
                    irregular = true;
                    break;
                }
                const uint32_t v = (uint32_t)a;
                if (idx < V) {
                    deg[idx] = v;
                    dmax = std::max(dmax, v);
                } else if (idx <= 2ull * V) {
                    off[idx - V] = v;
                } else {
                    nbr[idx - 2ull * V - 1] = v;
                }
                ++idx;
            }
        }
    }
    if (irregular) return read_serialized_serial(r, V, E, g);
    g.dmax = dmax;
    return OK;
}

// writeKCoreToDisk (reference src/graph.cpp:126-140): `{ ` then `"v": c, \n` for every
// non-zero core, then the last 3 bytes (", \n") are overwritten from their start by " }",
// leaving the final newline. With no non-zero core the seek fails and the file is `{ `.
int write_kcore_json(const std::string& path, uint32_t V, const uint32_t* core) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return ERR_IO;
    bool ok = std::fwrite("{ ", 1, 2, f) == 2;
    // every entry ends with ", \n": the last 3 bytes written so far are held back, so that the
    // final overwrite by " }" happens in memory
    char carry[3];
    bool any = false;
    format_parallel(
        V, 27, threads_or_default(0),
        [&](uint64_t i, char* q) {
            if (core[i] == 0) return q;
            *q++ = '"';
            q = std::to_chars(q, q + 10, i).ptr;
            *q++ = '"';
            *q++ = ':';
            *q++ = ' ';
            q = std::to_chars(q, q + 10, core[i]).ptr;
            *q++ = ',';
            *q++ = ' ';
            *q++ = '\n';
            return q;
        },
        [&](const char* p, size_t len) {
            if (len == 0) return;
            if (any) ok = ok && std::fwrite(carry, 1, 3, f) == 3;
            ok = ok && std::fwrite(p, 1, len - 3, f) == len - 3;
            std::memcpy(carry, p + len - 3, 3);
            any = true;
        });
    if (any) ok = ok && std::fwrite(" }\n", 1, 3, f) == 3;
    return (std::fclose(f) == 0 && ok) ? OK : ERR_IO;
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
