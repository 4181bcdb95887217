// bincache.cpp — binary CSR cache with a version header and content hashes.
//
// The reference caches a parsed edge list as text under ./output/serialized-<basename>
// (src/graph.cpp:9-54, keyed by basename only: two different files of the same name collide,
// and reading it back is a full text parse). This cache keeps that text format untouched for
// interop and adds a binary one:
//
//   file name  <outdir>bin-<basename>-<16 hex digits of the source file's content hash>.kcg
//   offset 0   char[8] magic "KCGBIN\0\1"
//          8   u32 format version (1)
//         12   u32 V
//         16   u64 arcs
//         24   u64 content hash of the source edge list (0 when built from no file)
//         32   u64 hash of the payload (offsets bytes, then neighbour bytes)
//         40   u64 reserved (0)
//         48   u32 offsets[V+1], then u32 neighbours[arcs], little-endian
//
// Degrees are recomputed from the offsets. A cache whose header, payload hash, offsets or
// source hash do not check out is not used: the loader parses the text and rewrites the cache.
//
// The 64-bit hash is this project's own: 8-byte little-endian words are folded into a state by
// xor, multiply by an odd constant and rotate; the tail bytes form one last word; the length
// and a splitmix64 finaliser close it. It detects corruption and keys files by content; it is
// not a cryptographic hash.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <vector>

#include "graph.h"

namespace kcg {

namespace {

constexpr char kMagic[8] = {'K', 'C', 'G', 'B', 'I', 'N', 0, 1};
constexpr uint32_t kVersion = 1;
constexpr uint64_t kMul = 0x9FB21C651E98DF25ull;

inline uint64_t rotl(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }

struct Hasher {
    uint64_t h;
    uint64_t n = 0;
    uint8_t tail[8];
    uint32_t nt = 0;
    explicit Hasher(uint64_t seed) : h(seed ^ 0x243F6A8885A308D3ull) {}
    void word(uint64_t w) { h = rotl((h ^ w) * kMul, 31); }
    void add(const void* p, size_t len) {
        const uint8_t* b = static_cast<const uint8_t*>(p);
        n += len;
        while (len && nt) {  // complete a pending word first
            tail[nt++] = *b++;
            --len;
            if (nt == 8) {
                uint64_t w;
                std::memcpy(&w, tail, 8);
                word(w);
                nt = 0;
            }
        }
        for (; len >= 8; len -= 8, b += 8) {
            uint64_t w;
            std::memcpy(&w, b, 8);
            word(w);
        }
        while (len--) tail[nt++] = *b++;
    }
    uint64_t done() {
        if (nt) {
            uint64_t w = 0;
            std::memcpy(&w, tail, nt);
            word(w ^ ((uint64_t)nt << 56));
        }
        return splitmix64(h ^ n);
    }
};

struct Header {
    char magic[8];
    uint32_t version;
    uint32_t V;
    uint64_t arcs;
    uint64_t src_hash;
    uint64_t payload_hash;
    uint64_t reserved;
};
static_assert(sizeof(Header) == 48, "binary cache header is 48 bytes");

}  // namespace

uint64_t hash_bytes(const void* p, size_t n, uint64_t seed) {
    Hasher hs(seed);
    hs.add(p, n);
    return hs.done();
}

int hash_file(const std::string& path, uint64_t* out) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return ERR_IO;
    Hasher hs(0);
    std::vector<char> buf(8 << 20);
    size_t got;
    while ((got = std::fread(buf.data(), 1, buf.size(), f)) > 0) hs.add(buf.data(), got);
    const bool bad = std::ferror(f) != 0;
    std::fclose(f);
    if (bad) return ERR_IO;
    *out = hs.done();
    return OK;
}

static uint64_t payload_hash(const Csr& g) {
    Hasher hs(1);
    hs.add(g.offsets.data(), g.offsets.size() * 4);
    hs.add(g.neighbors.data(), g.neighbors.size() * 4);
    return hs.done();
}

int write_bincache(const Csr& g, const std::string& path, uint64_t src_hash) {
    if (g.offsets.size() != (size_t)g.V + 1 || g.neighbors.size() != g.arcs) return ERR_ARG;
    Header h{};
    std::memcpy(h.magic, kMagic, 8);
    h.version = kVersion;
    h.V = g.V;
    h.arcs = g.arcs;
    h.src_hash = src_hash;
    h.payload_hash = payload_hash(g);
    // write to a temporary name and rename, so a reader never sees a half-written cache
    const std::string tmp = path + ".tmp";
    FILE* f = std::fopen(tmp.c_str(), "wb");
    if (!f) return ERR_IO;
    bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1;
    ok = ok && std::fwrite(g.offsets.data(), 4, g.offsets.size(), f) == g.offsets.size();
    if (g.arcs) ok = ok && std::fwrite(g.neighbors.data(), 4, g.arcs, f) == g.arcs;
    ok = (std::fclose(f) == 0) && ok;
    if (!ok || std::rename(tmp.c_str(), path.c_str()) != 0) {
        std::remove(tmp.c_str());
        return ERR_IO;
    }
    return OK;
}

int read_bincache(const std::string& path, Csr& g, uint64_t want_src_hash) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return ERR_IO;
    Header h{};
    int st = OK;
    if (std::fread(&h, sizeof(h), 1, f) != 1 || std::memcmp(h.magic, kMagic, 8) != 0 || h.version != kVersion ||
        h.arcs >= (1ull << 32) || (want_src_hash && h.src_hash != want_src_hash)) {
        st = ERR_FORMAT;
    }
    Csr c;
    if (st == OK) {
        c.V = h.V;
        c.arcs = h.arcs;
        c.offsets.resize((size_t)h.V + 1);
        c.neighbors.resize(h.arcs);
        if (std::fread(c.offsets.data(), 4, c.offsets.size(), f) != c.offsets.size() ||
            (h.arcs && std::fread(c.neighbors.data(), 4, h.arcs, f) != h.arcs) || std::fgetc(f) != EOF)
            st = ERR_FORMAT;
    }
    std::fclose(f);
    if (st != OK) return st;
    if (payload_hash(c) != h.payload_hash) return ERR_FORMAT;
    if (c.offsets[0] != 0 || c.offsets[c.V] != c.arcs) return ERR_FORMAT;
    c.degrees.resize(c.V);
    uint32_t dmax = 0;
    for (uint32_t v = 0; v < c.V; ++v) {
        if (c.offsets[v + 1] < c.offsets[v]) return ERR_FORMAT;
        c.degrees[v] = c.offsets[v + 1] - c.offsets[v];
        dmax = c.degrees[v] > dmax ? c.degrees[v] : dmax;
    }
    for (uint64_t j = 0; j < c.arcs; ++j)
        if (c.neighbors[j] >= c.V) return ERR_FORMAT;
    c.dmax = dmax;
    g = std::move(c);
    return OK;
}

std::string bincache_path(const std::string& path, const std::string& outdir, uint64_t src_hash) {
    char hex[17];
    std::snprintf(hex, sizeof(hex), "%016llx", (unsigned long long)src_hash);
    std::string dir = outdir;
    if (!dir.empty() && dir.back() != '/') dir += '/';
    return dir + "bin-" + basename_of(path) + "-" + hex + ".kcg";
}

int load_cached_bin(const std::string& path, const std::string& outdir, int nthreads, Csr& out, int* from_cache) {
    if (from_cache) *from_cache = 0;
    uint64_t hs = 0;
    int st = hash_file(path, &hs);
    if (st != OK) return st;
    const std::string cache = bincache_path(path, outdir, hs);
    if (read_bincache(cache, out, hs) == OK) {
        if (from_cache) *from_cache = 1;
        return OK;
    }
    st = load_edgelist(path, nthreads, out);
    if (st != OK) return st;
    (void)write_bincache(out, cache, hs);  // like the text cache, a failed write only costs a reparse
    return OK;
}

}  // namespace kcg
