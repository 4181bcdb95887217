// graph.h — host graph layer: reference-compatible `Graph` class plus the CSR builder,
// serialized cache and writers behind the C ABI in include/kcore_graph.h.
//
// `class Graph` keeps the public fields and member functions of the reference's
// inc/graph.h:7-23 so host code written against that class compiles unchanged; its
// behaviour (console lines, cache path ./output/serialized-<basename>, JSON writer)
// follows src/graph.cpp:9-161.
#pragma once
#include <cstddef>
#include <cstdint>
#include <memory>
#include <new>
#include <string>
#include <utility>
#include <vector>

namespace kcg {

// Allocator whose resize() leaves new elements uninitialised: the loader's arrays are written
// in parallel right after they are sized, so a serial zero fill would only cost time.
template <class T>
struct uninit_alloc : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = uninit_alloc<U>;
    };
    uninit_alloc() = default;
    template <class U>
    uninit_alloc(const uninit_alloc<U>&) noexcept {}
    template <class U>
    void construct(U* p) noexcept {
        ::new (static_cast<void*>(p)) U;
    }
    template <class U, class... A>
    void construct(U* p, A&&... a) {
        ::new (static_cast<void*>(p)) U(std::forward<A>(a)...);
    }
};
using u32vec = std::vector<uint32_t, uninit_alloc<uint32_t>>;

// A CSR graph as the loader builds it: V vertices, `arcs` = 2 * undirected edges,
// adjacency ascending and duplicate-free, no self loops (reference src/graph.cpp:97-122).
struct Csr {
    uint32_t V = 0;
    uint64_t arcs = 0;
    uint32_t dmax = 0;
    u32vec offsets;    // V + 1
    u32vec neighbors;  // arcs
    u32vec degrees;    // V
};

// Status codes shared with the C ABI.
enum Status { OK = 0, ERR_IO = 1, ERR_RANGE = 2, ERR_ARG = 3, ERR_FORMAT = 4 };

// Build a CSR from a pair list with the loader's rules: pairs with s == t are dropped,
// V = 1 + max id over the remaining pairs (1 when none remain), every pair is inserted in
// both directions, duplicates removed, adjacency sorted ascending.
int build_csr(const uint32_t* s, const uint32_t* t, uint64_t npairs, int nthreads, Csr& out);

// Parse a text edge list with the reference's per-line rules (src/graph.cpp:68-96):
// lines whose first char is alphabetic are skipped, each other line is read as
// `iss >> s >> t` (values kept from the previous line when extraction fails). Returns the
// pairs, which may include self loops (s == t: build_csr drops them, as the reference does).
int parse_edgelist(const std::string& path, int nthreads, u32vec& s, u32vec& t);

int load_edgelist(const std::string& path, int nthreads, Csr& out);
int read_serialized(const std::string& path, Csr& out);
int write_serialized(const Csr& g, const std::string& path);
int write_kcore_json(const std::string& path, uint32_t V, const uint32_t* core);
int write_core_u32(const std::string& path, uint32_t V, const uint32_t* core);
int write_core_text(const std::string& path, uint32_t V, const uint32_t* core);
int write_edgelist(const Csr& g, const std::string& path);

std::string basename_of(const std::string& path);

// ---- binary cache (bincache.cpp) ----
uint64_t hash_bytes(const void* p, size_t n, uint64_t seed);
int hash_file(const std::string& path, uint64_t* out);
int write_bincache(const Csr& g, const std::string& path, uint64_t src_hash);
int read_bincache(const std::string& path, Csr& g, uint64_t want_src_hash);
std::string bincache_path(const std::string& path, const std::string& outdir, uint64_t src_hash);
int load_cached_bin(const std::string& path, const std::string& outdir, int nthreads, Csr& out, int* from_cache);

// ---- generators (generators.cpp) ----
uint64_t splitmix64(uint64_t x);
int gen_er(uint32_t n, uint64_t m, uint64_t seed, int nthreads, Csr& out);
int gen_chung_lu(uint32_t n, double gamma, double wmin, double wmax, double avg_deg, uint64_t seed,
                 int nthreads, Csr& out);
int gen_rmat(uint32_t scale, uint64_t ndraws, double a, double b, double c, uint64_t seed,
             int nthreads, Csr& out);

}  // namespace kcg

// Reference-compatible graph object (inc/graph.h:7-23). Arrays are owned by the object.
class Graph {
public:
    unsigned int V = 0;
    unsigned int E = 0;  // arcs = 2 * undirected edges, as printed by the reference driver
    unsigned int AVG_DEGREE = 0;
    unsigned int* neighbors = nullptr;
    unsigned int* neighbors_offset = nullptr;
    unsigned int* degrees = nullptr;
    unsigned int kmax = 0, dmax = 0;

    explicit Graph(std::string input_file);
    Graph();
    Graph(const Graph&) = delete;
    Graph& operator=(const Graph&) = delete;
    bool readSerialized(std::string input_file);
    void writeSerialized(std::string input_file);
    // Reference readFile exits the process on a missing file; this one returns false after
    // printing the same line so the driver decides.
    bool readFile(std::string input_file);
    // Binary-cache variant of the constructor: the content-keyed binary cache first, else parse
    // the text and write that cache. Prints "Reading binary cache... " or the text-load lines.
    bool loadBinaryCached(std::string input_file);
    void writeKCoreToDisk(std::string file);
    ~Graph();

    // Directory holding the serialized cache and outputs (reference OUTPUT_LOC "./output/").
    static std::string output_dir;

private:
    kcg::Csr csr_;
    void adopt();
};
