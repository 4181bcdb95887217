// ref_loader_main.cpp — test harness linked against the reference's own loader source
// (src/graph.cpp, compiled where it lies by oracle/Makefile into oracle/_ref/). TEST
// INFRASTRUCTURE ONLY: it lets the loader tests compare this project's CSR, serialized cache
// and JSON core writer with the reference's on the same input files.
//
// usage: ref_loader <edgelist> <workdir> <dump_prefix> [core_u32_file json_name]
//   Runs `Graph g(<edgelist>)` with <workdir> as the current directory (so the reference's
//   ./output/ cache lives there), then writes <dump_prefix>.{meta,deg,off,nbr} as raw
//   little-endian uint32 arrays. With a core file, copies it into g.degrees and calls
//   g.writeKCoreToDisk(json_name).
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "graph.h"  // the reference's inc/graph.h (include path set by the Makefile)

static void dump(const std::string& p, const unsigned int* a, size_t n) {
    FILE* f = std::fopen(p.c_str(), "wb");
    if (!f) std::exit(3);
    if (n) std::fwrite(a, 4, n, f);
    std::fclose(f);
}

int main(int argc, char** argv) {
    if (argc < 4) return 2;
    std::string path = argv[1];
    if (chdir(argv[2]) != 0) return 4;
    std::string pre = argv[3];
    Graph g(path);
    unsigned int meta[2] = {g.V, g.E};
    dump(pre + ".meta", meta, 2);
    dump(pre + ".deg", g.degrees, g.V);
    dump(pre + ".off", g.neighbors_offset, (size_t)g.V + 1);
    dump(pre + ".nbr", g.neighbors, g.E);
    if (argc >= 6) {
        FILE* f = std::fopen(argv[4], "rb");
        if (!f) return 5;
        size_t got = std::fread(g.degrees, 4, g.V, f);
        std::fclose(f);
        if (got != g.V) return 6;
        g.writeKCoreToDisk(argv[5]);
    }
    return 0;
}
