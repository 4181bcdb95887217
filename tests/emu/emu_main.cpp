// emu_main.cpp — runs this project's peeling kernels on the CPU through the wave64 emulator
// (test infrastructure: tests/test_emu.py builds it with the host clang and compares every result
// with the oracle). The kernels are the unchanged device headers of kcoregpu_amd/csrc; the
// argument set-up mirrors what kcore_hip.hip and serve.hip do on the GPU.
//
//   emu peel <variant: persistent|prefetch|async|asyncnb> <graph> <out> [heavy chunk flags grid ratio]
//            (asyncnb: the cascade-queue kernel with the level hand-off, no grid barriers)
//   emu stepwise <graph> <out> [flags]  (the host level loop of kcore_hip.hip's peel_stepwise)
//   emu batch <graphs...> <out>
//   emu serve <graphs...> <out>          (every request published before the kernel starts)
//   emu lease                            (an idle serving kernel leaves when the heartbeat stops)
// <graph>: u32 V, u64 arcs, u32 offsets[V+1], u32 neighbours[arcs]. <out>: u32 kmax, u32 status,
// then u32 core[V] (peel); per graph u32 kmax, u32 status, u32 core[V_g] (batch, serve).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "emu_sched.h"
#include "../../kcoregpu_amd/csrc/batch_kernel.h"
#include "../../kcoregpu_amd/csrc/peel_async.h"
#include "../../kcoregpu_amd/csrc/serve_kernel.h"
#include "../../kcoregpu_amd/csrc/stepwise_kernel.h"

using namespace kcdev;

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

template <class T>
T* zalloc(size_t n) {
    // 256-byte aligned, zeroed, with slack as device allocations have; KCEMU_EXACT=1: exactly n
    // elements, so that AddressSanitizer reports any access past the end
    static const bool exact = std::getenv("KCEMU_EXACT") && *std::getenv("KCEMU_EXACT") == '1';
    const size_t m = exact ? (n ? n : 1) : n + 64;
    void* p = nullptr;
    if (posix_memalign(&p, 256, m * sizeof(T))) std::exit(4);
    std::memset(p, 0, m * sizeof(T));
    return static_cast<T*>(p);
}

constexpr uint64_t kTicksPerMs = 100000;  // the emulator's clock counts 10 ns ticks

int run_peel(int argc, char** argv) {
    const std::string variant = argv[2];
    G g = read_graph(argv[3]);
    const uint32_t heavy = argc > 5 ? (uint32_t)std::strtoul(argv[5], nullptr, 10) : 64;
    const uint32_t chunk = argc > 6 ? (uint32_t)std::strtoul(argv[6], nullptr, 10) : 512;
    const uint32_t flags = argc > 7 ? (uint32_t)std::strtoul(argv[7], nullptr, 10) : 0;
    const uint32_t grid = argc > 8 ? (uint32_t)std::strtoul(argv[8], nullptr, 10) : 4;
    const uint32_t ratio = argc > 9 ? (uint32_t)std::strtoul(argv[9], nullptr, 10) : 4;
    const uint32_t V = g.V, padV = (V + 255u) & ~255u;
    uint32_t dmax = 0;
    for (uint32_t v = 0; v < V; ++v) dmax = std::max(dmax, g.off[v + 1] - g.off[v]);
    const uint64_t chunk_cap = g.arcs / 16 + g.arcs / 64 + 64;
    uint32_t* deg0 = zalloc<uint32_t>(padV);
    uint32_t* core = zalloc<uint32_t>(padV);
    for (uint32_t v = 0; v < padV; ++v) deg0[v] = v < V ? g.off[v + 1] - g.off[v] : kSent;
    std::memcpy(core, deg0, (size_t)padV * 4);
    const bool async = variant == "async" || variant == "asyncnb";
    uint32_t* order = zalloc<uint32_t>(padV);
    if (async)
        for (uint32_t v = 0; v < padV; ++v) order[v] = kSent;
    PeelArgs a{};
    a.off = g.off.data();
    a.nbr = g.nbr.data();
    a.deg0 = deg0;
    a.core = core;
    a.order = order;
    a.work = zalloc<uint32_t>(padV);
    a.chunks = zalloc<uint64_t>(chunk_cap);
    a.shell = zalloc<uint32_t>(dmax + 3);
    a.wgmin = zalloc<uint32_t>(4096);
    a.trace = zalloc<uint64_t>(2);
    a.dbg = zalloc<uint32_t>(4096 * 8);
    a.ctrl = zalloc<uint32_t>((size_t)A_COUNT * kStride);
    a.V = V;
    a.nvec = padV / 4;
    a.dmax = dmax;
    a.heavy = heavy < 16 ? 16 : heavy;
    a.chunk = chunk < 64 ? 64 : chunk;
    a.chunk_cap = (uint32_t)chunk_cap;
    a.lq_depth = async ? 0 : WgState<256>::LQ;
    a.lq_arcs = 16384;
    a.lq_rounds = 4;
    a.flags = flags;
    a.trace_cap = 0;
    a.time_limit = 20000 * kTicksPerMs;
    a.reg_window = 200 * kTicksPerMs;  // OS threads start slower than workgroups
    uint32_t* alive = zalloc<uint32_t>((size_t)padV * 2);
    a.alive[0] = alive;
    a.alive[1] = alive + padV;
    a.async_mode = async ? 1 : 0;
    a.compact_ratio = ratio;
    a.poll_max = 2;
    if (variant == "asyncnb") {
        kcemu::launch(grid, 256, [&] { peel_async_nb<256, false>(a); });
    } else if (variant == "async") {
        kcemu::launch(grid, 256, [&] { peel_async<256, false>(a); });
    } else if (variant == "prefetch") {
        kcemu::launch(grid, 256, [&] { peel_persistent<256, true>(a); });
    } else {
        kcemu::launch(grid, 256, [&] { peel_persistent<256, false>(a); });
    }
    const uint32_t* r = a.ctrl + (size_t)C_RESULT * kStride;
    FILE* f = std::fopen(argv[4], "wb");
    uint32_t head[2] = {r[R_KMAX], r[R_ERR]};
    std::fwrite(head, 4, 2, f);
    std::fwrite(core, 4, V, f);
    std::fwrite(order, 4, V, f);
    std::fwrite(a.shell, 4, r[R_KMAX] + 2, f);
    std::fclose(f);
    std::printf("levels %u subrounds %u participants %u dense %u list %u surv %u chunks %u held %u\n", r[R_LEVELS],
                r[R_SUBROUNDS], r[R_NPART], r[R_WORDS], r[R_WORDS + 1], r[R_WORDS + 2], r[R_WORDS + 3],
                r[R_WORDS + 4]);
    return 0;
}

int run_stepwise(int argc, char** argv) {
    G g = read_graph(argv[2]);
    const uint32_t flags = argc > 4 ? (uint32_t)std::strtoul(argv[4], nullptr, 10) : 0;
    const uint32_t V = g.V;
    uint32_t dmax = 0;
    for (uint32_t v = 0; v < V; ++v) dmax = std::max(dmax, g.off[v + 1] - g.off[v]);
    uint32_t* core = zalloc<uint32_t>(V);
    for (uint32_t v = 0; v < V; ++v) core[v] = g.off[v + 1] - g.off[v];
    uint32_t* order = zalloc<uint32_t>(V);
    uint32_t* ctr = zalloc<uint32_t>(64);
    std::vector<uint32_t> shell(dmax + 3, 0);
    uint32_t L = 0, tail = 0, levels = 0, launches = 0, status = 0;
    while (true) {
        shell[L] = tail;
        ctr[1] = kSent;
        kcemu::launch(2, 256, [&] { step_select(core, V, L, order, ctr); });
        ++launches;
        ++levels;
        uint32_t beg = tail, end = ctr[0];
        while (beg < end) {
            kcemu::launch(3, 256, [&] { step_process(g.off.data(), g.nbr.data(), core, order, beg, end, L, ctr); });
            ++launches;
            beg = end;
            end = ctr[0];
        }
        tail = end;
        if (tail >= V) {
            shell[L + 1] = tail;
            break;
        }
        uint32_t next = L + 1;
        if (!(flags & 1u) && ctr[1] != kSent && ctr[1] > next) next = ctr[1];
        if (next > dmax) {
            status = 2;
            break;
        }
        for (uint32_t k = L + 1; k <= next; ++k) shell[k] = tail;
        L = next;
    }
    FILE* f = std::fopen(argv[3], "wb");
    uint32_t head[2] = {L, status};
    std::fwrite(head, 4, 2, f);
    std::fwrite(core, 4, V, f);
    std::fwrite(order, 4, V, f);
    std::fwrite(shell.data(), 4, L + 2, f);
    std::fclose(f);
    std::printf("levels %u launches %u\n", levels, launches);
    return 0;
}

int run_batch(int argc, char** argv) {
    std::vector<G> gs;
    for (int i = 2; i < argc - 1; ++i) gs.push_back(read_graph(argv[i]));
    const uint32_t Gn = (uint32_t)gs.size();
    std::vector<uint64_t> vo(Gn + 1, 0), ao(Gn + 1, 0);
    for (uint32_t i = 0; i < Gn; ++i) {
        vo[i + 1] = vo[i] + gs[i].V;
        ao[i + 1] = ao[i] + gs[i].arcs;
    }
    const uint64_t nv = vo[Gn], na = ao[Gn];
    uint32_t* off = zalloc<uint32_t>(nv + Gn);
    uint32_t* nbr = zalloc<uint32_t>(na + 1);
    uint32_t* core = zalloc<uint32_t>((nv + 255) & ~255ull);
    uint32_t* order = zalloc<uint32_t>(nv + 1);
    for (uint32_t i = 0; i < Gn; ++i) {
        std::memcpy(off + vo[i] + i, gs[i].off.data(), ((size_t)gs[i].V + 1) * 4);
        if (gs[i].arcs) std::memcpy(nbr + ao[i], gs[i].nbr.data(), gs[i].arcs * 4);
        for (uint32_t v = 0; v < gs[i].V; ++v) core[vo[i] + v] = gs[i].off[v + 1] - gs[i].off[v];
    }
    std::vector<uint32_t> sched(Gn);
    for (uint32_t i = 0; i < Gn; ++i) sched[i] = Gn - 1 - i;
    BatchArgs b{};
    b.vtx_off = vo.data();
    b.arc_off = ao.data();
    b.off = off;
    b.nbr = nbr;
    b.core = core;
    b.order = order;
    b.kmax = zalloc<uint32_t>(Gn);
    b.sched = sched.data();
    b.ctrl = zalloc<uint32_t>(1024);
    b.G = Gn;
    b.flags = 0;
    kcemu::launch(3, 256, [&] { peel_batch<256>(b); });
    FILE* f = std::fopen(argv[argc - 1], "wb");
    for (uint32_t i = 0; i < Gn; ++i) {
        uint32_t head[2] = {b.kmax[i], b.ctrl[192]};
        std::fwrite(head, 4, 2, f);
        std::fwrite(core + vo[i], 4, gs[i].V, f);
    }
    std::fclose(f);
    return 0;
}

int run_serve(int argc, char** argv) {
    std::vector<G> gs;
    for (int i = 2; i < argc - 1; ++i) gs.push_back(read_graph(argv[i]));
    const uint32_t n = (uint32_t)gs.size(), slots = 4;
    // arena regions back to back, like the first-fit allocator on an empty arena
    std::vector<uint64_t> base(n);
    uint64_t total = 0;
    for (uint32_t i = 0; i < n; ++i) {
        base[i] = total;
        total += serve_region(gs[i].V, (uint32_t)gs[i].arcs).bytes;
    }
    uint8_t* arena = zalloc<uint8_t>(total);
    ServeEntry* ring = zalloc<ServeEntry>(slots);
    ServeDone* done = zalloc<ServeDone>(slots);
    uint32_t* head = zalloc<uint32_t>(256 / 4);
    uint32_t* ctl = zalloc<uint32_t>(256 / 4);
    std::vector<std::vector<uint32_t>> out(n);
    for (uint32_t i = 0; i < n; ++i) {
        const ServeRegion r = serve_region(gs[i].V, (uint32_t)gs[i].arcs);
        std::memcpy(arena + base[i] + r.off, gs[i].off.data(), ((size_t)gs[i].V + 1) * 4);
        if (gs[i].arcs) std::memcpy(arena + base[i] + r.nbr, gs[i].nbr.data(), gs[i].arcs * 4);
        out[i].assign(gs[i].V, 0xDEADBEEFu);
    }
    ServeArgs a{};
    a.arena = arena;
    a.ring = ring;
    a.done = done;
    a.head = reinterpret_cast<uint64_t*>(head);
    a.stop = head + 32;
    a.hstop = ctl;
    a.hbeat = ctl + 16;
    a.expired = head + 48;
    a.lease_ticks = 60000 * kTicksPerMs;
    a.slots = slots;
    std::vector<uint32_t> kmax(n), status(n);
    std::thread kernel([&] { kcemu::launch(2, 256, [&] { serve_kernel<256>(a); }); });
    // the host: publish request i into slot i % slots once its slot's previous request is done
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t sl = i % slots;
        if (i >= slots) {
            while (__atomic_load_n(&done[sl].seq, __ATOMIC_SEQ_CST) != (uint64_t)(i - slots) + 1) std::this_thread::yield();
            kmax[i - slots] = done[sl].kmax;
            status[i - slots] = done[sl].err;
        }
        ServeEntry e{};
        e.base = base[i];
        e.V = gs[i].V;
        e.arcs = (uint32_t)gs[i].arcs;
        e.flags = 0;
        e.out = out[i].data();
        e.seq = (uint64_t)i + 1;
        ring[sl].base = e.base;
        ring[sl].V = e.V;
        ring[sl].arcs = e.arcs;
        ring[sl].flags = 0;
        ring[sl].out = e.out;
        __atomic_store_n(&ring[sl].seq, e.seq, __ATOMIC_SEQ_CST);
    }
    for (uint32_t i = (n > slots ? n - slots : 0); i < n; ++i) {
        const uint32_t sl = i % slots;
        while (__atomic_load_n(&done[sl].seq, __ATOMIC_SEQ_CST) != (uint64_t)i + 1) std::this_thread::yield();
        kmax[i] = done[sl].kmax;
        status[i] = done[sl].err;
    }
    __atomic_store_n(head + 32, 1u, __ATOMIC_SEQ_CST);  // stop
    kernel.join();
    FILE* f = std::fopen(argv[argc - 1], "wb");
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t h2[2] = {kmax[i], status[i]};
        std::fwrite(h2, 4, 2, f);
        std::fwrite(out[i].data(), 4, gs[i].V, f);
    }
    std::fclose(f);
    return 0;
}

int run_lease() {
    // no request ever comes and the heartbeat never moves: every workgroup must leave by itself
    ServeEntry* ring = zalloc<ServeEntry>(4);
    ServeDone* done = zalloc<ServeDone>(4);
    uint32_t* head = zalloc<uint32_t>(64);
    uint32_t* ctl = zalloc<uint32_t>(64);
    uint8_t* arena = zalloc<uint8_t>(1024);
    ServeArgs a{};
    a.arena = arena;
    a.ring = ring;
    a.done = done;
    a.head = reinterpret_cast<uint64_t*>(head);
    a.stop = head + 32;
    a.hstop = ctl;
    a.hbeat = ctl + 16;
    a.expired = head + 48;
    a.lease_ticks = 50 * kTicksPerMs;
    a.slots = 4;
    const uint64_t t0 = kcemu::clock_ticks();
    kcemu::launch(2, 256, [&] { serve_kernel<256>(a); });
    const double ms = (kcemu::clock_ticks() - t0) / (double)kTicksPerMs;
    std::printf("lease exit after %.1f ms, expired flag %u\n", ms, head[48]);
    return head[48] == 1 ? 0 : 1;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const std::string cmd = argv[1];
    if (cmd == "peel" && argc >= 5) return run_peel(argc, argv);
    if (cmd == "batch" && argc >= 4) return run_batch(argc, argv);
    if (cmd == "stepwise" && argc >= 4) return run_stepwise(argc, argv);
    if (cmd == "serve" && argc >= 4) return run_serve(argc, argv);
    if (cmd == "lease") return run_lease();
    return 2;
}
