// peel_async.h — the cascade-queue single-graph peeling kernel (KC_VARIANT_ASYNC).
//
// Same method as peel_persistent (the reference's level-synchronous peel, src/ours.cc:2-160) and
// the same per-arc rule, but a level has exactly two grid barriers, whatever its cascade depth:
//
//  * scan: every workgroup takes a static share of the vertices still to be peeled and records
//    those with degree == L in order[] (cutting heavy ones into arc chunks). Once few vertices
//    are left unpeeled (unpeeled * compact_ratio < entries a scan reads) a scan also writes the
//    vertices it finds with degree > L to a fresh alive list, and later scans read that list
//    (ids, then their degrees) instead of the whole degree array.
//  * barrier, then processing with no barrier inside the level: order[] and the chunk list are
//    work queues. A workgroup claims up to BLOCK entries of order[] or one chunk per wave with a
//    compare-and-swap on the queue head, processes them (decrements, undo, crossings), and its
//    crossings are appended to order[] (and chunked when heavy) where any workgroup can claim
//    them at once. Slots are pre-filled with a sentinel (order[]: ~0, chunks: 0), so a claimed
//    slot whose writer reserved it but has not stored it yet is simply waited for.
//  * termination: VDONE / CDONE count processed order[] entries / chunks over the whole peel.
//    After a workgroup's appends are complete it adds what it processed; whoever then reads
//    VDONE == OTAIL and CDONE == CTAIL (done counters read first, tails after) knows that
//    nothing is in flight and nothing more can be appended, and raises the level's done flag
//    (one copy per XCD group, tagged with the level's generation).
//  * barrier: the last workgroup computes the next level (smallest degree left above L), the
//    shell offsets and whether the next scan compacts.
// Cross-workgroup data uses the same device-scope (SC1) accesses as peel_persistent; there are
// no cache-wide fences.
#pragma once
#include "peel_kernel.h"

namespace kcdev {

enum : uint32_t {
    A_VHEAD = C_COUNT,  // next order[] entry to claim
    A_VDONE,            // order[] entries processed
    A_CHEAD,            // next chunk to claim
    A_CDONE,            // chunks processed
    A_MODE,             // 0: scans read the dense degree array, 1: an alive list
    A_LIST,             // current alive list (0/1)
    A_LEN,              // its length
    A_COMPACT,          // 1: the next scan writes a new alive list into A_DST
    A_DST,
    A_ALEN0,            // append counters of the two alive lists
    A_ALEN1,
    A_NEXT,             // next level
    A_PHASE,            // PH_NEXT_LEVEL / PH_DONE
    A_SCAN_DENSE,       // scan accounting: dense entries read
    A_SCAN_LIST,        // list entries read (id + gathered degree)
    A_SURV,             // alive-list entries written
    A_SVBEG,            // the level's scan frontier: order[] range and chunk range, dealt
    A_SVEND,            // statically over every wave before the queues open
    A_SCBEG,
    A_SCEND,
    A_SDONE,            // level hand-off: scan shares done over the whole peel (npart per level)
    A_GMIN0,            // level hand-off: smallest degree above L seen in even / odd levels,
    A_GMIN1,            // stored complemented (0 = none yet)
    A_LCLAIM,           // level hand-off: generation of the last level whose completion was won
    A_HELD,             // level hand-off: claims of the next level held for it (reported only)
    A_HINT,             // kXcds work hints: advanced after a workgroup appended crossings
    A_LDONE = A_HINT + kXcds,  // kXcds level-done flags
    A_GATE = A_LDONE + kXcds,  // level hand-off: kXcds flags, the generation whose scans are all counted
    A_COUNT = A_GATE + kXcds
};

template <int BLOCK>
struct WgStateA : WgState<BLOCK> {
    uint32_t surv[BLOCK / 64][kStageCap];
    uint32_t kind, beg, n;
    uint32_t vcount, ccount;
    uint32_t hkind, hbeg, hn;  // level hand-off: a claim of the next level, held for it (hkind 3: none)
};

__device__ __forceinline__ void order_point() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    drain_memory();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

__device__ __forceinline__ bool cas_rlx(uint32_t* p, uint32_t expect, uint32_t want) {
    return __hip_atomic_compare_exchange_strong(p, &expect, want, __ATOMIC_RELAXED, __ATOMIC_RELAXED, KC_AGENT);
}

// Alive list i (0/1). A select of the two fields, not an index: indexing the argument struct at
// run time would keep the whole struct in scratch memory.
__device__ __forceinline__ uint32_t* alive_list(const PeelArgs& a, uint32_t i) { return i ? a.alive[1] : a.alive[0]; }

// Survivors (degree > L) staged per wave and appended to the alive list being built.
template <int BLOCK>
__device__ void flush_surv(const PeelArgs& a, WgStateA<BLOCK>& S, uint32_t& n, uint32_t dst) {
    if (n == 0) return;
    const uint32_t lane = lane_id();
    const uint32_t* st = S.surv[threadIdx.x / 64];
    const uint32_t b = wave_reserve(cw(a, A_ALEN0 + dst), n);
    for (uint32_t i = lane; i < n; i += 64) st_rlx(alive_list(a, dst) + b + i, st[i]);
    n = 0;
    __builtin_amdgcn_wave_barrier();
}

template <int BLOCK>
__device__ __forceinline__ void surv_push(const PeelArgs& a, WgStateA<BLOCK>& S, uint32_t& n, bool pred, uint32_t v,
                                          uint32_t dst) {
    const uint64_t m = __ballot(pred);
    if (m == 0) return;
    const uint32_t cnt = __popcll(m);
    if (n + cnt > kStageCap) flush_surv(a, S, n, dst);
    if (pred) S.surv[threadIdx.x / 64][n + rank_below(m)] = v;
    n += cnt;
    __builtin_amdgcn_wave_barrier();
}

// Scan of the dense degree array (16-byte SC1 loads, four in flight per lane).
template <int BLOCK>
__device__ __forceinline__ void scan_dense(const PeelArgs& a, WgStateA<BLOCK>& S, uint32_t L, uint32_t& mn, bool compact,
                           uint32_t dst) {
    constexpr uint32_t WPB = BLOCK / 64;
    constexpr uint32_t U = 4;
    const uint32_t lane = lane_id();
    const uint32_t gw = S.vbid * WPB + threadIdx.x / 64;
    const uint32_t GW = S.npart * WPB;
    const __amdgpu_buffer_rsrc_t core = make_rsrc(a.core, a.nvec * 16u);
    uint32_t sn = 0, vn = 0;
    for (uint32_t t = gw * 64 * U; t < a.nvec; t += GW * 64 * U) {
        uint4 d[U];
#pragma unroll
        for (uint32_t k = 0; k < U; ++k) {
            const uint32_t i = t + k * 64 + lane;
            d[k] = i < a.nvec ? load_quad_coherent(core, i) : make_uint4(kSent, kSent, kSent, kSent);
        }
#pragma unroll
        for (uint32_t k = 0; k < U; ++k) {
            const uint32_t v0 = 4 * (t + k * 64 + lane);
            select_quad(a, S, L, d[k], v0, mn, sn);
            if (compact) {
                // padding entries (index >= V) hold the sentinel and are never survivors
                surv_push(a, S, vn, d[k].x > L && v0 < a.V, v0, dst);
                surv_push(a, S, vn, d[k].y > L && v0 + 1 < a.V, v0 + 1, dst);
                surv_push(a, S, vn, d[k].z > L && v0 + 2 < a.V, v0 + 2, dst);
                surv_push(a, S, vn, d[k].w > L && v0 + 3 < a.V, v0 + 3, dst);
            }
        }
    }
    flush_stage(a, S, sn, false, false);
    if (compact) flush_surv(a, S, vn, dst);
}

// Scan of an alive list: ids are read coalesced, their degrees gathered.
template <int BLOCK>
__device__ __forceinline__ void scan_list(const PeelArgs& a, WgStateA<BLOCK>& S, uint32_t L, uint32_t& mn, const uint32_t* list,
                          uint32_t len, bool compact, uint32_t dst) {
    constexpr uint32_t WPB = BLOCK / 64;
    constexpr uint32_t U = 4;
    const uint32_t lane = lane_id();
    const uint32_t gw = S.vbid * WPB + threadIdx.x / 64;
    const uint32_t GW = S.npart * WPB;
    uint32_t sn = 0, vn = 0;
    for (uint32_t t = gw * 64 * U; t < len; t += GW * 64 * U) {
        uint32_t id[U], d[U];
        // unconditional loads so each group of U leaves back to back: entries past the end read
        // the list's first slot (the lists hold padV entries) and their gathers the read-only
        // offsets; both are then replaced by the sentinel
#pragma unroll
        for (uint32_t k = 0; k < U; ++k) {
            const uint32_t i = t + k * 64 + lane;
            const uint32_t x = ld_rlx(list + (i < len ? i : 0u));
            id[k] = i < len ? x : kSent;
        }
#pragma unroll
        for (uint32_t k = 0; k < U; ++k) {
            const uint32_t x = ld_rlx(id[k] != kSent ? a.core + id[k] : a.off);
            d[k] = id[k] != kSent ? x : kSent;
        }
#pragma unroll
        for (uint32_t k = 0; k < U; ++k) {
            const bool valid = id[k] != kSent;
            if (valid && d[k] > L && d[k] < mn) mn = d[k];
            stage_push(a, S, sn, valid && d[k] == L, id[k], false, false);
            if (compact) surv_push(a, S, vn, valid && d[k] > L, id[k], dst);
        }
    }
    flush_stage(a, S, sn, false, false);
    if (compact) flush_surv(a, S, vn, dst);
}

__device__ __forceinline__ bool level_complete(const PeelArgs& a) {
    const uint32_t vd = ld_rlx(cw(a, A_VDONE));
    const uint32_t cd = ld_rlx(cw(a, A_CDONE));
    order_point();  // the done counters are read before the tails
    const uint32_t vt = ld_rlx(cw(a, C_OTAIL));
    const uint32_t ct = ld_rlx(cw(a, C_CTAIL));
    return vd == vt && cd == ct;
}

__device__ __forceinline__ void signal_level_done(const PeelArgs& a, uint32_t gen) {
    for (uint32_t x = 0; x < kXcds; ++x) st_rlx(cw(a, A_LDONE + x), gen);
}

// After a workgroup's appends are complete: wake the idle workgroups of every XCD group, which
// poll their group's hint instead of the shared queue counters.
template <int BLOCK>
__device__ __forceinline__ void publish_hint(const PeelArgs& a, WgStateA<BLOCK>& S) {
    if (S.appended) {
        S.appended = 0;
        for (uint32_t x = 0; x < kXcds; ++x) add_rlx(cw(a, A_HINT + x), 1u);
    }
}

__device__ __forceinline__ void idle_sleep(uint32_t nap, uint32_t poll_max) {
    if (nap < 8 || poll_max == 0) {
        __builtin_amdgcn_s_sleep(2);
    } else if (nap < 32 || poll_max == 1) {
        __builtin_amdgcn_s_sleep(8);
    } else if (nap < 128 || poll_max == 2) {
        __builtin_amdgcn_s_sleep(24);
    } else {
        __builtin_amdgcn_s_sleep(64);
    }
}

// The processing phase of level L for this workgroup. First its static share of the scan's
// frontier (chunks and 64-vertex batches dealt round-robin over every wave of the grid, as in
// peel_persistent's first sub-round); then claim, process, append, count, until the level's done
// flag carries generation `gen`.
template <int BLOCK, bool PREFETCH>
__device__ __forceinline__ void process_async(const PeelArgs& a, WgStateA<BLOCK>& S, uint32_t L, uint32_t gen, uint32_t& mn) {
    constexpr uint32_t WPB = BLOCK / 64;
    const uint32_t lane = lane_id();
    const uint32_t wid = threadIdx.x / 64;
    uint32_t* ldone = cw(a, A_LDONE + S.vbid % kXcds);
    uint32_t sn = 0;
    {
        const uint32_t gw = S.vbid * WPB + wid;
        const uint32_t GW = S.npart * WPB;
        const uint32_t vbeg = ld_rlx(cw(a, A_SVBEG)), vend = ld_rlx(cw(a, A_SVEND));
        const uint32_t cbeg = ld_rlx(cw(a, A_SCBEG)), cend = ld_rlx(cw(a, A_SCEND));
        if (threadIdx.x == 0) {
            S.vcount = 0;
            S.ccount = 0;
            S.appended = 0;
        }
        __syncthreads();
        uint32_t nc = 0, nv = 0;
        for (uint32_t c = cbeg + gw; c < cend; c += GW) {
            uint64_t ch = 0;
            if (lane == 0) ch = ld_rlx64(a.chunks + c);
            ch = (uint64_t)bcast_first((uint32_t)ch) | ((uint64_t)bcast_first((uint32_t)(ch >> 32)) << 32);
            process_chunk<BLOCK, PREFETCH>(a, S, L, ch, mn, sn, false);
            ++nc;
        }
        for (uint32_t b = vbeg + 64 * gw; b < vend; b += 64 * GW) {
            const bool mine = b + lane < vend;
            const uint32_t v = mine ? ld_rlx(a.order + b + lane) : 0;
            process_vertices<BLOCK, PREFETCH>(a, S, L, mine, v, mn, sn, false);
            nv += min(64u, vend - b);
        }
        flush_stage(a, S, sn, true, false);
        drain_memory();
        if (lane == 0 && (nc | nv)) {
            __hip_atomic_fetch_add(&S.vcount, nv, __ATOMIC_RELAXED, KC_WG);
            __hip_atomic_fetch_add(&S.ccount, nc, __ATOMIC_RELAXED, KC_WG);
        }
        __syncthreads();
        if (threadIdx.x == 0) publish_hint(a, S);
        if (threadIdx.x == 0 && (S.vcount | S.ccount)) {
            if (S.vcount) add_rlx(cw(a, A_VDONE), S.vcount);
            if (S.ccount) add_rlx(cw(a, A_CDONE), S.ccount);
            order_point();
            if (level_complete(a)) signal_level_done(a, gen);
        }
        __syncthreads();
    }
    while (true) {
        if (threadIdx.x == 0) {
            uint32_t kind = 3, beg = 0, n = 0, nap = 0;
            Waiter w;
            uint32_t* hint = cw(a, A_HINT + S.vbid % kXcds);
            const bool direct = (a.flags & FLAG_DIRECT_POLL) != 0;
            uint32_t seen = direct ? 0 : ld_rlx(hint);
            bool check = true;
            while (true) {
                if (check) {
                    const uint32_t ct = ld_rlx(cw(a, C_CTAIL));
                    const uint32_t ch = ld_rlx(cw(a, A_CHEAD));
                    if (ch < ct) {
                        const uint32_t k = min(WPB, ct - ch);
                        if (cas_rlx(cw(a, A_CHEAD), ch, ch + k)) {
                            kind = 1, beg = ch, n = k;
                            break;
                        }
                        continue;
                    }
                    const uint32_t vt = ld_rlx(cw(a, C_OTAIL));
                    const uint32_t vh = ld_rlx(cw(a, A_VHEAD));
                    if (vh < vt) {
                        const uint32_t k = min((uint32_t)BLOCK, vt - vh);
                        if (cas_rlx(cw(a, A_VHEAD), vh, vh + k)) {
                            kind = 2, beg = vh, n = k;
                            break;
                        }
                        continue;
                    }
                }
                if (ld_rlx(ldone) == gen) break;
                if ((++nap & 31) == 0 && level_complete(a)) {
                    signal_level_done(a, gen);
                    break;
                }
                if (w.expired(cw(a, C_ERR), a.time_limit)) break;
                idle_sleep(nap, a.poll_max);
                if (direct) {
                    check = true;
                } else {
                    // full check when the group's hint moved, and every 64 polls regardless
                    const uint32_t h = ld_rlx(hint);
                    check = h != seen || (nap & 63) == 0;
                    seen = h;
                }
            }
            S.kind = kind;
            S.beg = beg;
            S.n = n;
        }
        __syncthreads();
        const uint32_t kind = S.kind, beg = S.beg, n = S.n;
        if (kind == 3) break;
        if (kind == 1) {
            if (wid < n) {
                uint64_t ch = 0;
                if (lane == 0) {
                    Waiter w;
                    while ((ch = ld_rlx64(a.chunks + beg + wid)) == 0) {
                        if (w.expired(cw(a, C_ERR), a.time_limit)) break;
                        __builtin_amdgcn_s_sleep(1);
                    }
                }
                ch = (uint64_t)bcast_first((uint32_t)ch) | ((uint64_t)bcast_first((uint32_t)(ch >> 32)) << 32);
                if (ch) process_chunk<BLOCK, PREFETCH>(a, S, L, ch, mn, sn, false);
            }
        } else {
            for (uint32_t j0 = wid * 64; j0 < n; j0 += BLOCK) {
                const bool mine = j0 + lane < n;
                uint32_t v = kSent;
                if (mine) {
                    Waiter w;
                    while ((v = ld_rlx(a.order + beg + j0 + lane)) == kSent) {
                        if (w.expired(cw(a, C_ERR), a.time_limit)) break;
                        __builtin_amdgcn_s_sleep(1);
                    }
                }
                process_vertices<BLOCK, PREFETCH>(a, S, L, mine && v != kSent, v, mn, sn, false);
            }
        }
        // crossings go to order[] (and the chunk list) before this work counts as done
        flush_stage(a, S, sn, true, false);
        drain_memory();
        __syncthreads();
        if (threadIdx.x == 0) {
            publish_hint(a, S);
            add_rlx(cw(a, kind == 1 ? A_CDONE : A_VDONE), n);
            order_point();
            if (level_complete(a)) signal_level_done(a, gen);
        }
        // S.kind is rewritten only by thread 0 after this barrier's readers are done
        __syncthreads();
    }
}

// after the scan: account the scan, adopt a freshly built alive list; an empty level is done
__device__ __forceinline__ void hook_scan_async(const PeelArgs& a, uint32_t L, uint32_t gen) {
    if (threadIdx.x != 0) return;
    const uint32_t mode = ld_rlx(cw(a, A_MODE));
    if (mode == 0) {
        add_rlx(cw(a, A_SCAN_DENSE), a.nvec * 4);
    } else {
        add_rlx(cw(a, A_SCAN_LIST), ld_rlx(cw(a, A_LEN)));
    }
    if (ld_rlx(cw(a, A_COMPACT))) {
        const uint32_t dst = ld_rlx(cw(a, A_DST));
        const uint32_t len = ld_rlx(cw(a, A_ALEN0 + dst));
        add_rlx(cw(a, A_SURV), len);
        st_rlx(cw(a, A_MODE), 1u);
        st_rlx(cw(a, A_LIST), dst);
        st_rlx(cw(a, A_LEN), len);
        st_rlx(cw(a, A_COMPACT), 0u);
    }
    // the scan's frontier is dealt statically; the queues open behind it
    const uint32_t vh = ld_rlx(cw(a, A_VHEAD)), vt = ld_rlx(cw(a, C_OTAIL));
    const uint32_t ch = ld_rlx(cw(a, A_CHEAD)), ct = ld_rlx(cw(a, C_CTAIL));
    st_rlx(cw(a, A_SVBEG), vh);
    st_rlx(cw(a, A_SVEND), vt);
    st_rlx(cw(a, A_SCBEG), ch);
    st_rlx(cw(a, A_SCEND), ct);
    st_rlx(cw(a, A_VHEAD), vt);
    st_rlx(cw(a, A_CHEAD), ct);
    if (level_complete(a)) signal_level_done(a, gen);
    trace_rec(a, TR_SCAN, 0, L);
}

// level end: next level, shell offsets, compaction of the next scan
template <int BLOCK>
__device__ __forceinline__ void hook_level_async(const PeelArgs& a, WgStateA<BLOCK>& S, uint32_t L) {
    uint32_t m = kSent;
    for (uint32_t i = threadIdx.x; i < S.npart; i += BLOCK) m = min(m, ld_rlx(a.wgmin + i));
    m = wave_min(m);
    if (lane_id() == 0) S.red[threadIdx.x / 64] = m;
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int w = 0; w < BLOCK / 64; ++w) m = min(m, S.red[w]);
    const uint32_t t = ld_rlx(cw(a, C_OTAIL));
    uint32_t phase = PH_NEXT_LEVEL, next = L + 1;
    if (!(a.flags & FLAG_NO_LEVEL_SKIP) && m != kSent && m > next) next = m;
    add_rlx(cw(a, C_LEVELS), 1u);
    if (t >= a.V) {
        a.shell[L + 1] = t;
        phase = PH_DONE;
        next = L;
    } else if (next > a.dmax) {
        __hip_atomic_fetch_max(cw(a, C_ERR), 2u, __ATOMIC_RELAXED, KC_AGENT);
        phase = PH_DONE;
    } else {
        for (uint32_t k = L + 1; k <= next; ++k) a.shell[k] = t;
        // the next scan compacts when few of the entries it would read are still unpeeled
        const uint32_t mode = ld_rlx(cw(a, A_MODE));
        const uint64_t reads = mode ? ld_rlx(cw(a, A_LEN)) : (uint64_t)a.nvec * 4;
        const uint64_t unpeeled = a.V - t;
        if (a.compact_ratio && unpeeled * a.compact_ratio < reads) {
            const uint32_t dst = mode ? ld_rlx(cw(a, A_LIST)) ^ 1u : 0u;
            st_rlx(cw(a, A_ALEN0 + dst), 0u);
            st_rlx(cw(a, A_DST), dst);
            st_rlx(cw(a, A_COMPACT), 1u);
        }
    }
    st_rlx(cw(a, A_PHASE), phase);
    st_rlx(cw(a, A_NEXT), next);
    trace_rec(a, TR_LEVEL, 0, L);
}

template <int BLOCK, bool PREFETCH>
__global__ __launch_bounds__(BLOCK) void peel_async(PeelArgs a) {
    __shared__ WgStateA<BLOCK> S;
    bool registered;  // calls inlined at their call sites, as in peel_persistent
    [[clang::always_inline]] registered = register_workgroup<BLOCK>(a, S);
    if (!registered) return;
    uint32_t L = 0, gen = 0;
    uint32_t mn = kSent;
    if (threadIdx.x == 0) S.mn = kSent;
    __syncthreads();
    while (true) {
        // control words written by the hooks are read here, before this workgroup arrives at the
        // next barrier (the only place they change)
        const uint32_t mode = ld_rlx(cw(a, A_MODE));
        const uint32_t list = ld_rlx(cw(a, A_LIST));
        const uint32_t len = ld_rlx(cw(a, A_LEN));
        const bool compact = ld_rlx(cw(a, A_COMPACT)) != 0;
        const uint32_t dst = ld_rlx(cw(a, A_DST));
        if (threadIdx.x == 0) {
            st_rlx(a.dbg + S.vbid * 8 + 1, L);
            st_rlx(a.dbg + S.vbid * 8 + 2, 1u);
        }
        if (mode == 0) {
            scan_dense<BLOCK>(a, S, L, mn, compact, dst);
        } else {
            scan_list<BLOCK>(a, S, L, mn, alive_list(a, list), len, compact, dst);
        }
        ++gen;
        [[clang::always_inline]] grid_barrier<BLOCK>(a, S, [&] { hook_scan_async(a, L, gen); });
        if (ld_rlx(cw(a, C_ERR))) break;
        if (threadIdx.x == 0) st_rlx(a.dbg + S.vbid * 8 + 2, 2u);
        process_async<BLOCK, PREFETCH>(a, S, L, gen, mn);
        // publish this workgroup's smallest degree above L seen in the level
        mn = wave_min(mn);
        if (lane_id() == 0) __hip_atomic_fetch_min(&S.mn, mn, __ATOMIC_RELAXED, KC_WG);
        __syncthreads();
        if (threadIdx.x == 0) st_rlx(a.wgmin + S.vbid, S.mn);
        [[clang::always_inline]] grid_barrier<BLOCK>(a, S, [&] { hook_level_async<BLOCK>(a, S, L); });
        if (ld_rlx(cw(a, C_ERR))) break;
        if (ld_rlx(cw(a, A_PHASE)) == PH_DONE) break;
        L = ld_rlx(cw(a, A_NEXT));
        mn = kSent;
        __syncthreads();
        if (threadIdx.x == 0) S.mn = kSent;
        __syncthreads();
    }
    if (S.vbid == 0 && threadIdx.x == 0) {
        uint32_t* r = cw(a, C_RESULT);
        r[R_ERR] = ld_rlx(cw(a, C_ERR));
        r[R_KMAX] = L;
        r[R_LEVELS] = ld_rlx(cw(a, C_LEVELS));
        r[R_SUBROUNDS] = 0;
        r[R_NPART] = S.npart;
        r[R_NTRACE] = ld_rlx(cw(a, C_NTRACE));
        r[R_WORDS] = ld_rlx(cw(a, A_SCAN_DENSE));
        r[R_WORDS + 1] = ld_rlx(cw(a, A_SCAN_LIST));
        r[R_WORDS + 2] = ld_rlx(cw(a, A_SURV));
        r[R_WORDS + 3] = ld_rlx(cw(a, C_CTAIL));
    }
}

// ---------------------------------------------------------------------------------------
// peel_async_nb (KC_FLAG_LEVEL_HANDOFF): the cascade-queue kernel without grid barriers.
//  * A workgroup's scan share is a work item too: after its appends are complete it adds 1 to
//    A_SDONE, so a level can only be complete once npart scans of it are counted. The scan that
//    brings the count to npart * gen raises the level's gate flags (one per XCD group).
//  * Gate: no entry of a level is processed before its gate is open, so no scan of a level ever
//    overlaps that level's decrements. (A crossing L + 1 -> L would otherwise be selected again
//    by a later scan, and the undo of a decrement below L would hide a vertex of degree L.)
//  * Every wave folds the smallest degree above L it saw into the level's A_GMIN word (as a
//    complement, so a zeroed word means "none") before its workgroup counts work done.
//  * Whoever reads the level complete (scans, entries and chunks done == npart * gen, OTAIL,
//    CTAIL; done counters first) wins it with a compare-and-swap on A_LCLAIM, computes the next
//    level, the shell offsets and the next scan's compaction, and publishes them with the level's
//    generation in the level-done flags; each workgroup starts the next scan as soon as it sees
//    its group's flag. Nobody waits for the slowest workgroup to reach a barrier.
// Every entry, the scan's frontier included, is claimed from the queues.
//  * A workgroup still in level g may claim entries that a faster one, already scanning g + 1,
//    appended. After every claim it reads A_LCLAIM: if level g has been won, nothing of g is
//    left, so the claim is of g + 1. It is held: the workgroup leaves level g once its group's
//    flag says so, scans g + 1, passes that level's gate and processes the held claim first.
//    If g has not been won, no flag says g yet, so nothing of g + 1 has been appended.

__device__ __forceinline__ bool level_complete_nb(const PeelArgs& a, uint32_t npart, uint32_t gen) {
    const uint32_t sd = ld_rlx(cw(a, A_SDONE));
    const uint32_t vd = ld_rlx(cw(a, A_VDONE));
    const uint32_t cd = ld_rlx(cw(a, A_CDONE));
    order_point();  // the done counters are read before the tails
    const uint32_t vt = ld_rlx(cw(a, C_OTAIL));
    const uint32_t ct = ld_rlx(cw(a, C_CTAIL));
    return sd == npart * gen && vd == vt && cd == ct;
}

// This wave's smallest degree above L into the level's word; mn restarts. All lanes active.
__device__ __forceinline__ void fold_min(const PeelArgs& a, uint32_t gen, uint32_t& mn) {
    const uint32_t m = wave_min(mn);
    if (lane_id() == 0 && m != kSent)
        __hip_atomic_fetch_max(cw(a, A_GMIN0 + (gen & 1u)), ~m, __ATOMIC_RELAXED, KC_AGENT);
    mn = kSent;
}

// Called by one thread that read level `gen` complete: the winner of the level publishes what
// the next level needs, then the level-done flags.
__device__ __forceinline__ void hand_off_level(const PeelArgs& a, uint32_t L, uint32_t gen) {
    if (!cas_rlx(cw(a, A_LCLAIM), gen - 1, gen)) return;
    const uint32_t t = ld_rlx(cw(a, C_OTAIL));
    const uint32_t gm = ~ld_rlx(cw(a, A_GMIN0 + (gen & 1u)));
    uint32_t phase = PH_NEXT_LEVEL, next = L + 1;
    if (!(a.flags & FLAG_NO_LEVEL_SKIP) && gm != kSent && gm > next) next = gm;
    add_rlx(cw(a, C_LEVELS), 1u);
    // this level's scan, accounted in the mode it ran in; then the alive list it built adopted
    const uint32_t mode = ld_rlx(cw(a, A_MODE));
    if (mode == 0) {
        add_rlx(cw(a, A_SCAN_DENSE), a.nvec * 4);
    } else {
        add_rlx(cw(a, A_SCAN_LIST), ld_rlx(cw(a, A_LEN)));
    }
    if (ld_rlx(cw(a, A_COMPACT))) {
        const uint32_t dst = ld_rlx(cw(a, A_DST));
        const uint32_t len = ld_rlx(cw(a, A_ALEN0 + dst));
        add_rlx(cw(a, A_SURV), len);
        st_rlx(cw(a, A_MODE), 1u);
        st_rlx(cw(a, A_LIST), dst);
        st_rlx(cw(a, A_LEN), len);
        st_rlx(cw(a, A_COMPACT), 0u);
    }
    if (t >= a.V) {
        a.shell[L + 1] = t;
        phase = PH_DONE;
        next = L;
    } else if (next > a.dmax) {
        __hip_atomic_fetch_max(cw(a, C_ERR), 2u, __ATOMIC_RELAXED, KC_AGENT);
        phase = PH_DONE;
    } else {
        for (uint32_t k = L + 1; k <= next; ++k) a.shell[k] = t;
        // the next scan compacts when few of the entries it would read are still unpeeled
        const uint32_t m2 = ld_rlx(cw(a, A_MODE));
        const uint64_t reads = m2 ? ld_rlx(cw(a, A_LEN)) : (uint64_t)a.nvec * 4;
        const uint64_t unpeeled = a.V - t;
        if (a.compact_ratio && unpeeled * a.compact_ratio < reads) {
            const uint32_t dst = m2 ? ld_rlx(cw(a, A_LIST)) ^ 1u : 0u;
            st_rlx(cw(a, A_ALEN0 + dst), 0u);
            st_rlx(cw(a, A_DST), dst);
            st_rlx(cw(a, A_COMPACT), 1u);
        }
    }
    st_rlx(cw(a, A_GMIN0 + ((gen + 1) & 1u)), 0u);  // the next level's word starts empty
    st_rlx(cw(a, A_PHASE), phase);
    st_rlx(cw(a, A_NEXT), next);
    trace_rec(a, TR_LEVEL, 0, L);
    drain_memory();  // everything above is complete before the flags say so
    signal_level_done(a, gen);
}

// Gate, then claim, process, append, fold, count — until the level-done flag carries `gen`.
template <int BLOCK, bool PREFETCH>
__device__ __forceinline__ void process_nb(const PeelArgs& a, WgStateA<BLOCK>& S, uint32_t L, uint32_t gen) {
    constexpr uint32_t WPB = BLOCK / 64;
    const uint32_t lane = lane_id();
    const uint32_t wid = threadIdx.x / 64;
    uint32_t* ldone = cw(a, A_LDONE + S.vbid % kXcds);
    uint32_t sn = 0, mn = kSent;
    if (threadIdx.x == 0) {
        // Nothing of level g is processed before every scan of g is counted: a scan must see
        // the degrees as the previous level left them. A decrement that crosses L + 1 -> L
        // would otherwise be selected again by a later scan, and the undo of a decrement below
        // L would hide a vertex of degree L from it.
        Waiter w;
        uint32_t nap = 0;
        // the last scanner of the level raises one flag per XCD group, so the pollers spread
        uint32_t* gate = cw(a, A_GATE + S.vbid % kXcds);
        while (ld_rlx(gate) != gen) {
            if (w.expired(cw(a, C_ERR), a.time_limit)) break;
            idle_sleep(++nap, a.poll_max);
        }
        order_point();
        if (ld_rlx(cw(a, C_ERR))) S.hkind = 3;
    }
    while (true) {
        if (threadIdx.x == 0) {
            uint32_t kind = 3, beg = 0, n = 0, nap = 0;
            if (S.hkind != 3) {
                // the claim this workgroup made, and held, while leaving the previous level
                kind = S.hkind, beg = S.hbeg, n = S.hn;
                S.hkind = 3;
            } else {
                Waiter w;
                uint32_t* hint = cw(a, A_HINT + S.vbid % kXcds);
                const bool direct = (a.flags & FLAG_DIRECT_POLL) != 0;
                uint32_t seen = direct ? 0 : ld_rlx(hint);
                bool check = true;
                while (true) {
                    if (check) {
                        const uint32_t ct = ld_rlx(cw(a, C_CTAIL));
                        const uint32_t ch = ld_rlx(cw(a, A_CHEAD));
                        if (ch < ct) {
                            const uint32_t k = min(WPB, ct - ch);
                            if (cas_rlx(cw(a, A_CHEAD), ch, ch + k)) {
                                kind = 1, beg = ch, n = k;
                                break;
                            }
                            continue;
                        }
                        const uint32_t vt = ld_rlx(cw(a, C_OTAIL));
                        const uint32_t vh = ld_rlx(cw(a, A_VHEAD));
                        if (vh < vt) {
                            const uint32_t k = min((uint32_t)BLOCK, vt - vh);
                            if (cas_rlx(cw(a, A_VHEAD), vh, vh + k)) {
                                kind = 2, beg = vh, n = k;
                                break;
                            }
                            continue;
                        }
                    }
                    if (ld_rlx(ldone) == gen) break;
                    if ((++nap & 31) == 0 && level_complete_nb(a, S.npart, gen)) hand_off_level(a, L, gen);
                    if (w.expired(cw(a, C_ERR), a.time_limit)) break;
                    idle_sleep(nap, a.poll_max);
                    if (direct) {
                        check = true;
                    } else {
                        // full check when the group's hint moved, and every 64 polls regardless
                        const uint32_t h = ld_rlx(hint);
                        check = h != seen || (nap & 63) == 0;
                        seen = h;
                    }
                }
                if (kind != 3) {
                    // The claim is complete before the level word is read. A_LCLAIM, not this
                    // group's flag, decides: the winner sets it before writing any flag, and the
                    // flags of the XCD groups become visible one by one, so an entry appended in
                    // g + 1 by a workgroup that saw its own flag can be claimed here while this
                    // group's flag still reads g - 1.
                    order_point();
                    if (ld_rlx(cw(a, A_LCLAIM)) == gen) {
                        // a claim of g + 1: held until this workgroup has scanned g + 1 and the
                        // gate of g + 1 is passed; the level ends here once this group's flag,
                        // written after the next level's parameters, says so
                        S.hkind = kind, S.hbeg = beg, S.hn = n;
                        add_rlx(cw(a, A_HELD), 1u);
                        kind = 3;
                        Waiter w2;
                        while (ld_rlx(ldone) != gen) {
                            if (w2.expired(cw(a, C_ERR), a.time_limit)) break;
                            __builtin_amdgcn_s_sleep(1);
                        }
                    }
                }
            }
            S.kind = kind;
            S.beg = beg;
            S.n = n;
        }
        __syncthreads();
        const uint32_t kind = S.kind, beg = S.beg, n = S.n;
        if (kind == 3) break;
        if (kind == 1) {
            if (wid < n) {
                uint64_t ch = 0;
                if (lane == 0) {
                    Waiter w;
                    while ((ch = ld_rlx64(a.chunks + beg + wid)) == 0) {
                        if (w.expired(cw(a, C_ERR), a.time_limit)) break;
                        __builtin_amdgcn_s_sleep(1);
                    }
                }
                ch = (uint64_t)bcast_first((uint32_t)ch) | ((uint64_t)bcast_first((uint32_t)(ch >> 32)) << 32);
                if (ch) process_chunk<BLOCK, PREFETCH>(a, S, L, ch, mn, sn, false);
            }
        } else {
            for (uint32_t j0 = wid * 64; j0 < n; j0 += BLOCK) {
                const bool mine = j0 + lane < n;
                uint32_t v = kSent;
                if (mine) {
                    Waiter w;
                    while ((v = ld_rlx(a.order + beg + j0 + lane)) == kSent) {
                        if (w.expired(cw(a, C_ERR), a.time_limit)) break;
                        __builtin_amdgcn_s_sleep(1);
                    }
                }
                process_vertices<BLOCK, PREFETCH>(a, S, L, mine && v != kSent, v, mn, sn, false);
            }
        }
        // crossings go to order[] (and the chunk list), and the minimum to the level's word,
        // before this work counts as done
        flush_stage(a, S, sn, true, false);
        fold_min(a, gen, mn);
        drain_memory();
        __syncthreads();
        if (threadIdx.x == 0) {
            publish_hint(a, S);
            add_rlx(cw(a, kind == 1 ? A_CDONE : A_VDONE), n);
            order_point();
            if (level_complete_nb(a, S.npart, gen)) hand_off_level(a, L, gen);
        }
        // S.kind is rewritten only by thread 0 after this barrier's readers are done
        __syncthreads();
    }
}

template <int BLOCK, bool PREFETCH>
__global__ __launch_bounds__(BLOCK) void peel_async_nb(PeelArgs a) {
    __shared__ WgStateA<BLOCK> S;
    bool registered;  // calls inlined at their call sites, as in peel_persistent
    [[clang::always_inline]] registered = register_workgroup<BLOCK>(a, S);
    if (!registered) return;
    if (threadIdx.x == 0) {
        S.appended = 0;
        S.hkind = 3;
    }
    __syncthreads();
    uint32_t L = 0, gen = 0;
    while (true) {
        ++gen;
        // published with the previous level's done flag (or zero before the first level)
        const uint32_t mode = ld_rlx(cw(a, A_MODE));
        const uint32_t list = ld_rlx(cw(a, A_LIST));
        const uint32_t len = ld_rlx(cw(a, A_LEN));
        const bool compact = ld_rlx(cw(a, A_COMPACT)) != 0;
        const uint32_t dst = ld_rlx(cw(a, A_DST));
        uint32_t mn = kSent;
        if (mode == 0) {
            scan_dense<BLOCK>(a, S, L, mn, compact, dst);
        } else {
            scan_list<BLOCK>(a, S, L, mn, alive_list(a, list), len, compact, dst);
        }
        fold_min(a, gen, mn);
        drain_memory();
        __syncthreads();
        if (threadIdx.x == 0) {
            publish_hint(a, S);
            // the scan that completes the level's count opens the gate of every XCD group
            if (add_rlx(cw(a, A_SDONE), 1u) + 1 == S.npart * gen)
                for (uint32_t x = 0; x < kXcds; ++x) st_rlx(cw(a, A_GATE + x), gen);
            order_point();
            if (level_complete_nb(a, S.npart, gen)) hand_off_level(a, L, gen);
        }
        __syncthreads();
        process_nb<BLOCK, PREFETCH>(a, S, L, gen);
        if (ld_rlx(cw(a, C_ERR))) break;
        if (ld_rlx(cw(a, A_PHASE)) == PH_DONE) break;
        L = ld_rlx(cw(a, A_NEXT));
        __syncthreads();
    }
    if (S.vbid == 0 && threadIdx.x == 0) {
        uint32_t* r = cw(a, C_RESULT);
        r[R_ERR] = ld_rlx(cw(a, C_ERR));
        r[R_KMAX] = L;
        r[R_LEVELS] = ld_rlx(cw(a, C_LEVELS));
        r[R_SUBROUNDS] = 0;
        r[R_NPART] = S.npart;
        r[R_NTRACE] = ld_rlx(cw(a, C_NTRACE));
        r[R_WORDS] = ld_rlx(cw(a, A_SCAN_DENSE));
        r[R_WORDS + 1] = ld_rlx(cw(a, A_SCAN_LIST));
        r[R_WORDS + 2] = ld_rlx(cw(a, A_SURV));
        r[R_WORDS + 3] = ld_rlx(cw(a, C_CTAIL));
        r[R_WORDS + 4] = ld_rlx(cw(a, A_HELD));
    }
}

}  // namespace kcdev
