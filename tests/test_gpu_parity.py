"""GPU parity: the HIP peel (through the C ABI) against the oracle, bit-exact.

Small graphs are pinned by closed forms, brute force and networkx fixtures; the configured
workloads (ER 100K/1M, Chung-Lu 400K, R-MAT) by the Batagelj-Zaversnik oracle and the
level-synchronous restatement, plus the peel-order certificate (size-independent)."""
import os
import sys

import numpy as np
import pytest

import kcoregpu_amd as K
import oracle as O
from conftest import closed_form_graphs, csr_from_edges

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
# variant families: the sub-round kernel's three forms, the cascade-queue kernel, and the
# reference-shaped host level loop; the workload tests run once per family
FAMILIES = {"subround": ["wave", "lds", "prefetch"], "async": ["async"], "stepwise": ["stepwise"]}
VARIANTS = [v for f in FAMILIES.values() for v in f]


@pytest.fixture(params=list(FAMILIES))
def family(request):
    return request.param


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def peel_all(engine, g, variants, **kw):
    dg = engine.upload(g)
    out = {}
    for v in variants:
        r = dg.peel(v, **kw)
        order, shells = dg.order()
        out[v] = (r, order, shells)
    dg.free()
    return out


def check_against(g, res, core, kmax):
    for v, (r, order, shells) in res.items():
        assert r.kmax == kmax, v
        assert (r.core == core).all(), (v, int(np.flatnonzero(r.core != core)[0]))
        assert O.check_certificate(g.neighbors_offset, g.neighbors, r.core, order) == 0, v
        assert shells[0] == 0 and shells[-1] == g.V
        for k in (0, kmax // 2, kmax):
            assert (r.core[order[shells[k]:shells[k + 1]]] == k).all()


def test_copy_ceiling(engine):
    gbps = engine.copy_bandwidth(256 << 20, 5)
    say(f"copy kernel: {gbps:.0f} GB/s")
    assert 100 < gbps < 20000  # a plausible HBM copy rate, read + write


def test_selftest(engine):
    engine.selftest()


@pytest.mark.parametrize("case", closed_form_graphs(), ids=lambda c: c[0])
def test_closed_forms(engine, case, family):
    name, n, edges, expect = case
    say("closed form", name, family)
    off, nbr = csr_from_edges(n, edges)
    g = K.Graph.from_csr(off, nbr)
    expect = np.array(expect, np.uint32)
    check_against(g, peel_all(engine, g, FAMILIES[family]), expect, int(expect.max()))


@pytest.mark.parametrize("seed", range(24))
def test_random_small_vs_bruteforce(engine, seed, family):
    rng = np.random.default_rng(100 + seed)
    n = 5 + seed % 12
    p = [0.2, 0.4, 0.6][seed % 3]
    edges = [(i, j) for i in range(n) for j in range(i + 1, n) if rng.random() < p]
    off, nbr = csr_from_edges(n, edges)
    g = K.Graph.from_csr(off, nbr)
    bf, kb = O.bruteforce(off, nbr)
    check_against(g, peel_all(engine, g, FAMILIES[family]), bf, kb)


def test_networkx_fixtures(engine, family):
    z = np.load(f"{GOLDEN}/nx_cores.npz")
    for i in range(int(z["count"])):
        say("nx fixture", i, family)
        g = K.Graph.from_csr(z[f"off{i}"], z[f"nbr{i}"])
        expect = z[f"core{i}"]
        check_against(g, peel_all(engine, g, FAMILIES[family]), expect, int(expect.max()))


def test_er_100k_1m_all_variants(engine, family):
    say("ER n=100000 m=1000000 seed 1", family)
    g = K.Graph.erdos_renyi(100000, 1000000, 1)
    core, kmax = O.bz(g.neighbors_offset, g.neighbors)
    ls = O.levelsync(g.neighbors_offset, g.neighbors)
    assert (ls["core"] == core).all()
    check_against(g, peel_all(engine, g, FAMILIES[family]), core, kmax)
    if family != "subround":
        return
    # reference level loop (no level skipping) and other workgroup sizes give the same result
    res = peel_all(engine, g, ["wave"], flags=K.FLAG_NO_LEVEL_SKIP)
    check_against(g, res, core, kmax)
    assert res["wave"][0].stats["levels"] == kmax + 1
    for block in (512, 1024):
        check_against(g, peel_all(engine, g, ["wave"], block=block), core, kmax)


def test_chung_lu_400k_variants(engine, family):
    say("Chung-Lu n=400000 gamma=2.1 seed 2", family)
    g = K.Graph.chung_lu(400000, 2)
    core, kmax = O.bz(g.neighbors_offset, g.neighbors)
    check_against(g, peel_all(engine, g, FAMILIES[family]), core, kmax)
    if family != "subround":
        return
    # hub arcs cut into chunks at the smallest threshold, and never cut
    check_against(g, peel_all(engine, g, ["wave"], heavy_threshold=16, chunk_arcs=64), core, kmax)
    check_against(g, peel_all(engine, g, ["lds"], heavy_threshold=256, chunk_arcs=1000), core, kmax)
    check_against(g, peel_all(engine, g, ["wave"], heavy_threshold=0xFFFFFFFF), core, kmax)


def test_star_and_hub_chunks(engine):
    # a hub of degree 50,000 inside a clique: chunked arcs, crossing at several levels
    n_leaf = 50000
    src = [0] * n_leaf + [i for i in range(1, 9) for j in range(i + 1, 10)]
    dst = list(range(10, 10 + n_leaf)) + [j for i in range(1, 9) for j in range(i + 1, 10)]
    src += [0] * 9
    dst += list(range(1, 10))
    g = K.Graph.from_pairs(np.array(src), np.array(dst))
    core, kmax = O.bz(g.neighbors_offset, g.neighbors)
    assert kmax == 9
    for h, c in ((16, 64), (64, 512), (1000, 4096), (0xFFFFFFFF, 0)):
        check_against(g, peel_all(engine, g, ["wave", "prefetch"], heavy_threshold=h, chunk_arcs=c),
                      core, kmax)


def test_degenerate_inputs(engine, family):
    for name, (off, nbr) in {
        "one_vertex": (np.array([0, 0], np.uint32), np.zeros(0, np.uint32)),
        "edgeless_1000": (np.zeros(1001, np.uint32), np.zeros(0, np.uint32)),
        "single_edge": csr_from_edges(2, [(0, 1)]),
        "ragged_257": csr_from_edges(257, [(i, i + 1) for i in range(256)]),
    }.items():
        say("degenerate", name)
        g = K.Graph.from_csr(off, nbr)
        core, kmax = O.bz(off, nbr)
        check_against(g, peel_all(engine, g, FAMILIES[family]), core, kmax)


def test_repeat_peels_identical(engine):
    g = K.Graph.rmat(16, 17)
    core, kmax = O.bz(g.neighbors_offset, g.neighbors)
    dg = engine.upload(g)
    for i in range(5):
        r = dg.peel("wave")
        assert r.kmax == kmax and (r.core == core).all(), i
    # the uploaded degrees are untouched: a peel after a different variant still matches
    for v in ("prefetch", "lds", "wave"):
        r = dg.peel(v)
        assert (r.core == core).all() and r.kmax == kmax, v
    dg.free()


def test_repeat_peels_alternating_with_async(engine):
    """Variants alternate freely on one graph: the cascade-queue kernel re-clears the queue slots
    and chunk slots that earlier peels wrote."""
    g = K.Graph.rmat(16, 17)
    core, kmax = O.bz(g.neighbors_offset, g.neighbors)
    dg = engine.upload(g)
    for v in ("prefetch", "async", "lds", "async", "async", "wave", "stepwise", "async"):
        r = dg.peel(v)
        assert (r.core == core).all() and r.kmax == kmax, v
    dg.free()


def test_rmat20_levels_and_shells(engine, family):
    say("R-MAT scale 20 seed 3", family)
    g = K.Graph.rmat(20, 3)
    ls = O.levelsync(g.neighbors_offset, g.neighbors, want_order=True)
    core, kmax = ls["core"], ls["kmax"]
    res = peel_all(engine, g, FAMILIES[family])
    check_against(g, res, core, kmax)
    for v, (r, order, shells) in res.items():
        assert (shells == ls["shells"]).all(), v  # level boundaries are schedule-independent


@pytest.mark.slow
def test_rmat24_full_size(engine):
    say("R-MAT scale 24 edge factor 16 seed 3: generating")
    g = K.Graph.rmat(24, 3)
    say("V", g.V, "m", g.m, "dmax", g.dmax)
    dg = engine.upload(g)
    r = dg.peel("auto")
    order, shells = dg.order()
    say("peel ms", r.stats["ms_peel"], "kmax", r.kmax)
    assert O.check_certificate(g.neighbors_offset, g.neighbors, r.core, order) == 0
    core, kmax = O.bz(g.neighbors_offset, g.neighbors)
    assert r.kmax == kmax and (r.core == core).all()
    r2 = dg.peel("auto", flags=K.FLAG_NO_LEVEL_SKIP)
    assert (r2.core == core).all() and r2.stats["levels"] == kmax + 1
    dg.free()


@pytest.mark.late  # its candidates include the cascade-queue variant
def test_tuner_picks_valid_params_and_caches(engine, tmp_path):
    g = K.Graph.chung_lu(100000, 31)
    ref, km = O.bz(g.neighbors_offset, g.neighbors)
    dg = engine.upload(g)
    cache = str(tmp_path / "tune.txt")
    t = dg.tune(cache_path=cache)
    assert t["trials"] >= 10 and t["block"] in (256, 512, 1024)
    t2 = dg.tune(cache_path=cache)
    assert t2["trials"] == 0
    keys = ("variant", "block", "heavy_threshold", "chunk_arcs", "lds_depth", "compact_ratio", "poll_max")
    assert {k: t2[k] for k in keys} == {k: t[k] for k in keys}
    r = dg.peel(**{k: t[k] for k in keys})
    assert (r.core == ref).all() and r.kmax == km
    dg.free()
