"""GPU parity of batched peeling: every graph of a batch, peeled by the one persistent batch
kernel, matches the oracle and the single-graph kernel bit for bit; malformed input is rejected."""
import ctypes as C
import sys

import numpy as np
import pytest

import kcoregpu_amd as K
import oracle as O
from kcoregpu_amd import _native as N
from conftest import closed_form_graphs, csr_from_edges

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def check_batch(b, core, kmax, order=None):
    for g in range(b.G):
        gg = b.graph(g)
        ref, km = O.bz(gg.neighbors_offset, gg.neighbors)
        got = core[b.vertices(g)]
        assert (got == ref).all(), (g, int(np.flatnonzero(got != ref)[0]))
        assert kmax[g] == km, g
        if order is not None:
            assert O.check_certificate(gg.neighbors_offset, gg.neighbors, got, order[b.vertices(g)]) == 0, g


def test_batch_generated_mixed(engine):
    say("batch: 160 graphs, n in [200, 20000], deg in [4, 32]")
    b = K.Batch.generate(160, 4, nmin=200, nmax=20000)
    db = engine.upload_batch(b)
    for block in (256, 512, 1024):
        core, kmax, st = db.peel(block=block)
        check_batch(b, core, kmax, db.order())
    core2, kmax2, st2 = db.peel(flags=K.FLAG_NO_LEVEL_SKIP)
    assert (core2 == core).all() and (kmax2 == kmax).all()
    assert st2["levels"] >= st["levels"]
    db.free()


def test_batch_equals_single_graph_kernel(engine):
    b = K.Batch.generate(24, 9, nmin=1000, nmax=30000)
    db = engine.upload_batch(b)
    core, kmax, _ = db.peel()
    for g in range(b.G):
        dg = engine.upload(b.graph(g))
        r = dg.peel()
        assert (r.core == core[b.vertices(g)]).all() and r.kmax == kmax[g], g
        dg.free()
    db.free()


def test_batch_closed_forms_and_hubs(engine):
    graphs, expect = [], []
    for name, n, edges, ex in closed_form_graphs():
        off, nbr = csr_from_edges(n, edges)
        graphs.append(K.Graph.from_csr(off, nbr))
        expect.append(np.array(ex, np.uint32))
    # a hub of degree 60,000 inside a 12-clique: one tile holds it, all waves share its arcs
    n_leaf = 60000
    src = [0] * n_leaf + [i for i in range(12) for j in range(i + 1, 12)]
    dst = list(range(12, 12 + n_leaf)) + [j for i in range(12) for j in range(i + 1, 12)]
    hub = K.Graph.from_pairs(np.array(src), np.array(dst))
    graphs.append(hub)
    expect.append(O.bz(hub.neighbors_offset, hub.neighbors)[0])
    b = K.Batch.from_graphs(graphs)
    db = engine.upload_batch(b)
    core, kmax, _ = db.peel()
    for g, ex in enumerate(expect):
        assert (core[b.vertices(g)] == ex).all(), g
        assert kmax[g] == ex.max()
    check_batch(b, core, kmax, db.order())
    db.free()


def test_malformed_inputs_rejected(engine):
    # a neighbour id out of range must be refused at upload, before any peel kernel runs
    g = K.Graph.erdos_renyi(1000, 5000, 3)
    nbr = g.neighbors.copy()
    nbr[17] = 1000
    h = N.kc_csr(g.V, g.arcs, g.neighbors_offset.ctypes.data_as(N.u32p), nbr.ctypes.data_as(N.u32p),
                 g.degrees.ctypes.data_as(N.u32p))
    out = C.c_void_p()
    assert N.hip().kc_graph_upload(engine._c, C.byref(h), C.byref(out)) == 1  # KC_ERR_ARG
    b = K.Batch.generate(8, 2, nmin=100, nmax=500)
    bad = b.neighbors.copy()
    g3 = slice(int(b.arc_off[3]), int(b.arc_off[4]))
    bad[g3.start] = int(b.vtx_off[4] - b.vtx_off[3])  # = V_3: one past the graph's last vertex
    hb = N.kc_batch_csr(b.G, b.total_vertices, b.total_arcs, b.vtx_off.ctypes.data_as(N.u64p),
                        b.arc_off.ctypes.data_as(N.u64p), b.offsets.ctypes.data_as(N.u32p),
                        bad.ctypes.data_as(N.u32p), b.degrees.ctypes.data_as(N.u32p))
    assert N.hip().kc_batch_upload(engine._c, C.byref(hb), C.byref(out)) == 1
    vo = b.vtx_off.copy()
    vo[2], vo[3] = vo[3], vo[2]  # ranges out of order: refused on the host
    hb.vtx_off = vo.ctypes.data_as(N.u64p)
    hb.neighbors = b.neighbors.ctypes.data_as(N.u32p)
    assert N.hip().kc_batch_upload(engine._c, C.byref(hb), C.byref(out)) == 1


def test_batch_graph_replays_match_eager(engine):
    """The captured launch sequence replays correctly many times, with other eager work between
    the replays (no memset node is captured; the state reset is a fill kernel)."""
    b = K.Batch.generate(64, 21, nmin=500, nmax=20000)
    other = K.Batch.generate(16, 22, nmin=500, nmax=5000)
    db, do = engine.upload_batch(b), engine.upload_batch(other)
    core, kmax, _ = db.peel()
    ocore, okmax, _ = do.peel()
    for block in (512, 1024):
        for rep in range(4):
            c2, k2, st = db.peel(block=block, flags=K.FLAG_GRAPH)
            assert (c2 == core).all() and (k2 == kmax).all(), (block, rep)
            assert st["ms_kernel"] == st["ms_peel"]
            c3, k3, _ = do.peel()
            assert (c3 == ocore).all() and (k3 == okmax).all(), (block, rep)
    check_batch(b, core, kmax, db.order())
    db.free()
    do.free()


def test_batch_tuner_and_cache(engine, tmp_path):
    b = K.Batch.generate(48, 23, nmin=500, nmax=20000)
    db = engine.upload_batch(b)
    cache = str(tmp_path / "tune.txt")
    t = db.tune(cache_path=cache)
    assert t["block"] in (256, 512, 1024) and t["trials"] == 3 and t["ms"] > 0
    t2 = db.tune(cache_path=cache)
    assert t2["trials"] == 0 and t2["block"] == t["block"]
    core, kmax, _ = db.peel(block=t["block"])
    check_batch(b, core, kmax)
    db.free()
