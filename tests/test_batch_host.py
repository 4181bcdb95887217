"""Batch generator and packer (CPU): determinism, layout invariants, round trip."""
import numpy as np

import kcoregpu_amd as K


def test_batch_deterministic_across_threads():
    a = K.Batch.generate(40, 11, nmin=100, nmax=3000, nthreads=1)
    b = K.Batch.generate(40, 11, nmin=100, nmax=3000, nthreads=6)
    assert a.G == b.G == 40
    for x, y in ((a.vtx_off, b.vtx_off), (a.arc_off, b.arc_off), (a.offsets, b.offsets),
                 (a.neighbors, b.neighbors), (a.degrees, b.degrees)):
        assert (x == y).all()


def test_batch_layout_invariants():
    b = K.Batch.generate(60, 5, nmin=50, nmax=4000, dmin=4, dmax=32)
    assert b.vtx_off[0] == 0 and b.arc_off[0] == 0
    for g in range(b.G):
        V = int(b.vtx_off[g + 1] - b.vtx_off[g])
        A = int(b.arc_off[g + 1] - b.arc_off[g])
        off = b.offsets[b.vtx_off[g] + g: b.vtx_off[g] + g + V + 1]
        nbr = b.neighbors[b.arc_off[g]: b.arc_off[g + 1]]
        assert off[0] == 0 and off[-1] == A and (np.diff(off.astype(np.int64)) >= 0).all()
        assert (nbr < V).all()
        assert (b.degrees[b.vertices(g)] == np.diff(off)).all()
        gg = b.graph(g)
        assert gg.V == V and (gg.neighbors_offset == off).all() and (gg.neighbors == nbr).all()
        # n is log-uniform in [nmin, nmax] (V can be below n: trailing isolated ids do not load)
        assert V <= 4000
    # even graphs are Erdos-Renyi, odd ones Chung-Lu: the power law shows in the max degree
    er = [b.graph(g).dmax / max(1, b.graph(g).arcs / b.graph(g).V) for g in range(0, b.G, 2)]
    cl = [b.graph(g).dmax / max(1, b.graph(g).arcs / b.graph(g).V) for g in range(1, b.G, 2)]
    assert np.median(cl) > 2 * np.median(er)


def test_batch_from_graphs_roundtrip():
    gs = [K.Graph.erdos_renyi(300 + 50 * i, 900 + 100 * i, i) for i in range(6)]
    gs.append(K.Graph.from_csr(np.array([0, 0], np.uint32), np.zeros(0, np.uint32)))
    b = K.Batch.from_graphs(gs)
    assert b.G == 7
    for i, g in enumerate(gs):
        h = b.graph(i)
        assert h.V == g.V and (h.neighbors == g.neighbors).all()
