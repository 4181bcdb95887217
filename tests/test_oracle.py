"""Pins the CPU oracle: closed forms, the definition-level brute force, networkx fixtures,
cross-agreement of the two independent peels, the certificate checker, and that planted
mistakes are caught."""
import numpy as np
import pytest

import kcoregpu_amd as K
import oracle as O
from conftest import closed_form_graphs, csr_from_edges

GOLDEN = __import__("os").path.join(__import__("os").path.dirname(__file__), "golden")


@pytest.mark.parametrize("case", closed_form_graphs(), ids=lambda c: c[0])
def test_closed_forms(case):
    name, n, edges, expect = case
    off, nbr = csr_from_edges(n, edges)
    expect = np.array(expect, np.uint32)
    core, kmax = O.bz(off, nbr)
    assert (core == expect).all(), name
    assert kmax == expect.max()
    r = O.levelsync(off, nbr, want_order=True)
    assert (r["core"] == expect).all(), name
    assert r["kmax"] == expect.max()
    assert r["levels"] == expect.max() + 1
    assert O.check_certificate(off, nbr, r["core"], r["order"]) == 0
    if n <= 16:
        bf, _ = O.bruteforce(off, nbr)
        assert (bf == expect).all(), name


def _random_small(seed, n, p):
    rng = np.random.default_rng(seed)
    edges = [(i, j) for i in range(n) for j in range(i + 1, n) if rng.random() < p]
    return csr_from_edges(n, edges)


@pytest.mark.parametrize("seed", range(60))
def test_bruteforce_agrees(seed):
    n = 6 + seed % 9
    p = [0.15, 0.3, 0.5, 0.7][seed % 4]
    off, nbr = _random_small(seed, n, p)
    bf, kb = O.bruteforce(off, nbr)
    core, k = O.bz(off, nbr)
    r = O.levelsync(off, nbr, nthreads=1 + seed % 4, want_order=True)
    assert (bf == core).all() and (bf == r["core"]).all()
    assert kb == k == r["kmax"]
    assert O.check_certificate(off, nbr, core, r["order"]) == 0


def test_networkx_fixtures():
    z = np.load(f"{GOLDEN}/nx_cores.npz")
    for i in range(int(z["count"])):
        off, nbr, expect = z[f"off{i}"], z[f"nbr{i}"], z[f"core{i}"]
        core, kmax = O.bz(off, nbr)
        assert (core == expect).all()
        r = O.levelsync(off, nbr)
        assert (r["core"] == expect).all() and r["kmax"] == expect.max()


@pytest.mark.parametrize("gen", ["er", "cl", "rmat"])
def test_bz_vs_levelsync_medium(gen):
    if gen == "er":
        g = K.Graph.erdos_renyi(20000, 150000, 11)
    elif gen == "cl":
        g = K.Graph.chung_lu(30000, 12)
    else:
        g = K.Graph.rmat(14, 13)
    core, kmax = O.bz(g.neighbors_offset, g.neighbors)
    r = O.levelsync(g.neighbors_offset, g.neighbors, want_order=True, want_stats=True)
    assert (core == r["core"]).all() and kmax == r["kmax"]
    assert O.check_certificate(g.neighbors_offset, g.neighbors, core, r["order"]) == 0
    # the shells partition the order and each shell holds exactly the vertices of that core
    sh = r["shells"]
    assert sh[0] == 0 and sh[-1] == g.V
    for k in range(kmax + 1):
        assert (core[r["order"][sh[k]:sh[k + 1]]] == k).all()
    assert sum(s["peeled"] for s in r["stats"]) == g.V
    assert (core <= g.degrees).all()


def test_planted_mistakes_fail_checks():
    g = K.Graph.chung_lu(5000, 21, avg_deg=10.0)
    off, nbr = g.neighbors_offset, g.neighbors
    core, kmax = O.bz(off, nbr)
    r = O.levelsync(off, nbr, want_order=True)
    assert O.check_certificate(off, nbr, core, r["order"]) == 0
    rng = np.random.default_rng(0)
    for _ in range(20):
        bad = core.copy()
        v = int(rng.integers(g.V))
        while g.degrees[v] == 0:
            v = int(rng.integers(g.V))
        bad[v] = bad[v] + 1 if rng.random() < 0.5 or bad[v] == 0 else bad[v] - 1
        assert O.check_certificate(off, nbr, bad, r["order"]) != 0
    # a whole-shell under-estimate (the triangle-at-1 failure mode) is caught by the order test
    tri_off, tri_nbr = csr_from_edges(3, [(0, 1), (1, 2), (0, 2)])
    assert O.check_lower(tri_off, tri_nbr, np.array([1, 1, 1], np.uint32)) == 0
    assert O.check_certificate(tri_off, tri_nbr, np.array([1, 1, 1], np.uint32),
                               np.array([0, 1, 2], np.uint32)) != 0
    # a permutation error in the order is caught
    bad_order = r["order"].copy()
    bad_order[0] = bad_order[1]
    assert O.check_certificate(off, nbr, core, bad_order) == 2


def test_levelsync_thread_counts_agree():
    g = K.Graph.rmat(13, 5)
    ref = O.levelsync(g.neighbors_offset, g.neighbors, nthreads=1)["core"]
    for nt in (2, 3, 8):
        assert (O.levelsync(g.neighbors_offset, g.neighbors, nthreads=nt)["core"] == ref).all()
