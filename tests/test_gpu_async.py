"""GPU parity of the cascade-queue variant (KC_VARIANT_ASYNC) across its own settings: alive-list
compaction thresholds, idle poll steps, workgroup sizes, chunk thresholds and the reference level
loop. Every setting must give the oracle's core numbers and a valid peel order."""
import sys

import numpy as np
import pytest

import kcoregpu_amd as K
import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def check(g, dg, core, kmax, **kw):
    r = dg.peel("async", **kw)
    assert r.kmax == kmax, kw
    assert (r.core == core).all(), (kw, int(np.flatnonzero(r.core != core)[0]))
    order, shells = dg.order()
    assert O.check_certificate(g.neighbors_offset, g.neighbors, r.core, order) == 0, kw
    assert shells[0] == 0 and shells[-1] == g.V
    return r


def test_async_er_settings(engine):
    say("async: ER n=100000 m=1000000")
    g = K.Graph.erdos_renyi(100000, 1000000, 1)
    core, kmax = O.bz(g.neighbors_offset, g.neighbors)
    dg = engine.upload(g)
    for ratio in (1, 4, 64, 0xFFFFFFFF):
        check(g, dg, core, kmax, compact_ratio=ratio)
    for pm in (1, 2, 3):
        check(g, dg, core, kmax, poll_max=pm)
    check(g, dg, core, kmax, flags=K.FLAG_DIRECT_POLL)
    for block in (256, 512, 1024):
        check(g, dg, core, kmax, block=block)
    r = check(g, dg, core, kmax, flags=K.FLAG_NO_LEVEL_SKIP)
    assert r.stats["levels"] == kmax + 1
    dg.free()


def test_async_power_law_and_chunks(engine):
    say("async: Chung-Lu n=400000")
    g = K.Graph.chung_lu(400000, 2)
    core, kmax = O.bz(g.neighbors_offset, g.neighbors)
    dg = engine.upload(g)
    for h, c in ((16, 64), (64, 512), (256, 1000), (0xFFFFFFFF, 0)):
        check(g, dg, core, kmax, heavy_threshold=h, chunk_arcs=c)
    check(g, dg, core, kmax, compact_ratio=1, heavy_threshold=16, chunk_arcs=64, block=1024)
    dg.free()


def test_async_rmat20_shells(engine):
    say("async: R-MAT scale 20")
    g = K.Graph.rmat(20, 3)
    ls = O.levelsync(g.neighbors_offset, g.neighbors, want_order=True)
    dg = engine.upload(g)
    for ratio in (2, 4, 16):
        r = check(g, dg, ls["core"], ls["kmax"], compact_ratio=ratio)
        _, shells = dg.order()
        assert (shells == ls["shells"]).all()
    assert r.stats["scanned_entries"] < r.stats["levels"] * g.V  # the alive list shortened the scans
    dg.free()


# The level hand-off form of the kernel (KC_FLAG_LEVEL_HANDOFF, no grid barriers) has been checked
# only on the CPU (the wave64 emulator, with weak memory, random pauses and AddressSanitizer); it
# joins the GPU suite once a first run on the GPU has passed.
HANDOFF_ON_GPU = False


@pytest.mark.skipif(not HANDOFF_ON_GPU, reason="level hand-off kernel not yet run on a GPU")
@pytest.mark.parametrize("name", ["er", "chung_lu", "rmat20"])
def test_level_handoff_parity(engine, name):
    say("level hand-off:", name)
    g = {"er": lambda: K.Graph.erdos_renyi(100000, 1000000, 1),
         "chung_lu": lambda: K.Graph.chung_lu(400000, 2),
         "rmat20": lambda: K.Graph.rmat(20, 3)}[name]()
    ls = O.levelsync(g.neighbors_offset, g.neighbors, want_order=True)
    dg = engine.upload(g)
    ho = K.FLAG_LEVEL_HANDOFF
    for kw in (dict(flags=ho), dict(flags=ho, compact_ratio=1), dict(flags=ho, block=256, heavy_threshold=16,
                                                                     chunk_arcs=64),
               dict(flags=ho | K.FLAG_DIRECT_POLL, block=1024)):
        check(g, dg, ls["core"], ls["kmax"], **kw)
        _, shells = dg.order()
        assert (shells == ls["shells"]).all(), kw
    r = check(g, dg, ls["core"], ls["kmax"], flags=ho | K.FLAG_NO_LEVEL_SKIP)
    assert r.stats["levels"] == ls["kmax"] + 1
    dg.free()
