"""The Python package's GPU API (Engine, DeviceGraph, DeviceBatch, Server, tuners) over the CPU
build of libkcore_hip (the library's host sources against tests/hostemu's stand-in runtime, kernels
on the wave64 emulator), with every result checked against the oracle. Test infrastructure, run by
tests/test_hostemu.py in a process of its own:

    python py_api.py <emulated libkcore_hip.so> <scratch dir>

The path of the shared library is swapped in this process only, before the package loads it; the
package itself has no switch for it."""
import gc
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from kcoregpu_amd import _native as N  # noqa: E402

N.HIP_LIB = sys.argv[1]
import numpy as np  # noqa: E402

import kcoregpu_amd as K  # noqa: E402
import oracle as O  # noqa: E402

SCRATCH = sys.argv[2]


def say(*a):
    print(*a, flush=True)


def same(core, g):
    ref, km = O.bz(g.neighbors_offset, g.neighbors)
    return (core == ref).all(), km


def single_graph(eng):
    g = K.Graph.erdos_renyi(2000, 9000, 3)
    dg = eng.upload(g)
    for v in K.VARIANTS:
        r = dg.peel(v)
        ok, km = same(r.core, g)
        assert ok and r.kmax == km, v
        order, shells = dg.order()
        assert O.check_certificate(g.neighbors_offset, g.neighbors, r.core, order) == 0, v
        assert shells[0] == 0 and shells[-1] == g.V, v
    r = dg.peel("lds", flags=K.FLAG_TRACE)
    tr = dg.trace()
    assert tr and {k for k, _, _, _ in tr} <= {"scan", "sub", "level"}
    assert (dg.core() == r.core).all()
    t = dg.tune(reps=1, cache_path=os.path.join(SCRATCH, "tune.txt"))
    ms, trials = t.pop("ms"), t.pop("trials")
    assert trials > 10 and ms > 0
    t.pop("flags")
    variant = {v: k for k, v in K.VARIANTS.items()}[t.pop("variant")]
    r = dg.peel(variant, **t)
    assert same(r.core, g)[0]
    dg.free()
    r = K.kcore(K.Graph.rmat(10, 2), engine=eng)
    assert r.kmax > 0
    say("single graph: variants, orders, trace, tuner, kcore()")


def batches(eng):
    graphs = [K.Graph.erdos_renyi(300 + 50 * i, 1200 + 300 * i, 10 + i) for i in range(6)]
    graphs.append(K.Graph.rmat(9, 4))
    b = K.Batch.from_graphs(graphs)
    db = eng.upload_batch(b)
    for flags in (0, K.FLAG_GRAPH, K.FLAG_GRAPH):
        core, kmax, _ = db.peel(flags=flags)
        pos = 0
        for i, g in enumerate(graphs):
            ref, km = O.bz(g.neighbors_offset, g.neighbors)
            assert kmax[i] == km and (core[pos:pos + g.V] == ref).all(), (flags, i)
            pos += g.V
    t = db.tune(reps=1)
    assert t["block"] in (256, 512, 1024) and t["trials"] == 3
    db.free()
    say("batches: eager, captured replays, tuner")


def serving(eng):
    gs = [K.Graph.request(23, i, median_m=3000.0, sigma=1.2, mmin=200, mmax=30000) for i in range(30)]
    ex = [O.bz(g.neighbors_offset, g.neighbors) for g in gs]
    big = sum(g.arcs >= 20000 for g in gs)
    assert 1 <= big < len(gs)
    for pinned in (True, False):
        with eng.serve(workgroups=2, ring_slots=8, large_arcs=20000, large_lanes=2, pinned=pinned,
                       hbm_budget=24 << 20) as srv:
            for i, g in enumerate(gs[:15]):
                srv.submit(g, req_id=i)
            first = srv.drain(timeout_s=300)
            for c in first:
                assert c["status"] == 0 and (c["core"] == ex[c["req_id"]][0]).all(), c["req_id"]
            # results dropped with the kernel resident: their pinned buffers are released now
            del first, c
            gc.collect()
            dg = eng.upload(gs[0])  # a single-graph peel and free beside the resident kernel
            assert (dg.peel().core == ex[0][0]).all()
            dg.free()
            for i, g in enumerate(gs[15:], start=15):
                srv.submit(g, req_id=i)
            for c in srv.drain(timeout_s=300):
                assert c["status"] == 0 and (c["core"] == ex[c["req_id"]][0]).all(), c["req_id"]
                assert c["kmax"] == ex[c["req_id"]][1]
            assert srv.pending() == (0, 0, 0)
    say(f"serving: pinned and pageable outputs, {big} lane requests, frees while resident")


def main():
    eng = K.Engine(0)
    eng.selftest()
    single_graph(eng)
    batches(eng)
    serving(eng)
    eng.close()
    say("ok")


if __name__ == "__main__":
    main()
