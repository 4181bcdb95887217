"""GPU parity of served peeling: requests peeled by the resident serving kernel match the oracle bit
for bit, alone and in a mix with other requests, through ring wrap-around and arena pressure."""
import sys
import time

import numpy as np
import pytest

import kcoregpu_amd as K
import oracle as O
from conftest import closed_form_graphs, csr_from_edges

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def requests(seed, count, median=20000.0, mmax=400000):
    return [K.Graph.request(seed, i, median_m=median, sigma=1.5, mmax=mmax) for i in range(count)]


def expected(gs):
    return [O.bz(g.neighbors_offset, g.neighbors) for g in gs]


def test_serve_stream_matches_oracle(engine):
    gs = requests(11, 120)
    ex = expected(gs)
    say("serve: 120 requests, block 256")
    with engine.serve(block=256) as srv:
        for i, g in enumerate(gs):
            srv.submit(g, req_id=i)
        done = srv.drain(timeout_s=60)
        assert len(done) == len(gs)
        for c in done:
            i = c["req_id"]
            assert c["status"] == 0, i
            assert (c["core"] == ex[i][0]).all(), i
            assert c["kmax"] == ex[i][1], i
            assert c["t_submit"] <= c["t_admit"] <= c["t_done"]
        assert srv.pending() == (0, 0, 0)


@pytest.mark.parametrize("block", [512, 1024])
def test_serve_blocks(engine, block):
    gs = requests(12, 40)
    ex = expected(gs)
    say(f"serve: 40 requests, block {block}")
    with engine.serve(block=block) as srv:
        for i, g in enumerate(gs):
            srv.submit(g, req_id=i)
        for c in srv.drain(timeout_s=60):
            assert (c["core"] == ex[c["req_id"]][0]).all(), c["req_id"]


def test_serve_alone_equals_mix(engine):
    """Each request's result is the same when it runs alone on an idle server and when it runs
    among many others that arrive while it is on the GPU."""
    probe = requests(13, 12)
    ex = expected(probe)
    filler = requests(14, 200)
    say("serve: 12 probes alone, then inside 200 fillers")
    with engine.serve() as srv:
        alone = {}
        for i, g in enumerate(probe):
            srv.submit(g, req_id=i)
            (c,) = srv.drain(timeout_s=30)
            alone[i] = c["core"]
        mixed = {}
        rid = 1000
        for k, g in enumerate(filler):
            srv.submit(g, req_id=rid)
            rid += 1
            if k % 16 == 5 and k // 16 < len(probe):
                srv.submit(probe[k // 16], req_id=k // 16)
            for c in srv.poll():
                if c["req_id"] < 1000:
                    mixed[c["req_id"]] = c["core"]
        for c in srv.drain(timeout_s=60):
            if c["req_id"] < 1000:
                mixed[c["req_id"]] = c["core"]
    assert sorted(mixed) == list(range(len(probe)))
    for i in range(len(probe)):
        assert (alone[i] == ex[i][0]).all(), i
        assert (mixed[i] == alone[i]).all(), i


def test_serve_ring_wrap_and_arena_pressure(engine):
    gs = requests(15, 60, median=30000.0, mmax=200000)
    ex = expected(gs)
    need = max(4 * (g.V + 1) + 4 * g.arcs + 12 * g.V + 5 * 256 for g in gs)
    budget = 3 * need  # a few requests fit at once: the rest wait for memory
    say(f"serve: ring 4 slots, arena {budget} bytes")
    with engine.serve(ring_slots=4, hbm_budget=budget, workgroups=8) as srv:
        for i, g in enumerate(gs):
            srv.submit(g, req_id=i)
            q, f, used = srv.pending()
            assert f <= 4 and used <= budget
        seen = 0
        t0 = time.monotonic()
        while seen < len(gs):
            for c in srv.poll():
                assert (c["core"] == ex[c["req_id"]][0]).all(), c["req_id"]
                seen += 1
            q, f, used = srv.pending()
            assert f <= 4 and used <= budget
            assert time.monotonic() - t0 < 60
        # a request larger than the whole arena is refused, the server keeps working
        big = K.Graph.erdos_renyi(200000, 2000000, 3)
        with pytest.raises(K.KCoreError):
            srv.submit(big)
        srv.submit(gs[0], req_id=999)
        (c,) = srv.drain(timeout_s=30)
        assert (c["core"] == ex[0][0]).all()


def test_serve_closed_forms_and_hub(engine):
    gs, ex = [], []
    for name, n, edges, e in closed_form_graphs():
        off, nbr = csr_from_edges(n, edges)
        gs.append(K.Graph.from_csr(off, nbr))
        ex.append(np.array(e, np.uint32))
    n_leaf = 50000
    src = [0] * n_leaf + [i for i in range(10) for j in range(i + 1, 10)]
    dst = list(range(10, 10 + n_leaf)) + [j for i in range(10) for j in range(i + 1, 10)]
    hub = K.Graph.from_pairs(np.array(src), np.array(dst))
    gs.append(hub)
    ex.append(O.bz(hub.neighbors_offset, hub.neighbors)[0])
    with engine.serve() as srv:
        for i, g in enumerate(gs):
            srv.submit(g, req_id=i)
        for c in srv.drain(timeout_s=30):
            assert (c["core"] == ex[c["req_id"]]).all(), c["req_id"]
            assert c["kmax"] == ex[c["req_id"]].max()


def test_serve_no_level_skip_same_cores(engine):
    gs = requests(16, 20)
    ex = expected(gs)
    with engine.serve(flags=K.FLAG_NO_LEVEL_SKIP) as srv:
        for i, g in enumerate(gs):
            srv.submit(g, req_id=i)
        for c in srv.drain(timeout_s=60):
            assert (c["core"] == ex[c["req_id"]][0]).all()
            assert c["levels"] >= ex[c["req_id"]][1]


def test_serve_copy_path_and_malformed_request(engine):
    """Pageable output buffers take the copy-down path; a malformed request completes with
    KC_ERR_ARG without being peeled, and the server keeps serving."""
    gs = requests(17, 10)
    ex = expected(gs)
    bad = K.Graph.erdos_renyi(2000, 8000, 5)
    with engine.serve(pinned=False) as srv:
        for i, g in enumerate(gs[:5]):
            srv.submit(g, req_id=i)
        # neighbour id == V in one adjacency list: rejected by the kernel's check
        nbr = bad.neighbors.copy()
        nbr[100] = bad.V
        import ctypes as C
        from kcoregpu_amd import _native as N
        out = np.empty(bad.V, np.uint32)
        h = N.kc_csr(bad.V, bad.arcs, bad.neighbors_offset.ctypes.data_as(N.u32p),
                     nbr.ctypes.data_as(N.u32p), N.u32p())
        assert N.hip().kc_serve_submit(srv._s, 777, C.byref(h), out.ctypes.data_as(N.u32p)) == 0
        srv._live[777] = (bad, out)  # keeps the buffers alive until the completion
        for i, g in enumerate(gs[5:], start=5):
            srv.submit(g, req_id=i)
        got = {c["req_id"]: c for c in srv.drain(timeout_s=60)}
    assert got[777]["status"] == 1  # KC_ERR_ARG
    for i in range(len(gs)):
        assert got[i]["status"] == 0
        assert (got[i]["core"] == ex[i][0]).all(), i


def test_serve_survives_idle_longer_than_lease(engine):
    """The heartbeat keeps an idle server's kernel alive past its lease."""
    g = requests(18, 1)[0]
    ref = O.bz(g.neighbors_offset, g.neighbors)[0]
    with engine.serve(lease_ms=200, workgroups=16) as srv:
        srv.submit(g, req_id=0)
        (c,) = srv.drain(timeout_s=30)
        assert (c["core"] == ref).all()
        time.sleep(1.0)  # five leases without a request
        srv.submit(g, req_id=1)
        (c,) = srv.drain(timeout_s=30)
        assert (c["core"] == ref).all()


def test_serve_large_request_lane(engine):
    """Requests above large_arcs are peeled by the single-graph kernel on the large-request lanes,
    beside the resident kernel that keeps serving the small ones; all results match the oracle."""
    gs = requests(19, 60, median=30000.0, mmax=600000)
    ex = expected(gs)
    with engine.serve(large_arcs=100000, large_lanes=2) as srv:
        for i, g in enumerate(gs):
            srv.submit(g, req_id=i)
        got = {c["req_id"]: c for c in srv.drain(timeout_s=120)}
        assert srv.pending() == (0, 0, 0)
    big = [i for i, g in enumerate(gs) if g.arcs >= 100000]
    assert 5 <= len(big) < len(gs)
    for i in range(len(gs)):
        assert got[i]["status"] == 0, i
        assert (got[i]["core"] == ex[i][0]).all(), i
        assert got[i]["kmax"] == ex[i][1], i


def test_serve_frees_while_resident(engine):
    """hipFree and hipHostFree synchronise the whole device, so with the serving kernel resident
    they would wait for it forever. Pinned result buffers dropped mid-stream, and a graph uploaded,
    peeled and freed beside the server, must neither hang nor disturb it: the library defers its
    frees until no serving kernel is resident and hands returned pinned blocks out again."""
    import ctypes as C
    import gc

    from kcoregpu_amd import _native as N
    gs = requests(20, 40)
    ex = expected(gs)
    say("serve: frees with the kernel resident")
    with engine.serve() as srv:
        for i, g in enumerate(gs[:20]):
            srv.submit(g, req_id=i)
        for c in srv.drain(timeout_s=60):
            assert (c["core"] == ex[c["req_id"]][0]).all(), c["req_id"]
        gc.collect()  # the pinned result buffers are released here, with the kernel resident
        # a returned block is handed out again (64 MiB: no result buffer is in its size range);
        # a second free of it is refused
        p = C.c_void_p()
        assert N.hip().kc_host_alloc(64 << 20, C.byref(p)) == 0 and p.value
        assert N.hip().kc_host_free(p) == 0
        q = C.c_void_p()
        assert N.hip().kc_host_alloc((64 << 20) - 100, C.byref(q)) == 0 and q.value == p.value
        assert N.hip().kc_host_free(q) == 0
        assert N.hip().kc_host_free(q) != 0
        dg = engine.upload(gs[0])
        r = dg.peel()
        assert (r.core == ex[0][0]).all()
        dg.free()
        for i, g in enumerate(gs[20:], start=20):
            srv.submit(g, req_id=i)
        for c in srv.drain(timeout_s=60):
            assert (c["core"] == ex[c["req_id"]][0]).all(), c["req_id"]
