"""The device kernels run on the CPU through the wave64 emulator (tests/emu): the same headers the
GPU build compiles, with every wave operation, workgroup barrier and atomic emulated, several
workgroups running at once. Results must equal the oracle's. This checks the kernels' logic —
indexing, queue protocols, termination, level skipping, compaction, validation — without a GPU.
The weak-memory runs (KCEMU_WEAK=1) also hold every work-item's relaxed stores in a private store
buffer until it drains them (s_waitcnt, fences, release stores) or a random later point, so data
published without a drain is seen late or out of order, as in-flight stores on the GPU can be; a
planted missing drain is shown to be caught. What the host cannot show is the GPU's cache
behaviour itself (scopes, write-back, invalidation)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import kcoregpu_amd as K
import oracle as O
from conftest import closed_form_graphs, csr_from_edges

HERE = os.path.dirname(os.path.abspath(__file__))
CLANG = "/opt/rocm/llvm/bin/clang++"
pytestmark = pytest.mark.skipif(not os.path.exists(CLANG) and shutil.which("clang++") is None,
                                reason="needs clang++ (ext_vector_type)")


@pytest.fixture(scope="module")
def emu(tmp_path_factory):
    exe = tmp_path_factory.mktemp("emu") / "emu"
    cc = CLANG if os.path.exists(CLANG) else "clang++"
    subprocess.run([cc, "-O1", "-std=c++17", "-pthread", "-I", os.path.join(HERE, "emu"), "-o", str(exe),
                    os.path.join(HERE, "emu", "emu_main.cpp")], check=True)
    return str(exe)


def write_graph(path, off, nbr):
    with open(path, "wb") as f:
        f.write(np.uint32(len(off) - 1).tobytes())
        f.write(np.uint64(len(nbr)).tobytes())
        f.write(np.ascontiguousarray(off, np.uint32).tobytes())
        f.write(np.ascontiguousarray(nbr, np.uint32).tobytes())


WEAK = dict(os.environ, KCEMU_WEAK="1")


def run(emu, *args, env=None):
    r = subprocess.run([emu, *map(str, args)], capture_output=True, text=True, timeout=240, env=env)
    assert r.returncode == 0, r.stderr[-2000:]
    return r.stdout


def peel(emu, tmp_path, off, nbr, variant, *extra, env=None):
    gp, op = tmp_path / "g.bin", tmp_path / "o.bin"
    write_graph(gp, off, nbr)
    info = run(emu, "peel", variant, gp, op, *extra, env=env)
    d = np.fromfile(op, np.uint32)
    V = len(off) - 1
    kmax, err = int(d[0]), int(d[1])
    core, order, shells = d[2:2 + V], d[2 + V:2 + 2 * V], d[2 + 2 * V:]
    return kmax, err, core, order, shells, info


def check(off, nbr, kmax, err, core, order, shells):
    ref, km = O.bz(off, nbr)
    assert err == 0
    assert kmax == km
    assert (core == ref).all(), int(np.flatnonzero(core != ref)[0])
    assert O.check_certificate(off, nbr, core, order) == 0
    assert shells[0] == 0 and shells[kmax + 1] == len(off) - 1


def graphs():
    out = []
    for name, n, edges, _ in closed_form_graphs():
        off, nbr = csr_from_edges(n, edges)
        out.append((name, off, nbr))
    for seed, (n, m) in enumerate([(3000, 15000), (5000, 60000)]):
        g = K.Graph.erdos_renyi(n, m, seed + 1)
        out.append((f"er{n}", g.neighbors_offset.copy(), g.neighbors.copy()))
    g = K.Graph.chung_lu(8000, 3)
    out.append(("chung_lu8000", g.neighbors_offset.copy(), g.neighbors.copy()))
    g = K.Graph.rmat(12, 5)
    out.append(("rmat12", g.neighbors_offset.copy(), g.neighbors.copy()))
    # a hub of degree 3000 inside a 9-clique: chunked arcs
    src = [0] * 3000 + [i for i in range(1, 9) for j in range(i + 1, 10)] + [0] * 9
    dst = list(range(10, 3010)) + [j for i in range(1, 9) for j in range(i + 1, 10)] + list(range(1, 10))
    g = K.Graph.from_pairs(np.array(src), np.array(dst))
    out.append(("hub3000", g.neighbors_offset.copy(), g.neighbors.copy()))
    return out


GRAPHS = graphs()


@pytest.mark.parametrize("variant", ["persistent", "async", "asyncnb"])
@pytest.mark.parametrize("case", GRAPHS, ids=[c[0] for c in GRAPHS])
def test_single_graph_kernels(emu, tmp_path, variant, case):
    name, off, nbr = case
    check(off, nbr, *peel(emu, tmp_path, off, nbr, variant)[:5])


@pytest.mark.parametrize("settings", [
    dict(heavy=16, chunk=64, flags=0, grid=3, ratio=4),
    dict(heavy=64, chunk=512, flags=1, grid=4, ratio=4),      # reference level loop
    dict(heavy=64, chunk=512, flags=0, grid=2, ratio=1),      # compact at every chance
    dict(heavy=64, chunk=512, flags=0, grid=5, ratio=64),
    dict(heavy=0xFFFFFFFF, chunk=512, flags=1024, grid=4, ratio=4),  # never chunk; direct polling
])
@pytest.mark.parametrize("variant", ["async", "asyncnb"])
def test_async_settings(emu, tmp_path, settings, variant):
    g = K.Graph.rmat(13, 9)
    off, nbr = g.neighbors_offset.copy(), g.neighbors.copy()
    res = peel(emu, tmp_path, off, nbr, variant, settings["heavy"], settings["chunk"], settings["flags"],
               settings["grid"], settings["ratio"])
    check(off, nbr, *res[:5])
    info = dict(zip(res[5].split()[0::2], map(int, res[5].split()[1::2])))
    if settings["flags"] & 1:
        assert info["levels"] == res[0] + 1
    if settings["ratio"] == 1:
        assert info["list"] > 0 and info["surv"] > 0  # the alive list was built and scanned


@pytest.mark.parametrize("flags", [0, 1])
@pytest.mark.parametrize("name", ["K12", "union", "er3000", "rmat12", "hub3000", "isolated"])
def test_stepwise_kernels(emu, tmp_path, name, flags):
    _, off, nbr = next(c for c in GRAPHS if c[0] == name)
    gp, op = tmp_path / "g.bin", tmp_path / "o.bin"
    write_graph(gp, off, nbr)
    info = run(emu, "stepwise", gp, op, flags)
    d = np.fromfile(op, np.uint32)
    V = len(off) - 1
    check(off, nbr, int(d[0]), int(d[1]), d[2:2 + V], d[2 + V:2 + 2 * V], d[2 + 2 * V:])
    if flags:
        assert f"levels {int(d[0]) + 1} " in info  # the reference's loop: every level 0..kmax


def test_prefetch_variant(emu, tmp_path):
    g = K.Graph.chung_lu(6000, 8)
    off, nbr = g.neighbors_offset.copy(), g.neighbors.copy()
    check(off, nbr, *peel(emu, tmp_path, off, nbr, "prefetch")[:5])


def test_batch_kernel(emu, tmp_path, env=None):
    paths, refs = [], []
    cases = [c for c in GRAPHS if c[0] in ("K5", "cycle7", "star15", "isolated", "er3000", "chung_lu8000",
                                            "hub3000", "rmat12", "grid3x4", "union")]
    # graph arrays land at every alignment inside the packed arrays (quad-load shifts)
    for i, (name, off, nbr) in enumerate(cases):
        p = tmp_path / f"b{i}.bin"
        write_graph(p, off, nbr)
        paths.append(p)
        refs.append(O.bz(off, nbr))
    out = tmp_path / "bo.bin"
    run(emu, "batch", *paths, out, env=env)
    d = np.fromfile(out, np.uint32)
    pos = 0
    for (name, off, nbr), (ref, km) in zip(cases, refs):
        V = len(off) - 1
        assert d[pos] == km and d[pos + 1] == 0, name
        assert (d[pos + 2:pos + 2 + V] == ref).all(), name
        pos += 2 + V


def test_serve_kernel(emu, tmp_path, env=None):
    cases = [c for c in GRAPHS if c[0] in ("K12", "er3000", "rmat12", "hub3000", "isolated", "cycle16",
                                            "chung_lu8000", "K3,5", "path9")]
    paths, refs = [], []
    for i, (name, off, nbr) in enumerate(cases):
        p = tmp_path / f"s{i}.bin"
        write_graph(p, off, nbr)
        paths.append(p)
        refs.append(O.bz(off, nbr))
    # one malformed request: a neighbour id equal to V
    off, nbr = cases[1][1], cases[1][2].copy()
    nbr[7] = len(off) - 1
    bad = tmp_path / "bad.bin"
    write_graph(bad, off, nbr)
    out = tmp_path / "so.bin"
    run(emu, "serve", *paths[:4], bad, *paths[4:], out, env=env)
    d = np.fromfile(out, np.uint32)
    pos = 0
    order = list(range(4)) + [None] + list(range(4, len(cases)))
    for idx in order:
        if idx is None:
            V = len(off) - 1
            assert d[pos + 1] == 2  # SERVE_BAD_CSR, never peeled
            pos += 2 + V
            continue
        (name, o, n), (ref, km) = cases[idx], refs[idx]
        V = len(o) - 1
        assert d[pos] == km and d[pos + 1] == 0, name
        assert (d[pos + 2:pos + 2 + V] == ref).all(), name
        pos += 2 + V


def test_serve_lease(emu):
    assert "expired flag 1" in run(emu, "lease")


# ---- weak memory -------------------------------------------------------------------------

WEAK_CASES = [c for c in GRAPHS if c[0] in ("K12", "er5000", "rmat12", "hub3000", "chung_lu8000", "union")]


@pytest.mark.parametrize("variant", ["persistent", "async", "asyncnb", "prefetch"])
@pytest.mark.parametrize("case", WEAK_CASES, ids=[c[0] for c in WEAK_CASES])
def test_weak_memory_single_graph(emu, tmp_path, variant, case):
    name, off, nbr = case
    check(off, nbr, *peel(emu, tmp_path, off, nbr, variant, env=WEAK)[:5])


@pytest.mark.parametrize("settings", [
    dict(heavy=16, chunk=64, flags=0, grid=3, ratio=1),
    dict(heavy=64, chunk=512, flags=1024, grid=5, ratio=4),
])
@pytest.mark.parametrize("variant", ["async", "asyncnb"])
def test_weak_memory_async_settings(emu, tmp_path, settings, variant):
    g = K.Graph.rmat(13, 9)
    off, nbr = g.neighbors_offset.copy(), g.neighbors.copy()
    res = peel(emu, tmp_path, off, nbr, variant, settings["heavy"], settings["chunk"], settings["flags"],
               settings["grid"], settings["ratio"], env=WEAK)
    check(off, nbr, *res[:5])


def test_weak_memory_batch(emu, tmp_path):
    test_batch_kernel(emu, tmp_path, env=WEAK)


def test_weak_memory_serve(emu, tmp_path):
    test_serve_kernel(emu, tmp_path, env=WEAK)


def test_weak_memory_stepwise(emu, tmp_path):
    _, off, nbr = next(c for c in GRAPHS if c[0] == "rmat12")
    gp, op = tmp_path / "g.bin", tmp_path / "o.bin"
    write_graph(gp, off, nbr)
    run(emu, "stepwise", gp, op, 0, env=WEAK)
    d = np.fromfile(op, np.uint32)
    V = len(off) - 1
    check(off, nbr, int(d[0]), int(d[1]), d[2:2 + V], d[2 + V:2 + 2 * V], d[2 + 2 * V:])


def test_weak_memory_catches_a_missing_drain(tmp_path):
    """A copy of the batch kernel without the drain between the select's order[] stores and the
    barrier after which other work-items read them: correct in the ordinary emulation, wrong in
    the weak-memory one (the stores are still in flight when they are read)."""
    root = os.path.dirname(HERE)
    work = tmp_path / "planted"
    shutil.copytree(os.path.join(HERE, "emu"), work / "tests" / "emu")
    shutil.copytree(os.path.join(root, "kcoregpu_amd", "csrc"), work / "kcoregpu_amd" / "csrc",
                    ignore=shutil.ignore_patterns("*.so", "*.o"))
    hdr = work / "kcoregpu_amd" / "csrc" / "batch_kernel.h"
    text = hdr.read_text()
    drained = "        levels++;\n        drain_memory();\n        __syncthreads();"
    assert text.count(drained) == 1
    hdr.write_text(text.replace(drained, "        levels++;\n        __syncthreads();"))
    exe = work / "emu_planted"
    cc = CLANG if os.path.exists(CLANG) else "clang++"
    subprocess.run([cc, "-O1", "-std=c++17", "-pthread", "-I", str(work / "tests" / "emu"), "-o", str(exe),
                    str(work / "tests" / "emu" / "emu_main.cpp")], check=True)
    paths, refs = [], []
    for i, (n, m) in enumerate([(3000, 15000), (5000, 60000), (2000, 30000), (4000, 8000)]):
        g = K.Graph.erdos_renyi(n, m, i + 1)
        p = tmp_path / f"b{i}.bin"
        write_graph(p, g.neighbors_offset, g.neighbors)
        paths.append(p)
        refs.append(O.bz(g.neighbors_offset, g.neighbors))

    def mismatches(env):
        out = tmp_path / "o.bin"
        r = subprocess.run([str(exe), "batch", *map(str, paths), str(out)], capture_output=True, text=True,
                           timeout=240, env=env)
        if r.returncode != 0:
            return -1  # a planted fault may also crash the run: caught as well
        d = np.fromfile(out, np.uint32)
        pos, bad = 0, 0
        for ref, km in refs:
            V = len(ref)
            bad += int(d[pos] != km) + int((d[pos + 2:pos + 2 + V] != ref).sum())
            pos += 2 + V
        return bad

    assert mismatches(dict(os.environ, KCEMU_WEAK="0")) == 0
    assert mismatches(WEAK) != 0


@pytest.mark.parametrize("variant", ["persistent", "async", "asyncnb"])
def test_weak_memory_with_jitter(emu, tmp_path, variant):
    """Weak memory plus random pauses of whole workgroups (KCEMU_JITTER): workgroups overtake each
    other at random points of the claim, termination and barrier protocols."""
    env = dict(WEAK, KCEMU_JITTER="20")
    for name in ("rmat12", "hub3000", "chung_lu8000"):
        _, off, nbr = next(c for c in GRAPHS if c[0] == name)
        check(off, nbr, *peel(emu, tmp_path, off, nbr, variant, 16, 64, 0, 5, 2, env=env)[:5])


def test_handoff_held_claims(emu, tmp_path):
    """The level hand-off's rarest path: a workgroup that claims after the level was won holds
    entries of the next level, scans that level, and processes them after its gate. Small graphs
    with many levels, six workgroups and pauses of up to 5 ms (a preempted workgroup) make it
    happen; the kernel counts the holds (result word R_WORDS + 4) and every run must still be
    exact."""
    env = dict(os.environ, KCEMU_JITTER="100:5000")
    held = 0
    for seed in range(6):
        g = K.Graph.rmat(10, 40 + seed, edge_factor=12)
        off, nbr = g.neighbors_offset.copy(), g.neighbors.copy()
        res = peel(emu, tmp_path, off, nbr, "asyncnb", 16, 64, 0, 6, 4, env=env)
        check(off, nbr, *res[:5])
        info = dict(zip(res[5].split()[0::2], map(int, res[5].split()[1::2])))
        held += info["held"]
    assert held > 0
