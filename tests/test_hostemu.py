"""The library's real host code — kcore_hip.hip, serve.hip and tune.cpp, i.e. the C ABI of
include/kcore.h — built for the CPU against a stand-in HIP runtime (tests/hostemu/hip: in-order
stream threads, events, stream capture and graph replay, and hipFree / hipHostFree that wait for
every stream as HIP's do), with every kernel run by the wave64 emulator (tests/emu). Only the
`kernel<<<grid, block, shmem, stream>>>(args)` launches are rewritten (tests/hostemu/cpu_lib.py),
into a call that enqueues the emulated launch; nothing else in the sources changes. Results must
equal the oracle's.

This exercises what the device-code emulation alone cannot: argument set-up, buffer sizes and
alignment, queue and slot bookkeeping of the serving loop, the large-request lanes, the deferred
frees while a serving kernel is resident (a free that waits on it aborts the run), graph capture
and replay, the tuner and its cache."""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

import kcoregpu_amd as K
import oracle as O
from conftest import closed_form_graphs, csr_from_edges
from test_emu import CLANG, write_graph

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(HERE, "hostemu"))
import cpu_lib  # noqa: E402

rewrite_launches = cpu_lib.rewrite_launches
pytestmark = pytest.mark.skipif(not os.path.exists(CLANG) and shutil.which("clang++") is None,
                                reason="needs clang++ (ext_vector_type)")
ENV = dict(os.environ, KCEMU_WALL_KHZ="10000000", KCEMU_MAX_GRID="4")


def test_rewrite_launches():
    src = "  peel_async<BLOCK, PREFETCH><<<dim3(grid), dim3(BLOCK), 0, c->stream>>>(a);\n" \
          "  copy_kernel<<<2048, 256, 0, st>>>(reinterpret_cast<const uint4*>(g->d), f(x, y), 3);\n"
    got, n = rewrite_launches(src)
    assert n == 2
    assert got == ("  kcemu_launch(dim3(grid), dim3(BLOCK), c->stream, peel_async<BLOCK, PREFETCH>, a);\n"
                   "  kcemu_launch(2048, 256, st, copy_kernel, reinterpret_cast<const uint4*>(g->d), f(x, y), 3);\n")


CC = cpu_lib.CC


@pytest.fixture(scope="module")
def emu_lib(tmp_path_factory):
    """The library's host sources, launches rewritten, compiled against the stand-in runtime."""
    d = tmp_path_factory.mktemp("hostemu")
    return d, cpu_lib.build_objects(str(d))


@pytest.fixture(scope="module")
def host(emu_lib):
    d, objs = emu_lib
    exe = d / "host"
    subprocess.run([CC, "-O1", "-std=c++17", "-pthread", os.path.join(HERE, "hostemu", "host_main.cpp"), *objs,
                    "-o", str(exe)], check=True)
    return str(exe)


@pytest.fixture(scope="module")
def tools(emu_lib):
    """The `kcore` driver and the `kcore_serve` load generator, linked with the emulated library."""
    d, objs = emu_lib
    csrc = d / "kcoregpu_amd" / "csrc"
    host_objs = {}
    for f in ("graph.cpp", "generators.cpp", "graph_capi.cpp", "bincache.cpp", "kcore_main.cpp", "kcore_serve.cpp"):
        o = str(csrc / f) + ".o"
        subprocess.run(["g++", "-O1", "-std=c++17", "-fopenmp", "-c", str(csrc / f), "-o", o], check=True)
        host_objs[f] = o
    lib = [host_objs[f] for f in ("graph.cpp", "generators.cpp", "graph_capi.cpp", "bincache.cpp")]
    kcore, serve = str(d / "kcore"), str(d / "kcore_serve")
    subprocess.run(["g++", "-fopenmp", "-pthread", host_objs["kcore_main.cpp"], *lib, *objs, "-o", kcore], check=True)
    subprocess.run(["g++", "-fopenmp", "-pthread", host_objs["kcore_serve.cpp"], *lib, *objs, "-o", serve], check=True)
    return kcore, serve


def run(host, *args, timeout=600):
    r = subprocess.run([host, *map(str, args)], capture_output=True, text=True, timeout=timeout, env=ENV)
    assert r.returncode == 0, (r.returncode, r.stderr[-3000:])
    return r.stdout


def graph_set():
    out = []
    for name, n, edges, _ in closed_form_graphs():
        if name in ("K12", "union", "isolated", "grid3x4"):
            off, nbr = csr_from_edges(n, edges)
            out.append((name, off, nbr))
    g = K.Graph.erdos_renyi(3000, 15000, 1)
    out.append(("er3000", g.neighbors_offset.copy(), g.neighbors.copy()))
    g = K.Graph.rmat(11, 5)
    out.append(("rmat11", g.neighbors_offset.copy(), g.neighbors.copy()))
    src = [0] * 2500 + [i for i in range(1, 9) for j in range(i + 1, 10)] + [0] * 9
    dst = list(range(10, 2510)) + [j for i in range(1, 9) for j in range(i + 1, 10)] + list(range(1, 10))
    g = K.Graph.from_pairs(np.array(src), np.array(dst))
    out.append(("hub2500", g.neighbors_offset.copy(), g.neighbors.copy()))
    return out


GRAPHS = graph_set()
RUNS = ["defaults", "wave 256 chunks", "prefetch 1024", "lds no-skip", "stepwise", "stepwise no-skip",
        "async ratio 1", "async again", "async direct no-skip", "handoff ratio 1", "handoff no-skip",
        "defaults again"]


@pytest.mark.parametrize("case", GRAPHS, ids=[c[0] for c in GRAPHS])
def test_single_graph_c_abi(host, tmp_path, case):
    name, off, nbr = case
    gp, op = tmp_path / "g.bin", tmp_path / "o.bin"
    write_graph(gp, off, nbr)
    run(host, "peel", gp, op)
    d = np.fromfile(op, np.uint32)
    ref, km = O.bz(off, nbr)
    V = len(off) - 1
    pos = 0
    for label in RUNS:
        st, kmax, levels = (int(x) for x in d[pos:pos + 3])
        core, order = d[pos + 3:pos + 3 + V], d[pos + 3 + V:pos + 3 + 2 * V]
        shells = d[pos + 3 + 2 * V:pos + 5 + 2 * V + kmax]
        pos += 5 + 2 * V + kmax
        assert st == 0, label
        assert kmax == km, label
        assert (core == ref).all(), label
        assert O.check_certificate(off, nbr, core, order) == 0, label
        assert shells[0] == 0 and shells[kmax + 1] == V, label
        assert (np.diff(shells.astype(np.int64)) >= 0).all(), label
        if "no-skip" in label:
            assert levels == kmax + 1, label
    assert pos == len(d)


def test_batch_c_abi(host, tmp_path):
    paths, refs = [], []
    for i, (name, off, nbr) in enumerate(GRAPHS):
        p = tmp_path / f"b{i}.bin"
        write_graph(p, off, nbr)
        paths.append(p)
        refs.append(O.bz(off, nbr))
    out = tmp_path / "bo.bin"
    run(host, "batch", out, *paths)
    d = np.fromfile(out, np.uint32)
    # degree entries scanned: every graph at least once, at most once per level 0..kmax
    scanned = int(np.fromfile(str(out) + ".scan", np.uint64)[0])
    sizes = [(len(off) - 1, km) for (_, off, _), (_, km) in zip(GRAPHS, refs)]
    assert sum(V for V, _ in sizes) <= scanned <= sum(V * (km + 1) for V, km in sizes)
    pos = 0
    for rnd in ("eager", "replay 1", "replay 2", "replay 3", "tuned"):
        for (name, off, nbr), (ref, km) in zip(GRAPHS, refs):
            V = len(off) - 1
            assert d[pos] == km, (rnd, name)
            assert (d[pos + 1:pos + 1 + V] == ref).all(), (rnd, name)
            pos += 1 + V
    assert pos == len(d)


def test_serve_c_abi(host, tmp_path):
    """A served stream with requests of every size class: small ones on the resident kernel, large
    ones (>= 20,000 arcs) on the lanes, a malformed one, an oversized one; a small arena so requests
    wait for memory; pinned result buffers freed as they complete and a whole single-graph peel,
    all while the serving kernel is resident."""
    gs = [K.Graph.request(41, i, median_m=4000.0, sigma=1.2, mmin=200, mmax=40000) for i in range(24)]
    gs += [K.Graph.from_csr(off, nbr) for _, off, nbr in GRAPHS]
    paths, refs = [], []
    for i, g in enumerate(gs):
        p = tmp_path / f"s{i}.bin"
        write_graph(p, g.neighbors_offset, g.neighbors)
        paths.append(p)
        refs.append(O.bz(g.neighbors_offset, g.neighbors))
    large = 20000
    nbig = sum(g.arcs >= large for g in gs)
    assert 2 <= nbig < len(gs)
    # room for a few requests at once, including the largest lane request's whole state
    budget = 3 * max(48 * g.V + 6 * g.arcs + (2 << 20) for g in gs)
    out = tmp_path / "so.bin"
    run(host, "serve", out, large, budget, *paths, timeout=900)
    d = np.fromfile(out, np.uint32)
    pos = 0
    for i, (g, (ref, km)) in enumerate(zip(gs, refs)):
        assert d[pos] == 0, i
        assert d[pos + 1] == km, i
        assert (d[pos + 2:pos + 2 + g.V] == ref).all(), i
        pos += 2 + g.V
    assert d[pos] == nbig and d[pos + 1] == len(gs)
    side = np.fromfile(str(out) + ".side", np.uint32)
    assert side[0] == refs[0][1] and (side[1:] == refs[0][0]).all()


def test_tuner_c_abi(host, tmp_path):
    _, off, nbr = next(c for c in GRAPHS if c[0] == "rmat11")
    gp, op, cache = tmp_path / "g.bin", tmp_path / "t.bin", tmp_path / "tune.txt"
    write_graph(gp, off, nbr)
    run(host, "tune", gp, cache, op)
    d = np.fromfile(op, np.uint32)
    t1, t2, variant, block, st, kmax = (int(x) for x in d[:6])
    assert t1 > 10 and t2 == 0  # timed the first time, answered from the cache the second
    assert variant in (1, 2, 3, 5) and block in (256, 512, 1024) and st == 0
    ref, km = O.bz(off, nbr)
    assert kmax == km and (d[6:] == ref).all()
    line = cache.read_text().strip().split("\n")
    assert len(line) == 1 and line[0].startswith(f"peel {len(off) - 1} {len(nbr)} ")


# the reference's labels (main.cu:60-88), "SM:" with two spaces
LABELS = ["Ours: ", "SM:  ", "VP: ", "BC: ", "BC + SM: ", "BC + VP: ", "EC: ", "EC + SM: ", "EC + VP: "]


def test_driver_end_to_end(tools, tmp_path):
    """`kcore <edgelist>`: the reference's stdout lines, the text and binary caches, and every output
    file (raw, text and JSON core numbers, peel order, shell offsets) against the oracle, through
    the default, cascade-queue and reference-shaped variants."""
    import re
    kcore, _ = tools
    g = K.Graph.erdos_renyi(2500, 12000, 7)
    g.write_edgelist(str(tmp_path / "er.txt"))
    os.makedirs(tmp_path / "output")
    core, kmax = O.bz(g.neighbors_offset, g.neighbors)

    def drive(*args):
        r = subprocess.run([kcore, *args], cwd=tmp_path, capture_output=True, text=True, timeout=600, env=ENV)
        assert r.returncode == 0, (r.returncode, r.stderr[-2000:])
        return r.stdout.split("\n")

    out = drive("er.txt", "--core-out", "core.u32", "--core-text", "core.txt", "--json", "--order-out", "order.u32")
    assert out[:3] == ["Graph loading Started... ", "serialized-er.txt : File couldn't open", "Reading normal file... "]
    assert re.fullmatch(r"File Loaded in: \d+", out[3])
    assert out[4] == "******************  er.txt ****************** "
    assert out[5] == f"V: {g.V}" and out[6] == f"E: {g.arcs}"
    i = 7
    for k, lab in enumerate(LABELS):
        assert re.fullmatch(re.escape(lab) + r"\d+ ", out[i]), out[i]
        i += 1
        if k == 0:
            assert out[i] == f"Kmax: {kmax}"
            i += 1
    assert out[i] == "Deallocated... " and out[i + 1] == ""
    got = np.fromfile(tmp_path / "core.u32", np.uint32)
    assert (got == core).all()
    assert (tmp_path / "core.txt").read_text() == "".join(f"{c}\n" for c in core)
    order = np.fromfile(tmp_path / "order.u32", np.uint32)
    assert O.check_certificate(g.neighbors_offset, g.neighbors, got, order) == 0
    shells = np.fromfile(tmp_path / "order.u32.shells", np.uint32)
    assert shells[0] == 0 and shells[-1] == g.V and len(shells) == kmax + 2
    want = "{ " + "".join(f'"{v}": {c}, \n' for v, c in enumerate(core) if c)
    want = want[:-3] + " }\n" if want != "{ " else want
    assert (tmp_path / "output" / "pkc-kcore-er.txt").read_text() == want

    for variant, cache_args, cache_line in (("async", [], "Reading serialized file... "),
                                            ("stepwise", ["--binary-cache"], "Reading normal file... "),
                                            ("auto", ["--binary-cache"], "Reading binary cache... ")):
        out = drive("er.txt", "--variant", variant, *cache_args, "--core-out", f"{variant}.u32")
        assert out[:2] == ["Graph loading Started... ", cache_line], (variant, out[:3])
        assert (np.fromfile(tmp_path / f"{variant}.u32", np.uint32) == core).all(), variant
        assert f"Kmax: {kmax}" in out, variant


@pytest.mark.parametrize("mode", ["served", "plain"])
def test_kcore_serve_end_to_end(tools, tmp_path, mode):
    """`kcore_serve`, closed-loop, every result checked against the oracle's core numbers of the
    request pool (the file the benchmark's CPU leg writes); served with lane requests and a small
    arena."""
    import json
    _, serve = tools
    pool, seed, median, mmin, mmax = 10, 3, 3000.0, 200, 30000
    with open(tmp_path / "expect.u32", "wb") as f:
        for j in range(pool):
            g = K.Graph.request(seed, j, median_m=median, sigma=1.0, mmin=mmin, mmax=mmax)
            f.write(O.bz(g.neighbors_offset, g.neighbors)[0].tobytes())
    args = [serve, "--mode", mode, "--closed", "30", "--pool", str(pool), "--seed", str(seed), "--median", str(median),
            "--mmin", str(mmin), "--mmax", str(mmax), "--expect", str(tmp_path / "expect.u32"), "--seconds", "1",
            "--drain-limit", "300"]
    if mode == "served":
        args += ["--budget-gb", "0.02", "--workgroups", "2", "--slots", "8", "--large-arcs", "20000",
                 "--large-lanes", "2"]
    r = subprocess.run(args, capture_output=True, text=True, timeout=900, env=ENV)
    assert r.returncode == 0, (r.returncode, r.stdout[-1000:], r.stderr[-2000:])
    res = json.loads(r.stdout.strip().split("\n")[-1])
    assert res["requests"] == 30 and res["completed"] == 30 and res["unfinished"] == 0
    assert res["mismatches"] == 0 and res["device_error"] == 0 and res["dropped"] == 0
    assert res["checked_against"] == "expected core numbers"


def test_python_api_over_the_cpu_build(emu_lib, tmp_path):
    """The package's Engine / DeviceGraph / DeviceBatch / Server / tuner wrappers over the CPU build
    of the library (its path swapped in a separate process; the package has no switch for it):
    ctypes structures and calls, pinned result buffers released by the garbage collector while the
    serving kernel is resident, every result against the oracle."""
    d, objs = emu_lib
    so = cpu_lib.build_shared(objs, str(d / "libkcore_hip_emu.so"))
    r = subprocess.run(["python3", os.path.join(HERE, "hostemu", "py_api.py"), so, str(tmp_path)],
                       capture_output=True, text=True, timeout=1200, env=ENV)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-3000:])
    assert r.stdout.strip().endswith("ok")


@pytest.fixture(scope="module")
def host_asan(tmp_path_factory):
    """The same harness built with AddressSanitizer and UndefinedBehaviorSanitizer (host code and
    emulated kernels alike)."""
    d = tmp_path_factory.mktemp("hostemu_asan")
    flags = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer",
             "-gline-tables-only"]
    objs = cpu_lib.build_objects(str(d), flags)
    exe = str(d / "host")
    subprocess.run([CC, "-O1", "-std=c++17", "-pthread", *flags, os.path.join(HERE, "hostemu", "host_main.cpp"),
                    *objs, "-o", exe], check=True)
    return exe


ASAN_ENV = dict(ENV, KCEMU_EXACT="1", ASAN_OPTIONS="detect_leaks=0:detect_stack_use_after_return=0")


def test_no_out_of_bounds_access(host_asan, tmp_path):
    """Every device allocation exactly as large as the host code asks (KCEMU_EXACT), under
    AddressSanitizer and UndefinedBehaviorSanitizer: every single-graph variant on a hub graph (chunked arcs) and a served stream
    with lane requests touch nothing past an allocation's end (on the GPU such an access can fault
    the device)."""
    for name in ("hub2500",):
        _, off, nbr = next(c for c in GRAPHS if c[0] == name)
        gp, op = tmp_path / "g.bin", tmp_path / "o.bin"
        write_graph(gp, off, nbr)
        r = subprocess.run([host_asan, "peel", str(gp), str(op)], capture_output=True, text=True, timeout=600,
                           env=ASAN_ENV)
        assert r.returncode == 0, (name, r.stderr[-3000:])
    gs = [K.Graph.request(43, i, median_m=3000.0, sigma=1.2, mmin=200, mmax=30000) for i in range(8)]
    paths = []
    for i, g in enumerate(gs):
        p = tmp_path / f"s{i}.bin"
        write_graph(p, g.neighbors_offset, g.neighbors)
        paths.append(str(p))
    large = sorted(g.arcs for g in gs)[-2]
    budget = 3 * max(48 * g.V + 6 * g.arcs + (2 << 20) for g in gs)
    r = subprocess.run([host_asan, "serve", str(tmp_path / "so.bin"), str(large), str(budget), *paths],
                       capture_output=True, text=True, timeout=900, env=ASAN_ENV)
    assert r.returncode == 0, r.stderr[-3000:]
