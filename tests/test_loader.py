"""Loader, serialized cache, JSON writer and generators against the reference's own loader
(fixtures written by tests/golden/make_golden.py with oracle/_ref/ref_loader, built from the
reference's src/graph.cpp) and against properties of the generators."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import kcoregpu_amd as K
import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "loader")
CASES = sorted(f[:-4] for f in os.listdir(GOLDEN) if f.endswith(".txt"))


@pytest.mark.parametrize("case", CASES)
def test_loader_matches_reference_fixture(case, tmp_path):
    src = os.path.join(GOLDEN, case + ".txt")
    exp = os.path.join(GOLDEN, case + ".expected")
    z = np.load(os.path.join(exp, "csr.npz"))
    g = K.Graph.from_edgelist(src)
    assert g.V == int(z["V"]) and g.E == int(z["E"])
    assert (g.neighbors_offset == z["off"]).all()
    assert (g.neighbors == z["nbr"]).all()
    assert (g.degrees == z["deg"]).all()
    # serialized cache, byte for byte
    out = tmp_path / "ser"
    g.write_serialized(str(out))
    assert out.read_bytes() == open(os.path.join(exp, "serialized"), "rb").read()
    # the cache reads back to the same CSR
    h = K.Graph.read_serialized(str(out))
    assert h.V == g.V and (h.neighbors == g.neighbors).all() and (h.degrees == g.degrees).all()
    # JSON core writer, byte for byte
    j = tmp_path / "core.json"
    K.write_kcore_json(str(j), z["core"])
    assert j.read_bytes() == open(os.path.join(exp, "pkc-kcore.json"), "rb").read()
    # the text core dump has the form of the degree block the reference's writer produced:
    # written from the degrees, it is exactly that block of the reference's serialized bytes
    t = tmp_path / "deg.txt"
    K.write_core_text(str(t), z["deg"])
    ref = open(os.path.join(exp, "serialized"), "rb").read()
    head = f"{int(z['V'])}\n{int(z['E'])}\n".encode()
    assert ref.startswith(head + t.read_bytes()) and ref[len(head) + t.stat().st_size:].startswith(b"0 ")
    K.write_core_text(str(t), z["core"])
    assert [int(x) for x in t.read_text().split("\n")[:-1]] == list(z["core"])
    u = tmp_path / "core.u32"
    K.write_core_u32(str(u), z["core"])
    assert (np.fromfile(u, np.uint32) == z["core"]).all()


def test_load_cached_uses_basename_cache(tmp_path):
    src = tmp_path / "g.txt"
    src.write_text("0 1\n1 2\n")
    outdir = tmp_path / "output"
    outdir.mkdir()
    g = K.Graph.load(str(src), str(outdir))
    assert not g.from_cache and (outdir / "serialized-g.txt").exists()
    # the reference keys the cache on the basename only: a different file with the same name
    # in another directory hits the same cache (src/graph.cpp:10,36)
    other = tmp_path / "sub"
    other.mkdir()
    (other / "g.txt").write_text("5 6\n")
    g2 = K.Graph.load(str(other / "g.txt"), str(outdir))
    assert g2.from_cache and g2.V == 3


def test_missing_file_is_an_error(tmp_path):
    with pytest.raises(K.KCoreError):
        K.Graph.from_edgelist(str(tmp_path / "nope.txt"))


def test_json_writer_no_nonzero_core(tmp_path):
    j = tmp_path / "c.json"
    K.write_kcore_json(str(j), np.zeros(4, np.uint32))
    assert j.read_bytes() == b"{ "


def test_core_dumps_of_empty_and_large_values(tmp_path):
    t = tmp_path / "c.txt"
    K.write_core_text(str(t), np.zeros(0, np.uint32))
    assert t.read_bytes() == b""
    big = np.array([0, 4294967295, 7, 1000000], np.uint32)
    K.write_core_text(str(t), big)
    assert t.read_bytes() == b"0\n4294967295\n7\n1000000\n"
    with pytest.raises(K.KCoreError):
        K.write_core_text(str(tmp_path / "no" / "such" / "dir.txt"), big)


@pytest.mark.skipif(not os.path.exists(O.REF_LOADER), reason="reference loader not built here")
@pytest.mark.parametrize("gen", ["er", "cl", "rmat"])
def test_generated_text_through_reference_loader(gen, tmp_path):
    g = {"er": lambda: K.Graph.erdos_renyi(3000, 20000, 5),
         "cl": lambda: K.Graph.chung_lu(4000, 6),
         "rmat": lambda: K.Graph.rmat(12, 7)}[gen]()
    txt = tmp_path / f"{gen}.txt"
    g.write_edgelist(str(txt))
    work = tmp_path / "w"
    (work / "output").mkdir(parents=True)
    pre = str(work / "dump")
    subprocess.run([O.REF_LOADER, str(txt), str(work), pre], check=True, stdout=subprocess.DEVNULL)
    off = np.fromfile(pre + ".off", np.uint32)
    nbr = np.fromfile(pre + ".nbr", np.uint32)
    mine = K.Graph.from_edgelist(str(txt))
    # the generated edge list's max id can be below the generator's n: compare what loads
    assert (mine.neighbors_offset == off).all() and (mine.neighbors == nbr).all()
    ser = open(work / "output" / f"serialized-{gen}.txt", "rb").read()
    mine.write_serialized(str(tmp_path / "ser"))
    assert (tmp_path / "ser").read_bytes() == ser


def test_generators_deterministic_across_threads():
    for mk in (lambda nt: K.Graph.erdos_renyi(5000, 40000, 3, nthreads=nt),
               lambda nt: K.Graph.chung_lu(8000, 4, nthreads=nt),
               lambda nt: K.Graph.rmat(13, 9, nthreads=nt)):
        a, b = mk(1), mk(7)
        assert a.V == b.V and (a.neighbors_offset == b.neighbors_offset).all()
        assert (a.neighbors == b.neighbors).all()


def test_generator_shapes():
    g = K.Graph.erdos_renyi(100000, 1000000, 1)
    assert g.arcs == 2_000_000  # exactly m distinct undirected edges
    assert g.V <= 100000
    off, nbr = g.neighbors_offset, g.neighbors
    # symmetric, sorted, loop-free
    src = np.repeat(np.arange(g.V, dtype=np.uint32), np.diff(off))
    assert (src != nbr).all()
    fwd = np.sort(src.astype(np.uint64) << 32 | nbr)
    bwd = np.sort(nbr.astype(np.uint64) << 32 | src)
    assert (fwd == bwd).all()
    assert all((np.diff(nbr[off[v]:off[v + 1]].astype(np.int64)) > 0).all() for v in range(0, g.V, 997))
    cl = K.Graph.chung_lu(400000, 2)
    assert 2_000_000 < cl.m < 2_500_000 and cl.dmax <= 4000
    r = K.Graph.rmat(16, 3)
    assert r.V <= 1 << 16 and r.dmax > 1000


def test_text_roundtrip(tmp_path):
    g = K.Graph.rmat(12, 2)
    p = tmp_path / "g.txt"
    g.write_edgelist(str(p))
    h = K.Graph.from_edgelist(str(p))
    # ids beyond the last non-isolated vertex disappear in text form (V = 1 + max id)
    assert h.V <= g.V
    assert (h.neighbors_offset == g.neighbors_offset[:h.V + 1]).all()
    assert (h.neighbors == g.neighbors).all()


def test_pairs_rules():
    g = K.Graph.from_pairs([0, 1, 1, 3, 2], [1, 0, 1, 3, 4])
    assert g.V == 5  # self loops do not count towards V: max non-loop id is 4
    assert list(g.degrees) == [1, 1, 1, 0, 1]
    e = K.Graph.from_pairs([7], [7])
    assert e.V == 1 and e.arcs == 0


def test_networkx_crosscheck_tool(tmp_path):
    """oracle/nx_crosscheck.py (the reference notebook's check, offline) accepts the JSON core file
    written for the oracle's core numbers and rejects one with a changed core."""
    pytest.importorskip("networkx")
    import oracle.nx_crosscheck as X
    g = K.Graph.chung_lu(3000, 12)
    txt = tmp_path / "g.txt"
    g.write_edgelist(str(txt))
    core, _ = O.bz(g.neighbors_offset, g.neighbors)
    js = tmp_path / "k.json"
    K.write_kcore_json(str(js), core)
    assert X.crosscheck(str(txt), str(js)) == (True, 0)
    bad = core.copy()
    bad[int(np.flatnonzero(bad)[0])] += 1
    K.write_kcore_json(str(js), bad)
    assert X.crosscheck(str(txt), str(js)) == (False, 1)


def test_driver_error_lines_match_reference(tmp_path):
    """`kcore` without an argument, and with a missing file, prints the reference's lines
    (main.cu:45-48, src/graph.cpp:9-33,58-66) and exits with -1 before touching a GPU."""
    exe = os.path.join(os.path.dirname(K._native.HIP_LIB), "..", "bin", "kcore")
    if not os.path.exists(exe):
        pytest.skip("driver not built")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60, cwd=tmp_path)
    assert r.returncode == 255 and r.stdout == "Please provide data file\n"
    r = subprocess.run([exe, "missing.txt"], capture_output=True, text=True, timeout=60, cwd=tmp_path)
    assert r.returncode == 255
    assert r.stdout == ("Graph loading Started... \nserialized-missing.txt : File couldn't open\n"
                        "Reading normal file... \nload graph file failed \n")


def _serial_reader(data: bytes):
    """The serialized-cache reader's definition (src/graph.cpp:9-33 reads with `>>`): whitespace-
    separated decimal tokens V, E, V degrees, V + 1 offsets, E neighbours; a needed token must start
    with a digit and fit in uint32; whatever follows the last needed token is not read.
    Returns (degrees, offsets, neighbours) or None for a malformed file."""
    pos, n = 0, len(data)
    ws = b" \t\n\v\f\r"

    def nxt():
        nonlocal pos
        while pos < n and data[pos] in ws:
            pos += 1
        q = pos
        while pos < n and 48 <= data[pos] <= 57:
            pos += 1
        if pos == q or int(data[q:pos]) > 0xFFFFFFFF:
            return None
        return int(data[q:pos])

    V, E = nxt(), None
    if V is not None:
        E = nxt()
    if V is None or E is None:
        return None
    out = []
    for count in (V, V + 1, E):
        vals = []
        for _ in range(count):
            v = nxt()
            if v is None:
                return None
            vals.append(v)
        out.append(vals)
    return out


def test_serialized_reader_irregular_files(tmp_path):
    """The reader parses large regular caches in parallel and hands anything irregular to the serial
    reader: both must give exactly the definition's result on mutated caches (other whitespace,
    garbage glued to tokens, truncation, oversized values, trailing text)."""
    import random
    g = K.Graph.rmat(11, 7)
    p = tmp_path / "ser"
    g.write_serialized(str(p))
    base = p.read_bytes()
    rng = random.Random(5)
    seps = [b" ", b"\t", b"\n", b"  ", b"\r\n", b"\v", b"\f", b" \t "]
    for case in range(60):
        toks = base.split()
        kind = case % 6
        if kind == 0:    # other whitespace between tokens
            data = toks[0]
            for t in toks[1:]:
                data += rng.choice(seps) + t
        elif kind == 1:  # garbage glued to one token (accepted only on the last needed one)
            i = rng.choice([len(toks) - 1, rng.randrange(2, len(toks))])
            toks[i] = toks[i] + b"x"
            data = b" ".join(toks)
        elif kind == 2:  # truncated
            data = b" ".join(toks[:rng.randrange(2, len(toks))])
        elif kind == 3:  # a value beyond uint32
            toks[rng.randrange(2, len(toks))] = b"4294967296"
            data = b" ".join(toks)
        elif kind == 4:  # trailing text after the last needed token
            data = base + b"trailing words 12 x"
        else:            # a non-numeric token
            toks[rng.randrange(2, len(toks))] = b"-3"
            data = b" ".join(toks)
        q = tmp_path / f"case{case}"
        q.write_bytes(data)
        want = _serial_reader(data)
        if want is None:
            with pytest.raises(K.KCoreError):
                K.Graph.read_serialized(str(q))
            continue
        h = K.Graph.read_serialized(str(q))
        assert list(h.degrees) == want[0] and list(h.neighbors_offset) == want[1], case
        assert list(h.neighbors) == want[2], case


def test_core_writers_across_formatting_rounds(tmp_path):
    """The JSON and text core writers format in parallel rounds of 2^18 vertices per thread; long
    stretches of zero cores leave whole rounds empty. Bytes equal the reference writer's rules."""
    rng = np.random.default_rng(3)
    V = 2_500_000
    core = np.zeros(V, np.uint32)
    core[[7, 600_000, 2_400_000]] = [3, 4, 9]
    core[1_000_000:1_000_300] = rng.integers(1, 50, 300)
    j, t = tmp_path / "c.json", tmp_path / "c.txt"
    K.write_kcore_json(str(j), core)
    nz = np.flatnonzero(core)
    want = "{ " + "".join(f'"{v}": {core[v]}, \n' for v in nz)
    assert j.read_text() == want[:-3] + " }\n"
    K.write_core_text(str(t), core)
    assert t.read_bytes() == "".join(f"{c}\n" for c in core).encode()
