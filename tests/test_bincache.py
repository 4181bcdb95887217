"""The binary CSR cache (no reference counterpart; it sits beside the reference's text cache,
src/graph.cpp:9-54): round trip, content keying (no basename collisions), and refusal of any
cache whose header, hashes or arrays do not check out."""
import os

import numpy as np
import pytest

import kcoregpu_amd as K


def same_csr(a, b):
    return (a.V == b.V and a.arcs == b.arcs and (a.neighbors_offset == b.neighbors_offset).all()
            and (a.neighbors == b.neighbors).all() and (a.degrees == b.degrees).all() and a.dmax == b.dmax)


def test_hash_properties(tmp_path):
    data = bytes(range(256)) * 41 + b"tail"
    h = K.hash_bytes(data)
    assert h == K.hash_bytes(data) and h != K.hash_bytes(data, seed=1)
    # every single-bit flip, the length and the tail bytes all change the hash
    for i in (0, 7, 8, 100, len(data) - 1):
        flipped = bytearray(data)
        flipped[i] ^= 1
        assert K.hash_bytes(bytes(flipped)) != h
    assert K.hash_bytes(data + b"\0") != h
    assert len({K.hash_bytes(data[:n]) for n in range(0, 40)}) == 40
    # a file hashes like its bytes, whatever the read chunking (here several 8 MiB chunks)
    big = np.random.default_rng(3).integers(0, 256, (8 << 20) * 2 + 12345, np.uint8).tobytes()
    p = tmp_path / "f.bin"
    p.write_bytes(big)
    assert K.hash_file(str(p)) == K.hash_bytes(big)


def test_round_trip_and_cache_hit(tmp_path):
    g = K.Graph.erdos_renyi(5000, 40000, 9)
    src = tmp_path / "g.txt"
    g.write_edgelist(str(src))
    out = tmp_path / "out"
    out.mkdir()
    a = K.Graph.load_binary(str(src), str(out))
    assert not a.from_cache
    ref = K.Graph.from_edgelist(str(src))
    assert same_csr(a, ref)
    files = os.listdir(out)
    assert files == [f"bin-g.txt-{K.hash_file(str(src)):016x}.kcg"]
    b = K.Graph.load_binary(str(src), str(out))
    assert b.from_cache and same_csr(b, ref)
    # the file itself: 48-byte header + offsets + neighbours
    assert os.path.getsize(out / files[0]) == 48 + 4 * (ref.V + 1) + 4 * ref.arcs


def test_same_basename_different_content(tmp_path):
    out = tmp_path / "out"
    out.mkdir()
    (tmp_path / "a").mkdir()
    (tmp_path / "b").mkdir()
    g1 = K.Graph.erdos_renyi(1000, 5000, 1)
    g2 = K.Graph.erdos_renyi(1000, 5000, 2)
    g1.write_edgelist(str(tmp_path / "a" / "g.txt"))
    g2.write_edgelist(str(tmp_path / "b" / "g.txt"))
    x = K.Graph.load_binary(str(tmp_path / "a" / "g.txt"), str(out))
    y = K.Graph.load_binary(str(tmp_path / "b" / "g.txt"), str(out))
    assert not x.from_cache and not y.from_cache  # distinct keys: no collision
    assert same_csr(x, K.Graph.from_edgelist(str(tmp_path / "a" / "g.txt")))
    assert same_csr(y, K.Graph.from_edgelist(str(tmp_path / "b" / "g.txt")))
    assert len(os.listdir(out)) == 2


@pytest.mark.parametrize("damage", ["payload", "magic", "version", "truncate", "extra", "offsets"])
def test_damaged_cache_refused_and_rebuilt(tmp_path, damage):
    g = K.Graph.chung_lu(3000, 4)
    src = tmp_path / "g.txt"
    g.write_edgelist(str(src))
    out = tmp_path / "out"
    out.mkdir()
    ref = K.Graph.load_binary(str(src), str(out))
    path = out / os.listdir(out)[0]
    raw = bytearray(path.read_bytes())
    if damage == "payload":
        raw[48 + 4 * (ref.V + 1) + 17] ^= 0x40
    elif damage == "magic":
        raw[0] ^= 1
    elif damage == "version":
        raw[8] = 2
    elif damage == "truncate":
        raw = raw[:-4]
    elif damage == "extra":
        raw += b"\0\0\0\0"
    elif damage == "offsets":
        # a non-monotone offsets array with a matching payload hash must still be refused
        arr = np.frombuffer(bytes(raw[48:]), np.uint32).copy()
        arr[5], arr[6] = arr[6] + 1, arr[5]
        payload = arr.tobytes()
        raw = raw[:48] + payload
        # recompute the payload hash the way the writer does (offsets bytes, then neighbours)
        ph = K.hash_bytes(payload, seed=1)
        raw[32:40] = int(ph).to_bytes(8, "little")
    path.write_bytes(bytes(raw))
    with pytest.raises(K.KCoreError):
        K.Graph.read_bincache(str(path))
    again = K.Graph.load_binary(str(src), str(out))
    assert not again.from_cache and same_csr(again, ref)
    assert K.Graph.load_binary(str(src), str(out)).from_cache  # rewritten


def test_source_hash_must_match(tmp_path):
    g = K.Graph.erdos_renyi(500, 2000, 5)
    p = tmp_path / "c.kcg"
    g.write_bincache(str(p), src_hash=1234)
    assert same_csr(K.Graph.read_bincache(str(p)), g)
    assert same_csr(K.Graph.read_bincache(str(p), 1234), g)
    with pytest.raises(K.KCoreError):
        K.Graph.read_bincache(str(p), 999)
