"""networkx cross-check of a core-number JSON file — TEST INFRASTRUCTURE, offline only.

The reference's intended verification flow (its notebook, .ipynb_checkpoints/test-checkpoint.ipynb
cells 1-5) reads the edge list with networkx, computes nx.core_number and compares it with the
k-core output file. This tool does the same for the JSON dictionary that `kcore <edgelist> --json`
(and Graph::writeKCoreToDisk, src/graph.cpp:126-140) writes: vertices with a non-zero core number,
`{ "v": c, ... }`. Vertices absent from the file have core 0. It applies the reference loader's
rules to the edge list (lines starting with a letter skipped, self loops dropped) so the vertex set
matches the loaded graph.

    python oracle/nx_crosscheck.py <edgelist> <json>      -> prints "Test Passed" / "Test Failed!"

networkx is a third-party oracle (3.4.2 here); it is not installed on every machine and is not
needed by anything else in this repository.
"""
from __future__ import annotations

import json
import sys


def read_edges(path: str):
    """Pairs as `iss >> s >> t` reads them (src/graph.cpp:68-96): a field that fails to parse is
    set to 0 and stops the line, so a line without a leading number gives s = 0 and keeps the
    previous t; self loops are dropped."""
    edges = []
    s = t = 0
    with open(path) as f:
        for line in f:
            if line[:1].isalpha():
                continue
            parts = line.split()
            if parts and parts[0].isdigit():
                s = int(parts[0])
                t = int(parts[1]) if len(parts) > 1 and parts[1].isdigit() else 0
            else:
                s = 0
            if s != t:
                edges.append((s, t))
    return edges


def crosscheck(edgelist: str, json_path: str) -> tuple[bool, int]:
    import networkx as nx

    g = nx.Graph()
    edges = read_edges(edgelist)
    g.add_edges_from(edges)
    V = 1 + max((max(a, b) for a, b in edges), default=-1)
    g.add_nodes_from(range(max(V, 1)))
    want = nx.core_number(g)
    text = open(json_path).read().strip()
    got = {} if text in ("{", "{ }", "") else {int(k): int(v) for k, v in json.loads(text).items()}
    bad = sum(1 for v in range(max(V, 1)) if got.get(v, 0) != want[v])
    bad += sum(1 for v in got if v >= max(V, 1))
    return bad == 0, bad


def main(argv):
    if len(argv) != 3:
        print(__doc__)
        return 2
    ok, bad = crosscheck(argv[1], argv[2])
    print("Test Passed" if ok else f"Test Failed! ({bad} vertices differ)")
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main(sys.argv))
