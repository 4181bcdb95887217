"""Writes the committed golden fixtures under tests/golden/. Run where the reference sources and
networkx are available (the fixtures themselves need neither):

    python tests/golden/make_golden.py

1. nx_cores.npz — core numbers from networkx.core_number (networkx 3.4.2, the external oracle
   the reference's notebook used: .ipynb_checkpoints/test-checkpoint.ipynb cells 3-5) for
   small seeded graphs, stored as (offsets, neighbors, core) triples.
2. loader/<case>.txt + loader/<case>.expected/ — edge-list inputs covering the loader's rules
   (src/graph.cpp:58-124) and the outputs of the reference's own loader source, compiled by
   oracle/Makefile into oracle/_ref/ref_loader: the serialized cache bytes
   (src/graph.cpp:35-54), the CSR arrays, and the JSON core file of writeKCoreToDisk
   (src/graph.cpp:126-140) for the BZ core numbers of that graph.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import kcoregpu_amd as K  # noqa: E402
import oracle as O  # noqa: E402

LOADER_CASES = {
    # name: text content
    "basic": "0 1\n1 2\n2 0\n2 3\n",
    "selfloop_dup": "0 1\n1 0\n0 1\n2 2\n1 2\n3 3\n",
    "maxid_defines_v": "5 2\n2 7\n7 5\n",
    "alpha_comment": "Comment line\nV 9\n0 1\n1 2\n",
    "hash_comment_quirk": "0 5\n# a hash line keeps the previous target\n3 4\n",
    "percent_quirk": "2 6\n% percent line\n1 3\n",
    "blank_line_quirk": "0 1\n\n2 3\n\n",
    "crlf_and_tabs": "0\t1\r\n1  2 \r\n 2\t3\r\n",
    "extra_columns": "0 1 17\n1 2 0.5\n2 3 x\n",
    "unsorted_input": "9 3\n3 1\n9 1\n4 9\n1 4\n0 9\n",
    "no_trailing_newline": "0 1\n1 2",
    "isolated_gaps": "0 10\n10 20\n",
    "clique5_pendant": "0 1\n0 2\n0 3\n0 4\n1 2\n1 3\n1 4\n2 3\n2 4\n3 4\n4 5\n5 6\n5 7\n",
    "only_selfloops": "3 3\n",
}


def nx_fixtures():
    import networkx as nx
    graphs = []
    specs = [("er", 30, 60, 1), ("er", 60, 200, 2), ("er", 120, 700, 3), ("cl", 200, 4, 0),
             ("cl", 500, 5, 0), ("rmat", 8, 6, 0), ("rmat", 9, 7, 0), ("er", 16, 40, 8),
             ("er", 250, 3000, 9), ("rmat", 10, 10, 0)]
    for kind, a, b, c in specs:
        if kind == "er":
            g = K.Graph.erdos_renyi(a, b, c)
        elif kind == "cl":
            g = K.Graph.chung_lu(a, b, avg_deg=8.0, wmax=40.0)
        else:
            g = K.Graph.rmat(a, b, edge_factor=8)
        G = nx.Graph()
        G.add_nodes_from(range(g.V))
        off, nbr = g.neighbors_offset, g.neighbors
        for v in range(g.V):
            for u in nbr[off[v]:off[v + 1]]:
                G.add_edge(v, int(u))
        cn = nx.core_number(G)
        core = np.array([cn[v] for v in range(g.V)], np.uint32)
        graphs.append((off.copy(), nbr.copy(), core))
    arrs = {}
    for i, (o, n, c) in enumerate(graphs):
        arrs[f"off{i}"] = o
        arrs[f"nbr{i}"] = n
        arrs[f"core{i}"] = c
    np.savez_compressed(os.path.join(HERE, "nx_cores.npz"), count=len(graphs), **arrs)
    print("nx_cores.npz:", len(graphs), "graphs")


def loader_fixtures():
    if not os.path.exists(O.REF_LOADER):
        sys.exit(f"{O.REF_LOADER} missing: run `make -C oracle` with the reference present")
    ldir = os.path.join(HERE, "loader")
    os.makedirs(ldir, exist_ok=True)
    for name, text in LOADER_CASES.items():
        src = os.path.join(ldir, name + ".txt")
        with open(src, "w", newline="") as f:
            f.write(text)
        exp = os.path.join(ldir, name + ".expected")
        shutil.rmtree(exp, ignore_errors=True)
        os.makedirs(exp)
        with tempfile.TemporaryDirectory() as work:
            os.makedirs(os.path.join(work, "output"))
            pre = os.path.join(work, "dump")
            subprocess.run([O.REF_LOADER, src, work, pre], check=True, stdout=subprocess.DEVNULL)
            meta = np.fromfile(pre + ".meta", np.uint32)
            off = np.fromfile(pre + ".off", np.uint32)
            nbr = np.fromfile(pre + ".nbr", np.uint32)
            core, _ = O.bz(off, nbr)
            core.astype(np.uint32).tofile(os.path.join(work, "core.u32"))
            # the JSON writer needs ./output/pkc-kcore-<name>: a second run reads the cache
            subprocess.run([O.REF_LOADER, src, work, pre, os.path.join(work, "core.u32"),
                            name], check=True, stdout=subprocess.DEVNULL)
            shutil.copy(os.path.join(work, "output", "serialized-" + name + ".txt"),
                        os.path.join(exp, "serialized"))
            shutil.copy(os.path.join(work, "output", "pkc-kcore-" + name),
                        os.path.join(exp, "pkc-kcore.json"))
            np.savez(os.path.join(exp, "csr.npz"), V=meta[0], E=meta[1], off=off, nbr=nbr,
                     deg=np.fromfile(pre + ".deg", np.uint32), core=core)
    print("loader fixtures:", len(LOADER_CASES))


if __name__ == "__main__":
    nx_fixtures()
    loader_fixtures()
