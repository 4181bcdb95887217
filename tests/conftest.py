import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: full-size configuration")
    config.addinivalue_line("markers", "late: GPU test collected after the broad GPU checks")


# GPU tests run broadest first: single-graph parity of the sub-round kernel, batches, serving, the
# driver; then every case of the cascade-queue variant and of the reference-shaped host loop, and
# tests marked `late`; then any other GPU module (collected last, in name order). A failure still
# fails the run, but under `pytest -x` it cannot keep the broader checks from running. CPU tests
# keep their order.
_GPU_ORDER = ("test_gpu_parity", "test_gpu_batch", "test_gpu_serve", "test_gpu_driver", "test_gpu_async")


def pytest_collection_modifyitems(config, items):
    def rank(item):
        if item.get_closest_marker("gpu") is None:
            return (0, 0)
        mod = item.module.__name__.rsplit(".", 1)[-1] if item.module else ""
        if mod not in _GPU_ORDER:
            return (3, 0)
        late = (mod == "test_gpu_async" or "async" in item.nodeid or "stepwise" in item.nodeid
                or item.get_closest_marker("late") is not None)
        return (2 if late else 1, _GPU_ORDER.index(mod))

    items.sort(key=rank)  # stable: the order inside each group is kept


@pytest.fixture(scope="session")
def engine():
    import kcoregpu_amd as K
    eng = K.Engine(0)
    yield eng
    eng.close()


def csr_from_edges(n, edges):
    """Small helper: CSR (offsets, neighbors) of an undirected simple graph on n vertices."""
    adj = [set() for _ in range(n)]
    for a, b in edges:
        if a != b:
            adj[a].add(b)
            adj[b].add(a)
    off = np.zeros(n + 1, np.uint32)
    nbr = []
    for v in range(n):
        s = sorted(adj[v])
        nbr.extend(s)
        off[v + 1] = off[v] + len(s)
    return off, np.array(nbr, np.uint32)


def closed_form_graphs():
    """(name, n, edges, expected core) for graphs whose core numbers are known in closed form."""
    out = []
    for n in (1, 2, 5, 12):
        e = [(i, j) for i in range(n) for j in range(i + 1, n)]
        out.append((f"K{n}", n, e, [n - 1] * n))
    for n in (3, 7, 16):
        out.append((f"cycle{n}", n, [(i, (i + 1) % n) for i in range(n)], [2] * n))
    out.append(("path9", 9, [(i, i + 1) for i in range(8)], [1] * 9))
    out.append(("star15", 16, [(0, i) for i in range(1, 16)], [1] * 16))
    out.append(("isolated", 6, [], [0] * 6))
    for a, b in ((2, 3), (3, 5), (4, 4)):
        e = [(i, a + j) for i in range(a) for j in range(b)]
        out.append((f"K{a},{b}", a + b, e, [min(a, b)] * (a + b)))
    # r-regular circulants C_n(1..r/2)
    for n, r in ((10, 4), (13, 6), (16, 8)):
        e = {(min(i, (i + d) % n), max(i, (i + d) % n)) for i in range(n) for d in range(1, r // 2 + 1)}
        out.append((f"circ{n}r{r}", n, sorted(e), [r] * n))
    # grid 3x4: core 2 everywhere
    R, Cc = 3, 4
    e = []
    for r in range(R):
        for c in range(Cc):
            v = r * Cc + c
            if c + 1 < Cc:
                e.append((v, v + 1))
            if r + 1 < R:
                e.append((v, v + Cc))
    out.append(("grid3x4", R * Cc, e, [2] * (R * Cc)))
    # clique K6 with pendant trees hanging off two clique vertices
    e = [(i, j) for i in range(6) for j in range(i + 1, 6)]
    e += [(0, 6), (6, 7), (6, 8), (3, 9), (9, 10), (10, 11)]
    out.append(("K6+trees", 12, e, [5] * 6 + [1] * 6))
    # disjoint union: K4 + cycle5 + isolated + path3
    e = [(i, j) for i in range(4) for j in range(i + 1, 4)]
    e += [(4 + i, 4 + (i + 1) % 5) for i in range(5)]
    e += [(10, 11), (11, 12)]
    out.append(("union", 13, e, [3] * 4 + [2] * 5 + [0] + [1] * 3))
    return out
