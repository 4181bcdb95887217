"""kcoregpu_amd — k-core decomposition on AMD Instinct MI355X (gfx950).

Python host-side mirror of the reference's interface (paths relative to the reference repo):

* :class:`Graph` mirrors ``class Graph`` (inc/graph.h:7-23, src/graph.cpp): ``V``, ``E`` (arcs),
  ``neighbors``, ``neighbors_offset``, ``degrees``, the text loader with the serialized
  ``./output/serialized-<basename>`` cache, ``writeSerialized`` and ``writeKCoreToDisk``.
* :func:`kcore` mirrors ``int kcore(Graph&)`` (src/ours.cc:117-160) but returns the core numbers
  the reference leaves on the device, with the peel time and ``kmax``.
* :class:`Engine` / :class:`DeviceGraph` expose the C ABI of ``libkcore_hip.so``
  (include/kcore.h): a resident device graph peeled as often as wanted.

Every compute call goes through the HIP library; nothing here computes core numbers itself.
"""
from __future__ import annotations

import atexit
import time
import ctypes as C
import os
import weakref
from dataclasses import dataclass

import numpy as np

from . import _native as N

VARIANTS = {"auto": 0, "wave": 1, "lds": 2, "prefetch": 3, "stepwise": 4, "async": 5}
FLAG_NO_LEVEL_SKIP = 1
FLAG_TRACE = 1 << 8
FLAG_GRAPH = 1 << 9
FLAG_DIRECT_POLL = 1 << 10
FLAG_LEVEL_HANDOFF = 1 << 11

KCG_ERRORS = {1: "I/O error", 2: "vertex id or arc count beyond uint32", 3: "bad argument",
              4: "malformed serialized cache"}


class KCoreError(RuntimeError):
    def __init__(self, status: int, what: str):
        self.status = status
        super().__init__(f"{what}: {status}")


def _u32(a: np.ndarray):
    return a.ctypes.data_as(N.u32p)


class Graph:
    """Host CSR graph with the reference loader's semantics (src/graph.cpp:58-124)."""

    def __init__(self, handle: int):
        self._h = C.c_void_p(handle)
        L = N.host()
        V = C.c_uint32()
        arcs = C.c_uint64()
        dmax = C.c_uint32()
        L.kcg_graph_info(self._h, C.byref(V), C.byref(arcs), C.byref(dmax))
        self.V = V.value
        self.arcs = arcs.value
        self.dmax = dmax.value
        off = N.u32p()
        nbr = N.u32p()
        deg = N.u32p()
        L.kcg_graph_arrays(self._h, C.byref(off), C.byref(nbr), C.byref(deg))
        self.neighbors_offset = np.ctypeslib.as_array(off, shape=(self.V + 1,))
        self.neighbors = (np.ctypeslib.as_array(nbr, shape=(self.arcs,)) if self.arcs
                          else np.zeros(0, np.uint32))
        self.degrees = np.ctypeslib.as_array(deg, shape=(self.V,))
        self.kmax = 0

    # the reference's E is the arc count (src/graph.cpp:114, printed by main.cu:55)
    @property
    def E(self) -> int:
        return self.arcs

    @property
    def m(self) -> int:
        return self.arcs // 2

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            N.host().kcg_graph_free(h)
            self._h = C.c_void_p(0)

    @staticmethod
    def _make(fn, *args, kind_ptr=None) -> "Graph":
        out = C.c_void_p()
        st = fn(*args, C.byref(out)) if kind_ptr is None else fn(*args, C.byref(out), kind_ptr)
        if st != 0:
            raise KCoreError(st, KCG_ERRORS.get(st, "graph error"))
        return Graph(out.value)

    # ---- loaders -------------------------------------------------------------------
    @classmethod
    def from_edgelist(cls, path: str, nthreads: int = 0) -> "Graph":
        return cls._make(N.host().kcg_load_edgelist, os.fsencode(path), nthreads)

    @classmethod
    def load(cls, path: str, output_dir: str = "./output/", nthreads: int = 0) -> "Graph":
        """Reference constructor: serialized cache first, else parse and write the cache."""
        L = N.host()
        out = C.c_void_p()
        hit = C.c_int(0)
        st = L.kcg_load_cached(os.fsencode(path), os.fsencode(output_dir), nthreads, C.byref(out),
                               C.byref(hit))
        if st != 0:
            raise KCoreError(st, KCG_ERRORS.get(st, "graph error"))
        g = Graph(out.value)
        g.from_cache = bool(hit.value)
        return g

    @classmethod
    def load_binary(cls, path: str, output_dir: str = "./output/", nthreads: int = 0) -> "Graph":
        """Like load(), with the content-keyed binary cache (bin-<basename>-<hash>.kcg)."""
        out = C.c_void_p()
        hit = C.c_int(0)
        st = N.host().kcg_load_cached_bin(os.fsencode(path), os.fsencode(output_dir), nthreads,
                                          C.byref(out), C.byref(hit))
        if st != 0:
            raise KCoreError(st, KCG_ERRORS.get(st, "graph error"))
        g = Graph(out.value)
        g.from_cache = bool(hit.value)
        return g

    @classmethod
    def read_bincache(cls, path: str, src_hash: int = 0) -> "Graph":
        return cls._make(N.host().kcg_read_bincache, os.fsencode(path), src_hash)

    def write_bincache(self, path: str, src_hash: int = 0) -> None:
        st = N.host().kcg_write_bincache(self._h, os.fsencode(path), src_hash)
        if st:
            raise KCoreError(st, "write_bincache")

    @classmethod
    def read_serialized(cls, path: str) -> "Graph":
        return cls._make(N.host().kcg_read_serialized, os.fsencode(path))

    @classmethod
    def from_pairs(cls, src, dst, nthreads: int = 0) -> "Graph":
        s = np.ascontiguousarray(src, dtype=np.uint32)
        t = np.ascontiguousarray(dst, dtype=np.uint32)
        if s.shape != t.shape:
            raise ValueError("src and dst differ in length")
        return cls._make(N.host().kcg_from_pairs, _u32(s), _u32(t), s.size, nthreads)

    @classmethod
    def from_csr(cls, offsets, neighbors) -> "Graph":
        o = np.ascontiguousarray(offsets, dtype=np.uint32)
        n = np.ascontiguousarray(neighbors, dtype=np.uint32)
        return cls._make(N.host().kcg_from_csr, o.size - 1, n.size, _u32(o), _u32(n))

    # ---- generators ----------------------------------------------------------------
    @classmethod
    def erdos_renyi(cls, n: int, m: int, seed: int, nthreads: int = 0) -> "Graph":
        return cls._make(N.host().kcg_gen_er, n, m, seed, nthreads)

    @classmethod
    def chung_lu(cls, n: int, seed: int, gamma: float = 2.1, wmin: float = 1.0,
                 wmax: float = 2800.0, avg_deg: float = 12.0, nthreads: int = 0) -> "Graph":
        return cls._make(N.host().kcg_gen_chung_lu, n, gamma, wmin, wmax, avg_deg, seed, nthreads)

    @classmethod
    def rmat(cls, scale: int, seed: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19,
             c: float = 0.19, ndraws: int | None = None, nthreads: int = 0) -> "Graph":
        nd = ndraws if ndraws is not None else edge_factor << scale
        return cls._make(N.host().kcg_gen_rmat, scale, nd, a, b, c, seed, nthreads)

    @classmethod
    def request(cls, seed: int, i: int, median_m: float = 50000.0, sigma: float = 1.0,
                mmin: int = 1000, mmax: int = 50_000_000, nthreads: int = 0) -> "Graph":
        """Request i of a served-workload stream (R-MAT or Chung-Lu by a seeded coin flip, edge
        count lognormal around median_m); .kind is 0 for R-MAT, 1 for Chung-Lu."""
        kind = C.c_int()
        g = cls._make(N.host().kcg_gen_request, seed, i, median_m, sigma, mmin, mmax, nthreads,
                      kind_ptr=C.byref(kind))
        g.kind = kind.value
        return g

    @staticmethod
    def request_shape(seed: int, i: int, median_m: float = 50000.0, sigma: float = 1.0,
                      mmin: int = 1000, mmax: int = 50_000_000):
        m, kind = C.c_uint64(), C.c_int()
        st = N.host().kcg_request_shape(seed, i, median_m, sigma, mmin, mmax, C.byref(m), C.byref(kind))
        if st:
            raise KCoreError(st, "kcg_request_shape")
        return m.value, kind.value

    # ---- writers (reference src/graph.cpp:35-54, :126-140) ---------------------------
    def write_serialized(self, path: str) -> None:
        st = N.host().kcg_write_serialized(self._h, os.fsencode(path))
        if st:
            raise KCoreError(st, "writeSerialized")

    def write_edgelist(self, path: str) -> None:
        st = N.host().kcg_write_edgelist(self._h, os.fsencode(path))
        if st:
            raise KCoreError(st, "write_edgelist")


class Batch:
    """Many graphs packed into one CSR (layout of kc_batch_csr, include/kcore.h)."""

    def __init__(self, handle: int):
        self._h = C.c_void_p(handle)
        L = N.host()
        G = C.c_uint32()
        nv = C.c_uint64()
        na = C.c_uint64()
        L.kcg_batch_info(self._h, C.byref(G), C.byref(nv), C.byref(na))
        self.G, self.total_vertices, self.total_arcs = G.value, nv.value, na.value
        vo, ao, of, nb, dg = N.u64p(), N.u64p(), N.u32p(), N.u32p(), N.u32p()
        L.kcg_batch_arrays(self._h, C.byref(vo), C.byref(ao), C.byref(of), C.byref(nb), C.byref(dg))
        self.vtx_off = np.ctypeslib.as_array(vo, shape=(self.G + 1,))
        self.arc_off = np.ctypeslib.as_array(ao, shape=(self.G + 1,))
        self.offsets = np.ctypeslib.as_array(of, shape=(self.total_vertices + self.G,))
        self.neighbors = (np.ctypeslib.as_array(nb, shape=(self.total_arcs,)) if self.total_arcs
                          else np.zeros(0, np.uint32))
        self.degrees = (np.ctypeslib.as_array(dg, shape=(self.total_vertices,)) if self.total_vertices
                        else np.zeros(0, np.uint32))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            N.host().kcg_batch_free(h)
            self._h = C.c_void_p(0)

    @classmethod
    def generate(cls, G: int, seed: int, nmin: int = 1000, nmax: int = 200000, dmin: float = 4.0,
                 dmax: float = 32.0, gamma: float = 2.1, nthreads: int = 0) -> "Batch":
        out = C.c_void_p()
        st = N.host().kcg_gen_batch(G, nmin, nmax, dmin, dmax, gamma, seed, nthreads, C.byref(out))
        if st:
            raise KCoreError(st, "kcg_gen_batch")
        return cls(out.value)

    @classmethod
    def from_graphs(cls, graphs: list) -> "Batch":
        arr = (C.c_void_p * len(graphs))(*[g._h.value for g in graphs])
        out = C.c_void_p()
        st = N.host().kcg_batch_from_graphs(arr, len(graphs), C.byref(out))
        if st:
            raise KCoreError(st, "kcg_batch_from_graphs")
        return cls(out.value)

    def graph(self, g: int) -> Graph:
        return Graph._make(N.host().kcg_batch_graph, self._h, g)

    def vertices(self, g: int) -> slice:
        return slice(int(self.vtx_off[g]), int(self.vtx_off[g + 1]))


def hash_file(path: str) -> int:
    """The binary cache's 64-bit content hash of a file."""
    h = C.c_uint64()
    st = N.host().kcg_hash_file(os.fsencode(path), C.byref(h))
    if st:
        raise KCoreError(st, "hash_file")
    return h.value


def hash_bytes(data: bytes, seed: int = 0) -> int:
    return N.host().kcg_hash_bytes(data, len(data), seed)


def write_kcore_json(path: str, core: np.ndarray) -> None:
    c = np.ascontiguousarray(core, dtype=np.uint32)
    st = N.host().kcg_write_kcore_json(os.fsencode(path), c.size, _u32(c))
    if st:
        raise KCoreError(st, "writeKCoreToDisk")


def write_core_u32(path: str, core: np.ndarray) -> None:
    """Raw little-endian uint32[V] core dump."""
    c = np.ascontiguousarray(core, dtype=np.uint32)
    st = N.host().kcg_write_core_u32(os.fsencode(path), c.size, _u32(c))
    if st:
        raise KCoreError(st, "write_core_u32")


def write_core_text(path: str, core: np.ndarray) -> None:
    """One core number per line (the form of the serialized cache's degree block)."""
    c = np.ascontiguousarray(core, dtype=np.uint32)
    st = N.host().kcg_write_core_text(os.fsencode(path), c.size, _u32(c))
    if st:
        raise KCoreError(st, "write_core_text")


def splitmix64(x: int) -> int:
    return N.host().kcg_splitmix64(x & 0xFFFFFFFFFFFFFFFF)


# ---- GPU engine ----------------------------------------------------------------------

@dataclass
class PeelResult:
    core: np.ndarray | None
    kmax: int
    stats: dict


def _check(st: int, what: str):
    if st != 0:
        msg = N.hip().kc_status_str(st).decode()
        raise KCoreError(st, f"{what}: {msg}")


_ENGINES = weakref.WeakSet()


@atexit.register
def _close_engines():
    # release device state before the HIP runtime's own exit handlers run
    for e in list(_ENGINES):
        e.close()


class Engine:
    """One HIP device context (kc_open). Not shared across threads."""

    def __init__(self, device: int = 0):
        self._c = C.c_void_p()
        self._graphs = weakref.WeakSet()
        _check(N.hip().kc_open(device, C.byref(self._c)), "kc_open")
        _ENGINES.add(self)

    def close(self):
        # device graphs hold the context's stream: release them first
        for g in list(self._graphs):
            g.free()
        if self._c.value:
            N.hip().kc_close(self._c)
            self._c = C.c_void_p(0)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def selftest(self) -> None:
        _check(N.hip().kc_selftest(self._c), "kc_selftest")

    def copy_bandwidth(self, nbytes: int = 1 << 30, reps: int = 10) -> float:
        """Empirical HBM ceiling in GB/s: the degree-copy kernel over two buffers of nbytes
        (bytes read plus bytes written per second)."""
        g = C.c_double()
        _check(N.hip().kc_copy_bandwidth(self._c, nbytes, reps, C.byref(g)), "kc_copy_bandwidth")
        return g.value

    def upload(self, g: Graph) -> "DeviceGraph":
        h = N.kc_csr(g.V, g.arcs, _u32(g.neighbors_offset),
                     _u32(g.neighbors) if g.arcs else N.u32p(), _u32(g.degrees))
        out = C.c_void_p()
        _check(N.hip().kc_graph_upload(self._c, C.byref(h), C.byref(out)), "kc_graph_upload")
        dg = DeviceGraph(self, out, g.V)
        self._graphs.add(dg)
        return dg

    def upload_batch(self, b: Batch) -> DeviceBatch:
        h = N.kc_batch_csr(b.G, b.total_vertices, b.total_arcs, b.vtx_off.ctypes.data_as(N.u64p),
                           b.arc_off.ctypes.data_as(N.u64p), _u32(b.offsets),
                           _u32(b.neighbors) if b.total_arcs else N.u32p(),
                           _u32(b.degrees) if b.total_vertices else N.u32p())
        out = C.c_void_p()
        _check(N.hip().kc_batch_upload(self._c, C.byref(h), C.byref(out)), "kc_batch_upload")
        db = DeviceBatch(self, out, b)
        self._graphs.add(db)
        return db

    def serve(self, **cfg) -> "Server":
        srv = Server(self, **cfg)
        self._graphs.add(srv)
        return srv


def _tune(fn, eng, handle, reps, flags, cache_path, single):
    cfg = N.kc_tune_cfg(reps, flags, os.fsencode(cache_path) if cache_path else None)
    best, ms, n = N.kc_params(), C.c_double(), C.c_uint32()
    _check(fn(eng._c, handle, C.byref(cfg), C.byref(best), C.byref(ms), C.byref(n)), fn.__name__)
    out = {"block": best.block, "flags": best.flags, "ms": ms.value, "trials": n.value}
    if single:
        out.update(variant=best.variant, heavy_threshold=best.heavy_threshold,
                   chunk_arcs=best.chunk_arcs, lds_depth=best.lds_depth,
                   compact_ratio=best.compact_ratio, poll_max=best.poll_max)
    return out


class _Pinned:
    """Owner of one kc_host_alloc block; freed when the last array viewing it is collected."""

    def __init__(self, nbytes: int):
        self.p = C.c_void_p()
        _check(N.hip().kc_host_alloc(nbytes, C.byref(self.p)), "kc_host_alloc")

    def __del__(self):
        if self.p.value:
            N.hip().kc_host_free(self.p)
            self.p = C.c_void_p(0)


def pinned_empty(n: int, dtype=np.uint32) -> np.ndarray:
    """A numpy array in pinned host memory (device-accessible; kc_host_alloc)."""
    dt = np.dtype(dtype)
    owner = _Pinned(max(n, 1) * dt.itemsize)
    buf = (C.c_char * (max(n, 1) * dt.itemsize)).from_address(owner.p.value)
    buf._owner = owner  # the ctypes buffer keeps the allocation alive
    return np.frombuffer(buf, dtype=dt, count=n)


class Server:
    """Served peeling (kc_serve_*): a resident persistent kernel that peels requests as they
    arrive, each in its own region of an HBM arena. submit() never blocks on the GPU; poll()
    returns finished requests with their core numbers. With pinned=True the core numbers land
    in pinned buffers written by the kernel itself; otherwise they are copied down."""

    def __init__(self, eng: "Engine", hbm_budget: int = 0, ring_slots: int = 0, block: int = 0,
                 workgroups: int = 0, flags: int = 0, lease_ms: int = 0, pinned: bool = True,
                 large_arcs: int = 0, large_lanes: int = 0):
        self.eng = eng
        self._s = C.c_void_p()
        self.pinned = pinned
        cfg = N.kc_serve_cfg(hbm_budget, ring_slots, block, workgroups, flags, lease_ms, large_arcs,
                             large_lanes)
        _check(N.hip().kc_serve_open(eng._c, C.byref(cfg), C.byref(self._s)), "kc_serve_open")
        self._live = {}  # req_id -> (graph, core buffer): kept alive until completion
        self._next = 0

    def submit(self, g: Graph, req_id: int | None = None) -> int:
        if req_id is None:
            req_id = self._next
        self._next = max(self._next, req_id + 1)
        if req_id in self._live:
            raise ValueError(f"request id {req_id} is still in flight")
        core = pinned_empty(g.V) if self.pinned else np.empty(g.V, np.uint32)
        h = N.kc_csr(g.V, g.arcs, _u32(g.neighbors_offset),
                     _u32(g.neighbors) if g.arcs else N.u32p(), N.u32p())
        _check(N.hip().kc_serve_submit(self._s, req_id, C.byref(h), _u32(core)), "kc_serve_submit")
        self._live[req_id] = (g, core)
        return req_id

    def poll(self, cap: int = 256) -> list:
        buf = (N.kc_completion * cap)()
        n = C.c_uint32()
        _check(N.hip().kc_serve_poll(self._s, buf, cap, C.byref(n)), "kc_serve_poll")
        out = []
        for i in range(n.value):
            c = buf[i]
            _, core = self._live.pop(c.req_id)
            out.append({"req_id": c.req_id, "status": c.status, "core": core, "kmax": c.kmax,
                        "levels": c.levels, "rounds": c.rounds, "t_submit": c.t_submit,
                        "t_admit": c.t_admit, "t_done": c.t_done})
        return out

    def pending(self):
        q, f, used = C.c_uint32(), C.c_uint32(), C.c_uint64()
        _check(N.hip().kc_serve_pending(self._s, C.byref(q), C.byref(f), C.byref(used)), "kc_serve_pending")
        return q.value, f.value, used.value

    def drain(self, timeout_s: float = 60.0) -> list:
        """Polls until every submitted request has completed."""
        out = []
        t0 = time.monotonic()
        while self._live:
            out += self.poll()
            if time.monotonic() - t0 > timeout_s:
                raise TimeoutError(f"{len(self._live)} requests unfinished after {timeout_s} s")
        return out

    def close(self):
        if self._s.value:
            s, self._s = self._s, C.c_void_p(0)
            _check(N.hip().kc_serve_close(s), "kc_serve_close")

    free = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBatch:
    def __init__(self, eng: "Engine", handle: C.c_void_p, batch: Batch):
        self.eng = eng
        self._b = handle
        self.G = batch.G
        self.total_vertices = batch.total_vertices

    def free(self):
        if self._b.value:
            N.hip().kc_batch_free(self._b)
            self._b = C.c_void_p(0)

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def tune(self, reps: int = 2, flags: int = 0, cache_path: str | None = None) -> dict:
        """Autotuned batch settings (kc_tune_batch): {'block', 'flags', 'ms', 'trials'}."""
        return _tune(N.hip().kc_tune_batch, self.eng, self._b, reps, flags, cache_path, single=False)

    def peel(self, block: int = 0, flags: int = 0, download: bool = True):
        p = N.kc_params(0, block, 0, 0, flags, 0)
        st = N.kc_stats()
        core = np.empty(self.total_vertices, np.uint32) if download else None
        kmax = np.empty(self.G, np.uint32)
        _check(N.hip().kc_batch_peel(self.eng._c, self._b, C.byref(p),
                                     _u32(core) if download else N.u32p(), _u32(kmax),
                                     C.byref(st)), "kc_batch_peel")
        return core, kmax, st.as_dict()

    def order(self) -> np.ndarray:
        out = np.empty(self.total_vertices, np.uint32)
        _check(N.hip().kc_batch_download_order(self.eng._c, self._b, _u32(out)), "batch order")
        return out


class DeviceGraph:
    def __init__(self, eng: Engine, handle: C.c_void_p, V: int):
        self.eng = eng
        self._g = handle
        self.V = V
        self.kmax = None

    def free(self):
        if self._g.value:
            N.hip().kc_graph_free(self._g)
            self._g = C.c_void_p(0)

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def peel(self, variant: str = "auto", block: int = 0, lds_depth: int = 0,
             heavy_threshold: int = 0, flags: int = 0, download: bool = True,
             chunk_arcs: int = 0, compact_ratio: int = 0, poll_max: int = 0) -> PeelResult:
        v = VARIANTS[variant] if isinstance(variant, str) else int(variant)
        p = N.kc_params(v, block, lds_depth, heavy_threshold, flags, chunk_arcs, compact_ratio, poll_max)
        st = N.kc_stats()
        kmax = C.c_uint32()
        core = np.empty(self.V, np.uint32) if download else None
        _check(N.hip().kc_peel(self.eng._c, self._g, C.byref(p),
                               _u32(core) if download else N.u32p(), C.byref(kmax),
                               C.byref(st)), "kc_peel")
        self.kmax = kmax.value
        return PeelResult(core, kmax.value, st.as_dict())

    def tune(self, reps: int = 2, flags: int = 0, cache_path: str | None = None) -> dict:
        """Autotuned peel settings (kc_tune_peel): a dict of peel() keyword arguments, plus
        'ms' (median device time of the winner) and 'trials'."""
        return _tune(N.hip().kc_tune_peel, self.eng, self._g, reps, flags, cache_path, single=True)

    def core(self) -> np.ndarray:
        out = np.empty(self.V, np.uint32)
        _check(N.hip().kc_graph_download_core(self.eng._c, self._g, _u32(out)), "download_core")
        return out

    def trace(self) -> list[tuple[str, int, int, float]]:
        """Barrier timestamps of the last peel run with FLAG_TRACE: (kind, level, subround, us)."""
        cap = 1 << 16
        rec = np.zeros(2 * cap, np.uint64)
        n = C.c_uint32()
        _check(N.hip().kc_graph_download_trace(self.eng._c, self._g, rec.ctypes.data_as(N.u64p),
                                                cap, C.byref(n)), "download_trace")
        khz = N.hip().kc_wall_clock_khz(self.eng._c)
        kinds = {1: "scan", 2: "sub", 3: "level"}
        out = []
        for i in range(n.value):
            w, t = int(rec[2 * i]), int(rec[2 * i + 1])
            out.append((kinds.get(w >> 56, "?"), w & 0xFFFFFFFF, (w >> 32) & 0xFFFFFF, t * 1e3 / khz))
        out.sort(key=lambda r: r[3])
        return out

    def order(self) -> tuple[np.ndarray, np.ndarray]:
        if self.kmax is None:
            raise RuntimeError("peel first")
        order = np.empty(self.V, np.uint32)
        shells = np.empty(self.kmax + 2, np.uint32)
        _check(N.hip().kc_graph_download_order(self.eng._c, self._g, _u32(order), _u32(shells)),
               "download_order")
        return order, shells


def kcore(g: Graph, variant: str = "auto", engine: Engine | None = None) -> PeelResult:
    """``int kcore(Graph&)`` (src/ours.cc:117-160), returning core numbers, kmax and timings."""
    eng = engine or Engine()
    dg = eng.upload(g)
    try:
        r = dg.peel(variant)
    finally:
        dg.free()
    g.kmax = r.kmax
    return r
