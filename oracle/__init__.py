"""CPU oracle for k-core peeling — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this package,
and only as the checker or the CPU baseline. It binds liboracle.so (oracle.c, built by
oracle/Makefile). Function-by-function pinning is described in oracle.h.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LOADER = os.path.join(HERE, "_ref", "ref_loader")

u32p = C.POINTER(C.c_uint32)


class kco_level_stat(C.Structure):
    _fields_ = [("level", C.c_uint32), ("selected", C.c_uint32), ("peeled", C.c_uint32),
                ("rounds", C.c_uint32), ("arcs", C.c_uint64), ("decrements", C.c_uint64)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} not built; run `make -C {HERE} liboracle.so`")
        L = C.CDLL(LIB)
        L.kco_bz.argtypes = [C.c_uint32, u32p, u32p, u32p]
        L.kco_bz.restype = C.c_uint32
        L.kco_levelsync.argtypes = [C.c_uint32, u32p, u32p, u32p, C.c_int, u32p, u32p, u32p, u32p,
                                    C.POINTER(kco_level_stat), C.c_uint32, u32p]
        L.kco_bruteforce.argtypes = [C.c_uint32, u32p, u32p, u32p]
        L.kco_bruteforce.restype = C.c_uint32
        L.kco_check_certificate.argtypes = [C.c_uint32, u32p, u32p, u32p, u32p]
        L.kco_check_lower.argtypes = [C.c_uint32, u32p, u32p, u32p]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(u32p)


def _csr(off, nbr):
    off = np.ascontiguousarray(off, dtype=np.uint32)
    nbr = np.ascontiguousarray(nbr, dtype=np.uint32)
    if nbr.size == 0:
        nbr = np.zeros(1, np.uint32)
    return off, nbr


def bz(off, nbr):
    """Batagelj-Zaversnik core numbers -> (core, kmax)."""
    off, nbr = _csr(off, nbr)
    V = off.size - 1
    core = np.empty(max(V, 1), np.uint32)
    kmax = lib().kco_bz(V, _p(off), _p(nbr), _p(core))
    return core[:V], int(kmax)


def levelsync(off, nbr, degrees=None, nthreads=0, want_order=False, want_stats=False):
    """Reference-method level-synchronous peel (src/ours.cc) -> dict."""
    off, nbr = _csr(off, nbr)
    V = off.size - 1
    deg = (np.diff(off).astype(np.uint32) if degrees is None
           else np.ascontiguousarray(degrees, dtype=np.uint32))
    dmax = int(deg.max()) if V else 0
    core = np.empty(max(V, 1), np.uint32)
    kmax = C.c_uint32()
    nlev = C.c_uint32()
    order = np.empty(max(V, 1), np.uint32) if want_order else None
    shells = np.empty(dmax + 2, np.uint32) if want_order else None
    cap = dmax + 2
    stats = (kco_level_stat * cap)() if want_stats else None
    rc = lib().kco_levelsync(V, _p(off), _p(nbr), _p(deg), nthreads, _p(core), C.byref(kmax),
                             _p(order) if want_order else None,
                             _p(shells) if want_order else None,
                             stats, cap if want_stats else 0, C.byref(nlev))
    if rc != 0:
        raise RuntimeError("levelsync: no-progress guard tripped")
    out = {"core": core[:V], "kmax": kmax.value, "levels": nlev.value}
    if want_order:
        out["order"] = order[:V]
        out["shells"] = shells[:kmax.value + 2]
    if want_stats:
        out["stats"] = [dict((k, getattr(stats[i], k)) for k, _ in kco_level_stat._fields_)
                        for i in range(nlev.value)]
    return out


def bruteforce(off, nbr):
    off, nbr = _csr(off, nbr)
    V = off.size - 1
    if V > 16:
        raise ValueError("bruteforce needs n <= 16")
    core = np.empty(max(V, 1), np.uint32)
    kmax = lib().kco_bruteforce(V, _p(off), _p(nbr), _p(core))
    return core[:V], int(kmax)


def check_certificate(off, nbr, core, order) -> int:
    off, nbr = _csr(off, nbr)
    c = np.ascontiguousarray(core, dtype=np.uint32)
    o = np.ascontiguousarray(order, dtype=np.uint32)
    return lib().kco_check_certificate(off.size - 1, _p(off), _p(nbr), _p(c), _p(o))


def check_lower(off, nbr, core) -> int:
    off, nbr = _csr(off, nbr)
    c = np.ascontiguousarray(core, dtype=np.uint32)
    return lib().kco_check_lower(off.size - 1, _p(off), _p(nbr), _p(c))
