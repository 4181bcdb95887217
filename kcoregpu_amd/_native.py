"""ctypes loading of the in-tree native libraries (built by kcoregpu_amd/csrc/Makefile).

The libraries are never replaced by a Python fallback: a missing library raises, naming the
build command, so a GPU run can only ever execute the HIP path.
"""
from __future__ import annotations

import ctypes as C
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(PKG_DIR, "lib")
BIN_DIR = os.path.join(PKG_DIR, "bin")
HOST_LIB = os.path.join(LIB_DIR, "libkcore_host.so")
HIP_LIB = os.path.join(LIB_DIR, "libkcore_hip.so")

u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)


class NativeMissing(RuntimeError):
    pass


def _load(path: str) -> C.CDLL:
    if not os.path.exists(path):
        raise NativeMissing(
            f"{path} is not built; run `make -C {os.path.join(PKG_DIR, 'csrc')}` "
            "(or __graft_entry__.build())")
    return C.CDLL(path)


_host = None
_hip = None


class kc_csr(C.Structure):
    _fields_ = [("V", C.c_uint32), ("arcs", C.c_uint64), ("offsets", u32p),
                ("neighbors", u32p), ("degrees", u32p)]


class kc_batch_csr(C.Structure):
    _fields_ = [("G", C.c_uint32), ("total_vertices", C.c_uint64), ("total_arcs", C.c_uint64),
                ("vtx_off", u64p), ("arc_off", u64p), ("offsets", u32p), ("neighbors", u32p),
                ("degrees", u32p)]


class kc_params(C.Structure):
    _fields_ = [("variant", C.c_int), ("block", C.c_uint32), ("lds_depth", C.c_uint32),
                ("heavy_threshold", C.c_uint32), ("flags", C.c_uint32), ("chunk_arcs", C.c_uint32),
                ("compact_ratio", C.c_uint32), ("poll_max", C.c_uint32)]


class kc_stats(C.Structure):
    _fields_ = [("ms_peel", C.c_double), ("ms_h2d", C.c_double), ("ms_d2h", C.c_double),
                ("ms_kernel", C.c_double),
                ("kmax", C.c_uint32), ("levels", C.c_uint32), ("subrounds", C.c_uint32),
                ("scanned_entries", C.c_uint64),
                ("bytes_alg", C.c_uint64), ("bytes_scan", C.c_uint64), ("bytes_edge", C.c_uint64),
                ("grid", C.c_uint32), ("block", C.c_uint32), ("participants", C.c_uint32)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class kc_serve_cfg(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("ring_slots", C.c_uint32), ("block", C.c_uint32),
                ("workgroups", C.c_uint32), ("flags", C.c_uint32), ("lease_ms", C.c_uint32),
                ("large_arcs", C.c_uint32), ("large_lanes", C.c_uint32)]


class kc_completion(C.Structure):
    _fields_ = [("req_id", C.c_uint64), ("status", C.c_int), ("kmax", C.c_uint32),
                ("levels", C.c_uint32), ("rounds", C.c_uint32), ("t_submit", C.c_double),
                ("t_admit", C.c_double), ("t_done", C.c_double)]


class kc_tune_cfg(C.Structure):
    _fields_ = [("reps", C.c_uint32), ("flags", C.c_uint32), ("cache_path", C.c_char_p)]


def host() -> C.CDLL:
    global _host
    if _host is None:
        L = _load(HOST_LIB)
        vp = C.c_void_p
        pp = C.POINTER(C.c_void_p)
        L.kcg_load_edgelist.argtypes = [C.c_char_p, C.c_int, pp]
        L.kcg_read_serialized.argtypes = [C.c_char_p, pp]
        L.kcg_write_serialized.argtypes = [vp, C.c_char_p]
        L.kcg_load_cached.argtypes = [C.c_char_p, C.c_char_p, C.c_int, pp, C.POINTER(C.c_int)]
        L.kcg_from_pairs.argtypes = [u32p, u32p, C.c_uint64, C.c_int, pp]
        L.kcg_from_csr.argtypes = [C.c_uint32, C.c_uint64, u32p, u32p, pp]
        L.kcg_graph_free.argtypes = [vp]
        L.kcg_graph_free.restype = None
        L.kcg_graph_info.argtypes = [vp, u32p, u64p, u32p]
        L.kcg_graph_info.restype = None
        L.kcg_graph_arrays.argtypes = [vp, C.POINTER(u32p), C.POINTER(u32p), C.POINTER(u32p)]
        L.kcg_graph_arrays.restype = None
        L.kcg_write_kcore_json.argtypes = [C.c_char_p, C.c_uint32, u32p]
        L.kcg_write_core_u32.argtypes = [C.c_char_p, C.c_uint32, u32p]
        L.kcg_write_core_text.argtypes = [C.c_char_p, C.c_uint32, u32p]
        L.kcg_gen_er.argtypes = [C.c_uint32, C.c_uint64, C.c_uint64, C.c_int, pp]
        L.kcg_gen_chung_lu.argtypes = [C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_double,
                                       C.c_uint64, C.c_int, pp]
        L.kcg_gen_rmat.argtypes = [C.c_uint32, C.c_uint64, C.c_double, C.c_double, C.c_double,
                                   C.c_uint64, C.c_int, pp]
        L.kcg_write_edgelist.argtypes = [vp, C.c_char_p]
        L.kcg_splitmix64.argtypes = [C.c_uint64]
        L.kcg_splitmix64.restype = C.c_uint64
        L.kcg_gen_batch.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_double, C.c_double,
                                    C.c_double, C.c_uint64, C.c_int, pp]
        L.kcg_batch_from_graphs.argtypes = [C.POINTER(vp), C.c_uint32, pp]
        L.kcg_batch_free.argtypes = [vp]
        L.kcg_batch_free.restype = None
        L.kcg_batch_info.argtypes = [vp, u32p, u64p, u64p]
        L.kcg_batch_info.restype = None
        L.kcg_batch_arrays.argtypes = [vp, C.POINTER(u64p), C.POINTER(u64p), C.POINTER(u32p),
                                       C.POINTER(u32p), C.POINTER(u32p)]
        L.kcg_batch_arrays.restype = None
        L.kcg_batch_graph.argtypes = [vp, C.c_uint32, pp]
        L.kcg_hash_file.argtypes = [C.c_char_p, u64p]
        L.kcg_hash_bytes.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64]
        L.kcg_hash_bytes.restype = C.c_uint64
        L.kcg_write_bincache.argtypes = [vp, C.c_char_p, C.c_uint64]
        L.kcg_read_bincache.argtypes = [C.c_char_p, C.c_uint64, pp]
        L.kcg_load_cached_bin.argtypes = [C.c_char_p, C.c_char_p, C.c_int, pp, C.POINTER(C.c_int)]
        L.kcg_request_shape.argtypes = [C.c_uint64, C.c_uint64, C.c_double, C.c_double, C.c_uint64,
                                        C.c_uint64, u64p, C.POINTER(C.c_int)]
        L.kcg_gen_request.argtypes = [C.c_uint64, C.c_uint64, C.c_double, C.c_double, C.c_uint64,
                                      C.c_uint64, C.c_int, pp, C.POINTER(C.c_int)]
        _host = L
    return _host


def hip() -> C.CDLL:
    global _hip
    if _hip is None:
        L = _load(HIP_LIB)
        vp = C.c_void_p
        pp = C.POINTER(C.c_void_p)
        L.kc_status_str.argtypes = [C.c_int]
        L.kc_status_str.restype = C.c_char_p
        L.kc_open.argtypes = [C.c_int, pp]
        L.kc_close.argtypes = [vp]
        L.kc_close.restype = None
        L.kc_graph_upload.argtypes = [vp, C.POINTER(kc_csr), pp]
        L.kc_graph_free.argtypes = [vp]
        L.kc_graph_free.restype = None
        L.kc_graph_info.argtypes = [vp, u32p, u64p, u32p]
        L.kc_graph_info.restype = None
        L.kc_peel.argtypes = [vp, vp, C.POINTER(kc_params), u32p, u32p, C.POINTER(kc_stats)]
        L.kc_graph_download_core.argtypes = [vp, vp, u32p]
        L.kc_graph_download_order.argtypes = [vp, vp, u32p, u32p]
        L.kc_graph_core_device.argtypes = [vp]
        L.kc_graph_core_device.restype = vp
        L.kc_selftest.argtypes = [vp]
        L.kc_graph_download_trace.argtypes = [vp, vp, u64p, C.c_uint32, u32p]
        L.kc_wall_clock_khz.argtypes = [vp]
        L.kc_wall_clock_khz.restype = C.c_uint64
        L.kc_occupancy.argtypes = [vp, C.c_uint32, C.c_int, u32p, u32p]
        L.kc_copy_bandwidth.argtypes = [vp, C.c_uint64, C.c_uint32, C.POINTER(C.c_double)]
        L.kc_batch_upload.argtypes = [vp, C.POINTER(kc_batch_csr), pp]
        L.kc_batch_free.argtypes = [vp]
        L.kc_batch_free.restype = None
        L.kc_batch_peel.argtypes = [vp, vp, C.POINTER(kc_params), u32p, u32p, C.POINTER(kc_stats)]
        L.kc_batch_download_order.argtypes = [vp, vp, u32p]
        L.kc_graph_download_debug.argtypes = [vp, vp, u32p, C.c_uint32, u32p, u32p, C.c_uint32]
        L.kc_serve_open.argtypes = [vp, C.POINTER(kc_serve_cfg), pp]
        L.kc_serve_submit.argtypes = [vp, C.c_uint64, C.POINTER(kc_csr), u32p]
        L.kc_serve_poll.argtypes = [vp, C.POINTER(kc_completion), C.c_uint32, u32p]
        L.kc_serve_pending.argtypes = [vp, u32p, u32p, u64p]
        L.kc_serve_close.argtypes = [vp]
        L.kc_serve_clock.argtypes = []
        L.kc_serve_clock.restype = C.c_double
        L.kc_host_alloc.argtypes = [C.c_uint64, pp]
        L.kc_host_free.argtypes = [vp]
        L.kc_batch_info.argtypes = [vp, u32p, u64p, u64p]
        L.kc_batch_info.restype = None
        L.kc_tune_peel.argtypes = [vp, vp, C.POINTER(kc_tune_cfg), C.POINTER(kc_params),
                                   C.POINTER(C.c_double), u32p]
        L.kc_tune_batch.argtypes = [vp, vp, C.POINTER(kc_tune_cfg), C.POINTER(kc_params),
                                    C.POINTER(C.c_double), u32p]
        _hip = L
    return _hip
