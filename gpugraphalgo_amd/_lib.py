"""ctypes binding of libgga.so (the C-ABI declared in include/gga.h).

The shared library is built in-tree (gpugraphalgo_amd/lib/libgga.so) by
gpugraphalgo_amd.build.build().  Loading fails loudly when it is missing:
there is no Python or CPU fallback for any traversal.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libgga.so")
CLI_PATH = os.path.join(_HERE, "bin", "gga_run")

u32 = ctypes.c_uint32
u64 = ctypes.c_uint64
i32 = ctypes.c_int32
P = ctypes.POINTER
vp = ctypes.c_void_p

GGA_OK = 0
GGA_EBUDGET = -9
ERRORS = {
    -1: "GGA_EINVAL", -2: "GGA_EIO", -3: "GGA_EFORMAT", -4: "GGA_ERANGE", -5: "GGA_ENOMEM",
    -6: "GGA_EHIP", -7: "GGA_ENOSPLIT", -8: "GGA_ETIMEDOUT", -9: "GGA_EBUDGET", -10: "GGA_ENODEV",
}
INF = 0xFFFFFFFF
REF_INF = 1000000000

MODE_TOPO, MODE_WORKLIST, MODE_PERSISTENT, MODE_AUTO = 0, 1, 2, 3
MODES = {"topo": MODE_TOPO, "worklist": MODE_WORKLIST, "persistent": MODE_PERSISTENT,
         "auto": MODE_AUTO}


class Csr(ctypes.Structure):
    _fields_ = [("n", u32), ("m", u32), ("row_ptr", P(u32)), ("dst", P(u32)), ("w", P(u32)),
                ("has_weights", u32)]


class SplitStats(ctypes.Structure):
    _fields_ = [("nnode", u32), ("nedge", u32), ("ns_nnode", u32), ("ns_nedge", u32),
                ("extranode", u32), ("maxdeg", u32), ("level", i32), ("bucket", u32), ("mda", u32),
                ("max_chunk", u32), ("split_ms", ctypes.c_float), ("exec_chunks", u32)]


class WgRecord(ctypes.Structure):
    _fields_ = [("start", u64), ("ticks", u32), ("place", u32)]


class ConvRow(ctypes.Structure):
    _fields_ = [("step", u32), ("diff_count", u32), ("norm2", ctypes.c_double)]


class Opts(ctypes.Structure):
    _fields_ = [("mode", i32), ("inf", u32), ("max_iters", u32), ("block", u32),
                ("sync_every", u32), ("use_graph", u32), ("count", u32), ("grid", u32),
                ("dense_min", u32), ("pull_alpha", u32), ("ms_layout", u32), ("delta", u32),
                ("timeline", ctypes.c_void_p), ("timeline_cap", u32), ("conv_cap", u32),
                ("conv_final", ctypes.c_void_p), ("conv_rows", ctypes.c_void_p)]


class Stats(ctypes.Structure):
    _fields_ = [("iterations", u32), ("reached", u32), ("time_ms", ctypes.c_float),
                ("ktime_ms", ctypes.c_float), ("sktime_ms", ctypes.c_float),
                ("ctime_ms", ctypes.c_float), ("mode", i32),
                ("edges_inspected", u64), ("chunks_expanded", u64), ("relaxations", u64),
                ("m_reach", u64), ("launches", u64), ("pull_levels", u32), ("timeline_records", u32),
                ("timeline_grid", u32), ("conv_steps", u32)]


class ServerOpts(ctypes.Structure):
    _fields_ = [("hbm_budget", u64), ("steps_per_round", u32), ("lanes", u32), ("use_graph", u32),
                ("dense_min", u32), ("host_results", u32), ("dry_run", u32), ("dry_steps", u32),
                ("layout", u32), ("engines_per_graph", u32), ("mesh_reserve", ctypes.c_int32)]


class Query(ctypes.Structure):
    _fields_ = [("graph", u32), ("src", u32), ("bfs", u32), ("out_on_device", u32),
                ("arrival", ctypes.c_double), ("out", ctypes.c_void_p)]


class QueryResult(ctypes.Structure):
    _fields_ = [("id", u64), ("graph", u32), ("src", u32), ("bfs", u32), ("lane", u32),
                ("status", i32), ("engine", u32), ("arrival", ctypes.c_double),
                ("admitted", ctypes.c_double), ("done", ctypes.c_double), ("checksum", u64)]


class ServerCounters(ctypes.Structure):
    _fields_ = [("submitted", u64), ("completed", u64), ("rejected", u64), ("rounds", u64),
                ("evictions", u64), ("engines_created", u64), ("bytes_in_use", u64),
                ("peak_bytes", u64), ("engines_resident", u32), ("peak_running", u32),
                ("joined", u64)]


class TuneResult(ctypes.Structure):
    _fields_ = [("mda", u32), ("grid", u32), ("dense_min", u32), ("layout", u32), ("mode", i32), ("delta", u32),
                ("seed_mda", u32),
                ("ntried", u32), ("seed_ms", ctypes.c_float), ("best_ms", ctypes.c_float),
                ("layout_ms", ctypes.c_float * 2), ("tried_mda", u32 * 16), ("tried_ms", ctypes.c_float * 16),
                ("block", u32), ("block_ms", ctypes.c_float * 3)]


class VerifyResult(ctypes.Structure):
    _fields_ = [("nerr", u64), ("untight", u64), ("src_ok", u32), ("reached", u32)]


class ResultRow(ctypes.Structure):
    _fields_ = [("exename", ctypes.c_char * 100), ("algo", ctypes.c_char * 10),
                ("graphfile", ctypes.c_char * 200), ("bucket", u32), ("nnode", u32),
                ("nedge", u32), ("ns_nnode", u32), ("ns_nedge", u32), ("extranode", u32),
                ("maxdeg", u32), ("level", i32), ("mda", u32), ("iterations", u32),
                ("runtime_ms", u32), ("time_ms", ctypes.c_float), ("ktime_ms", ctypes.c_float),
                ("sktime_ms", ctypes.c_float)]


# (name, restype, argtypes) for every symbol include/gga.h declares
SIGNATURES = [
    ("gga_strerror", ctypes.c_char_p, [ctypes.c_int]),
    ("gga_last_error", ctypes.c_char_p, []),
    ("gga_version", ctypes.c_char_p, []),
    ("gga_device_count", ctypes.c_int, []),
    ("gga_csr_read_gr", ctypes.c_int, [ctypes.c_char_p, P(Csr)]),
    ("gga_csr_write_gr", ctypes.c_int, [ctypes.c_char_p, P(Csr)]),
    ("gga_csr_from_row_end", ctypes.c_int, [u32, u32, vp, vp, vp, P(Csr)]),
    ("gga_csr_free", None, [P(Csr)]),
    ("gga_csr_maxdeg", u32, [P(Csr), P(u32)]),
    ("gga_gen_rmat", ctypes.c_int, [u32, u32, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                    u64, u32, u32, ctypes.c_int, P(Csr)]),
    ("gga_gen_grid", ctypes.c_int, [u32, u32, ctypes.c_double, u64, u32, u32, P(Csr)]),
    ("gga_graph_from_csr", ctypes.c_int, [u32, u32, vp, vp, vp, ctypes.c_int, P(vp)]),
    ("gga_graph_from_csr32", ctypes.c_int, [u32, u32, vp, vp, vp, ctypes.c_int, P(vp)]),
    ("gga_graph_load_gr", ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, P(vp)]),
    ("gga_graph_free", ctypes.c_int, [vp]),
    ("gga_graph_n", u32, [vp]),
    ("gga_graph_m", u32, [vp]),
    ("gga_graph_device_bytes", u64, [vp]),
    ("gga_graph_split", ctypes.c_int, [vp, u32, i32, u32, P(SplitStats)]),
    ("gga_graph_chunks", ctypes.c_int, [vp, vp, vp]),
    ("gga_opts_init", None, [P(Opts)]),
    ("gga_bfs", ctypes.c_int, [vp, u32, vp, P(Opts), P(Stats)]),
    ("gga_sssp", ctypes.c_int, [vp, u32, vp, P(Opts), P(Stats)]),
    ("gga_graph_copy_dist", ctypes.c_int, [vp, vp]),
    ("gga_sssp_noalloc", ctypes.c_int, [vp, u32, vp, P(Opts), P(Stats)]),
    ("gga_write_degree_dist", ctypes.c_int, [ctypes.c_char_p, u32, vp]),
    ("gga_graph_dist_device", vp, [vp]),
    ("gga_verify", ctypes.c_int, [u32, vp, vp, vp, u32, vp, u32, P(VerifyResult)]),
    ("gga_graph_verify", ctypes.c_int, [vp, u32, ctypes.c_int, u32, P(VerifyResult)]),
    ("gga_graph_checksum", ctypes.c_int, [vp, P(u64)]),
    ("gga_timeline_makespan", ctypes.c_double, [vp, u32, u32, u32]),
    ("gga_write_conv", ctypes.c_int, [ctypes.c_char_p, i32, u32, ctypes.c_char_p, ctypes.c_char_p, vp, u32, u32]),
    ("gga_write_dist", ctypes.c_int, [ctypes.c_char_p, vp, u32]),
    ("gga_write_result_row", ctypes.c_int, [ctypes.c_char_p, P(ResultRow)]),
    ("gga_graph_name", None, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_size_t]),
    ("gga_ms_create", ctypes.c_int, [vp, u32, u32, P(vp)]),
    ("gga_ms_free", ctypes.c_int, [vp]),
    ("gga_ms_set_delta", ctypes.c_int, [vp, u32]),
    ("gga_ms_graph_bytes", u64, [vp]),
    ("gga_ms_start", ctypes.c_int, [vp, u32, vp, u32]),
    ("gga_ms_run", ctypes.c_int, [vp, u32, P(Opts), P(u32)]),
    ("gga_ms_done", u32, [vp]),
    ("gga_ms_launch", ctypes.c_int, [vp, u32, P(Opts)]),
    ("gga_ms_collect", ctypes.c_int, [vp, P(u32)]),
    ("gga_ms_ready", ctypes.c_int, [vp]),
    ("gga_ms_checksum", ctypes.c_int, [vp, u32, vp]),
    ("gga_ms_bytes", u64, [u32, u32]),
    ("gga_checksum", u64, [vp, u32]),
    ("gga_ms_fetch", ctypes.c_int, [vp, u32, vp, ctypes.c_int]),
    ("gga_ms_stats", ctypes.c_int, [vp, P(Stats)]),
    ("gga_multi_sp", ctypes.c_int, [vp, vp, u32, u32, vp, P(Opts), P(Stats)]),
    ("gga_multi_bfs", ctypes.c_int, [vp, vp, u32, vp, P(Opts), P(Stats)]),
    ("gga_tune", ctypes.c_int, [vp, u32, u32, P(TuneResult)]),
    ("gga_tune_layout", u32, [vp]),
    ("gga_server_create", ctypes.c_int, [P(ServerOpts), P(vp)]),
    ("gga_server_free", ctypes.c_int, [vp]),
    ("gga_server_add_graph", ctypes.c_int, [vp, vp, P(u32)]),
    ("gga_server_add_graph_dry", ctypes.c_int, [vp, u32, u32, P(u32)]),
    ("gga_server_now", ctypes.c_double, [vp]),
    ("gga_server_submit", ctypes.c_int, [vp, P(Query), P(u64)]),
    ("gga_server_pump", ctypes.c_int, [vp, P(u32)]),
    ("gga_server_poll", ctypes.c_int, [vp, P(QueryResult), u32, P(u32)]),
    ("gga_server_drain", ctypes.c_int, [vp]),
    ("gga_server_replay", ctypes.c_int, [vp, P(Query), u32, P(QueryResult)]),
    ("gga_server_replay_plain", ctypes.c_int, [vp, P(Query), u32, i32, P(QueryResult)]),
    ("gga_server_stats", ctypes.c_int, [vp, P(ServerCounters)]),
]

_LIB = None


class GgaError(RuntimeError):
    def __init__(self, rc: int, where: str):
        self.rc = rc
        detail = ""
        if _LIB is not None:
            detail = (_LIB.gga_last_error() or b"").decode(errors="replace")
        super().__init__(f"{where}: {ERRORS.get(rc, rc)} ({detail})")


def load() -> ctypes.CDLL:
    """Loads libgga.so; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing: run gpugraphalgo_amd.build.build() "
                              "(there is no fallback path)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = lib
    return _LIB


def check(rc: int, where: str) -> None:
    if rc != GGA_OK:
        raise GgaError(rc, where)
