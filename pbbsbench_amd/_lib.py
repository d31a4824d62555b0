"""ctypes declarations of the C ABI in include/pbbs_bfs.h.

The shared library is built in-tree (``pbbsbench_amd/libpbbs_bfs.so``).  There
is no fallback: if it is missing or fails to load, every entry point raises.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# PBBS_BFS_LIB names another build of the library (e.g. one compiled with
# different launch-shape constants, for timing it against this one).
LIB_PATH = os.environ.get("PBBS_BFS_LIB") or os.path.join(_HERE, "libpbbs_bfs.so")

PBBS_OK = 0
PBBS_ERR_INVALID = -1
PBBS_ERR_HIP = -2
PBBS_ERR_NOMEM = -3
PBBS_ERR_LIMIT = -4
PBBS_ERR_FORMAT = -5
PBBS_ERR_IO = -6
PBBS_ERR_STATE = -7

PBBS_GRAPH_OFFSETS_I64 = 0x1
PBBS_GRAPH_DEVICE_INPUT = 0x2
PBBS_GRAPH_ASSUME_SYMMETRIC = 0x8
PBBS_OUT_DEVICE = 0x1
PBBS_BFS_ANY_PARENT = 0x4
PBBS_PREPARE_SINGLE = 0x1
PBBS_PREPARE_BATCH = 0x2


class Tuning(C.Structure):
    _fields_ = [
        ("td_to_bu_div", C.c_int32),
        ("bu_to_td_div", C.c_int32),
        ("grid_blocks", C.c_int32),
        ("bu_serial_depth", C.c_int32),
        ("chunk_levels", C.c_int32),
        ("use_graphs", C.c_int32),
        ("batch_pull_div", C.c_int32),
        ("bu_row_tops", C.c_int32),
        ("queue_div", C.c_int32),
        ("persist", C.c_int32),
        ("persist_wait_us", C.c_int32),
    ]


class GraphInfo(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("m", C.c_int64),
        ("offsets_i64", C.c_int32),
        ("sorted", C.c_int32),
        ("symmetric", C.c_int32),
        ("pad", C.c_int32),
        ("isolated", C.c_int64),
        ("max_degree", C.c_int64),
        ("device_bytes", C.c_int64),
    ]


class BfsStats(C.Structure):
    _fields_ = [
        ("levels", C.c_int32),
        ("td_levels", C.c_int32),
        ("bu_levels", C.c_int32),
        ("launches", C.c_int32),
        ("reached", C.c_int64),
        ("edges_reached", C.c_int64),
        ("device_ms", C.c_double),
        ("copy_ms", C.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class MisStats(C.Structure):
    _fields_ = [
        ("rounds", C.c_int32),
        ("pad", C.c_int32),
        ("members", C.c_int64),
        ("device_ms", C.c_double),
        ("copy_ms", C.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "pad"}


class StStats(C.Structure):
    _fields_ = [
        ("rounds", C.c_int32),
        ("pad", C.c_int32),
        ("edges", C.c_int64),
        ("device_ms", C.c_double),
        ("copy_ms", C.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "pad"}


class MisCheckResult(C.Structure):
    _fields_ = [
        ("valid", C.c_int32),
        ("pad", C.c_int32),
        ("first_bad", C.c_int64),
        ("bad_edges", C.c_int64),
        ("bad_vertices", C.c_int64),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "pad"}


class LevelProfile(C.Structure):
    _fields_ = [
        ("level", C.c_int32),
        ("mode", C.c_int32),
        ("frontier", C.c_int64),
        ("frontier_edges", C.c_int64),
        ("ms", C.c_double),
    ]


class Csr(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("m", C.c_int64),
        ("offsets", C.POINTER(C.c_int64)),
        ("adj", C.POINTER(C.c_int32)),
    ]


class ServerConfig(C.Structure):
    _fields_ = [
        ("slots", C.c_int32),
        ("max_batch", C.c_int32),
        ("hbm_budget", C.c_int64),
        ("reserved", C.c_int32 * 4),
    ]


class ServerResult(C.Structure):
    _fields_ = [
        ("ticket", C.c_uint64),
        ("graph", C.c_int32),
        ("source", C.c_int32),
        ("status", C.c_int32),
        ("batch", C.c_int32),
        ("latency_ms", C.c_double),
        ("queue_ms", C.c_double),
        ("device_ms", C.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


PBBS_CHECK_DEVICE_INPUT = 0x1
PBBS_EDGES_DEVICE = 0x2
PBBS_ST_ANY_FOREST = 0x8


class CheckResult(C.Structure):
    _fields_ = [
        ("valid", C.c_int32),
        ("cycle", C.c_int32),
        ("first_bad", C.c_int64),
        ("bad", C.c_int64),
        ("max_rule_bad", C.c_int64),
        ("depth_bad", C.c_int64),
        ("reached", C.c_int64),
        ("max_level", C.c_int32),
        ("pad", C.c_int32),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "pad"}


# (name, restype, argtypes) for every symbol include/pbbs_bfs.h declares.
_P = C.c_void_p
_I32P = C.POINTER(C.c_int32)
SIGNATURES = [
    ("pbbs_last_error", C.c_char_p, []),
    ("pbbs_version", C.c_char_p, []),
    ("pbbs_ctx_create", C.c_int, [C.c_int, C.POINTER(_P)]),
    ("pbbs_ctx_destroy", C.c_int, [_P]),
    ("pbbs_tuning_get", C.c_int, [_P, C.POINTER(Tuning)]),
    ("pbbs_tuning_set", C.c_int, [_P, C.POINTER(Tuning)]),
    ("pbbs_graph_upload", C.c_int,
     [_P, C.c_int64, C.c_int64, _P, _P, C.c_uint32, C.POINTER(_P)]),
    ("pbbs_graph_info_get", C.c_int, [_P, C.POINTER(GraphInfo)]),
    ("pbbs_graph_free", C.c_int, [_P]),
    ("pbbs_graph_prepare", C.c_int, [_P, _P, C.c_uint32, C.POINTER(C.c_double)]),
    ("pbbs_bfs_multi", C.c_int, [_P, C.c_int32, _P, _P, _P, _P, C.c_uint32,
                                 C.POINTER(BfsStats)]),
    ("pbbs_bfs", C.c_int, [_P, _P, C.c_int32, _P, _P, C.c_uint32, C.POINTER(BfsStats)]),
    ("pbbs_bfs_noalloc", C.c_int, [_P, _P, C.c_int32, _P, _P, C.POINTER(BfsStats)]),
    ("pbbs_mis", C.c_int, [_P, _P, _P, C.c_uint32, C.POINTER(MisStats)]),
    ("pbbs_spanning_forest", C.c_int,
     [_P, C.c_int64, C.c_int64, _P, _P, C.c_uint32, _P, C.POINTER(C.c_int64),
      C.POINTER(StStats)]),
    ("pbbs_bfs_batch", C.c_int,
     [_P, _P, _P, C.c_int32, _P, _P, C.c_uint32, C.POINTER(BfsStats)]),
    ("pbbs_bfs_profile", C.c_int,
     [_P, _P, C.c_int32, C.POINTER(LevelProfile), C.c_int32, C.POINTER(C.c_int32),
      C.POINTER(C.c_double)]),
    ("pbbs_graph_gen_rmat", C.c_int,
     [_P, C.c_int32, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_uint64, C.c_uint64,
      C.c_uint32, C.POINTER(_P)]),
    ("pbbs_graph_gen_grid3d", C.c_int, [_P, C.c_int32, C.c_uint32, C.POINTER(_P)]),
    ("pbbs_graph_download", C.c_int, [_P, C.POINTER(Csr)]),
    ("pbbs_graph_edges", C.c_int,
     [_P, _P, C.c_uint64, _P, _P, C.c_int64, C.POINTER(C.c_int64)]),
    ("pbbs_ctx_outputs", C.c_int, [_P, C.c_int, C.POINTER(_P), C.POINTER(_P)]),
    ("pbbs_check_bfs", C.c_int,
     [_P, _P, C.c_int32, _P, _P, C.c_uint32, C.POINTER(CheckResult)]),
    ("pbbs_check_mis", C.c_int, [_P, _P, _P, C.c_uint32, C.POINTER(MisCheckResult)]),
    ("pbbs_server_create", C.c_int, [C.c_int, C.POINTER(ServerConfig), C.POINTER(_P)]),
    ("pbbs_server_destroy", C.c_int, [_P]),
    ("pbbs_server_add_graph", C.c_int,
     [_P, C.c_int64, C.c_int64, _P, _P, C.c_uint32, C.POINTER(C.c_int32)]),
    ("pbbs_server_submit", C.c_int,
     [_P, C.c_int32, C.c_int32, _P, _P, C.POINTER(C.c_uint64)]),
    ("pbbs_server_poll", C.c_int,
     [_P, C.POINTER(ServerResult), C.c_int32, C.POINTER(C.c_int32), C.c_int32]),
    ("pbbs_server_load", C.c_int,
     [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    ("pbbs_csr_free", None, [C.POINTER(Csr)]),
    ("pbbs_gen_rmat", C.c_int,
     [C.c_int32, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_uint64, C.c_uint64,
      C.POINTER(Csr)]),
    ("pbbs_gen_grid3d", C.c_int, [C.c_int32, C.POINTER(Csr)]),
    ("pbbs_gen_randlocal", C.c_int, [C.c_int64, C.c_int32, C.c_uint64, C.POINTER(Csr)]),
    ("pbbs_csr_from_edges", C.c_int,
     [C.c_int64, C.c_int64, _P, _P, C.c_int, C.POINTER(Csr)]),
    ("pbbs_csr_permute", C.c_int, [C.POINTER(Csr), C.c_uint64, C.POINTER(Csr)]),
    ("pbbs_read_adjacency_graph", C.c_int, [C.c_char_p, C.POINTER(Csr)]),
    ("pbbs_write_adjacency_graph", C.c_int, [C.c_char_p, C.POINTER(Csr)]),
    ("pbbs_write_graph_binary", C.c_int, [C.c_char_p, C.POINTER(Csr)]),
    ("pbbs_read_graph_binary", C.c_int, [C.c_char_p, C.POINTER(Csr)]),
    ("pbbs_graph_read", C.c_int, [_P, C.c_char_p, C.c_uint32, C.POINTER(_P)]),
    ("pbbs_write_sequence_int", C.c_int, [C.c_char_p, _P, C.c_int64]),
    ("pbbs_read_sequence_int", C.c_int, [C.c_char_p, C.POINTER(_I32P), C.POINTER(C.c_int64)]),
    ("pbbs_read_edge_array", C.c_int,
     [C.c_char_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(_I32P),
      C.POINTER(_I32P)]),
    ("pbbs_write_edge_array", C.c_int, [C.c_char_p, C.c_int64, _P, _P]),
    ("pbbs_free", None, [_P]),
]

_lib = None


def lib():
    """Loads libpbbs_bfs.so once; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build the HIP extension first "
                "(python -c 'import __graft_entry__ as g; g.build()')")
        h = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


class PbbsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def check(rc, what):
    if rc != PBBS_OK:
        msg = lib().pbbs_last_error().decode(errors="replace")
        raise PbbsError(rc, f"{what}: {msg}")
