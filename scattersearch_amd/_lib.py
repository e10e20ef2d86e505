"""ctypes binding of libscattersearch.so (the C ABI in include/scattersearch.h).

The library is built in-tree by ``__graft_entry__.build()`` (or
``make -C scattersearch_amd/csrc``).  There is no fallback: if the shared
library is missing, importing this module raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libscattersearch.so")

SS_OK = 0
SS_EINVAL = -22
SS_ENOMEM = -12
SS_EIO = -5
SS_EDEVICE = -19
SS_ESTATE = -71
SS_EBUSY = -16
SS_EINTERNAL = -131
SS_QUERY_OK = 0
SS_QUERY_REJECTED = 0xFFFFFFFF
SS_QUERY_FAULT = 0x0BAD0001
SS_NONE = 0xFFFFFFFF
SS_MAX_L = 512
SS_MAX_L_LARGE = 16384
SS_MAX_K = 512
SS_MAX_RPAD = 64

SS_MODE_BATCHED = 0
SS_MODE_PER_QUERY = 1

SS_ORDER_AUTO = 0
SS_ORDER_INDEX = 1
SS_ORDER_LOCALITY = 2
SS_ORDER_MIN_BATCH = 4096
SS_NOALLOC_MAX_QUERIES = 131072

SS_POOL_AUTO = 0
SS_POOL_REGISTERS = 1
SS_POOL_LDS = 2
SS_POOL_SHIFT = 3

SS_CLASSES_AUTO = 0
SS_CLASSES_OFF = 1
SS_CLASSES_SERIAL = 2

SS_VISITED_TWO_CHOICE = 0x100


class SearchParams(ctypes.Structure):
    _fields_ = [
        ("mode", ctypes.c_uint32),
        ("waves_per_block", ctypes.c_uint32),
        ("visited_slots", ctypes.c_uint32),
        ("blocks", ctypes.c_uint32),
        ("use_graph", ctypes.c_uint32),
        ("expand_width", ctypes.c_uint32),
        ("waves_per_query", ctypes.c_uint32),
        ("query_order", ctypes.c_uint32),
        ("team_size", ctypes.c_uint32),
        ("pool_layout", ctypes.c_uint32),
        ("pool_classes", ctypes.c_uint32),
        ("visited_bucket", ctypes.c_uint32),
    ]


class QueryStats(ctypes.Structure):
    _fields_ = [
        ("ndc", ctypes.c_uint32),
        ("hops", ctypes.c_uint32),
        ("lossy", ctypes.c_uint32),
        ("status", ctypes.c_uint32),
        ("t_start", ctypes.c_uint64),
        ("t_end", ctypes.c_uint64),
    ]


class ServeConfig(ctypes.Structure):
    _fields_ = [
        ("k_max", ctypes.c_uint32),
        ("l_max", ctypes.c_uint32),
        ("ring_slots", ctypes.c_uint32),
        ("visited_slots", ctypes.c_uint32),
        ("waves_per_block", ctypes.c_uint32),
        ("blocks", ctypes.c_uint32),
        ("streams", ctypes.c_uint32),
        ("linger_us", ctypes.c_uint32),
        ("admit_us", ctypes.c_uint32),
        ("standby_waves", ctypes.c_uint32),
    ]


class SSError(RuntimeError):
    """A C-ABI call returned a negative status."""

    def __init__(self, code, msg):
        super().__init__(f"{msg} (status {code})")
        self.code = code


if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is not built; run __graft_entry__.build() or "
        "make -C scattersearch_amd/csrc")

# PyTorch-ROCm ships its own HIP runtime (its libamdhip64 has the same soname
# as the system one).  Loading torch first makes this library bind to that
# runtime, so torch tensors and streams and this library share one HIP
# runtime in the process; loading this library first would start a second
# runtime and torch would then find no GPU.
try:
    import torch  # noqa: F401
except ImportError:
    torch = None

lib = ctypes.CDLL(LIB_PATH)

u32 = ctypes.c_uint32
u32p = ctypes.POINTER(ctypes.c_uint32)
f32p = ctypes.POINTER(ctypes.c_float)
vp = ctypes.c_void_p
cint = ctypes.c_int
cstr = ctypes.c_char_p

_SIGS = {
    "ss_last_error": (cstr, []),
    "ss_create": (cint, [u32, u32, f32p, cint, ctypes.POINTER(vp)]),
    "ss_create_from_device": (cint, [u32, u32, vp, cint, ctypes.POINTER(vp)]),
    "ss_destroy": (cint, [vp]),
    "ss_dim_padded": (u32, [u32]),
    "ss_wall_clock_khz": (cint, [vp, u32p]),
    "ss_index_info": (cint, [vp, u32p, u32p, u32p, u32p, ctypes.POINTER(cint)]),
    "ss_load_graph": (cint, [vp, cstr]),
    "ss_set_graph": (cint, [vp, u32p, u32, u32]),
    "ss_get_graph": (cint, [vp, u32p, u32p, u32p]),
    "ss_search": (cint, [vp, f32p, u32, u32p, u32p, u32, u32, u32,
                         ctypes.POINTER(SearchParams), u32p, f32p,
                         ctypes.POINTER(QueryStats)]),
    "ss_search_path": (cint, [vp, f32p, u32, u32, u32, u32, ctypes.POINTER(SearchParams), u32,
                              u32p, f32p, ctypes.POINTER(QueryStats), u32p]),
    "ss_search_device": (cint, [vp, vp, u32, vp, vp, u32, u32, u32,
                                ctypes.POINTER(SearchParams), vp, vp, vp, vp]),
    "ss_search_device_noalloc": (cint, [vp, vp, u32, vp, vp, u32, u32, u32,
                                        ctypes.POINTER(SearchParams), vp, vp, vp, vp]),
    "ss_bruteforce": (cint, [vp, f32p, u32, u32, u32p, u32p, f32p]),
    "ss_bruteforce_device": (cint, [vp, vp, u32, u32, cint, u32, vp, vp, vp]),
    "ss_knn_graph": (cint, [vp, u32, u32p, f32p]),
    "ss_nn_descent": (cint, [vp, u32, u32, u32, u32p, f32p, u32p]),
    "ss_autotune": (cint, [vp, f32p, u32, u32, u32, ctypes.c_float, ctypes.POINTER(SearchParams),
                           ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]),
    "ss_serve_start": (cint, [vp, ctypes.POINTER(ServeConfig), ctypes.POINTER(vp)]),
    "ss_submit": (cint, [vp, f32p, u32, u32, ctypes.POINTER(ctypes.c_uint64)]),
    "ss_poll": (cint, [vp, u32, ctypes.POINTER(ctypes.c_uint64), u32p, f32p,
                       ctypes.POINTER(QueryStats), u32p]),
    "ss_serve_stop": (cint, [vp]),
    "ss_serve_replay": (cint, [vp, ctypes.POINTER(ServeConfig), f32p, u32, u32p, u32p,
                               ctypes.POINTER(ctypes.c_double), u32, u32p, f32p,
                               ctypes.POINTER(ctypes.c_double), ctypes.POINTER(QueryStats), vp]),
    "ss_fbin_header": (cint, [cstr, u32p, u32p]),
    "ss_fbin_read": (cint, [cstr, f32p, u32, u32]),
    "ss_fbin_write": (cint, [cstr, f32p, u32, u32]),
    "ss_vecs_header": (cint, [cstr, u32, u32p, u32p]),
    "ss_fvecs_read": (cint, [cstr, f32p, u32, u32]),
    "ss_ivecs_read": (cint, [cstr, u32p, u32, u32]),
    "ss_bvecs_read": (cint, [cstr, f32p, u32, u32]),
    "ss_fvecs_write": (cint, [cstr, f32p, u32, u32]),
    "ss_ivecs_write": (cint, [cstr, u32p, u32, u32]),
    "ss_gt_header": (cint, [cstr, u32p, u32p]),
    "ss_gt_read": (cint, [cstr, u32p, u32, u32]),
    "ss_gt_write": (cint, [cstr, u32p, u32, u32]),
    "ss_nsg_header": (cint, [cstr, u32p, u32p, u32p]),
    "ss_nsg_read": (cint, [cstr, u32p, u32, u32]),
    "ss_nsg_write": (cint, [cstr, u32p, u32, u32, u32]),
    "ss_medoid": (cint, [f32p, u32, u32, u32p]),
    "ss_build_graph": (cint, [f32p, u32, u32, cint, u32p, u32, u32p, u32, u32,
                              ctypes.c_float, u32, cint, u32p, u32p]),
    "ss_graph_reachable": (cint, [u32p, u32, u32, u32, u32p]),
    "ss_build_graph_gpu": (cint, [vp, f32p, u32p, u32, u32p, u32, u32, ctypes.c_float, u32, u32p, u32p]),
}

for _name, (_res, _args) in _SIGS.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args


def check(rc):
    """Raise SSError for a negative status."""
    if rc != SS_OK:
        msg = lib.ss_last_error()
        raise SSError(rc, msg.decode() if msg else "error")
    return rc


def fptr(a):
    return a.ctypes.data_as(f32p) if a is not None else None


def uptr(a):
    return a.ctypes.data_as(u32p) if a is not None else None
