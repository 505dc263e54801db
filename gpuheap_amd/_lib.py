"""ctypes binding of libbgpq.so (the C ABI in include/bgpq.h).

The library is built in-tree by ``__graft_entry__.build()`` (or
``make -C gpuheap_amd/csrc``).  Loading fails loudly when it is missing:
there is no CPU fallback for any operation.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libbgpq.so")

_lib = None

c_int = ctypes.c_int
c_i64 = ctypes.c_int64
c_u64 = ctypes.c_uint64
c_void_p = ctypes.c_void_p
c_size_t = ctypes.c_size_t
c_float_p = ctypes.POINTER(ctypes.c_float)
c_int_p = ctypes.POINTER(ctypes.c_int)
c_i64_p = ctypes.POINTER(ctypes.c_int64)

# name -> (restype, argtypes); mirrors include/bgpq.h
SIGNATURES = {
    "bgpq_create": (c_int, [ctypes.POINTER(c_void_p), c_int, c_int, c_int, c_int, c_u64]),
    "bgpq_destroy": (None, [c_void_p]),
    "bgpq_reset": (c_int, [c_void_p, c_void_p]),
    "bgpq_check": (c_int, [c_void_p]),
    "bgpq_item_count": (c_int, [c_void_p, c_i64_p]),
    "bgpq_node_count": (c_int, [c_void_p, c_i64_p]),
    "bgpq_is_empty": (c_int, [c_void_p]),
    "bgpq_insert": (c_int, [c_void_p, c_void_p, c_i64, c_int, c_int, c_void_p]),
    "bgpq_delete": (c_int, [c_void_p, c_void_p, c_i64, c_i64, c_int, c_void_p]),
    "bgpq_set_cursor": (c_int, [c_void_p, c_i64]),
    "bgpq_concurrent": (c_int, [c_void_p, c_void_p, c_i64, c_int, c_void_p, c_i64, c_i64, c_int, c_void_p]),
    "bgpq_heapsort": (c_int, [c_void_p, c_void_p, c_void_p, c_i64, c_int, c_int, c_int, c_float_p, c_float_p, c_void_p]),
    "bgpq_view": (c_int, [c_void_p, c_void_p]),
    "bgpq_vote_reset": (c_int, [c_void_p, c_int, c_void_p]),
    "bgpq_terminated": (c_int, [c_void_p]),
    "bgpq_clear": (c_int, [c_void_p, c_void_p]),
    "bgpq_snapshot": (c_int, [c_void_p, c_void_p, c_int_p, c_int_p]),
    "bgpq_error_flags": (c_int, [c_void_p, c_int_p]),
    "bgpq_threads": (c_int, [c_int]),
    "bgpq_lds_bytes": (c_int, [c_int, c_int]),
    "bgpq_default_grid": (c_int, [c_int]),
    "bgpq_test_sort": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_u64, c_int]),
    "bgpq_test_merge": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int]),
    "bgpq_gen_keys": (c_int, [c_int, c_u64, c_i64, c_void_p, c_void_p]),
    "bgpq_last_error": (ctypes.c_char_p, []),
    "bgpq_dev_malloc": (c_int, [ctypes.POINTER(c_void_p), c_size_t]),
    "bgpq_dev_free": (c_int, [c_void_p]),
    "bgpq_h2d": (c_int, [c_void_p, c_void_p, c_size_t]),
    "bgpq_d2h": (c_int, [c_void_p, c_void_p, c_size_t]),
    "bgpq_sync": (c_int, []),
    "bgpq_device_name": (c_int, [ctypes.c_char_p, c_int]),
    # include/bgpq_apps.h
    "bgpq_sssp": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_int32, ctypes.c_int32, c_void_p, c_int,
                          c_void_p, c_void_p]),
    "bgpq_astar": (c_int, [c_void_p, c_void_p, ctypes.c_int32, ctypes.c_int32, c_void_p, c_int,
                           ctypes.POINTER(ctypes.c_uint32), c_void_p, c_void_p]),
    "bgpq_knapsack": (c_int, [c_void_p, ctypes.c_int32, c_void_p, c_void_p, c_u64, c_i64, c_int,
                              ctypes.POINTER(c_u64), c_void_p, c_void_p]),
    "bgpq_knapsack_noalloc": (c_int, [c_void_p, ctypes.c_int32, c_void_p, c_void_p, c_u64, c_i64, c_int,
                                      ctypes.POINTER(c_u64), c_void_p, c_void_p]),
    "bgpq_label_digest": (c_int, [c_void_p, c_i64, c_void_p, c_void_p]),
    "bgpq_gen_rmat": (c_int, [c_int, c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, c_u64,
                              ctypes.c_uint32, ctypes.c_uint32, c_void_p, c_void_p, c_void_p]),
    "bgpq_gen_grid": (c_int, [c_int, c_int, ctypes.c_double, c_u64, c_void_p]),
    "bgpq_gen_knapsack": (c_int, [c_int, c_u64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, c_void_p,
                                  c_void_p, ctypes.POINTER(c_u64)]),
    # include/bgpq_serve.h
    "bgpq_now": (ctypes.c_double, []),
    "bgpq_query_bytes": (c_i64, [c_void_p]),
    "bgpq_server_create": (c_int, [ctypes.POINTER(c_void_p), c_i64, c_int, c_int]),
    "bgpq_server_destroy": (None, [c_void_p]),
    "bgpq_server_start": (c_int, [c_void_p]),
    "bgpq_server_set_idle_limit": (c_int, [c_void_p, c_i64]),
    "bgpq_server_live": (c_int, [c_void_p]),
    "bgpq_server_submit": (c_int, [c_void_p, c_void_p]),
    "bgpq_server_poll": (c_int, [c_void_p, c_void_p, c_int]),
    "bgpq_server_inflight": (c_int, [c_void_p]),
    "bgpq_server_used_bytes": (c_i64, [c_void_p]),
    "bgpq_server_close": (c_int, [c_void_p]),
    "bgpq_batch_run": (c_int, [c_void_p, c_int, c_i64, c_int, c_int, c_void_p, c_float_p]),
}


class View(ctypes.Structure):
    """bgpq_view_t (include/bgpq.h)."""
    _fields_ = [("items", c_void_p), ("status", c_void_p), ("ctr", c_void_p), ("init_limits", c_u64),
                ("batch_num", ctypes.c_int32), ("batch_size", ctypes.c_int32), ("variant", ctypes.c_int32),
                ("key_type", ctypes.c_int32), ("word_stride", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class AppStats(ctypes.Structure):
    """bgpq_app_stats (include/bgpq_apps.h)."""
    _fields_ = [("popped", ctypes.c_int64), ("expanded", ctypes.c_int64), ("stale", ctypes.c_int64),
                ("pruned", ctypes.c_int64), ("pushed", ctypes.c_int64), ("error_flags", ctypes.c_int32),
                ("ms", ctypes.c_float)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class BgpqError(RuntimeError):
    """A C-ABI call returned a negative status."""

    def __init__(self, code, msg):
        super().__init__(f"bgpq error {code}: {msg}")
        self.code = code


def lib():
    """Load libbgpq.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "or `make -C gpuheap_amd/csrc` (no CPU fallback exists)")
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def last_error():
    msg = lib().bgpq_last_error()
    return msg.decode() if msg else ""


def check(rc):
    """Raise BgpqError for a negative status, return rc otherwise."""
    if rc < 0:
        raise BgpqError(rc, last_error())
    return rc
