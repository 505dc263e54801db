"""ctypes binding of the CPU oracle (oracle/_build/liborc.so).

Test infrastructure only: the product never imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(REPO, "oracle")
LIB = os.path.join(ORACLE_DIR, "_build", "liborc.so")

VARIANTS = {"buitdd": 0, "buitdd_np": 1, "tditdd": 2, "tditdd_np": 3}
U32_MAX = 0xFFFFFFFF
U64_MAX = 0xFFFFFFFFFFFFFFFF

_u64p = ctypes.POINTER(ctypes.c_uint64)


class _OrcHeap(ctypes.Structure):
    _fields_ = [
        ("variant", ctypes.c_int), ("batch_num", ctypes.c_int), ("B", ctypes.c_int),
        ("init", ctypes.c_uint64), ("items", _u64p),
        ("batch_count", ctypes.c_int), ("partial", ctypes.c_int),
        ("delete_count", ctypes.c_int), ("error", ctypes.c_int),
        ("x0", _u64p), ("x1", _u64p), ("x2", _u64p), ("tmp", _u64p),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.run(["make", "-C", ORACLE_DIR], check=True, capture_output=True)
        L = ctypes.CDLL(LIB)
        P = ctypes.POINTER(_OrcHeap)
        L.orc_rev_idx.restype = ctypes.c_uint32
        L.orc_rev_idx.argtypes = [ctypes.c_uint32]
        L.orc_parent.restype = ctypes.c_int
        L.orc_parent.argtypes = [ctypes.c_int, ctypes.c_int]
        L.orc_create.restype = P
        L.orc_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64]
        L.orc_destroy.argtypes = [P]
        L.orc_reset.argtypes = [P]
        L.orc_insert.restype = ctypes.c_int
        L.orc_insert.argtypes = [P, _u64p, ctypes.c_int]
        L.orc_delete_root.restype = ctypes.c_int
        L.orc_delete_root.argtypes = [P, _u64p, ctypes.POINTER(ctypes.c_int)]
        L.orc_delete_update.argtypes = [P]
        L.orc_take_last.restype = ctypes.c_int
        L.orc_take_last.argtypes = [P, _u64p]
        L.orc_root_out.restype = ctypes.c_int
        L.orc_root_out.argtypes = [P, _u64p, _u64p, ctypes.c_int]
        L.orc_sift_step.restype = ctypes.c_int
        L.orc_sift_step.argtypes = [P, ctypes.c_int, _u64p]
        L.orc_check.restype = ctypes.c_int
        L.orc_check.argtypes = [P, ctypes.c_char_p, ctypes.c_int]
        L.orc_heapsort.restype = ctypes.c_int
        L.orc_heapsort.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, _u64p,
                                   ctypes.c_int64, ctypes.c_int, _u64p]
        L.orc_splitmix64.restype = ctypes.c_uint64
        L.orc_splitmix64.argtypes = [_u64p]
        L.orc_keys_u32.argtypes = [ctypes.c_uint64, ctypes.c_int64, ctypes.POINTER(ctypes.c_uint32)]
        L.orc_keys_u64.argtypes = [ctypes.c_uint64, ctypes.c_int64, _u64p]
        vp = ctypes.c_void_p
        L.orc_dijkstra.restype = ctypes.c_int
        L.orc_dijkstra.argtypes = [ctypes.c_int32, vp, vp, vp, ctypes.c_int32, vp, ctypes.POINTER(ctypes.c_int64)]
        L.orc_astar.restype = ctypes.c_int
        L.orc_astar.argtypes = [ctypes.c_int32, ctypes.c_int32, vp, ctypes.POINTER(ctypes.c_uint32),
                                ctypes.POINTER(ctypes.c_int64)]
        L.orc_knapsack_dp.restype = ctypes.c_int
        L.orc_knapsack_dp.argtypes = [ctypes.c_int32, vp, vp, ctypes.c_uint64, _u64p]
        L.orc_knapsack_bnb.restype = ctypes.c_int
        L.orc_knapsack_bnb.argtypes = [ctypes.c_int32, vp, vp, ctypes.c_uint64, ctypes.c_int64, ctypes.c_int,
                                       _u64p, ctypes.POINTER(ctypes.c_int64)]
        _lib = L
    return _lib


def _p64(a):
    return a.ctypes.data_as(_u64p)


class Model:
    """One sequential BGPQ (oracle/bgpq_model.c)."""

    def __init__(self, variant, batch_num, batch_size, init_limits):
        self.variant = variant
        self.B = batch_size
        self.batch_num = batch_num
        self.init = init_limits
        self._m = lib().orc_create(VARIANTS[variant], batch_num, batch_size, init_limits)
        if not self._m:
            raise ValueError("bad model parameters")
        self.out = np.zeros(0, dtype=np.uint64)

    def __del__(self):
        if getattr(self, "_m", None):
            lib().orc_destroy(self._m)
            self._m = None

    @property
    def st(self):
        return self._m.contents

    def insert(self, keys):
        k = np.ascontiguousarray(keys, dtype=np.uint64)
        return lib().orc_insert(self._m, _p64(k), len(k))

    def ensure_out(self, n):
        if len(self.out) < n:
            o = np.zeros(n, dtype=np.uint64)
            o[:len(self.out)] = self.out
            self.out = o

    def delete(self):
        """deleteRoot (+deleteUpdate when it returns true); returns the size."""
        self.ensure_out(self.st.delete_count + self.B)
        size = ctypes.c_int()
        more = lib().orc_delete_root(self._m, _p64(self.out), ctypes.byref(size))
        if more:
            lib().orc_delete_update(self._m)
        return size.value

    # ---- one deletion in the steps a concurrent deleter takes, for scripted
    # interleavings (orc_take_last / orc_root_out / orc_sift_step)
    def take_last(self):
        """The last batch as a new candidate (numpy array), or None."""
        cand = np.zeros(self.B, dtype=np.uint64)
        return cand if lib().orc_take_last(self._m, _p64(cand)) else None

    def root_out(self, cand=None, check=True):
        """The root leaves at the cursor; with a candidate taken early, the
        root check merges them.  Returns True when it merged."""
        self.ensure_out(self.st.delete_count + self.B)
        c = _p64(cand) if cand is not None else None
        return lib().orc_root_out(self._m, _p64(self.out), c, 1 if check else 0) == 1

    def sift_step(self, cur, cand):
        """One level step at node cur; the next node, or 0 when done."""
        return lib().orc_sift_step(self._m, cur, _p64(cand))

    def set_items(self, slot, keys):
        items = np.ctypeslib.as_array(self.st.items, shape=((self.batch_num + 1) * self.B,))
        items[slot * self.B:(slot + 1) * self.B] = np.asarray(keys, dtype=np.uint64)

    def check(self):
        buf = ctypes.create_string_buffer(256)
        ok = lib().orc_check(self._m, buf, 256)
        return ok == 1, buf.value.decode()

    def items(self):
        n = (self.batch_num + 1) * self.B
        return np.ctypeslib.as_array(self.st.items, shape=(n,)).copy()

    def counters(self):
        s = self.st
        return s.batch_count, s.partial, s.delete_count, s.error


def heapsort(variant, batch_size, batch_num, init_limits, keys, chunk=0):
    k = np.ascontiguousarray(keys, dtype=np.uint64)
    out = np.zeros(len(k), dtype=np.uint64)
    rc = lib().orc_heapsort(VARIANTS[variant], batch_size, batch_num, init_limits, _p64(k), len(k), chunk,
                            _p64(out))
    if rc != 0:
        raise RuntimeError(f"orc_heapsort failed: {rc}")
    return out


def rev_idx(x):
    return lib().orc_rev_idx(x)


def parent(variant, i):
    return lib().orc_parent(VARIANTS[variant], i)


def splitmix64_stream(seed, n):
    st = ctypes.c_uint64(seed)
    return [lib().orc_splitmix64(ctypes.byref(st)) for _ in range(n)]


def keys_u32(seed, n):
    out = np.empty(n, dtype=np.uint32)
    lib().orc_keys_u32(seed, n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
    return out


def keys_u64(seed, n):
    out = np.empty(n, dtype=np.uint64)
    lib().orc_keys_u64(seed, n, _p64(out))
    return out


def harness_keys(seed, n):
    """The reference harness's key split (test/heap_test.cu:111-116): the first
    half below 2^31, the second half at or above it, so keys inserted during
    the concurrent phase are larger than every key present before it."""
    k = keys_u32(seed, n).astype(np.uint64)
    half = n // 2
    lo = (k[:half] >> np.uint64(1)).astype(np.uint32)
    hi = ((k[half:] >> np.uint64(1)) | np.uint64(1 << 31)).astype(np.uint32)
    hi = np.minimum(hi, np.uint32(U32_MAX - 1))
    return np.concatenate([lo, hi])


def _vp(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def dijkstra(rowptr, col, w, source):
    """orc_dijkstra: uint32 distances (U32_MAX = unreachable)."""
    n = len(rowptr) - 1
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    w = np.ascontiguousarray(w, np.uint32)
    dist = np.empty(n, np.uint32)
    pops = ctypes.c_int64()
    rc = lib().orc_dijkstra(n, _vp(rowptr), _vp(col), _vp(w), source, _vp(dist), ctypes.byref(pops))
    assert rc == 0, rc
    return dist


def astar(free, width, height):
    """orc_astar: (optimal cost or U32_MAX, expanded nodes)."""
    free = np.ascontiguousarray(free, np.uint8)
    cost = ctypes.c_uint32()
    exp = ctypes.c_int64()
    rc = lib().orc_astar(width, height, _vp(free), ctypes.byref(cost), ctypes.byref(exp))
    assert rc == 0, rc
    return cost.value, exp.value


def knapsack_dp(w, v, cap):
    w = np.ascontiguousarray(w, np.uint32)
    v = np.ascontiguousarray(v, np.uint32)
    best = np.zeros(1, np.uint64)
    assert lib().orc_knapsack_dp(len(w), _vp(w), _vp(v), cap, _p64(best)) == 0
    return int(best[0])


def knapsack_bnb(w, v, cap, dominance=True, max_nodes=1 << 26):
    """(optimum, nodes pushed, rc); rc -3 = node limit reached."""
    w = np.ascontiguousarray(w, np.uint32)
    v = np.ascontiguousarray(v, np.uint32)
    best = np.zeros(1, np.uint64)
    nodes = ctypes.c_int64()
    rc = lib().orc_knapsack_bnb(len(w), _vp(w), _vp(v), cap, max_nodes, 1 if dominance else 0, _p64(best),
                                ctypes.byref(nodes))
    return int(best[0]), nodes.value, rc
