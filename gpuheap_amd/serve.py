"""Batched and served heap solves (include/bgpq_serve.h).

One persistent kernel runs many independent queries, each on its own
device-resident heap and solved by a team of workgroups; the host posts new
queries while others run, within an HBM budget.  ``batch_run`` solves a list
of queries at once through one captured hipGraph.
"""
import ctypes

import numpy as np

from ._lib import AppStats, check, lib

SSSP, ASTAR, KNAPSACK = 0, 1, 2
EAGAIN = 1


class Query(ctypes.Structure):
    """bgpq_query."""
    _fields_ = [
        ("type", ctypes.c_int32), ("batch_size", ctypes.c_int32), ("team", ctypes.c_int32),
        ("reserved", ctypes.c_int32), ("heap_keys", ctypes.c_int64),
        ("rowptr", ctypes.c_void_p), ("col", ctypes.c_void_p), ("weight", ctypes.c_void_p),
        ("n_vertices", ctypes.c_int32), ("source", ctypes.c_int32), ("n_edges", ctypes.c_int64),
        ("free_cells", ctypes.c_void_p), ("width", ctypes.c_int32), ("height", ctypes.c_int32),
        ("item_weight", ctypes.c_void_p), ("item_value", ctypes.c_void_p), ("n_items", ctypes.c_int32),
        ("reserved2", ctypes.c_int32), ("capacity", ctypes.c_uint64), ("pool_nodes", ctypes.c_int64),
        ("labels_out", ctypes.c_void_p), ("tag", ctypes.c_uint64),
    ]


class Result(ctypes.Structure):
    """bgpq_result."""
    _fields_ = [
        ("tag", ctypes.c_uint64), ("type", ctypes.c_int32), ("status", ctypes.c_int32),
        ("value", ctypes.c_uint64), ("digest", ctypes.c_uint64), ("stats", AppStats),
        ("t_submit", ctypes.c_double), ("t_done", ctypes.c_double),
    ]


_keep = {}  # id(query) -> host buffers that must outlive its submission


def _hold(q, *bufs):
    _keep[id(q)] = bufs
    return q


def sssp_query(graph, source, batch_size=1024, team=1, tag=0, labels_out=None):
    """A query on an apps.Graph (device CSR).  labels_out: optional uint32
    host array of n_vertices that receives the distances."""
    q = Query(type=SSSP, batch_size=batch_size, team=team, rowptr=graph.d_rowptr.ptr, col=graph.d_col.ptr,
              weight=graph.d_w.ptr, n_vertices=graph.n, source=source, n_edges=graph.m, tag=tag)
    if labels_out is not None:
        q.labels_out = labels_out.ctypes.data
    return _hold(q, labels_out)


def astar_query(gmap, batch_size=256, team=1, tag=0, labels_out=None):
    """A query on an apps.GridMap (device obstacle map)."""
    q = Query(type=ASTAR, batch_size=batch_size, team=team, free_cells=gmap.d_free.ptr, width=gmap.width,
              height=gmap.height, tag=tag)
    if labels_out is not None:
        q.labels_out = labels_out.ctypes.data
    return _hold(q, labels_out)


def knapsack_query(weights, values, capacity, batch_size=256, team=1, tag=0, pool_nodes=0):
    w = np.ascontiguousarray(weights, np.uint32)
    v = np.ascontiguousarray(values, np.uint32)
    q = Query(type=KNAPSACK, batch_size=batch_size, team=team, item_weight=w.ctypes.data, item_value=v.ctypes.data,
              n_items=len(w), capacity=capacity, pool_nodes=pool_nodes, tag=tag)
    return _hold(q, w, v)


def query_bytes(q):
    """HBM footprint of a query in the server's arena."""
    return check(lib().bgpq_query_bytes(ctypes.byref(q)))


def now():
    """The server's host clock (monotonic seconds)."""
    return lib().bgpq_now()


def sssp_digest(dist):
    """sum over v of (dist[v] + 1) * (2v + 1) mod 2^64: the digest the serve
    kernel reports for an SSSP query."""
    d = np.asarray(dist).astype(np.uint64) + np.uint64(1)
    k = 2 * np.arange(len(d), dtype=np.uint64) + np.uint64(1)
    with np.errstate(over="ignore"):
        return int(np.sum(d * k, dtype=np.uint64))


def _result(r):
    return {"tag": r.tag, "type": r.type, "status": r.status, "value": r.value, "digest": r.digest,
            "stats": r.stats.as_dict(), "t_submit": r.t_submit, "t_done": r.t_done}


class Server:
    """Persistent serve kernel over an HBM arena of `budget` bytes.  Inputs
    (graphs, grids) must be on the device before start()."""

    def __init__(self, budget, workgroups=0, queue_slots=0):
        h = ctypes.c_void_p()
        check(lib().bgpq_server_create(ctypes.byref(h), int(budget), workgroups, queue_slots))
        self._h = h
        self._pending = {}

    def set_idle_limit(self, polls):
        """Idle polls before a workgroup leaves the kernel (0: default)."""
        check(lib().bgpq_server_set_idle_limit(self._h, int(polls)))

    def live(self):
        """Workgroups of the running kernel that have not left it."""
        return check(lib().bgpq_server_live(self._h))

    def start(self):
        check(lib().bgpq_server_start(self._h))
        return self

    def submit(self, q):
        """True if posted, False if it does not fit right now (EAGAIN)."""
        rc = check(lib().bgpq_server_submit(self._h, ctypes.byref(q)))
        if rc == EAGAIN:
            return False
        self._pending[q.tag] = q
        return True

    def poll(self, max_results=256):
        out = (Result * max_results)()
        n = check(lib().bgpq_server_poll(self._h, out, max_results))
        res = [_result(out[i]) for i in range(n)]
        for r in res:
            q = self._pending.pop(r["tag"], None)
            if q is not None:
                _keep.pop(id(q), None)
        return res

    def inflight(self):
        return lib().bgpq_server_inflight(self._h)

    def used_bytes(self):
        return lib().bgpq_server_used_bytes(self._h)

    def close(self):
        if self._h:
            check(lib().bgpq_server_close(self._h))

    def destroy(self):
        if getattr(self, "_h", None):
            lib().bgpq_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self.start()

    def __exit__(self, *exc):
        try:
            self.close()
        finally:
            self.destroy()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def batch_run(queries, budget, workgroups=0, repeats=1):
    """Solve every query at once with one captured hipGraph, replayed
    `repeats` times.  Returns (results of the last replay, ms per replay)."""
    arr = (Query * len(queries))(*queries)
    out = (Result * len(queries))()
    ms = ctypes.c_float()
    check(lib().bgpq_batch_run(arr, len(queries), int(budget), workgroups, repeats, out, ctypes.byref(ms)))
    return [_result(out[i]) for i in range(len(queries))], ms.value
