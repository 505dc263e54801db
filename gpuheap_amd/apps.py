"""Heap consumers on the GPU (include/bgpq_apps.h): SSSP, grid A* and 0/1
knapsack branch-and-bound, each one persistent kernel over a u64 Heap, and
the generators of their synthetic inputs.

Counterparts: applications/sssp_longfei_new/main.cu (SSSP),
artifacts/astar/astar_main.cu (A*), artifacts/knapsack/main.cu (knapsack).
"""
import ctypes

import numpy as np

from ._lib import AppStats, check, lib
from .heap import DeviceBuffer, Heap

RMAT_ABC = (0.57, 0.19, 0.19)


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def rmat_csr(scale, degree=16, seed=0, abc=RMAT_ABC, wmin=1, wmax=1000):
    """R-MAT graph as CSR (rowptr int32[V+1], col int32[E], w uint32[E])."""
    nv, ne = 1 << scale, degree << scale
    rowptr = np.empty(nv + 1, np.int32)
    col = np.empty(ne, np.int32)
    w = np.empty(ne, np.uint32)
    check(lib().bgpq_gen_rmat(scale, degree, abc[0], abc[1], abc[2], seed, wmin, wmax, _p(rowptr), _p(col), _p(w)))
    return rowptr, col, w


def grid_map(width, height, p=0.25, seed=0):
    """Obstacle map: uint8[height*width], 1 = free; corners forced free."""
    free = np.empty(width * height, np.uint8)
    check(lib().bgpq_gen_grid(width, height, p, seed, _p(free)))
    return free


def knapsack_instance(n, seed=0, wmin=1, wmax=1000, value_add=100):
    """(weights, values, capacity): values = weight + value_add, capacity =
    half the total weight."""
    w = np.empty(n, np.uint32)
    v = np.empty(n, np.uint32)
    cap = ctypes.c_uint64()
    check(lib().bgpq_gen_knapsack(n, seed, wmin, wmax, value_add, _p(w), _p(v), ctypes.byref(cap)))
    return w, v, cap.value


def _stats(st):
    return st.as_dict()


def _heap_handle(heap):
    if not isinstance(heap, Heap):
        raise TypeError("heap must be a gpuheap_amd.Heap")
    return heap._h


class Graph:
    """A CSR graph resident on the device."""

    def __init__(self, rowptr, col, w):
        self.n = len(rowptr) - 1
        self.m = len(col)
        self.d_rowptr = DeviceBuffer.from_numpy(np.ascontiguousarray(rowptr, np.int32))
        self.d_col = DeviceBuffer.from_numpy(np.ascontiguousarray(col, np.int32))
        self.d_w = DeviceBuffer.from_numpy(np.ascontiguousarray(w, np.uint32))
        self.d_dist = DeviceBuffer(4 * self.n)


def sssp(heap, graph, source=0, grid=0, fetch=True):
    """Distances from `source` (uint32, 0xFFFFFFFF = unreachable) and the
    solve's statistics; with fetch=False the distances stay on the device
    (graph.d_dist) and None is returned in their place."""
    st = AppStats()
    check(lib().bgpq_sssp(_heap_handle(heap), graph.d_rowptr.ptr, graph.d_col.ptr, graph.d_w.ptr, graph.n,
                          source, graph.d_dist.ptr, grid, ctypes.byref(st), None))
    return (graph.d_dist.to_numpy(np.uint32, graph.n) if fetch else None), _stats(st)


def label_digest(graph):
    """(digest, reached) of the distances in graph.d_dist, computed on the
    device (bgpq_label_digest); 16 bytes come back."""
    if getattr(graph, "_digest_buf", None) is None:
        graph._digest_buf = DeviceBuffer(16)
    check(lib().bgpq_label_digest(graph.d_dist.ptr, graph.n, graph._digest_buf.ptr, None))
    d = graph._digest_buf.to_numpy(np.uint64, 2)
    return int(d[0]), int(d[1])


class GridMap:
    """An obstacle grid resident on the device (with its g-label scratch)."""

    def __init__(self, free, width, height):
        self.width, self.height = width, height
        self.d_free = DeviceBuffer.from_numpy(np.ascontiguousarray(free, np.uint8))
        self.d_g = DeviceBuffer(4 * width * height)


def astar(heap, gmap, grid=0):
    """Optimal cost from cell 0 to the last cell (0xFFFFFFFF = unreachable)
    and the solve's statistics."""
    st = AppStats()
    cost = ctypes.c_uint32()
    check(lib().bgpq_astar(_heap_handle(heap), gmap.d_free.ptr, gmap.width, gmap.height, gmap.d_g.ptr, grid,
                           ctypes.byref(cost), ctypes.byref(st), None))
    return cost.value, _stats(st)


def knapsack(heap, weights, values, capacity, grid=0, pool_nodes=0):
    """Optimal 0/1 knapsack value and the solve's statistics."""
    w = np.ascontiguousarray(weights, np.uint32)
    v = np.ascontiguousarray(values, np.uint32)
    st = AppStats()
    best = ctypes.c_uint64()
    check(lib().bgpq_knapsack(_heap_handle(heap), len(w), _p(w), _p(v), capacity, pool_nodes, grid,
                              ctypes.byref(best), ctypes.byref(st), None))
    return best.value, _stats(st)


def app_heap(batch_size=1024, batch_num=1 << 14, variant="buitdd"):
    """A u64 heap for the drivers (variant with a partial buffer)."""
    return Heap(batch_num, batch_size, variant=variant, key="u64")
