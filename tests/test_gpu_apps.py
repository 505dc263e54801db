"""GPU parity of the heap consumers against their CPU oracles, through the
C ABI (include/bgpq_apps.h): SSSP distances equal Dijkstra's bit for bit,
A* costs equal the sequential A*'s, knapsack optima equal the DP's; each
for both partial-buffer variants and several batch sizes.  After a solve
the heap is empty and serves the next one.
"""
import numpy as np
import pytest

import _oracle as O
from gpuheap_amd import apps as A
from gpuheap_amd import Heap

pytestmark = pytest.mark.gpu


def _heap(B, keys_hint, variant):
    bn = 1
    while bn * B < 2 * keys_hint + 16 * B:
        bn <<= 1
    return Heap(bn, B, variant=variant, key="u64")


@pytest.mark.parametrize("variant", ["buitdd", "tditdd"])
@pytest.mark.parametrize("B", [256, 1024, 2048])
@pytest.mark.parametrize("scale,seed", [(10, 0), (14, 1)])
def test_sssp_matches_dijkstra(gpu, variant, B, scale, seed):
    print(f"[case] sssp {variant} B={B} scale={scale} seed={seed}", flush=True)
    rowptr, col, w = A.rmat_csr(scale, 16, seed)
    g = A.Graph(rowptr, col, w)
    with _heap(B, len(col), variant) as h:
        dist, st = A.sssp(h, g, source=0, grid=64)
        assert st["error_flags"] == 0 and h.isEmpty()
        want = O.dijkstra(rowptr, col, w, 0)
        assert np.array_equal(dist, want), f"{np.count_nonzero(dist != want)} distances differ"
        assert st["popped"] == st["expanded"] + st["stale"]
        # the same heap serves a second solve from another source
        dist2, _ = A.sssp(h, g, source=7, grid=32)
        assert np.array_equal(dist2, O.dijkstra(rowptr, col, w, 7))


def test_sssp_disconnected(gpu):
    # 0 -> 1 -> 2, vertex 3 isolated, 4 -> 0
    rowptr = np.array([0, 1, 2, 2, 2, 3], np.int32)
    col = np.array([1, 2, 0], np.int32)
    w = np.array([7, 5, 1], np.uint32)
    with _heap(256, 16, "buitdd") as h:
        dist, _ = A.sssp(h, A.Graph(rowptr, col, w), source=0, grid=4)
    assert list(dist) == [0, 7, 12, O.U32_MAX, O.U32_MAX]


@pytest.mark.parametrize("variant", ["buitdd", "tditdd"])
@pytest.mark.parametrize("B", [256, 1024, 2048])
@pytest.mark.parametrize("size,p,seed", [(96, 0.25, 0), (256, 0.25, 1), (200, 0.35, 2)])
def test_astar_matches_cpu(gpu, variant, B, size, p, seed):
    print(f"[case] astar {variant} B={B} {size}x{size} p={p} seed={seed}", flush=True)
    free = A.grid_map(size, size, p, seed)
    gm = A.GridMap(free, size, size)
    with _heap(B, 8 * size * size, variant) as h:
        cost, st = A.astar(h, gm, grid=64)
        assert st["error_flags"] == 0 and h.isEmpty()
    want, _ = O.astar(free, size, size)
    assert cost == want


def test_astar_unreachable_goal(gpu):
    size = 64
    free = A.grid_map(size, size, 0.0, 0)
    free[(size - 2) * size + size - 2:(size - 2) * size + size] = 0  # wall off the goal corner
    free[(size - 1) * size + size - 2] = 0
    with _heap(256, 8 * size * size, "buitdd") as h:
        cost, _ = A.astar(h, A.GridMap(free, size, size), grid=16)
    assert cost == O.U32_MAX == O.astar(free, size, size)[0]


@pytest.mark.parametrize("variant", ["buitdd", "tditdd"])
@pytest.mark.parametrize("B", [256, 1024, 2048])
@pytest.mark.parametrize("n,seed", [(40, 0), (120, 1), (200, 2)])
def test_knapsack_matches_dp(gpu, variant, B, n, seed):
    print(f"[case] knapsack {variant} B={B} n={n} seed={seed}", flush=True)
    w, v, cap = A.knapsack_instance(n, seed)
    with _heap(B, 1 << 20, variant) as h:
        best, st = A.knapsack(h, w, v, cap, grid=64)
        assert st["error_flags"] == 0 and h.isEmpty()
    assert best == O.knapsack_dp(w, v, cap)


def test_driver_rejects_bad_heaps(gpu):
    import gpuheap_amd as G
    w, v, cap = A.knapsack_instance(10, 0)
    with Heap(64, 256, variant="buitdd_np", key="u64") as h:
        with pytest.raises(G.BgpqError, match="variant"):
            A.knapsack(h, w, v, cap)
    with Heap(64, 256, variant="buitdd", key="u32") as h:
        with pytest.raises(G.BgpqError, match="u64"):
            A.knapsack(h, w, v, cap)
