"""Parity at the configurations' full sizes, through the C ABI.

* Heap sort of 2^28 keys (the at-scale configuration), each output compared
  key by key with numpy's sort of the oracle's keys (the reference checks
  against std::sort, test/heap_test.cu:118,165-170): u32 random through
  buitdd at B=1024, u32 in descending order (the adversarial order of
  artifacts/synthetic/sort.cu:115-119) through tditdd at B=2048, and u64
  keys with their input position as payload through buitdd_np at B=1024.
* The descending and u64 key sets at 2^22 keys for every variant and the
  configuration's batch sizes 256..2048.
* SSSP on R-MAT scale 22 (4.2 M vertices, 67 M edges) at B=1024, equal to
  the CPU Dijkstra bit for bit; A* on the 4096 x 4096 grid, equal to the
  CPU A*'s optimal cost.
"""
import numpy as np
import pytest

import _oracle as O
import gpuheap_amd as G
from gpuheap_amd import DeviceBuffer, Heap, batch_num_for
from gpuheap_amd import apps as A

pytestmark = pytest.mark.gpu

N28 = 1 << 28
N22 = 1 << 22


def _keys(key, order, n):
    if key == "u64":
        k = O.keys_u64(0, n)
    else:
        k = O.keys_u32(0, n)
    if order == "descending":
        k = np.sort(k)[::-1].copy()
    return k


def _heapsort(keys, variant, B):
    n = len(keys)
    key = "u64" if keys.dtype == np.uint64 else "u32"
    d = DeviceBuffer.from_numpy(keys)
    with Heap(batch_num_for(n, B), B, variant=variant, key=key) as h:
        t_ins, t_del = h.heapsort(d, n)
        assert h.error_flags() == 0 and h.isEmpty()
    out = d.to_numpy(keys.dtype, n)
    d.free()
    return out, t_ins, t_del


@pytest.mark.parametrize("key,order,variant,B", [
    ("u32", "random", "buitdd", 1024),
    ("u32", "descending", "tditdd", 2048),
    ("u64", "random", "buitdd_np", 1024),
])
def test_heapsort_2p28_equals_sort(gpu, key, order, variant, B):
    print(f"[case] heap sort 2^28 {key} {order} {variant} B={B}", flush=True)
    keys = _keys(key, order, N28)
    out, t_ins, t_del = _heapsort(keys, variant, B)
    print(f"[case] GPU {N28 / ((t_ins + t_del) / 1e3):.3e} keys/s; sorting on the CPU", flush=True)
    keys.sort()
    assert np.array_equal(out, keys), f"{np.count_nonzero(out != keys)} keys differ"


@pytest.mark.parametrize("variant", list(G.VARIANTS))
@pytest.mark.parametrize("B", [256, 512, 1024, 2048])
@pytest.mark.parametrize("key,order", [("u32", "descending"), ("u64", "random")])
def test_heapsort_2p22_key_sets(gpu, variant, B, key, order):
    print(f"[case] heap sort 2^22 {key} {order} {variant} B={B}", flush=True)
    keys = _keys(key, order, N22)
    out, _, _ = _heapsort(keys, variant, B)
    assert np.array_equal(out, np.sort(keys))


def test_sssp_rmat22_b1024(gpu):
    print("[case] SSSP R-MAT scale 22 B=1024", flush=True)
    rowptr, col, w = A.rmat_csr(22, 16, 0)
    g = A.Graph(rowptr, col, w)
    with Heap(batch_num_for(len(col), 1024), 1024, variant="buitdd", key="u64") as h:
        dist, st = A.sssp(h, g, source=0, grid=32)
        assert st["error_flags"] == 0 and h.isEmpty()
    print(f"[case] GPU solve {st['ms']:.1f} ms; Dijkstra on the CPU", flush=True)
    want = O.dijkstra(rowptr, col, w, 0)
    assert np.array_equal(dist, want), f"{np.count_nonzero(dist != want)} distances differ"


def test_astar_4096(gpu):
    print("[case] A* 4096 x 4096 p=0.25", flush=True)
    size = 4096
    free = A.grid_map(size, size, 0.25, 0)
    gm = A.GridMap(free, size, size)
    with Heap(batch_num_for(4 * size * size, 256), 256, variant="buitdd", key="u64") as h:
        cost, st = A.astar(h, gm, grid=4)
        assert st["error_flags"] == 0 and h.isEmpty()
    want, expanded = O.astar(free, size, size)
    print(f"[case] GPU {st['ms']:.1f} ms, {st['expanded']} expansions (CPU {expanded})", flush=True)
    assert cost == want
