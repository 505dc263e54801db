"""GPU results against tests/golden/fixtures.json: the device key generator,
the heap sort of 2^20 keys (every variant), and the drivers."""
import hashlib
import json
import os

import numpy as np
import pytest

import gpuheap_amd as G
from gpuheap_amd import apps as A
from gpuheap_amd._lib import check, lib

pytestmark = pytest.mark.gpu

FX = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fixtures.json")))


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _dev_keys(kt, n, dtype):
    d = G.DeviceBuffer(n * np.dtype(dtype).itemsize)
    check(lib().bgpq_gen_keys(kt, 0, n, d.ptr, None))
    G.sync()
    return d


def test_device_key_generator(gpu):
    d = _dev_keys(0, 1 << 20, np.uint32)
    assert sha(d.to_numpy(np.uint32, 1 << 20)) == FX["keys"]["u32_seed0_2^20"]["sha256"]
    d = _dev_keys(1, 1 << 12, np.uint64)
    assert sha(d.to_numpy(np.uint64, 1 << 12)) == FX["keys"]["u64_seed0_2^12"]["sha256"]


@pytest.mark.parametrize("variant", list(G.VARIANTS))
@pytest.mark.parametrize("B", [1024, 8192])
def test_heapsort_2p20(gpu, variant, B):
    print(f"[case] golden heapsort {variant} B={B}", flush=True)
    n = 1 << 20
    d = _dev_keys(0, n, np.uint32)
    out = G.DeviceBuffer(4 * n)
    with G.Heap(G.batch_num_for(n, B), B, variant=variant) as h:
        h.heapsort(d, n, d_out=out)
    assert sha(out.to_numpy(np.uint32, n)) == FX["keys"]["u32_seed0_2^20"]["sorted_sha256"]


def test_drivers(gpu):
    f = FX["sssp_rmat10_seed0_src0"]
    rowptr, col, w = A.rmat_csr(10, 16, 0)
    with G.Heap(64, 256, variant="buitdd", key="u64") as h:
        dist, _ = A.sssp(h, A.Graph(rowptr, col, w), 0, grid=16)
        assert sha(dist) == f["dist_sha256"]
    for name, fa in FX["astar_p025"].items():
        side, seed = int(name.split("x")[0]), int(name.split("seed")[1])
        with G.Heap(256, 256, variant="tditdd", key="u64") as h:
            cost, _ = A.astar(h, A.GridMap(A.grid_map(side, side, 0.25, seed), side, side), grid=16)
        assert cost == fa["cost"]
    for seed in range(3):
        w, v, cap = A.knapsack_instance(200, seed)
        with G.Heap(4096, 256, variant="buitdd", key="u64") as h:
            best, _ = A.knapsack(h, w, v, cap, grid=32)
        assert best == FX["knapsack"][f"seed{seed}"]["n200_optimum"]
