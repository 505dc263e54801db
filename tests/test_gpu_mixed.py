"""The mixed workload end to end at a small size: the plain
one-heap-per-launch path and the persistent serve kernel (closed-loop and at
a Poisson rate) give identical answers, equal to the CPU oracles."""
import numpy as np
import pytest

import _oracle as O
from gpuheap_amd import mixed as M
from gpuheap_amd import serve as S

pytestmark = pytest.mark.gpu


def _oracle(cat, kind, key):
    if kind == "sssp":
        rowptr, col, w = cat.graphs[key[0]]["csr"]
        dist = O.dijkstra(rowptr, col, w, key[1])
        return int(np.count_nonzero(dist != O.U32_MAX)), S.sssp_digest(dist)
    if kind == "astar":
        side = cat.grids[key]["dev"].width
        return O.astar(cat.grids[key]["free"], side, side)[0], 0
    w, v, cap = cat.knaps[key]
    return O.knapsack_dp(w, v, cap), 0


def test_mixed_plain_vs_served(gpu):
    print("[case] mixed small", flush=True)
    cat = M.Catalog(scales=(8, 9, 10, 11), sides=(64, 128), grid_seeds=2, knap_instances=4)
    stream = M.make_stream(cat, 60, seed=5)
    want = {q: _oracle(cat, *q) for q in set(stream)}
    plain, _, _ = M.run_plain(cat, stream)
    table = {"sssp:8": {"B": 256, "team": 1}, "sssp:11": {"B": 1024, "team": 4},
             "astar:64": {"B": 256, "team": 2}, "knapsack:200": {"B": 512, "team": 3}}
    closed, _, _ = M.run_served(cat, stream, table, rate=0.0, budget=2 << 30, workgroups=48)
    timed, lat, _ = M.run_served(cat, stream, table, rate=2000.0, budget=256 << 20, workgroups=48)
    assert np.all(lat > 0)
    for q, a, b, c in zip(stream, plain, closed, timed):
        assert b["status"] == 0 and c["status"] == 0
        assert M.answer_of(*q, a) == M.answer_of(*q, b) == M.answer_of(*q, c) == want[q], q
