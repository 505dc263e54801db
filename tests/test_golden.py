"""The oracle and the input generators against tests/golden/fixtures.json
(written by tests/golden/make_fixtures.py from independent methods)."""
import hashlib
import json
import os

import numpy as np
import pytest

import _oracle as O
from gpuheap_amd import apps as A

FX = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fixtures.json")))


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_splitmix64():
    assert [hex(x) for x in O.splitmix64_stream(0, 8)] == FX["splitmix64_state0"]


def test_rev_idx_table():
    assert [O.lib().orc_rev_idx(x) for x in range(1, 65)] == FX["rev_idx_1_64"]


@pytest.mark.parametrize("n_log2", [12, 20])
def test_key_stream_u32(n_log2):
    k = O.keys_u32(0, 1 << n_log2)
    f = FX["keys"][f"u32_seed0_2^{n_log2}"]
    assert [int(x) for x in k[:8]] == f["first8"] and sha(k) == f["sha256"]


def test_key_stream_u64():
    k = O.keys_u64(0, 1 << 12)
    f = FX["keys"]["u64_seed0_2^12"]
    assert [hex(int(x)) for x in k[:4]] == f["first4"] and sha(k) == f["sha256"]


@pytest.mark.parametrize("variant", list(O.VARIANTS))
def test_model_heapsort_matches_sorted_fixture(variant):
    k = O.keys_u32(0, 1 << 12)
    B = 64
    out = O.heapsort(variant, B, 256, O.U32_MAX, k.astype(np.uint64), chunk=B)
    assert sha(out.astype(np.uint32)) == FX["keys"]["u32_seed0_2^12"]["sorted_sha256"]


def test_rmat_and_dijkstra():
    f = FX["sssp_rmat10_seed0_src0"]
    rowptr, col, w = A.rmat_csr(10, 16, 0)
    assert (sha(rowptr), sha(col), sha(w)) == (f["rowptr_sha256"], f["col_sha256"], f["w_sha256"])
    assert sha(O.dijkstra(rowptr, col, w, 0)) == f["dist_sha256"]


@pytest.mark.parametrize("name", sorted(FX["astar_p025"]))
def test_grid_and_astar(name):
    f = FX["astar_p025"][name]
    side, seed = int(name.split("x")[0]), int(name.split("seed")[1])
    free = A.grid_map(side, side, 0.25, seed)
    assert sha(free) == f["free_sha256"]
    assert O.astar(free, side, side)[0] == f["cost"]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_knapsack(seed):
    f = FX["knapsack"][f"seed{seed}"]
    w, v, cap = A.knapsack_instance(14, seed)
    assert cap == f["n14_capacity"] and O.knapsack_dp(w, v, cap) == f["n14_optimum"]
    w, v, cap = A.knapsack_instance(200, seed)
    assert cap == f["n200_capacity"] and sha(w) == f["n200_weights_sha256"]
    assert O.knapsack_dp(w, v, cap) == f["n200_optimum"]
    assert O.knapsack_bnb(w, v, cap)[0] == f["n200_optimum"]
