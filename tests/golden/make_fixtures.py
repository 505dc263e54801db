"""Writes tests/golden/fixtures.json: small input/output vectors that pin the
oracle, the generators and the GPU results.  Run from the repository root:

    python tests/golden/make_fixtures.py

Each expected output comes from a method independent of the code it pins:
  * SplitMix64: Vigna's published first outputs for state 0, the rest from
    a pure-Python restatement;
  * key streams and heap-sort outputs: numpy's sort of the generated keys
    (SHA-256 of the little-endian bytes);
  * rev_idx: bit reversal by string manipulation;
  * SSSP distances / A* costs: scipy.sparse.csgraph.dijkstra;
  * knapsack optima: brute force (n = 14) and numpy DP (n = 200).
"""
import hashlib
import itertools
import json
import os
import sys

import numpy as np
import scipy.sparse as sp
from scipy.sparse.csgraph import dijkstra

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

import _oracle as O  # noqa: E402
from gpuheap_amd import apps as A  # noqa: E402


def splitmix_py(state):
    m = (1 << 64) - 1
    state = (state + 0x9E3779B97F4A7C15) & m
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    return state, z ^ (z >> 31)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def rev_str(x):
    if x <= 1:
        return x
    b = bin(x)[3:]  # bits below the leading one
    return int("1" + b[::-1], 2)


def sssp_scipy(rowptr, col, w, src):
    n = len(rowptr) - 1
    s = np.repeat(np.arange(n), np.diff(rowptr))
    order = np.lexsort((w, col, s))
    s, c, ww = s[order], col[order], w[order].astype(float)
    keep = np.ones(len(s), bool)
    keep[1:] = (s[1:] != s[:-1]) | (c[1:] != c[:-1])
    g = sp.csr_matrix((ww[keep], (s[keep], c[keep])), shape=(n, n))
    d = dijkstra(g, directed=True, indices=src)
    out = np.full(n, 0xFFFFFFFF, np.uint64)
    out[np.isfinite(d)] = d[np.isfinite(d)].astype(np.uint64)
    return out.astype(np.uint32)


def astar_scipy(free, width, height):
    rows, cols, ws = [], [], []
    for y in range(height):
        for x in range(width):
            if not free[y * width + x]:
                continue
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    nx, ny = x + dx, y + dy
                    if (dx or dy) and 0 <= nx < width and 0 <= ny < height and free[ny * width + nx]:
                        rows.append(y * width + x)
                        cols.append(ny * width + nx)
                        ws.append(14 if dx and dy else 10)
    g = sp.csr_matrix((np.array(ws, float), (rows, cols)), shape=(width * height,) * 2)
    d = dijkstra(g, directed=True, indices=0)[-1]
    return int(d) if np.isfinite(d) else 0xFFFFFFFF


def knap_dp_numpy(w, v, cap):
    f = np.zeros(cap + 1, np.int64)
    for wi, vi in zip(w.astype(np.int64), v.astype(np.int64)):
        if wi <= cap:
            f[wi:] = np.maximum(f[wi:], f[:cap + 1 - wi] + vi)
    return int(f[cap])


def main():
    fx = {}
    # the first two are the published outputs for state 0; the rest follow
    # from the same arithmetic in pure Python (a second implementation)
    st, outs = 0, []
    for _ in range(8):
        st, z = splitmix_py(st)
        outs.append(hex(z))
    assert outs[:2] == ["0xe220a8397b1dcdaf", "0x6e789e6aa1b965f4"]
    fx["splitmix64_state0"] = outs
    fx["rev_idx_1_64"] = [rev_str(x) for x in range(1, 65)]
    keys = {}
    for n_log2 in (12, 20):
        k32 = O.keys_u32(0, 1 << n_log2)
        keys[f"u32_seed0_2^{n_log2}"] = {"first8": [int(x) for x in k32[:8]], "sha256": sha(k32),
                                        "sorted_sha256": sha(np.sort(k32))}
    k64 = O.keys_u64(0, 1 << 12)
    keys["u64_seed0_2^12"] = {"first4": [hex(int(x)) for x in k64[:4]], "sha256": sha(k64),
                              "sorted_sha256": sha(np.sort(k64))}
    fx["keys"] = keys
    rowptr, col, w = A.rmat_csr(10, 16, 0)
    d = sssp_scipy(rowptr, col, w, 0)
    fx["sssp_rmat10_seed0_src0"] = {"edges": int(len(col)), "rowptr_sha256": sha(rowptr), "col_sha256": sha(col),
                                    "w_sha256": sha(w), "dist_sha256": sha(d),
                                    "reachable": int(np.count_nonzero(d != 0xFFFFFFFF))}
    astar = {}
    for side, seed in ((64, 0), (64, 1), (100, 2)):
        free = A.grid_map(side, side, 0.25, seed)
        astar[f"{side}x{side}_seed{seed}"] = {"free_sha256": sha(free), "cost": astar_scipy(free, side, side)}
    fx["astar_p025"] = astar
    knap = {}
    for seed in range(3):
        w14, v14, c14 = A.knapsack_instance(14, seed)
        best = 0
        for mask in itertools.product((0, 1), repeat=14):
            m = np.array(mask, bool)
            if w14[m].sum() <= c14:
                best = max(best, int(v14[m].sum()))
        w2, v2, c2 = A.knapsack_instance(200, seed)
        knap[f"seed{seed}"] = {"n14_capacity": int(c14), "n14_optimum": best, "n200_capacity": int(c2),
                               "n200_optimum": knap_dp_numpy(w2, v2, c2), "n200_weights_sha256": sha(w2)}
    fx["knapsack"] = knap
    with open(os.path.join(HERE, "fixtures.json"), "w") as f:
        json.dump(fx, f, indent=1, sort_keys=True)
    print("wrote", os.path.join(HERE, "fixtures.json"))


if __name__ == "__main__":
    main()
