"""CPU checks of the driver oracles and of the input generators.

* orc_dijkstra against scipy.sparse.csgraph.dijkstra (an independent
  implementation) on R-MAT graphs;
* orc_astar's optimal cost against Dijkstra on the same grid graph (an
  admissible heuristic leaves the optimum unchanged), including an
  unreachable goal;
* orc_knapsack_dp against brute-force enumeration, and the best-first
  branch-and-bound (with and without dominance) against the DP;
* the host generators of libbgpq.so against a pure-Python restatement of
  their documented streams (include/bgpq_apps.h), plus their statistics.
These run without a GPU: the generators are host code.
"""
import itertools

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import dijkstra as sp_dijkstra

import _oracle as O
from gpuheap_amd import apps as A

GOLDEN = 0x9E3779B97F4A7C15
M64 = (1 << 64) - 1


def _splitmix(state):
    state = (state + GOLDEN) & M64
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return state, z ^ (z >> 31)


def _scipy_dist(rowptr, col, w, source):
    n = len(rowptr) - 1
    src = np.repeat(np.arange(n), np.diff(rowptr))
    # parallel edges: keep the lightest (a sparse matrix would add them up)
    order = np.lexsort((w, col, src))
    s, c, ww = src[order], col[order], w[order].astype(np.float64)
    first = np.ones(len(s), bool)
    first[1:] = (s[1:] != s[:-1]) | (c[1:] != c[:-1])
    g = sp.csr_matrix((ww[first], (s[first], c[first])), shape=(n, n))
    d = sp_dijkstra(g, directed=True, indices=source)
    out = np.full(n, O.U32_MAX, np.uint64)
    fin = np.isfinite(d)
    out[fin] = d[fin].astype(np.uint64)
    return out.astype(np.uint32)


@pytest.mark.parametrize("scale,degree,seed", [(8, 4, 0), (10, 8, 1), (11, 16, 2)])
def test_dijkstra_matches_scipy(scale, degree, seed):
    rowptr, col, w = A.rmat_csr(scale, degree, seed)
    for source in (0, 5):
        assert np.array_equal(O.dijkstra(rowptr, col, w, source), _scipy_dist(rowptr, col, w, source))


def test_dijkstra_unreachable_and_isolated():
    # 0 -> 1 -> 2, vertex 3 isolated, 4 -> 0 only
    rowptr = np.array([0, 1, 2, 2, 2, 3], np.int32)
    col = np.array([1, 2, 0], np.int32)
    w = np.array([7, 5, 1], np.uint32)
    d = O.dijkstra(rowptr, col, w, 0)
    assert list(d) == [0, 7, 12, O.U32_MAX, O.U32_MAX]


def _grid_graph(free, width, height):
    rows, cols, ws = [], [], []
    for y in range(height):
        for x in range(width):
            c = y * width + x
            if not free[c]:
                continue
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    if dx == 0 and dy == 0:
                        continue
                    nx, ny = x + dx, y + dy
                    if 0 <= nx < width and 0 <= ny < height and free[ny * width + nx]:
                        rows.append(c)
                        cols.append(ny * width + nx)
                        ws.append(14 if dx and dy else 10)
    return sp.csr_matrix((np.array(ws, float), (rows, cols)), shape=(width * height,) * 2)


@pytest.mark.parametrize("width,height,p,seed", [(40, 40, 0.25, 0), (64, 48, 0.25, 1), (33, 70, 0.3, 2),
                                                 (30, 30, 0.6, 3)])
def test_astar_cost_matches_dijkstra(width, height, p, seed):
    free = A.grid_map(width, height, p, seed)
    cost, expanded = O.astar(free, width, height)
    d = sp_dijkstra(_grid_graph(free, width, height), directed=True, indices=0)[-1]
    want = O.U32_MAX if not np.isfinite(d) else int(d)
    assert cost == want
    assert expanded >= 1


def test_astar_open_grid_closed_form():
    # no obstacles: the octile distance itself
    free = np.ones(50 * 20, np.uint8)
    cost, _ = O.astar(free, 50, 20)
    assert cost == 14 * 19 + 10 * (49 - 19)


@pytest.mark.parametrize("seed", range(6))
def test_knapsack_dp_matches_brute_force(seed):
    w, v, cap = A.knapsack_instance(12, seed)
    best = 0
    for mask in itertools.product((0, 1), repeat=12):
        m = np.array(mask, bool)
        if w[m].sum() <= cap:
            best = max(best, int(v[m].sum()))
    assert O.knapsack_dp(w, v, cap) == best


@pytest.mark.parametrize("n,seed", [(30, 0), (60, 1), (100, 2), (200, 0), (200, 1), (200, 2)])
def test_knapsack_bnb_matches_dp(n, seed):
    w, v, cap = A.knapsack_instance(n, seed)
    best, nodes, rc = O.knapsack_bnb(w, v, cap, dominance=True)
    assert rc == 0 and best == O.knapsack_dp(w, v, cap)
    assert nodes <= (n + 1) * (cap + 1)


def test_knapsack_bnb_without_dominance_small():
    w, v, cap = A.knapsack_instance(30, 5)
    best, _, rc = O.knapsack_bnb(w, v, cap, dominance=False)
    assert rc == 0 and best == O.knapsack_dp(w, v, cap)


# ---- generators against their documented streams


def _rmat_spec(scale, degree, seed, abc, wmin, wmax):
    ta, tab, tabc = (int(x * 4294967296.0) for x in (abc[0], abc[0] + abc[1], sum(abc)))
    edges = []
    for e in range(degree << scale):
        st = (seed + (e + 1) * GOLDEN) & M64
        s = d = 0
        bits = 0
        for lvl in range(scale):
            if lvl % 2 == 0:
                st, bits = _splitmix(st)
                x = bits >> 32
            else:
                x = bits & 0xFFFFFFFF
            sb = 1 if x >= tab else 0
            db = 1 if (ta <= x < tab) or x >= tabc else 0
            s, d = (s << 1) | sb, (d << 1) | db
        st, r = _splitmix(st)
        edges.append((s, d, wmin + r % (wmax - wmin + 1)))
    return edges


def test_rmat_follows_spec():
    scale, degree, seed = 5, 3, 11
    rowptr, col, w = A.rmat_csr(scale, degree, seed)
    edges = _rmat_spec(scale, degree, seed, A.RMAT_ABC, 1, 1000)
    n = 1 << scale
    want_rowptr = np.zeros(n + 1, np.int64)
    for s, _, _ in edges:
        want_rowptr[s + 1] += 1
    assert np.array_equal(rowptr, np.cumsum(want_rowptr))
    # within a vertex, generation order
    got = [(int(s), int(c), int(x)) for s in range(n) for c, x in zip(col[rowptr[s]:rowptr[s + 1]],
                                                                      w[rowptr[s]:rowptr[s + 1]])]
    want = sorted(edges, key=lambda t: t[0])  # stable: generation order per source
    assert got == want


def test_rmat_statistics():
    rowptr, col, w = A.rmat_csr(14, 16, 0)
    ne = 16 << 14
    assert rowptr[0] == 0 and rowptr[-1] == ne and np.all(np.diff(rowptr) >= 0)
    assert col.min() >= 0 and col.max() < (1 << 14)
    assert w.min() >= 1 and w.max() <= 1000
    src = np.repeat(np.arange(1 << 14), np.diff(rowptr))
    top_src = np.mean(src < (1 << 13))  # a + b of the first level
    top_dst = np.mean(col < (1 << 13))  # a + c
    assert abs(top_src - 0.76) < 0.01 and abs(top_dst - 0.76) < 0.01
    r2 = A.rmat_csr(14, 16, 0)
    assert np.array_equal(r2[1], col) and not np.array_equal(A.rmat_csr(14, 16, 1)[1], col)


def test_grid_follows_spec():
    width, height, p, seed = 7, 5, 0.25, 3
    free = A.grid_map(width, height, p, seed)
    t = int(p * 4294967296.0)
    want = []
    for i in range(width * height):
        _, z = _splitmix((seed + (i + 1) * GOLDEN) & M64)
        want.append(0 if (z >> 32) < t else 1)
    want[0] = want[-1] = 1
    assert list(free) == want


def test_grid_statistics():
    free = A.grid_map(512, 512, 0.25, 0)
    assert free[0] == 1 and free[-1] == 1
    assert abs(1.0 - free.mean() - 0.25) < 0.005


def test_knapsack_instance_follows_spec():
    w, v, cap = A.knapsack_instance(9, 4)
    want = []
    for i in range(9):
        _, z = _splitmix((4 + (i + 1) * GOLDEN) & M64)
        want.append(1 + z % 1000)
    assert list(w) == want
    assert np.array_equal(v, w + 100)
    assert cap == int(w.sum()) // 2
