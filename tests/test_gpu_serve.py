"""GPU tests of the batched and served solves (include/bgpq_serve.h).

* A mixed batch (SSSP, A*, knapsack; several batch sizes and team sizes)
  through the captured hipGraph, replayed: every result equals its CPU oracle
  (SSSP distances bit for bit, A* costs, knapsack optima).
* The persistent serve kernel: each query gives the same result alone (the
  only query in flight) and in a mix submitted all at once, and both equal
  the oracle.
* Admission: with a budget that holds one query at a time, submissions
  wait (EAGAIN) until completions free the arena, and all complete.
"""
import time

import numpy as np
import pytest

import _oracle as O
from gpuheap_amd import apps as A
from gpuheap_amd import serve as S

pytestmark = pytest.mark.gpu

GiB = 1 << 30


def _catalog():
    graphs = [(A.Graph(*A.rmat_csr(s, 16, s)), A.rmat_csr(s, 16, s)) for s in (8, 10, 12)]
    grids = []
    for size, seed in ((64, 0), (96, 1), (128, 2)):
        free = A.grid_map(size, size, 0.25, seed)
        grids.append((A.GridMap(free, size, size), free, size))
    knaps = [A.knapsack_instance(n, seed) for n, seed in ((30, 0), (60, 1), (100, 2))]
    return graphs, grids, knaps


def _queries(graphs, grids, knaps, with_labels=True):
    qs, want = [], []
    tag = 0
    for (g, (rp, col, w)), B, team in zip(graphs, (256, 512, 1024), (1, 2, 4)):
        for src in (0, 3):
            out = np.zeros(g.n, np.uint32) if with_labels else None
            qs.append(S.sssp_query(g, src, batch_size=B, team=team, tag=tag, labels_out=out))
            dist = O.dijkstra(rp, col, w, src)
            want.append(("sssp", S.sssp_digest(dist), int(np.count_nonzero(dist != O.U32_MAX)), out, dist))
            tag += 1
    for (gm, free, size), B, team in zip(grids, (256, 512, 256), (1, 2, 3)):
        qs.append(S.astar_query(gm, batch_size=B, team=team, tag=tag))
        want.append(("astar", None, O.astar(free, size, size)[0], None, None))
        tag += 1
    for (w, v, cap), B, team in zip(knaps, (256, 1024, 512), (1, 2, 2)):
        qs.append(S.knapsack_query(w, v, cap, batch_size=B, team=team, tag=tag))
        want.append(("knapsack", None, O.knapsack_dp(w, v, cap), None, None))
        tag += 1
    return qs, want


def _check(res, want):
    kind, digest, value, out, dist = want
    assert res["status"] == 0, res
    assert res["value"] == value, (kind, res["value"], value)
    if kind == "sssp":
        assert res["digest"] == digest
        if out is not None:
            assert np.array_equal(out, dist)


def test_batch_graph_mixed(gpu):
    print("[case] batch graph mixed", flush=True)
    graphs, grids, knaps = _catalog()
    qs, want = _queries(graphs, grids, knaps)
    res, ms = S.batch_run(qs, 4 * GiB, workgroups=64, repeats=4)
    assert ms > 0
    for r, w in zip(res, want):
        _check(r, w)


def _drain(srv, n, timeout=60.0):
    got = {}
    t0 = time.time()
    while len(got) < n:
        for r in srv.poll():
            got[r["tag"]] = r
        assert time.time() - t0 < timeout, f"{n - len(got)} queries still running"
        time.sleep(0.001)
    return got


def test_server_alone_vs_mixed(gpu):
    print("[case] server alone vs mixed", flush=True)
    graphs, grids, knaps = _catalog()
    alone_q, want = _queries(graphs, grids, knaps, with_labels=False)
    with S.Server(4 * GiB, workgroups=64) as srv:
        alone = {}
        for q in alone_q:
            assert srv.submit(q)
            alone.update(_drain(srv, 1))
        mixed_q, _ = _queries(graphs, grids, knaps, with_labels=False)
        for q in mixed_q:
            while not srv.submit(q):
                time.sleep(0.001)
        mixed = _drain(srv, len(mixed_q))
    for tag, w in enumerate(want):
        _check(alone[tag], w)
        _check(mixed[tag], w)
        assert (alone[tag]["value"], alone[tag]["digest"]) == (mixed[tag]["value"], mixed[tag]["digest"])


def test_server_budget_admission(gpu):
    print("[case] server admission", flush=True)
    graphs, grids, knaps = _catalog()
    qs, want = _queries(graphs, grids, knaps, with_labels=False)
    big = max(S.query_bytes(q) for q in qs)
    with S.Server(big + (1 << 20), workgroups=32) as srv:
        waited = 0
        got = {}
        for q in qs:
            while not srv.submit(q):
                waited += 1
                for r in srv.poll():
                    got[r["tag"]] = r
                time.sleep(0.0005)
            assert srv.used_bytes() <= big + (1 << 20)
        got.update(_drain(srv, len(qs) - len(got)))
    assert waited > 0
    for tag, w in enumerate(want):
        _check(got[tag], w)


def test_server_rejects_bad_queries(gpu):
    import gpuheap_amd as G
    w, v, cap = A.knapsack_instance(10, 0)
    with S.Server(1 << 28, workgroups=4) as srv:
        with pytest.raises(G.BgpqError, match="batch size"):
            srv.submit(S.knapsack_query(w, v, cap, batch_size=2048))
        with pytest.raises(G.BgpqError, match="team"):
            srv.submit(S.knapsack_query(w, v, cap, team=0))


def test_server_idle_limit_refuses_after_workgroups_leave(gpu):
    """Idle workgroups take no ring slot: once all of them have left at a
    short idle limit, the kernel has ended normally, a submission is refused
    with an error instead of being queued for no one, and close() is clean.
    A server whose workgroups are still there serves a query posted after a
    long idle wait."""
    print("[case] serve idle limit", flush=True)
    gm = A.GridMap(A.grid_map(64, 64, 0.25, 0), 64, 64)
    want = O.astar(A.grid_map(64, 64, 0.25, 0), 64, 64)[0]
    srv = S.Server(1 * GiB, workgroups=4)
    try:
        srv.set_idle_limit(200)  # about a millisecond of polling
        srv.start()
        t0 = time.time()
        while srv.live() > 0 and time.time() - t0 < 30:
            time.sleep(0.01)
        assert srv.live() == 0
        with pytest.raises(Exception, match="no workgroup left"):
            srv.submit(S.astar_query(gm, batch_size=256, team=1, tag=1))
        srv.close()
    finally:
        srv.destroy()
    # the default limit: workgroups stay, and a query posted late is served
    srv = S.Server(1 * GiB, workgroups=4)
    try:
        srv.start()
        time.sleep(0.5)
        assert srv.live() == 2 * 4  # both tiers' kernels, 4 workgroups each
        assert srv.submit(S.astar_query(gm, batch_size=256, team=2, tag=2))
        got = []
        t0 = time.time()
        while not got and time.time() - t0 < 60:
            got = srv.poll()
            time.sleep(0.002)
        assert got and got[0]["status"] == 0 and got[0]["value"] == want
        srv.close()
    finally:
        srv.destroy()


def test_server_labels_out_while_serving(gpu):
    """A served SSSP query's labels come back through labels_out while the
    serve kernels keep running at full size (one workgroup per CU and
    tier), promptly, and equal Dijkstra's distances."""
    print("[case] serve labels_out", flush=True)
    rp, col, w = A.rmat_csr(12, 16, 12)
    g = A.Graph(rp, col, w)
    dist = O.dijkstra(rp, col, w, 5)
    with S.Server(1 * GiB) as srv:
        for tag, (B, team) in enumerate(((1024, 4), (256, 2))):
            out = np.zeros(g.n, np.uint32)
            t0 = time.time()
            assert srv.submit(S.sssp_query(g, 5, batch_size=B, team=team, tag=tag, labels_out=out))
            got = _drain(srv, 1, timeout=20.0)
            assert time.time() - t0 < 5.0, "a served query with labels_out took too long"
            assert got[tag]["status"] == 0
            assert np.array_equal(out, dist)
