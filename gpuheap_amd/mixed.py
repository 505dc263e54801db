"""The batched + served mixed workload: a stream of independent SSSP, A* and
knapsack queries, each with its own device heap.

* ``Catalog`` holds the inputs on the device: R-MAT graphs of 2^10..2^20
  vertices (degree 16, weights 1..1000), obstacle grids of 256^2..2048^2
  (p = 0.25), knapsack instances (n = 200, values = weight + 100, capacity
  half the weight).  Queries draw from it: an SSSP query is a (graph,
  source) pair, an A* query a grid, a knapsack query an instance.
* ``make_stream`` draws the 40 / 40 / 20 mix with log-uniform sizes.
* ``run_plain`` is the one-heap-per-launch path: one query at a time, each a
  single-solve launch (bgpq_sssp / bgpq_astar / bgpq_knapsack) on a reused
  heap, as the reference's applications run.
* ``run_served`` feeds the persistent serve kernel with Poisson arrivals at
  a given rate and records each query's latency (arrival to completion).
* ``autotune`` picks the batch size and team size of every size class from
  the latency of the class's costliest query alone on the serve kernel,
  among the configurations that use their workgroups efficiently.
"""
import json
import time

import numpy as np

from . import apps as A
from . import serve as S
from .heap import Heap

SSSP_SCALES = tuple(range(10, 21))
ASTAR_SIDES = (256, 512, 1024, 2048)
KNAP_ITEMS = 200
MIX = (0.4, 0.4, 0.2)


def size_class(kind, key):
    """The autotuner's class of a query: ('sssp', scale), ('astar', side), ('knapsack', n)."""
    if kind == "sssp":
        return "sssp", key[0]
    if kind == "astar":
        return "astar", key[0]
    return "knapsack", KNAP_ITEMS


class Catalog:
    """Device-resident inputs of the mixed workload and their CPU answers
    (computed on demand by `oracle`, which the caller supplies)."""

    def __init__(self, scales=SSSP_SCALES, sides=ASTAR_SIDES, grid_seeds=2, sources_per_graph=4, knap_instances=64,
                 seed=0):
        rng = np.random.default_rng(seed)
        self.graphs = {}
        for s in scales:
            rowptr, col, w = A.rmat_csr(s, 16, 1000 + s)
            deg = np.diff(rowptr)
            cand = np.flatnonzero(deg > 0)
            src = rng.choice(cand, size=min(sources_per_graph, len(cand)), replace=False)
            self.graphs[s] = {"dev": A.Graph(rowptr, col, w), "csr": (rowptr, col, w),
                              "sources": [int(x) for x in src]}
        self.grids = {}
        for side in sides:
            for k in range(grid_seeds):
                free = A.grid_map(side, side, 0.25, 2000 + 16 * side + k)
                self.grids[(side, k)] = {"dev": A.GridMap(free, side, side), "free": free}
        self.knaps = [A.knapsack_instance(KNAP_ITEMS, 3000 + i) for i in range(knap_instances)]
        self.scales, self.sides, self.grid_seeds = tuple(scales), tuple(sides), grid_seeds

    def keys(self):
        """Every distinct problem of the catalog."""
        out = [("sssp", (s, src)) for s in self.scales for src in self.graphs[s]["sources"]]
        out += [("astar", g) for g in self.grids]
        out += [("knapsack", i) for i in range(len(self.knaps))]
        return out

    def query(self, kind, key, B, team, tag):
        if kind == "sssp":
            return S.sssp_query(self.graphs[key[0]]["dev"], key[1], batch_size=B, team=team, tag=tag)
        if kind == "astar":
            return S.astar_query(self.grids[key]["dev"], batch_size=B, team=team, tag=tag)
        w, v, cap = self.knaps[key]
        return S.knapsack_query(w, v, cap, batch_size=B, team=team, tag=tag)


def make_stream(cat, n, seed=1, stratified=False):
    """n (kind, key) draws: 40% SSSP (scale uniform over the catalog's
    scales, i.e. log-uniform vertices), 40% A* (side log-uniform), 20%
    knapsack.  stratified: exactly those shares of the n queries, in a
    random order."""
    rng = np.random.default_rng(seed)
    if stratified:
        counts = [int(round(MIX[0] * n)), int(round(MIX[1] * n))]
        counts.append(n - sum(counts))
        kinds = rng.permutation(np.repeat(np.arange(3), counts))
    else:
        kinds = rng.choice(3, size=n, p=MIX)
    out = []
    for k in kinds:
        if k == 0:
            s = cat.scales[rng.integers(len(cat.scales))]
            src = cat.graphs[s]["sources"][rng.integers(len(cat.graphs[s]["sources"]))]
            out.append(("sssp", (s, src)))
        elif k == 1:
            side = cat.sides[rng.integers(len(cat.sides))]
            out.append(("astar", (side, int(rng.integers(cat.grid_seeds)))))
        else:
            out.append(("knapsack", int(rng.integers(len(cat.knaps)))))
    return out


def answer_of(kind, key, result):
    """What is compared with the oracle: (value, digest)."""
    return int(result["value"]), int(result["digest"]) if kind == "sssp" else 0


# ---------------------------------------------------------------- plain path
class PlainRunner:
    """One query at a time through the single-solve drivers, each on a heap
    kept for its kind (emptied by every solve)."""

    # workgroups per driver as measured best for single solves (fewer
    # workgroups pop fewer keys out of order)
    def __init__(self, cat, sssp_B=1024, astar_B=256, knap_B=256, grid_sssp=32, grid_astar=4, grid_knap=8):
        emax = max(g["dev"].m for g in cat.graphs.values())
        cmax = max(g["dev"].width * g["dev"].height for g in cat.grids.values())
        self.cat = cat
        self.h = {"sssp": Heap(_bn(emax, sssp_B), sssp_B, variant="buitdd", key="u64"),
                  "astar": Heap(_bn(4 * cmax, astar_B), astar_B, variant="buitdd", key="u64"),
                  "knapsack": Heap(_bn(1 << 20, knap_B), knap_B, variant="buitdd", key="u64")}
        self.grid = {"sssp": grid_sssp, "astar": grid_astar, "knapsack": grid_knap}

    def solve(self, kind, key):
        if kind == "sssp":
            # the labels are summarised on the device, as the served path does
            g = self.cat.graphs[key[0]]["dev"]
            _, st = A.sssp(self.h[kind], g, key[1], grid=self.grid[kind], fetch=False)
            digest, reached = A.label_digest(g)
            return {"value": reached, "digest": digest, "stats": st}
        if kind == "astar":
            cost, st = A.astar(self.h[kind], self.cat.grids[key]["dev"], grid=self.grid[kind])
            return {"value": cost, "digest": 0, "stats": st}
        w, v, cap = self.cat.knaps[key]
        best, st = A.knapsack(self.h[kind], w, v, cap, grid=self.grid[kind])
        return {"value": best, "digest": 0, "stats": st}

    def close(self):
        for h in self.h.values():
            h.close()


def _bn(keys, B):
    bn = 1
    while bn * B < keys + 16 * B:
        bn <<= 1
    return bn


def run_plain(cat, stream, runner=None, log=None):
    """Sequential solves; returns (results, per-query seconds, total seconds)."""
    own = runner is None
    runner = runner or PlainRunner(cat)
    res, lat = [], []
    t0 = time.perf_counter()
    for i, (kind, key) in enumerate(stream):
        a = time.perf_counter()
        res.append(runner.solve(kind, key))
        lat.append(time.perf_counter() - a)
        if log and (i + 1) % 250 == 0:
            log(f"[plain] {i + 1}/{len(stream)} queries, {time.perf_counter() - t0:.1f} s")
    total = time.perf_counter() - t0
    if own:
        runner.close()
    return res, np.array(lat), total


# ---------------------------------------------------------------- served path
DEFAULT_TABLE = {"sssp": {"B": 1024, "team": 4}, "astar": {"B": 256, "team": 2}, "knapsack": {"B": 256, "team": 2}}


def config_for(table, kind, key):
    cls = size_class(kind, key)
    t = table.get(f"{cls[0]}:{cls[1]}") or DEFAULT_TABLE[kind]
    return t["B"], t["team"]


def run_served(cat, stream, table, rate, budget, workgroups=0, seed=2, timeout=600.0):
    """Poisson arrivals at `rate` queries/s (rate <= 0: all at time 0).  A
    query is submitted at its arrival time or as soon after as the arena
    admits it.  Returns (results by index, latencies in s, makespan in s)."""
    rng = np.random.default_rng(seed)
    n = len(stream)
    gaps = rng.exponential(1.0 / rate, size=n) if rate > 0 else np.zeros(n)
    arrive_rel = np.cumsum(gaps) - (gaps[0] if n else 0)
    results = [None] * n
    lat = np.zeros(n)
    with S.Server(budget, workgroups=workgroups) as srv:
        t0 = S.now()
        arrive = t0 + arrive_rel
        nxt = 0
        done = 0
        pending = None
        while done < n:
            now = S.now()
            # submit everything that has arrived, as far as the arena admits
            while nxt < n and arrive[nxt] <= now:
                if pending is None:
                    kind, key = stream[nxt]
                    B, team = config_for(table, kind, key)
                    pending = cat.query(kind, key, B, team, tag=nxt)
                if not srv.submit(pending):
                    break
                pending = None
                nxt += 1
            for r in srv.poll(1024):
                i = int(r["tag"])
                results[i] = r
                lat[i] = r["t_done"] - arrive[i]
                done += 1
            if S.now() - t0 > timeout:
                raise RuntimeError(f"served run exceeded {timeout} s with {n - done} queries open")
            if nxt < n and pending is None and arrive[nxt] > S.now():
                time.sleep(min(0.0005, max(0.0, arrive[nxt] - S.now())))
        makespan = S.now() - t0
    return results, lat, makespan


# ---------------------------------------------------------------- autotuner
def _solve_alone(srv, q, timeout=120.0):
    """Submit one query to an idle server and wait for it: its latency (s)
    and its result."""
    while not srv.submit(q):
        srv.poll(64)
        time.sleep(0.0005)
    t0 = S.now()
    while True:
        got = srv.poll(8)
        if got:
            return got[0]["t_done"] - got[0]["t_submit"], got[0]
        if S.now() - t0 > timeout:
            raise RuntimeError("autotune: a query did not complete")
        time.sleep(0.0002)


def autotune(cat, budget, batch_sizes=(256, 1024), teams=(1, 2, 4, 8, 16, 32, 64), workgroups=0, work_slack=1.5,
             reps=2, log=None):
    """For every size class, the batch size and team that solve the class's
    costliest sample query (each sample timed once at B = 1024, team 8) with
    the lowest latency, among the configurations whose work (team x latency,
    in workgroup-seconds) is within `work_slack` of the least.  Large queries
    get the workgroups that shorten them while the workgroups stay busy with
    useful work; that keeps both the stream's tail latency and its makespan
    short.  Each configuration is timed `reps` times, alone on the serve
    kernel (the median counts)."""
    classes = {}
    for kind, key in cat.keys():
        classes.setdefault(size_class(kind, key), []).append((kind, key))
    table = {}
    with S.Server(budget, workgroups=workgroups) as srv:
        for cls, members in sorted(classes.items()):
            kind = members[0][0]
            # the costliest member is the sample
            sample, worst = members[0], -1.0
            if len(members) > 1:
                for k, key in members:
                    t, _ = _solve_alone(srv, cat.query(k, key, 1024, 8, 0))
                    if t > worst:
                        sample, worst = (k, key), t
            trials = []
            for B in batch_sizes:
                for team in teams:
                    ts = []
                    for r in range(reps):
                        t, res = _solve_alone(srv, cat.query(kind, sample[1], B, team, r))
                        if res["status"] != 0:
                            ts = None
                            break
                        ts.append(t)
                    if ts is None:
                        continue
                    t = float(np.median(ts))
                    trials.append((t, team * t, B, team))
                    if log:
                        log(f"[autotune] {cls} {sample[1]} B={B} team={team}: {t * 1e3:.2f} ms")
            least = min(w for _, w, _, _ in trials)
            t, w, B, team = min(x for x in trials if x[1] <= work_slack * least)
            table[f"{cls[0]}:{cls[1]}"] = {"B": B, "team": team, "ms": round(t * 1e3, 3), "wg_ms": round(w * 1e3, 3)}
    return table


def save_table(table, path):
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)


def load_table(path):
    with open(path) as f:
        return json.load(f)
