"""Benchmark of the batched GPU priority queue on one MI355X.

One step = one heap sort of N synthetic keys through the heap: insert all
keys in batch-sized insertion() calls, then deleteRoot/deleteUpdate until the
heap is empty (the reference harness, test/heap_test.cu:135-162).  The keys
are generated on the device before timing and stay resident; every step reads
them and writes the sorted output to a second buffer.

Prints one JSON line (the driver's contract) with keys/s, the roofline of the
dominant kernel and the CPU baseline (std::sort and std::priority_queue on a
bounded sample, timed on this host).

    python bench.py [--steps K] [--warmup W] [--n LOG2|N] [--variant V] [--batch B]
    python bench.py --sweep            # the heap-sort sweep, one JSON line per config
"""
import argparse
import ctypes
import json
import os
import subprocess
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

import gpuheap_amd as G  # noqa: E402
from gpuheap_amd import DeviceBuffer, Heap, batch_num_for  # noqa: E402
from gpuheap_amd._lib import check, lib  # noqa: E402

METRIC = "keys/s insert+delete (2^28 u32) & %HBM peak; batched solves/s + p99 latency"
HBM_PEAK_GBS = 8000.0
KEYS = {"u32": (0, np.uint32, 4), "u64": (1, np.uint64, 8)}


def levels_sum(m):
    """S(M) = sum_{k=1..M} floor(log2 k): tree levels below the root summed
    over the heap sizes 1..M batches."""
    total, lvl, start = 0, 0, 1
    while start <= m:
        end = min(2 * start - 1, m)
        total += lvl * (end - start + 1)
        lvl += 1
        start *= 2
    return total


def alg_bytes(n, b, s):
    """Algorithmic HBM bytes of a heap sort of n keys, batch b, key size s:
    insert reads the input once and reads+writes one batch per level; delete
    writes the output once, reads two children and writes two nodes per level."""
    m = n // b
    sm = levels_sum(m)
    return {"insert": n * s + 2 * b * s * sm, "delete": n * s + 4 * b * s * sm,
            "total": 2 * n * s + 6 * b * s * sm}


def parse_n(v):
    v = int(v)
    return 1 << v if v < 64 else v


def fingerprint(a):
    """Order-independent fingerprint of a key multiset (sum, xor, sum of
    squares, all mod 2^64)."""
    x = a.astype(np.uint64)
    with np.errstate(over="ignore"):
        return (int(np.sum(x, dtype=np.uint64)), int(np.bitwise_xor.reduce(x)),
                int(np.sum(x * x, dtype=np.uint64)))


def make_input(n, key, order="random", seed=0):
    """Device buffer of n keys: SplitMix64 keys generated on the device, in
    generation order or sorted ascending / descending (the adversarial order
    of artifacts/synthetic/sort.cu:115-119)."""
    kt, dtype, ks = KEYS[key]
    d_in = DeviceBuffer(n * ks)
    check(lib().bgpq_gen_keys(kt, seed, n, d_in.ptr, None))
    if order != "random":
        host = np.sort(d_in.to_numpy(dtype, n))
        if order == "descending":
            host = host[::-1].copy()
        d_in.upload(host)
        del host
    G.sync()
    return d_in


def run_config(n, key, variant, batch, gi, gd, steps, warmup, order="random", seed=0, verify=True, d_in=None,
               fp_in=None):
    kt, dtype, ks = KEYS[key]
    own_in = d_in is None
    if own_in:
        d_in = make_input(n, key, order, seed)
    d_out = DeviceBuffer(n * ks)
    h = Heap(batch_num_for(n, batch), batch, variant=variant, key=key)
    for _ in range(warmup):
        h.heapsort(d_in, n, grid_insert=gi, grid_delete=gd, d_out=d_out)
    G.sync()
    ins, dele = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        a, b = h.heapsort(d_in, n, grid_insert=gi, grid_delete=gd, d_out=d_out)
        ins.append(a)
        dele.append(b)
    G.sync()
    wall = time.perf_counter() - t0
    ok = None
    if verify:
        if fp_in is None:
            src = d_in.to_numpy(dtype, n)
            fp_in = fingerprint(src)
            del src
        out = d_out.to_numpy(dtype, n)
        ok = bool(np.all(out[1:] >= out[:-1])) and fp_in == fingerprint(out)
        del out
    h.close()
    if own_in:
        d_in.free()
    d_out.free()
    t_ins, t_del = float(np.mean(ins)), float(np.mean(dele))
    return {"ms_insert": t_ins, "ms_delete": t_del, "ms_step_wall": 1e3 * wall / max(steps, 1),
            "keys_per_s": n / ((t_ins + t_del) / 1e3), "ok": ok}


def cpu_model():
    """The host CPU's model name (what lscpu prints) and its usable threads."""
    name = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return name, len(os.sched_getaffinity(0))


PQ_LOG2 = 26  # std::priority_queue leg: the first 2^26 keys of the array


def start_cpu_baseline(key, n):
    """Start the CPU legs on the same keys as the GPU run (SplitMix64 seed
    0): std::sort of the whole array, median of 5 after a warm-up, and a
    std::priority_queue push-all / pop-all of its first 2^26 keys, median
    of 3.  Each runs single-threaded in a process pinned to a core of its
    own, while the GPU legs run; finish_cpu_baseline collects them."""
    exe = os.path.join(REPO, "oracle", "_build", "cpu_baseline")
    if not os.path.exists(exe):
        subprocess.run(["make", "-C", os.path.join(REPO, "oracle")], check=True, capture_output=True)
    cores = sorted(os.sched_getaffinity(0))
    c_sort, c_pq = cores[-1], cores[-2] if len(cores) > 1 else cores[-1]
    pq_n = min(n, 1 << PQ_LOG2)

    def pinned(core):
        return lambda: os.sched_setaffinity(0, {core})

    sort_p = subprocess.Popen([exe, key, str(n), "0", "5"], stdout=subprocess.PIPE, text=True,
                              preexec_fn=pinned(c_sort))
    pq_p = subprocess.Popen([exe, key, str(pq_n), "0", "0", str(pq_n), "3"], stdout=subprocess.PIPE, text=True,
                            preexec_fn=pinned(c_pq))
    return sort_p, pq_p, n, pq_n


def finish_cpu_baseline(legs, key, gpu_keys_per_s):
    sort_p, pq_p, n, pq_n = legs
    out_s, _ = sort_p.communicate()
    out_q, _ = pq_p.communicate()
    js = json.loads(out_s.strip().splitlines()[-1])
    jq = json.loads(out_q.strip().splitlines()[-1])
    model, threads = cpu_model()
    return {"value": js["std_sort_keys_per_s"], "unit": "keys/s", "cores": 1, "kind": "port",
            "sample": (f"std::sort of the whole {n}-key {key} array the GPU sorts (SplitMix64 seed 0), median of "
                       f"5 timed runs after a warm-up, 1 of {threads} host threads (pinned), while the GPU legs "
                       f"ran; std::priority_queue push-all/pop-all of its first {pq_n} keys, median of 3"),
            "cpu_model": model, "host_threads": threads, "std_sort_s": js["std_sort_s"],
            "pq_keys_per_s": jq["pq_keys_per_s"], "pq_n": pq_n,
            "gpu_over_cpu": gpu_keys_per_s / js["std_sort_keys_per_s"], "ok": bool(js["ok"] and jq["ok"])}


def _orc():
    """The CPU oracle library (oracle/_build/liborc.so), for the baseline leg."""
    path = os.path.join(REPO, "oracle", "_build", "liborc.so")
    if not os.path.exists(path):
        subprocess.run(["make", "-C", os.path.join(REPO, "oracle")], check=True, capture_output=True)
    L = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    L.orc_dijkstra.argtypes = [ctypes.c_int32, vp, vp, vp, ctypes.c_int32, vp, ctypes.POINTER(ctypes.c_int64)]
    L.orc_astar.argtypes = [ctypes.c_int32, ctypes.c_int32, vp, ctypes.POINTER(ctypes.c_uint32),
                            ctypes.POINTER(ctypes.c_int64)]
    L.orc_knapsack_dp.argtypes = [ctypes.c_int32, vp, vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]
    return L


def _vp(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def app_bench(args):
    """One driver solve at the configured size, timed with HIP events inside
    the library, next to the CPU oracle on the same input (one thread)."""
    from gpuheap_amd import apps as A
    B = args.batch
    reps = max(args.steps, 1)
    if args.app == "sssp":
        scale = args.scale
        t0 = time.perf_counter()
        rowptr, col, w = A.rmat_csr(scale, 16, 0)
        print(f"[bench] R-MAT scale {scale}: {len(col)} edges generated in {time.perf_counter() - t0:.1f} s",
              file=sys.stderr, flush=True)
        g = A.Graph(rowptr, col, w)
        h = Heap(batch_num_for(len(col), B), B, variant=args.variant, key="u64")
        for _ in range(args.warmup):
            A.sssp(h, g, 0, grid=args.grid_delete)
        ms = []
        for _ in range(reps):
            dist, st = A.sssp(h, g, 0, grid=args.grid_delete)
            ms.append(st["ms"])
        L = _orc()
        ref = np.empty(len(rowptr) - 1, np.uint32)
        pops = ctypes.c_int64()
        t0 = time.perf_counter()
        L.orc_dijkstra(len(rowptr) - 1, _vp(rowptr), _vp(col), _vp(w), 0, _vp(ref), ctypes.byref(pops))
        cpu_s = time.perf_counter() - t0
        ok = bool(np.array_equal(dist, ref))
        t = float(np.median(ms)) / 1e3
        line = {"metric": "SSSP solve (R-MAT, Dijkstra frontier on the heap)", "value": 1.0 / t, "unit": "solves/s",
                "ms_per_solve": t * 1e3, "edges_per_s": len(col) / t, "n_gpus": 1, "steps": reps,
                "warmup": args.warmup, "higher_is_better": True, "dtype": "u64",
                "config": {"workload": f"SSSP on R-MAT scale {scale} (a,b,c = 0.57,0.19,0.19), degree 16, "
                                       f"weights 1..1000, source 0", "batch_size": B, "variant": args.variant,
                           "grid": args.grid_delete},
                "stats": st, "check": "distances equal CPU Dijkstra: " + ("ok" if ok else "FAILED"),
                "cpu_baseline": {"value": 1.0 / cpu_s, "unit": "solves/s", "cores": 1, "kind": "port",
                                 "sample": f"binary-heap Dijkstra (oracle/apps.c) on the same graph, one solve, "
                                           f"{cpu_s:.2f} s"}}
    elif args.app == "astar":
        size = args.size
        free = A.grid_map(size, size, 0.25, 0)
        gm = A.GridMap(free, size, size)
        h = Heap(batch_num_for(4 * size * size, B), B, variant=args.variant, key="u64")
        for _ in range(args.warmup):
            A.astar(h, gm, grid=args.grid_delete)
        ms = []
        for _ in range(reps):
            cost, st = A.astar(h, gm, grid=args.grid_delete)
            ms.append(st["ms"])
        L = _orc()
        c = ctypes.c_uint32()
        exp = ctypes.c_int64()
        t0 = time.perf_counter()
        L.orc_astar(size, size, _vp(free), ctypes.byref(c), ctypes.byref(exp))
        cpu_s = time.perf_counter() - t0
        ok = cost == c.value
        t = float(np.median(ms)) / 1e3
        line = {"metric": "A* solve (8-connected grid, octile heuristic)", "value": 1.0 / t, "unit": "solves/s",
                "ms_per_solve": t * 1e3, "optimal_cost": cost, "cpu_expanded": exp.value, "n_gpus": 1,
                "steps": reps, "warmup": args.warmup, "higher_is_better": True, "dtype": "u64",
                "config": {"workload": f"A* on a {size}x{size} grid, obstacles p=0.25 seed 0, costs 10/14",
                           "batch_size": B, "variant": args.variant, "grid": args.grid_delete},
                "stats": st, "check": "cost equals CPU A*: " + ("ok" if ok else "FAILED"),
                "cpu_baseline": {"value": 1.0 / cpu_s, "unit": "solves/s", "cores": 1, "kind": "port",
                                 "sample": f"sequential A* (oracle/apps.c) on the same grid, one solve, "
                                           f"{cpu_s:.2f} s"}}
    else:
        n = args.items
        w, v, cap = A.knapsack_instance(n, 0)
        h = Heap(1 << 12, B, variant=args.variant, key="u64")
        for _ in range(args.warmup):
            A.knapsack(h, w, v, cap, grid=args.grid_delete)
        ms = []
        for _ in range(reps):
            best, st = A.knapsack(h, w, v, cap, grid=args.grid_delete)
            ms.append(st["ms"])
        L = _orc()
        ref = ctypes.c_uint64()
        t0 = time.perf_counter()
        L.orc_knapsack_dp(n, _vp(w), _vp(v), cap, ctypes.byref(ref))
        cpu_s = time.perf_counter() - t0
        ok = best == ref.value
        t = float(np.median(ms)) / 1e3
        line = {"metric": "0/1 knapsack branch-and-bound solve", "value": 1.0 / t, "unit": "solves/s",
                "ms_per_solve": t * 1e3, "optimum": best, "n_gpus": 1, "steps": reps, "warmup": args.warmup,
                "higher_is_better": True, "dtype": "u64",
                "config": {"workload": f"knapsack n={n}, weights 1..1000, values = weight + 100, capacity half",
                           "batch_size": B, "variant": args.variant, "grid": args.grid_delete},
                "stats": st, "check": "optimum equals CPU DP: " + ("ok" if ok else "FAILED"),
                "cpu_baseline": {"value": 1.0 / cpu_s, "unit": "solves/s", "cores": 1, "kind": "port",
                                 "sample": f"O(n*C) dynamic program (oracle/apps.c), one solve, {cpu_s * 1e3:.1f} ms"}}
    print(json.dumps(line), flush=True)
    return 0 if ok else 1


def mixed_run(queries, max_scale, max_side, budget_gib, rate=0.0, retune=False, table_out=None, cpu=True,
              stratified=False, catalog=None):
    """The mixed SSSP / A* / knapsack stream: the plain one-heap-per-launch
    path, then the persistent serve kernel closed-loop (everything arrives at
    once: solves/s) and at a Poisson rate (p50 / p99 latency).  Every served
    result must equal the plain path's, and (cpu) the CPU oracles'."""
    from gpuheap_amd import mixed as M
    t0 = time.perf_counter()
    cat = catalog or M.Catalog(scales=tuple(range(10, max_scale + 1)),
                               sides=tuple(x for x in M.ASTAR_SIDES if x <= max_side),
                               grid_seeds=4, sources_per_graph=8, knap_instances=128)
    stream = M.make_stream(cat, queries, stratified=stratified)
    print(f"[bench] catalog and a {len(stream)}-query stream ready in {time.perf_counter() - t0:.1f} s",
          file=sys.stderr, flush=True)
    budget = int(budget_gib * (1 << 30))
    tpath = os.path.join(REPO, "profiles", "autotune_mixed.json")
    if os.path.exists(tpath) and not retune:
        table = M.load_table(tpath)
    else:
        table = M.autotune(cat, budget, log=lambda m: print(m, file=sys.stderr, flush=True))
        if retune:
            M.save_table(table, table_out or tpath)
    plain, plain_lat, plain_s = M.run_plain(cat, stream, log=lambda m: print(m, file=sys.stderr, flush=True))
    print(f"[bench] plain path: {len(stream) / plain_s:.1f} q/s", file=sys.stderr, flush=True)
    served, _, makespan = M.run_served(cat, stream, table, rate=0.0, budget=budget)
    qps = len(stream) / makespan
    print(f"[bench] served closed-loop: {qps:.1f} q/s", file=sys.stderr, flush=True)
    rate = rate if rate > 0 else 0.5 * qps
    served_r, lat, _ = M.run_served(cat, stream, table, rate=rate, budget=budget)
    same = all(M.answer_of(k, key, a) == M.answer_of(k, key, b) and b["status"] == 0
               for (k, key), a, b in zip(stream, plain, served))
    same = same and all(M.answer_of(k, key, a) == M.answer_of(k, key, b) and b["status"] == 0
                        for (k, key), a, b in zip(stream, plain, served_r))
    kinds = {k: sum(1 for x, _ in stream if x == k) for k in ("sssp", "astar", "knapsack")}
    res = {"queries": len(stream), "mix": kinds, "distinct_problems": len(set(stream)),
           "served_solves_per_s": qps,
           "served_at_rate": {"rate_qps": rate, "p50_ms": float(np.percentile(lat, 50) * 1e3),
                              "p99_ms": float(np.percentile(lat, 99) * 1e3)},
           "plain_one_heap_per_launch": {"solves_per_s": len(stream) / plain_s,
                                         "p50_ms": float(np.percentile(plain_lat, 50) * 1e3),
                                         "p99_ms": float(np.percentile(plain_lat, 99) * 1e3)},
           "speedup_vs_plain": qps / (len(stream) / plain_s),
           "workload": f"{len(stream)} queries: 40% SSSP on R-MAT 2^10..2^{max_scale}, 40% A* on "
                       f"256^2..{max_side}^2 grids (p=0.25), 20% knapsack n=200",
           "hbm_budget_gib": budget_gib, "size_class_table": table,
           "check": "served results equal the plain path's: " + ("ok" if same else "FAILED")}
    if cpu:
        res["cpu_baseline"], ok = mixed_cpu_baseline(cat, stream, plain)
        res["check"] += "; plain results equal the CPU oracle: " + ("ok" if ok else "FAILED")
        same = same and ok
    res["ok"] = same
    return res


def mixed_bench(args):
    """--app mixed: the mixed stream alone, as its own JSON line."""
    r = mixed_run(args.queries, args.max_scale, args.max_side, args.budget_gib, rate=args.rate, retune=args.retune,
                  table_out=args.table_out, cpu=not args.no_cpu)
    line = {"metric": "mixed SSSP/A*/knapsack stream: solves/s (persistent serve kernel) and p99 latency",
            "value": r["served_solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": 1, "warmup": 0,
            "ms_per_step": 1e3 * r["queries"] / r["served_solves_per_s"], "higher_is_better": True, "dtype": "u64",
            "config": {"workload": r["workload"], "hbm_budget_gib": r["hbm_budget_gib"],
                       "size_class_table": r["size_class_table"]},
            "plain_one_heap_per_launch": r["plain_one_heap_per_launch"], "speedup_vs_plain": r["speedup_vs_plain"],
            "served_at_rate": r["served_at_rate"], "check": r["check"]}
    if "cpu_baseline" in r:
        line["cpu_baseline"] = r["cpu_baseline"]
    print(json.dumps(line), flush=True)
    return 0 if r["ok"] else 1


def mixed_cpu_baseline(cat, stream, plain):
    """CPU oracle on every distinct problem of the stream (one thread): its
    time, scaled to the stream, is the CPU baseline, and its answers check
    the GPU results."""
    from gpuheap_amd import serve as S
    L = _orc()
    answers, secs = {}, {}
    for kind, key in sorted(set(stream), key=str):
        t0 = time.perf_counter()
        if kind == "sssp":
            rowptr, col, w = cat.graphs[key[0]]["csr"]
            dist = np.empty(len(rowptr) - 1, np.uint32)
            pops = ctypes.c_int64()
            L.orc_dijkstra(len(rowptr) - 1, _vp(rowptr), _vp(col), _vp(w), key[1], _vp(dist), ctypes.byref(pops))
            ans = (int(np.count_nonzero(dist != 0xFFFFFFFF)), S.sssp_digest(dist))
        elif kind == "astar":
            g = cat.grids[key]
            c = ctypes.c_uint32()
            exp = ctypes.c_int64()
            side = g["dev"].width
            L.orc_astar(side, side, _vp(g["free"]), ctypes.byref(c), ctypes.byref(exp))
            ans = (c.value, 0)
        else:
            w, v, cap = cat.knaps[key]
            best = ctypes.c_uint64()
            L.orc_knapsack_dp(len(w), _vp(w), _vp(v), cap, ctypes.byref(best))
            ans = (best.value, 0)
        secs[(kind, key)] = time.perf_counter() - t0
        answers[(kind, key)] = ans
    from gpuheap_amd import mixed as M
    ok = all(M.answer_of(k, key, r) == answers[(k, key)] for (k, key), r in zip(stream, plain))
    total = sum(secs[q] for q in stream)
    return {"value": len(stream) / total, "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": f"{len(answers)} distinct problems solved once each (Dijkstra, A*, knapsack DP; "
                      f"oracle/apps.c), time scaled to the {len(stream)}-query stream"}, ok


def prims_bench(B, nb):
    """The batch primitives alone, for counter collection under rocprofv3:
    the group merge, the two half merges of a sift step and the low-register
    merge path (bgpq_test_merge algo 2, 4, 5), then the pad-and-sort of the
    insert path (bgpq_test_sort), each on nb random u32 batches, one
    workgroup per batch; every result is checked against numpy."""
    rng = np.random.default_rng(1)
    a = np.sort(rng.integers(0, 2**32 - 1, size=(nb, B), dtype=np.uint32), axis=1)
    b = np.sort(rng.integers(0, 2**32 - 1, size=(nb, B), dtype=np.uint32), axis=1)
    da, db = DeviceBuffer.from_numpy(a), DeviceBuffer.from_numpy(b)
    small, large = DeviceBuffer(a.nbytes), DeviceBuffer(a.nbytes)
    ok = True
    u = np.sort(np.concatenate([a, b], axis=1), axis=1)
    for algo in (2, 4, 5):
        check(lib().bgpq_test_merge(0, B, da.ptr, db.ptr, small.ptr, large.ptr, nb, algo))
        G.sync()
        ok = ok and np.array_equal(small.to_numpy(np.uint32).reshape(nb, B), u[:, :B])
        ok = ok and np.array_equal(large.to_numpy(np.uint32).reshape(nb, B), u[:, B:])
    sizes = DeviceBuffer.from_numpy(np.full(nb, B, dtype=np.int32))
    check(lib().bgpq_test_sort(0, B, da.ptr, sizes.ptr, small.ptr, 0xFFFFFFFF, nb))
    G.sync()
    ok = ok and np.array_equal(small.to_numpy(np.uint32).reshape(nb, B), a)
    print(json.dumps({"prims": {"batch_size": B, "batches": nb, "merges": [2, 4, 5], "sort": True,
                                "check": "ok" if ok else "FAILED"}}), flush=True)
    return 0 if ok else 1


def traffic_for(workload):
    """HBM bytes per launch of the dominant kernel, from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (profiles/hbm_traffic.json)."""
    p = os.path.join(REPO, "profiles", "hbm_traffic.json")
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            return json.load(f).get(workload)
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=parse_n, default=1 << 28, help="keys (or log2 of keys if < 64)")
    ap.add_argument("--key", choices=sorted(KEYS), default="u32")
    ap.add_argument("--variant", choices=sorted(G.VARIANTS), default="buitdd_np")
    ap.add_argument("--batch", type=int, default=None,
                    help="batch size (heap sort: 8192; --app: sssp 1024, astar 256, knapsack 256)")
    ap.add_argument("--grid-insert", type=int, default=256)
    ap.add_argument("--grid-delete", type=int, default=None,
                    help="delete workgroups (heap sort: 64) or driver workgroups (--app: sssp 32, astar 4, knapsack 8)")
    ap.add_argument("--order", choices=["random", "ascending", "descending"], default="random")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sweep", action="store_true")
    ap.add_argument("--sweep-sets", default="u32:random,u32:descending,u64:random",
                    help="key sets of the sweep, key:order pairs")
    ap.add_argument("--calibrate", action="store_true",
                    help="also write and read back 1 GiB with plain kernels (HBM counter calibration under rocprofv3)")
    ap.add_argument("--queries", type=int, default=2000, help="--app mixed: queries in the stream")
    ap.add_argument("--line-queries", type=int, default=1000,
                    help="queries of the stratified mixed stream in the default line (0: leave it out)")
    ap.add_argument("--extra-batches", default="1024,2048",
                    help="batch sizes also timed (1 step) in the default line ('' for none)")
    ap.add_argument("--max-scale", type=int, default=20, help="mixed: largest SSSP scale")
    ap.add_argument("--max-side", type=int, default=2048, help="mixed: largest A* grid side")
    ap.add_argument("--budget-gib", type=float, default=64.0, help="mixed: HBM budget of the serve arena")
    ap.add_argument("--rate", type=float, default=0.0, help="mixed: Poisson rate (0: half the closed-loop rate)")
    ap.add_argument("--retune", action="store_true", help="mixed: run the autotuner instead of the committed table")
    ap.add_argument("--table-out", default=None,
                    help="mixed --retune: where to write the new table (default: profiles/autotune_mixed.json)")
    ap.add_argument("--prims", type=int, nargs=2, metavar=("B", "BATCHES"),
                    help="run the batch primitives alone (merges, sort) for counter collection")
    ap.add_argument("--app", choices=["sssp", "astar", "knapsack", "mixed"],
                    help="time one driver solve instead of the heap sort (--batch, --variant, --grid-delete apply)")
    ap.add_argument("--scale", type=int, default=22, help="SSSP: log2 of the vertex count")
    ap.add_argument("--size", type=int, default=4096, help="A*: grid side")
    ap.add_argument("--items", type=int, default=200, help="knapsack: items")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(a multi-GPU heap sort would route keys to per-GPU heaps by sampled splitters "
              "with one all-to-all, and the mixed stream would split its queries)", file=sys.stderr)
        sys.exit(2)
    print(f"[bench] device: {G.device_name()}", file=sys.stderr, flush=True)
    if args.prims:
        sys.exit(prims_bench(*args.prims))
    if args.app == "mixed":
        sys.exit(mixed_bench(args))
    if args.app:
        # per-driver defaults: the fastest measured configuration of each solve
        # (R-MAT scale 22, 4096^2 grid, n = 200) on one MI355X
        # measured best (batch size, workgroups) per driver: fewer workgroups
        # pop fewer keys out of order, hence less stale and repeated work
        dflt = {"sssp": (1024, 32), "astar": (256, 4), "knapsack": (256, 8)}.get(args.app, (1024, 64))
        args.batch = args.batch or dflt[0]
        args.grid_delete = args.grid_delete or dflt[1]
        if args.variant not in ("buitdd", "tditdd"):
            args.variant = "buitdd"
        if args.batch > 2048:
            args.batch = 1024
        sys.exit(app_bench(args))
    args.batch = args.batch or 8192
    args.grid_delete = args.grid_delete or 64

    if args.sweep:
        # the heap-sort sweep: every variant and batch size on three key sets,
        # each generated once and reused (u64 keys stop at B = 4096)
        for key, order in (tuple(x.split(":")) for x in args.sweep_sets.split(",")):
            d_in = make_input(args.n, key, order)
            fp_in = fingerprint(d_in.to_numpy(KEYS[key][1], args.n))
            for variant in ("buitdd", "buitdd_np", "tditdd", "tditdd_np"):
                for b in (256, 512, 1024, 2048, 4096, 8192):
                    if key == "u64" and b > 4096:
                        continue
                    r = run_config(args.n, key, variant, b, args.grid_insert, args.grid_delete, 1, 0, order=order,
                                   d_in=d_in, fp_in=fp_in)
                    ab = alg_bytes(args.n, b, KEYS[key][2])
                    r.update({"n": args.n, "key": key, "order": order, "variant": variant, "batch": b,
                              "hbm_frac_total": ab["total"] / ((r["ms_insert"] + r["ms_delete"]) / 1e3) / 8e12})
                    print(json.dumps(r), flush=True)
            d_in.free()
        return

    print(f"[bench] heap sort n={args.n} {args.key} {args.variant} B={args.batch} "
          f"grids {args.grid_insert}/{args.grid_delete}", file=sys.stderr, flush=True)
    cpu_legs = None
    if not args.no_cpu:
        # the CPU legs run on cores of their own while the GPU legs below run
        print("[bench] CPU baseline (std::sort, std::priority_queue) started", file=sys.stderr, flush=True)
        cpu_legs = start_cpu_baseline(args.key, args.n)
    d_in = make_input(args.n, args.key, args.order)
    fp_in = fingerprint(d_in.to_numpy(KEYS[args.key][1], args.n))
    r = run_config(args.n, args.key, args.variant, args.batch, args.grid_insert, args.grid_delete, args.steps,
                   args.warmup, order=args.order, d_in=d_in, fp_in=fp_in)
    if not r["ok"]:
        print("[bench] output check FAILED", file=sys.stderr, flush=True)
    s = KEYS[args.key][2]
    ab = alg_bytes(args.n, args.batch, s)
    dom = "delete" if r["ms_delete"] >= r["ms_insert"] else "insert"
    dom_ms = r["ms_" + dom]
    achieved = ab[dom] / (dom_ms / 1e3) / 1e9
    workload = f"heapsort_{args.key}_n{args.n}_{args.variant}_b{args.batch}_{args.order}"
    measured = traffic_for(workload)
    roofline = {"bound": "hbm", "kernel": f"k_{dom}", "achieved": achieved, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                "traffic": measured.get(f"k_{dom}") if measured else None,
                "alg_bytes_per_launch": ab[dom]}
    ok = bool(r["ok"])
    line = {
        "metric": METRIC,
        "value": r["keys_per_s"],
        "unit": "keys/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_step_wall"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.key,
        "data": f"synthetic: SplitMix64 seed 0 {args.key} keys, {args.order} order, generated on device",
        "config": {"workload": f"heap sort (insert all, delete until empty) of {args.n} {args.key} keys",
                   "n_keys": args.n, "key": args.key, "variant": args.variant, "batch_size": args.batch,
                   "grid_insert": args.grid_insert, "grid_delete": args.grid_delete, "order": args.order},
        "phases_ms": {"insert": r["ms_insert"], "delete": r["ms_delete"]},
        "hbm_frac_total": ab["total"] / ((r["ms_insert"] + r["ms_delete"]) / 1e3) / (HBM_PEAK_GBS * 1e9),
        "check": "output sorted and multiset fingerprint equal to input: " + ("ok" if r["ok"] else "FAILED"),
        "roofline": roofline,
    }
    # the configuration's own batch sizes (its sweep stops at 2048), same keys
    extras = {}
    for b in (int(x) for x in args.extra_batches.split(",") if x.strip()):
        if b == args.batch:
            continue
        print(f"[bench] heap sort at B={b}", file=sys.stderr, flush=True)
        e = run_config(args.n, args.key, args.variant, b, args.grid_insert, args.grid_delete, 1, 1,
                       order=args.order, d_in=d_in, fp_in=fp_in)
        ok = ok and bool(e["ok"])
        ab_b = alg_bytes(args.n, b, s)
        extras[str(b)] = {"keys_per_s": e["keys_per_s"], "ms_insert": e["ms_insert"], "ms_delete": e["ms_delete"],
                          "hbm_frac_total": ab_b["total"] / ((e["ms_insert"] + e["ms_delete"]) / 1e3) /
                          (HBM_PEAK_GBS * 1e9), "check": "ok" if e["ok"] else "FAILED"}
    if extras:
        line["other_batch_sizes"] = extras
    d_in.free()
    if args.line_queries > 0:
        # the metric's second half: batched / served solves/s and p99 latency
        print(f"[bench] mixed stream of {args.line_queries} queries", file=sys.stderr, flush=True)
        m = mixed_run(args.line_queries, args.max_scale, args.max_side, args.budget_gib, cpu=not args.no_cpu,
                      stratified=True)
        ok = ok and m.pop("ok")
        m.pop("size_class_table", None)
        line["mixed"] = m
    if args.calibrate:
        # kernels of known traffic to scale the FETCH_SIZE / WRITE_SIZE
        # counters: the key generator writes 1 GiB (k_gen_u32), the digest
        # kernel reads the same 1 GiB back (k_digest)
        nb = 1 << 30
        cal = DeviceBuffer(nb)
        dig = DeviceBuffer(16)
        check(lib().bgpq_gen_keys(0, 1, nb // 4, cal.ptr, None))
        check(lib().bgpq_label_digest(cal.ptr, nb // 4, dig.ptr, None))
        G.sync()
        cal.free()
        dig.free()
        line["calibration_bytes"] = nb
    if cpu_legs is not None:
        print("[bench] waiting for the CPU baseline", file=sys.stderr, flush=True)
        line["cpu_baseline"] = finish_cpu_baseline(cpu_legs, args.key, r["keys_per_s"])
    print(json.dumps(line), flush=True)
    if not ok:
        sys.exit(1)


if __name__ == "__main__":
    main()
