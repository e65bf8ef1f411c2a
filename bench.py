"""Benchmark: exact max-flow queries/sec on one MI355X, batched vs plain.

Workload (default): R-MAT scale 22 (a,b,c,d = .57,.19,.19,.05; 16*2^22
samples, duplicates merged by summing log-uniform [1,1e6] capacities, graph
seed 3001), --steps x --batch queries with s from the top 1% of vertices by
out-degree and t uniform over vertices with in-degree > 0 (query seed 3002).
The timed region submits them all at once to --engines engines sharing the
--batch slots (each engine on its own host thread and HIP stream, through
mf_serve_lanes), which refill a slot as soon as its query finishes
(continuous batching); ms_per_step is the wall time / --steps.
Every GPU flow carries the GPU's min-cut certificate; a sample is also
compared bit for bit with the CPU oracle (Dinic), whose time on the same
sample is the cpu_baseline.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import mscresearch_amd as M  # noqa: E402

METRIC = "max-flow queries/sec + p50/p99 latency, served/batched vs plain, bit-exact"
HBM_PEAK = 8.0e12
# algorithmic bytes per unit: edge scan, vertex scan, push, relabel/discovery
UNIT_BYTES = np.array([16.0, 16.0, 64.0, 8.0])
FAMILIES = ["init", "bfs", "active", "discharge", "rm", "cert", "bfs_bu"]


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def pct(v, p):
    v = sorted(v)
    if not v:
        return 0.0
    k = max(1, min(len(v), int(np.ceil(p * len(v)))))
    return float(v[k - 1])


def build_workload(args):
    t0 = time.time()
    if args.workload == "rmat":
        e = M.Edges.rmat(args.scale, 10**6, args.graph_seed, threads=min(16, os.cpu_count() or 4))
        qs, qt = e.hub_queries(args.queries + args.batch * args.warmup, args.query_seed)
        name = f"rmat{args.scale}_hub_queries"
        desc = (f"R-MAT scale {args.scale} (a=.57,b=.19,c=.19), 16*2^{args.scale} samples, duplicates merged, "
                f"caps log-uniform [1,1e6]; s in top 1% out-degree, t in-degree>0")
    elif args.workload == "grid":
        side = args.grid
        e = M.Edges.grid(side, side, 1000, args.graph_seed)
        qs, qt = M.grid_queries(side, side, args.queries + args.batch * args.warmup, args.query_seed)
        name = f"grid{side}x{side}_left_right"
        desc = f"{side}x{side} 4-neighbour grid, both directions, caps U[1,1000]; s col 0, t col {side - 1}"
    else:
        e = M.Edges.uniform(10000, 100000, 100, args.graph_seed)
        qs, qt = M.uniform_queries(10000, args.queries + args.batch * args.warmup, args.query_seed)
        name = "uniform_n10000_m100000"
        desc = "uniform n=1e4 m=1e5 caps U[1,100]"
    n, cnt = e.info()
    log(f"generated {name}: n={n} lines={cnt} in {time.time() - t0:.1f}s")
    t0 = time.time()
    hg = M.HostGraph.from_edgelist(e).partition(args.R)
    log(f"CSR + partition R={args.R} in {time.time() - t0:.1f}s")
    d = hg.export()
    nid = d["new_id"]
    qs2 = nid[qs].astype(np.int32)
    qt2 = nid[qt].astype(np.int32)
    return e, hg, d, qs2, qt2, name, desc


def run_engines(engs, qs, qt, lo, hi):
    """Queries lo..hi-1 over several engines, query i on engine (i - lo) % k,
    all arriving at once, through mf_serve_lanes (one host thread and HIP
    stream per engine, each keeping its slots full)."""
    k = len(engs)
    idx = np.arange(lo, hi)
    lane = (idx - lo) % k
    gids = np.array([engs[j][1] for j in lane], np.int32)
    res_all = M.serve_lanes([e for e, _ in engs], lane, gids, np.asarray(qs)[idx], np.asarray(qt)[idx],
                            np.zeros(len(idx)))
    lat, svc, flows = [], [], {}
    for r in res_all:
        if r.status != 0 or r.cert_ok != 1:
            raise RuntimeError(f"query {lo + r.tag}: status {r.status} cert {r.cert_ok}: {M.lib().mf_last_error().decode()}")
        lat.append(r.t_end - r.t_arrive)
        svc.append(r.t_end - r.t_start)
        flows[lo + int(r.tag)] = r.flow
    return res_all, lat, svc, flows


def cpu_baseline(d, qs, qt, target_s, threads):
    """Oracle Dinic on a bounded sample of the same queries (CPU baseline)."""
    import oracle as O

    og = O.Graph(len(d["new_id"]), d["row_ptr"], d["col"], d["cap"], d["rev"])
    t0 = time.time()
    f0, sec0 = O.maxflow_many(og, qs[:1], qt[:1], threads=1)
    per = max(sec0, 1e-3)
    k = int(max(threads, min(len(qs), threads * max(1, int(target_s / per)))))
    k = min(k, len(qs))
    flows, sec = O.maxflow_many(og, qs[:k], qt[:k], threads=threads)
    log(f"cpu oracle: 1 query {sec0:.2f}s single-thread; {k} queries on {threads} threads in {sec:.2f}s")
    return flows, sec, k, sec0


def serve_bench(dg22, hg22, e22, args):
    """Serving loop: grid 1024x1024, R-MAT 20 and R-MAT 22 resident, one engine
    lane each (own host thread and HIP stream, slab arena sized in slabs of
    its graph); Poisson arrivals at each offered load, graph chosen 0.3 /
    0.4 / 0.3; each lane admits by slab bytes and batches continuously."""
    t0 = time.time()
    grid = M.Edges.grid(1024, 1024, 1000, 2001)
    hgg = M.HostGraph.from_edgelist(grid).partition(512)
    dgg = M.DeviceGraph(hgg)
    e20 = M.Edges.rmat(20, 10**6, 4001, threads=min(16, os.cpu_count() or 4))
    hg20 = M.HostGraph.from_edgelist(e20).partition(512)
    dg20 = M.DeviceGraph(hg20)
    log(f"serving graphs ready in {time.time() - t0:.1f}s")
    lanes = [("grid1024", dgg, hgg, args.serve_grid_slots, M.default_params(fuse=32, small_wgs=args.serve_grid_wgs)),
             ("rmat20", dg20, hg20, 32, M.default_params()),
             ("rmat22", dg22, hg22, 48, M.default_params())]
    engines, budgets = [], []
    for name, dg, hg, slots, prm in lanes:
        eng = M.Engine(slots, prm)
        eng.add_graph(dg)
        slab = dg.slab_bytes()
        nslab = args.serve_slabs if name == "rmat22" else slots
        eng.set_budget(int(nslab * slab))
        engines.append(eng)
        budgets.append(int(nslab * slab))
    nid = [hgg.export()["new_id"], hg20.export()["new_id"], hg22.export()["new_id"]]
    nmax = 4096
    pools = [M.grid_queries(1024, 1024, nmax, 2002), e20.hub_queries(nmax, 4002), e22.hub_queries(nmax, 4003)]
    pools = [(nid[k][a].astype(np.int32), nid[k][b].astype(np.int32)) for k, (a, b) in enumerate(pools)]
    rng = np.random.default_rng(4004)
    # warm: reachability classes and pruned subgraphs of the first queries
    wq = [(k, j) for k in range(3) for j in range(4)]
    M.serve_lanes(engines, np.array([k for k, _ in wq]), np.zeros(len(wq)), np.array([pools[k][0][j] for k, j in wq]),
                  np.array([pools[k][1][j] for k, j in wq]), np.zeros(len(wq)))
    levels = []
    used = [4, 4, 4]
    for qps in args.serve_qps:
        nq = int(min(args.serve_max, max(20, qps * args.serve_seconds)))
        g = rng.choice(3, nq, p=[0.3, 0.4, 0.3])
        arrive = np.cumsum(rng.exponential(1.0 / qps, nq))
        s, t = np.empty(nq, np.int32), np.empty(nq, np.int32)
        for i in range(nq):
            k = g[i]
            j = used[k] % nmax
            used[k] += 1
            s[i], t[i] = pools[k][0][j], pools[k][1][j]
        before = [e.mem_stats()["admit_waits"] for e in engines]
        res = M.serve_lanes(engines, g, np.zeros(nq), s, t, arrive)
        lat = [r.t_end - r.t_arrive for r in res]
        span = max(r.t_end for r in res) - min(r.t_arrive for r in res)
        bad = sum(1 for r in res if r.status != 0 or r.cert_ok != 1)
        per = {}
        for k, (nm, *_rest) in enumerate(lanes):
            lk = [r.t_end - r.t_arrive for r in res if g[r.tag] == k]
            per[nm] = {"n": len(lk), "p50_ms": round(pct(lk, .5) * 1e3, 3), "p99_ms": round(pct(lk, .99) * 1e3, 3),
                       "admission_waits": engines[k].mem_stats()["admit_waits"] - before[k]}
        lv = {"offered_qps": qps, "queries": nq, "sustained_qps": round(nq / span, 2),
              "p50_ms": round(pct(lat, .5) * 1e3, 3), "p99_ms": round(pct(lat, .99) * 1e3, 3),
              "failed_or_uncertified": bad, "per_graph": per}
        log(f"serve {qps} q/s offered: {lv}")
        levels.append(lv)
    return {"graphs": ["grid 1024x1024 caps U[1,1000] seed 2001", "R-MAT 20 seed 4001", "R-MAT 22 (the main workload)"],
            "mix": [0.3, 0.4, 0.3], "lanes": [{"graph": nm, "slots": sl, "slab_budget_bytes": b,
                                             "slab_peak_bytes": e.mem_stats()["peak"]}
                                            for (nm, _d, _h, sl, _p), b, e in zip(lanes, budgets, engines)],
            "levels": levels}


def dynamic_bench(args):
    """Dynamic mode with batched re-solve: R-MAT scale --scale (23), 32 hub
    queries pinned to slots, --rounds rounds each changing a uniformly chosen
    1% of the input edges by a factor U[0.5, 2] (rounded, floored at 1); each
    round re-solves all pinned queries warm (from their residuals, repaired
    by the updates) and cold (static re-solve on the updated graph); the
    first --oracle-rounds rounds are also checked against the CPU oracle."""
    t0 = time.time()
    e = M.Edges.rmat(args.scale, 10**6, 5001, threads=min(16, os.cpu_count() or 4))
    hg = M.HostGraph.from_edgelist(e).partition(args.R)
    d = hg.export()
    n = len(d["new_id"])
    cap = d["cap"].copy()
    fwd = np.nonzero(cap > 0)[0].astype(np.int32)
    qs, qt = e.hub_queries(32, 5002)
    q = [(int(d["new_id"][a]), int(d["new_id"][b])) for a, b in zip(qs, qt)]
    dg = M.DeviceGraph(hg)
    log(f"R-MAT {args.scale}: n={n} E={len(cap)} ready in {time.time() - t0:.1f}s")
    warm = M.Engine(32, M.default_params(prune=0))
    gid = warm.add_graph(dg)
    cold = M.Engine(8, M.default_params(prune=0))
    cgid = cold.add_graph(dg)
    for i, (s, t) in enumerate(q):
        warm.pin(i, gid, s, t, i)
    base = {r.tag: r for r in warm.solve_pinned(warm=False)}
    rng = np.random.default_rng(5003)
    k = max(1, len(fwd) // 100)
    rounds = []
    tw_all = tc_all = 0.0
    mism_static = mism_oracle = 0
    checked = 0
    threads = max(1, min(16, os.cpu_count() or 1))
    for rnd in range(args.rounds):
        pick = rng.choice(fwd, k, replace=False)
        new = np.maximum(1, np.rint(cap[pick] * rng.uniform(0.5, 2.0, k))).astype(np.int64)
        t1 = time.perf_counter()
        warm.update_caps(gid, pick, new)
        tu = time.perf_counter() - t1
        cap[pick] = new
        t1 = time.perf_counter()
        wres = {r.tag: r for r in warm.solve_pinned(warm=True)}
        tw = time.perf_counter() - t1
        for i, (s, t) in enumerate(q):
            cold.submit(cgid, s, t, i)
        t1 = time.perf_counter()
        cres = {r.tag: r for r in cold.run()}
        tc = time.perf_counter() - t1
        tw_all += tw
        tc_all += tc
        bad = sum(1 for i in range(32) if wres[i].flow != cres[i].flow or wres[i].cert_ok != 1 or cres[i].cert_ok != 1)
        mism_static += bad
        line = {"round": rnd + 1, "updates": k, "update_ms": round(tu * 1e3, 2), "warm_ms": round(tw * 1e3, 2),
                "static_ms": round(tc * 1e3, 2), "warm_vs_static_mismatches": bad,
                "drains": int(sum(wres[i].drains for i in range(32)))}
        if rnd < args.oracle_rounds:
            import oracle as O

            og = O.Graph(n, d["row_ptr"], d["col"], cap, d["rev"])
            idx = list(range(0, 32, max(1, 32 // args.oracle_sample)))[:args.oracle_sample]
            flows, _ = O.maxflow_many(og, np.array([q[i][0] for i in idx], np.int32),
                                      np.array([q[i][1] for i in idx], np.int32), threads=threads)
            om = sum(int(flows[j] != wres[i].flow) for j, i in enumerate(idx))
            mism_oracle += om
            checked += len(idx)
            line["oracle_checked"] = len(idx)
            line["oracle_mismatches"] = om
        log(f"dynamic round {line}")
        rounds.append(line)
    out = {"metric": METRIC, "value": round(32 * args.rounds / tw_all, 3), "unit": "queries/s",
           "n_gpus": 1, "steps": args.rounds, "warmup": 0, "ms_per_step": round(tw_all / args.rounds * 1e3, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int64", "data": "synthetic",
           "config": {"workload": f"rmat{args.scale}_dynamic_32_pinned", "n": n, "csr_entries": int(len(cap)),
                      "updates_per_round": k, "rounds": args.rounds, "parallelism": "one GPU, 32 pinned slots"},
           "static_resolve_qps": round(32 * args.rounds / tc_all, 3),
           "warm_vs_static_speedup": round(tc_all / tw_all, 3),
           "warm_vs_static_mismatches": mism_static, "oracle_checked": checked, "oracle_mismatches": mism_oracle,
           "base_certified": sum(int(r.cert_ok) for r in base.values()), "rounds": rounds}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16, help="timed steps of --batch queries each")
    ap.add_argument("--warmup", type=int, default=1, help="untimed steps (distinct queries)")
    ap.add_argument("--batch", type=int, default=192, help="engine slots = queries per step")
    ap.add_argument("--engines", type=int, default=2,
                    help="engines sharing the batch slots, one host thread and HIP stream each")
    ap.add_argument("--workload", choices=["rmat", "grid", "uniform", "dynamic"], default="rmat")
    ap.add_argument("--rounds", type=int, default=20, help="dynamic: update rounds")
    ap.add_argument("--oracle-rounds", type=int, default=2, help="dynamic: rounds also checked on the CPU oracle")
    ap.add_argument("--oracle-sample", type=int, default=4, help="dynamic: queries per checked round")
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--grid", type=int, default=1024)
    ap.add_argument("--R", type=int, default=512)
    ap.add_argument("--graph-seed", type=int, default=3001)
    ap.add_argument("--query-seed", type=int, default=3002)
    ap.add_argument("--plain", type=int, default=64, help="queries timed on the plain (one-at-a-time) solver")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-qps", type=float, nargs="+", default=[50.0, 200.0, 800.0])
    ap.add_argument("--serve-seconds", type=float, default=4.0)
    ap.add_argument("--serve-max", type=int, default=1200)
    ap.add_argument("--serve-grid-slots", type=int, default=32, help="slots of the grid lane")
    ap.add_argument("--serve-grid-wgs", type=int, default=8,
                    help="workgroups per slot of the grid lane's deep relabel BFS")
    ap.add_argument("--serve-slabs", type=float, default=24.0,
                    help="slab arena budget in full R-MAT 22 slabs")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run 3 device-to-device copies of BYTES (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch

        a = torch.empty(args.calibrate_copy, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        a.fill_(1)
        torch.cuda.synchronize()
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": args.calibrate_copy, "copies": 3}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU (one MI355X); across GPUs this path shards by query, each GPU "
              "running its own engines on its own share of the queries; run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    args.queries = args.steps * args.batch
    if M.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(3)
    if args.workload == "dynamic":
        if args.scale == 22 and "--scale" not in sys.argv:
            args.scale = 23
        dynamic_bench(args)
        return

    e, hg, d, qs, qt, name, desc = build_workload(args)
    n = len(d["new_id"])
    E = len(d["col"])
    t0 = time.time()
    dg = M.DeviceGraph(hg)
    log(f"upload {dg.info()} in {time.time() - t0:.1f}s")
    timed_q = slice(args.batch * args.warmup, args.batch * args.warmup + args.queries)
    tq_s, tq_t = qs[timed_q], qt[timed_q]

    # ---- batched (the measured value): --engines engines share the --batch
    # slots, each driven by its own host thread on its own HIP stream
    if args.batch % args.engines:
        raise SystemExit("--batch must be a multiple of --engines")
    engs = []
    for _ in range(args.engines):
        eng = M.Engine(args.batch // args.engines)
        engs.append((eng, eng.add_graph(dg)))
    if args.warmup:
        run_engines(engs, qs, qt, 0, args.batch * args.warmup)
    for eng, _ in engs:
        eng.profile(True)
    t0 = time.perf_counter()
    res, lat, svc, flows = run_engines(engs, qs, qt, args.batch * args.warmup,
                                       args.batch * args.warmup + args.queries)
    wall = time.perf_counter() - t0
    ms, launches, work = 0.0, 0, 0
    hits = misses = 0
    for eng, _ in engs:
        m_, l_, w_ = eng.profile_read()
        eng.profile(False)
        ms, launches, work = ms + m_, launches + l_, work + w_
        h_, mi_ = eng.cache_stats()
        hits, misses = hits + h_, misses + mi_
    qps = args.queries / wall
    e_run = sum(r.E_run for r in res)
    log(f"batched: {args.queries} queries in {wall:.3f}s = {qps:.1f} q/s; cache hits {hits} misses {misses}")
    st = {k: float(np.mean([getattr(r, k) for r in res])) for k in
          ("iters", "steps", "grs", "bfs_levels", "drains", "pruned", "n_run", "E_run", "kept")}
    log("per-query means: " + ", ".join(f"{k}={v:.1f}" for k, v in st.items()))
    log("family ms: " + ", ".join(f"{FAMILIES[f]}={ms[f]:.1f}/{launches[f]}" for f in range(len(FAMILIES))))

    # roofline of the dominant kernel family (HIP events on the engine stream)
    fam = int(np.argmax(ms))
    algo_bytes = float((work[fam].astype(np.float64) * UNIT_BYTES).sum())
    achieved = algo_bytes / (ms[fam] * 1e-3) if ms[fam] > 0 else 0.0
    alg_per_launch = algo_bytes / max(1, int(launches[fam]))
    from mscresearch_amd.profiling import traffic_for

    traffic = traffic_for(name, FAMILIES[fam])
    traffic = round(traffic, 1) if traffic is not None else None
    fam_table = {FAMILIES[f]: {"ms": round(float(ms[f]), 3), "launches": int(launches[f]),
                               "alg_bytes": float((work[f].astype(np.float64) * UNIT_BYTES).sum())}
                 for f in range(len(FAMILIES))}

    # ---- plain: one query at a time on one slot
    plain_lat = []
    plain_flows = {}
    if args.plain > 0:
        peng = M.Engine(1)
        pgid = peng.add_graph(dg)
        k = min(args.plain, args.queries)
        peng.solve(pgid, tq_s[0], tq_t[0])  # warm the subgraph cache like the batched warmup
        peng.profile(True)
        t0 = time.perf_counter()
        for i in range(k):
            r = peng.solve(pgid, tq_s[i], tq_t[i])
            plain_lat.append(r.t_end - r.t_start)
            plain_flows[i] = r.flow
        pwall = time.perf_counter() - t0
        pms, pla, _ = peng.profile_read()
        plain_qps = k / pwall
        del peng
        log(f"plain: {k} queries in {pwall:.3f}s = {plain_qps:.1f} q/s, p50 {pct(plain_lat, .5) * 1e3:.2f} ms")
        log("plain family ms: " + ", ".join(f"{FAMILIES[f]}={pms[f]:.1f}/{pla[f]}" for f in range(len(FAMILIES))))
        for i in range(k):
            assert plain_flows[i] == flows[args.batch * args.warmup + i], "plain and batched flows differ"
    else:
        plain_qps = None

    del engs, eng  # their slab arenas are not needed any more
    serving = None
    if not args.no_serve and args.workload == "rmat" and args.scale == 22:
        serving = serve_bench(dg, hg, e, args)
        if plain_lat:
            serving["plain_p50_ms_rmat22"] = round(pct(plain_lat, .5) * 1e3, 3)

    # ---- CPU oracle baseline on a bounded sample; bit-exact check of the sample
    cpu = None
    exact = None
    if not args.no_cpu:
        threads = max(1, min(16, os.cpu_count() or 1))
        oflows, osec, k, sec1 = cpu_baseline(d, tq_s, tq_t, args.cpu_seconds, threads)
        mism = sum(int(oflows[i] != flows[args.batch * args.warmup + i]) for i in range(k))
        exact = {"checked": k, "mismatches": mism}
        cpu = {"value": round(k / osec, 3), "unit": "queries/s", "cores": threads, "kind": "port",
               "sample": f"first {k} timed queries of the same workload, oracle Dinic, one query per thread; "
                         f"single-thread {1.0 / max(sec1, 1e-9):.3f} q/s",
               "model": open("/proc/cpuinfo").read().split("model name")[1].split("\n")[0].strip(": ")
               if os.path.exists("/proc/cpuinfo") and "model name" in open("/proc/cpuinfo").read() else "unknown"}
        if mism:
            log(f"EXACTNESS FAILURE: {mism} of {k} flows differ from the oracle")

    out = {
        "metric": METRIC,
        "value": round(qps, 3),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic",
        "config": {"workload": name, "graph": desc, "n": n, "csr_entries": E, "queries": args.queries,
                   "batch": args.batch, "engines": args.engines, "regions": args.R,
                   "parallelism": f"one GPU, {args.engines} engine(s) x {args.batch // args.engines} batched slots, "
                                  "one host thread and HIP stream each"},
        "plain_qps": round(plain_qps, 3) if plain_qps else None,
        "speedup_vs_plain": round(qps / plain_qps, 3) if plain_qps else None,
        "p50_ms": round(pct(lat, .5) * 1e3, 3),
        "p99_ms": round(pct(lat, .99) * 1e3, 3),
        "latency_note": "p50/p99: submit to result with all timed queries submitted at once; "
                        "service_*: admission to result",
        "service_p50_ms": round(pct(svc, .5) * 1e3, 3),
        "service_p99_ms": round(pct(svc, .99) * 1e3, 3),
        "plain_p50_ms": round(pct(plain_lat, .5) * 1e3, 3) if plain_lat else None,
        "plain_p99_ms": round(pct(plain_lat, .99) * 1e3, 3) if plain_lat else None,
        "residual_edges_per_s": round(e_run / wall, 1),
        "certified": sum(int(r.cert_ok) for r in res),
        "exact_sample": exact,
        "subgraph_cache": {"hits": hits, "misses": misses},
        "kernel_families": fam_table,
        "roofline": {"bound": "hbm", "kernel_family": FAMILIES[fam], "achieved": round(achieved / 1e9, 3),
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 5),
                     "traffic": traffic, "alg_bytes_per_launch": round(alg_per_launch, 1),
                     "ms_per_launch": round(float(ms[fam]) / max(1, int(launches[fam])), 4),
                     "traffic_source": f"profiles/{name}_hbm_traffic.json (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, "
                                       "copy-calibrated)" if traffic is not None else None},
        "cpu_baseline": cpu,
        "serving": serving,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
