"""Benchmark: exact max-flow queries/sec on one MI355X, batched vs plain.

Headline (the timed region): R-MAT scale 22 (a,b,c,d = .57,.19,.19,.05;
16*2^22 samples, duplicates merged by summing log-uniform [1,1e6] caps,
graph seed 3001), its 1,024 hub queries (s from the top 1% by out-degree, t
with in-degree > 0, query seed 3002).  One step is one pass over the 1,024
queries, all submitted at once to --engines engines sharing --batch slots
(each on its own host thread and HIP stream, through mf_serve_lanes), which
refill a slot as soon as its query finishes (continuous batching).  `value`
is queries/s over --steps timed passes with the graph resident, kernel
profiling off.  Every timed flow is compared with the CPU oracle's (the
committed fixture tests/golden/config_flows.json) and carries the GPU's
min-cut certificate; the bench exits 1 after its line if any differs.

Beside it (not timed): the plain solver (one slot, one query at a time,
profiling off) on the first --plain queries; a profiled pass for the kernel
family times behind `roofline`; an open-loop Poisson sweep on R-MAT 22 at
fractions of the batched rate (p50/p99 against the plain median); and, unless
switched off, the other workloads as sub-records, each oracle-checked: c2
(grid 1024x1024, R 128 and 512), c4 (serving three resident graphs), c5
(R-MAT 23 dynamic, 20 rounds); and the CPU oracle baseline.

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
from mscresearch_amd import workloads as W  # noqa: E402

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


def fixture(name):
    path = W.fixture_path()
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get(name)


class Workload:
    """A generated workload, its host graph at R, its device graph, and its
    queries (region-ordered ids) with the oracle's flows when committed."""

    def __init__(self, name, R, tile_max=4096, e=None):
        t0 = time.time()
        self.name = name
        e = W.edges(name) if e is None else e
        self.edges = e
        self.n, self.lines = e.info()
        fx = fixture(name)
        if fx is not None:
            _n, us, vs, cs = e.arrays()
            if W.edge_hash(us, vs, cs) != fx["edge_hash"]:
                raise SystemExit(f"{name}: the generated graph differs from the oracle fixture's")
            s0, t0_ = np.asarray(fx["s"], np.int64), np.asarray(fx["t"], np.int64)
            self.flows = fx["flows"]
            self.us, self.vs, self.cs = us, vs, cs
        else:
            s0, t0_ = W.queries(name, e)
            self.flows = None
        self.hg = M.HostGraph.from_edgelist(e).partition(R)
        nid = self.hg.new_ids()
        self.s = nid[s0].astype(np.int32)
        self.t = nid[t0_].astype(np.int32)
        self.dg = M.DeviceGraph(self.hg, tile_max)
        self.E = self.dg.info()["E"]
        log(f"{name} R={R}: n={self.n} lines={self.lines} E={self.E} ready in {time.time() - t0:.1f}s"
            f"{'' if self.flows is not None else ' (no oracle fixture)'}")


def mismatches(res, want, idx_of_tag=None):
    """(differing or uncertified, checked)."""
    bad = 0
    for r in res:
        i = r.tag if idx_of_tag is None else idx_of_tag[r.tag]
        if r.status != 0 or r.cert_ok != 1 or (want is not None and r.flow != want[i]):
            bad += 1
    return bad, (len(res) if want is not None else 0)


def pass_over(engs, s, t, arrive=None):
    """All queries of one pass over several engines (query i on engine i % k),
    through mf_serve_lanes.  Results in completion order, tag = i."""
    k = len(engs)
    q = len(s)
    lane = (np.arange(q) % k).astype(np.int32)
    return M.serve_lanes(engs, lane, np.zeros(q, np.int32), s, t, np.zeros(q) if arrive is None else arrive)


def family_table(engs):
    ms, launches, work = np.zeros(7), np.zeros(7, np.int64), np.zeros((7, 4), np.uint64)
    for e in engs:
        m_, l_, w_ = e.profile_read()
        ms, launches, work = ms + m_, launches + l_, work + w_
    return {FAMILIES[f]: {"ms": round(float(ms[f]), 3), "launches": int(launches[f]),
                          "alg_bytes": float((work[f].astype(np.float64) * UNIT_BYTES).sum())}
            for f in range(7)}


def roofline_of(fam_table, workload_name):
    from mscresearch_amd.profiling import traffic_for

    fam = max(fam_table, key=lambda k: fam_table[k]["ms"])
    st = fam_table[fam]
    per_launch_ms = st["ms"] / max(1, st["launches"])
    alg = st["alg_bytes"] / max(1, st["launches"])
    achieved = alg / (per_launch_ms * 1e-3) if per_launch_ms > 0 else 0.0
    tr = traffic_for(workload_name, fam, full=True)
    out = {"bound": "hbm", "kernel_family": fam, "achieved": round(achieved / 1e9, 3), "peak": HBM_PEAK / 1e9,
           "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 5), "traffic": None,
           "alg_bytes_per_launch": round(alg, 1), "ms_per_launch": round(per_launch_ms, 4)}
    if tr is not None:
        out["traffic"] = round(tr["hbm_bytes_per_launch"], 1)
        out["traffic_run_alg_bytes_per_launch"] = round(tr["alg_bytes_per_launch"], 1)
        out["traffic_over_alg"] = round(tr["hbm_bytes_per_launch"] / max(1.0, tr["alg_bytes_per_launch"]), 3)
        out["traffic_source"] = (f"profiles/{workload_name}_hbm_traffic.json: rocprofv3 --pmc FETCH_SIZE and "
                                 "WRITE_SIZE runs of the same bench (warmed), copy-calibrated; traffic and its "
                                 "algorithmic bytes per launch from those runs")
    return out


# ------------------------------------------------------------------ c3
def headline(args, out):
    e3 = W.edges("c3")
    # the autotuner's recorded choice for this graph class and engine size
    # (mscresearch_amd/tuned_params.txt, written by python -m mscresearch_amd.tune)
    tv = None
    if not args.no_tune:
        from mscresearch_amd.tune import cached

        tv = cached(M.HostGraph.from_edgelist(e3), args.tune_batch)
    R, tile, slots = args.R, 4096, args.batch // args.engines
    prm = M.default_params()
    if tv is not None:
        # the tuner times one engine; its slots are shared by the bench's engines
        R, tile, slots = tv["R"], tv["tile_max"], max(1, tv["slots"] // args.engines)
        prm = M.default_params(K=tv["K"], gr_freq=tv["gr_freq"], gr_freq_min=min(tv["gr_freq"], prm.gr_freq_min),
                               gr_freq_max=max(2 * tv["gr_freq"], prm.gr_freq_max), tile_threads=tv["tile_threads"],
                               tile_split=tv.get("tile_split", 1))
        log(f"tuned parameters for {tv['graph_class']}: {tv}")
    args.R = R
    c3 = Workload("c3", R, tile, e=e3)
    q = len(c3.s)
    engs = []
    for _ in range(args.engines):
        e = M.Engine(slots, prm)
        e.add_graph(c3.dg)
        engs.append(e)
    bad_total = checked = 0
    for w in range(args.warmup):
        res = pass_over(engs, c3.s, c3.t)
        b, c = mismatches(res, c3.flows)
        bad_total, checked = bad_total + b, checked + c
    # the timed passes: profiling off
    t0 = time.perf_counter()
    all_res = []
    for k in range(args.steps):
        all_res.append(pass_over(engs, c3.s, c3.t))
    wall = time.perf_counter() - t0
    for res in all_res:
        b, c = mismatches(res, c3.flows)
        bad_total, checked = bad_total + b, checked + c
    qps = args.steps * q / wall
    res = all_res[-1]
    lat = [r.t_end - r.t_arrive for rr in all_res for r in rr]
    svc = [r.t_end - r.t_start for rr in all_res for r in rr]
    e_run = sum(r.E_run for rr in all_res for r in rr)
    st = {k: round(float(np.mean([getattr(r, k) for r in res])), 2) for k in
          ("iters", "steps", "grs", "bfs_levels", "drains", "pruned", "n_run", "E_run", "kept")}
    log(f"batched: {args.steps} x {q} queries in {wall:.3f}s = {qps:.1f} q/s; per query {st}")
    hits = sum(e.cache_stats()[0] for e in engs)
    misses = sum(e.cache_stats()[1] for e in engs)
    del engs, e

    # roofline pass: the same queries through ONE engine holding every batched
    # slot, HIP events per kernel family on its stream.  Concurrent engines
    # stretch each other's launches, so their per-launch durations measure
    # contention; one engine's launches measure the kernels.  The trace
    # markers bracket this pass alone (the profiles' rocprof runs summarise
    # the same dispatches).
    reng = M.Engine(args.engines * slots, prm)
    reng.add_graph(c3.dg)
    pass_over([reng], c3.s, c3.t)  # slabs allocated, residual resets by marked entries
    M.trace_marker()
    reng.profile(True)
    t1 = time.perf_counter()
    pres = pass_over([reng], c3.s, c3.t)
    pwall = time.perf_counter() - t1
    b, c = mismatches(pres, c3.flows)
    bad_total, checked = bad_total + b, checked + c
    fams = family_table([reng])
    M.trace_marker()
    reng.profile(False)
    log(f"roofline pass (1 engine x {args.engines * slots} slots): {q / pwall:.1f} q/s; family ms: " +
        ", ".join(f"{k}={v['ms']:.1f}/{v['launches']}" for k, v in fams.items()))
    del reng

    # plain: one slot, one query at a time, profiling off (the same parameters)
    peng = M.Engine(1, prm)
    pgid = peng.add_graph(c3.dg)
    kq = min(args.plain, q)
    peng.solve(pgid, c3.s[0], c3.t[0])  # the reachability classes, as the batched warm-up
    plain_lat, plain_bad = [], 0
    t1 = time.perf_counter()
    for i in range(kq):
        r = peng.solve(pgid, c3.s[i], c3.t[i])
        plain_lat.append(r.t_end - r.t_start)
        if r.cert_ok != 1 or (c3.flows is not None and r.flow != c3.flows[i]):
            plain_bad += 1
    plain_wall = time.perf_counter() - t1
    plain_qps = kq / plain_wall if kq else None
    bad_total += plain_bad
    checked += kq if c3.flows is not None else 0
    if kq:
        log(f"plain: {kq} queries in {plain_wall:.3f}s = {plain_qps:.1f} q/s, p50 {pct(plain_lat, .5) * 1e3:.3f} ms")
    del peng

    # open-loop Poisson sweep on R-MAT 22 alone, offered at fractions of the batched rate
    sweep = None
    if not args.no_sweep and kq:
        sweep = {"offered_fractions": args.sweep, "plain_p50_ms": round(pct(plain_lat, .5) * 1e3, 3),
                 "target_p99_ms": round(2 * pct(plain_lat, .5) * 1e3, 3), "levels": []}
        engs = []
        for _ in range(args.engines):
            e = M.Engine(slots, prm)
            e.add_graph(c3.dg)
            engs.append(e)
        pass_over(engs, c3.s[:args.engines * slots], c3.t[:args.engines * slots])  # slabs in place
        rng = np.random.default_rng(3003)
        for frac in args.sweep:
            rate = frac * qps
            nq = int(min(q, max(64, rate * args.sweep_seconds)))
            pick = rng.integers(0, q, nq)
            arrive = np.cumsum(rng.exponential(1.0 / rate, nq))
            res = pass_over(engs, c3.s[pick], c3.t[pick], arrive)
            b, c = mismatches(res, None if c3.flows is None else [c3.flows[i] for i in pick])
            bad_total, checked = bad_total + b, checked + c
            lat_s = [r.t_end - r.t_arrive for r in res]
            span = max(r.t_end for r in res) - min(r.t_arrive for r in res)
            lv = {"fraction": frac, "offered_qps": round(rate, 1), "queries": nq,
                  "sustained_qps": round(nq / span, 1), "p50_ms": round(pct(lat_s, .5) * 1e3, 3),
                  "p99_ms": round(pct(lat_s, .99) * 1e3, 3)}
            lv["p99_within_2x_plain_p50"] = bool(lv["p99_ms"] <= sweep["target_p99_ms"])
            log(f"sweep {lv}")
            sweep["levels"].append(lv)
        del engs

    out.update({
        "value": round(qps, 3),
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "config": {"workload": "rmat22_hub_queries", "graph": "R-MAT scale 22 (a=.57,b=.19,c=.19,d=.05), 16*2^22 "
                   "samples, duplicates merged by summing log-uniform [1,1e6] caps, seed 3001",
                   "queries_per_step": q, "query_rule": "s uniform over the top 1% by out-degree, t uniform over "
                   "in-degree > 0, seed 3002", "n": c3.n, "csr_entries": c3.E, "batch": args.engines * slots,
                   "engines": args.engines, "regions": R, "tile_max": tile, "K": prm.K, "gr_freq": prm.gr_freq,
                   "tile_threads": prm.tile_threads, "tile_split": prm.tile_split,
                   "tuned": None if tv is None else {k: tv[k] for k in ("graph_class", "R", "tile_max", "K", "gr_freq",
                                                                        "tile_threads", "slots", "tile_split",
                                                                        "ms_per_query",
                                                                        "default_ms_per_query")},
                   "parallelism": f"one GPU, {args.engines} engine(s) x {slots} batched slots, "
                                  "one host thread and HIP stream each"},
        "plain_qps": round(plain_qps, 3) if plain_qps else None,
        "plain_queries": kq,
        "speedup_vs_plain": round(qps / plain_qps, 3) if plain_qps else None,
        "p50_ms": round(pct(lat, .5) * 1e3, 3),
        "p99_ms": round(pct(lat, .99) * 1e3, 3),
        "latency_note": "p50/p99: submit to result with each step's 1,024 queries submitted at once; service_*: "
                        "admission to result; open-loop latency under load in latency_sweep",
        "service_p50_ms": round(pct(svc, .5) * 1e3, 3),
        "service_p99_ms": round(pct(svc, .99) * 1e3, 3),
        "plain_p50_ms": round(pct(plain_lat, .5) * 1e3, 3) if plain_lat else None,
        "plain_p99_ms": round(pct(plain_lat, .99) * 1e3, 3) if plain_lat else None,
        "residual_edges_per_s": round(e_run / wall, 1),
        "per_query_means": st,
        "subgraph_cache": {"hits": hits, "misses": misses},
        "kernel_families": fams,
        "roofline_pass": {"engines": 1, "slots": args.engines * slots, "qps": round(q / pwall, 3),
                          "note": "kernel_families and roofline come from this pass: the same queries through one "
                                  "engine holding every batched slot, HIP events on its stream, between the trace "
                                  "markers"},
        "roofline": roofline_of(fams, "rmat22_hub_queries"),
        "latency_sweep": sweep,
    })
    return c3, bad_total, checked


# ------------------------------------------------------------------ c2
def grid_record(args):
    rec = {"workload": "grid1024x1024_left_right", "queries": 0, "per_R": []}
    bad_total = checked = 0
    for R in args.grid_R:
        g = Workload("c2", R)
        eng = M.Engine(args.grid_slots, M.default_params(fuse=64))
        gid = eng.add_graph(g.dg)
        q = min(args.grid_queries, len(g.s))
        # warm-up over the first slots' worth of queries: slabs allocated and
        # reset once, the reachability class cached, as in serving
        for i in range(min(q, args.grid_slots)):
            eng.submit(gid, g.s[i], g.t[i], i)
        wres = eng.run()
        wb, wc = mismatches(wres, g.flows)
        bad_total, checked = bad_total + wb, checked + wc
        if args.profile_all:
            M.trace_marker()
            eng.profile(True)
        t0 = time.perf_counter()
        for i in range(q):
            eng.submit(gid, g.s[i], g.t[i], i)
        res = eng.run()
        wall = time.perf_counter() - t0
        b, c = mismatches(res, g.flows)
        bad_total, checked = bad_total + b, checked + c
        if not args.profile_all:
            # kernel families from a profiled pass over the first queries
            M.trace_marker()
            eng.profile(True)
            pq = min(q, args.grid_slots)
            for i in range(pq):
                eng.submit(gid, g.s[i], g.t[i], i)
            pres = eng.run()
            pb, pc = mismatches(pres, g.flows)
            bad_total, checked = bad_total + pb, checked + pc
            b, c = b + pb, c + pc
        fams = family_table([eng])
        M.trace_marker()
        eng.profile(False)
        lat = [r.t_end - r.t_arrive for r in res]
        e_run = sum(r.E_run for r in res)
        roof = roofline_of(fams, "grid1024x1024_left_right")
        # the plain solver (one slot, one query at a time) on the first queries
        plain = {}
        if args.grid_plain > 0:
            pe = M.Engine(1, M.default_params(fuse=64))
            pg = pe.add_graph(g.dg)
            pq = min(args.grid_plain, len(g.s))
            pe.solve(pg, g.s[0], g.t[0])  # warm: slab allocated, pruning cached
            t1 = time.perf_counter()
            pres = [pe.solve(pg, g.s[i], g.t[i]) for i in range(pq)]
            pw = time.perf_counter() - t1
            for i, r in enumerate(pres):
                r.tag = i
            pb, pc = mismatches(pres, g.flows)
            bad_total, checked = bad_total + pb, checked + pc
            b, c = b + pb, c + pc
            plat = [r.t_end - r.t_start for r in pres]
            plain = {"plain_queries": pq, "plain_qps": round(pq / pw, 3), "plain_p50_ms": round(pct(plat, .5) * 1e3, 3)}
            del pe
        rr = {"R": R, "slots": args.grid_slots, "queries": q, "qps": round(q / wall, 3), **plain,
              "p50_ms": round(pct(lat, .5) * 1e3, 3), "p99_ms": round(pct(lat, .99) * 1e3, 3),
              "residual_edges_per_s": round(e_run / wall, 1),
              "mean_bfs_levels": round(float(np.mean([r.bfs_levels for r in res])), 1),
              "oracle_mismatches": b, "oracle_checked": c, "kernel_families": fams, "roofline": roof}
        log(f"c2 R={R}: {q} queries {rr['qps']} q/s, {b} of {c} differ, roofline {roof['kernel_family']} "
            f"{roof['achieved']} GB/s")
        rec["per_R"].append(rr)
        rec["queries"] += q
        del eng, g
    return rec, bad_total, checked


# ------------------------------------------------------------------ c4
def serve_record(args, c3):
    """Serving: grid 1024x1024, R-MAT 20 and R-MAT 22 resident, one engine lane
    each (own host thread and HIP stream, slab arena sized in slabs of its
    graph); Poisson arrivals at each offered load, graph chosen 0.3 / 0.4 /
    0.3 from each workload's oracle-checked query pool; each lane admits by
    slab bytes and batches continuously."""
    gw = Workload("c2", 512)
    r20 = Workload("r20", 512)
    ws = [gw, r20, c3]
    lanes = [("grid1024", args.serve_grid_slots, M.default_params(fuse=64, small_wgs=args.serve_grid_wgs)),
             ("rmat20", 96, M.default_params()), ("rmat22", 48, M.default_params())]
    engines, budgets = [], []
    for (nm, slots, prm), w in zip(lanes, ws):
        eng = M.Engine(slots, prm)
        eng.add_graph(w.dg)
        slab = w.dg.slab_bytes()
        nslab = args.serve_slabs if nm == "rmat22" else slots
        eng.set_budget(int(nslab * slab))
        engines.append(eng)
        budgets.append(int(nslab * slab))
    rng = np.random.default_rng(4004)
    wq = [(k, j) for k in range(3) for j in range(4)]  # reachability classes of the first queries
    M.serve_lanes(engines, np.array([k for k, _ in wq], np.int32), np.zeros(len(wq), np.int32),
                  np.array([ws[k].s[j] for k, j in wq], np.int32), np.array([ws[k].t[j] for k, j in wq], np.int32),
                  np.zeros(len(wq)))
    used = [4, 4, 4]
    levels = []
    bad_total = checked = 0
    for qps in args.serve_qps:
        nq = int(min(args.serve_max, max(20, qps * args.serve_seconds)))
        g = rng.choice(3, nq, p=[0.3, 0.4, 0.3])
        arrive = np.cumsum(rng.exponential(1.0 / qps, nq))
        idx = np.empty(nq, np.int64)
        for i in range(nq):
            idx[i] = used[g[i]] % len(ws[g[i]].s)
            used[g[i]] += 1
        s = np.array([ws[g[i]].s[idx[i]] for i in range(nq)], np.int32)
        t = np.array([ws[g[i]].t[idx[i]] for i in range(nq)], np.int32)
        before = [e.mem_stats()["admit_waits"] for e in engines]
        res = M.serve_lanes(engines, g.astype(np.int32), np.zeros(nq, np.int32), s, t, arrive)
        want = None if any(w.flows is None for w in ws) else [ws[g[i]].flows[idx[i]] for i in range(nq)]
        b, c = mismatches(res, want)
        bad_total, checked = bad_total + b, checked + c
        lat = [r.t_end - r.t_arrive for r in res]
        span = max(r.t_end for r in res) - min(r.t_arrive for r in res)
        per = {}
        for k, (nm, *_rest) in enumerate(lanes):
            lk = [r.t_end - r.t_arrive for r in res if g[r.tag] == k]
            per[nm] = {"n": len(lk), "p50_ms": round(pct(lk, .5) * 1e3, 3), "p99_ms": round(pct(lk, .99) * 1e3, 3),
                       "admission_waits": engines[k].mem_stats()["admit_waits"] - before[k]}
        lv = {"offered_qps": qps, "queries": nq, "sustained_qps": round(nq / span, 2),
              "p50_ms": round(pct(lat, .5) * 1e3, 3), "p99_ms": round(pct(lat, .99) * 1e3, 3),
              "oracle_mismatches_or_uncertified": b, "oracle_checked": c, "per_graph": per}
        log(f"serve {qps} q/s offered: {lv}")
        levels.append(lv)
    rec = {"graphs": ["grid 1024x1024 caps U[1,1000] seed 2001 (c2 queries)",
                      "R-MAT 20 seed 4001 (hub queries seed 4002)", "R-MAT 22 (the c3 graph and queries)"],
           "mix": [0.3, 0.4, 0.3],
           "lanes": [{"graph": nm, "slots": sl, "slab_budget_bytes": b, "slab_peak_bytes": e.mem_stats()["peak"]}
                     for (nm, sl, _p), b, e in zip(lanes, budgets, engines)],
           "levels": levels}
    return rec, bad_total, checked


# ------------------------------------------------------------------ c5
def dynamic_record(args):
    """R-MAT 23, 32 hub queries pinned to slots, 20 rounds each changing the
    capacity of a uniformly chosen 1% of the input lines by a factor U[0.5, 2]
    (rounded, floored at 1); every round re-solves all pinned queries warm
    (from their residuals, repaired by the updates) and cold (a static
    re-solve on the updated graph), both against the oracle's flows."""
    w = Workload("c5", args.R)
    spec = W.WORKLOADS["c5"]["updates"]
    if w.flows is None:
        _n, w.us, w.vs, w.cs = w.edges.arrays()
    ent = w.hg.line_entries(w.us, w.vs)
    warm = M.Engine(32, M.default_params(prune=0))
    gid = warm.add_graph(w.dg)
    cold = M.Engine(16)
    cgid = cold.add_graph(w.dg)
    for i in range(len(w.s)):
        warm.pin(i, gid, w.s[i], w.t[i], i)
    base = warm.solve_pinned(warm=False)
    bad_total, checked = mismatches(base, None if w.flows is None else w.flows[0])
    rounds = []
    tw_all = tc_all = tu_all = 0.0
    for rnd, (lines, new) in enumerate(W.capacity_update_rounds(w.cs, spec["rounds"], spec["seed"], spec["frac"]), 1):
        t1 = time.perf_counter()
        warm.update_caps(gid, ent[lines], new)
        tu = time.perf_counter() - t1
        t1 = time.perf_counter()
        wres = warm.solve_pinned(warm=True)
        tw = time.perf_counter() - t1
        for i in range(len(w.s)):
            cold.submit(cgid, w.s[i], w.t[i], i)
        t1 = time.perf_counter()
        cres = cold.run()
        tc = time.perf_counter() - t1
        tw_all, tc_all, tu_all = tw_all + tw, tc_all + tc, tu_all + tu
        want = None if w.flows is None else w.flows[rnd]
        bw, cw = mismatches(wres, want)
        bc, cc = mismatches(cres, want)
        cf = {r.tag: r.flow for r in cres}
        ws_bad = sum(1 for r in wres if r.flow != cf.get(r.tag))
        bad_total, checked = bad_total + bw + bc + ws_bad, checked + cw + cc
        rounds.append({"round": rnd, "updates": int(len(lines)), "update_ms": round(tu * 1e3, 2),
                       "warm_ms": round(tw * 1e3, 2), "static_ms": round(tc * 1e3, 2),
                       "warm_vs_static_mismatches": ws_bad, "oracle_mismatches": bw + bc,
                       "drains": int(sum(r.drains for r in wres))})
    n_round = len(w.s) * len(rounds)
    rec = {"workload": "rmat23_dynamic_32_pinned", "n": w.n, "csr_entries": w.E,
           "updates_per_round": rounds[0]["updates"] if rounds else 0, "rounds": len(rounds),
           "warm_qps": round(n_round / tw_all, 3) if tw_all else None,
           "static_resolve_qps": round(n_round / tc_all, 3) if tc_all else None,
           "warm_vs_static_speedup": round(tc_all / tw_all, 3) if tw_all else None,
           "update_ms_per_round": round(tu_all / max(1, len(rounds)) * 1e3, 2),
           "oracle_checked": checked, "oracle_mismatches_or_uncertified": bad_total, "per_round": rounds}
    log(f"c5: warm {rec['warm_qps']} q/s, static {rec['static_resolve_qps']} q/s, {bad_total} of {checked} differ")
    return rec, bad_total, checked


# ------------------------------------------------------------------ CPU
def cpu_share():
    try:
        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        return os.cpu_count() or 1


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_record(c3, args):
    """The oracle (exact int64 Dinic, -O3, pruned to each query's mask as
    BC:1391-1445 prunes) on a bounded sample of the same queries, one query
    per thread on the host's CPU share (at most 16), plus one thread alone."""
    import oracle as O

    d = c3.hg.export()
    og = O.Graph(c3.n, d["row_ptr"], d["col"], d["cap"], d["rev"])
    del d
    threads = max(1, min(16, cpu_share()))
    _f, sec1 = O.maxflow_many(og, c3.s[:1], c3.t[:1], threads=1, prune=True)
    per = max(sec1, 1e-3)
    k = int(min(len(c3.s), max(threads, threads * args.cpu_seconds / per)))
    flows, sec = O.maxflow_many(og, c3.s[:k], c3.t[:k], threads=threads, prune=True)
    bad = 0 if c3.flows is None else sum(int(flows[i] != c3.flows[i]) for i in range(k))
    log(f"cpu oracle: 1 query {sec1:.2f}s single-thread; {k} queries on {threads} threads in {sec:.2f}s")
    return {"value": round(k / sec, 3), "unit": "queries/s", "cores": threads, "kind": "port",
            "sample": f"first {k} of the 1,024 c3 queries: oracle Dinic (C, -O3) on each query's pruning mask, "
                      f"one query per thread on {threads} threads; one thread alone {1.0 / per:.3f} q/s",
            "single_thread_qps": round(1.0 / per, 4), "host_cpus": os.cpu_count(), "cpu_share": cpu_share(),
            "model": cpu_model(), "oracle_mismatches": bad}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4, help="timed passes over the 1,024 R-MAT 22 queries")
    ap.add_argument("--warmup", type=int, default=1, help="untimed passes before them")
    ap.add_argument("--batch", type=int, default=192, help="engine slots over all engines")
    ap.add_argument("--engines", type=int, default=4,
                    help="engines sharing the batch slots, one host thread and HIP stream each "
                         "(192 slots on R-MAT 22: 1, 2, 3, 4, 6, 8 engines gave 1712, 2089, 2163, 2206, 2010, 1850 q/s)")
    ap.add_argument("--R", type=int, default=512, help="regions when no tuned value is recorded")
    ap.add_argument("--tune-batch", type=int, default=96,
                    help="the autotuner's batched entry to use (its engine size; the tuned slots are shared by --engines)")
    ap.add_argument("--no-tune", action="store_true", help="ignore the autotuner's recorded parameters")
    ap.add_argument("--plain", type=int, default=256, help="queries timed on the plain (one-at-a-time) solver")
    ap.add_argument("--sweep", type=float, nargs="*", default=[0.1, 0.25, 0.5, 0.75],
                    help="open-loop Poisson loads on R-MAT 22, as fractions of the batched rate")
    ap.add_argument("--sweep-seconds", type=float, default=2.0)
    ap.add_argument("--grid-queries", type=int, default=256)
    ap.add_argument("--grid-slots", type=int, default=256)
    ap.add_argument("--grid-plain", type=int, default=48, help="grid queries timed on the plain (one-at-a-time) solver")
    ap.add_argument("--grid-R", type=int, nargs="+", default=[128, 512], help="region counts of the c2 sub-record")
    ap.add_argument("--serve-qps", type=float, nargs="+", default=[50.0, 200.0, 800.0])
    ap.add_argument("--serve-seconds", type=float, default=4.0)
    ap.add_argument("--serve-max", type=int, default=1200)
    ap.add_argument("--serve-grid-slots", type=int, default=512, help="slots of the grid lane")
    ap.add_argument("--serve-grid-wgs", type=int, default=4,
                    help="workgroups per slot of the grid lane's deep relabel BFS")
    ap.add_argument("--serve-slabs", type=float, default=48.0, help="slab arena budget in full R-MAT 22 slabs")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--no-grid", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-dynamic", action="store_true")
    ap.add_argument("--only-headline", action="store_true", help="skip every sub-record and the CPU baseline")
    ap.add_argument("--only-grid", action="store_true", help="only the c2 sub-record (for its profiles)")
    ap.add_argument("--profile-all", action="store_true",
                    help="c2: kernel-family timing over the timed pass too (counter runs: the families then "
                         "cover every dispatch between the run's trace markers); the headline's roofline pass "
                         "is always the one pass between its markers")
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
    if args.batch % args.engines:
        raise SystemExit("--batch must be a multiple of --engines")
    if M.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(3)
    if args.only_headline:
        args.no_cpu = args.no_grid = args.no_serve = args.no_dynamic = True
    out = {"metric": METRIC, "value": None, "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "int64", "data": "synthetic"}
    bad = checked = 0
    if args.only_grid:
        rec, b, c = grid_record(args)
        out.update({"config": {"workload": "grid1024x1024_left_right"}, "c2_grid": rec})
        print(json.dumps(out), flush=True)
        sys.exit(1 if b else 0)
    c3, b, c = headline(args, out)
    bad, checked = bad + b, checked + c
    if not args.no_grid:
        rec, b, c = grid_record(args)
        out["c2_grid"] = rec
        bad, checked = bad + b, checked + c
    if not args.no_serve:
        rec, b, c = serve_record(args, c3)
        out["c4_serving"] = rec
        bad, checked = bad + b, checked + c
    if not args.no_cpu:
        cpu = cpu_record(c3, args)
        out["cpu_baseline"] = cpu
        bad += cpu["oracle_mismatches"]
    else:
        out["cpu_baseline"] = None
    del c3
    if not args.no_dynamic:
        rec, b, c = dynamic_record(args)
        out["c5_dynamic"] = rec
        bad, checked = bad + b, checked + c
    out["exactness"] = {"oracle_checked": checked, "mismatches_or_uncertified": bad,
                        "note": "every flow of every leg compared with the CPU oracle's (committed fixtures) "
                                "and certified on the GPU"}
    print(json.dumps(out), flush=True)
    if bad:
        log(f"EXACTNESS FAILURE: {bad} answers differ from the oracle or failed their certificate")
        sys.exit(1)


if __name__ == "__main__":
    main()
