"""The benchmark workloads at their stated sizes, every flow against the
oracle's (committed fixtures, tests/golden/config_flows.json, written by
tests/golden/make_config_flows.py with the CPU oracle's Dinic).

  c2  grid 1024x1024, 256 left -> right queries, R = 128 and R = 512
  c3  R-MAT 22, all 1,024 hub queries, batched over two engines
  c4  serving: grid + R-MAT 20 + R-MAT 22 resident, one lane each, Poisson
      arrivals at 50 queries/s, graph mix 0.3 / 0.4 / 0.3
  c5  R-MAT 23, 32 pinned queries, 20 rounds of 1% capacity changes: every
      round's warm re-solve and a static re-solve against the oracle

The reference's static driver answers one query after another
(Region_Aware_PR_with_Pruning_BoundaryCore.cu:2628-2843) and its dynamic
driver re-solves from the kept state (:2273-2335); the flows are unique, so
each must equal the oracle's exactly.  Graphs are regenerated here from the
same seeds; the edge-list hash recorded with the fixture must match first.
"""
from __future__ import annotations

import json
import os
import time

import numpy as np
import pytest

import mscresearch_amd as M
from mscresearch_amd import workloads as W
from gpu_util import log

pytestmark = pytest.mark.gpu

_FIX = None
_BUILT = {}


def fixture(name):
    global _FIX
    if _FIX is None:
        path = W.fixture_path()
        _FIX = json.load(open(path)) if os.path.exists(path) else {}
    if name not in _FIX:
        pytest.skip(f"no oracle fixture for {name} (run tests/golden/make_config_flows.py)")
    return _FIX[name]


class Built:
    """A workload's edges, host graph partitioned at R, and device graph."""

    def __init__(self, name, R):
        t0 = time.time()
        fx = fixture(name)
        e = W.edges(name)
        n, us, vs, cs = e.arrays()
        assert n == fx["n"] and len(us) == fx["lines"]
        assert W.edge_hash(us, vs, cs) == fx["edge_hash"], f"{name}: generated graph differs from the fixture's"
        self.n, self.us, self.vs, self.cs = n, us, vs, cs
        self.hg = M.HostGraph.from_edgelist(e).partition(R)
        self.new_id = self.hg.new_ids()
        self.dg = M.DeviceGraph(self.hg)
        self.s = self.new_id[np.asarray(fx["s"], np.int64)].astype(np.int32)
        self.t = self.new_id[np.asarray(fx["t"], np.int64)].astype(np.int32)
        self.flows = fx["flows"]
        log(f"{name} R={R}: n={n} lines={len(us)} ready in {time.time() - t0:.1f}s")


def built(name, R):
    key = (name, R)
    if key not in _BUILT:
        for k in list(_BUILT):  # one large graph at a time besides the serving set
            if k[0] == "c5" or name == "c5":
                del _BUILT[k]
        _BUILT[key] = Built(name, R)
    return _BUILT[key]


def check(results, flows, what):
    bad = [(r.tag, r.flow, flows[r.tag]) for r in results if r.status != 0 or r.cert_ok != 1 or r.flow != flows[r.tag]]
    assert len(results) == len(flows), (what, len(results), len(flows))
    assert not bad, f"{what}: {len(bad)} of {len(flows)} differ, first {bad[:3]}"


@pytest.mark.timeout(600)
@pytest.mark.parametrize("R", [128, 512])
def test_c2_grid_1024_all_queries(R):
    b = built("c2", R)
    eng = M.Engine(64, M.default_params(fuse=32))
    gid = eng.add_graph(b.dg)
    t0 = time.time()
    for i in range(len(b.s)):
        eng.submit(gid, b.s[i], b.t[i], i)
    res = eng.run()
    log(f"c2 R={R}: {len(res)} queries in {time.time() - t0:.1f}s")
    check(res, b.flows, f"c2 R={R}")


@pytest.mark.timeout(600)
def test_c3_rmat22_all_1024_queries():
    b = built("c3", 512)
    engines = [M.Engine(96) for _ in range(2)]
    for e in engines:
        e.add_graph(b.dg)
    q = len(b.s)
    t0 = time.time()
    res = M.serve_lanes(engines, np.arange(q) % 2, np.zeros(q), b.s, b.t, np.zeros(q))
    log(f"c3: {q} queries in {time.time() - t0:.2f}s")
    check(res, b.flows, "c3")


@pytest.mark.timeout(600)
def test_c4_serving_three_graphs_poisson():
    lanes = [("c2", 512, 32, M.default_params(fuse=32)), ("r20", 512, 32, M.default_params()),
             ("c3", 512, 48, M.default_params())]
    bs = [built(nm, R) for nm, R, _k, _p in lanes]
    engines = []
    for (nm, _R, slots, prm), b in zip(lanes, bs):
        e = M.Engine(slots, prm)
        e.add_graph(b.dg)
        engines.append(e)
    rng = np.random.default_rng(4004)
    nq = 150
    g = rng.choice(3, nq, p=[0.3, 0.4, 0.3])
    arrive = np.cumsum(rng.exponential(1.0 / 50.0, nq))
    used = [0, 0, 0]
    idx = np.empty(nq, np.int64)
    for i in range(nq):
        idx[i] = used[g[i]] % len(bs[g[i]].s)
        used[g[i]] += 1
    s = np.array([bs[g[i]].s[idx[i]] for i in range(nq)], np.int32)
    t = np.array([bs[g[i]].t[idx[i]] for i in range(nq)], np.int32)
    t0 = time.time()
    res = M.serve_lanes(engines, g.astype(np.int32), np.zeros(nq, np.int32), s, t, arrive)
    log(f"c4: {nq} arrivals over {arrive[-1]:.1f}s served in {time.time() - t0:.1f}s")
    want = [bs[g[i]].flows[idx[i]] for i in range(nq)]
    check(res, want, "c4")


@pytest.mark.timeout(900)
def test_c5_rmat23_dynamic_20_rounds():
    b = built("c5", 512)
    w = W.WORKLOADS["c5"]["updates"]
    ent = b.hg.line_entries(b.us, b.vs)
    warm = M.Engine(32)
    gid = warm.add_graph(b.dg)
    cold = M.Engine(16)
    cgid = cold.add_graph(b.dg)
    for i in range(len(b.s)):
        warm.pin(i, gid, b.s[i], b.t[i], i)
    base = warm.solve_pinned(warm=False)
    check(base, b.flows[0], "c5 round 0")
    for r, (lines, new) in enumerate(W.capacity_update_rounds(b.cs, w["rounds"], w["seed"], w["frac"]), 1):
        t0 = time.time()
        warm.update_caps(gid, ent[lines], new)
        wres = warm.solve_pinned(warm=True)
        for i in range(len(b.s)):
            cold.submit(cgid, b.s[i], b.t[i], i)
        cres = cold.run()
        log(f"c5 round {r}: {len(lines)} updates, warm + static re-solve in {time.time() - t0:.2f}s")
        check(wres, b.flows[r], f"c5 round {r} warm")
        check(cres, b.flows[r], f"c5 round {r} static")
