"""Serving loop on the GPU: several resident graphs, queries arriving over
time, admission by slab bytes, concurrent execution (continuous batching).

Each served flow must equal the same query solved alone and the oracle's
(the max-flow value is unique, so "the same alone and in the mix" is exact
equality).  Admission must never put more slab bytes in use than the budget.
"""
import numpy as np
import pytest

import mscresearch_amd as M
import oracle as O
from gpu_util import Case, log

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def graphs():
    log("serving graphs: grid 40x40, rmat 12, uniform 3000/30000")
    return [
        Case(M.Edges.grid(40, 40, 1000, 6001), R=16, tile_max=256),
        Case(M.Edges.rmat(12, 10**6, 6002), R=32),
        Case(M.Edges.uniform(3000, 30000, 100, 6003), R=16),
    ]


def stream(graphs, nq, qps, seed):
    rng = np.random.default_rng(seed)
    gsel = rng.choice(3, nq, p=[0.3, 0.4, 0.3])
    arrive = np.cumsum(rng.exponential(1.0 / qps, nq))
    qs, qt = [], []
    pools = [M.grid_queries(40, 40, nq, seed + 1), graphs[1].edges.hub_queries(nq, seed + 2),
             M.uniform_queries(3000, nq, seed + 3)]
    for i, g in enumerate(gsel):
        s, t = graphs[g].q(int(pools[g][0][i]), int(pools[g][1][i]))
        qs.append(s)
        qt.append(t)
    return gsel.astype(np.int32), np.array(qs, np.int32), np.array(qt, np.int32), arrive


def slab_bytes(case):
    return case.dg.slab_bytes()


@pytest.mark.parametrize("budget_slabs", [0, 2.5])
def test_serve_mixed_graphs_exact_and_within_budget(graphs, budget_slabs):
    eng = M.Engine(8)
    gids = [eng.add_graph(c.dg) for c in graphs]
    big = max(slab_bytes(c) for c in graphs)
    if budget_slabs:
        eng.set_budget(int(budget_slabs * big))
    g, s, t, arrive = stream(graphs, 60, 400.0, 6100)
    res = eng.serve(np.array(gids, np.int32)[g], s, t, arrive)
    assert len(res) == 60
    by = {r.tag: r for r in res}
    for i in range(60):
        r = by[i]
        want = graphs[g[i]].oracle_flow(int(s[i]), int(t[i]))
        assert r.status == 0 and r.cert_ok == 1, (i, r.status)
        assert r.flow == want, (i, r.flow, want)
        assert r.t_start >= r.t_arrive - 1e-9 and r.t_end >= r.t_start
    st = eng.mem_stats()
    log(f"budget {budget_slabs} slabs: {st}")
    assert st["peak"] <= st["budget"]
    if budget_slabs:
        assert st["budget"] <= int(budget_slabs * big) + 256
    # alone: the same queries one at a time on a fresh engine give the same flows
    alone = M.Engine(1)
    agids = [alone.add_graph(c.dg) for c in graphs]
    for i in range(0, 60, 7):
        assert alone.solve(agids[g[i]], int(s[i]), int(t[i])).flow == by[i].flow


def test_serve_reports_queries_that_can_never_fit(graphs):
    eng = M.Engine(4)
    gids = [eng.add_graph(c.dg) for c in graphs]
    room = max(slab_bytes(graphs[0]), slab_bytes(graphs[2]))
    need_rmat = slab_bytes(graphs[1])
    assert need_rmat > 1.2 * room
    eng.set_budget(int(room * 1.2))  # room for a grid or a uniform slab, not for the R-MAT one
    hs, ht = graphs[1].edges.hub_queries(1, 7)
    pairs = [graphs[0].q(0, 39), graphs[1].q(int(hs[0]), int(ht[0])), graphs[2].q(5, 9), graphs[0].q(40, 79)]
    g = np.array([0, 1, 2, 0], np.int32)
    s = np.array([p[0] for p in pairs], np.int32)
    t = np.array([p[1] for p in pairs], np.int32)
    res = {r.tag: r for r in eng.serve(np.array(gids, np.int32)[g], s, t, np.zeros(4))}
    assert len(res) == 4
    assert res[1].status == M.MF_ENOMEM
    for i in (0, 2, 3):
        assert res[i].status == 0
        assert res[i].flow == graphs[g[i]].oracle_flow(int(s[i]), int(t[i]))


def test_serve_lanes_one_engine_per_graph(graphs):
    """One engine lane per graph (own thread and stream): the same flows as
    the oracle, every lane's slab use within its budget."""
    engines = []
    for c in graphs:
        eng = M.Engine(4, M.default_params(fuse=16))
        eng.add_graph(c.dg)
        eng.set_budget(int(2.5 * slab_bytes(c)))
        engines.append(eng)
    g, s, t, arrive = stream(graphs, 60, 300.0, 6200)
    res = M.serve_lanes(engines, g, np.zeros(len(g)), s, t, arrive)
    assert len(res) == 60
    assert all(res[k].t_end <= res[k + 1].t_end for k in range(59))
    for r in res:
        i = r.tag
        assert r.status == 0 and r.cert_ok == 1
        assert r.flow == graphs[g[i]].oracle_flow(int(s[i]), int(t[i])), i
    for e in engines:
        st = e.mem_stats()
        assert st["peak"] <= st["budget"]


def test_lanes_share_one_graph_across_many_classes():
    """Three engines, each on its own host thread, share one freshly uploaded
    graph, so its per-SCC forward-set cache starts empty; their first queries
    come from many SCC classes at once, and the engines build and insert
    forward sets concurrently.  Every flow equals the oracle's."""
    case = Case(M.Edges.rmat(13, 10**6, 6301), R=32)
    scc, _wcc = case.hg.components()
    n, E, _R, _l = case.hg.info()
    og = case.og
    owner = np.repeat(np.arange(n), np.diff(og.row_ptr))
    outdeg = np.bincount(owner[og.cap > 0], minlength=n)
    indeg = np.bincount(og.col[og.cap > 0], minlength=n)
    rng = np.random.default_rng(6302)
    # one source per distinct SCC among vertices with out-edges, up to 72
    cand = rng.permutation(np.nonzero(outdeg > 0)[0])
    seen, srcs = set(), []
    for v in cand:
        if scc[v] not in seen:
            seen.add(scc[v])
            srcs.append(int(v))
        if len(srcs) == 72:
            break
    tgt = np.nonzero(indeg > 0)[0]
    s = np.array(srcs, np.int32)
    t = rng.choice(tgt, len(s)).astype(np.int32)
    t = np.where(t == s, tgt[0] if tgt[0] != s[0] else tgt[1], t).astype(np.int32)
    log(f"shared graph: {len(s)} queries from {len(set(scc[s].tolist()))} SCC classes over 3 lanes")
    dg = M.DeviceGraph(case.hg)
    engines = [M.Engine(6) for _ in range(3)]
    gids = [e.add_graph(dg) for e in engines]
    assert gids == [0, 0, 0]
    lane = (np.arange(len(s)) % 3).astype(np.int32)
    res = M.serve_lanes(engines, lane, np.zeros(len(s), np.int32), s, t, np.zeros(len(s)))
    assert len(res) == len(s)
    flows, _sec = O.maxflow_many(og, s, t, threads=4)
    for r in res:
        assert r.status == 0 and r.cert_ok == 1, (r.tag, r.status)
        assert r.flow == flows[r.tag], (r.tag, r.flow, flows[r.tag])
