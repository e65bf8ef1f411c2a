"""Exact SCC routing in the engine (scc_route = 1): a query whose s and t
share a strongly connected component runs on that component's induced
subgraph; its flow must equal the oracle's on the full graph."""
from __future__ import annotations

import numpy as np
import pytest

import mscresearch_amd as M
import oracle as O
from gpu_util import Case, log

pytestmark = pytest.mark.gpu


class LinesCase:
    """Case for a hand-built line list (0-based)."""

    def __init__(self, n, us, vs, cs, R=16):
        self.n = n
        self.hg = M.HostGraph.from_edges(n, us, vs, cs).partition(R)
        d = self.hg.export()
        self.new_id = d["new_id"]
        self.og = O.Graph(n, d["row_ptr"], d["col"], d["cap"], d["rev"])
        self.dg = M.DeviceGraph(self.hg, 4096)

    def q(self, s, t):
        return int(self.new_id[s]), int(self.new_id[t])


def blocks(n=20000, seed=8101):
    """Blocks of 64 vertices, dense inside (one SCC each), wired forward in
    a chain with a few backward lines that merge some neighbouring blocks."""
    rng = np.random.default_rng(seed)
    nb = n // 64
    b = rng.integers(0, nb, 12 * n)
    u = 64 * b + rng.integers(0, 64, len(b))
    v = 64 * b + rng.integers(0, 64, len(b))
    fwd = rng.random(len(b)) < 0.15
    v[fwd] = np.minimum(n - 1, 64 * (b[fwd] + rng.integers(1, 4, fwd.sum())) + rng.integers(0, 64, fwd.sum()))
    back = rng.random(len(b)) < 0.0005
    v[back] = np.maximum(0, u[back] - 64)
    keep = u != v
    u, v = u[keep], v[keep]
    c = rng.integers(1, 10**6, len(u))
    return LinesCase(n, u.astype(np.int32), v.astype(np.int32), c.astype(np.int64))


@pytest.fixture(scope="module")
def cases():
    log("cases: blocks n=20000, R-MAT 12 (seed 8102), uniform n=10000 m=100000 seed 1001")
    return {"blocks": blocks(), "rmat12": Case(M.Edges.rmat(12, 10**6, 8102), R=64),
            "c1": Case(M.Edges.uniform(10000, 100000, 100, 1001), R=128)}


def queries(which, case, rng):
    scc, _ = case.hg.components()
    if which == "rmat12":
        s0, t0 = case.edges.hub_queries(32, 8103)
        pairs = [case.q(a, b) for a, b in zip(s0, t0)]
    else:
        pairs = [(int(a), int(b)) for a, b in zip(rng.integers(0, case.n, 24), rng.integers(0, case.n, 24))]
    # queries inside components of several vertices
    sizes = np.bincount(scc)
    big = np.flatnonzero(sizes[scc] > 2)
    for x in rng.choice(big, 16):
        ys = np.flatnonzero(scc == scc[x])
        pairs.append((int(x), int(rng.choice(ys))))
    return [(s, t) for s, t in pairs if s != t], scc, sizes


@pytest.mark.parametrize("which", ["blocks", "rmat12", "c1"])
def test_scc_route_plain_and_batched_are_exact(cases, which):
    case = cases[which]
    rng = np.random.default_rng(8104)
    pairs, scc, sizes = queries(which, case, rng)
    want = [O.maxflow(case.og, s, t) for s, t in pairs]
    eng = M.Engine(1, M.default_params(scc_route=1))
    gid = eng.add_graph(case.dg)
    intra = 0
    for (s, t), w in zip(pairs, want):
        log(f"scc_route {which} plain s={s} t={t}")
        r = eng.solve(gid, s, t)
        assert r.flow == w and r.cert_ok == 1 and r.status == 0, (s, t)
        if scc[s] == scc[t]:
            intra += 1
            assert r.kept == sizes[scc[s]]
            if sizes[scc[s]] < 0.995 * case.n:
                assert r.pruned == 1 and r.n_run == sizes[scc[s]]
    assert intra >= 16
    log(f"scc_route {which} batched {len(pairs)} queries")
    beng = M.Engine(8, M.default_params(scc_route=1))
    bgid = beng.add_graph(case.dg)
    for i, (s, t) in enumerate(pairs):
        beng.submit(bgid, s, t, i)
    got = {r.tag: r for r in beng.run()}
    assert [got[i].flow for i in range(len(pairs))] == want
    assert all(got[i].cert_ok == 1 for i in range(len(pairs)))
