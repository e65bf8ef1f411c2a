"""GPU parity of the deterministic pieces, bit-exact against the oracle.

Residual init + preflow (BC:151-200), global-relabel BFS heights
(BC:203-233, BC:322-348), the pruning mask (BC:1391-1445) and the induced
subgraph (BC:2477-2551) are deterministic, so the HIP results must equal the
oracle's array for array.
"""
import numpy as np
import pytest

import mscresearch_amd as M
import oracle as O
from gpu_util import Case, engine_with, log

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1():
    log("case: uniform n=10000 m=100000 seed 1001, R=128")
    return Case(M.Edges.uniform(10000, 100000, 100, 1001), R=128)


@pytest.fixture(scope="module")
def rmat12():
    log("case: rmat scale 12 seed 1201, R=64")
    return Case(M.Edges.rmat(12, 10**6, 1201), R=64)


@pytest.fixture(scope="module")
def grid():
    log("case: grid 48x64 seed 1301, R=32")
    return Case(M.Edges.grid(48, 64, 1000, 1301), R=32, tile_max=256)


def test_init_preflow_bit_exact(c1):
    e, gid = engine_with(c1)
    s_, t_ = M.uniform_queries(c1.n, 8, 1002)
    for a, b in zip(s_, t_):
        s, t = c1.q(a, b)
        res, ex = e.dbg_init_preflow(gid, s, t, c1.n, c1.E)
        ores, oex = O.init_preflow(c1.og, s)
        assert np.array_equal(res, ores)
        assert np.array_equal(ex, oex)


@pytest.mark.parametrize("which", ["c1", "rmat12", "grid"])
def test_global_relabel_heights_bit_exact(which, request):
    case = request.getfixturevalue(which)
    e, gid = engine_with(case)
    rng = np.random.default_rng(77)
    if which == "rmat12":
        s_, t_ = case.edges.hub_queries(4, 1202)
    elif which == "grid":
        s_, t_ = M.grid_queries(48, 64, 4, 1302)
    else:
        s_, t_ = M.uniform_queries(case.n, 4, 1003)
    for a, b in zip(s_, t_):
        s, t = case.q(a, b)
        log(f"GR {which} s={s} t={t}")
        res, ex = O.init_preflow(case.og, s)
        h = e.dbg_global_relabel(gid, s, t, res, ex)
        assert np.array_equal(h, O.bfs_heights(case.og, res, ex, s, t))
        # a later residual (the oracle's maximum flow) plus planted deficits
        f, res2 = O.maxflow(case.og, s, t, want_residual=True)
        ex2 = np.zeros(case.n, np.int64)
        pick = rng.choice(case.n, 5, replace=False)
        ex2[pick] = -3
        h2 = e.dbg_global_relabel(gid, s, t, res2, ex2)
        assert np.array_equal(h2, O.bfs_heights(case.og, res2, ex2, s, t))


@pytest.mark.parametrize("which", ["c1", "rmat12", "grid"])
def test_prune_mask_bit_exact(which, request):
    case = request.getfixturevalue(which)
    e, gid = engine_with(case)
    if which == "rmat12":
        s_, t_ = case.edges.hub_queries(6, 1203)
    elif which == "grid":
        s_, t_ = M.grid_queries(48, 64, 3, 1303)
    else:
        s_, t_ = M.uniform_queries(case.n, 3, 1004)
    for a, b in zip(s_, t_):
        s, t = case.q(a, b)
        mask, kept = e.dbg_prune_mask(gid, s, t, case.n)
        om = O.prune_mask(case.og, s, t)
        assert np.array_equal(mask, om)
        assert kept == int(om.sum())


def test_induced_subgraph_bit_exact(rmat12):
    e, gid = engine_with(rmat12)
    s_, t_ = rmat12.edges.hub_queries(3, 1204)
    rng = np.random.default_rng(5)
    masks = [O.prune_mask(rmat12.og, *rmat12.q(a, b)) for a, b in zip(s_, t_)]
    masks.append((rng.random(rmat12.n) < 0.5).astype(np.int32))
    for mask in masks:
        got = e.dbg_induced(gid, mask, rmat12.E)
        sub, o2s, _ = O.induced(rmat12.og, mask)
        assert got["n"] == sub.n and got["E"] == sub.E
        for k in ("row_ptr", "col", "cap", "rev"):
            assert np.array_equal(got[k], getattr(sub, k)), k
        assert np.array_equal(got["old2sub"], o2s)
