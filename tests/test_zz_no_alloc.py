"""mf_solve_noalloc (no device allocation: the query's slab sits in a static
per-device array) against mf_solve, the entry point it is a variant of, and
the oracle: the same int64 flows, certified, on two calls in a row, the
second on a smaller graph; a graph whose slab does not fit is refused."""
import numpy as np
import pytest

import mscresearch_amd as M
import oracle as O
from gpu_util import Case, engine_with, log

pytestmark = pytest.mark.gpu


def test_noalloc_equals_solve_then_smaller_graph():
    big = Case(M.Edges.grid(256, 256, 1000, 8101), R=64, tile_max=1024)
    small = Case(M.Edges.uniform(2000, 20000, 100, 8102), R=16)
    assert big.dg.slab_bytes() > small.dg.slab_bytes()
    for case, (qs, qt) in ((big, M.grid_queries(256, 256, 3, 8103)), (small, M.uniform_queries(2000, 3, 8104))):
        assert case.dg.slab_bytes() <= M.NOALLOC_SCRATCH_BYTES
        eng, gid = engine_with(case)
        for a, b in zip(qs, qt):
            s, t = case.q(a, b)
            log(f"noalloc n={case.n} s={s} t={t}")
            want = case.oracle_flow(s, t)
            r0 = eng.solve(gid, s, t)
            r1 = eng.solve_noalloc(gid, s, t)
            assert r1.status == 0 and r1.cert_ok == 1
            assert r1.flow == r0.flow == want
            assert r1.pruned == 0 and r1.n_run == case.n
            # the engine's own slab is back in place for ordinary solves
            assert eng.solve(gid, s, t).flow == want


def test_noalloc_refuses_a_graph_beyond_the_scratch():
    e = M.Edges.uniform(1 << 20, 1 << 24, 100, 8105)
    hg = M.HostGraph.from_edgelist(e).partition(64)
    dg = M.DeviceGraph(hg)
    assert dg.slab_bytes() > M.NOALLOC_SCRATCH_BYTES
    eng = M.Engine(1)
    gid = eng.add_graph(dg)
    with pytest.raises(M.MaxFlowError) as ei:
        eng.solve_noalloc(gid, 0, 1)
    assert ei.value.code == M.MF_ENOMEM
    assert "bytes" in str(ei.value)
