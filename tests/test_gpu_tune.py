"""Autotuner on the GPU: picks R, tile size, K and GR_FREQ for a graph class
by timing the plain solver on sample queries, never changes a flow, and
answers the second call for the same class from its cache file."""
import numpy as np
import pytest

import mscresearch_amd as M
import oracle as O
from gpu_util import log

pytestmark = pytest.mark.gpu


def test_autotune_keeps_flows_and_caches(tmp_path):
    e = M.Edges.rmat(12, 10**6, 7001)
    s, t = e.hub_queries(4, 7002)
    cache = str(tmp_path / "tune.txt")
    r = M.autotune(M.HostGraph.from_edgelist(e), s, t, cache=cache)
    log(f"tuned: {r}")
    assert r["evaluations"] > 1
    assert r["R"] in (128, 256, 512, 1024) and r["tile_max"] in (1024, 2048, 4096)
    assert r["K"] in (0, 1, 2, 4, 8) and r["gr_freq"] in (64, 128, 256, 512)
    assert r["ms_per_query"] <= r["default_ms_per_query"]
    assert r["graph_class"].startswith("n12-")
    # the tuned configuration solves exactly
    hg = M.HostGraph.from_edgelist(e).partition(r["R"])
    d = hg.export()
    og = O.Graph(len(d["new_id"]), d["row_ptr"], d["col"], d["cap"], d["rev"])
    eng = M.Engine(1, M.default_params(K=r["K"], gr_freq=r["gr_freq"]))
    gid = eng.add_graph(M.DeviceGraph(hg, r["tile_max"]))
    qs, qt = e.hub_queries(8, 7003)
    for a, b in zip(qs, qt):
        sa, tb = int(d["new_id"][a]), int(d["new_id"][b])
        assert eng.solve(gid, sa, tb).flow == O.maxflow(og, sa, tb)
    # same class again: from the cache, no timing
    r2 = M.autotune(M.HostGraph.from_edgelist(e), s, t, cache=cache)
    assert r2["evaluations"] == 0
    assert (r2["R"], r2["tile_max"], r2["K"], r2["gr_freq"]) == (r["R"], r["tile_max"], r["K"], r["gr_freq"])


def test_autotune_rejects_a_partitioned_graph():
    e = M.Edges.uniform(500, 3000, 100, 7101)
    hg = M.HostGraph.from_edgelist(e).partition(8)
    with pytest.raises(M.MaxFlowError):
        M.autotune(hg, np.array([1], np.int32), np.array([2], np.int32))
