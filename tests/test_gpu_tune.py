"""Autotuner on the GPU: picks R, tile size, K, GR_FREQ, the tile discharge
workgroup size and (batched) slots per engine for a graph class by timing
the plain solver or the batched engine on sample queries, never changes a
flow, and answers the second call for the same class and mode from its
cache file."""
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
    assert r["tile_threads"] in (256, 512, 1024) and r["slots"] == 0
    assert r["tile_split"] in (1, 2, 4, 8)
    assert r["ms_per_query"] <= r["default_ms_per_query"]
    assert r["graph_class"].startswith("n12-")
    # the tuned configuration solves exactly
    hg = M.HostGraph.from_edgelist(e).partition(r["R"])
    d = hg.export()
    og = O.Graph(len(d["new_id"]), d["row_ptr"], d["col"], d["cap"], d["rev"])
    eng = M.Engine(1, M.default_params(K=r["K"], gr_freq=r["gr_freq"], tile_threads=r["tile_threads"],
                                       tile_split=r["tile_split"]))
    gid = eng.add_graph(M.DeviceGraph(hg, r["tile_max"]))
    qs, qt = e.hub_queries(8, 7003)
    for a, b in zip(qs, qt):
        sa, tb = int(d["new_id"][a]), int(d["new_id"][b])
        assert eng.solve(gid, sa, tb).flow == O.maxflow(og, sa, tb)
    # same class again: from the cache, no timing
    r2 = M.autotune(M.HostGraph.from_edgelist(e), s, t, cache=cache)
    assert r2["evaluations"] == 0
    assert (r2["R"], r2["tile_max"], r2["K"], r2["gr_freq"], r2["tile_threads"], r2["tile_split"]) == \
        (r["R"], r["tile_max"], r["K"], r["gr_freq"], r["tile_threads"], r["tile_split"])


def test_autotune_batched_engine(tmp_path):
    """Batched mode times the engine with every sample query submitted at
    once and adds slots per engine ({b/2, b, 2b}); its cache entry is its own
    (the plain entry of the same class does not answer it)."""
    e = M.Edges.rmat(12, 10**6, 7011)
    s, t = e.hub_queries(24, 7012)
    cache = str(tmp_path / "tune.txt")
    rp = M.autotune(M.HostGraph.from_edgelist(e), s[:4], t[:4], cache=cache)
    rb = M.autotune(M.HostGraph.from_edgelist(e), s, t, cache=cache, batch=8)
    log(f"batched tuned: {rb}")
    assert rp["evaluations"] > 1 and rb["evaluations"] > 1
    assert rb["slots"] in (4, 8, 16)
    assert rb["ms_per_query"] <= rb["default_ms_per_query"]
    rb2 = M.autotune(M.HostGraph.from_edgelist(e), s, t, cache=cache, batch=8)
    assert rb2["evaluations"] == 0 and rb2["slots"] == rb["slots"]
    # the tuned batched configuration is exact
    hg = M.HostGraph.from_edgelist(e).partition(rb["R"])
    d = hg.export()
    og = O.Graph(len(d["new_id"]), d["row_ptr"], d["col"], d["cap"], d["rev"])
    eng = M.Engine(rb["slots"], M.default_params(K=rb["K"], gr_freq=rb["gr_freq"], tile_threads=rb["tile_threads"],
                                                 tile_split=rb["tile_split"]))
    gid = eng.add_graph(M.DeviceGraph(hg, rb["tile_max"]))
    qs, qt = e.hub_queries(16, 7013)
    pairs = [(int(d["new_id"][a]), int(d["new_id"][b])) for a, b in zip(qs, qt)]
    for i, (a, b) in enumerate(pairs):
        eng.submit(gid, a, b, i)
    for r in eng.run():
        a, b = pairs[r.tag]
        assert r.flow == O.maxflow(og, a, b)


def test_autotune_rejects_a_partitioned_graph():
    e = M.Edges.uniform(500, 3000, 100, 7101)
    hg = M.HostGraph.from_edgelist(e).partition(8)
    with pytest.raises(M.MaxFlowError):
        M.autotune(hg, np.array([1], np.int32), np.array([2], np.int32))
