"""The five benchmark workloads, defined once for the bench, the tests and the
fixture script.

Each workload is a seeded graph from the library's generators (splitmix64 ->
xoshiro256** streams; R-MAT draws fixed 2^20-sample blocks from their own
streams, so the edge list does not depend on the thread count) plus seeded
queries.  Ids here are those of the generated edge list (0-based, before the
region partition renumbers vertices).

  c1  uniform n=10,000, m=100,000, caps U[1,100], 64 queries, R=128
  c2  1024x1024 4-neighbour grid, both directions, caps U[1,1000],
      256 queries col 0 -> col 1023, R in {128, 512}
  c3  R-MAT scale 22 (a,b,c = .57,.19,.19; 16*2^22 samples, duplicates merged
      by summing log-uniform [1,1e6] caps), 1,024 queries: s from the top 1%
      by out-degree, t with in-degree > 0
  r20 R-MAT scale 20 (same generator), the third graph of the serving mix
  c5  R-MAT scale 23 (same generator), 32 queries as in c3, 20 rounds each
      changing the capacity of a uniformly chosen 1% of the input lines by a
      factor U[0.5, 2] (rounded, floored at 1)

The reference takes its inputs as graph.txt / test.txt
(Region_Aware_PR_with_Pruning_BoundaryCore.cu:1044-1126, :2603-2621); these
generators write exactly those files through Edges.write / write_queries.
"""
from __future__ import annotations

import os

import numpy as np

WORKLOADS = {
    "c1": {"kind": "uniform", "n": 10000, "m": 100000, "cap_max": 100, "seed": 1001,
           "queries": {"kind": "uniform", "count": 64, "seed": 1002}, "R": [128]},
    "c2": {"kind": "grid", "rows": 1024, "cols": 1024, "cap_max": 1000, "seed": 2001,
           "queries": {"kind": "grid", "count": 256, "seed": 2002}, "R": [128, 512]},
    "c3": {"kind": "rmat", "scale": 22, "cap_max": 10**6, "seed": 3001,
           "queries": {"kind": "hub", "count": 1024, "seed": 3002}, "R": [512]},
    "r20": {"kind": "rmat", "scale": 20, "cap_max": 10**6, "seed": 4001,
            "queries": {"kind": "hub", "count": 1024, "seed": 4002}, "R": [512]},
    "c5": {"kind": "rmat", "scale": 23, "cap_max": 10**6, "seed": 5001,
           "queries": {"kind": "hub", "count": 32, "seed": 5002}, "R": [512],
           "updates": {"rounds": 20, "frac": 0.01, "seed": 5003}},
}


def gen_threads() -> int:
    """Generator threads: the host's CPU share, at most 16 (the output does not
    depend on it)."""
    try:
        k = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        k = os.cpu_count() or 4
    return max(1, min(16, k))


def edges(name):
    """The workload's generated edge list (mscresearch_amd.Edges)."""
    import mscresearch_amd as M

    w = WORKLOADS[name]
    if w["kind"] == "uniform":
        return M.Edges.uniform(w["n"], w["m"], w["cap_max"], w["seed"])
    if w["kind"] == "grid":
        return M.Edges.grid(w["rows"], w["cols"], w["cap_max"], w["seed"])
    return M.Edges.rmat(w["scale"], w["cap_max"], w["seed"], threads=gen_threads())


def queries(name, e, count=None):
    """(s, t) of the workload in edge-list ids; the first `count` of a longer
    draw are the same pairs (one sequential stream)."""
    import mscresearch_amd as M

    w = WORKLOADS[name]
    q = w["queries"]
    k = q["count"] if count is None else count
    if q["kind"] == "uniform":
        return M.uniform_queries(w["n"], k, q["seed"])
    if q["kind"] == "grid":
        return M.grid_queries(w["rows"], w["cols"], k, q["seed"])
    return e.hub_queries(k, q["seed"])


def edge_hash(us, vs, caps) -> str:
    """64-bit fingerprint of an edge list (order-sensitive), hex."""
    us = np.asarray(us, np.uint64)
    vs = np.asarray(vs, np.uint64)
    cs = np.asarray(caps, np.int64).view(np.uint64)
    i = np.arange(len(us), dtype=np.uint64)
    with np.errstate(over="ignore"):
        x = (us * np.uint64(0x9E3779B97F4A7C15)) ^ (vs * np.uint64(0xC2B2AE3D27D4EB4F)) ^ \
            (cs * np.uint64(0x165667B19E3779F9)) ^ (i * np.uint64(0x27D4EB2F165667C5))
        x ^= x >> np.uint64(29)
        x *= np.uint64(0xBF58476D1CE4E5B9)
        h = int(np.bitwise_xor.reduce(x)) if len(x) else 0
        h ^= int(x.sum(dtype=np.uint64)) if len(x) else 0
    return f"{h & 0xFFFFFFFFFFFFFFFF:016x}"


def capacity_update_rounds(caps, rounds, seed, frac=0.01):
    """Yield (lines, new_caps) for each round: floor(frac * m) distinct input
    lines chosen uniformly get cap <- max(1, rint(cap * U[0.5, 2])); caps
    compound over the rounds.  `lines` ascending, int64."""
    rng = np.random.default_rng(seed)
    cur = np.array(caps, dtype=np.int64, copy=True)
    m = len(cur)
    k = max(1, int(m * frac))
    for _ in range(rounds):
        idx = np.sort(rng.choice(m, k, replace=False)).astype(np.int64)
        f = rng.uniform(0.5, 2.0, k)
        new = np.maximum(1, np.rint(cur[idx].astype(np.float64) * f)).astype(np.int64)
        cur[idx] = new
        yield idx, new


def fixture_path():
    return os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests", "golden",
                        "config_flows.json")
