"""Autotune a benchmark workload and record the result in a cache file.

    python -m mscresearch_amd.tune --workload c3 --batch 96 --queries 384 \\
        --cache mscresearch_amd/tuned_params.txt

Runs mf_autotune (the library's coordinate descent over R, tile size, K,
GR_FREQ, the tile discharge workgroup size, workgroups per tile by excess
score and, with --batch, slots per engine) on the workload's first --queries queries, and appends the chosen
values for the graph's class to the cache.  The bench reads that file: a
class found there is not timed again.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

import mscresearch_amd as M
from mscresearch_amd import workloads as W

DEFAULT_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned_params.txt")


def cached(hg_unpartitioned, batch, cache=DEFAULT_CACHE):
    """The cache's entry for this graph's class and mode (the last one), or
    None: {"R", "tile_max", "K", "gr_freq", "tile_threads", "slots", "tile_split",
    "ms_per_query", "default_ms_per_query", "graph_class"}."""
    if not cache or not os.path.exists(cache):
        return None
    cls = hg_unpartitioned.graph_class()
    key = cls + (f":b{batch}" if batch > 0 else ":plain")
    found = None
    with open(cache) as f:
        for line in f:
            p = line.split()
            if len(p) in (9, 10) and p[0] == key:
                found = p
    if found is None:
        return None
    names = ["R", "tile_max", "K", "gr_freq", "tile_threads", "slots"]
    out = {k: int(v) for k, v in zip(names, found[1:7])}
    out.update(tile_split=int(found[9]) if len(found) == 10 else 1, ms_per_query=float(found[7]),
               default_ms_per_query=float(found[8]), graph_class=cls, evaluations=0)
    return out


def tuned(name, e, batch, cache=DEFAULT_CACHE, queries=384, R0=512, tile0=4096, params=None):
    """The cached (or, when the class is missing, freshly tuned) values for
    workload `name` with generated edges `e`."""
    s, t = W.queries(name, e, count=max(queries, W.WORKLOADS[name]["queries"]["count"]))
    s, t = s[:queries], t[:queries]
    hg = M.HostGraph.from_edgelist(e)
    return M.autotune(hg, np.asarray(s, np.int32), np.asarray(t, np.int32), params=params, R=R0, tile=tile0,
                      cache=cache, batch=batch)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="c3", choices=sorted(W.WORKLOADS))
    ap.add_argument("--batch", type=int, default=96, help="slots per engine to start from (0 = plain solver)")
    ap.add_argument("--queries", type=int, default=384)
    ap.add_argument("--cache", default=DEFAULT_CACHE)
    args = ap.parse_args()
    t0 = time.time()
    e = W.edges(args.workload)
    r = tuned(args.workload, e, args.batch, args.cache, args.queries)
    r["workload"] = args.workload
    r["seconds"] = round(time.time() - t0, 1)
    print(json.dumps(r), flush=True)


if __name__ == "__main__":
    sys.exit(main())
