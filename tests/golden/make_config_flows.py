"""Oracle max-flow fixtures for the benchmark workloads at their stated sizes.

Run in a CPU container (no GPU needed), about an hour on 8 cores:

    python tests/golden/make_config_flows.py [--threads N] [--only c2,c3,r20,c5]

Writes tests/golden/config_flows.json (merging into an existing file, one
workload at a time).  Each graph comes from the library's seeded generators
(mscresearch_amd.workloads); the CSR is the oracle's own (orc_build_csr, the
reference's construction Region_Aware_PR_with_Pruning_BoundaryCore.cu:1078-1125)
and every flow is the oracle's exact int64 Dinic on it.  Nothing from the
GPU path is used.  An order-sensitive hash of the edge list is recorded so a
consumer can tell a changed generator from a wrong flow.

c5 (dynamic): flows[0] are the 32 queries on the generated graph, flows[r]
after the r-th round of capacity_update_rounds (lines changed in place, the
CSR forward entry of each line found by orc_line_entries).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from mscresearch_amd import workloads as W  # noqa: E402


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make(name, threads):
    w = W.WORKLOADS[name]
    t0 = time.time()
    e = W.edges(name)
    n, us, vs, cs = e.arrays()
    qs, qt = W.queries(name, e)
    del e
    g = O.build_csr(n, us, vs, cs)
    log(f"{name}: n={n} lines={len(us)} E={g.E} built in {time.time() - t0:.1f}s")
    rec = {"workload": {k: v for k, v in w.items()}, "n": n, "lines": int(len(us)), "E": int(g.E),
           "edge_hash": W.edge_hash(us, vs, cs), "s": [int(x) for x in qs], "t": [int(x) for x in qt]}
    t1 = time.time()
    f, sec = O.maxflow_many(g, qs, qt, threads=threads)
    log(f"{name}: {len(qs)} flows in {sec:.1f}s")
    if "updates" not in w:
        rec["flows"] = [int(x) for x in f]
    else:
        u = w["updates"]
        ent = O.line_entries(g, us, vs)
        flows = [[int(x) for x in f]]
        for r, (lines, new) in enumerate(W.capacity_update_rounds(cs, u["rounds"], u["seed"], u["frac"])):
            g.cap[ent[lines]] = new
            f, sec = O.maxflow_many(g, qs, qt, threads=threads)
            flows.append([int(x) for x in f])
            log(f"{name}: round {r + 1}: {len(lines)} updates, flows in {sec:.1f}s")
        rec["flows"] = flows
    rec["seconds"] = round(time.time() - t1, 1)
    return rec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 4)
    ap.add_argument("--only", default="c2,c3,r20,c5")
    ap.add_argument("--out", default=W.fixture_path())
    args = ap.parse_args()
    data = {}
    if os.path.exists(args.out):
        with open(args.out) as f:
            data = json.load(f)
    for name in args.only.split(","):
        data[name] = make(name, args.threads)
        tmp = args.out + ".tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, separators=(",", ":"))
        os.replace(tmp, args.out)
        log(f"{name}: written to {args.out}")


if __name__ == "__main__":
    main()
