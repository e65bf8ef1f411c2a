"""Writes the golden max-flow fixtures under tests/golden/.

The reference ships no tests, data or result files for its max-flow path, so
the fixtures come from independent in-container libraries: networkx
``maximum_flow_value`` (exact Python integers) and scipy
``csgraph.maximum_flow`` (int32 capacities, used where every capacity sum fits).
Graphs are produced by the package's seeded generators and written as the
reference's 1-based ``graph.txt`` format; each fixture records the SHA-256 of
that text so a changed generator is detected, plus the queries and flows.

Run once:  PYTHONPATH=. python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import tempfile

import networkx as nx
import numpy as np
import scipy.sparse as sp
from scipy.sparse.csgraph import maximum_flow

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import mscresearch_amd as M  # noqa: E402


def graph_text_sha(e: M.Edges) -> str:
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "graph.txt")
        e.write(p)
        return hashlib.sha256(open(p, "rb").read()).hexdigest()


def nx_flows(n, us, vs, cs, qs, qt):
    G = nx.DiGraph()
    G.add_nodes_from(range(n))
    for u, v, c in zip(us.tolist(), vs.tolist(), cs.tolist()):
        if G.has_edge(u, v):
            G[u][v]["capacity"] += c
        else:
            G.add_edge(u, v, capacity=c)
    return [int(nx.maximum_flow_value(G, int(a), int(b))) for a, b in zip(qs, qt)]


def scipy_flows(n, us, vs, cs, qs, qt):
    if int(cs.sum()) >= 2**31:
        return None
    A = sp.coo_matrix((cs.astype(np.int32), (us, vs)), shape=(n, n)).tocsr()  # duplicates summed
    A.sum_duplicates()
    return [int(maximum_flow(A, int(a), int(b)).flow_value) for a, b in zip(qs, qt)]


def fixture(name, e: M.Edges, qs, qt, note):
    n, us, vs, cs = e.arrays()
    f_nx = nx_flows(n, us, vs, cs, qs, qt)
    f_sp = scipy_flows(n, us, vs, cs, qs, qt)
    if f_sp is not None:
        assert f_sp == f_nx, (name, f_sp, f_nx)
    return {
        "name": name, "note": note, "n": int(n), "lines": int(len(us)),
        "graph_txt_sha256": graph_text_sha(e),
        "queries": [[int(a), int(b)] for a, b in zip(qs, qt)],
        "flows": f_nx, "scipy_agrees": f_sp is not None,
    }


def main():
    out = []
    # small uniform graphs, several densities
    for i, (n, m, cap) in enumerate([(12, 40, 9), (40, 200, 20), (150, 1200, 100), (300, 3000, 100)]):
        e = M.Edges.uniform(n, m, cap, 7000 + i)
        qs, qt = M.uniform_queries(n, 16, 7100 + i)
        out.append(fixture(f"uniform_n{n}_m{m}", e, qs, qt, "uniform random digraph"))
    # small grids (left column -> right column queries)
    for i, (r, c) in enumerate([(4, 6), (16, 16), (32, 48)]):
        e = M.Edges.grid(r, c, 1000, 7200 + i)
        qs, qt = M.grid_queries(r, c, 12, 7300 + i)
        out.append(fixture(f"grid_{r}x{c}", e, qs, qt, "4-neighbour grid, both directions"))
    # small R-MAT graphs with hub sources
    for i, sc in enumerate([8, 10]):
        e = M.Edges.rmat(sc, 10**6, 7400 + i)
        qs, qt = e.hub_queries(16, 7500 + i)
        out.append(fixture(f"rmat_s{sc}", e, qs, qt, "R-MAT a=.57 b=.19 c=.19, log-uniform caps"))
    # the uniform n=10,000 m=100,000 workload (seeds 1001/1002)
    e = M.Edges.uniform(10000, 100000, 100, 1001)
    qs, qt = M.uniform_queries(10000, 64, 1002)
    out.append(fixture("uniform_n10000_m100000", e, qs, qt, "uniform n=1e4 m=1e5 caps U[1,100], 64 queries"))
    path = os.path.join(HERE, "maxflow_golden.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(path, sum(len(x["flows"]) for x in out), "flows")


if __name__ == "__main__":
    main()
