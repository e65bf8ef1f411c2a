"""Helpers shared by the GPU tests: build a graph once on both sides.

The product path (C++ host code + HIP kernels through libmscmf.so) and the
oracle see the same region-ordered CSR, so vertex and entry ids agree.
"""
from __future__ import annotations

import numpy as np

import mscresearch_amd as M
import oracle as O


class Case:
    def __init__(self, edges: M.Edges, R: int = 128, tile_max: int = 4096):
        self.edges = edges
        n, us, vs, cs = edges.arrays()
        self.n = n
        self.us, self.vs, self.cs = us, vs, cs
        self.hg = M.HostGraph.from_edgelist(edges).partition(R)
        d = self.hg.export()
        self.d = d
        self.new_id = d["new_id"]
        self.old_id = d["old_id"]
        self.og = O.Graph(n, d["row_ptr"], d["col"], d["cap"], d["rev"])
        self.E = self.og.E
        self.dg = M.DeviceGraph(self.hg, tile_max)

    def q(self, s, t):
        """Original ids -> region-ordered ids."""
        return int(self.new_id[s]), int(self.new_id[t])

    def oracle_flow(self, s, t):
        return O.maxflow(self.og, s, t)


def engine_with(case: Case, nslots=1, **params):
    e = M.Engine(nslots, M.default_params(**params))
    gid = e.add_graph(case.dg)
    return e, gid


def log(msg):
    print(msg, flush=True)
