"""CPU oracle (test infrastructure only).

A ctypes wrapper over ``oracle/liboracle.so`` (oracle.c), the CPU restatement
of the reference's max-flow path plus an exact Dinic solver and a min-cut
certificate checker.  Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py use it, and only as the checker or the timed
CPU baseline; the product (mscresearch_amd) never imports it.

Reference citations ("BC:line") are to Region_Aware_PR_with_Pruning_BoundaryCore.cu.
See oracle.h for what pins each function.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_DIR, "liboracle.so")
_lib = None


def build() -> str:
    """Build liboracle.so, and the reference's SCC tools into _ref/ when their sources are present."""
    subprocess.run(["make", "-s", "-C", _DIR], check=True)
    if os.path.isdir("/root/reference/SCC_CONTRACT"):
        subprocess.run(["make", "-s", "-C", _DIR, "ref"], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.orc_parse_header.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int64)]
        L.orc_parse_edges.argtypes = [C.c_char_p, C.c_int64, P, P, P]
        L.orc_parse_edges.restype = C.c_int64
        L.orc_parse_queries.argtypes = [C.c_char_p, C.c_int64, P, P]
        L.orc_parse_queries.restype = C.c_int64
        L.orc_build_csr.argtypes = [C.c_int, C.c_int64, P, P, P, P, P, P, P, P, P, P, P]
        L.orc_line_entries.argtypes = [C.c_int, C.c_int64, P, P, P, P]
        L.orc_csr_stats.argtypes = [C.c_int, P, P, P, P, P, P, P]
        L.orc_partition.argtypes = [C.c_int, C.c_int, P, P, P, P, P, P]
        L.orc_reorder.argtypes = [C.c_int, C.c_int, P, P, P, P, P, P, P, P, P, P, P, P, P, P]
        L.orc_prune_mask.argtypes = [C.c_int, P, P, P, C.c_int, C.c_int, P]
        L.orc_induced.argtypes = [C.c_int, P, P, P, P, P, P, P, P, P, P, P, P]
        L.orc_induced.restype = C.c_int
        L.orc_core_mask.argtypes = [C.c_int, C.c_int, P, P, P, C.c_int, C.c_int, C.c_int, P, P]
        L.orc_init_residual.argtypes = [C.c_int, C.c_int64, P, P, P]
        L.orc_preflow.argtypes = [C.c_int, P, P, P, P, P]
        L.orc_bfs_heights.argtypes = [C.c_int, P, P, P, P, P, P, C.c_int, C.c_int, P]
        L.orc_bfs_heights.restype = C.c_int64
        L.orc_dinic.argtypes = [C.c_int, P, P, P, P, C.c_int, C.c_int, P]
        L.orc_dinic.restype = C.c_int64
        L.orc_certify.argtypes = [C.c_int, P, P, P, P, P, P, C.c_int, C.c_int, C.c_int64]
        L.orc_certify.restype = C.c_int
        L.orc_dinic_many.argtypes = [C.c_int, P, P, P, P, C.c_int, P, P, C.c_int, P]
        L.orc_dinic_many.restype = C.c_double
        L.orc_dinic_many_ex.argtypes = [C.c_int, P, P, P, P, C.c_int, P, P, C.c_int, C.c_int, P]
        L.orc_dinic_many_ex.restype = C.c_double
        L.orc_scc_kosaraju.argtypes = [C.c_int, C.c_int64, P, P, P]
        L.orc_scc_kosaraju.restype = C.c_int
        L.orc_scc_contract.argtypes = [C.c_int64, P, P, P, P, P, P, P]
        L.orc_scc_contract.restype = C.c_int64
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


class Graph:
    """Oracle CSR: row_ptr, col, cap, rev (+ stats for the input-line variant)."""

    def __init__(self, n, row_ptr, col, cap, rev, stats=None):
        self.n = int(n)
        self.row_ptr = _c(row_ptr, np.int32)
        self.col = _c(col, np.int32)
        self.cap = _c(cap, np.int64)
        self.rev = _c(rev, np.int32)
        self.stats = stats

    @property
    def E(self):
        return len(self.col)


def parse_graph(path) -> tuple[int, np.ndarray, np.ndarray, np.ndarray]:
    """BC:1044-1076 restated.  Returns (n, us, vs, caps) of accepted lines."""
    n, m = C.c_int(0), C.c_int64(0)
    rc = lib().orc_parse_header(str(path).encode(), C.byref(n), C.byref(m))
    if rc != 0:
        raise IOError(f"oracle: cannot parse header of {path} ({rc})")
    mm = max(int(m.value), 0)
    us = np.empty(mm, np.int32)
    vs = np.empty(mm, np.int32)
    cs = np.empty(mm, np.int64)
    cnt = lib().orc_parse_edges(str(path).encode(), mm, _p(us), _p(vs), _p(cs))
    if cnt < 0:
        raise IOError(f"oracle: cannot parse {path}")
    return n.value, us[:cnt].copy(), vs[:cnt].copy(), cs[:cnt].copy()


def parse_queries(path, cap=1 << 24):
    s = np.empty(cap, np.int32)
    t = np.empty(cap, np.int32)
    q = lib().orc_parse_queries(str(path).encode(), cap, _p(s), _p(t))
    if q < 0:
        raise IOError(f"oracle: cannot open {path}")
    return s[:q].copy(), t[:q].copy()


def build_csr(n, us, vs, caps) -> Graph:
    """BC:1078-1125 restated."""
    us, vs, caps = _c(us, np.int32), _c(vs, np.int32), _c(caps, np.int64)
    cnt = len(us)
    rp = np.empty(n + 1, np.int32)
    col = np.empty(2 * cnt, np.int32)
    cap = np.empty(2 * cnt, np.int64)
    rev = np.empty(2 * cnt, np.int32)
    so, si = np.empty(n, np.int64), np.empty(n, np.int64)
    do, di = np.empty(n, np.int32), np.empty(n, np.int32)
    lib().orc_build_csr(n, cnt, _p(us), _p(vs), _p(caps), _p(rp), _p(col), _p(cap), _p(rev), _p(so), _p(si),
                        _p(do), _p(di))
    return Graph(n, rp, col, cap, rev, stats=(so, si, do, di))


def line_entries(g: Graph, us, vs):
    """Forward entry of every accepted line in build_csr's layout."""
    us, vs = _c(us, np.int32), _c(vs, np.int32)
    out = np.empty(len(us), np.int32)
    lib().orc_line_entries(g.n, len(us), _p(us), _p(vs), _p(g.row_ptr), _p(out))
    return out


def csr_stats(g: Graph):
    n = g.n
    so, si = np.empty(n, np.int64), np.empty(n, np.int64)
    do, di = np.empty(n, np.int32), np.empty(n, np.int32)
    lib().orc_csr_stats(n, _p(g.row_ptr), _p(g.col), _p(g.cap), _p(so), _p(si), _p(do), _p(di))
    return so, si, do, di


def partition(n, R, so, si, do, di):
    """BC:1130-1178 restated (ties by id).  Returns (part, score)."""
    part = np.empty(n, np.int32)
    score = np.empty(n, np.float64)
    lib().orc_partition(n, R, _p(_c(so, np.int64)), _p(_c(si, np.int64)), _p(_c(do, np.int32)),
                        _p(_c(di, np.int32)), _p(score), _p(part))
    return part, score


def reorder(g: Graph, R, part):
    """BC:1182-1276 restated.  Returns (Graph', new_id, old_id, rstart, rend, region_id)."""
    n, E = g.n, g.E
    new_id, old_id = np.empty(n, np.int32), np.empty(n, np.int32)
    rstart, rend = np.empty(R, np.int32), np.empty(R, np.int32)
    rid = np.empty(n, np.int32)
    rp2 = np.empty(n + 1, np.int32)
    col2, cap2, rev2 = np.empty(E, np.int32), np.empty(E, np.int64), np.empty(E, np.int32)
    lib().orc_reorder(n, R, _p(_c(part, np.int32)), _p(g.row_ptr), _p(g.col), _p(g.cap), _p(g.rev), _p(new_id),
                      _p(old_id), _p(rstart), _p(rend), _p(rid), _p(rp2), _p(col2), _p(cap2), _p(rev2))
    return Graph(n, rp2, col2, cap2, rev2), new_id, old_id, rstart, rend, rid


def prune_mask(g: Graph, s, t):
    """BC:1391-1445 restated."""
    mask = np.empty(g.n, np.int32)
    lib().orc_prune_mask(g.n, _p(g.row_ptr), _p(g.col), _p(g.cap), int(s), int(t), _p(mask))
    return mask


def induced(g: Graph, mask):
    """BC:2477-2551 restated.  Returns (Graph', old2sub, sub2old)."""
    n, E = g.n, g.E
    o2s = np.empty(n, np.int32)
    s2o = np.empty(n, np.int32)
    rp2 = np.empty(n + 1, np.int32)
    col2, cap2, rev2 = np.empty(max(E, 1), np.int32), np.empty(max(E, 1), np.int64), np.empty(max(E, 1), np.int32)
    m = C.c_int64(0)
    k = lib().orc_induced(n, _p(g.row_ptr), _p(g.col), _p(g.cap), _p(g.rev), _p(_c(mask, np.int32)), _p(o2s),
                          _p(s2o), _p(rp2), _p(col2), _p(cap2), _p(rev2), C.byref(m))
    mm = m.value
    return Graph(k, rp2[: k + 1], col2[:mm], cap2[:mm], rev2[:mm]), o2s, s2o[:k].copy()


def core_mask(g: Graph, R, region_id, s, t, hop_limit=2):
    """BC:2340-2475 restated (the reference's boundary-aware core)."""
    mask = np.empty(g.n, np.int32)
    kept = C.c_int64(0)
    lib().orc_core_mask(g.n, R, _p(g.row_ptr), _p(g.col), _p(_c(region_id, np.int32)), int(s), int(t),
                        hop_limit, _p(mask), C.byref(kept))
    return mask, kept.value


def init_preflow(g: Graph, s):
    """BC:151-200 restated: residual init then saturate the row of s."""
    res = np.empty(g.E, np.int64)
    ex = np.empty(g.n, np.int64)
    lib().orc_init_residual(g.n, g.E, _p(g.cap), _p(res), _p(ex))
    lib().orc_preflow(int(s), _p(g.row_ptr), _p(g.col), _p(g.rev), _p(res), _p(ex))
    return res, ex


def bfs_heights(g: Graph, res, excess, s, t, mask=None):
    """BC:203-233 + BC:322-348 restated (s never labelled)."""
    h = np.empty(g.n, np.int32)
    lib().orc_bfs_heights(g.n, _p(g.row_ptr), _p(g.col), _p(g.rev), _p(_c(res, np.int64)),
                          _p(_c(excess, np.int64)), _p(None if mask is None else _c(mask, np.int32)), int(s),
                          int(t), _p(h))
    return h


def maxflow(g: Graph, s, t, want_residual=False):
    """Exact int64 max flow (Dinic)."""
    res = np.empty(g.E, np.int64) if want_residual else None
    f = lib().orc_dinic(g.n, _p(g.row_ptr), _p(g.col), _p(g.cap), _p(g.rev), int(s), int(t), _p(res))
    return (f, res) if want_residual else f


def certify(g: Graph, res, s, t, flow, excess=None) -> int:
    """Min-cut certificate; 0 = the residual proves `flow` is maximum."""
    return lib().orc_certify(g.n, _p(g.row_ptr), _p(g.col), _p(g.cap), _p(g.rev), _p(_c(res, np.int64)),
                             _p(None if excess is None else _c(excess, np.int64)), int(s), int(t), int(flow))


def maxflow_many(g: Graph, s, t, threads, prune=False):
    """Dinic over many queries, one query per thread; prune=True restricts each
    query to its pruning mask (BC:1391-1445) first.  Returns (flows, seconds)."""
    s, t = _c(s, np.int32), _c(t, np.int32)
    flows = np.empty(len(s), np.int64)
    sec = lib().orc_dinic_many_ex(g.n, _p(g.row_ptr), _p(g.col), _p(g.cap), _p(g.rev), len(s), _p(s), _p(t),
                                  int(threads), 1 if prune else 0, _p(flows))
    return flows, sec


def scc_kosaraju(n, us, vs):
    """SCC_CONTRACT/scc_contract.cpp:44-105 restated: (component ids, count)."""
    us, vs = _c(us, np.int32), _c(vs, np.int32)
    comp = np.empty(n, np.int32)
    C_ = lib().orc_scc_kosaraju(int(n), len(us), _p(us), _p(vs), _p(comp))
    return comp, C_


def scc_contract(us, vs, caps, comp):
    """scc_contract.cpp:107-125 restated: (cu, cv, cap) sorted by (cu, cv)."""
    us, vs, caps = _c(us, np.int32), _c(vs, np.int32), _c(caps, np.int64)
    m = len(us)
    cu, cv, cc = np.empty(max(m, 1), np.int32), np.empty(max(m, 1), np.int32), np.empty(max(m, 1), np.int64)
    k = lib().orc_scc_contract(m, _p(us), _p(vs), _p(caps), _p(_c(comp, np.int32)), _p(cu), _p(cv), _p(cc))
    return cu[:k].copy(), cv[:k].copy(), cc[:k].copy()
