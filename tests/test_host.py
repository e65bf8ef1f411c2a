"""Product host code (C++ behind the C ABI) against the oracle, on the CPU.

Parsing, CSR, the capacity-degree partition, the region reorder, query
files and the seeded generators are deterministic integer (or
deterministic double) work, so they must match the oracle exactly.
"""
import hashlib
import os
import subprocess

import numpy as np
import pytest

import mscresearch_amd as M
import oracle as O


def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


GRAPH_FIXTURES = {
    "skip_rules": "4 9\n1 2 5\n1 5 3\n2 2 4\n2 3 0\n3 4 -3\n3 4 7\n0 1 2\n4 1 1\n2 x 9\n1 3 8\n",
    "short": "3 10\n1 2 1\n2 3 1\n",
    "m_limits": "3 2\n1 2 1\n2 3 1\n1 3 1\n",
    "whitespace": "  5\t6\n\n1 2 3\r\n2 3 4\n  3 4 +5\n4 5 6 7\n5 1 9\n",
    "dups_antiparallel": "4 6\n1 2 3\n1 2 4\n2 1 5\n2 3 6\n3 4 1\n4 3 2\n",
    "empty_body": "7 0\n",
    "truncated": "3 3\n1 2 4\n2 3\n",
}


@pytest.mark.parametrize("name", sorted(GRAPH_FIXTURES))
def test_parse_and_csr_match_oracle(tmp_path, name):
    p = write(tmp_path, "graph.txt", GRAPH_FIXTURES[name])
    n, us, vs, cs = O.parse_graph(p)
    og = O.build_csr(n, us, vs, cs)
    hg = M.HostGraph.load(p)
    nn, E, R, lines = hg.info()
    assert (nn, E, lines) == (n, 2 * len(us), len(us))
    d = hg.export()
    for k in ("row_ptr", "col", "cap", "rev"):
        assert np.array_equal(d[k], getattr(og, k)), k
    for a, b in zip(hg.stats(), og.stats):
        assert np.array_equal(a, b)


def test_parse_errors(tmp_path):
    with pytest.raises(M.MaxFlowError):
        M.HostGraph.load(str(tmp_path / "missing.txt"))
    with pytest.raises(M.MaxFlowError):
        M.HostGraph.load(write(tmp_path, "bad.txt", "n m\n1 2 3\n"))


def test_query_file(tmp_path):
    p = write(tmp_path, "test.txt", "1 2\n3 4\n0 9\n5 5\n7\n")
    s, t = M.read_queries(p)
    os_, ot = O.parse_queries(p)
    assert s.tolist() == os_.tolist() == [0, 2, -1, 4]
    assert t.tolist() == ot.tolist() == [1, 3, 8, 4]
    s, t = M.read_queries(write(tmp_path, "empty.txt", ""))
    assert len(s) == 0


@pytest.mark.parametrize("gen", ["uniform", "grid", "rmat"])
@pytest.mark.parametrize("R", [1, 7, 128])
def test_partition_and_reorder_match_oracle(gen, R):
    if gen == "uniform":
        e = M.Edges.uniform(5000, 40000, 100, 17)
    elif gen == "grid":
        e = M.Edges.grid(40, 50, 1000, 18)
    else:
        e = M.Edges.rmat(11, 10**6, 19)
    n, us, vs, cs = e.arrays()
    og = O.build_csr(n, us, vs, cs)
    part_o, _ = O.partition(n, R, *og.stats)
    assert np.array_equal(M.partition(R, *og.stats), part_o)
    hg = M.HostGraph.from_edgelist(e).partition(R)
    d = hg.export()
    g2, new_id, old_id, rs, re_, rid = O.reorder(og, R, part_o)
    for k in ("row_ptr", "col", "cap", "rev"):
        assert np.array_equal(d[k], getattr(g2, k)), k
    assert np.array_equal(d["new_id"], new_id)
    assert np.array_equal(d["old_id"], old_id)
    assert np.array_equal(d["rstart"][:-1], rs) and np.array_equal(d["rstart"][1:], re_)
    assert np.array_equal(d["region_id"], rid)


def test_partition_stats_after_prune_variant():
    # per-query stats over CSR entries with cap > 0 (BC:2707-2720) equal the
    # input-line stats on an unpruned graph
    e = M.Edges.uniform(800, 5000, 50, 23)
    n, us, vs, cs = e.arrays()
    og = O.build_csr(n, us, vs, cs)
    for a, b in zip(O.csr_stats(og), og.stats):
        assert np.array_equal(a, b)


def test_generators_deterministic_and_shaped(tmp_path):
    a = M.Edges.uniform(1000, 5000, 100, 5).arrays()
    b = M.Edges.uniform(1000, 5000, 100, 5).arrays()
    c = M.Edges.uniform(1000, 5000, 100, 6).arrays()
    for x, y in zip(a[1:], b[1:]):
        assert np.array_equal(x, y)
    assert not np.array_equal(a[1], c[1])
    n, us, vs, cs = a
    assert np.all(us != vs) and cs.min() >= 1 and cs.max() <= 100 and len(us) <= 5000
    # grid: both directions of every 4-neighbour pair
    n, us, vs, cs = M.Edges.grid(6, 9, 1000, 7).arrays()
    assert n == 54 and len(us) == 2 * (6 * 8 + 9 * 5)
    pairs = set(zip(us.tolist(), vs.tolist()))
    for u, v in list(pairs):
        assert (v, u) in pairs
        assert abs(u - v) in (1, 9)
    assert cs.min() >= 1 and cs.max() <= 1000
    # R-MAT: no self-loops, unique sorted (u,v), log-uniform caps merged
    e = M.Edges.rmat(10, 10**6, 8)
    n, us, vs, cs = e.arrays()
    assert n == 1024 and np.all(us != vs)
    key = us.astype(np.int64) * n + vs
    assert np.all(np.diff(key) > 0)
    assert cs.min() >= 1
    # same output whatever the thread count
    e1 = M.Edges.rmat(10, 10**6, 8, threads=1).arrays()
    for x, y in zip(e1[1:], (us, vs, cs)):
        assert np.array_equal(x, y)
    # hub queries: s from the top 1% by out-degree, t with in-degree > 0
    s, t = e.hub_queries(200, 9)
    dout = np.bincount(us, minlength=n)
    din = np.bincount(vs, minlength=n)
    order = sorted(range(n), key=lambda v: (-dout[v], v))
    top = set(order[: max(1, int(np.floor(0.01 * n)))])
    assert all(x in top for x in s.tolist())
    assert np.all(din[t] > 0) and np.all(s != t)
    s, t = M.grid_queries(10, 12, 50, 3)
    assert np.all(s % 12 == 0) and np.all(t % 12 == 11)


def test_graph_and_query_files_round_trip(tmp_path):
    e = M.Edges.uniform(300, 2000, 100, 31)
    p = str(tmp_path / "graph.txt")
    e.write(p)
    n, us, vs, cs = e.arrays()
    n2, us2, vs2, cs2 = O.parse_graph(p)
    assert n2 == n and np.array_equal(us, us2) and np.array_equal(vs, vs2) and np.array_equal(cs, cs2)
    s, t = M.uniform_queries(300, 10, 32)
    q = str(tmp_path / "test.txt")
    M.write_queries(q, s, t)
    s2, t2 = O.parse_queries(q)
    assert np.array_equal(s, s2) and np.array_equal(t, t2)


def test_workload_sizes_match_the_configs():
    # uniform n=1e4 m=1e5 drops only self-loops
    n, us, _, _ = M.Edges.uniform(10000, 100000, 100, 1001).arrays()
    assert 99900 < len(us) <= 100000
    # grid 1024x1024: 4,190,208 lines -> 8,380,416 CSR entries
    n, us, _, _ = M.Edges.grid(1024, 1024, 1000, 2001).arrays()
    assert n == 1 << 20 and len(us) == 4190208


def test_find_entry():
    hg = M.HostGraph.from_edges(4, [0, 1, 0], [1, 2, 1], [5, 6, 7]).partition(1)
    d = hg.export()
    nid = d["new_id"]
    e = hg.find_entry(int(nid[0]), int(nid[1]))
    assert e >= 0 and d["cap"][e] == 5 and d["col"][e] == nid[1]
    assert hg.find_entry(int(nid[1]), int(nid[0])) == -1  # reverse entry only (cap 0)


def _reach_sets(n, us, vs):
    import scipy.sparse as sp
    from scipy.sparse.csgraph import connected_components

    A = sp.coo_matrix((np.ones(len(us)), (us, vs)), shape=(n, n)).tocsr()
    ns, scc = connected_components(A, directed=True, connection="strong")
    nw, wcc = connected_components(A, directed=False)
    return scc, wcc


@pytest.mark.parametrize("gen", ["uniform", "rmat", "grid"])
def test_components_match_scipy(gen):
    if gen == "uniform":
        e = M.Edges.uniform(3000, 4000, 10, 41)  # sparse: many small components
    elif gen == "rmat":
        e = M.Edges.rmat(12, 10**6, 42)
    else:
        e = M.Edges.grid(30, 30, 1000, 43)
    n, us, vs, cs = e.arrays()
    hg = M.HostGraph.from_edgelist(e)
    scc, wcc = hg.components()
    ref_s, ref_w = _reach_sets(n, us, vs)
    # same partitions of the vertex set (labels may differ)
    for mine, ref in ((scc, ref_s), (wcc, ref_w)):
        pairs = set(zip(mine.tolist(), ref.tolist()))
        assert len(pairs) == len(set(mine.tolist())) == len(set(ref.tolist()))


def test_class_mask_equals_prune_mask():
    """fwd(scc(s)) AND wcc == wcc(t) is the mask of BC:1391-1445 whenever t is
    reachable from s; otherwise t is outside fwd(s) and the flow is 0."""
    e = M.Edges.rmat(11, 10**6, 44)
    n, us, vs, cs = e.arrays()
    og = O.build_csr(n, us, vs, cs)
    hg = M.HostGraph.from_edgelist(e)
    scc, wcc = hg.components()
    rng = np.random.default_rng(45)
    hub_s, hub_t = e.hub_queries(20, 46)
    pairs = list(zip(hub_s.tolist(), hub_t.tolist())) + [tuple(x) for x in rng.integers(0, n, (20, 2)).tolist()]
    for s, t in pairs:
        if s == t:
            continue
        mask = O.prune_mask(og, s, t)
        # forward set of another member of s's SCC; prune_mask(r, r) is
        # fwd(r) AND wcc(r) = fwd(r) because reachable vertices share r's WCC
        rep = int(np.nonzero(scc == scc[s])[0][0])
        fwd = O.prune_mask(og, rep, rep).astype(bool)
        if not fwd[t]:
            assert O.maxflow(og, s, t) == 0
            continue
        cls = fwd & (wcc == wcc[t])
        assert np.array_equal(cls.astype(np.int32), mask), (s, t)


@pytest.mark.parametrize("gen", ["uniform", "rmat", "grid"])
def test_line_entries_after_partition(gen):
    """mf_hgraph_line_entries gives, in the region-ordered CSR, the forward
    entry of every input line: the oracle's entry of the same line (its CSR
    in input order) mapped through the relabel."""
    if gen == "uniform":
        e = M.Edges.uniform(500, 4000, 60, 91)   # repeated (u, v) pairs included
    elif gen == "rmat":
        e = M.Edges.rmat(10, 10**6, 92)
    else:
        e = M.Edges.grid(12, 17, 1000, 93)
    n, us, vs, cs = e.arrays()
    hg = M.HostGraph.from_edgelist(e).partition(16)
    d = hg.export()
    ent = hg.line_entries(us, vs)
    nid = d["new_id"]
    owner = np.repeat(np.arange(n), np.diff(d["row_ptr"]))
    assert np.array_equal(owner[ent], nid[us])
    assert np.array_equal(d["col"][ent], nid[vs])
    assert np.array_equal(d["cap"][ent], cs)
    assert len(np.unique(ent)) == len(ent)
    # same rank within the row as the oracle's entry in input order
    og = O.build_csr(n, us, vs, cs)
    oent = O.line_entries(og, us, vs)
    assert np.array_equal(ent - d["row_ptr"][nid[us]], oent - og.row_ptr[us])
    with pytest.raises(M.MaxFlowError):
        hg.line_entries(us[::-1].copy(), vs[::-1].copy())  # not the graph's line order


def test_capacity_update_rounds_shape_and_determinism():
    from mscresearch_amd import workloads as W

    caps = np.random.default_rng(3).integers(1, 10**6, 5000)
    a = list(W.capacity_update_rounds(caps, 4, 77, 0.01))
    b = list(W.capacity_update_rounds(caps, 4, 77, 0.01))
    cur = caps.copy()
    for (la, na), (lb, nb) in zip(a, b):
        assert np.array_equal(la, lb) and np.array_equal(na, nb)
        assert len(la) == 50 and len(np.unique(la)) == 50 and np.all(np.diff(la) > 0)
        assert np.all(na >= 1)
        ratio = na / cur[la]
        assert np.all(ratio >= 0.5 - 0.5 / cur[la]) and np.all(ratio <= 2.0 + 0.5 / cur[la])
        cur[la] = na


def test_edge_hash_sees_any_change():
    from mscresearch_amd import workloads as W

    us = np.array([0, 1, 2], np.int32)
    vs = np.array([1, 2, 0], np.int32)
    cs = np.array([5, 6, 7], np.int64)
    h = W.edge_hash(us, vs, cs)
    assert h == W.edge_hash(us.copy(), vs.copy(), cs.copy())
    assert h != W.edge_hash(us, vs, np.array([5, 6, 8]))
    assert h != W.edge_hash(vs, us, cs)
    assert h != W.edge_hash(us[[1, 0, 2]], vs[[1, 0, 2]], cs[[1, 0, 2]])


@pytest.mark.parametrize("gen,R", [("rmat", 16), ("rmat", 4), ("uniform", 8), ("sparse", 8)])
def test_reference_core_decision_matches_oracle(gen, R):
    """The CLI's diagnostic of the reference's boundary-core step
    (BC:2666-2753): prune mask, induced subgraph, re-partition into
    min(R, n') regions and the boundary-aware core, each restated by the
    oracle; the kept counts and the take decision must agree."""
    if gen == "rmat":
        e = M.Edges.rmat(9, 10**6, 95)
        qs, qt = e.hub_queries(10, 96)
    elif gen == "uniform":
        e = M.Edges.uniform(600, 3000, 40, 97)
        qs, qt = M.uniform_queries(600, 10, 98)
    else:  # sparse: many small components, so pruning bites
        e = M.Edges.uniform(2000, 1500, 40, 99)
        qs, qt = M.uniform_queries(2000, 10, 100)
    n, us, vs, cs = e.arrays()
    hg = M.HostGraph.from_edgelist(e).partition(R)
    d = hg.export()
    og = O.Graph(n, d["row_ptr"], d["col"], d["cap"], d["rev"])
    takes = 0
    for a, b in zip(qs, qt):
        s, t = int(d["new_id"][a]), int(d["new_id"][b])
        got = hg.core_decision(s, t)
        mask = O.prune_mask(og, s, t)
        pn = int(mask.sum())
        assert got["pruned_n"] == pn
        if not (0 < pn < 0.995 * n):
            assert got["run_n"] == n and not got["would_take"]
            continue
        sub, o2s, _ = O.induced(og, mask)
        runR = min(R, sub.n)
        so, si, do, di = O.csr_stats(sub)
        part, _ = O.partition(sub.n, runR, so, si, do, di)
        g2, nid, _oid, _rs, _re, rid = O.reorder(sub, runR, part)
        core, _kr = O.core_mask(g2, runR, rid, int(nid[o2s[s]]), int(nid[o2s[t]]), 2)
        kept = int(core.sum())
        assert got["run_n"] == sub.n
        assert got["core_kept"] == kept
        assert got["would_take"] == (0 < kept < 0.98 * sub.n)
        takes += int(got["would_take"])
