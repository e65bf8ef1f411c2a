"""Dynamic mode and the drop-in command line on the GPU.

Dynamic (reference: OnAdd_kernel BC:959-1000, incremental update
BC:2273-2335; the reference handles capacity increases only and never calls
its update path from main): pinned queries keep their residual, capacity
updates (increases and decreases) are applied to it, and the warm re-solve
must give exactly the flow of a cold solve on the updated graph and of the
oracle.

CLI (reference main, BC:2561-2933): same arguments, inputs and res_mt.txt /
res_dynamic.txt lines; flows compared with the oracle on the original ids.
"""
import os
import subprocess

import numpy as np
import pytest

import mscresearch_amd as M
import oracle as O
from gpu_util import Case, engine_with, log

pytestmark = pytest.mark.gpu


def updated_round(rng, cap, fwd_entries, frac):
    k = max(1, int(frac * len(fwd_entries)))
    pick = rng.choice(fwd_entries, k, replace=False)
    factor = rng.uniform(0.5, 2.0, k)
    new = np.maximum(1, np.rint(cap[pick] * factor)).astype(np.int64)
    return pick.astype(np.int32), new


@pytest.mark.parametrize("gen", ["rmat", "grid", "uniform"])
def test_warm_resolve_matches_cold_and_oracle(gen):
    rng = np.random.default_rng(7)
    if gen == "rmat":
        e = M.Edges.rmat(12, 10**6, 5001)
        qs, qt = e.hub_queries(8, 5002)
    elif gen == "grid":
        e = M.Edges.grid(40, 40, 1000, 5003)
        qs, qt = M.grid_queries(40, 40, 8, 5004)
    else:
        e = M.Edges.uniform(3000, 30000, 100, 5005)
        qs, qt = M.uniform_queries(3000, 8, 5006)
    hg = M.HostGraph.from_edgelist(e).partition(64)
    d = hg.export()
    n = len(d["new_id"])
    cap = d["cap"].copy()
    fwd = np.nonzero(cap > 0)[0]
    dg = M.DeviceGraph(hg, 512)
    eng = M.Engine(8)
    gid = eng.add_graph(dg)
    q = [(int(d["new_id"][a]), int(d["new_id"][b])) for a, b in zip(qs, qt)]
    for i, (s, t) in enumerate(q):
        eng.pin(i, gid, s, t, i)
    res = {r.tag: r for r in eng.solve_pinned(warm=False)}
    og = O.Graph(n, d["row_ptr"], d["col"], cap, d["rev"])
    for i, (s, t) in enumerate(q):
        assert res[i].flow == O.maxflow(og, s, t)
    for rnd in range(4):
        ent, new = updated_round(rng, cap, fwd, 0.05)
        eng.update_caps(gid, ent, new)
        cap[ent] = new
        og = O.Graph(n, d["row_ptr"], d["col"], cap, d["rev"])
        warm = {r.tag: r for r in eng.solve_pinned(warm=True)}
        # cold: a fresh upload of the updated graph, solved from scratch
        hg2 = M.HostGraph.from_edges(n, *edges_from_csr(d, cap)).partition(64)
        d2 = hg2.export()
        ceng = M.Engine(1)
        cgid = ceng.add_graph(M.DeviceGraph(hg2, 512))
        for i, (s, t) in enumerate(q):
            want = O.maxflow(og, s, t)
            log(f"{gen} round {rnd} q{i}: warm={warm[i].flow} oracle={want} drains={warm[i].drains}")
            assert warm[i].flow == want, (gen, rnd, i)
            assert warm[i].cert_ok == 1
            cold = ceng.solve(cgid, int(d2["new_id"][s]), int(d2["new_id"][t]))
            assert cold.flow == want


def edges_from_csr(d, cap):
    """Forward entries of a CSR as an edge list (region-ordered ids)."""
    rp, col = d["row_ptr"], d["col"]
    owner = np.repeat(np.arange(len(rp) - 1), np.diff(rp)).astype(np.int32)
    fwd = cap > 0
    return owner[fwd], col[fwd].astype(np.int32), cap[fwd]


def test_cold_solve_after_update_equals_oracle():
    """Queries submitted after an update (not pinned) see the new capacities."""
    rng = np.random.default_rng(11)
    e = M.Edges.uniform(2000, 16000, 100, 5101)
    hg = M.HostGraph.from_edgelist(e).partition(32)
    d = hg.export()
    n = len(d["new_id"])
    cap = d["cap"].copy()
    dg = M.DeviceGraph(hg, 512)
    eng = M.Engine(4)
    gid = eng.add_graph(dg)
    ent, new = updated_round(rng, cap, np.nonzero(cap > 0)[0], 0.1)
    eng.update_caps(gid, ent, new)
    cap[ent] = new
    og = O.Graph(n, d["row_ptr"], d["col"], cap, d["rev"])
    qs, qt = M.uniform_queries(n, 10, 5102)
    for a, b in zip(qs, qt):
        s, t = int(d["new_id"][a]), int(d["new_id"][b])
        assert eng.solve(gid, s, t).flow == O.maxflow(og, s, t)


def run_cli(tmp_path, mode, env=None):
    out = tmp_path / ("out_" + mode)
    r = subprocess.run([M.CLI_PATH, str(tmp_path / "in"), str(out), mode], capture_output=True, text=True,
                       env={**os.environ, **(env or {})}, timeout=300)
    return r, out


def write_inputs(tmp_path, e, qs, qt, extra=""):
    (tmp_path / "in").mkdir(exist_ok=True)
    e.write(str(tmp_path / "in" / "graph.txt"))
    lines = "".join(f"{a + 1} {b + 1}\n" for a, b in zip(qs, qt)) + extra
    (tmp_path / "in" / "test.txt").write_text(lines)


@pytest.mark.parametrize("slots", ["1", "8"])
def test_cli_static_matches_oracle(tmp_path, slots):
    e = M.Edges.uniform(10000, 100000, 100, 1001)
    qs, qt = M.uniform_queries(10000, 24, 1002)
    # plus invalid queries: out of range, id 0, s == t
    write_inputs(tmp_path, e, qs, qt, extra="10001 5\n0 3\n7 7\n")
    r, out = run_cli(tmp_path, "static", {"MSCMF_SLOTS": slots})
    assert r.returncode == 0, r.stderr[-2000:]
    lines = (out / "res_mt.txt").read_text().splitlines()
    assert len(lines) == 27
    n, us, vs, cs = O.parse_graph(str(tmp_path / "in" / "graph.txt"))
    og = O.build_csr(n, us, vs, cs)
    for i, (a, b) in enumerate(zip(qs, qt)):
        flow, ms = map(int, lines[i].split())
        assert flow == O.maxflow(og, a, b), i
        assert ms >= 0
    assert lines[24:] == ["0 0", "0 0", "0 0"]
    assert (out / "metrics.json").exists()
    assert "Total time:" in r.stderr


def test_cli_static_graph_cache(tmp_path):
    """MSCMF_GRAPH_CACHE: the first run writes the prepared graph, the next
    reads it back; both answer like the oracle.  A rewritten graph.txt makes
    the cache stale: the run rebuilds it from the text and answers for the
    new graph."""
    cache = tmp_path / "graph.mfg"
    env = {"MSCMF_GRAPH_CACHE": str(cache), "MSCMF_R": "64"}
    for seed, expect in ((5301, "[GraphCache] wrote"), (5301, "[GraphCache] loaded"), (5311, "graph cache is stale")):
        e = M.Edges.rmat(11, 10**6, seed)
        qs, qt = e.hub_queries(6, seed + 1)
        if "loaded" not in expect:
            write_inputs(tmp_path, e, qs, qt)
        log(f"cli with graph cache: seed {seed}, expect {expect}")
        r, out = run_cli(tmp_path, "static", env)
        assert r.returncode == 0, r.stderr[-2000:]
        assert expect in r.stderr, r.stderr[-2000:]
        n, us, vs, cs = O.parse_graph(str(tmp_path / "in" / "graph.txt"))
        og = O.build_csr(n, us, vs, cs)
        lines = (out / "res_mt.txt").read_text().splitlines()
        assert [int(l.split()[0]) for l in lines] == [O.maxflow(og, a, b) for a, b in zip(qs, qt)]
    assert cache.exists()


def parse_ref_log(text):
    """Per-query records from stderr as the reference's comparison scripts
    read them (COMPARISON/ECL_RA/plot_three_datasets_ecl_labels.py:149-215):
    a "Query i:" opens a record, the "[Init] kernel_1(...)=x ms, kernel_7(...)=y ms"
    line gives k1/k7, each "Iteration j: init_hl=a ms, bfs=b ms, push_relabel=c ms
    ... remove_invalid=d ms" line one iteration."""
    import re
    text = re.sub(r"(?=Query \d+:)", "\n", text)
    text = re.sub(r"(?=\[Init\])", "\n", text)
    text = re.sub(r"(?=Iteration \d+:)", "\n", text)
    q, cur = {}, None
    for line in text.splitlines():
        m = re.search(r"Query\s+(\d+):", line)
        if m:
            cur = int(m.group(1))
            q.setdefault(cur, {"iters": [], "k1": None, "k7": None, "flow": None})
        if cur is None:
            continue
        m = re.search(r"kernel_1\(init residual/excess\)=([0-9.]+)\s*ms,\s*kernel_7\(preflow\)=([0-9.]+)\s*ms", line)
        if m:
            q[cur]["k1"], q[cur]["k7"] = float(m.group(1)), float(m.group(2))
        m = re.search(r"Iteration\s+(\d+):\s*init_hl=([0-9.]+)\s*ms,\s*bfs=([0-9.]+)\s*ms,\s*"
                      r"push_relabel=([0-9.]+)\s*ms.*?remove_invalid=([0-9.]+)\s*ms", line)
        if m:
            q[cur]["iters"].append(tuple(float(m.group(k)) for k in range(2, 6)))
        m = re.search(r"\[QueryPrune\+RegionBoundary\].*Flow=(\d+), Time=(\d+) ms", line)
        if m:
            q[cur]["flow"] = int(m.group(1))
    return q


@pytest.mark.parametrize("gen", ["rmat", "grid"])
def test_cli_reference_log_lines(tmp_path, gen):
    """MSCMF_LOG=1: every valid query gets a kernel_1/kernel_7 line and
    timed iteration lines the reference's scripts parse, and a closing line
    whose flow is the res_mt.txt flow; invalid queries get no record."""
    if gen == "rmat":
        e = M.Edges.rmat(11, 10**6, 5401)
        qs, qt = e.hub_queries(5, 5402)
    else:
        e = M.Edges.grid(48, 48, 1000, 5403)
        qs, qt = M.grid_queries(48, 48, 3, 5404)
    write_inputs(tmp_path, e, qs, qt, extra="7 7\n")
    r, out = run_cli(tmp_path, "static", {"MSCMF_LOG": "1"})
    assert r.returncode == 0, r.stderr[-2000:]
    flows = [int(l.split()[0]) for l in (out / "res_mt.txt").read_text().splitlines()]
    recs = parse_ref_log(r.stderr)
    assert sorted(recs) == list(range(1, len(qs) + 2))
    for i in range(1, len(qs) + 1):
        rec = recs[i]
        assert rec["k1"] is not None and rec["k1"] >= 0 and rec["k7"] >= 0, (i, r.stderr[-3000:])
        assert rec["iters"], i
        assert all(min(x) >= 0 for x in rec["iters"])
        assert sum(x[0] + x[1] + x[2] for x in rec["iters"]) > 0  # timed
        assert rec["flow"] == flows[i - 1]
    last = recs[len(qs) + 1]
    assert last["k1"] is None and not last["iters"] and "Invalid" in r.stderr
    assert "Total time:" in r.stderr


def test_cli_dynamic_with_updates(tmp_path):
    rng = np.random.default_rng(3)
    e = M.Edges.rmat(11, 10**6, 5201)
    qs, qt = e.hub_queries(12, 5202)
    write_inputs(tmp_path, e, qs, qt)
    n, us, vs, cs = e.arrays()
    cap = cs.copy()
    upd = []
    rounds = 3
    caps_after = {}
    for rnd in range(1, rounds + 1):
        pick = rng.choice(len(us), max(1, len(us) // 50), replace=False)
        newc = np.maximum(1, np.rint(cap[pick] * rng.uniform(0.5, 2.0, len(pick)))).astype(np.int64)
        for k, c in zip(pick, newc):
            upd.append(f"{rnd} {us[k] + 1} {vs[k] + 1} {c}\n")
        cap[pick] = newc
        caps_after[rnd] = cap.copy()
    (tmp_path / "in" / "updates.txt").write_text("".join(upd))
    r, out = run_cli(tmp_path, "dynamic")
    assert r.returncode == 0, r.stderr[-2000:]
    base = (out / "res_dynamic.txt").read_text().splitlines()
    og0 = O.build_csr(n, us, vs, cs)
    for i, (a, b) in enumerate(zip(qs, qt)):
        assert int(base[i].split()[0]) == O.maxflow(og0, a, b)
    rows = [list(map(int, l.split())) for l in (out / "res_dynamic_rounds.txt").read_text().splitlines()]
    # windows of 10 queries each restart from graph.txt and replay the rounds
    assert len(rows) == rounds * len(qs)
    for rnd, want_caps in caps_after.items():
        og = O.build_csr(n, us, vs, want_caps)
        for rd, qi, flow, ms in rows:
            if rd == rnd:
                assert flow == O.maxflow(og, qs[qi - 1], qt[qi - 1]), (rnd, qi)


def test_engines_sharing_a_graph_see_capacity_updates():
    """Engine B keeps residuals of a graph in its slabs; engine A updates the
    graph's capacities; B's next solves must start from the new capacities."""
    rng = np.random.default_rng(13)
    e = M.Edges.uniform(3000, 30000, 100, 5301)
    hg = M.HostGraph.from_edgelist(e).partition(32)
    d = hg.export()
    n = len(d["new_id"])
    cap = d["cap"].copy()
    dg = M.DeviceGraph(hg, 512)
    a, b = M.Engine(2), M.Engine(2, M.default_params(prune=0))
    ga, gb = a.add_graph(dg), b.add_graph(dg)
    qs, qt = M.uniform_queries(n, 6, 5302)
    q = [(int(d["new_id"][x]), int(d["new_id"][y])) for x, y in zip(qs, qt)]
    og = O.Graph(n, d["row_ptr"], d["col"], cap, d["rev"])
    for s, t in q:
        assert b.solve(gb, s, t).flow == O.maxflow(og, s, t)
    for rnd in range(2):
        ent, new = updated_round(rng, cap, np.nonzero(cap > 0)[0], 0.2)
        a.update_caps(ga, ent, new)
        cap[ent] = new
        og = O.Graph(n, d["row_ptr"], d["col"], cap, d["rev"])
        for s, t in q:
            r = b.solve(gb, s, t)
            assert r.flow == O.maxflow(og, s, t), (rnd, s, t)
            assert r.cert_ok == 1


def test_update_caps_rejects_reverse_and_repeated_entries():
    """A capacity update may only name forward entries (cap > 0), each once;
    a bad batch is refused whole and leaves every capacity as it was."""
    case = Case(M.Edges.rmat(11, 10**6, 7101), R=16)
    og = case.og
    eng, gid = engine_with(case)
    fwd = np.nonzero(og.cap > 0)[0]
    rev = np.nonzero(og.cap == 0)[0]
    s_, t_ = case.edges.hub_queries(2, 7102)
    pairs = [case.q(a, b) for a, b in zip(s_, t_)]
    before = [eng.solve(gid, s, t).flow for s, t in pairs]
    for bad in (np.array([fwd[0], rev[0]], np.int32), np.array([fwd[1], fwd[2], fwd[1]], np.int32)):
        with pytest.raises(M.MaxFlowError) as ei:
            eng.update_caps(gid, bad, np.full(len(bad), 7, np.int64))
        assert ei.value.code == M.MF_EINVAL
    # nothing changed: the same flows, equal to the oracle's on the old capacities
    for (s, t), f in zip(pairs, before):
        r = eng.solve(gid, s, t)
        assert r.flow == f == O.maxflow(og, s, t)
    # a valid batch still applies
    pick = fwd[:50]
    eng.update_caps(gid, pick, np.full(50, 3, np.int64))
    cap2 = og.cap.copy()
    cap2[pick] = 3
    og2 = O.Graph(og.n, og.row_ptr, og.col, cap2, og.rev)
    for s, t in pairs:
        assert eng.solve(gid, s, t).flow == O.maxflow(og2, s, t)


def test_pruned_subgraphs_of_another_engine_follow_capacity_updates():
    """Engine C prunes (R-MAT: the mask drops the isolated vertices) and caches
    induced subgraphs holding copies of the capacities; engine A updates the
    shared graph.  C's next solves must not run on its stale copies."""
    case = Case(M.Edges.rmat(12, 10**6, 7201), R=32)
    og = case.og
    cap = og.cap.copy()
    a, ga = engine_with(case)
    c, gc = engine_with(case)
    s_, t_ = case.edges.hub_queries(6, 7202)
    pairs = [case.q(x, y) for x, y in zip(s_, t_)]
    for s, t in pairs:
        r = c.solve(gc, s, t)
        assert r.flow == O.maxflow(og, s, t)
    assert any(c.solve(gc, s, t).pruned for s, t in pairs)
    rng = np.random.default_rng(7203)
    for rnd in range(2):
        ent, new = updated_round(rng, cap, np.nonzero(cap > 0)[0], 0.3)
        a.update_caps(ga, ent, new)
        cap[ent] = new
        og2 = O.Graph(og.n, og.row_ptr, og.col, cap, og.rev)
        for s, t in pairs:
            r = c.solve(gc, s, t)
            assert r.cert_ok == 1
            assert r.flow == O.maxflow(og2, s, t), (rnd, s, t)
