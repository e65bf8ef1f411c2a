"""End-to-end GPU flows, bit-exact against the oracle and the golden fixtures.

Discharge, path push, RM and the hot sweep are nondeterministic, so they are
pinned through the final int64 flow (unique), the GPU's own min-cut
certificate and an independent CPU certificate on the downloaded residual.
"""
import numpy as np
import pytest

import mscresearch_amd as M
import oracle as O
from gpu_util import Case, engine_with, log

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1():
    log("case: uniform n=10000 m=100000 seed 1001, R=128")
    return Case(M.Edges.uniform(10000, 100000, 100, 1001), R=128)


def test_golden_flows(golden):
    from test_oracle import regenerate

    for fx in golden:
        log(f"golden {fx['name']}")
        e = regenerate(fx)
        case = Case(e, R=min(128, fx["n"]))
        eng, gid = engine_with(case)
        for (a, b), f in zip(fx["queries"], fx["flows"]):
            s, t = case.q(a, b)
            r = eng.solve(gid, s, t)
            assert r.flow == f, (fx["name"], a, b, r.flow, f)
            assert r.cert_ok == 1


def test_c1_flows_and_cpu_certificate(c1):
    eng, gid = engine_with(c1)
    qs, qt = M.uniform_queries(c1.n, 64, 1002)
    for i, (a, b) in enumerate(zip(qs, qt)):
        s, t = c1.q(a, b)
        r, res, ex, h = eng.dbg_solve_state(gid, s, t, c1.n, c1.E)
        f = c1.oracle_flow(s, t)
        assert r.flow == f, (i, r.flow, f)
        assert O.certify(c1.og, res, s, t, r.flow, excess=ex) == 0
        # the GPU's T = {h < n} is a closed set holding t and not s, so it
        # holds every vertex with a residual path to t
        T = h < c1.n
        assert T[t] and not T[s]
        owner = np.repeat(np.arange(c1.n), np.diff(c1.og.row_ptr))
        into_T = (res > 0) & ~T[owner] & T[c1.og.col]
        assert not into_T.any()
        hb = O.bfs_heights(c1.og, res, np.zeros(c1.n, np.int64), s, t)
        assert np.all(T[hb < c1.n])


@pytest.mark.parametrize("rows,cols,fuse,extra", [
    (64, 64, 8, {}), (128, 96, 8, {}), (128, 96, 1, {}),
    (128, 96, 8, dict(small_wgs=1)),                   # deep relabels in one workgroup
    (128, 96, 8, dict(small_wgs=2, small_levels=3)),   # two workgroups, a launch every 3 levels
])
def test_grid_flows(rows, cols, fuse, extra):
    """Grids are deep (BFS depth ~ rows + cols): fused discharge cycles run,
    and relabel BFS levels past depth 64 run on several workgroups that meet
    at a barrier after each level."""
    log(f"grid {rows}x{cols} fuse {fuse} {extra}")
    case = Case(M.Edges.grid(rows, cols, 1000, 2001), R=64, tile_max=512)
    eng, gid = engine_with(case, fuse=fuse, **extra)
    qs, qt = M.grid_queries(rows, cols, 8, 2002)
    for a, b in zip(qs, qt):
        s, t = case.q(a, b)
        r = eng.solve(gid, s, t)
        assert r.flow == case.oracle_flow(s, t)
        assert r.cert_ok == 1


@pytest.mark.parametrize("nslots", [1, 4])
def test_barrier_timeout_restarts_single_workgroup(nslots):
    """A multi-workgroup BFS barrier that waits past its bound flags the slot,
    and the BFS starts again in one workgroup.  Forced from the host with a
    bound of one clock tick (every wait ends at once, so every kernel exits):
    the restart path runs and every flow stays exact and certified."""
    rows, cols = 128, 96
    log(f"grid {rows}x{cols}, {nslots} slot(s), barrier bound 1 tick")
    case = Case(M.Edges.grid(rows, cols, 1000, 2001), R=64, tile_max=512)
    eng, gid = engine_with(case, nslots=nslots, small_wgs=4)
    assert eng.dbg_barrier_ticks(1) == 0
    qs, qt = M.grid_queries(rows, cols, 8, 2002)
    q = [case.q(a, b) for a, b in zip(qs, qt)]
    for i, (s, t) in enumerate(q):
        eng.submit(gid, s, t, i)
    res = eng.run()
    assert len(res) == len(q)
    for r in res:
        assert r.status == 0 and r.cert_ok == 1
        assert r.flow == case.oracle_flow(*q[r.tag]), r.tag
    assert eng.dbg_barrier_ticks(0) > 0  # the restart path ran


@pytest.mark.parametrize("scale", [12, 14])
def test_rmat_flows_pruned(scale):
    log(f"rmat scale {scale}")
    case = Case(M.Edges.rmat(scale, 10**6, 3000 + scale), R=128)
    eng, gid = engine_with(case)
    qs, qt = case.edges.hub_queries(24, 3100 + scale)
    pruned = 0
    for a, b in zip(qs, qt):
        s, t = case.q(a, b)
        r = eng.solve(gid, s, t)
        assert r.flow == case.oracle_flow(s, t), (a, b)
        assert r.cert_ok == 1
        pruned += r.pruned
    assert pruned > 0  # R-MAT has isolated vertices: pruning shrinks the graph
    hits, misses = eng.cache_stats()
    assert hits + misses == pruned


def test_edge_cases():
    # 0: s, 5: t; parallel and antiparallel lines; an isolated vertex 6;
    # 7 -> 8 unreachable component
    lines = [(0, 1, 3), (0, 1, 4), (1, 0, 2), (1, 5, 5), (0, 2, 9), (2, 5, 1), (2, 3, 6), (3, 5, 2),
             (5, 3, 7), (7, 8, 5), (4, 0, 3)]
    n = 9
    us = np.array([l[0] for l in lines], np.int32)
    vs = np.array([l[1] for l in lines], np.int32)
    cs = np.array([l[2] for l in lines], np.int64)
    hg = M.HostGraph.from_edges(n, us, vs, cs).partition(2)
    d = hg.export()
    og = O.Graph(n, d["row_ptr"], d["col"], d["cap"], d["rev"])
    dg = M.DeviceGraph(hg, 4)
    eng = M.Engine(1)
    gid = eng.add_graph(dg)
    nid = d["new_id"]
    pairs = [(0, 5), (0, 8), (7, 8), (5, 0), (6, 5), (4, 5), (0, 3), (8, 7), (1, 2)]
    for a, b in pairs:
        s, t = int(nid[a]), int(nid[b])
        r = eng.solve(gid, s, t)
        assert r.flow == O.maxflow(og, s, t), (a, b)
    with pytest.raises(M.MaxFlowError):
        eng.solve(gid, 1, 1)
    with pytest.raises(M.MaxFlowError):
        eng.solve(gid, 0, n)


@pytest.mark.parametrize("variant", [
    dict(tile_frac=0.0),                        # region (tile) discharge every iteration
    dict(tile_frac=1e9),                        # active-list discharge every iteration
    dict(path_switch=1 << 20, p_count0=1),      # path push whenever anything is active
    dict(h_cut=0, rm_freq0=1, rm_freq_max=1),   # RM repair on every iteration
    dict(hot=1, hot_min_active=1, hot_k=4),            # hot sweep on every even iteration
    dict(gr_freq=1, gr_freq_min=1, gr_freq_max=1),
    dict(K=1, gr_freq=1000, gr_freq_min=1000, gr_freq_max=1000),
    dict(prune=0),
    dict(local_relabel=1),
    dict(local_relabel=1, local_band=1 << 30, gr_stall=0, gr_work=0.0),  # local relabel between all discharges
    dict(local_relabel=1, h_cut=0, rm_freq0=1, rm_freq_max=1),       # ... and RM on every iteration
    dict(cert=2),                               # certificate pair sums entry by entry
    dict(cert=1),                               # fingerprint over every pair (not only the marked ones)
    dict(spec_cert=0),                          # terminate through an exact global relabel only
    dict(sparse_count=0),                       # count actives over every vertex after each discharge
    dict(hot=1, sparse_count=1, hot_min_active=1, hot_k=4, tile_frac=0.0),  # sparse counts after tile + hot sweeps
    dict(small_edges=0),                        # every BFS level grid-wide (top-down or bottom-up)
    dict(small_edges=1 << 30),                  # single-workgroup levels up to kSmallFrontier vertices
    dict(mark_list=0),                          # resets and certificates from the mark map
    dict(mark_list=16),                         # ... once the list of marked entries overflows
    dict(inject=1),                             # bounded injection at s instead of a saturated row
    dict(inject=1, cert=1),                     # ... with the fingerprint over every pair
    dict(inject=1, cert=2, sparse_count=0),     # ... and entry-by-entry pair sums
    dict(tile_threads=256, tile_frac=0.0),      # tile discharge in 256-thread workgroups, every step
    dict(tile_threads=1024, tile_frac=0.0),     # ... and in 1024-thread workgroups
    dict(tile_split=4, tile_frac=0.0),          # up to 4 workgroups per tile by excess score, every step
    dict(tile_split=8, tile_threads=256, tile_frac=0.0, tile_rounds=1),  # ... 8 of 256 threads, one round each
    dict(hot=1, tile_split=2, sparse_count=1, hot_min_active=1, hot_k=4, tile_frac=0.0),  # ... and in the hot sweep
    dict(bu_skip=0),                            # bottom-up relabel levels scan every unvisited vertex
    dict(sink_cut=0),                           # no early stop when excess(t) fills t's in-capacity
    dict(bu_pick=0),                            # bottom-up relabel levels read rows only
    dict(push_first=1),                         # lane/wave discharge: push first, minimum only to relabel
    dict(push_first=1, tile_frac=0.0),          # ... in tiles
    dict(disc_band=1),                          # discharge the lowest two active levels only
    dict(disc_band=2, sink_cut=0, tile_frac=0.0),  # ... three, to the end, in tiles
    dict(bu_skip=1, small_edges=0),             # ... skip those that cannot reach t's SCC, every level grid-wide
    dict(bu_bits=0),                            # bottom-up levels test parents by their levels, not a bitmap
    dict(bu_bits=0, small_edges=0, bu_alpha=1.0),  # ... every level grid-wide, bottom-up early
    dict(small_edges=0, bu_alpha=1.0, bu_beta=4.0),  # bitmaps over runs of bottom-up levels between top-down ones
    dict(dis_queue=1),                          # list discharge as one launch over a device work queue
    dict(bu_recv=0),                            # relabel levels read every row's residuals (no fresh-row tests)
    dict(bu_recv=0, small_edges=0, bu_alpha=1.0),  # ... every level grid-wide, bottom-up early
    dict(small_edges=0, bu_alpha=1.0, h_cut=0, rm_freq0=1, rm_freq_max=1),  # fresh rows beside RM repairs
])
def test_scheduler_variants_stay_exact(c1, variant):
    log(f"variant {variant}")
    eng, gid = engine_with(c1, **variant)
    qs, qt = M.uniform_queries(c1.n, 12, 1005)
    for a, b in zip(qs, qt):
        s, t = c1.q(a, b)
        r = eng.solve(gid, s, t)
        assert r.flow == c1.oracle_flow(s, t), (variant, a, b)
        assert r.cert_ok == 1


def test_sink_cut_stops_early_and_stays_exact(c1):
    """Where t's in-entries are the minimum cut, the solve stops as soon as
    excess(t) fills them, certified on T = {t}; flows equal the oracle's and
    the full solve's."""
    qs, qt = M.uniform_queries(c1.n, 24, 1021)
    on, gid = engine_with(c1, sink_cut=1)
    off, gid2 = engine_with(c1, sink_cut=0)
    early = 0
    for a, b in zip(qs, qt):
        s, t = c1.q(a, b)
        r1, r0 = on.solve(gid, s, t), off.solve(gid2, s, t)
        assert r1.flow == r0.flow == c1.oracle_flow(s, t), (a, b)
        assert r1.cert_ok == 1 and r0.sink_cut == 0
        early += r1.sink_cut
    assert early > 0


@pytest.mark.parametrize("extra", [{}, dict(dis_queue=1)])
def test_batched_equals_alone(c1, extra):
    """Each result is the same alone and in the mix (continuous batching);
    with dis_queue=1 the mix's list discharge is one launch over a device work
    queue of every slot's blocks."""
    qs, qt = M.uniform_queries(c1.n, 64, 1002)
    q = [c1.q(a, b) for a, b in zip(qs, qt)]
    alone_eng, gid = engine_with(c1)
    alone = [alone_eng.solve(gid, s, t).flow for s, t in q]
    mix_eng, gid2 = engine_with(c1, nslots=24, **extra)
    for i, (s, t) in enumerate(q):
        mix_eng.submit(gid2, s, t, i)
    res = mix_eng.run()
    assert len(res) == len(q)
    got = {r.tag: r.flow for r in res}
    assert [got[i] for i in range(len(q))] == alone
    assert all(r.cert_ok == 1 and r.status == 0 for r in res)


def test_mixed_graphs_in_one_engine():
    g1 = Case(M.Edges.grid(40, 40, 1000, 4001), R=32, tile_max=256)
    g2 = Case(M.Edges.rmat(11, 10**6, 4002), R=64)
    eng = M.Engine(8)
    a = eng.add_graph(g1.dg)
    b = eng.add_graph(g2.dg)
    qa = [g1.q(x, y) for x, y in zip(*M.grid_queries(40, 40, 10, 4003))]
    qb = [g2.q(x, y) for x, y in zip(*g2.edges.hub_queries(10, 4004))]
    for i in range(10):
        eng.submit(a, *qa[i], 2 * i)
        eng.submit(b, *qb[i], 2 * i + 1)
    res = {r.tag: r for r in eng.run()}
    for i in range(10):
        assert res[2 * i].flow == g1.oracle_flow(*qa[i])
        assert res[2 * i + 1].flow == g2.oracle_flow(*qb[i])


def test_giant_rows():
    """Rows longer than the giant threshold (2048) go through the workgroup
    paths: top-down/bottom-up heavy BFS, block discharge, heavy certificate."""
    rng = np.random.default_rng(99)
    L = 6000
    s, hub, t = 0, 1, 2
    leaves = np.arange(3, 3 + L)
    us = [s, s] + [hub] * L + leaves.tolist() + leaves[:L // 2].tolist()
    vs = [hub, 3] + leaves.tolist() + [t] * L + [hub] * (L // 2)
    cs = [10**9, 7] + rng.integers(1, 50, L).tolist() + rng.integers(1, 50, L).tolist() + rng.integers(1, 9, L // 2).tolist()
    n = 3 + L
    hg = M.HostGraph.from_edges(n, us, vs, cs).partition(4)
    d = hg.export()
    og = O.Graph(n, d["row_ptr"], d["col"], d["cap"], d["rev"])
    dg = M.DeviceGraph(hg, 1024)
    for variant in (dict(), dict(tile_frac=0.0), dict(tile_frac=1e9), dict(prune=0),
                    dict(giant_rounds=1), dict(giant_rounds=1, tile_frac=0.0)):
        eng = M.Engine(1, M.default_params(**variant))
        gid = eng.add_graph(dg)
        for a, b in ((s, t), (hub, t), (s, int(leaves[5]))):
            sa, tb = int(d["new_id"][a]), int(d["new_id"][b])
            r = eng.solve(gid, sa, tb)
            assert r.flow == O.maxflow(og, sa, tb), (variant, a, b)
            assert r.cert_ok == 1
    # planted mistakes in and beside the giant hub row: the fingerprint mode
    # (giant rows by exact row sums, the others by vertex terms) rejects each
    eng = M.Engine(1)
    gid = eng.add_graph(dg)
    sa, tb = int(d["new_id"][s]), int(d["new_id"][t])
    hb = int(d["new_id"][hub])
    r, res, ex, h = eng.dbg_solve_state(gid, sa, tb, n, len(d["col"]))
    for mode in (1, 2, 3):
        assert eng.dbg_certify(gid, sa, tb, res, ex, h, mode) == (0, r.flow, r.flow, True)
    rp = d["row_ptr"]
    assert rp[hb + 1] - rp[hb] > 2048
    e_hub = rp[hb] + int(np.nonzero(res[rp[hb]:rp[hb + 1]] > 0)[0][0])
    leaf = int(d["col"][e_hub])
    assert rp[leaf + 1] - rp[leaf] <= 32
    e_leaf = int(rp[leaf])
    for e_bad, row in ((e_hub, hb), (e_leaf, leaf)):
        # pair sum off, row sum kept: only the pair side of the fingerprint
        bad_res, bad_ex = res.copy(), ex.copy()
        bad_res[e_bad] += 1
        bad_ex[row] += 1
        for mode in (1, 3):
            b1 = eng.dbg_certify(gid, sa, tb, bad_res, bad_ex, h, mode)
            assert b1[0] > 0 or not b1[3], (row, mode)
        # stored excess off alone
        bad_ex = ex.copy()
        bad_ex[row] -= 1
        for mode in (1, 3):
            b1 = eng.dbg_certify(gid, sa, tb, res, bad_ex, h, mode)
            assert b1[0] > 0 or not b1[3], (row, mode)


def test_gpu_certificate_catches_planted_mistakes(c1):
    """The GPU certificate (fingerprint over every pair or over the changed
    pairs, and entry-by-entry pair sums) accepts a correct final state and
    rejects planted mistakes, like the oracle's."""
    eng, gid = engine_with(c1)
    qs, qt = M.uniform_queries(c1.n, 4, 1007)
    og = c1.og
    n = c1.n
    owner = np.repeat(np.arange(n), np.diff(og.row_ptr))
    rng = np.random.default_rng(5)
    for a, b in zip(qs, qt):
        s, t = c1.q(a, b)
        r, res, ex, h = eng.dbg_solve_state(gid, s, t, c1.n, c1.E)
        assert O.certify(og, res, s, t, r.flow, excess=ex) == 0
        for mode in (1, 2, 3, 4):
            bad, cut, ext, fz = eng.dbg_certify(gid, s, t, res, ex, h, mode)
            assert (bad, fz, cut, ext) == (0, True, r.flow, r.flow), mode
        # (a) two pair sums off by +1 / -1 inside one row: the recomputed
        # excess is unchanged, only the pair-sum check can see it
        fwd = np.nonzero((og.cap > 0) & (res > 0) & (owner != s) & (owner != t))[0]
        rows, cnt = np.unique(owner[fwd], return_counts=True)
        v = rng.choice(rows[cnt >= 2])
        e1, e2 = fwd[owner[fwd] == v][:2]
        bad_res = res.copy()
        bad_res[e1] += 1
        bad_res[e2] -= 1
        assert O.certify(og, bad_res, s, t, r.flow, excess=ex) != 0
        assert not eng.dbg_certify(gid, s, t, bad_res, ex, h, 1)[3]
        assert eng.dbg_certify(gid, s, t, bad_res, ex, h, 2)[0] > 0
        assert not eng.dbg_certify(gid, s, t, bad_res, ex, h, 3)[3]
        assert not eng.dbg_certify(gid, s, t, bad_res, ex, h, 4)[3]
        # (b) stored excess off by one at an inner vertex: the full mode
        # counts it; the fingerprint mode's vertex terms leave a nonzero sum
        # (or count it, for a vertex of T or of a giant row)
        w = int(rng.choice(np.setdiff1d(np.arange(n), [s, t])))
        bad_ex = ex.copy()
        bad_ex[w] += 1
        assert O.certify(og, res, s, t, r.flow, excess=bad_ex) != 0
        assert eng.dbg_certify(gid, s, t, res, bad_ex, h, 2)[0] > 0
        for mode in (1, 3, 4):
            b1 = eng.dbg_certify(gid, s, t, res, bad_ex, h, mode)
            assert b1[0] > 0 or not b1[3], mode
        # (b') a residual off by one with the stored excess of its row
        # following it: pair sums and row sums disagree, only the pair side
        # of the fingerprint sees it
        e3 = int(fwd[owner[fwd] == v][0])
        bad_res = res.copy()
        bad_res[e3] += 1
        bad_ex = ex.copy()
        bad_ex[owner[e3]] += 1
        for mode in (1, 3, 4):
            b1 = eng.dbg_certify(gid, s, t, bad_res, bad_ex, h, mode)
            assert b1[0] > 0 or not b1[3], mode
        assert eng.dbg_certify(gid, s, t, bad_res, bad_ex, h, 2)[0] > 0
        # (c) a negative residual
        bad_res = res.copy()
        bad_res[int(rng.choice(np.nonzero(res > 0)[0]))] = -1
        for mode in (1, 2, 3, 4):
            assert eng.dbg_certify(gid, s, t, bad_res, ex, h, mode)[0] > 0
        # (d) heights claiming s reaches t (s inside T)
        bad_h = h.copy()
        bad_h[s] = 0
        for mode in (1, 2, 3, 4):
            assert eng.dbg_certify(gid, s, t, res, ex, bad_h, mode)[0] > 0
        # (e) one unit pushed back across the cut (pair sum and stored
        # excesses follow it): a residual entry now enters T
        T = h < n
        back = np.nonzero(T[owner] & ~T[og.col] & (res > 0) & (res[og.rev] == 0))[0]
        if len(back):
            e = int(back[0])
            rv = int(og.rev[e])
            bad_res, bad_ex = res.copy(), ex.copy()
            bad_res[e] -= 1
            bad_res[rv] += 1
            bad_ex[owner[e]] -= 1
            bad_ex[og.col[e]] += 1
            for mode in (1, 2, 3, 4):
                got = eng.dbg_certify(gid, s, t, bad_res, bad_ex, h, mode)
                assert got != (0, r.flow, r.flow, True), mode


def test_fused_cycles_replay_captured_graphs():
    """Deep graphs fuse discharge cycles; a repeated set of fused slots is
    replayed from a captured HIP graph.  Flows equal the oracle's and the
    same solve without graphs or without fusing."""
    case = Case(M.Edges.grid(96, 96, 1000, 2101), R=32, tile_max=512)
    qs, qt = M.grid_queries(96, 96, 4, 2102)
    q = [case.q(a, b) for a, b in zip(qs, qt)]
    want = [case.oracle_flow(s, t) for s, t in q]
    # the reference's path-push switch on deep graphs too, so that the
    # discharge runs fused cycles long enough to repeat
    ref_paths = dict(path_switch_deep=0, p_count_deep=64)
    for prm in (dict(), dict(graphs=0), dict(fuse=1)):
        eng, gid = engine_with(case, **ref_paths, **prm)
        got = [eng.solve(gid, s, t) for s, t in q]
        assert [r.flow for r in got] == want, prm
        assert all(r.cert_ok == 1 for r in got)
        cap, rep = eng.graph_stats()
        log(f"{prm}: graphs captured {cap}, replays {rep}")
        if not prm:
            assert rep > 0
        else:
            assert rep == 0
    # batched: several slots fusing together
    eng, gid = engine_with(case, nslots=4)
    for i, (s, t) in enumerate(q):
        eng.submit(gid, s, t, i)
    res = {r.tag: r.flow for r in eng.run()}
    assert [res[i] for i in range(4)] == want


def test_step_log_reports_each_step(c1):
    """The per-step log: one record per slot and step with the query's tag,
    what the step ran (pruning, init + preflow + relabel seeding first, then
    relabel levels, counts and discharges, the certificate last), counters
    and, while profiling, the step's stage times."""
    eng, gid = engine_with(c1, nslots=2)
    seen = []
    eng.set_step_log(lambda L: seen.append((L.tag, L.slot, L.steps, L.iters, L.active, L.K, L.gr_freq,
                                            [L.ms[i] for i in range(7)])))
    eng.profile(True)
    qs, qt = M.uniform_queries(c1.n, 2, 1011)
    for k, (a, b) in enumerate(zip(qs, qt)):
        eng.submit(gid, *c1.q(a, b), 70 + k)
    res = {r.tag: r for r in eng.run()}
    eng.set_step_log(None)
    eng.profile(False)
    assert sorted(res) == [70, 71]
    for tag in (70, 71):
        mine = [x for x in seen if x[0] == tag]
        bits = [x[2] for x in mine]
        first = next(i for i, b in enumerate(bits) if b & M.MF_STEP_PREFLOW)
        assert bits[first] & M.MF_STEP_INIT and bits[first] & M.MF_STEP_SEED
        assert all(b & (M.MF_STEP_PRUNE | M.MF_STEP_PREFLOW) for b in bits[:first + 1])
        assert any(b & M.MF_STEP_DISCHARGE for b in bits)
        assert bits[-1] & M.MF_STEP_CERT
        assert len({x[1] for x in mine}) == 1 and mine[-1][3] == res[tag].iters
        assert all(x[5] >= 1 and x[6] >= 1 for x in mine)
        assert all(min(x[7]) >= 0 for x in mine) and sum(sum(x[7]) for x in mine) > 0
        assert res[tag].flow == c1.oracle_flow(*c1.q(qs[tag - 70], qt[tag - 70]))
    seen.clear()
    eng.solve(gid, *c1.q(qs[0], qt[0]))
    assert not seen  # log off
