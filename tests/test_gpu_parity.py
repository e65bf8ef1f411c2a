"""GPU parity of the deterministic pieces, bit-exact against the oracle.

Residual init + preflow (BC:151-200), global-relabel BFS heights
(BC:203-233, BC:322-348), the pruning mask (BC:1391-1445) and the induced
subgraph (BC:2477-2551) are deterministic, so the HIP results must equal the
oracle's array for array.
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


@pytest.fixture(scope="module")
def rmat12():
    log("case: rmat scale 12 seed 1201, R=64")
    return Case(M.Edges.rmat(12, 10**6, 1201), R=64)


@pytest.fixture(scope="module")
def grid():
    log("case: grid 48x64 seed 1301, R=32")
    return Case(M.Edges.grid(48, 64, 1000, 1301), R=32, tile_max=256)


def test_init_preflow_bit_exact(c1):
    e, gid = engine_with(c1)
    s_, t_ = M.uniform_queries(c1.n, 8, 1002)
    for a, b in zip(s_, t_):
        s, t = c1.q(a, b)
        res, ex = e.dbg_init_preflow(gid, s, t, c1.n, c1.E)
        ores, oex = O.init_preflow(c1.og, s)
        assert np.array_equal(res, ores)
        assert np.array_equal(ex, oex)


DIRECTIONS = {
    "auto": {},
    # every level bottom-up (switch at once, never back): runs of bottom-up
    # levels read the frontier bitmaps the levels before them wrote
    "bottom_up": dict(bu_alpha=1e12, bu_beta=1e12),
    # ... testing candidate parents by their levels instead of the bitmaps
    "bottom_up_levels": dict(bu_alpha=1e12, bu_beta=1e12, bu_bits=0),
}


@pytest.mark.parametrize("direction", list(DIRECTIONS))
@pytest.mark.parametrize("which", ["c1", "rmat12", "grid"])
def test_global_relabel_heights_bit_exact(which, direction, request):
    """Relabel heights equal the oracle's BFS distances, with the direction-
    optimizing switch and with every level bottom-up."""
    case = request.getfixturevalue(which)
    e, gid = engine_with(case, **DIRECTIONS[direction])
    rng = np.random.default_rng(77)
    if which == "rmat12":
        s_, t_ = case.edges.hub_queries(4, 1202)
    elif which == "grid":
        s_, t_ = M.grid_queries(48, 64, 4, 1302)
    else:
        s_, t_ = M.uniform_queries(case.n, 4, 1003)
    for a, b in zip(s_, t_):
        s, t = case.q(a, b)
        log(f"GR {which} {direction} s={s} t={t}")
        res, ex = O.init_preflow(case.og, s)
        h = e.dbg_global_relabel(gid, s, t, res, ex)
        assert np.array_equal(h, O.bfs_heights(case.og, res, ex, s, t))
        # a later residual (the oracle's maximum flow) plus planted deficits
        f, res2 = O.maxflow(case.og, s, t, want_residual=True)
        ex2 = np.zeros(case.n, np.int64)
        pick = rng.choice(case.n, 5, replace=False)
        ex2[pick] = -3
        h2 = e.dbg_global_relabel(gid, s, t, res2, ex2)
        assert np.array_equal(h2, O.bfs_heights(case.og, res2, ex2, s, t))


@pytest.mark.parametrize("which", ["c1", "rmat12", "grid"])
def test_prune_mask_bit_exact(which, request):
    case = request.getfixturevalue(which)
    e, gid = engine_with(case)
    if which == "rmat12":
        s_, t_ = case.edges.hub_queries(6, 1203)
    elif which == "grid":
        s_, t_ = M.grid_queries(48, 64, 3, 1303)
    else:
        s_, t_ = M.uniform_queries(case.n, 3, 1004)
    for a, b in zip(s_, t_):
        s, t = case.q(a, b)
        mask, kept = e.dbg_prune_mask(gid, s, t, case.n)
        om = O.prune_mask(case.og, s, t)
        assert np.array_equal(mask, om)
        assert kept == int(om.sum())


def unreachable_pairs(case, k, seed):
    """(s, t) in region-ordered ids with t not reachable from s over cap > 0:
    s without a cap > 0 out-entry, or t without a cap > 0 in-entry."""
    og = case.og
    owner = np.repeat(np.arange(og.n), np.diff(og.row_ptr))
    fwd = og.cap > 0
    outdeg = np.bincount(owner[fwd], minlength=og.n)
    indeg = np.bincount(og.col[fwd], minlength=og.n)
    rng = np.random.default_rng(seed)
    pairs = []
    sinks = np.nonzero((outdeg == 0) & (indeg > 0))[0]
    sources = np.nonzero((indeg == 0) & (outdeg > 0))[0]
    hub = int(np.argmax(outdeg))
    for v in rng.permutation(sinks)[:k]:
        pairs.append((int(v), hub if hub != v else int((v + 1) % og.n)))
    for v in rng.permutation(sources)[:k]:
        pairs.append((hub if hub != v else int((v + 1) % og.n), int(v)))
    return pairs


@pytest.mark.parametrize("which", ["c1", "rmat12", "grid"])
def test_class_mask_bit_exact(which, request):
    """The mask the product prunes with (the forward set of s's SCC AND the
    WCC of t) equals the reference mask of BC:1391-1445 whenever t is
    reachable from s; when it is not, the product answers 0 without a mask,
    and the oracle's flow is 0.  Each pair runs twice: the second call finds
    s's forward set in the graph's cache.  A solve through the scheduler
    reports the same kept count."""
    case = request.getfixturevalue(which)
    e, gid = engine_with(case)
    if which == "rmat12":
        s_, t_ = case.edges.hub_queries(6, 1205)
    elif which == "grid":
        s_, t_ = M.grid_queries(48, 64, 4, 1305)
    else:
        s_, t_ = M.uniform_queries(case.n, 4, 1005)
    pairs = [case.q(a, b) for a, b in zip(s_, t_)] + unreachable_pairs(case, 3, 1206)
    n_unreach = 0
    for s, t in pairs:
        om = O.prune_mask(case.og, s, t)
        f = O.maxflow(case.og, s, t)
        for rep in range(2):
            log(f"class mask {which} s={s} t={t} pass {rep}")
            mask, kept, reach = e.dbg_class_mask(gid, s, t, case.n)
            assert reach == (f > 0), (s, t, f)
            if reach:
                assert np.array_equal(mask, om)
                assert kept == int(om.sum())
        n_unreach += 0 if f > 0 else 1
        r = e.solve(gid, s, t)
        assert r.flow == f and r.cert_ok == 1
        if f > 0:
            assert r.kept == int(om.sum())
    if which == "rmat12":
        assert n_unreach > 0  # the unreachable branch was exercised


def test_induced_subgraph_bit_exact(rmat12):
    e, gid = engine_with(rmat12)
    s_, t_ = rmat12.edges.hub_queries(3, 1204)
    rng = np.random.default_rng(5)
    masks = [O.prune_mask(rmat12.og, *rmat12.q(a, b)) for a, b in zip(s_, t_)]
    masks.append((rng.random(rmat12.n) < 0.5).astype(np.int32))
    for mask in masks:
        got = e.dbg_induced(gid, mask, rmat12.E)
        sub, o2s, _ = O.induced(rmat12.og, mask)
        assert got["n"] == sub.n and got["E"] == sub.E
        for k in ("row_ptr", "col", "cap", "rev"):
            assert np.array_equal(got[k], getattr(sub, k)), k
        assert np.array_equal(got["old2sub"], o2s)


def tile_of(case, tile_max):
    """Tiles as the upload cuts them: each region in runs of tile_max."""
    rs = case.d["rstart"]
    starts = np.concatenate([np.arange(rs[r], rs[r + 1], tile_max) for r in range(len(rs) - 1)])
    return np.searchsorted(starts, np.arange(case.n), side="right") - 1, len(starts)


def local_relabel_expected(og, res, ex, h, tof, active, s, t, lcap):
    """Largest valid labelling of R' = {v in active tiles, h(v) < lcap, v != s}
    below lcap, heights outside R' fixed, by Dijkstra with unit weights from
    the seeds (h == 0: t or a deficit) and from the fixed heights around R'."""
    import heapq

    n = og.n
    rp, col, rev = og.row_ptr, og.col, og.rev
    inR = active[tof] > 0
    Rp = inR & (h < lcap)
    Rp[s] = False
    seed = Rp & (h == 0) & ((np.arange(n) == t) | (ex < 0))
    marked = Rp & ~seed
    dist = np.full(n, np.iinfo(np.int64).max, np.int64)
    heap = []
    for u in np.nonzero(marked)[0]:
        best = dist[u]
        for e in range(rp[u], rp[u + 1]):
            x = col[e]
            if res[e] > 0 and not marked[x]:
                best = min(best, int(h[x]) + 1)
        dist[u] = best
        if best < lcap:
            heap.append((best, int(u)))
    for u in np.nonzero(seed)[0]:
        dist[u] = 0
        heap.append((0, int(u)))
    heapq.heapify(heap)
    while heap:
        d, v = heapq.heappop(heap)
        if d != dist[v] or d + 1 >= lcap:
            continue
        for e in range(rp[v], rp[v + 1]):
            w = col[e]
            if marked[w] and res[rev[e]] > 0 and dist[w] > d + 1:
                dist[w] = d + 1
                heapq.heappush(heap, (d + 1, int(w)))
    out = h.astype(np.int64).copy()
    m = np.nonzero(marked)[0]
    out[m] = np.maximum(h[m], np.minimum(dist[m], lcap))
    return out.astype(np.int32), marked


@pytest.mark.parametrize("which", ["c1", "rmat12", "grid"])
def test_local_relabel_bit_exact(which, request):
    """The local relabel only raises heights, keeps a valid labelling valid
    and gives the largest valid labelling of the relabelled set below lcap."""
    case = request.getfixturevalue(which)
    tmax = 256 if which == "grid" else 4096
    tof, ntiles = tile_of(case, tmax)
    assert ntiles == case.dg.info()["ntiles"]
    e, gid = engine_with(case)
    rng = np.random.default_rng(91)
    if which == "rmat12":
        s_, t_ = case.edges.hub_queries(3, 1204)
    elif which == "grid":
        s_, t_ = M.grid_queries(48, 64, 3, 1304)
    else:
        s_, t_ = M.uniform_queries(case.n, 3, 1004)
    n = case.n
    og = case.og
    owner = np.repeat(np.arange(n), np.diff(og.row_ptr))
    for a, b in zip(s_, t_):
        s, t = case.q(a, b)
        for state in ("preflow", "maxflow"):
            if state == "preflow":
                res, ex = O.init_preflow(og, s)
            else:
                _, res = O.maxflow(og, s, t, want_residual=True)
                ex = np.zeros(n, np.int64)
                ex[rng.choice(n, 4, replace=False)] = -2
            hgr = O.bfs_heights(og, res, ex, s, t)
            for shrink in ("half", "clip"):
                h0 = (hgr // 2) if shrink == "half" else np.minimum(hgr, 3)
                h0 = h0.astype(np.int32)
                h0[s] = n
                for act, lcap in (("some", 5), ("some", min(256, n)), ("all", min(256, n))):
                    active = (rng.random(ntiles) < 0.6).astype(np.int32) if act == "some" else np.ones(ntiles, np.int32)
                    log(f"local {which} s={s} t={t} {state} {shrink} {act} lcap={lcap}")
                    h1, lvl = e.dbg_local_relabel(gid, s, t, res, ex, h0, active, lcap)
                    assert lvl == lcap
                    want, marked = local_relabel_expected(og, res, ex, h0, tof, active, s, t, lcap)
                    assert np.array_equal(h1, want)
                    assert np.all(h1 >= h0)
                    # validity: every residual entry u -> v off s has h(u) <= h(v) + 1
                    live = (res > 0) & (owner != s)
                    assert np.all(h1[owner[live]] <= h1[og.col[live]] + 1)
                    # fixpoint of the relabelled vertices
                    lo = np.full(n, np.iinfo(np.int64).max, np.int64)
                    np.minimum.at(lo, owner[res > 0], h1[og.col[res > 0]].astype(np.int64) + 1)
                    m = np.nonzero(marked)[0]
                    assert np.array_equal(h1[m], np.maximum(h0[m], np.minimum(lo[m], lcap)))
                    if act == "all":  # one region: the exact global relabel, capped
                        assert np.array_equal(h1[m], np.minimum(hgr[m], lcap))


@pytest.mark.parametrize("n,tile_max", [(4096, 64), (20000, 2)])
def test_hot_select_is_the_sorted_choice(n, tile_max):
    """The hot sweep's region choice (BC:1931-1946 sorts the region counts
    and takes the top 32) equals the first k tiles in (count descending,
    index ascending) order among nonzero counts; the bisection kernel must
    give exactly that set, ties included.  The second shape has more tiles
    than the kernel keeps in LDS."""
    case = Case(M.Edges.uniform(n, 4 * n, 100, 1401), R=16, tile_max=tile_max)
    eng, gid = engine_with(case)
    T = case.dg.info()["ntiles"]
    log(f"hot select: {T} tiles")
    rng = np.random.default_rng(1402)
    trials = [np.zeros(T, np.int64), np.full(T, 3)]
    trials += [rng.integers(0, 6, T) * (rng.random(T) < 0.5) for _ in range(3)]
    trials += [rng.integers(0, 4097, T) for _ in range(2)]
    for cnt in trials:
        order = sorted(np.nonzero(cnt)[0].tolist(), key=lambda i: (-int(cnt[i]), i))
        for k in (1, 4, 32, T + 3):
            want = np.zeros(T, np.int32)
            want[order[:k]] = 1
            got = eng.dbg_hot_select(gid, cnt, k)
            assert np.array_equal(got, want), (k, int(cnt.max()))
