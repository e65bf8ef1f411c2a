"""SCC contraction, query split/convert and exact routing (CPU).

The oracle's Kosaraju restatement is pinned three ways: the reference's own
tools built from their sources (oracle/_ref, byte-identical .map and
graph_scc.txt; skipped where they are not built), scipy's strong components
(the same partition) and the topological order Kosaraju ids carry (every
contracted entry has cu < cv), with brute-force closures on tiny graphs.  The
library (libmscmf, host code) and bin/mscmf_scc are then held to the oracle
and to the reference's files byte for byte.  Exact routing is checked
against the oracle's max flow on the full graph and on the component alone.
"""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components

import mscresearch_amd as M
import oracle as O

REF = os.path.join(os.path.dirname(O.__file__), "_ref")


def ref_tool(name):
    p = os.path.join(REF, name)
    if not os.path.exists(p):
        pytest.skip(f"reference tool {name} not built (make -C oracle ref)")
    return p


def write_graph(path, n, lines, m_header=None):
    with open(path, "w") as f:
        f.write(f"{n} {len(lines) if m_header is None else m_header}\n")
        for u, v, c in lines:
            f.write(f"{u} {v} {c}\n")


def random_lines(rng, n, m, kind):
    """1-based lines with the reader's edge cases mixed in."""
    if kind == "dag":
        u = rng.integers(1, n, m)
        v = u + rng.integers(1, 5, m)
    elif kind == "blocks":
        # dense blocks of 8 vertices wired in a chain of blocks, some back edges
        b = rng.integers(0, n // 8, m)
        u = 8 * b + rng.integers(1, 9, m)
        v = 8 * b + rng.integers(1, 9, m)
        jump = rng.random(m) < 0.2
        v[jump] = np.minimum(n, u[jump] + 8 * rng.integers(1, 3, jump.sum()))
        back = rng.random(m) < 0.01
        v[back] = np.maximum(1, u[back] - 8 * rng.integers(1, 4, back.sum()))
    else:
        u = rng.integers(1, n + 1, m)
        v = rng.integers(1, n + 1, m)
    c = rng.integers(1, 1000, m)
    lines = [(int(a), int(b), int(x)) for a, b, x in zip(u, v, c)]
    # self-loops, zero and negative capacities, ids out of range, duplicates
    lines += [(3, 3, 5), (1, 2, 0), (2, 1, -4), (0, 2, 7), (n + 1, 1, 9), (1, n + 5, 2)]
    lines += lines[: m // 10]
    order = rng.permutation(len(lines))
    return [lines[i] for i in order]


def parsed(path):
    n, us, vs, cs = O.parse_graph(path)
    return n, us, vs, cs


CASES = [("uniform", 60, 150), ("uniform", 300, 700), ("dag", 200, 600), ("blocks", 400, 2500),
         ("uniform", 2000, 2600), ("blocks", 4000, 30000)]


@pytest.mark.parametrize("kind,n,m", CASES)
def test_oracle_kosaraju_matches_reference_tool(tmp_path, kind, n, m):
    tool = ref_tool("scc_contract")
    rng = np.random.default_rng(n * 7 + m)
    g = tmp_path / "graph.txt"
    write_graph(g, n, random_lines(rng, n, m, kind))
    out = tmp_path / "graph_scc.txt"
    p = subprocess.run([tool, str(g), str(out)], capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
    nn, us, vs, cs = parsed(str(g))
    comp, C = O.scc_kosaraju(nn, us, vs)
    want_map = "".join(f"{v + 1} {comp[v] + 1}\n" for v in range(nn))
    assert open(str(out) + ".map").read() == want_map
    cu, cv, cc = O.scc_contract(us, vs, cs, comp)
    want = f"{C} {len(cu)}\n" + "".join(f"{a + 1} {b + 1} {x}\n" for a, b, x in zip(cu, cv, cc))
    assert open(out).read() == want


def test_oracle_reads_header_stop_and_skip_rules_like_the_tool(tmp_path):
    """Fewer lines than the header says (the read stops), a malformed line
    (it stops there too) and more lines than the header (the rest ignored)."""
    tool = ref_tool("scc_contract")
    rng = np.random.default_rng(5)
    for variant in ("short", "malformed", "long"):
        lines = random_lines(rng, 40, 120, "uniform")
        g = tmp_path / f"g_{variant}.txt"
        if variant == "short":
            write_graph(g, 40, lines, m_header=len(lines) + 50)
        elif variant == "long":
            write_graph(g, 40, lines, m_header=len(lines) // 2)
        else:
            write_graph(g, 40, lines)
            txt = open(g).read().splitlines()
            txt.insert(30, "7 x 3")
            open(g, "w").write("\n".join(txt) + "\n")
        out = tmp_path / f"o_{variant}.txt"
        assert subprocess.run([tool, str(g), str(out)], capture_output=True).returncode == 0
        nn, us, vs, cs = parsed(str(g))
        comp, C = O.scc_kosaraju(nn, us, vs)
        assert open(str(out) + ".map").read() == "".join(f"{v + 1} {comp[v] + 1}\n" for v in range(nn))


def brute_scc(n, us, vs):
    """Reachability closure: u ~ v iff each reaches the other."""
    R = np.eye(n, dtype=bool)
    R[us, vs] = True
    for k in range(n):
        R |= R[:, [k]] & R[[k], :]
    return R & R.T


@pytest.mark.parametrize("seed", range(6))
def test_oracle_kosaraju_pins(seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(5, 40))
    m = int(rng.integers(0, 3 * n))
    us = rng.integers(0, n, m).astype(np.int32)
    vs = rng.integers(0, n, m).astype(np.int32)
    keep = us != vs
    us, vs = us[keep], vs[keep]
    caps = rng.integers(1, 50, len(us))
    comp, C = O.scc_kosaraju(n, us, vs)
    assert comp.min() >= 0 and comp.max() == C - 1
    # same partition as the closure and as scipy
    same = brute_scc(n, us, vs)
    assert np.array_equal(same, comp[:, None] == comp[None, :])
    A = sp.csr_matrix((np.ones(len(us)), (us, vs)), shape=(n, n))
    k, lab = connected_components(A, directed=True, connection="strong")
    assert k == C
    pairs = set(zip(comp.tolist(), lab.tolist()))
    assert len(pairs) == C
    # ids are a topological order of the condensation; sums are brute force
    cu, cv, cc = O.scc_contract(us, vs, caps, comp)
    assert np.all(cu < cv)
    want = {}
    for a, b, x in zip(us, vs, caps):
        if comp[a] != comp[b]:
            want[(comp[a], comp[b])] = want.get((comp[a], comp[b]), 0) + int(x)
    assert sorted(want) == list(zip(cu.tolist(), cv.tolist()))
    assert [want[k_] for k_ in zip(cu.tolist(), cv.tolist())] == cc.tolist()


def test_oracle_kosaraju_planted_mistake_is_caught():
    """Numbering in finishing order instead of reverse finishing order breaks
    the topological property the pins check."""
    us = np.array([0, 1, 2], np.int32)
    vs = np.array([1, 2, 0], np.int32)
    us = np.concatenate([us, [2, 3]]).astype(np.int32)
    vs = np.concatenate([vs, [3, 4]]).astype(np.int32)
    comp, C = O.scc_kosaraju(5, us, vs)
    assert C == 3
    wrong = (C - 1) - comp  # the same partition, ids reversed
    cu, cv, _ = O.scc_contract(us, vs, np.ones(len(us), np.int64), wrong)
    assert not np.all(cu < cv)


@pytest.mark.parametrize("which", ["rmat12", "grid", "blocks"])
def test_library_scc_matches_oracle(which):
    if which == "rmat12":
        n, us, vs, cs = M.Edges.rmat(12, 10**6, 7101).arrays()
    elif which == "grid":
        n, us, vs, cs = M.Edges.grid(40, 50, 100, 7102).arrays()
    else:
        rng = np.random.default_rng(7103)
        lines = random_lines(rng, 3000, 20000, "blocks")
        lines = [(u - 1, v - 1, c) for u, v, c in lines if 1 <= u <= 3000 and 1 <= v <= 3000 and u != v and c > 0]
        n = 3000
        us, vs, cs = (np.array(x) for x in zip(*lines))
    sc = M.SccContract.from_lines(n, us, vs, cs)
    comp, C = O.scc_kosaraju(n, us, vs)
    assert sc.ncomp == C
    assert np.array_equal(sc.comp(), comp)
    got = sc.edges()
    want = O.scc_contract(us, vs, cs, comp)
    for a, b in zip(got, want):
        assert np.array_equal(a, b)


def write_queries(path, pairs):
    with open(path, "w") as f:
        for a, b in pairs:
            f.write(f"{a} {b}\n")


@pytest.mark.parametrize("kind,n,m", [("blocks", 400, 2500), ("uniform", 300, 500)])
def test_scc_cli_is_byte_identical_to_the_reference_tools(tmp_path, kind, n, m):
    tools = {k: ref_tool(k) for k in ("scc_contract", "split_queries_by_scc", "convert_test_to_scc")}
    rng = np.random.default_rng(n + m)
    g = tmp_path / "graph.txt"
    write_graph(g, n, random_lines(rng, n, m, kind))
    pairs = [(int(a), int(b)) for a, b in zip(rng.integers(1, n + 1, 200), rng.integers(1, n + 1, 200))]
    pairs += [(0, 5), (n + 3, 1), (2, n + 1), (-1, 4), (7, 7)]
    q = tmp_path / "test.txt"
    write_queries(q, pairs)
    cli = M.SCC_CLI_PATH
    runs = {}
    for who, cmd in (("ref", [tools["scc_contract"]]), ("ours", [cli, "contract"])):
        d = tmp_path / who
        d.mkdir()
        p = subprocess.run(cmd + [str(g), str(d / "graph_scc.txt")], capture_output=True, text=True)
        assert p.returncode == 0, p.stderr
        err = p.stderr.replace(str(d), "D")
        p2 = subprocess.run(([tools["convert_test_to_scc"]] if who == "ref" else [cli, "convert"]) +
                            [str(q), str(d / "graph_scc.txt.map"), str(d / "test_scc.txt")],
                            capture_output=True, text=True)
        assert p2.returncode == 0, p2.stderr
        p3 = subprocess.run(([tools["split_queries_by_scc"]] if who == "ref" else [cli, "split"]) +
                            [str(q), str(d / "graph_scc.txt.map"), str(d / "inter.txt"), str(d / "intra.txt")],
                            capture_output=True, text=True)
        assert p3.returncode == 0, p3.stderr
        runs[who] = (d, err + p2.stderr.replace(str(d), "D") + p3.stderr)
    for f in ("graph_scc.txt", "graph_scc.txt.map", "test_scc.txt", "inter.txt", "intra.txt", "inter.txt.index"):
        assert (runs["ref"][0] / f).read_bytes() == (runs["ours"][0] / f).read_bytes(), f
    assert runs["ref"][1] == runs["ours"][1]


def test_scc_cli_errors_like_the_reference(tmp_path):
    cli = M.SCC_CLI_PATH
    assert subprocess.run([cli], capture_output=True).returncode == 1
    p = subprocess.run([cli, "contract", str(tmp_path / "missing.txt"), str(tmp_path / "o.txt")],
                       capture_output=True, text=True)
    assert p.returncode == 1 and "Cannot open" in p.stderr
    bad = tmp_path / "bad.txt"
    bad.write_text("x y\n")
    p = subprocess.run([cli, "contract", str(bad), str(tmp_path / "o.txt")], capture_output=True, text=True)
    assert p.returncode == 1 and "Invalid header" in p.stderr


def reach(og, s):
    seen = np.zeros(og.n, bool)
    seen[s] = True
    st = [s]
    while st:
        x = st.pop()
        for k in range(og.row_ptr[x], og.row_ptr[x + 1]):
            y = og.col[k]
            if og.cap[k] > 0 and not seen[y]:
                seen[y] = True
                st.append(y)
    return seen


@pytest.mark.parametrize("kind,n,m", [("blocks", 400, 2500), ("uniform", 200, 260), ("dag", 150, 400)])
def test_route_is_exact(kind, n, m):
    """zero => max flow 0; intra => the component alone has the same max flow
    as the graph; inter => t reachable, another component."""
    rng = np.random.default_rng(3 * n + m)
    lines = [(u - 1, v - 1, c) for u, v, c in random_lines(rng, n, m, kind)
             if 1 <= u <= n and 1 <= v <= n and u != v and c > 0]
    us, vs, cs = (np.array(x) for x in zip(*lines))
    sc = M.SccContract.from_lines(n, us, vs, cs)
    comp = sc.comp()
    og = O.build_csr(n, us, vs, cs)
    qs = rng.integers(0, n, 120).astype(np.int32)
    qt = rng.integers(0, n, 120).astype(np.int32)
    # make sure some queries fall inside a component of several vertices
    big = np.flatnonzero(np.bincount(comp)[comp] > 2)
    if len(big):
        a = rng.choice(big, 20)
        for i, x in enumerate(a):
            ys = np.flatnonzero(comp == comp[x])
            qs[i], qt[i] = x, rng.choice(ys)
    qs[-1], qt[-1] = 0, n  # out of range
    r = sc.route(qs, qt)
    assert r[-1] == -1
    kinds = set()
    for s, t, k in zip(qs[:-1], qt[:-1], r[:-1]):
        s, t = int(s), int(t)
        if s == t:
            assert k == -1
            continue
        kinds.add(int(k))
        f = O.maxflow(og, s, t)
        assert k == (0 if not reach(og, s)[t] else 1 if comp[s] == comp[t] else 2)
        if k == 0:
            assert f == 0
        if k == 1:
            sub, o2s, _ = O.induced(og, (comp == comp[s]).astype(np.int32))
            assert O.maxflow(sub, int(o2s[s]), int(o2s[t])) == f
    assert kinds & {1, 2}


def test_route_cli(tmp_path):
    rng = np.random.default_rng(77)
    n = 400
    lines = random_lines(rng, n, 2500, "blocks")
    g = tmp_path / "graph.txt"
    write_graph(g, n, lines)
    pairs = [(int(a), int(b)) for a, b in zip(rng.integers(1, n + 1, 100), rng.integers(1, n + 1, 100))]
    q = tmp_path / "test.txt"
    write_queries(q, pairs)
    p = subprocess.run([M.SCC_CLI_PATH, "route", str(g), str(q), str(tmp_path / "routes.txt")],
                       capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
    sc = M.SccContract.load(str(g))
    r = sc.route(np.array([a - 1 for a, _ in pairs]), np.array([b - 1 for _, b in pairs]))
    names = {-1: "invalid", 0: "zero", 1: "intra", 2: "inter"}
    assert open(tmp_path / "routes.txt").read() == "".join(f"{i + 1} {names[int(k)]}\n" for i, k in enumerate(r))
    assert p.stderr.startswith(f"[Route] total queries={len(pairs)}, ")
