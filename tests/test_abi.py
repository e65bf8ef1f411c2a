"""The C-ABI library loads and exports every symbol include/mscmf.h declares;
error paths that need no GPU behave (CPU only, no compute calls)."""
import ctypes as C
import os
import re
import subprocess

import pytest

import mscresearch_amd as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "mscmf.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mf_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = M.lib()
    names = declared_symbols()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", M.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (mf_\w+)", out))
    assert set(names) <= exported


def test_bindings_cover_the_header():
    assert set(declared_symbols()) == set(M.SIGNATURES)


def test_no_gpu_errors_are_reported_not_aborted():
    if M.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(M.MaxFlowError) as ei:
        M.Engine(1)
    assert ei.value.code == M.MF_ENODEV
    hg = M.HostGraph.from_edges(3, [0, 1], [1, 2], [1, 1])
    with pytest.raises(M.MaxFlowError) as ei:
        M.DeviceGraph(hg)
    assert ei.value.code == M.MF_ENODEV


def test_invalid_arguments():
    with pytest.raises(M.MaxFlowError) as ei:
        M.HostGraph.from_edges(3, [0, 1], [1, 1], [1, 1])  # self-loop
    assert ei.value.code == M.MF_EINVAL
    with pytest.raises(M.MaxFlowError):
        M.HostGraph.from_edges(3, [0], [5], [1])
    with pytest.raises(M.MaxFlowError):
        M.HostGraph.from_edges(3, [0], [1], [0])
    hg = M.HostGraph.from_edges(3, [0], [1], [2]).partition(2)
    with pytest.raises(M.MaxFlowError):
        hg.partition(2)  # already partitioned
    assert M.lib().mf_version() == 1


def test_params_defaults_are_the_reference_constants():
    p = M.default_params()
    assert (p.gr_freq, p.h_cut, p.stagnate_rounds, p.path_switch, p.path_len) == (128, 100, 128, 8, 256)
    assert (p.hot_k, p.hot_min_active, p.rm_freq0, p.rm_margin, p.iter_max) == (32, 64, 64, 4, 20000000)
    assert p.prune == 1 and abs(p.prune_ratio - 0.995) < 1e-12 and p.cert == 3
    assert (p.local_relabel, p.local_band) == (0, 128)  # band limit of BC:1611; off by default
    # additions without a reference constant: fused cycles, graph replay,
    # speculative certificate, sparse counts
    assert (p.fuse, p.graphs, p.spec_cert, p.sparse_count, p.small_wgs) == (8, 1, 1, 1, 4)
    assert p.small_edges == 1 << 13 and p.small_edges_deep == 1 << 17 and p.mark_list == 1 << 20
    # fresh-row relabel tests on, the list discharge as a grid per slot
    assert (p.bu_recv, p.dis_queue, p.giant_rounds) == (1, 0, 2)


def test_cli_usage_and_errors(tmp_path):
    cli = M.CLI_PATH
    r = subprocess.run([cli], capture_output=True, text=True)
    assert r.returncode == 1 and "Usage" in r.stderr
    (tmp_path / "in").mkdir()
    (tmp_path / "in" / "graph.txt").write_text("3 2\n1 2 5\n2 3 4\n")
    r = subprocess.run([cli, str(tmp_path / "in"), str(tmp_path / "out")], capture_output=True, text=True)
    assert r.returncode == 1 and "Cannot open" in r.stderr  # no test.txt
    (tmp_path / "in" / "test.txt").write_text("1 3\n")
    r = subprocess.run([cli, str(tmp_path / "in"), str(tmp_path / "out"), "bogus"], capture_output=True, text=True)
    assert r.returncode == 1 and "Unknown mode" in r.stderr
    r = subprocess.run([cli, str(tmp_path / "nope"), str(tmp_path / "out")], capture_output=True, text=True)
    assert r.returncode == 1
    assert (tmp_path / "out").is_dir()  # output dir created, as the reference does


def test_struct_layouts_match_the_library():
    p, r = C.c_int64(0), C.c_int64(0)
    M.lib().mf_struct_sizes(C.byref(p), C.byref(r))
    assert (p.value, r.value) == (C.sizeof(M.Params), C.sizeof(M.Result))
