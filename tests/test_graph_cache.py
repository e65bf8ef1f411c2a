"""Binary graph cache: the prepared graph (CSR, stats, partition, id maps)
written once and read back through mmap must equal the graph built from the
text file, and a cache that no longer matches its text file, its region
count or its own hash must be refused.  Host code only (no GPU)."""
import os

import numpy as np
import pytest

import mscresearch_amd as M


def prepared(tmp_path, R):
    e = M.Edges.rmat(10, 10**6, 7301)
    src = tmp_path / "graph.txt"
    e.write(str(src))
    hg = M.HostGraph.load(str(src))
    if R:
        hg.partition(R)
    return src, hg


def same(a, b):
    assert a.info() == b.info()
    da, db = a.export(), b.export()
    assert da.keys() == db.keys()
    for k in da:
        np.testing.assert_array_equal(da[k], db[k], err_msg=k)
    for x, y in zip(a.stats(), b.stats()):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("R", [0, 32])
def test_round_trip_equals_the_text_build(tmp_path, R):
    src, hg = prepared(tmp_path, R)
    cache = tmp_path / "graph.mfg"
    hg.save(str(cache), source=str(src))
    assert os.path.getsize(cache) % 64 == 0
    same(hg, M.HostGraph.open(str(cache), source=str(src), R=R))
    same(hg, M.HostGraph.open(str(cache)))  # no checks asked for


def test_stale_source_or_region_count_is_refused(tmp_path):
    src, hg = prepared(tmp_path, 32)
    cache = tmp_path / "graph.mfg"
    hg.save(str(cache), source=str(src))
    with pytest.raises(M.MaxFlowError) as ei:
        M.HostGraph.open(str(cache), source=str(src), R=64)
    assert ei.value.code == M.MF_ESTALE
    st = os.stat(src)
    os.utime(src, ns=(st.st_atime_ns, st.st_mtime_ns + 1_000_000_000))
    with pytest.raises(M.MaxFlowError) as ei:
        M.HostGraph.open(str(cache), source=str(src), R=32)
    assert ei.value.code == M.MF_ESTALE
    # saved without a source: any source is a mismatch
    hg.save(str(cache))
    with pytest.raises(M.MaxFlowError) as ei:
        M.HostGraph.open(str(cache), source=str(src), R=32)
    assert ei.value.code == M.MF_ESTALE


def test_corrupted_truncated_or_foreign_files_are_refused(tmp_path):
    src, hg = prepared(tmp_path, 32)
    cache = tmp_path / "graph.mfg"
    hg.save(str(cache), source=str(src))
    raw = bytearray(cache.read_bytes())
    for name, data in [("flipped", raw[:4096] + bytes([raw[4096] ^ 1]) + raw[4097:]),
                       ("truncated", raw[:len(raw) // 2]),
                       ("foreign", b"not a graph" + raw[11:]),
                       ("empty", b"")]:
        p = tmp_path / f"{name}.mfg"
        p.write_bytes(bytes(data))
        with pytest.raises(M.MaxFlowError) as ei:
            M.HostGraph.open(str(p), source=str(src), R=32)
        assert ei.value.code == M.MF_EIO, name
    with pytest.raises(M.MaxFlowError) as ei:
        M.HostGraph.open(str(tmp_path / "missing.mfg"))
    assert ei.value.code == M.MF_EIO
    with pytest.raises(M.MaxFlowError) as ei:
        hg.save(str(tmp_path / "no_such_dir" / "g.mfg"))
    assert ei.value.code == M.MF_EIO
