"""MI355X-native exact integer max flow (push-relabel) behind a C ABI.

Python bindings over ``libmscmf.so`` (include/mscmf.h).  The library is the
product: graph preparation runs in its C++ host code, every solve runs in the
HIP kernels on the GPU.  There is no CPU fallback; without the built library
or without a GPU the calls raise.

Reference interface mirrored: xiyuweiguang/MScResearch,
Region_Aware_PR_with_Pruning_BoundaryCore.cu (build_graph BC:1044,
partition BC:1130, partition_and_reorder BC:1182, init_global_graph_on_device
BC:1293, run_gpu_static_with_state BC:2119, incremental update BC:2273).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libmscmf.so")
CLI_PATH = os.path.join(_PKG, "bin", "mscmf")
SCC_CLI_PATH = os.path.join(_PKG, "bin", "mscmf_scc")

MF_OK, MF_EINVAL, MF_ENOMEM, MF_EHIP, MF_EIO, MF_EITER, MF_ECERT, MF_ENODEV, MF_EBUSY, MF_ESTALE = (
    0, -1, -2, -3, -4, -5, -6, -7, -8, -9)
NOALLOC_SCRATCH_BYTES = 256 << 20  # MF_NOALLOC_SCRATCH_BYTES

_lib = None


class MaxFlowError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [
        ("K", C.c_int), ("gr_freq", C.c_int), ("gr_freq_min", C.c_int), ("gr_freq_max", C.c_int),
        ("gr_stall", C.c_int), ("h_cut", C.c_int), ("stagnate_rounds", C.c_int),
        ("path_switch", C.c_int), ("path_len", C.c_int), ("path_attempts", C.c_int), ("p_count0", C.c_int),
        ("hot", C.c_int), ("hot_k", C.c_int), ("hot_min_active", C.c_int), ("hot_rounds", C.c_int),
        ("rm", C.c_int), ("rm_freq0", C.c_int), ("rm_freq_max", C.c_int), ("rm_margin", C.c_int),
        ("rm_tau", C.c_double), ("tile_rounds", C.c_int), ("tile_frac", C.c_double), ("cert", C.c_int),
        ("iter_max", C.c_longlong), ("bu_alpha", C.c_double), ("bu_beta", C.c_double),
        ("prune", C.c_int), ("prune_ratio", C.c_double), ("gr_work", C.c_double), ("K_heavy", C.c_int),
        ("local_relabel", C.c_int), ("local_band", C.c_int), ("fuse", C.c_int), ("graphs", C.c_int),
        ("spec_cert", C.c_int), ("sparse_count", C.c_int), ("small_wgs", C.c_int),
        ("small_edges", C.c_longlong), ("mark_list", C.c_int), ("inject", C.c_int),
        ("tile_threads", C.c_int), ("path_switch_deep", C.c_int), ("p_count_deep", C.c_int),
        ("tile_split", C.c_int), ("scc_route", C.c_int), ("bu_skip", C.c_int), ("sink_cut", C.c_int),
        ("disc_band", C.c_int), ("bu_pick", C.c_int), ("small_levels", C.c_int),
        ("push_first", C.c_int), ("slot_blocks", C.c_int), ("launch_blocks", C.c_int),
        ("bu_light", C.c_int), ("bu_bits", C.c_int), ("gr_work_deep", C.c_double),
        ("small_edges_deep", C.c_longlong), ("giant_rounds", C.c_int),
        ("bu_recv", C.c_int), ("dis_queue", C.c_int),
    ]


class StepLog(C.Structure):
    """mf_step_log: one slot's engine step (MF_STEP_* bits, counters, the
    step's stage times init / preflow / seeding / BFS / push-relabel / RM /
    certificate in ms while profiling)."""
    _fields_ = [("tag", C.c_int64), ("slot", C.c_int32), ("steps", C.c_uint32), ("iters", C.c_int64),
                ("active", C.c_int32), ("active_high", C.c_int32), ("fixed_edges", C.c_int64),
                ("K", C.c_int32), ("gr_freq", C.c_int32), ("path", C.c_int32), ("cycles", C.c_int32),
                ("ms", C.c_double * 7)]


STEP_FN = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(StepLog))
MF_STEP_INIT, MF_STEP_PREFLOW, MF_STEP_SEED, MF_STEP_BFS, MF_STEP_DISCHARGE = 1, 2, 4, 8, 16
MF_STEP_HOT, MF_STEP_RM, MF_STEP_CERT, MF_STEP_PRUNE, MF_STEP_COUNT = 32, 64, 128, 256, 512


class TuneResult(C.Structure):
    _fields_ = [("R", C.c_int), ("tile_max", C.c_int), ("K", C.c_int), ("gr_freq", C.c_int),
                ("tile_threads", C.c_int), ("slots", C.c_int), ("tile_split", C.c_int),
                ("ms_per_query", C.c_double), ("default_ms_per_query", C.c_double), ("evaluations", C.c_int),
                ("graph_class", C.c_char * 64)]


class Result(C.Structure):
    _fields_ = [
        ("tag", C.c_int64), ("flow", C.c_int64), ("status", C.c_int), ("cert_ok", C.c_int),
        ("pruned", C.c_int), ("n_run", C.c_int), ("E_run", C.c_int64), ("kept", C.c_int64),
        ("iters", C.c_int64), ("steps", C.c_int64), ("grs", C.c_int64), ("bfs_levels", C.c_int64),
        ("drains", C.c_int64), ("rm_fixed", C.c_int64),
        ("edge_scans", C.c_uint64), ("vertex_scans", C.c_uint64), ("pushes", C.c_uint64),
        ("relabels", C.c_uint64),
        ("t_arrive", C.c_double), ("t_start", C.c_double), ("t_end", C.c_double),
        ("work", (C.c_uint64 * 4) * 7), ("sink_cut", C.c_int64),
    ]

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "work"}
        d["work"] = [[int(self.work[f][k]) for k in range(4)] for f in range(7)]
        return d


_P = C.c_void_p
_I32P = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_I64P = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")

# name -> (restype, argtypes); also the list of symbols include/mscmf.h declares
SIGNATURES = {
    "mf_last_error": (C.c_char_p, []),
    "mf_version": (C.c_int, []),
    "mf_struct_sizes": (None, [C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "mf_hgraph_load": (C.c_int, [C.c_char_p, C.POINTER(_P)]),
    "mf_hgraph_from_edges": (C.c_int, [C.c_int, C.c_int64, _P, _P, _P, C.POINTER(_P)]),
    "mf_hgraph_from_edgelist": (C.c_int, [_P, C.POINTER(_P)]),
    "mf_hgraph_partition": (C.c_int, [_P, C.c_int]),
    "mf_hgraph_save": (C.c_int, [_P, C.c_char_p, C.c_char_p]),
    "mf_hgraph_open": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int, C.POINTER(_P)]),
    "mf_hgraph_info": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int64), C.POINTER(C.c_int), C.POINTER(C.c_int64)]),
    "mf_hgraph_export": (C.c_int, [_P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "mf_hgraph_stats": (C.c_int, [_P, _P, _P, _P, _P]),
    "mf_hgraph_free": (None, [_P]),
    "mf_hgraph_find_entry": (C.c_int, [_P, C.c_int, C.c_int, C.POINTER(C.c_int32)]),
    "mf_hgraph_components": (C.c_int, [_P, _P, _P]),
    "mf_hgraph_line_entries": (C.c_int, [_P, C.c_int64, _P, _P, _P]),
    "mf_hgraph_core_decision": (C.c_int, [_P, C.c_int, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                          C.POINTER(C.c_int64), C.POINTER(C.c_int)]),
    "mf_partition": (C.c_int, [C.c_int, C.c_int, _P, _P, _P, _P, _P]),
    "mf_trace_marker": (C.c_int, []),
    "mf_scc_build": (C.c_int, [C.c_int, C.c_int64, _P, _P, _P, C.POINTER(_P)]),
    "mf_scc_load": (C.c_int, [C.c_char_p, C.POINTER(_P)]),
    "mf_scc_info": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int32), C.POINTER(C.c_int64)]),
    "mf_scc_comp": (C.c_int, [_P, _P]),
    "mf_scc_edges": (C.c_int, [_P, _P, _P, _P]),
    "mf_scc_route": (C.c_int, [_P, C.c_int, _P, _P, _P]),
    "mf_scc_free": (None, [_P]),
    "mf_read_queries": (C.c_int, [C.c_char_p, C.c_int64, _P, _P, C.POINTER(C.c_int64)]),
    "mf_gen_uniform": (C.c_int, [C.c_int, C.c_int64, C.c_int64, C.c_uint64, C.POINTER(_P)]),
    "mf_gen_grid": (C.c_int, [C.c_int, C.c_int, C.c_int64, C.c_uint64, C.POINTER(_P)]),
    "mf_gen_rmat": (C.c_int, [C.c_int, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_int64, C.c_uint64, C.c_int, C.POINTER(_P)]),
    "mf_edges_info": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int64)]),
    "mf_edges_copy": (C.c_int, [_P, _P, _P, _P]),
    "mf_edges_write": (C.c_int, [_P, C.c_char_p]),
    "mf_edges_free": (None, [_P]),
    "mf_gen_queries_uniform": (C.c_int, [C.c_int, C.c_int, C.c_uint64, _P, _P]),
    "mf_gen_queries_grid": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_uint64, _P, _P]),
    "mf_gen_queries_hub": (C.c_int, [_P, C.c_double, C.c_int, C.c_uint64, _P, _P]),
    "mf_write_queries": (C.c_int, [C.c_char_p, C.c_int, _P, _P]),
    "mf_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "mf_graph_upload": (C.c_int, [_P, C.c_int, C.POINTER(_P)]),
    "mf_graph_info": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int64), C.POINTER(C.c_int), C.POINTER(C.c_int64)]),
    "mf_graph_slab_bytes": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "mf_graph_free": (None, [_P]),
    "mf_params_default": (None, [C.POINTER(Params)]),
    "mf_engine_create": (C.c_int, [C.c_int, C.POINTER(Params), C.POINTER(_P)]),
    "mf_engine_add_graph": (C.c_int, [_P, _P, C.POINTER(C.c_int)]),
    "mf_engine_free": (None, [_P]),
    "mf_solve": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.POINTER(Result)]),
    "mf_solve_noalloc": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.POINTER(Result)]),
    "mf_engine_submit": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int64, C.c_double]),
    "mf_engine_run": (C.c_int, [_P, C.c_int64, C.POINTER(Result), C.c_int, C.POINTER(C.c_int)]),
    "mf_engine_live": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int64)]),
    "mf_engine_set_budget": (C.c_int, [_P, C.c_int64]),
    "mf_engine_mem_stats": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "mf_engine_serve": (C.c_int, [_P, C.c_int, _P, _P, _P, _P, C.POINTER(Result), C.POINTER(C.c_int)]),
    "mf_serve_lanes": (C.c_int, [C.c_int, _P, C.c_int, _P, _P, _P, _P, _P, C.POINTER(Result), C.POINTER(C.c_int)]),
    "mf_engine_now": (C.c_double, [_P]),
    "mf_engine_cache_stats": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "mf_engine_profile": (C.c_int, [_P, C.c_int]),
    "mf_engine_set_step_log": (C.c_int, [_P, STEP_FN, C.c_void_p]),
    "mf_engine_graph_stats": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "mf_graph_class": (C.c_int, [_P, C.c_char_p, C.c_int]),
    "mf_autotune": (C.c_int, [_P, C.c_int, _P, _P, C.POINTER(Params), C.c_int, C.c_int, C.c_int, C.c_char_p,
                              C.POINTER(TuneResult)]),
    "mf_engine_profile_read": (C.c_int, [_P, _P, _P, _P]),
    "mf_engine_pin": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int64]),
    "mf_engine_solve_pinned": (C.c_int, [_P, C.c_int, C.POINTER(Result), C.c_int, C.POINTER(C.c_int)]),
    "mf_graph_update_caps": (C.c_int, [_P, C.c_int, C.c_int, _P, _P]),
    "mf_dbg_init_preflow": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, _P, _P]),
    "mf_dbg_global_relabel": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, _P, _P, _P]),
    "mf_dbg_local_relabel": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, _P, _P, _P, _P, C.c_int, _P]),
    "mf_dbg_certify": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, _P, _P, _P, C.c_int, _P, _P, _P, _P]),
    "mf_dbg_prune_mask": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, _P, C.POINTER(C.c_int)]),
    "mf_dbg_hot_select": (C.c_int, [_P, C.c_int, _P, C.c_int, _P]),
    "mf_dbg_class_mask": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, _P, C.POINTER(C.c_int64), C.POINTER(C.c_int)]),
    "mf_dbg_induced": (C.c_int, [_P, C.c_int, _P, C.POINTER(C.c_int), C.POINTER(C.c_int64), _P, _P, _P, _P, _P]),
    "mf_dbg_solve_state": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.POINTER(Result), _P, _P, _P]),
    "mf_dbg_barrier_ticks": (C.c_int, [_P, C.c_longlong, C.POINTER(C.c_longlong)]),
}


def lib() -> C.CDLL:
    """Load libmscmf.so (built in-tree).  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MaxFlowError(MF_EIO, f"{LIB_PATH} not built; run python -m mscresearch_amd.build")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(rc: int) -> None:
    if rc != MF_OK:
        raise MaxFlowError(rc, lib().mf_last_error().decode())


def device_count() -> int:
    c = C.c_int(0)
    _check(lib().mf_device_count(C.byref(c)))
    return c.value


def trace_marker() -> None:
    """One empty kernel (k_trace_marker), waited for: splits kernel traces
    and counter runs into phases."""
    _check(lib().mf_trace_marker())


def default_params(**over) -> Params:
    p = Params()
    lib().mf_params_default(C.byref(p))
    names = {f for f, _ in Params._fields_}
    for k, v in over.items():
        if k not in names:
            raise KeyError(f"mf_params has no field {k!r}")
        setattr(p, k, v)
    return p


class Edges:
    """A generated edge list (0-based ids, emission order)."""

    def __init__(self, h):
        self._h = h

    @staticmethod
    def uniform(n, m, cap_max, seed):
        h = _P()
        _check(lib().mf_gen_uniform(n, m, cap_max, seed, C.byref(h)))
        return Edges(h)

    @staticmethod
    def grid(rows, cols, cap_max, seed):
        h = _P()
        _check(lib().mf_gen_grid(rows, cols, cap_max, seed, C.byref(h)))
        return Edges(h)

    @staticmethod
    def rmat(scale, cap_max, seed, a=0.57, b=0.19, c=0.19, samples=None, threads=8):
        h = _P()
        samples = 16 * (1 << scale) if samples is None else samples
        _check(lib().mf_gen_rmat(scale, samples, a, b, c, cap_max, seed, threads, C.byref(h)))
        return Edges(h)

    def info(self):
        n, cnt = C.c_int(0), C.c_int64(0)
        _check(lib().mf_edges_info(self._h, C.byref(n), C.byref(cnt)))
        return n.value, cnt.value

    def arrays(self):
        n, cnt = self.info()
        us = np.empty(cnt, np.int32)
        vs = np.empty(cnt, np.int32)
        cs = np.empty(cnt, np.int64)
        _check(lib().mf_edges_copy(self._h, _ptr(us), _ptr(vs), _ptr(cs)))
        return n, us, vs, cs

    def write(self, path):
        _check(lib().mf_edges_write(self._h, str(path).encode()))

    def hub_queries(self, q, seed, top_frac=0.01):
        s = np.empty(q, np.int32)
        t = np.empty(q, np.int32)
        _check(lib().mf_gen_queries_hub(self._h, top_frac, q, seed, _ptr(s), _ptr(t)))
        return s, t

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.mf_edges_free(self._h)
            self._h = None


def autotune(hg_unpartitioned, s, t, params=None, R=128, tile=4096, cache=None, batch=0) -> dict:
    """Pick R, tile size, K, GR_FREQ, the tile workgroup size, workgroups per
    tile by excess score and (batch > 0) slots per engine for this graph's
    class by timing the plain solver (batch = 0) or the batched engine on
    sample queries (ids of the unpartitioned graph); every configuration must
    give the same flows.  Returns the chosen values."""
    s = np.ascontiguousarray(s, np.int32)
    t = np.ascontiguousarray(t, np.int32)
    p = params if params is not None else default_params()
    out = TuneResult()
    _check(lib().mf_autotune(hg_unpartitioned._h, len(s), _ptr(s), _ptr(t), C.byref(p), int(R), int(tile), int(batch),
                             cache.encode() if cache else None, C.byref(out)))
    return {"R": out.R, "tile_max": out.tile_max, "K": out.K, "gr_freq": out.gr_freq,
            "tile_threads": out.tile_threads, "slots": out.slots, "tile_split": out.tile_split,
            "ms_per_query": out.ms_per_query, "default_ms_per_query": out.default_ms_per_query,
            "evaluations": out.evaluations, "graph_class": out.graph_class.decode()}


def serve_lanes(engines, lane, gids, s, t, arrive):
    """Serving over several engines (one host thread and HIP stream each):
    query i goes to engines[lane[i]], graph gids[i] of that engine, arriving
    arrive[i] s after the call (ascending).  Results (tag = i) in completion
    order."""
    arr = (C.c_void_p * len(engines))(*[e._h.value if isinstance(e._h, C.c_void_p) else e._h for e in engines])
    lane = np.ascontiguousarray(lane, np.int32)
    gids = np.ascontiguousarray(gids, np.int32)
    s = np.ascontiguousarray(s, np.int32)
    t = np.ascontiguousarray(t, np.int32)
    arrive = np.ascontiguousarray(arrive, np.float64)
    nq = len(s)
    out = (Result * max(nq, 1))()
    k = C.c_int(0)
    _check(lib().mf_serve_lanes(len(engines), C.cast(arr, C.c_void_p), nq, _ptr(lane), _ptr(gids), _ptr(s), _ptr(t),
                                _ptr(arrive), out, C.byref(k)))
    return [out[i] for i in range(k.value)]


def uniform_queries(n, q, seed):
    s = np.empty(q, np.int32)
    t = np.empty(q, np.int32)
    _check(lib().mf_gen_queries_uniform(n, q, seed, _ptr(s), _ptr(t)))
    return s, t


def grid_queries(rows, cols, q, seed):
    s = np.empty(q, np.int32)
    t = np.empty(q, np.int32)
    _check(lib().mf_gen_queries_grid(rows, cols, q, seed, _ptr(s), _ptr(t)))
    return s, t


def write_queries(path, s, t):
    s = np.ascontiguousarray(s, np.int32)
    t = np.ascontiguousarray(t, np.int32)
    _check(lib().mf_write_queries(str(path).encode(), len(s), _ptr(s), _ptr(t)))


def read_queries(path):
    cnt = C.c_int64(0)
    _check(lib().mf_read_queries(str(path).encode(), 0, None, None, C.byref(cnt)))
    s = np.empty(cnt.value, np.int32)
    t = np.empty(cnt.value, np.int32)
    _check(lib().mf_read_queries(str(path).encode(), cnt.value, _ptr(s), _ptr(t), C.byref(cnt)))
    return s, t


def partition(R, so, si, dout, din):
    n = len(so)
    part = np.empty(n, np.int32)
    so, si = np.ascontiguousarray(so, np.int64), np.ascontiguousarray(si, np.int64)
    dout, din = np.ascontiguousarray(dout, np.int32), np.ascontiguousarray(din, np.int32)
    _check(lib().mf_partition(n, R, _ptr(so), _ptr(si), _ptr(dout), _ptr(din), _ptr(part)))
    return part


class SccContract:
    """Strongly connected components of input lines in the reference's ids
    (SCC_CONTRACT/scc_contract.cpp), the contracted graph and exact routing.

    ``comp`` holds 0-based ids (the reference's .map is v+1, comp+1); the
    contracted graph (``edges()``) sums capacities per component pair and is
    not flow-exact.  ``route(s, t)``: 0 = t unreachable from s (max flow 0),
    1 = same component (solve on it alone), 2 = other component, reachable,
    -1 = bad query."""

    def __init__(self, h):
        self._h = h
        n, c, m = C.c_int(0), C.c_int32(0), C.c_int64(0)
        _check(lib().mf_scc_info(h, C.byref(n), C.byref(c), C.byref(m)))
        self.n, self.ncomp, self.m = n.value, c.value, m.value

    @classmethod
    def from_lines(cls, n, us, vs, caps):
        us = np.ascontiguousarray(us, np.int32)
        vs = np.ascontiguousarray(vs, np.int32)
        caps = np.ascontiguousarray(caps, np.int64)
        h = _P()
        _check(lib().mf_scc_build(int(n), len(us), _ptr(us), _ptr(vs), _ptr(caps), C.byref(h)))
        return cls(h)

    @classmethod
    def load(cls, path):
        h = _P()
        _check(lib().mf_scc_load(str(path).encode(), C.byref(h)))
        return cls(h)

    def comp(self):
        out = np.empty(self.n, np.int32)
        _check(lib().mf_scc_comp(self._h, _ptr(out)))
        return out

    def edges(self):
        cu, cv, cap = np.empty(self.m, np.int32), np.empty(self.m, np.int32), np.empty(self.m, np.int64)
        _check(lib().mf_scc_edges(self._h, _ptr(cu), _ptr(cv), _ptr(cap)))
        return cu, cv, cap

    def route(self, s, t):
        s = np.ascontiguousarray(s, np.int32)
        t = np.ascontiguousarray(t, np.int32)
        out = np.empty(len(s), np.int32)
        _check(lib().mf_scc_route(self._h, len(s), _ptr(s), _ptr(t), _ptr(out)))
        return out

    def __del__(self):
        if getattr(self, "_h", None):
            lib().mf_scc_free(self._h)
            self._h = None


class HostGraph:
    """Host CSR with reverse entries (+ region partition once partitioned)."""

    def __init__(self, h):
        self._h = h

    @staticmethod
    def load(path):
        h = _P()
        _check(lib().mf_hgraph_load(str(path).encode(), C.byref(h)))
        return HostGraph(h)

    @staticmethod
    def from_edges(n, us, vs, caps):
        us = np.ascontiguousarray(us, np.int32)
        vs = np.ascontiguousarray(vs, np.int32)
        caps = np.ascontiguousarray(caps, np.int64)
        h = _P()
        _check(lib().mf_hgraph_from_edges(n, len(us), _ptr(us), _ptr(vs), _ptr(caps), C.byref(h)))
        return HostGraph(h)

    @staticmethod
    def from_edgelist(e: Edges):
        h = _P()
        _check(lib().mf_hgraph_from_edgelist(e._h, C.byref(h)))
        return HostGraph(h)

    @staticmethod
    def open(path, source=None, R=-1):
        """Load a binary graph cache; MaxFlowError(MF_ESTALE) when `source`
        (the text file) or `R` no longer match it."""
        h = _P()
        _check(lib().mf_hgraph_open(str(path).encode(), None if source is None else str(source).encode(), int(R),
                                    C.byref(h)))
        return HostGraph(h)

    def save(self, path, source=None):
        _check(lib().mf_hgraph_save(self._h, str(path).encode(), None if source is None else str(source).encode()))
        return self

    def partition(self, R):
        _check(lib().mf_hgraph_partition(self._h, R))
        return self

    def info(self):
        n, E, R, lines = C.c_int(0), C.c_int64(0), C.c_int(0), C.c_int64(0)
        _check(lib().mf_hgraph_info(self._h, C.byref(n), C.byref(E), C.byref(R), C.byref(lines)))
        return n.value, E.value, R.value, lines.value

    def export(self):
        n, E, R, _ = self.info()
        d = {
            "row_ptr": np.empty(n + 1, np.int32), "col": np.empty(E, np.int32),
            "cap": np.empty(E, np.int64), "rev": np.empty(E, np.int32),
        }
        if R > 0:
            d.update(new_id=np.empty(n, np.int32), old_id=np.empty(n, np.int32),
                     rstart=np.empty(R + 1, np.int32), region_id=np.empty(n, np.int32))
        _check(lib().mf_hgraph_export(self._h, _ptr(d["row_ptr"]), _ptr(d["col"]), _ptr(d["cap"]),
                                      _ptr(d["rev"]), _ptr(d.get("new_id")), _ptr(d.get("old_id")),
                                      _ptr(d.get("rstart")), _ptr(d.get("region_id"))))
        return d

    def new_ids(self):
        """Original id -> region-ordered id (after partition)."""
        n, _E, R, _ = self.info()
        if R <= 0:
            return np.arange(n, dtype=np.int32)
        out = np.empty(n, np.int32)
        _check(lib().mf_hgraph_export(self._h, None, None, None, None, _ptr(out), None, None, None))
        return out

    def stats(self):
        n = self.info()[0]
        so, si = np.empty(n, np.int64), np.empty(n, np.int64)
        do, di = np.empty(n, np.int32), np.empty(n, np.int32)
        _check(lib().mf_hgraph_stats(self._h, _ptr(so), _ptr(si), _ptr(do), _ptr(di)))
        return so, si, do, di

    def components(self):
        n = self.info()[0]
        scc = np.empty(n, np.int32)
        wcc = np.empty(n, np.int32)
        _check(lib().mf_hgraph_components(self._h, _ptr(scc), _ptr(wcc)))
        return scc, wcc

    def core_decision(self, s, t):
        """The reference driver's pruning / boundary-core decision for (s, t)
        (region-ordered ids): pruned_n, run_n, core_kept, would_take."""
        a, b, c, w = C.c_int64(0), C.c_int64(0), C.c_int64(0), C.c_int(0)
        _check(lib().mf_hgraph_core_decision(self._h, int(s), int(t), C.byref(a), C.byref(b), C.byref(c), C.byref(w)))
        return {"pruned_n": a.value, "run_n": b.value, "core_kept": c.value, "would_take": bool(w.value)}

    def graph_class(self) -> str:
        """The class the autotuner keys its cache by."""
        buf = C.create_string_buffer(64)
        _check(lib().mf_graph_class(self._h, buf, 64))
        return buf.value.decode()

    def line_entries(self, us, vs):
        """Forward CSR entry (current ids) of every line the graph was built
        from, given those lines in their original ids."""
        us = np.ascontiguousarray(us, np.int32)
        vs = np.ascontiguousarray(vs, np.int32)
        out = np.empty(len(us), np.int32)
        _check(lib().mf_hgraph_line_entries(self._h, len(us), _ptr(us), _ptr(vs), _ptr(out)))
        return out

    def find_entry(self, u, v):
        e = C.c_int32(-1)
        _check(lib().mf_hgraph_find_entry(self._h, u, v, C.byref(e)))
        return e.value

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.mf_hgraph_free(self._h)
            self._h = None


class DeviceGraph:
    """A region-ordered graph resident in HBM."""

    def __init__(self, hg: HostGraph, tile_max: int = 4096):
        self._h = _P()
        _check(lib().mf_graph_upload(hg._h, tile_max, C.byref(self._h)))

    def info(self):
        n, E, T, b = C.c_int(0), C.c_int64(0), C.c_int(0), C.c_int64(0)
        _check(lib().mf_graph_info(self._h, C.byref(n), C.byref(E), C.byref(T), C.byref(b)))
        return {"n": n.value, "E": E.value, "ntiles": T.value, "bytes": b.value}

    def slab_bytes(self) -> int:
        """Device bytes of one query slab on this graph (engine arena unit)."""
        b = C.c_int64(0)
        _check(lib().mf_graph_slab_bytes(self._h, C.byref(b)))
        return b.value

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.mf_graph_free(self._h)
            self._h = None


class Engine:
    """Query slots + scheduler + HIP stream (plain solve = one slot)."""

    def __init__(self, nslots: int = 1, params: Params | None = None):
        self._h = _P()
        self._graphs = []
        p = params if params is not None else default_params()
        _check(lib().mf_engine_create(nslots, C.byref(p), C.byref(self._h)))
        self.nslots = nslots

    def add_graph(self, g: DeviceGraph) -> int:
        gid = C.c_int(0)
        _check(lib().mf_engine_add_graph(self._h, g._h, C.byref(gid)))
        self._graphs.append(g)
        return gid.value

    def solve(self, gid, s, t) -> Result:
        r = Result()
        rc = lib().mf_solve(self._h, gid, int(s), int(t), C.byref(r))
        _check(rc)
        return r

    def solve_noalloc(self, gid, s, t) -> Result:
        """mf_solve without device allocation (slab in the per-device static
        scratch array; full graph)."""
        r = Result()
        _check(lib().mf_solve_noalloc(self._h, gid, int(s), int(t), C.byref(r)))
        return r

    def submit(self, gid, s, t, tag, t_arrive=0.0):
        _check(lib().mf_engine_submit(self._h, gid, int(s), int(t), int(tag), float(t_arrive)))

    def run(self, max_steps=-1, cap=1 << 16):
        out = (Result * cap)()
        k = C.c_int(0)
        _check(lib().mf_engine_run(self._h, max_steps, out, cap, C.byref(k)))
        return [out[i] for i in range(k.value)]

    def graph_stats(self):
        """(graphs captured, fused steps replayed from a captured graph)."""
        a, b = C.c_int64(0), C.c_int64(0)
        _check(lib().mf_engine_graph_stats(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_budget(self, nbytes: int):
        """Slab arena size in bytes (0 = nslots full slabs of the largest graph)."""
        _check(lib().mf_engine_set_budget(self._h, int(nbytes)))

    def mem_stats(self):
        v = [C.c_int64(0) for _ in range(4)]
        _check(lib().mf_engine_mem_stats(self._h, *[C.byref(x) for x in v]))
        return {"budget": v[0].value, "in_use": v[1].value, "peak": v[2].value, "admit_waits": v[3].value}

    def serve(self, gids, s, t, arrive):
        """Serving loop: query i arrives arrive[i] s after the call (ascending);
        returns the results (tag = i) in completion order."""
        gids = np.ascontiguousarray(gids, np.int32)
        s = np.ascontiguousarray(s, np.int32)
        t = np.ascontiguousarray(t, np.int32)
        arrive = np.ascontiguousarray(arrive, np.float64)
        nq = len(s)
        out = (Result * max(nq, 1))()
        k = C.c_int(0)
        _check(lib().mf_engine_serve(self._h, nq, _ptr(gids), _ptr(s), _ptr(t), _ptr(arrive), out, C.byref(k)))
        return [out[i] for i in range(k.value)]

    def live(self):
        a, b = C.c_int(0), C.c_int64(0)
        _check(lib().mf_engine_live(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def now(self):
        return lib().mf_engine_now(self._h)

    def cache_stats(self):
        a, b = C.c_int64(0), C.c_int64(0)
        _check(lib().mf_engine_cache_stats(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def profile(self, on: bool):
        _check(lib().mf_engine_profile(self._h, 1 if on else 0))

    def set_step_log(self, fn):
        """fn(StepLog) after every engine step, per slot that ran something;
        None turns it off."""
        self._step_cb = None if fn is None else STEP_FN(lambda _u, p: fn(p.contents))
        _check(lib().mf_engine_set_step_log(self._h, self._step_cb if fn is not None else STEP_FN(), None))

    def profile_read(self):
        ms = np.zeros(7, np.float64)
        la = np.zeros(7, np.int64)
        wk = np.zeros(28, np.uint64)
        _check(lib().mf_engine_profile_read(self._h, _ptr(ms), _ptr(la), _ptr(wk)))
        return ms, la, wk.reshape(7, 4)

    def pin(self, slot, gid, s, t, tag):
        _check(lib().mf_engine_pin(self._h, slot, gid, int(s), int(t), int(tag)))

    def solve_pinned(self, warm: bool):
        out = (Result * self.nslots)()
        k = C.c_int(0)
        _check(lib().mf_engine_solve_pinned(self._h, 1 if warm else 0, out, self.nslots, C.byref(k)))
        return [out[i] for i in range(k.value)]

    def update_caps(self, gid, entries, new_caps):
        entries = np.ascontiguousarray(entries, np.int32)
        new_caps = np.ascontiguousarray(new_caps, np.int64)
        _check(lib().mf_graph_update_caps(self._h, gid, len(entries), _ptr(entries), _ptr(new_caps)))

    # parity hooks ---------------------------------------------------------
    def dbg_init_preflow(self, gid, s, t, n, E):
        res = np.empty(E, np.int64)
        ex = np.empty(n, np.int64)
        _check(lib().mf_dbg_init_preflow(self._h, gid, int(s), int(t), _ptr(res), _ptr(ex)))
        return res, ex

    def dbg_global_relabel(self, gid, s, t, res, excess):
        res = np.ascontiguousarray(res, np.int64)
        excess = np.ascontiguousarray(excess, np.int64)
        h = np.empty(len(excess), np.int32)
        _check(lib().mf_dbg_global_relabel(self._h, gid, int(s), int(t), _ptr(res), _ptr(excess), _ptr(h)))
        return h

    def dbg_local_relabel(self, gid, s, t, res, excess, h, tile_active, lcap):
        """Local relabel on the given state; returns (new heights, level)."""
        res = np.ascontiguousarray(res, np.int64)
        excess = np.ascontiguousarray(excess, np.int64)
        h = np.array(h, np.int32)
        ta = np.ascontiguousarray(tile_active, np.int32)
        lvl = C.c_int(0)
        _check(lib().mf_dbg_local_relabel(self._h, gid, int(s), int(t), _ptr(res), _ptr(excess), _ptr(h),
                                          _ptr(ta), int(lcap), C.byref(lvl)))
        return h, lvl.value

    def dbg_barrier_ticks(self, ticks):
        """Set the multi-workgroup BFS barrier's wait bound (ticks <= 0: the
        default); returns the BFS restarts of the queries retired so far."""
        r = C.c_longlong(0)
        _check(lib().mf_dbg_barrier_ticks(self._h, int(ticks), C.byref(r)))
        return r.value

    def dbg_hot_select(self, gid, tile_cnt, k):
        """The hot sweep's choice of k tiles on given per-tile active counts."""
        tile_cnt = np.ascontiguousarray(tile_cnt, np.int32)
        hot = np.empty(len(tile_cnt), np.int32)
        _check(lib().mf_dbg_hot_select(self._h, gid, _ptr(tile_cnt), int(k), _ptr(hot)))
        return hot

    def dbg_certify(self, gid, s, t, res, excess, h, mode):
        """GPU certificate of a given state: (bad, cut, excess_t, fingerprint_zero)."""
        res = np.ascontiguousarray(res, np.int64)
        excess = np.ascontiguousarray(excess, np.int64)
        h = np.ascontiguousarray(h, np.int32)
        bad, cut, ext, fz = C.c_int64(0), C.c_int64(0), C.c_int64(0), C.c_int(0)
        _check(lib().mf_dbg_certify(self._h, gid, int(s), int(t), _ptr(res), _ptr(excess), _ptr(h), int(mode),
                                    C.byref(bad), C.byref(cut), C.byref(ext), C.byref(fz)))
        return bad.value, cut.value, ext.value, bool(fz.value)

    def dbg_prune_mask(self, gid, s, t, n):
        mask = np.empty(n, np.int32)
        k = C.c_int(0)
        _check(lib().mf_dbg_prune_mask(self._h, gid, int(s), int(t), _ptr(mask), C.byref(k)))
        return mask, k.value

    def dbg_class_mask(self, gid, s, t, n):
        """The product's class-path pruning mask: (mask or None, kept, reachable)."""
        mask = np.zeros(n, np.int32)
        kept, reach = C.c_int64(0), C.c_int(0)
        _check(lib().mf_dbg_class_mask(self._h, gid, int(s), int(t), _ptr(mask), C.byref(kept), C.byref(reach)))
        return (mask if reach.value else None), kept.value, bool(reach.value)

    def dbg_induced(self, gid, mask, E_parent):
        mask = np.ascontiguousarray(mask, np.int32)
        n = len(mask)
        nk, Ek = C.c_int(0), C.c_int64(0)
        rp = np.empty(n + 1, np.int32)
        col = np.empty(max(E_parent, 1), np.int32)
        cap = np.empty(max(E_parent, 1), np.int64)
        rev = np.empty(max(E_parent, 1), np.int32)
        o2s = np.empty(n, np.int32)
        _check(lib().mf_dbg_induced(self._h, gid, _ptr(mask), C.byref(nk), C.byref(Ek), _ptr(rp), _ptr(col),
                                    _ptr(cap), _ptr(rev), _ptr(o2s)))
        k, m = nk.value, Ek.value
        return {"n": k, "E": m, "row_ptr": rp[: k + 1].copy(), "col": col[:m].copy(), "cap": cap[:m].copy(),
                "rev": rev[:m].copy(), "old2sub": o2s}

    def dbg_solve_state(self, gid, s, t, n, E):
        r = Result()
        res = np.empty(E, np.int64)
        ex = np.empty(n, np.int64)
        h = np.empty(n, np.int32)
        _check(lib().mf_dbg_solve_state(self._h, gid, int(s), int(t), C.byref(r), _ptr(res), _ptr(ex), _ptr(h)))
        return r, res, ex, h

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.mf_engine_free(self._h)
            self._h = None
