// mf_capi.cpp -- the extern "C" boundary declared in include/mscmf.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mscmf.h"
#include "host_graph.h"
#include "mf_engine.h"
#include "mf_handles.h"
#include "mf_kernels.h"
#include "mf_tune.h"

using namespace mf;

namespace mf {
void result_to_c(const Result& r, mf_result* o);
}

namespace {

thread_local std::string g_capi_err;

int fail(int code, const std::string& msg) {
    set_error(msg);
    return code;
}

void to_c(const Result& r, mf_result* o) { mf::result_to_c(r, o); }

}  // namespace

void mf::result_to_c(const Result& r, mf_result* o) {
    memset(o, 0, sizeof(*o));
    o->tag = r.tag;
    o->flow = r.st.flow;
    o->status = r.st.status;
    o->cert_ok = r.st.cert_ok;
    o->pruned = r.st.pruned;
    o->n_run = r.st.n_run;
    o->E_run = r.st.E_run;
    o->kept = r.st.kept;
    o->iters = r.st.iters;
    o->steps = r.st.steps;
    o->grs = r.st.grs;
    o->bfs_levels = r.st.bfs_levels;
    o->drains = r.st.drains;
    o->sink_cut = (r.st.sink_cuts > 0 && r.st.spec_fails == 0) ? 1 : 0;
    o->rm_fixed = r.st.rm_fixed;
    o->edge_scans = r.st.edge_scans;
    o->vertex_scans = r.st.vertex_scans;
    o->pushes = r.st.pushes;
    o->relabels = r.st.relabels;
    o->t_arrive = r.st.t_submit;
    o->t_start = r.st.t_start;
    o->t_end = r.st.t_end;
    for (int f = 0; f < kFamilies; ++f)
        for (int k = 0; k < 4; ++k) o->work[f][k] = r.st.wk[f][k];
}

namespace {

Params from_c(const mf_params* p) {
    Params P;
    if (!p) return P;
    P.K = p->K;
    P.gr_freq = p->gr_freq;
    P.gr_freq_min = p->gr_freq_min;
    P.gr_freq_max = p->gr_freq_max;
    P.gr_stall = p->gr_stall;
    P.h_cut = p->h_cut;
    P.stagnate_rounds = p->stagnate_rounds;
    P.path_switch = p->path_switch;
    P.path_len = p->path_len;
    P.path_attempts = p->path_attempts;
    P.p_count0 = p->p_count0;
    P.hot = p->hot;
    P.hot_k = p->hot_k;
    P.hot_min_active = p->hot_min_active;
    P.hot_rounds = p->hot_rounds;
    P.rm = p->rm;
    P.rm_freq0 = p->rm_freq0;
    P.rm_freq_max = p->rm_freq_max;
    P.rm_margin = p->rm_margin;
    P.rm_tau = p->rm_tau;
    P.tile_rounds = p->tile_rounds;
    P.tile_frac = p->tile_frac;
    P.cert = p->cert;
    P.iter_max = p->iter_max;
    P.bu_alpha = p->bu_alpha;
    P.bu_beta = p->bu_beta;
    P.prune = p->prune;
    P.prune_ratio = p->prune_ratio;
    P.gr_work = p->gr_work;
    P.K_heavy = p->K_heavy;
    P.local_relabel = p->local_relabel;
    P.local_band = p->local_band;
    P.fuse = p->fuse;
    P.graphs = p->graphs;
    P.spec_cert = p->spec_cert;
    P.sparse_count = p->sparse_count;
    P.small_wgs = p->small_wgs;
    P.small_edges = p->small_edges;
    P.mark_list = p->mark_list;
    P.inject = p->inject;
    P.tile_threads = p->tile_threads;
    P.path_switch_deep = p->path_switch_deep;
    P.p_count_deep = p->p_count_deep;
    P.tile_split = p->tile_split;
    P.scc_route = p->scc_route;
    P.bu_skip = p->bu_skip;
    P.sink_cut = p->sink_cut;
    P.disc_band = p->disc_band;
    P.bu_pick = p->bu_pick;
    P.small_levels = p->small_levels;
    P.push_first = p->push_first;
    P.slot_blocks = p->slot_blocks;
    P.launch_blocks = p->launch_blocks;
    P.bu_light = p->bu_light;
    P.bu_bits = p->bu_bits;
    P.gr_work_deep = p->gr_work_deep;
    P.small_edges_deep = p->small_edges_deep;
    P.giant_rounds = p->giant_rounds;
    P.bu_recv = p->bu_recv;
    P.dis_queue = p->dis_queue;
    return P;
}

#define HIPCK(expr)                                                             \
    do {                                                                        \
        hipError_t e_ = (expr);                                                 \
        if (e_ != hipSuccess) return fail(MF_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

}  // namespace

extern "C" {

const char* mf_last_error(void) {
    g_capi_err = last_error();
    return g_capi_err.c_str();
}

int mf_version(void) { return 1; }

void mf_struct_sizes(int64_t* params_bytes, int64_t* result_bytes) {
    if (params_bytes) *params_bytes = (int64_t)sizeof(mf_params);
    if (result_bytes) *result_bytes = (int64_t)sizeof(mf_result);
}

// ------------------------------------------------------------- host graph

int mf_hgraph_load(const char* graph_path, mf_hgraph** out) {
    if (!graph_path || !out) return fail(MF_EINVAL, "null argument");
    std::vector<int32_t> us, vs;
    std::vector<int64_t> caps;
    int n = 0;
    std::string err;
    const int rc = parse_graph_file(graph_path, us, vs, caps, n, err);
    if (rc != 0) return fail(MF_EIO, err);
    mf_hgraph* h = new (std::nothrow) mf_hgraph();
    if (!h) return fail(MF_ENOMEM, "out of host memory");
    build_csr(n, us, vs, caps, h->g);
    *out = h;
    return MF_OK;
}

int mf_hgraph_from_edges(int n, int64_t cnt, const int32_t* us, const int32_t* vs, const int64_t* caps,
                         mf_hgraph** out) {
    if (n < 0 || cnt < 0 || !out || (cnt > 0 && (!us || !vs || !caps))) return fail(MF_EINVAL, "bad argument");
    for (int64_t i = 0; i < cnt; ++i)
        if (us[i] < 0 || us[i] >= n || vs[i] < 0 || vs[i] >= n || us[i] == vs[i] || caps[i] <= 0)
            return fail(MF_EINVAL, "edge " + std::to_string(i) + " out of range, self-loop or cap <= 0");
    if (2 * cnt >= INT32_MAX) return fail(MF_EINVAL, "too many edges for 32-bit entry ids");
    mf_hgraph* h = new (std::nothrow) mf_hgraph();
    if (!h) return fail(MF_ENOMEM, "out of host memory");
    std::vector<int32_t> u(us, us + cnt), v(vs, vs + cnt);
    std::vector<int64_t> c(caps, caps + cnt);
    build_csr(n, u, v, c, h->g);
    *out = h;
    return MF_OK;
}

int mf_hgraph_from_edgelist(const mf_edges* e, mf_hgraph** out) {
    if (!e || !out) return fail(MF_EINVAL, "null argument");
    mf_hgraph* h = new (std::nothrow) mf_hgraph();
    if (!h) return fail(MF_ENOMEM, "out of host memory");
    build_csr(e->n, e->us, e->vs, e->caps, h->g);
    *out = h;
    return MF_OK;
}

int mf_hgraph_save(const mf_hgraph* h, const char* path, const char* source_path) {
    if (!h || !path) return fail(MF_EINVAL, "null argument");
    std::string err;
    if (save_graph_binary(h->g, path, source_path ? source_path : "", err) != 0) return fail(MF_EIO, err);
    return MF_OK;
}

int mf_hgraph_open(const char* path, const char* source_path, int R, mf_hgraph** out) {
    if (!path || !out) return fail(MF_EINVAL, "null argument");
    *out = nullptr;
    auto* h = new (std::nothrow) mf_hgraph;
    if (!h) return fail(MF_ENOMEM, "out of host memory");
    std::string err;
    const int rc = load_graph_binary(path, source_path ? source_path : "", R, h->g, err);
    if (rc != 0) {
        delete h;
        if (rc > 0) return fail(MF_ESTALE, std::string("graph cache is stale: ") + path);
        return fail(MF_EIO, err);
    }
    *out = h;
    return MF_OK;
}

int mf_hgraph_partition(mf_hgraph* h, int R) {
    if (!h || R < 1) return fail(MF_EINVAL, "bad argument");
    if (!h->g.rstart.empty()) return fail(MF_EINVAL, "graph already partitioned");
    std::vector<int32_t> part;
    capacity_degree_partition(h->g.n, R, h->g.sum_out, h->g.sum_in, h->g.deg_out, h->g.deg_in, part);
    partition_and_reorder(h->g, R, part);
    return MF_OK;
}

int mf_hgraph_info(const mf_hgraph* h, int* n, int64_t* E, int* R, int64_t* lines) {
    if (!h) return fail(MF_EINVAL, "null graph");
    if (n) *n = h->g.n;
    if (E) *E = h->g.E();
    if (R) *R = h->g.R;
    if (lines) *lines = h->g.lines;
    return MF_OK;
}

int mf_hgraph_export(const mf_hgraph* h, int32_t* row_ptr, int32_t* col, int64_t* cap, int32_t* rev,
                     int32_t* new_id, int32_t* old_id, int32_t* rstart, int32_t* region_id) {
    if (!h) return fail(MF_EINVAL, "null graph");
    const HostGraph& g = h->g;
    if (row_ptr) std::copy(g.row_ptr.begin(), g.row_ptr.end(), row_ptr);
    if (col) std::copy(g.col.begin(), g.col.end(), col);
    if (cap) std::copy(g.cap.begin(), g.cap.end(), cap);
    if (rev) std::copy(g.rev.begin(), g.rev.end(), rev);
    if ((new_id || old_id || rstart || region_id) && g.rstart.empty())
        return fail(MF_EINVAL, "graph is not partitioned");
    if (new_id) std::copy(g.new_id.begin(), g.new_id.end(), new_id);
    if (old_id) std::copy(g.old_id.begin(), g.old_id.end(), old_id);
    if (rstart) std::copy(g.rstart.begin(), g.rstart.end(), rstart);
    if (region_id) std::copy(g.region_id.begin(), g.region_id.end(), region_id);
    return MF_OK;
}

int mf_hgraph_stats(const mf_hgraph* h, int64_t* so, int64_t* si, int32_t* dout, int32_t* din) {
    if (!h) return fail(MF_EINVAL, "null graph");
    const HostGraph& g = h->g;
    if (so) std::copy(g.sum_out.begin(), g.sum_out.end(), so);
    if (si) std::copy(g.sum_in.begin(), g.sum_in.end(), si);
    if (dout) std::copy(g.deg_out.begin(), g.deg_out.end(), dout);
    if (din) std::copy(g.deg_in.begin(), g.deg_in.end(), din);
    return MF_OK;
}

void mf_hgraph_free(mf_hgraph* h) { delete h; }

int mf_hgraph_find_entry(const mf_hgraph* h, int u, int v, int32_t* entry) {
    if (!h || !entry) return fail(MF_EINVAL, "null argument");
    *entry = -1;
    if (u < 0 || u >= h->g.n || v < 0 || v >= h->g.n) return fail(MF_EINVAL, "vertex out of range");
    for (int e = h->g.row_ptr[u]; e < h->g.row_ptr[u + 1]; ++e)
        if (h->g.col[e] == v && h->g.cap[e] > 0) { *entry = e; break; }
    return MF_OK;
}

int mf_hgraph_line_entries(const mf_hgraph* h, int64_t cnt, const int32_t* us, const int32_t* vs, int32_t* entry) {
    if (!h || (cnt > 0 && (!us || !vs || !entry))) return fail(MF_EINVAL, "null argument");
    const HostGraph& g = h->g;
    if (cnt != g.lines) return fail(MF_EINVAL, "line count differs from the graph's accepted lines");
    const int n = g.n;
    const bool reord = (int)g.new_id.size() == n && n > 0;
    // row lengths of the original order are those of the current rows
    std::vector<int32_t> cur(n, 0);
    for (int64_t i = 0; i < cnt; ++i) {
        const int u = us[i], v = vs[i];
        if (u < 0 || u >= n || v < 0 || v >= n) return fail(MF_EINVAL, "line endpoint out of range");
        const int ru = reord ? g.new_id[u] : u, rv = reord ? g.new_id[v] : v;
        const int64_t e = (int64_t)g.row_ptr[ru] + cur[u]++;
        cur[v]++;
        if (e >= g.row_ptr[ru + 1] || g.col[e] != rv || g.cap[e] <= 0)
            return fail(MF_EINVAL, "line " + std::to_string(i) + " is not the graph's line (order or endpoints differ)");
        entry[i] = (int32_t)e;
    }
    return MF_OK;
}

int mf_hgraph_components(const mf_hgraph* h, int32_t* scc, int32_t* wcc) {
    if (!h) return fail(MF_EINVAL, "null graph");
    std::vector<int32_t> a;
    if (scc) { strong_components(h->g, a); std::copy(a.begin(), a.end(), scc); }
    if (wcc) { weak_components(h->g, a); std::copy(a.begin(), a.end(), wcc); }
    return MF_OK;
}

int mf_hgraph_core_decision(const mf_hgraph* h, int s, int t, int64_t* pruned_n, int64_t* run_n,
                            int64_t* core_kept, int* would_take) {
    if (!h) return fail(MF_EINVAL, "null graph");
    if (h->g.rstart.empty()) return fail(MF_EINVAL, "graph is not partitioned");
    if (s < 0 || t < 0 || s >= h->g.n || t >= h->g.n || s == t) return fail(MF_EINVAL, "bad query");
    const CoreDecision d = reference_core_decision(h->g, s, t);
    if (pruned_n) *pruned_n = d.pruned_n;
    if (run_n) *run_n = d.run_n;
    if (core_kept) *core_kept = d.core_kept;
    if (would_take) *would_take = d.would_take ? 1 : 0;
    return MF_OK;
}

int mf_partition(int n, int R, const int64_t* so, const int64_t* si, const int32_t* dout,
                 const int32_t* din, int32_t* part) {
    if (n < 0 || R < 1 || (n > 0 && (!so || !si || !dout || !din || !part))) return fail(MF_EINVAL, "bad argument");
    std::vector<int64_t> a(so, so + n), b(si, si + n);
    std::vector<int32_t> c(dout, dout + n), d(din, din + n), p;
    capacity_degree_partition(n, R, a, b, c, d, p);
    std::copy(p.begin(), p.end(), part);
    return MF_OK;
}

int mf_read_queries(const char* path, int64_t cap, int32_t* s, int32_t* t, int64_t* count) {
    if (!path || !count) return fail(MF_EINVAL, "null argument");
    std::vector<int32_t> a, b;
    std::string err;
    if (parse_query_file(path, a, b, err) != 0) return fail(MF_EIO, err);
    const int64_t q = std::min<int64_t>(cap, (int64_t)a.size());
    for (int64_t i = 0; i < q; ++i) { if (s) s[i] = a[i]; if (t) t[i] = b[i]; }
    *count = (cap <= 0) ? (int64_t)a.size() : q;
    return MF_OK;
}

// ------------------------------------------------------------- generators

int mf_gen_uniform(int n, int64_t m, int64_t cap_max, uint64_t seed, mf_edges** out) {
    if (n < 2 || m < 0 || cap_max < 1 || !out) return fail(MF_EINVAL, "bad argument");
    mf_edges* e = new mf_edges();
    e->n = n;
    gen_uniform(n, m, cap_max, seed, e->us, e->vs, e->caps);
    *out = e;
    return MF_OK;
}

int mf_gen_grid(int rows, int cols, int64_t cap_max, uint64_t seed, mf_edges** out) {
    if (rows < 1 || cols < 1 || cap_max < 1 || !out) return fail(MF_EINVAL, "bad argument");
    mf_edges* e = new mf_edges();
    e->n = rows * cols;
    gen_grid(rows, cols, cap_max, seed, e->us, e->vs, e->caps);
    *out = e;
    return MF_OK;
}

int mf_gen_rmat(int scale, int64_t samples, double a, double b, double c, int64_t cap_max, uint64_t seed,
                int threads, mf_edges** out) {
    if (scale < 1 || scale > 30 || samples < 0 || cap_max < 1 || !out) return fail(MF_EINVAL, "bad argument");
    mf_edges* e = new mf_edges();
    e->n = 1 << scale;
    gen_rmat(scale, samples, a, b, c, cap_max, seed, threads, e->us, e->vs, e->caps);
    *out = e;
    return MF_OK;
}

int mf_edges_info(const mf_edges* e, int* n, int64_t* cnt) {
    if (!e) return fail(MF_EINVAL, "null edges");
    if (n) *n = e->n;
    if (cnt) *cnt = (int64_t)e->us.size();
    return MF_OK;
}

int mf_edges_copy(const mf_edges* e, int32_t* us, int32_t* vs, int64_t* caps) {
    if (!e) return fail(MF_EINVAL, "null edges");
    if (us) std::copy(e->us.begin(), e->us.end(), us);
    if (vs) std::copy(e->vs.begin(), e->vs.end(), vs);
    if (caps) std::copy(e->caps.begin(), e->caps.end(), caps);
    return MF_OK;
}

int mf_edges_write(const mf_edges* e, const char* path) {
    if (!e || !path) return fail(MF_EINVAL, "null argument");
    if (write_graph_file(path, e->n, e->us, e->vs, e->caps) != 0) return fail(MF_EIO, std::string("cannot write ") + path);
    return MF_OK;
}

void mf_edges_free(mf_edges* e) { delete e; }

int mf_gen_queries_uniform(int n, int q, uint64_t seed, int32_t* s, int32_t* t) {
    if (n < 2 || q < 0 || (q > 0 && (!s || !t))) return fail(MF_EINVAL, "bad argument");
    std::vector<int32_t> a, b;
    gen_queries_uniform(n, q, seed, a, b);
    std::copy(a.begin(), a.end(), s);
    std::copy(b.begin(), b.end(), t);
    return MF_OK;
}

int mf_gen_queries_grid(int rows, int cols, int q, uint64_t seed, int32_t* s, int32_t* t) {
    if (rows < 1 || cols < 2 || q < 0 || (q > 0 && (!s || !t))) return fail(MF_EINVAL, "bad argument");
    std::vector<int32_t> a, b;
    gen_queries_grid(rows, cols, q, seed, a, b);
    std::copy(a.begin(), a.end(), s);
    std::copy(b.begin(), b.end(), t);
    return MF_OK;
}

int mf_gen_queries_hub(const mf_edges* e, double top_frac, int q, uint64_t seed, int32_t* s, int32_t* t) {
    if (!e || q < 0 || (q > 0 && (!s || !t))) return fail(MF_EINVAL, "bad argument");
    std::vector<int32_t> a, b;
    gen_queries_hub(e->n, e->us, e->vs, top_frac, q, seed, a, b);
    if ((int)a.size() != q) return fail(MF_EINVAL, "graph has no vertex with in-degree > 0");
    std::copy(a.begin(), a.end(), s);
    std::copy(b.begin(), b.end(), t);
    return MF_OK;
}

int mf_write_queries(const char* path, int q, const int32_t* s, const int32_t* t) {
    if (!path || q < 0) return fail(MF_EINVAL, "bad argument");
    std::vector<int32_t> a(s, s + q), b(t, t + q);
    if (write_query_file(path, a, b) != 0) return fail(MF_EIO, std::string("cannot write ") + path);
    return MF_OK;
}

// ----------------------------------------------------------- device graph

int mf_device_count(int* count) {
    if (!count) return fail(MF_EINVAL, "null argument");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return MF_OK;
}

int mf_trace_marker(void) {
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || nd < 1) return fail(MF_ENODEV, "no HIP device");
    launch_trace_marker(nullptr);
    if (hipDeviceSynchronize() != hipSuccess) return fail(MF_EHIP, "trace marker failed");
    return MF_OK;
}

int mf_graph_upload(const mf_hgraph* h, int tile_max, mf_graph** out) {
    if (!h || !out) return fail(MF_EINVAL, "null argument");
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || nd < 1) return fail(MF_ENODEV, "no HIP device");
    auto dg = std::make_shared<DeviceGraph>();
    const int rc = upload_graph(h->g, tile_max <= 0 ? kTileMax : tile_max, *dg);
    if (rc != MF_OK) return rc;
    mf_graph* g = new mf_graph();
    g->dg = dg;
    *out = g;
    return MF_OK;
}

int mf_graph_info(const mf_graph* g, int* n, int64_t* E, int* ntiles, int64_t* bytes) {
    if (!g) return fail(MF_EINVAL, "null graph");
    if (n) *n = g->dg->g.n;
    if (E) *E = g->dg->g.E;
    if (ntiles) *ntiles = g->dg->g.ntiles;
    if (bytes) *bytes = (int64_t)g->dg->bytes;
    return MF_OK;
}

int mf_graph_slab_bytes(const mf_graph* g, int64_t* bytes) {
    if (!g || !bytes) return fail(MF_EINVAL, "null argument");
    *bytes = (int64_t)Slab::bytes_for(g->dg->g.n, g->dg->g.E, g->dg->g.ntiles);
    return MF_OK;
}

void mf_graph_free(mf_graph* g) { delete g; }

// ------------------------------------------------------------------ engine

void mf_params_default(mf_params* p) {
    if (!p) return;
    Params P;
    memset(p, 0, sizeof(*p));
    p->K = P.K;
    p->gr_freq = P.gr_freq;
    p->gr_freq_min = P.gr_freq_min;
    p->gr_freq_max = P.gr_freq_max;
    p->gr_stall = P.gr_stall;
    p->h_cut = P.h_cut;
    p->stagnate_rounds = P.stagnate_rounds;
    p->path_switch = P.path_switch;
    p->path_len = P.path_len;
    p->path_attempts = P.path_attempts;
    p->p_count0 = P.p_count0;
    p->hot = P.hot;
    p->hot_k = P.hot_k;
    p->hot_min_active = P.hot_min_active;
    p->hot_rounds = P.hot_rounds;
    p->rm = P.rm;
    p->rm_freq0 = P.rm_freq0;
    p->rm_freq_max = P.rm_freq_max;
    p->rm_margin = P.rm_margin;
    p->rm_tau = P.rm_tau;
    p->tile_rounds = P.tile_rounds;
    p->tile_frac = P.tile_frac;
    p->cert = P.cert;
    p->iter_max = P.iter_max;
    p->bu_alpha = P.bu_alpha;
    p->bu_beta = P.bu_beta;
    p->prune = P.prune;
    p->prune_ratio = P.prune_ratio;
    p->gr_work = P.gr_work;
    p->K_heavy = P.K_heavy;
    p->local_relabel = P.local_relabel;
    p->local_band = P.local_band;
    p->fuse = P.fuse;
    p->graphs = P.graphs;
    p->spec_cert = P.spec_cert;
    p->sparse_count = P.sparse_count;
    p->small_wgs = P.small_wgs;
    p->small_edges = P.small_edges;
    p->mark_list = P.mark_list;
    p->inject = P.inject;
    p->tile_threads = P.tile_threads;
    p->path_switch_deep = P.path_switch_deep;
    p->p_count_deep = P.p_count_deep;
    p->tile_split = P.tile_split;
    p->scc_route = P.scc_route;
    p->bu_skip = P.bu_skip;
    p->sink_cut = P.sink_cut;
    p->disc_band = P.disc_band;
    p->bu_pick = P.bu_pick;
    p->small_levels = P.small_levels;
    p->push_first = P.push_first;
    p->slot_blocks = P.slot_blocks;
    p->launch_blocks = P.launch_blocks;
    p->bu_light = P.bu_light;
    p->bu_bits = P.bu_bits;
    p->gr_work_deep = P.gr_work_deep;
    p->small_edges_deep = P.small_edges_deep;
    p->giant_rounds = P.giant_rounds;
    p->bu_recv = P.bu_recv;
    p->dis_queue = P.dis_queue;
}

int mf_engine_create(int nslots, const mf_params* p, mf_engine** out) {
    if (!out || nslots < 1) return fail(MF_EINVAL, "bad argument");
    mf_engine* e = new (std::nothrow) mf_engine();
    if (!e) return fail(MF_ENOMEM, "out of host memory");
    const int rc = e->eng.init(nslots, from_c(p));
    if (rc != MF_OK) { delete e; return rc; }
    *out = e;
    return MF_OK;
}

int mf_engine_add_graph(mf_engine* e, mf_graph* g, int* gid) {
    if (!e || !g) return fail(MF_EINVAL, "null argument");
    const int id = e->eng.add_graph(g->dg);
    if (id < 0) return id;
    if (gid) *gid = id;
    return MF_OK;
}

void mf_engine_free(mf_engine* e) { delete e; }

static int check_query(mf_engine* e, int gid, int s, int t) {
    if (!e) return fail(MF_EINVAL, "null engine");
    if (gid < 0 || gid >= e->eng.ngraphs()) return fail(MF_EINVAL, "bad graph index");
    const int n = e->eng.graph(gid)->g.n;
    if (s < 0 || t < 0 || s >= n || t >= n) return fail(MF_EINVAL, "vertex out of range");
    if (s == t) return fail(MF_EINVAL, "s == t");
    return MF_OK;
}

int mf_solve(mf_engine* e, int gid, int s, int t, mf_result* r) {
    int rc = check_query(e, gid, s, t);
    if (rc != MF_OK) return rc;
    if (e->eng.live() || e->eng.queued()) return fail(MF_EBUSY, "engine has queued work");
    Query q;
    q.graph = gid; q.s = s; q.t = t; q.tag = -1;
    rc = e->eng.submit(q);
    if (rc != MF_OK) return rc;
    std::vector<Result> out;
    rc = e->eng.run(out);
    if (rc != MF_OK) return rc;
    if (out.size() != 1) return fail(MF_EHIP, "solve produced no result");
    if (r) to_c(out[0], r);
    if (out[0].st.status != MF_OK) return out[0].st.status;
    return MF_OK;
}

int mf_engine_submit(mf_engine* e, int gid, int s, int t, int64_t tag, double t_arrive) {
    int rc = check_query(e, gid, s, t);
    if (rc != MF_OK) return rc;
    Query q;
    q.graph = gid; q.s = s; q.t = t; q.tag = tag; q.t_arrive = t_arrive;
    return e->eng.submit(q);
}

int mf_engine_run(mf_engine* e, int64_t max_steps, mf_result* out, int cap, int* nout) {
    if (!e) return fail(MF_EINVAL, "null engine");
    const int rc = e->eng.run(e->done, max_steps);
    int k = 0;
    while (k < cap && !e->done.empty()) {
        to_c(e->done.front(), &out[k]);
        e->done.erase(e->done.begin());
        ++k;
    }
    if (nout) *nout = k;
    return rc;
}

int mf_engine_set_budget(mf_engine* e, int64_t bytes) {
    if (!e) return fail(MF_EINVAL, "null engine");
    if (bytes < 0) return fail(MF_EINVAL, "negative budget");
    return e->eng.set_budget((size_t)bytes);
}

int mf_engine_mem_stats(const mf_engine* e, int64_t* budget, int64_t* in_use, int64_t* peak,
                        int64_t* admit_waits) {
    if (!e) return fail(MF_EINVAL, "null engine");
    size_t b = 0, u = 0, p = 0;
    int64_t w = 0;
    e->eng.mem_stats(&b, &u, &p, &w);
    if (budget) *budget = (int64_t)b;
    if (in_use) *in_use = (int64_t)u;
    if (peak) *peak = (int64_t)p;
    if (admit_waits) *admit_waits = w;
    return MF_OK;
}

int mf_engine_serve(mf_engine* e, int nq, const int32_t* gid, const int32_t* s, const int32_t* t,
                    const double* arrive, mf_result* out, int* nout) {
    if (!e) return fail(MF_EINVAL, "null engine");
    if (nq < 0 || (nq > 0 && (!gid || !s || !t || !arrive || !out))) return fail(MF_EINVAL, "bad arguments");
    std::vector<Query> qs(nq);
    for (int i = 0; i < nq; ++i) {
        if (gid[i] < 0 || gid[i] >= e->eng.ngraphs()) return fail(MF_EINVAL, "bad graph index");
        const int n = e->eng.graph(gid[i])->g.n;
        if (s[i] < 0 || t[i] < 0 || s[i] >= n || t[i] >= n || s[i] == t[i]) return fail(MF_EINVAL, "bad query");
        qs[i].graph = gid[i];
        qs[i].s = s[i];
        qs[i].t = t[i];
        qs[i].tag = i;
        qs[i].t_arrive = arrive[i];
    }
    std::vector<Result> got;
    const int rc = e->eng.serve(std::move(qs), got);
    int k = 0;
    for (auto& r : got) {
        if (k >= nq) break;
        to_c(r, &out[k++]);
    }
    if (nout) *nout = k;
    return rc;
}

int mf_serve_lanes(int nengines, mf_engine* const* engines, int nq, const int32_t* lane, const int32_t* gid,
                   const int32_t* s, const int32_t* t, const double* arrive, mf_result* out, int* nout) {
    if (nengines < 1 || !engines) return fail(MF_EINVAL, "no engines");
    for (int l = 0; l < nengines; ++l) {
        if (!engines[l]) return fail(MF_EINVAL, "null engine");
        for (int k = 0; k < l; ++k)
            if (engines[k] == engines[l]) return fail(MF_EINVAL, "an engine serves one lane");
    }
    if (nq < 0 || (nq > 0 && (!lane || !gid || !s || !t || !arrive || !out))) return fail(MF_EINVAL, "bad arguments");
    std::vector<std::vector<Query>> per(nengines);
    for (int i = 0; i < nq; ++i) {
        if (lane[i] < 0 || lane[i] >= nengines) return fail(MF_EINVAL, "bad lane");
        Engine& eng = engines[lane[i]]->eng;
        if (gid[i] < 0 || gid[i] >= eng.ngraphs()) return fail(MF_EINVAL, "bad graph index");
        const int n = eng.graph(gid[i])->g.n;
        if (s[i] < 0 || t[i] < 0 || s[i] >= n || t[i] >= n || s[i] == t[i]) return fail(MF_EINVAL, "bad query");
        if (i > 0 && arrive[i] < arrive[i - 1]) return fail(MF_EINVAL, "arrivals must be ascending");
        Query q;
        q.graph = gid[i];
        q.s = s[i];
        q.t = t[i];
        q.tag = i;
        q.t_arrive = arrive[i];
        per[lane[i]].push_back(q);
    }
    // one host thread and one HIP stream per lane: the lanes' kernels overlap
    // on the GPU and a lane of shallow queries never waits for a deep one
    std::vector<std::vector<Result>> got(nengines);
    std::vector<int> rc(nengines, MF_OK);
    std::vector<std::string> err(nengines);
    const double T0 = engines[0]->eng.now() + 1e-3;
    std::vector<std::thread> th;
    for (int l = 0; l < nengines; ++l) {
        if (per[l].empty()) continue;
        th.emplace_back([&, l]() {
            rc[l] = engines[l]->eng.serve(per[l], got[l], T0);
            if (rc[l] != MF_OK) err[l] = last_error();
        });
    }
    for (auto& x : th) x.join();
    std::vector<const Result*> all;
    for (auto& g : got)
        for (auto& r : g) all.push_back(&r);
    std::sort(all.begin(), all.end(), [](const Result* a, const Result* b) { return a->st.t_end < b->st.t_end; });
    int k = 0;
    for (auto* r : all) {
        if (k >= nq) break;
        to_c(*r, &out[k++]);
    }
    if (nout) *nout = k;
    for (int l = 0; l < nengines; ++l)
        if (rc[l] != MF_OK) return fail(rc[l], err[l]);
    return MF_OK;
}

int mf_engine_live(const mf_engine* e, int* live, int64_t* queued) {
    if (!e) return fail(MF_EINVAL, "null engine");
    if (live) *live = e->eng.live();
    if (queued) *queued = (int64_t)e->eng.queued() + (int64_t)e->done.size() * 0;
    return MF_OK;
}

double mf_engine_now(const mf_engine* e) { return e ? e->eng.now() : 0.0; }

int mf_engine_cache_stats(const mf_engine* e, int64_t* hits, int64_t* misses) {
    if (!e) return fail(MF_EINVAL, "null engine");
    if (hits) *hits = e->eng.cache_hits;
    if (misses) *misses = e->eng.cache_misses;
    return MF_OK;
}

int mf_autotune(const mf_hgraph* h, int nq, const int32_t* s, const int32_t* t, const mf_params* base, int R0,
                int tile0, int batch, const char* cache_path, mf_tune_result* out) {
    if (!h || nq < 1 || !s || !t || !out) return fail(MF_EINVAL, "bad argument");
    if (R0 < 1 || tile0 < 1 || tile0 > kTileMax) return fail(MF_EINVAL, "bad R0 or tile0");
    std::vector<int32_t> vs(s, s + nq), vt(t, t + nq);
    TuneResult r;
    const int rc = autotune(h->g, vs, vt, from_c(base), R0, tile0, batch, cache_path ? cache_path : "", r);
    if (rc != MF_OK) return rc;
    out->R = r.R;
    out->tile_max = r.tile_max;
    out->K = r.K;
    out->gr_freq = r.gr_freq;
    out->tile_threads = r.tile_threads;
    out->tile_split = r.tile_split;
    out->slots = r.slots;
    out->ms_per_query = r.ms;
    out->default_ms_per_query = r.ms_default;
    out->evaluations = r.evaluations;
    std::snprintf(out->graph_class, sizeof(out->graph_class), "%s", r.cls.c_str());
    return MF_OK;
}

int mf_scc_build(int n, int64_t m, const int32_t* us, const int32_t* vs, const int64_t* caps, mf_scc** out) {
    if (n < 0 || m < 0 || !out || (m > 0 && (!us || !vs || !caps))) return fail(MF_EINVAL, "bad argument");
    for (int64_t i = 0; i < m; ++i)
        if (us[i] < 0 || us[i] >= n || vs[i] < 0 || vs[i] >= n || us[i] == vs[i] || caps[i] <= 0)
            return fail(MF_EINVAL, "line " + std::to_string(i) + " out of range, self-loop or cap <= 0");
    mf_scc* h = new (std::nothrow) mf_scc();
    if (!h) return fail(MF_ENOMEM, "out of host memory");
    scc_contract(n, m, us, vs, caps, h->c);
    *out = h;
    return MF_OK;
}

int mf_scc_load(const char* graph_path, mf_scc** out) {
    if (!graph_path || !out) return fail(MF_EINVAL, "null argument");
    std::vector<int32_t> us, vs;
    std::vector<int64_t> caps;
    int n = 0;
    std::string err;
    if (parse_graph_file(graph_path, us, vs, caps, n, err) != 0) return fail(MF_EIO, err);
    return mf_scc_build(n, (int64_t)us.size(), us.data(), vs.data(), caps.data(), out);
}

int mf_scc_info(const mf_scc* h, int* n, int32_t* ncomp, int64_t* m_contracted) {
    if (!h) return fail(MF_EINVAL, "null handle");
    if (n) *n = h->c.n;
    if (ncomp) *ncomp = h->c.ncomp;
    if (m_contracted) *m_contracted = (int64_t)h->c.cu.size();
    return MF_OK;
}

int mf_scc_comp(const mf_scc* h, int32_t* comp) {
    if (!h || (!comp && h->c.n > 0)) return fail(MF_EINVAL, "bad argument");
    std::copy(h->c.comp.begin(), h->c.comp.end(), comp);
    return MF_OK;
}

int mf_scc_edges(const mf_scc* h, int32_t* cu, int32_t* cv, int64_t* cap) {
    if (!h) return fail(MF_EINVAL, "null handle");
    if (cu) std::copy(h->c.cu.begin(), h->c.cu.end(), cu);
    if (cv) std::copy(h->c.cv.begin(), h->c.cv.end(), cv);
    if (cap) std::copy(h->c.ccap.begin(), h->c.ccap.end(), cap);
    return MF_OK;
}

int mf_scc_route(const mf_scc* h, int nq, const int32_t* s, const int32_t* t, int32_t* route) {
    if (!h || nq < 0 || (nq > 0 && (!s || !t || !route))) return fail(MF_EINVAL, "bad argument");
    scc_route(h->c, nq, s, t, route);
    return MF_OK;
}

void mf_scc_free(mf_scc* h) { delete h; }

int mf_graph_class(const mf_hgraph* h, char* out, int cap) {
    if (!h || !out || cap < 1) return fail(MF_EINVAL, "bad argument");
    std::snprintf(out, (size_t)cap, "%s", graph_class(h->g).c_str());
    return MF_OK;
}

int mf_engine_graph_stats(const mf_engine* e, int64_t* captures, int64_t* replays) {
    if (!e) return fail(MF_EINVAL, "null engine");
    if (captures) *captures = e->eng.graph_captures;
    if (replays) *replays = e->eng.graph_replays;
    return MF_OK;
}

int mf_engine_profile(mf_engine* e, int on) {
    if (!e) return fail(MF_EINVAL, "null engine");
    if (on) e->eng.prof_reset();
    e->eng.profile = on != 0;
    return MF_OK;
}

int mf_engine_set_step_log(mf_engine* e, mf_step_fn fn, void* user) {
    if (!e) return fail(MF_EINVAL, "null engine");
    e->eng.step_fn = fn;
    e->eng.step_user = user;
    return MF_OK;
}

int mf_engine_profile_read(const mf_engine* e, double* ms, int64_t* launches, uint64_t* work) {
    if (!e) return fail(MF_EINVAL, "null engine");
    for (int f = 0; f < kFamilies; ++f) {
        if (ms) ms[f] = e->eng.fam_ms[f];
        if (launches) launches[f] = e->eng.fam_launches[f];
        for (int k = 0; k < 4; ++k) if (work) work[f * 4 + k] = e->eng.fam_work[f][k];
    }
    return MF_OK;
}

// ------------------------------------------------------------------ dynamic

int mf_engine_pin(mf_engine* e, int slot, int gid, int s, int t, int64_t tag) {
    int rc = check_query(e, gid, s, t);
    if (rc != MF_OK) return rc;
    Query q;
    q.graph = gid; q.s = s; q.t = t; q.tag = tag;
    return e->eng.pin(slot, q);
}

int mf_engine_solve_pinned(mf_engine* e, int warm, mf_result* out, int cap, int* nout) {
    if (!e) return fail(MF_EINVAL, "null engine");
    std::vector<Result> res;
    const int rc = e->eng.solve_pinned(warm != 0, res);
    int k = 0;
    for (auto& r : res) if (k < cap) to_c(r, &out[k++]);
    if (nout) *nout = k;
    if (rc != MF_OK) return rc;
    for (auto& r : res) if (r.st.status != MF_OK) return r.st.status;
    return MF_OK;
}

int mf_graph_update_caps(mf_engine* e, int gid, int n, const int32_t* entries, const int64_t* new_cap) {
    if (!e || n < 0 || (n > 0 && (!entries || !new_cap))) return fail(MF_EINVAL, "bad argument");
    std::vector<int32_t> seen(entries, entries + n);
    std::sort(seen.begin(), seen.end());
    if (std::adjacent_find(seen.begin(), seen.end()) != seen.end()) return fail(MF_EINVAL, "repeated entry in update batch");
    return e->eng.update_caps(gid, entries, new_cap, n);
}

// ------------------------------------------------------------- parity hooks

int mf_dbg_init_preflow(mf_engine* e, int gid, int s, int t, int64_t* res, int64_t* excess) {
    int rc = check_query(e, gid, s, t);
    if (rc != MF_OK) return rc;
    rc = e->eng.dbg_begin(gid, s, t);
    if (rc != MF_OK) return rc;
    Plan p;
    p.stages = ST_RESET_PRE | ST_INIT | ST_PREFLOW;
    p.reset_pre = RST_BFS | RST_ACTIVE | RST_CERT | RST_WORK | RST_ERR;
    rc = e->eng.dbg_plan(p);
    const DGraph& g = e->eng.view(0).g;
    if (rc == MF_OK && res) HIPCK(hipMemcpy(res, e->eng.view(0).res, sizeof(int64_t) * g.E, hipMemcpyDeviceToHost));
    if (rc == MF_OK && excess) HIPCK(hipMemcpy(excess, e->eng.view(0).excess, sizeof(int64_t) * g.n, hipMemcpyDeviceToHost));
    e->eng.dbg_end();
    return rc;
}

// Drive the BFS state machine of a fresh scheduler over hand-built plans.
static int dbg_bfs(mf_engine* e, Plan first, int mode) {
    int rc = e->eng.dbg_plan(first);
    if (rc != MF_OK) return rc;
    SlotSched sched;
    const Params& P = e->eng.params();
    (void)P;
    // Level loop with the same direction heuristic the scheduler uses.
    int64_t medges = 0;
    bool bu = false;
    const DGraph& g = e->eng.view(0).g;
    for (int guard = 0; guard < 4 * g.n + 8; ++guard) {
        const Ctr& c = e->eng.ctr(0);
        const int L = c.level;
        const int f = c.fsize[L & 3];
        if (f == 0) return MF_OK;
        medges += c.fedges[L & 3];
        Plan p;
        p.stages = ST_BFS;
        p.bfs_mode = mode;
        const int64_t mu = std::max<int64_t>(0, (int64_t)g.E - medges);
        if (!bu && (double)c.fedges[L & 3] > (double)mu / P.bu_alpha) {
            bu = true;
            p.bu_bits = true;
        }
        else if (bu && (double)f < (double)g.n / P.bu_beta) bu = false;
        if (bu) { p.bfs_kind = BK_BU; p.bfs_grid = std::max(1, std::min(8192, (g.n + 255) / 256)); }
        else if (f <= kSmallFrontier && c.fedges[L & 3] <= P.small_edges) p.bfs_kind = BK_SMALL;
        else { p.bfs_kind = BK_TD; p.bfs_grid = (f + 255) / 256; }
        rc = e->eng.dbg_plan(p);
        if (rc != MF_OK) return rc;
    }
    return fail(MF_EITER, "BFS did not finish");
}

int mf_dbg_global_relabel(mf_engine* e, int gid, int s, int t, const int64_t* res, const int64_t* excess,
                          int32_t* height) {
    int rc = check_query(e, gid, s, t);
    if (rc != MF_OK) return rc;
    if (!res || !excess || !height) return fail(MF_EINVAL, "null argument");
    rc = e->eng.dbg_begin(gid, s, t);
    if (rc != MF_OK) return rc;
    SlotView& v = e->eng.view(0);
    const DGraph& g = v.g;
    HIPCK(hipMemcpy(v.res, res, sizeof(int64_t) * g.E, hipMemcpyHostToDevice));
    e->eng.slab(0)->res_uid = 0;  // res no longer follows the slab's marks
    e->eng.slab(0)->marks_ok = false;
    HIPCK(hipMemcpy(v.excess, excess, sizeof(int64_t) * g.n, hipMemcpyHostToDevice));
    // rows of the given residual that still hold their capacities take the
    // relabel's fresh-row tests, as in a solve
    if (e->eng.params().bu_recv && g.fwdm) {
        launch_recv_from_res(nullptr, g, v.res, v.recv);
        HIPCK(hipGetLastError());
        HIPCK(hipDeviceSynchronize());
        v.recv_ok = 1;
    }
    Plan p;
    p.stages = ST_RESET_PRE | ST_GR_SEED | ST_BFS;
    p.reset_pre = RST_BFS | RST_WORK | RST_ERR;
    p.bfs_kind = BK_SMALL;
    p.bfs_mode = BFS_GR;
    rc = dbg_bfs(e, p, BFS_GR);
    if (rc == MF_OK) HIPCK(hipMemcpy(height, v.h, sizeof(int32_t) * g.n, hipMemcpyDeviceToHost));
    e->eng.dbg_end();
    return rc;
}

int mf_dbg_local_relabel(mf_engine* e, int gid, int s, int t, const int64_t* res, const int64_t* excess,
                         int32_t* height, const int32_t* tile_active, int lcap, int* level) {
    int rc = check_query(e, gid, s, t);
    if (rc != MF_OK) return rc;
    if (!res || !excess || !height || !tile_active || !level) return fail(MF_EINVAL, "null argument");
    if (lcap < 1 || lcap > kLocalLevels) return fail(MF_EINVAL, "lcap out of range");
    rc = e->eng.dbg_begin(gid, s, t);
    if (rc != MF_OK) return rc;
    SlotView& v = e->eng.view(0);
    const DGraph& g = v.g;
    HIPCK(hipMemcpy(v.res, res, sizeof(int64_t) * g.E, hipMemcpyHostToDevice));
    e->eng.slab(0)->res_uid = 0;  // res no longer follows the slab's marks
    e->eng.slab(0)->marks_ok = false;
    HIPCK(hipMemcpy(v.excess, excess, sizeof(int64_t) * g.n, hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(v.h, height, sizeof(int32_t) * g.n, hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(v.tile_cnt, tile_active, sizeof(int32_t) * g.ntiles, hipMemcpyHostToDevice));
    Plan p;
    p.stages = ST_RESET_PRE | ST_LOCAL;
    p.reset_pre = RST_BFS | RST_WORK | RST_ERR;
    p.local_cap = lcap;
    rc = e->eng.dbg_plan(p);
    if (rc == MF_OK) {
        HIPCK(hipMemcpy(height, v.h, sizeof(int32_t) * g.n, hipMemcpyDeviceToHost));
        *level = e->eng.ctr(0).level;
    }
    e->eng.dbg_end();
    return rc;
}

int mf_dbg_certify(mf_engine* e, int gid, int s, int t, const int64_t* res, const int64_t* excess,
                   const int32_t* height, int mode, int64_t* bad, int64_t* cut, int64_t* excess_t,
                   int* fp_zero) {
    int rc = check_query(e, gid, s, t);
    if (rc != MF_OK) return rc;
    if (!res || !excess || !height || !bad || !cut || !excess_t || !fp_zero) return fail(MF_EINVAL, "null argument");
    if (mode < 1 || mode > 4) return fail(MF_EINVAL, "mode must be 1, 2, 3 or 4");
    // mode 4: mode 3 finding the cut entries in the rows of V \ T
    const bool sside = mode == 4;
    if (sside) mode = 3;
    rc = e->eng.dbg_begin(gid, s, t);
    if (rc != MF_OK) return rc;
    SlotView& v = e->eng.view(0);
    const DGraph& g = v.g;
    HIPCK(hipMemcpy(v.res, res, sizeof(int64_t) * g.E, hipMemcpyHostToDevice));
    e->eng.slab(0)->res_uid = 0;  // res no longer follows the slab's marks
    e->eng.slab(0)->marks_ok = false;
    if (mode == 3) {
        // sparse mode: mark every entry whose residual left its capacity
        std::vector<int64_t> cap((size_t)g.E);
        HIPCK(hipMemcpy(cap.data(), g.cap, sizeof(int64_t) * g.E, hipMemcpyDeviceToHost));
        std::vector<uint8_t> mk((size_t)g.E);
        for (int64_t i = 0; i < g.E; ++i) mk[i] = res[i] != cap[i] ? 1 : 0;
        HIPCK(hipMemcpy(v.emark, mk.data(), (size_t)g.E, hipMemcpyHostToDevice));
        const int32_t over = v.mcap + 1;  // no list: the certificate reads the map
        HIPCK(hipMemcpy(v.mlist, &over, sizeof(int32_t), hipMemcpyHostToDevice));
        e->eng.slab(0)->marks_ok = true;
    }
    HIPCK(hipMemcpy(v.excess, excess, sizeof(int64_t) * g.n, hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(v.h, height, sizeof(int32_t) * g.n, hipMemcpyHostToDevice));
    const int saved = e->eng.params().cert;
    e->eng.set_cert_mode(mode);
    v.cert_sside = sside ? 1 : 0;
    Plan p;
    p.stages = ST_RESET_PRE | ST_CERT;
    p.reset_pre = RST_CERT | RST_WORK | RST_ERR;
    rc = e->eng.dbg_plan(p);
    e->eng.set_cert_mode(saved);
    if (rc == MF_OK) {
        const Ctr& c = e->eng.ctr(0);
        *bad = c.cert_bad;
        *cut = c.cut;
        *excess_t = c.ex_t;
        *fp_zero = c.fp == 0 ? 1 : 0;
    }
    e->eng.dbg_end();
    return rc;
}

int mf_dbg_barrier_ticks(mf_engine* e, long long ticks, long long* restarts) {
    if (!e) return fail(MF_EINVAL, "null engine");
    e->eng.set_barrier_ticks(ticks);
    if (restarts) *restarts = e->eng.bfs_restarts;
    return MF_OK;
}

int mf_dbg_hot_select(mf_engine* e, int gid, const int32_t* tile_cnt, int k, int32_t* tile_hot) {
    if (!e || !tile_cnt || !tile_hot || k < 1) return fail(MF_EINVAL, "bad argument");
    int rc = check_query(e, gid, 0, 1);
    if (rc != MF_OK) return rc;
    rc = e->eng.dbg_begin(gid, 0, 1);
    if (rc != MF_OK) return rc;
    SlotView& v = e->eng.view(0);
    const DGraph& g = v.g;
    // no excess anywhere: the hot sweep's tile pass after the choice moves nothing
    HIPCK(hipMemset(v.excess, 0, sizeof(int64_t) * g.n));
    HIPCK(hipMemcpy(v.tile_cnt, tile_cnt, sizeof(int32_t) * g.ntiles, hipMemcpyHostToDevice));
    const int saved = e->eng.params().hot_k;
    e->eng.params_mut().hot_k = k;
    Plan p;
    p.stages = ST_HOT;
    rc = e->eng.dbg_plan(p);
    e->eng.params_mut().hot_k = saved;
    if (rc == MF_OK) HIPCK(hipMemcpy(tile_hot, v.tile_hot, sizeof(int32_t) * g.ntiles, hipMemcpyDeviceToHost));
    e->eng.dbg_end();
    return rc;
}

int mf_dbg_prune_mask(mf_engine* e, int gid, int s, int t, int32_t* mask, int* kept) {
    int rc = check_query(e, gid, s, t);
    if (rc != MF_OK) return rc;
    rc = e->eng.dbg_begin(gid, s, t);
    if (rc != MF_OK) return rc;
    SlotView& v = e->eng.view(0);
    Plan p;
    p.stages = ST_RESET_PRE | ST_BFS_SEED1 | ST_BFS;
    p.reset_pre = RST_BFS | RST_WORK | RST_ERR;
    p.seed_mode = BFS_FWD;
    p.bfs_kind = BK_SMALL;
    p.bfs_mode = BFS_FWD;
    rc = dbg_bfs(e, p, BFS_FWD);
    if (rc == MF_OK) {
        p.seed_mode = BFS_UND;
        p.bfs_mode = BFS_UND;
        p.reset_pre = RST_BFS;
        rc = dbg_bfs(e, p, BFS_UND);
    }
    if (rc == MF_OK) {
        Plan m;
        m.stages = ST_RESET_PRE | ST_MASK;
        m.reset_pre = RST_BFS;
        rc = e->eng.dbg_plan(m);
    }
    if (rc == MF_OK) {
        if (kept) *kept = e->eng.ctr(0).visited;
        if (mask) HIPCK(hipMemcpy(mask, v.aux, sizeof(int32_t) * v.g.n, hipMemcpyDeviceToHost));
    }
    e->eng.dbg_end();
    return rc;
}

int mf_dbg_class_mask(mf_engine* e, int gid, int s, int t, int32_t* mask, int64_t* kept, int* reachable) {
    int rc = check_query(e, gid, s, t);
    if (rc != MF_OK) return rc;
    if (!mask || !kept || !reachable) return fail(MF_EINVAL, "null argument");
    rc = e->eng.dbg_begin(gid, s, t);
    if (rc != MF_OK) return rc;
    DeviceGraph* G = e->eng.graph(gid);
    if (G->classes && !G->fwd_find(G->scc[s])) {
        // the forward BFS from s the scheduler's S_PRUNE_FWD step runs
        Plan p;
        p.stages = ST_RESET_PRE | ST_BFS_SEED1 | ST_BFS;
        p.reset_pre = RST_BFS | RST_WORK | RST_ERR;
        p.seed_mode = BFS_FWD;
        p.bfs_kind = BK_SMALL;
        p.bfs_mode = BFS_FWD;
        rc = dbg_bfs(e, p, BFS_FWD);
    }
    if (rc == MF_OK) rc = e->eng.dbg_class_mask(gid, s, t, mask, kept, reachable);
    e->eng.dbg_end();
    return rc;
}

int mf_dbg_induced(mf_engine* e, int gid, const int32_t* mask, int* n_out, int64_t* E_out, int32_t* row_ptr,
                   int32_t* col, int64_t* cap, int32_t* rev, int32_t* old2sub) {
    if (!e || !mask) return fail(MF_EINVAL, "null argument");
    if (gid < 0 || gid >= e->eng.ngraphs()) return fail(MF_EINVAL, "bad graph index");
    DeviceGraph* P = e->eng.graph(gid);
    const int n = P->g.n;
    int kept = 0;
    for (int i = 0; i < n; ++i) kept += mask[i] ? 1 : 0;
    int32_t* d_mask = nullptr;
    HIPCK(hipMalloc(&d_mask, sizeof(int32_t) * std::max(n, 1)));
    HIPCK(hipMemcpy(d_mask, mask, sizeof(int32_t) * n, hipMemcpyHostToDevice));
    DeviceGraph out;
    int rc = build_induced_gpu(*P, d_mask, kept, e->eng.stream(), out);
    (void)hipFree(d_mask);
    if (rc != MF_OK) return rc;
    if (n_out) *n_out = out.g.n;
    if (E_out) *E_out = out.g.E;
    if (row_ptr) HIPCK(hipMemcpy(row_ptr, out.d_row_ptr, sizeof(int32_t) * (out.g.n + 1), hipMemcpyDeviceToHost));
    if (col && out.g.E) HIPCK(hipMemcpy(col, out.d_col, sizeof(int32_t) * out.g.E, hipMemcpyDeviceToHost));
    if (cap && out.g.E) HIPCK(hipMemcpy(cap, out.d_cap, sizeof(int64_t) * out.g.E, hipMemcpyDeviceToHost));
    if (rev && out.g.E) HIPCK(hipMemcpy(rev, out.d_rev, sizeof(int32_t) * out.g.E, hipMemcpyDeviceToHost));
    if (old2sub) HIPCK(hipMemcpy(old2sub, out.d_old2sub, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    return MF_OK;
}

int mf_dbg_solve_state(mf_engine* e, int gid, int s, int t, mf_result* r, int64_t* res, int64_t* excess,
                       int32_t* height) {
    int rc = check_query(e, gid, s, t);
    if (rc != MF_OK) return rc;
    // the reference's preflow, on the full graph: the state is the
    // reference algorithm's and an independent CPU certificate reads it
    const int saved = e->eng.params().prune, saved_inj = e->eng.params().inject;
    e->eng.params_mut().prune = 0;
    e->eng.params_mut().inject = 0;
    mf_result tmp;
    rc = mf_solve(e, gid, s, t, &tmp);
    e->eng.params_mut().prune = saved;
    e->eng.params_mut().inject = saved_inj;
    if (r) *r = tmp;
    if (rc != MF_OK) return rc;
    // the plain solve ran in slot 0 on the full graph
    Slab* sl = e->eng.slab(0);
    const DGraph& g = e->eng.graph(gid)->g;
    if (res) HIPCK(hipMemcpy(res, sl->res, sizeof(int64_t) * g.E, hipMemcpyDeviceToHost));
    if (excess) HIPCK(hipMemcpy(excess, sl->excess, sizeof(int64_t) * g.n, hipMemcpyDeviceToHost));
    if (height) HIPCK(hipMemcpy(height, sl->h, sizeof(int32_t) * g.n, hipMemcpyDeviceToHost));
    return MF_OK;
}

}  // extern "C"
