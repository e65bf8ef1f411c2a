// mf_engine.cpp -- resident graphs, slabs, scheduler and step executor.
#include "mf_engine.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <array>
#include <chrono>
#include <climits>
#include <cstring>
#include <mutex>
#include <thread>

#include "mf_kernels.h"

namespace mf {

// ------------------------------------------------------------------ errors
namespace {
thread_local std::string g_err;
}
void set_error(const std::string& msg) { g_err = msg; }
const std::string& last_error() { return g_err; }

#define HIPCHK(expr)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess) {                                                          \
            set_error(std::string(#expr) + ": " + hipGetErrorString(e_));                \
            return e_ == hipErrorOutOfMemory ? MF_ENOMEM : MF_EHIP;                      \
        }                                                                                \
    } while (0)

double wall_seconds() {
    using namespace std::chrono;
    return duration_cast<duration<double>>(steady_clock::now().time_since_epoch()).count();
}

// Every engine's clock starts at the same instant, so times from engines
// serving side by side compare directly.
static double clock_epoch() {
    static const double t0 = wall_seconds();
    return t0;
}

// Engines alive in this process.  Each may run a multi-workgroup BFS on its
// main and its side stream at once, and every workgroup of such a launch must
// be resident together (they meet at barriers): the launches share a budget
// of kMultiBudget workgroups of 256 threads, half of what 256 CUs hold at 8
// such workgroups (32 waves) each, so other kernels always keep room to
// finish and free the slots a launch still waits for.
namespace {
std::atomic<int> g_engines{0};
constexpr int kMultiBudget = 1024;
}  // namespace

int multi_wg_cap() { return std::max(32, kMultiBudget / (2 * std::max(1, g_engines.load()))); }

// ----------------------------------------------------------- device graph
uint64_t next_graph_uid() {
    static std::atomic<uint64_t> next{1};
    return next.fetch_add(1);
}

DeviceGraph::~DeviceGraph() {
    (void)hipFree(d_row_ptr); (void)hipFree(d_col); (void)hipFree(d_cap); (void)hipFree(d_rev);
    (void)hipFree(d_tile_start); (void)hipFree(d_tile_of);
    (void)hipFree(d_sub2old); (void)hipFree(d_old2sub); (void)hipFree(d_mask);
    (void)hipFree(d_wcc);
    (void)hipFree(d_scc);
    (void)hipFree(d_live);
    (void)hipFree(d_capsum); (void)hipFree(d_cert_fp); (void)hipFree(d_pcap);
    (void)hipFree(d_fstart);
    (void)hipFree(d_bu_pick);
    (void)hipFree(d_fwdm);
}

DeviceGraph::FwdSet::~FwdSet() { (void)hipFree(d_mask); }

std::shared_ptr<DeviceGraph::FwdSet> DeviceGraph::fwd_find(int32_t cls) const {
    std::lock_guard<std::mutex> lk(fwd_mu);
    auto it = fwd_cache.find(cls);
    return it == fwd_cache.end() ? nullptr : it->second;
}

std::shared_ptr<DeviceGraph::FwdSet> DeviceGraph::fwd_insert(int32_t cls, std::shared_ptr<FwdSet> set) {
    std::lock_guard<std::mutex> lk(fwd_mu);
    auto ins = fwd_cache.emplace(cls, std::move(set));
    return ins.first->second;
}

void make_tiles(const std::vector<int32_t>& rstart, int tile_max, std::vector<int32_t>& ts) {
    ts.clear();
    const int R = (int)rstart.size() - 1;
    for (int r = 0; r < R; ++r)
        for (int a = rstart[r]; a < rstart[r + 1]; a += tile_max) ts.push_back(a);
    ts.push_back(R >= 0 ? rstart[R] : 0);
}

int upload_graph(const HostGraph& hg, int tile_max, DeviceGraph& dg) {
    if (tile_max < 1 || tile_max > kTileMax) { set_error("tile_max out of range"); return MF_EINVAL; }
    const int n = hg.n;
    const int64_t E = hg.E();
    if (E >= INT32_MAX) { set_error("graph has 2^31 or more CSR entries"); return MF_EINVAL; }
    std::vector<int32_t> rstart = hg.rstart;
    if (rstart.empty()) rstart = {0, n};  // unpartitioned: one region
    std::vector<int32_t> ts;
    make_tiles(rstart, tile_max, ts);
    const int ntiles = (int)ts.size() - 1;
    std::vector<int32_t> tile_of(n);
    for (int k = 0; k < ntiles; ++k)
        for (int v = ts[k]; v < ts[k + 1]; ++v) tile_of[v] = k;
    HIPCHK(hipMalloc(&dg.d_row_ptr, sizeof(int32_t) * ((size_t)n + 1)));
    HIPCHK(hipMalloc(&dg.d_col, sizeof(int32_t) * std::max<int64_t>(E, 1)));
    HIPCHK(hipMalloc(&dg.d_cap, sizeof(int64_t) * std::max<int64_t>(E, 1)));
    HIPCHK(hipMalloc(&dg.d_rev, sizeof(int32_t) * std::max<int64_t>(E, 1)));
    HIPCHK(hipMalloc(&dg.d_tile_start, sizeof(int32_t) * ((size_t)ntiles + 1)));
    HIPCHK(hipMalloc(&dg.d_tile_of, sizeof(int32_t) * std::max(n, 1)));
    HIPCHK(hipMemcpy(dg.d_row_ptr, hg.row_ptr.data(), sizeof(int32_t) * ((size_t)n + 1), hipMemcpyHostToDevice));
    if (E > 0) {
        HIPCHK(hipMemcpy(dg.d_col, hg.col.data(), sizeof(int32_t) * E, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dg.d_cap, hg.cap.data(), sizeof(int64_t) * E, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dg.d_rev, hg.rev.data(), sizeof(int32_t) * E, hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpy(dg.d_tile_start, ts.data(), sizeof(int32_t) * ts.size(), hipMemcpyHostToDevice));
    if (n > 0) HIPCHK(hipMemcpy(dg.d_tile_of, tile_of.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    dg.g.n = n;
    dg.g.E = (int)E;
    dg.g.ntiles = ntiles;
    for (int v = 0; v < n; ++v) dg.g.max_deg = std::max(dg.g.max_deg, hg.row_ptr[v + 1] - hg.row_ptr[v]);
    dg.g.R = (int)rstart.size() - 1;
    dg.g.row_ptr = dg.d_row_ptr;
    dg.g.col = dg.d_col;
    dg.g.cap = dg.d_cap;
    dg.g.rev = dg.d_rev;
    dg.g.tile_start = dg.d_tile_start;
    dg.g.tile_of = dg.d_tile_of;
    dg.tile_max = tile_max;
    dg.bytes = sizeof(int32_t) * ((size_t)n + 1) + (size_t)E * 16 + sizeof(int32_t) * (ts.size() + n);
    // reachability classes for pruning
    strong_components(hg, dg.scc);
    weak_components(hg, dg.wcc);
    HIPCHK(hipMalloc(&dg.d_wcc, sizeof(int32_t) * std::max(n, 1)));
    if (n > 0) HIPCHK(hipMemcpy(dg.d_wcc, dg.wcc.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&dg.d_scc, sizeof(int32_t) * std::max(n, 1)));
    if (n > 0) HIPCHK(hipMemcpy(dg.d_scc, dg.scc.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    dg.scc_size.assign(n, 0);
    for (int v = 0; v < n; ++v) ++dg.scc_size[dg.scc[v]];
    // vertices that can reach the largest SCC: a reverse search over the
    // forward entries (an entry with cap 0 in w's row is the reverse of the
    // forward entry col -> w)
    if (n > 0) {
        dg.giant_scc = (int32_t)(std::max_element(dg.scc_size.begin(), dg.scc_size.end()) - dg.scc_size.begin());
        std::vector<uint8_t> live(n, 0);
        std::vector<int32_t> q;
        q.reserve(n);
        for (int v = 0; v < n; ++v)
            if (dg.scc[v] == dg.giant_scc) { live[v] = 1; q.push_back(v); }
        for (size_t h = 0; h < q.size(); ++h) {
            const int w = q[h];
            for (int e = hg.row_ptr[w]; e < hg.row_ptr[w + 1]; ++e)
                if (hg.cap[e] == 0 && !live[hg.col[e]]) { live[hg.col[e]] = 1; q.push_back(hg.col[e]); }
        }
        HIPCHK(hipMalloc(&dg.d_live, n));
        HIPCHK(hipMemcpy(dg.d_live, live.data(), n, hipMemcpyHostToDevice));
        dg.bytes += (size_t)n;
    }
    dg.classes = true;
    dg.bytes += 2 * sizeof(int32_t) * (size_t)n;
    dg.fwd_bits.assign(((size_t)E + 63) / 64, 0);
    for (int64_t e = 0; e < E; ++e)
        if (hg.cap[e] > 0) dg.fwd_bits[(size_t)e >> 6] |= 1ull << (e & 63);
    return init_cert_consts(dg, nullptr);
}

int init_cert_consts(DeviceGraph& dg, hipStream_t st) {
    static uint64_t graphs_made = 0;
    const int n = dg.g.n;
    HIPCHK(hipMalloc(&dg.d_capsum, sizeof(int64_t) * std::max(n, 1)));
    HIPCHK(hipMalloc(&dg.d_cert_fp, sizeof(uint64_t)));
    HIPCHK(hipMalloc(&dg.d_pcap, sizeof(int64_t) * std::max(dg.g.E, 1)));
    HIPCHK(hipMalloc(&dg.d_fstart, sizeof(int32_t) * std::max(n, 1)));
    HIPCHK(hipMemsetAsync(dg.d_cert_fp, 0, sizeof(uint64_t), st));
    dg.g.capsum = dg.d_capsum;
    dg.g.cert_fp = dg.d_cert_fp;
    dg.g.pcap = dg.d_pcap;
    launch_first_forward(st, dg.g, dg.d_fstart);
    dg.g.fstart = dg.d_fstart;
    HIPCHK(hipMalloc(&dg.d_bu_pick, sizeof(int32_t) * 2 * (size_t)std::max(n, 1)));
    launch_bu_pick(st, dg.g, dg.d_bu_pick);
    dg.g.bu_pick = dg.d_bu_pick;
    HIPCHK(hipMalloc(&dg.d_fwdm, sizeof(uint64_t) * (((size_t)std::max(dg.g.E, 1) + 63) / 64)));
    launch_fwd_mask(st, dg.g, dg.d_fwdm);
    dg.g.fwdm = dg.d_fwdm;
    dg.g.cert_seed = 0x9e3779b97f4a7c15ULL * (++graphs_made) ^ ((uint64_t)(uint32_t)n << 32) ^ (uint64_t)(uint32_t)dg.g.E;
    launch_cert_consts(st, dg.g, dg.d_capsum, dg.d_cert_fp);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    dg.bytes += sizeof(int64_t) * ((size_t)n + (size_t)dg.g.E) + sizeof(uint64_t) + 3 * sizeof(int32_t) * (size_t)n;
    return MF_OK;
}

// ------------------------------------------------------------------- slabs
size_t Slab::layout(int n, int64_t E, int ntiles, char* base, Slab* sl) {
    const size_t nn = (size_t)std::max(n, 1), ee = (size_t)std::max<int64_t>(E, 1),
                 tt = (size_t)std::max(ntiles, 1);
    // rows longer than kGiantDeg: at most E / (kGiantDeg + 1) of them
    const size_t gcap = ee / (kGiantDeg + 1) + 1;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t at = off;
        off += (bytes + 255) & ~(size_t)255;
        return base ? base + at : nullptr;
    };
    char* p_res = take(8 * ee);
    char* p_ex = take(8 * nn);
    char* p_h = take(4 * nn);
    char* p_fa = take(4 * nn);
    char* p_fb = take(4 * nn);
    char* p_lock = take(4 * nn);
    char* p_aux = take(4 * nn);
    char* p_hl = take(4 * nn);
    char* p_tc = take(4 * tt);
    char* p_th = take(4 * tt);
    char* p_tx = take(8 * tt);
    char* p_ts = take(sizeof(int32_t) * (1 + 3 * (size_t)kSplitMax * tt));
    char* p_lb = take(sizeof(int32_t) * 2 * kLocalLevels);
    char* p_gi = take(sizeof(int32_t) * (giant_task_offset((int64_t)gcap) + 4 * (ee / kGiantChunk + gcap + 1)));
    char* p_ga = take(sizeof(int64_t) * gcap);
    char* p_fb2 = take(16 * ((nn + 63) / 64));
    char* p_recv = take(nn);
    char* p_em = take(ee);
    const size_t mcap = std::min(ee, (size_t)kMarkListMax);
    char* p_ml = take(sizeof(int32_t) * (mcap + 1));
    if (sl) {
        sl->res = (int64_t*)p_res;
        sl->excess = (int64_t*)p_ex;
        sl->h = (int32_t*)p_h;
        sl->fa = (int32_t*)p_fa;
        sl->fb = (int32_t*)p_fb;
        sl->lock = (int32_t*)p_lock;
        sl->aux = (int32_t*)p_aux;
        sl->hl = (int32_t*)p_hl;
        sl->tile_cnt = (int32_t*)p_tc;
        sl->tile_hot = (int32_t*)p_th;
        sl->tile_ex = (int64_t*)p_tx;
        sl->tsplit = (int32_t*)p_ts;
        sl->lbkt = (int32_t*)p_lb;
        sl->giant = (int32_t*)p_gi;
        sl->gacc = (int64_t*)p_ga;
        sl->fbits = (uint64_t*)p_fb2;
        sl->recv = (uint8_t*)p_recv;
        sl->emark = (uint8_t*)p_em;
        sl->mlist = (int32_t*)p_ml;
        sl->mcap = (int)mcap;
        sl->res_uid = 0;
        sl->marks_ok = false;
        sl->gcap = (int)gcap;
        sl->n_cap = n;
        sl->E_cap = (int)E;
        sl->tiles_cap = ntiles;
    }
    return off;
}

int SlabArena::create(size_t bytes) {
    destroy();
    bytes = (bytes + 255) & ~(size_t)255;
    if (bytes == 0) bytes = 256;
    HIPCHK(hipMalloc(&base, bytes));
    size = bytes;
    used = peak = 0;
    free_.clear();
    free_[0] = bytes;
    return MF_OK;
}

void SlabArena::destroy() {
    if (base) (void)hipFree(base);
    base = nullptr;
    size = used = 0;
    free_.clear();
}

bool SlabArena::alloc(size_t bytes, size_t* off) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second < bytes) continue;
        *off = it->first;
        const size_t rest = it->second - bytes;
        const size_t at = it->first + bytes;
        free_.erase(it);
        if (rest) free_[at] = rest;
        used += bytes;
        peak = std::max(peak, used);
        return true;
    }
    return false;
}

void SlabArena::release(size_t off, size_t bytes) {
    used -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
        it->second += nx->second;
        free_.erase(nx);
    }
    if (it != free_.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            free_.erase(it);
        }
    }
}

// --------------------------------------------------------------- scheduler
namespace {
inline int cdiv(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }
inline int clampi(int64_t x, int lo, int hi) { return (int)std::max<int64_t>(lo, std::min<int64_t>(hi, x)); }
}  // namespace

void SlotSched::begin(const Params* P, const DGraph& g, int s, int t, bool prune_first, bool warm) {
    P_ = P;
    g_ = g;
    s_ = s;
    t_ = t;
    status_ = MF_OK;
    err_.clear();
    phase_ = 0;
    stats = QueryStats();
    stats.n_run = g.n;
    stats.E_run = g.E;
    const int avg = g.n > 0 ? (int)(g.E / g.n) : 0;
    K_ = P->K > 0 ? P->K : std::max(4, avg);  // BC:1465
    GR_FREQ_ = P->gr_freq;
    RM_FREQ_ = P->rm_freq0;
    iter_ = 0;
    last_gr_iter_ = 0;
    gr_depth_ = 0;
    depth_hint_ = 0;
    spec_ = false;
    spec_failed_ = false;
    disc_cap_ = 0x7fffffff;
    stag_ = no_stag_ = 0;
    last_pos_ = INT64_MAX;
    pos_full_ = ext_full_ = 0;
    counted_sparse_ = false;
    force_gr_ = false;
    path_stable_ = 0;
    p_count_ = P->p_count0;
    path_rounds_ = 0;
    last_ext_ = -1;
    stall_ = 0;
    drain_loops_ = 0;
    rm_off_ = false;
    just_gr_ = false;
    after_local_ = false;
    work_reset_ = false;
    fwd_only_ = false;
    deficits_ = warm;
    no_multi_ = false;
    st_ = prune_first ? S_PRUNE_FWD : (warm ? S_WARM_START : S_START);
}

void SlotSched::resume_after_prune(const DGraph& g, int s, int t) {
    g_ = g;
    s_ = s;
    t_ = t;
    stats.n_run = g.n;
    stats.E_run = g.E;
    const int avg = g.n > 0 ? (int)(g.E / g.n) : 0;
    K_ = P_->K > 0 ? P_->K : std::max(4, avg);
    st_ = S_START;
}

void SlotSched::finish_trivial(int64_t flow) {
    stats.flow = flow;
    stats.cert_ok = 1;
    stats.n_run = 0;
    stats.E_run = 0;
    st_ = S_DONE;
}

void SlotSched::fail(int code, const std::string& msg) {
    status_ = code;
    err_ = msg;
    st_ = S_ERROR;
    stats.status = code;
}

Plan SlotSched::start_gr(const Ctr* c) {
    Plan p;
    scans_at_gr_ = c ? c->wk[WF_DISCHARGE][0] : 0;
    // deficits come only from RM repairs and capacity updates (warm starts):
    // when neither can have left one, t is the one seed
    p.seed_t_only = c && phase_ == 0 && !deficits_;
    p.stages = ST_RESET_PRE | ST_GR_SEED | ST_BFS;
    p.reset_pre = RST_BFS;
    p.bfs_kind = BK_SMALL;
    p.bfs_mode = BFS_GR;
    p.small_edges = deep_bfs(0) ? P_->small_edges_deep : P_->small_edges;
    bfs_mode_ = BFS_GR;
    bfs_medges_ = 0;
    bfs_bottom_up_ = false;
    last_gr_iter_ = iter_;
    force_gr_ = false;
    stall_ = 0;
    stats.grs++;
    st_ = S_BFS;
    return p;
}

// One BFS level decision.  Direction-optimizing (Beamer): go bottom-up when
// the frontier's rows outweigh the unvisited rows / alpha, back to top-down
// when the frontier is below n / beta; small frontiers run many levels in
// one single-workgroup launch.
Plan SlotSched::bfs_step(const Ctr& c, int mode) {
    Plan p;
    const int L = c.level;
    const int f = c.fsize[L & 3];
    const int64_t mf = c.fedges[L & 3];
    p.stages = ST_BFS;
    p.bfs_mode = mode;
    const int64_t mu = std::max<int64_t>(0, (int64_t)g_.E - bfs_medges_);
    if (!bfs_bottom_up_ && (double)mf > (double)mu / P_->bu_alpha) {
        bfs_bottom_up_ = true;
        p.bu_bits = true;  // level L came from a top-down or single-workgroup level
    } else if (bfs_bottom_up_ && (double)f < (double)g_.n / P_->bu_beta)
        bfs_bottom_up_ = false;
    if (bfs_bottom_up_) {
        p.bfs_kind = BK_BU;
        p.bfs_grid = clampi(cdiv(g_.n, 256), 1, 8192);
    } else if (f <= kSmallFrontier && mf <= (deep_bfs(L) ? P_->small_edges_deep : P_->small_edges)) {
        p.bfs_kind = BK_SMALL;
        p.small_edges = deep_bfs(L) ? P_->small_edges_deep : P_->small_edges;
        if (!no_multi_ && deep_bfs(L)) p.small_wgs = std::max(1, P_->small_wgs);
    } else {
        p.bfs_kind = BK_TD;
        p.bfs_grid = std::max(1, cdiv(f, 256));
    }
    return p;
}

// A multi-workgroup BFS whose barrier timed out left a partial level: the
// BFS starts again from its seeds, single-workgroup.  A global relabel and
// the pruning BFS both recompute everything they write from the seeds.
Plan SlotSched::restart_bfs(const Ctr& c) {
    no_multi_ = true;
    stats.bfs_restarts++;
    if (st_ == S_BFS && bfs_mode_ == BFS_GR) {
        stats.grs--;  // the restart is the same relabel
        Plan p = start_gr(&c);
        p.reset_pre |= RST_ERR;
        return p;
    }
    st_ = S_PRUNE_FWD;
    Plan p = plan(c);  // S_PRUNE_FWD resets the error word
    return p;
}

Plan SlotSched::plan(const Ctr& c) {
    Plan p;
    // a kernel flagged the slot: bit 4 = a multi-workgroup BFS barrier gave
    // up (retried single-workgroup); anything else fails the query
    if (c.err && st_ != S_START && st_ != S_WARM_START && st_ != S_PRUNE_FWD && st_ != S_DONE) {
        if (c.err == 4 && (st_ == S_BFS || st_ == S_PRUNE_UND || st_ == S_PRUNE_MASK) && stats.bfs_restarts < 4)
            return restart_bfs(c);
        fail(MF_EHIP, "device error flag " + std::to_string(c.err));
        return p;
    }
    switch (st_) {
    case S_PRUNE_FWD: {
        // forward BFS from s over cap > 0 (BC:1402-1416)
        p.stages = ST_RESET_PRE | ST_BFS_SEED1 | ST_BFS;
        p.reset_pre = RST_BFS | RST_WORK | RST_ERR;
        work_reset_ = true;
        p.seed_mode = BFS_FWD;
        p.bfs_kind = BK_SMALL;
        p.bfs_mode = BFS_FWD;
        bfs_mode_ = BFS_FWD;
        bfs_medges_ = 0;
        bfs_bottom_up_ = false;
        st_ = S_PRUNE_UND;
        return p;
    }
    case S_PRUNE_UND:
    case S_PRUNE_MASK: {
        const int L = c.level;
        if (c.fsize[L & 3] > 0) {
            bfs_medges_ += c.fedges[L & 3];
            return bfs_step(c, bfs_mode_);
        }
        if (st_ == S_PRUNE_UND && fwd_only_) {
            st_ = S_PRUNE_DECIDE;  // the engine combines h with t's WCC
            return p;
        }
        if (st_ == S_PRUNE_UND) {
            // undirected BFS from t (BC:1419-1433)
            p.stages = ST_RESET_PRE | ST_BFS_SEED1 | ST_BFS;
            p.reset_pre = RST_BFS;
            p.seed_mode = BFS_UND;
            p.bfs_kind = BK_SMALL;
            p.bfs_mode = BFS_UND;
            bfs_mode_ = BFS_UND;
            bfs_medges_ = 0;
            bfs_bottom_up_ = false;
            st_ = S_PRUNE_MASK;
            return p;
        }
        p.stages = ST_RESET_PRE | ST_MASK;
        p.reset_pre = RST_BFS;
        st_ = S_PRUNE_DECIDE;
        return p;
    }
    case S_PRUNE_DECIDE:
        // the engine must call resume_after_prune first
        return p;
    case S_START:
    case S_WARM_START: {
        // cold: residual init (BC:151-170) + preflow (BC:173-200);
        // warm: the slab keeps the residual, preflow saturates whatever new
        // residual s has (BC:2304-2322).
        const bool cold = (st_ == S_START);
        p = start_gr();
        p.stages |= ST_PREFLOW | (cold ? ST_INIT : 0u);
        p.reset_pre = RST_BFS | RST_ACTIVE | RST_CERT | RST_RM | RST_ERR | (work_reset_ ? 0u : (unsigned)RST_WORK);
        work_reset_ = true;
        stats.grs = 1;
        return p;
    }
    case S_BFS: {
        const int L = c.level;
        if (c.fsize[L & 3] > 0) {
            bfs_medges_ += c.fedges[L & 3];
            return bfs_step(c, bfs_mode_);
        }
        stats.bfs_levels += L;
        if (bfs_mode_ == BFS_GR) {
            gr_depth_ = L;
            gr_visited_ = c.visited;
        }
        p.stages = ST_RESET_ACT | ST_ACTIVE;
        p.reset_act = RST_ACTIVE;
        just_gr_ = true;
        st_ = S_COUNTED;
        return p;
    }
    case S_COUNTED: {
        const bool exact = just_gr_;
        just_gr_ = false;
        const bool sparse = counted_sparse_;
        counted_sparse_ = false;
        int64_t pos = c.pos_excess;
        if (sparse) {
            pos = pos_full_ - (c.ex_t - ext_full_);
        } else {
            pos_full_ = c.pos_excess;
            ext_full_ = c.ex_t;
            deficits_ = c.neg_count > 0;
        }
        // counts right after a local relabel follow no discharge: they do not
        // feed the stagnation and stall monitors
        const bool was_local = after_local_;
        if (c.err) { fail(MF_EHIP, "device error flag"); return p; }
        const int active = c.active;
        if (active == 0) {
            if (!exact) {
                // The certificate proves optimality for any closed T = {h < n}:
                // try it on the discharge's labels first; a failure falls back
                // to the exact global relabel.
                if (P_->cert && P_->spec_cert && phase_ == 0 && !spec_failed_ && !deficits_) {
                    p.stages = ST_RESET_PRE | ST_CERT;
                    p.reset_pre = RST_CERT;
                    p.cert_sside = 2 * gr_visited_ > g_.n;
                    spec_ = true;
                    st_ = S_CERT;
                    return p;
                }
                force_gr_ = false;
                return start_gr(&c);
            }
            if (phase_ == 0 && c.neg_count > 0) {
                // Deficits left (RM or a capacity decrease made them) and no
                // positive excess can reach t or a deficit: drain them from t.
                if (++drain_loops_ > 64) { fail(MF_EITER, "deficit drain did not converge"); return p; }
                phase_ = 1;
                rm_off_ = true;
                stats.drains++;
                return start_gr(&c);
            }
            if (phase_ == 1) {
                if (c.neg_count > 0) { fail(MF_EITER, "deficits unreachable from t"); return p; }
                phase_ = 0;
                return start_gr(&c);
            }
            // Exact termination: no active vertex can reach t.
            if (P_->cert) {
                p.stages = ST_RESET_PRE | ST_CERT;
                p.reset_pre = RST_CERT;
                p.cert_sside = 2 * (int64_t)c.visited > g_.n;
                st_ = S_CERT;
                return p;
            }
            stats.flow = c.ex_t;
            st_ = S_DONE;
            return p;
        }
        // excess(t) fills t's in-capacity: every entry into t is saturated,
        // so the cut around t has capacity excess(t) and no flow can exceed
        // it, whatever excess is still held elsewhere; certified with
        // T = {t}.  Or it equals s's out-capacity: everything s sent has
        // reached t, and the cut around s (saturated since the preflow)
        // proves it, with T = V \ {s}.  A failure falls back to the global
        // relabel.
        const bool at_t = c.ex_t == c.incap_t, at_s = c.ex_t == c.outcap_s;
        if (P_->sink_cut && P_->cert && phase_ == 0 && !deficits_ && !spec_failed_ && c.incap_t >= 0 &&
            (at_t || at_s)) {
            p.stages = ST_RESET_PRE | (at_t ? ST_SINK : ST_SOURCE) | ST_CERT;
            p.reset_pre = RST_CERT;
            p.cert_sside = !at_t;  // T = V \ {s}: scan s's row
            spec_ = true;
            stats.sink_cuts++;
            st_ = S_CERT;
            return p;
        }
        if (iter_ > P_->iter_max) { fail(MF_EITER, "iteration watchdog fired"); return p; }
        // stagnation monitor on the positive excess sum (BC:1851-1891)
        bool hotspot_triggered = false;
        if (!exact && !was_local && iter_ > 0) {
            if (last_pos_ != INT64_MAX && pos >= last_pos_) ++stag_;
            else stag_ = 0;
            last_pos_ = pos;
            if (stag_ >= P_->stagnate_rounds) {
                hotspot_triggered = true;
                force_gr_ = true;
                stag_ = 0;
                GR_FREQ_ = std::min(P_->gr_freq_max, GR_FREQ_ * 2);
                K_ = std::min(16, K_ * 2);
                no_stag_ = 0;
            } else if (++no_stag_ >= P_->stagnate_rounds) {
                GR_FREQ_ = std::max(P_->gr_freq_min, GR_FREQ_ / 2);
                K_ = std::max(4, K_ / 2);
                no_stag_ = 0;
            }
            if (c.ex_t == last_ext_) ++stall_;
            else stall_ = 0;
            last_ext_ = c.ex_t;
            // forced GR when most active labels are high (BC:1845-1849)
            const float frac_high = active > 0 ? (float)c.active_high / (float)active : 0.f;
            if (frac_high > 0.8f && iter_ % 2048 == 0) force_gr_ = true;
            if (P_->rm && !rm_off_) {
                // RM feedback (BC:2083-2093)
                if (rm_last_) {
                    const double frac_fix = g_.E > 0 ? (double)c.fixed_edges / (double)g_.E : 0.0;
                    if (frac_fix < P_->rm_tau) RM_FREQ_ = std::min(RM_FREQ_ * 2, P_->rm_freq_max);
                    else RM_FREQ_ = std::max(P_->rm_freq0, RM_FREQ_ / 2);
                    stats.rm_fixed += c.fixed_edges;
                }
            }
        } else if (!was_local) {
            last_pos_ = pos;
            last_ext_ = c.ex_t;
        }
        rm_last_ = false;
        // GR when due by count (BC:1496), forced, when the flow into t has
        // stalled, or when the discharge has done about a GR's worth of scans.
        // A stall only counts once the excess has had a quarter as many
        // iterations as the last GR had levels: excess moves about one hop
        // per iteration, so on deep graphs (grids) it is still in transit.
        const uint64_t scans = c.wk[WF_DISCHARGE][0] - scans_at_gr_;
        const double gr_work = gr_depth_ >= 64 ? P_->gr_work_deep : P_->gr_work;
        const bool gr_due = force_gr_ || (iter_ - last_gr_iter_ >= std::max<int64_t>(GR_FREQ_, gr_depth_ / 4)) ||
                            (P_->gr_stall > 0 && stall_ >= P_->gr_stall && 4 * (iter_ - last_gr_iter_) >= gr_depth_) ||
                            (gr_work > 0 && (double)scans >= gr_work * (double)g_.E);
        if (gr_due && !exact) return start_gr(&c);
        // local relabel (BC:1582-1721): BFS from t and the deficits through
        // the tiles holding active vertices only, for at most band + 2 levels,
        // when the active heights span a narrow band; vertices it does not
        // reach keep their height.  The discharge follows.
        const int lcap = std::min(c.hmax + 2, std::min(kLocalLevels, g_.n));
        if (P_->local_relabel && !was_local && !exact && c.hmax >= c.hmin &&
            c.hmax - c.hmin + 1 <= P_->local_band && c.hmax + 2 <= lcap) {
            p.stages = ST_RESET_PRE | ST_LOCAL | ST_RESET_ACT | ST_ACTIVE;
            p.reset_pre = RST_BFS;
            p.reset_act = RST_ACTIVE;
            p.local_cap = lcap;
            after_local_ = true;
            stats.local_relabels++;
            return p;
        }
        after_local_ = false;
        // discharge step
        // deep graphs (last exact relabel >= 64 levels): excess crosses the
        // graph about one hop per discharge iteration, so path pushes start
        // earlier and with more active vertices
        const bool deep = gr_depth_ >= 64;
        const int pswitch = deep ? std::max(P_->path_switch, P_->path_switch_deep) : P_->path_switch;
        const int pc0 = deep ? std::min(P_->p_count0, P_->p_count_deep) : P_->p_count0;
        if (active > 0 && active <= pswitch) ++path_stable_;
        else { path_stable_ = 0; p_count_ = P_->p_count0; }
        const int pneed = deep ? std::min(p_count_, pc0) : p_count_;
        const bool use_path = active > 0 && active <= pswitch && path_stable_ >= pneed;
        p.stages = ST_DISCHARGE | ST_RESET_ACT | ST_ACTIVE;
        p.reset_act = RST_ACTIVE;
        // lowest band first on shallow graphs: the excess nearest t moves
        // first, and with the terminal-cut stop the rest may never need to
        disc_cap_ = (P_->disc_band > 0 && !deep && phase_ == 0 && c.hmin <= c.hmax) ? c.hmin + P_->disc_band
                                                                                   : 0x7fffffff;
        if (use_path) {
            p.dis_kind = DK_PATH;
            p.dis_grid = std::max(1, cdiv(active, 4));
            path_stable_ = 0;
            if (++path_rounds_ >= 4) { path_rounds_ = 0; p_count_ = 2; }  // BC:1827-1829
        } else if ((double)active >= P_->tile_frac * (double)g_.n) {
            p.dis_kind = DK_TILE;
            p.dis_grid = std::max(1, g_.ntiles);
        } else {
            p.dis_kind = DK_LIST;
            p.dis_grid = std::max(1, cdiv(c.nlight, 256) + cdiv(c.nheavy, 4));
        }
        // hot-region extra sweep (BC:1898-1984)
        if (P_->hot && !use_path && active >= P_->hot_min_active && (iter_ % 2 == 0) &&
            !hotspot_triggered)
            p.stages |= ST_HOT;
        // RM repair (BC:1990-2097)
        const float frac_high = active > 0 ? (float)c.active_high / (float)active : 0.f;
        if (P_->rm && !rm_off_ && phase_ == 0 && frac_high >= 0.5f && iter_ != 0 &&
            (iter_ % RM_FREQ_ == 0)) {
            p.stages |= ST_RM;
            p.reset_act |= RST_RM;
            rm_last_ = true;
            deficits_ = true;
        }
        // sparse count: every kernel of this step that pushes records the
        // vertices it pushes into, so those and the previous active list hold
        // every vertex that can be active after it
        if (P_->sparse_count && !use_path && !(p.stages & ST_RM) && phase_ == 0 && !deficits_) {
            p.sparse = true;
            p.stages |= ST_RESET_PRE;
            p.reset_pre |= RST_TOUCH;
            p.reset_act = RST_SPARSE;
            counted_sparse_ = true;
        }
        // deep graphs: several cycles per host round trip, up to the next
        // count-due global relabel.  The extra cycles count sparsely, so only
        // a step whose own count is sparse (no RM, no deficit) fuses: the
        // counters the host reads then always come from a sparse count.
        if (!use_path && p.sparse && P_->fuse > 1 && gr_depth_ >= 64) {
            const int64_t left = std::max<int64_t>(1, std::max<int64_t>(GR_FREQ_, gr_depth_ / 4) - (iter_ - last_gr_iter_));
            p.cycles = (int)std::min<int64_t>(P_->fuse, left);
        }
        iter_ += p.cycles;
        stats.iters = iter_;
        return p;
    }
    case S_CERT: {
        stats.cert_ok = (c.cert_bad == 0 && c.fp == 0 && c.cut == c.ex_t) ? 1 : 0;
        stats.flow = c.ex_t;
        if (!stats.cert_ok && spec_) {  // labels not exact enough: relabel and go on
            spec_ = false;
            spec_failed_ = true;
            stats.spec_fails++;
            return start_gr(&c);
        }
        spec_ = false;
        if (!stats.cert_ok) {
            fail(MF_ECERT, "min-cut certificate failed: bad=" + std::to_string(c.cert_bad) +
                               " fingerprint=" + std::to_string(c.fp) + " cut=" + std::to_string(c.cut) +
                               " excess_t=" + std::to_string(c.ex_t));
            return p;
        }
        st_ = S_DONE;
        return p;
    }
    case S_DONE:
    case S_ERROR:
        return p;
    }
    return p;
}

// ------------------------------------------------------------------ engine
Engine::~Engine() {
    if (counted_) g_engines.fetch_sub(1);
    if (stream_) (void)hipStreamSynchronize(stream_);
    slots_.clear();
    arena_.destroy();
    cache_.clear();
    graphs_.clear();
    (void)hipFree(d_views_);
    (void)hipFree(d_hash_);
    (void)hipFree(d_ctrs_);
    for (auto& kv : graph_cache_) {
        (void)hipGraphExecDestroy(kv.second.x);
        (void)hipGraphDestroy(kv.second.g);
    }
    graph_cache_.clear();
    (void)hipFree(cyc_d_);
    if (cyc_h_) (void)hipHostFree(cyc_h_);
    if (h_stage_) (void)hipHostFree(h_stage_);
    if (h_ctr_) (void)hipHostFree(h_ctr_);
    if (h_hash_) (void)hipHostFree(h_hash_);
    for (auto& ev : ev_) (void)hipEventDestroy(ev);
    if (stream_) (void)hipStreamDestroy(stream_);
    if (side_) (void)hipStreamDestroy(side_);
    if (ev_up_) (void)hipEventDestroy(ev_up_);
    if (ev_side_) (void)hipEventDestroy(ev_side_);
}

double Engine::now() const { return wall_seconds() - t0_; }


long long barrier_spin_ticks() {
    int dev = 0, khz = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess ||
        khz <= 0)
        khz = 100000;  // 100 MHz
    return 2000LL * khz;
}

int Engine::init(int nslots, const Params& P) {
    if (nslots < 1) { set_error("nslots must be >= 1"); return MF_EINVAL; }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) { set_error("no HIP device"); return MF_ENODEV; }
    P_ = P;
    nslots_ = nslots;
    HIPCHK(hipGetDevice(&device_));
    spin_ticks_ = barrier_spin_ticks();
    HIPCHK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&ev_up_, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&ev_side_, hipEventDisableTiming));
    // views followed by the launch lists: a slot sits in at most one group per
    // stage plus three per extra fused cycle
    // (+ the list discharge's work queue: 3 + nslots words)
    list_cap_ = (size_t)nslots * (33 + 3 * (size_t)std::max(1, P.fuse)) + 3;
    HIPCHK(hipMalloc(&d_views_, sizeof(SlotView) * nslots + sizeof(int) * list_cap_));
    HIPCHK(hipMalloc(&d_hash_, sizeof(unsigned long long) * nslots));
    HIPCHK(hipHostMalloc(&h_stage_, sizeof(SlotView) * nslots + sizeof(int) * list_cap_));
    HIPCHK(hipHostMalloc((void**)&h_ctr_, sizeof(Ctr) * nslots));
    HIPCHK(hipMalloc((void**)&d_ctrs_, sizeof(Ctr) * nslots));
    HIPCHK(hipMemset(d_ctrs_, 0, sizeof(Ctr) * nslots));
    HIPCHK(hipHostMalloc((void**)&h_hash_, sizeof(unsigned long long) * nslots));
    {
        const size_t cyc = (size_t)std::max(1, P.fuse) * 3 * (size_t)nslots;
        HIPCHK(hipMalloc(&cyc_d_, sizeof(int32_t) * cyc));
        HIPCHK(hipHostMalloc((void**)&cyc_h_, sizeof(int32_t) * cyc));
    }
    memset(h_ctr_, 0, sizeof(Ctr) * nslots);
    ev_.resize(256);  // timed launch ranges per step (profiling)
    for (auto& ev : ev_) HIPCHK(hipEventCreate(&ev));
    slots_.resize(nslots);
    t0_ = clock_epoch();
    g_engines.fetch_add(1);
    counted_ = true;
    return MF_OK;
}

int Engine::add_graph(std::shared_ptr<DeviceGraph> g) {
    graphs_.push_back(g);
    if (arena_.base && budget_ == 0) arena_stale_ = true;  // resized when nothing holds a slab
    return (int)graphs_.size() - 1;
}

int Engine::make_arena() {
    size_t bytes = budget_;
    if (bytes == 0) {
        size_t full = 0;
        for (auto& g : graphs_) full = std::max(full, Slab::bytes_for(g->g.n, g->g.E, g->g.ntiles));
        bytes = full * (size_t)nslots_;
    }
    arena_stale_ = false;
    return arena_.create(bytes);
}

int Engine::set_budget(size_t bytes) {
    for (auto& S : slots_)
        if (S.live || S.pinned) { set_error("set_budget: slots are in use"); return MF_EBUSY; }
    for (int i = 0; i < nslots_; ++i) drop_slab(i);
    arena_.destroy();
    budget_ = bytes;
    return MF_OK;
}

void Engine::mem_stats(size_t* budget, size_t* in_use, size_t* peak, int64_t* waits) const {
    if (budget) *budget = arena_.size;
    if (in_use) *in_use = arena_.used;
    if (peak) *peak = arena_.peak;
    if (waits) *waits = admit_waits_;
}

void Engine::drop_slab(int i) {
    Slab& sl = slots_[i].slab;
    if (sl.held) arena_.release(sl.off, sl.bytes);
    sl = Slab();
}

int Engine::ensure_slab(int i, int n, int64_t E, int ntiles) {
    if (slots_[i].slab.fits(n, E, ntiles)) return MF_OK;
    drop_slab(i);
    if (arena_stale_) {
        bool busy = false;
        for (auto& S : slots_) busy = busy || S.live || S.pinned;
        if (!busy) {
            for (int j = 0; j < nslots_; ++j) drop_slab(j);
            arena_.destroy();
        }
    }
    if (!arena_.base) {
        int rc = make_arena();
        if (rc != MF_OK) return rc;
    }
    const size_t need = Slab::bytes_for(n, E, ntiles);
    size_t off = 0;
    if (!arena_.alloc(need, &off)) {
        // slabs kept by idle slots go back first
        for (int j = 0; j < nslots_; ++j)
            if (j != i && !slots_[j].live && !slots_[j].pinned) drop_slab(j);
        if (!arena_.alloc(need, &off)) return kNoRoom;
    }
    Slab& sl = slots_[i].slab;
    Slab::layout(n, E, ntiles, arena_.base + off, &sl);
    sl.off = off;
    sl.bytes = need;
    sl.held = true;
    return MF_OK;
}

void Engine::run_dims(const Query& q, int* n, int64_t* E, int* ntiles) const {
    const DeviceGraph& G = *graphs_[q.graph];
    *n = G.g.n;
    *E = G.g.E;
    *ntiles = G.g.ntiles;
    if (!P_.prune || !G.classes) return;
    const int cls = G.scc[q.s];
    if (!intra_route(G, q)) {
        auto fwd = G.fwd_find(cls);
        if (!fwd || !fwd->mask[q.t]) return;
    }
    auto ct = cache_.find(route_key(G, q));
    if (ct == cache_.end() || !ct->second.g || ct->second.version != G.cap_version.load()) return;
    const DGraph& r = ct->second.g->g;
    *n = r.n;
    *E = r.E;
    *ntiles = r.ntiles;
}


int Engine::submit(const Query& q) {
    if (q.graph < 0 || q.graph >= (int)graphs_.size()) { set_error("bad graph index"); return MF_EINVAL; }
    queue_.push_back(q);
    if (queue_.back().t_arrive <= 0) queue_.back().t_arrive = now();
    return MF_OK;
}

int Engine::live() const {
    int k = 0;
    for (auto& s : slots_) k += s.live ? 1 : 0;
    return k;
}

int Engine::admit(int i, const Query& q, bool warm, bool prune_ok) {
    SlotState& S = slots_[i];
    S.live = true;
    S.q = q;
    S.run_graph.reset();
    const DeviceGraph& G = *graphs_[q.graph];
    S.parent_n = G.g.n;
    Slab& sl = S.slab;
    SlotView& v = S.view;
    v = SlotView();
    v.g = G.g;
    v.res = sl.res; v.excess = sl.excess; v.h = sl.h; v.fa = sl.fa; v.fb = sl.fb;
    v.lock = sl.lock; v.aux = sl.aux; v.hl = sl.hl; v.tile_cnt = sl.tile_cnt; v.tile_hot = sl.tile_hot;
    v.tile_ex = sl.tile_ex; v.tsplit = sl.tsplit;
    v.lbkt = sl.lbkt;
    v.giant = sl.giant; v.gacc = sl.gacc; v.gcap = sl.gcap;
    v.emark = sl.emark;
    v.fbits = P_.bu_bits ? sl.fbits : nullptr;
    v.recv = sl.recv;
    v.recv_ok = (P_.bu_recv && !warm) ? 1 : 0;
    v.mlist = sl.mlist;
    v.mcap = std::max(0, std::min(sl.mcap, P_.mark_list));
    v.ctr = d_ctrs_ + i;
    v.s = q.s; v.t = q.t;
    v.H_cut = P_.h_cut;
    v.Kh = std::max(1, P_.K_heavy);
    v.inject = (P_.inject && !S.pinned && !warm) ? 1 : 0;
    v.push_first = P_.push_first ? 1 : 0;
    v.bu_light = std::max(1, std::min(P_.bu_light, kGiantDeg));
    const bool prune = P_.prune != 0 && prune_ok && !warm;
    S.sched.begin(&P_, G.g, q.s, q.t, prune, warm);
    S.sched.set_depth_hint(G.depth_hint.load(std::memory_order_relaxed));
    S.sched.stats.t_submit = q.t_arrive;
    S.sched.stats.t_start = now();
    v.K = S.sched.K();
    if (prune && G.classes) {
        // with the forward set of s's SCC at hand (or s and t in one SCC
        // under SCC routing) the prune needs no BFS
        if (intra_route(G, q) || G.fwd_find(G.scc[q.s])) return class_decide(i);
        S.sched.set_fwd_only();
    }
    return MF_OK;
}

// Pruning from reachability classes: mask = fwd(scc(s)) AND wcc == wcc(t)
// is the mask of BC:1391-1445 whenever t is reachable from s (then s and t
// are in it); when t is not reachable the flow is 0 (fwd(s) is a closed set
// holding s but not t: an s-t cut of capacity 0).
int Engine::fwd_set(DeviceGraph& G, int cls, const int32_t* d_h, std::shared_ptr<DeviceGraph::FwdSet>& out) {
    out = G.fwd_find(cls);
    if (out) return MF_OK;
    // d_h holds a BFS_FWD from a vertex of the class: h < n is its forward set
    auto fwd = std::make_shared<DeviceGraph::FwdSet>();
    const int n = G.g.n;
    HIPCHK(hipMalloc(&fwd->d_mask, std::max(n, 1)));
    launch_fwd_from_h(stream_, d_h, n, fwd->d_mask);
    fwd->mask.resize(n);
    if (n > 0) HIPCHK(hipMemcpyAsync(fwd->mask.data(), fwd->d_mask, n, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    int64_t cnt = 0;
    for (int v = 0; v < n; ++v) cnt += fwd->mask[v];
    fwd->count = cnt;
    // another engine may have cached the same class meanwhile: keep one
    out = G.fwd_insert(cls, fwd);
    return MF_OK;
}

int64_t Engine::class_kept(const DeviceGraph& G, const DeviceGraph::FwdSet& fwd, int wt) {
    int64_t kept = 0;
    for (int v = 0; v < G.g.n; ++v) kept += (fwd.mask[v] && G.wcc[v] == wt) ? 1 : 0;
    return kept;
}

int Engine::dbg_class_mask(int gid, int s, int t, int32_t* mask, int64_t* kept, int* reachable) {
    DeviceGraph& G = *graphs_[gid];
    if (!G.classes) { set_error("graph has no reachability classes"); return MF_EINVAL; }
    std::shared_ptr<DeviceGraph::FwdSet> fwd;
    int rc = fwd_set(G, G.scc[s], slots_[0].view.h, fwd);
    if (rc != MF_OK) return rc;
    *reachable = fwd->mask[t] ? 1 : 0;
    *kept = 0;
    if (!*reachable) return MF_OK;
    const int wt = G.wcc[t];
    *kept = class_kept(G, *fwd, wt);
    launch_class_mask(stream_, fwd->d_mask, G.d_wcc, wt, G.g.n, slots_[0].view.aux);
    HIPCHK(hipMemcpyAsync(mask, slots_[0].view.aux, sizeof(int32_t) * G.g.n, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    return MF_OK;
}

bool Engine::intra_route(const DeviceGraph& G, const Query& q) const {
    return P_.scc_route != 0 && G.scc[q.s] == G.scc[q.t];
}

std::string Engine::route_key(const DeviceGraph& G, const Query& q) const {
    if (intra_route(G, q)) return std::to_string(q.graph) + ":i" + std::to_string(G.scc[q.s]);
    return std::to_string(q.graph) + ":c" + std::to_string(G.scc[q.s]) + ":" + std::to_string(G.wcc[q.t]);
}

int Engine::class_decide(int i) {
    SlotState& S = slots_[i];
    DeviceGraph& G = *graphs_[S.q.graph];
    const int cls = G.scc[S.q.s];
    const bool intra = intra_route(G, S.q);
    std::shared_ptr<DeviceGraph::FwdSet> fwd;
    if (!intra) {
        // after S_PRUNE_FWD (when the class was not cached) the slot's h
        // holds the forward BFS from s
        int frc = fwd_set(G, cls, S.view.h, fwd);
        if (frc != MF_OK) return frc;
        if (!fwd->mask[S.q.t]) {
            S.sched.finish_trivial(0);
            return MF_OK;
        }
    }
    const int wt = G.wcc[S.q.t];
    const std::string key = route_key(G, S.q);
    std::shared_ptr<DeviceGraph> run;
    bool known = false;
    auto ct = cache_.find(key);
    const uint64_t ver = G.cap_version.load();
    if (ct != cache_.end() && ct->second.version != ver) {
        // built before a capacity update (by any engine sharing the graph)
        if (ct->second.g) cache_bytes_ -= ct->second.g->bytes;
        cache_.erase(ct);
        ct = cache_.end();
    }
    if (ct != cache_.end()) {
        run = ct->second.g;
        ct->second.last_use = ++use_clock_;
        known = true;
        ++cache_hits;
        S.sched.stats.kept = (int)ct->second.kept;
    }
    int64_t kept = 0;
    if (!known) {
        ++cache_misses;
        const int n = G.g.n;
        kept = intra ? G.scc_size[cls] : class_kept(G, *fwd, wt);
        S.sched.stats.kept = (int)kept;
        if (kept > 0 && (double)kept < P_.prune_ratio * (double)n) {
            // an SCC is inside its own forward set: the intra mask is the SCC
            if (intra) launch_eq_mask(stream_, G.d_scc, cls, n, S.view.aux);
            else launch_class_mask(stream_, fwd->d_mask, G.d_wcc, wt, n, S.view.aux);
            auto ng = std::make_shared<DeviceGraph>();
            int rc = build_induced_gpu(G, S.view.aux, (int)kept, stream_, *ng);
            if (rc != MF_OK) return rc;
            run = ng;
            cache_bytes_ += ng->bytes;
        }
        while (cache_bytes_ > cache_budget && !cache_.empty()) {
            auto lru = cache_.begin();
            for (auto jt = cache_.begin(); jt != cache_.end(); ++jt)
                if (jt->second.last_use < lru->second.last_use) lru = jt;
            if (lru->second.g) cache_bytes_ -= lru->second.g->bytes;
            cache_.erase(lru);
        }
        cache_[key] = CacheEntry{run, ++use_clock_, kept, ver};
    }
    // t in the largest SCC: the relabel's bottom-up levels skip vertices
    // that cannot reach it (no flow reaches them but from s, so they never
    // have a residual path to t); not with injection, where s is relabelled
    const bool skip = P_.bu_skip && !S.view.inject && G.giant_scc >= 0 && G.scc[S.q.t] == G.giant_scc;
    if (run) {
        const int s = run->old2sub_host.empty() ? -1 : run->old2sub_host[S.q.s];
        const int t = run->old2sub_host.empty() ? -1 : run->old2sub_host[S.q.t];
        if (s >= 0 && t >= 0) {
            S.run_graph = run;
            S.view.g = run->g;
            S.view.s = s;
            S.view.t = t;
            S.view.live = skip ? run->d_live : nullptr;
            S.sched.stats.pruned = 1;
            S.sched.resume_after_prune(run->g, s, t);
            S.sched.set_depth_hint(run->depth_hint.load(std::memory_order_relaxed));
            S.view.K = S.sched.K();
            return MF_OK;
        }
    }
    S.view.live = skip ? G.d_live : nullptr;
    S.sched.resume_after_prune(G.g, S.q.s, S.q.t);
    S.view.K = S.sched.K();
    return MF_OK;
}

// Prune decision (BC:2678-2698): keep the induced subgraph when the mask
// drops at least (1 - prune_ratio) of the vertices.
int Engine::prune_decide(int i) {
    SlotState& S = slots_[i];
    if (graphs_[S.q.graph]->classes) return class_decide(i);
    const DeviceGraph& G = *graphs_[S.q.graph];
    const Ctr& c = h_ctr_[i];
    const int kept = c.visited;
    S.sched.stats.kept = kept;
    if (kept > 0 && (double)kept < P_.prune_ratio * (double)G.g.n) {
        const unsigned long long hsh = h_hash_[i];
        const std::string key = std::to_string(S.q.graph) + ":" + std::to_string(hsh) + ":" + std::to_string(kept);
        std::shared_ptr<DeviceGraph> run;
        auto it = cache_.find(key);
        const uint64_t ver = G.cap_version.load();
        if (it != cache_.end() && it->second.version != ver) {
            cache_bytes_ -= it->second.g->bytes;
            cache_.erase(it);
            it = cache_.end();
        }
        if (it != cache_.end()) {
            run = it->second.g;
            it->second.last_use = ++use_clock_;
            ++cache_hits;
        } else {
            auto ng = std::make_shared<DeviceGraph>();
            int rc = build_induced_gpu(G, S.view.aux, kept, stream_, *ng);
            if (rc != MF_OK) return rc;
            // evict least recently used entries over budget
            cache_bytes_ += ng->bytes;
            while (cache_bytes_ > cache_budget && !cache_.empty()) {
                auto lru = cache_.begin();
                for (auto jt = cache_.begin(); jt != cache_.end(); ++jt)
                    if (jt->second.last_use < lru->second.last_use) lru = jt;
                cache_bytes_ -= lru->second.g->bytes;
                cache_.erase(lru);
            }
            cache_[key] = CacheEntry{ng, ++use_clock_, kept, ver};
            run = ng;
            ++cache_misses;
        }
        // map s, t into the run graph
        int ids[2] = {-1, -1};
        HIPCHK(hipMemcpyAsync(&ids[0], run->d_old2sub + S.q.s, sizeof(int), hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipMemcpyAsync(&ids[1], run->d_old2sub + S.q.t, sizeof(int), hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
        if (ids[0] >= 0 && ids[1] >= 0) {
            S.run_graph = run;
            S.view.g = run->g;
            S.view.s = ids[0];
            S.view.t = ids[1];
            S.sched.stats.pruned = 1;
            S.sched.resume_after_prune(run->g, ids[0], ids[1]);
            S.sched.set_depth_hint(run->depth_hint.load(std::memory_order_relaxed));
            S.view.K = S.sched.K();
            return MF_OK;
        }
    }
    S.sched.resume_after_prune(G.g, S.q.s, S.q.t);
    S.view.K = S.sched.K();
    return MF_OK;
}

int Engine::step(std::vector<Result>& out) {
    // refill free slots (continuous batching), first come first served, each
    // query admitted when its slab fits in the arena
    for (int i = 0; i < nslots_ && !queue_.empty(); ++i) {
        if (slots_[i].live || slots_[i].pinned) continue;
        const Query q = queue_.front();
        int n = 0, ntiles = 0;
        int64_t E = 0;
        run_dims(q, &n, &E, &ntiles);
        int rc = ensure_slab(i, n, E, ntiles);
        if (rc == kNoRoom) {
            bool busy = false;
            for (auto& S : slots_) busy = busy || S.live || S.pinned;
            if (busy) {  // wait for running queries to give bytes back
                ++admit_waits_;
                break;
            }
            queue_.pop_front();  // cannot fit even in an empty arena
            Result r;
            r.tag = q.tag;
            r.st.status = MF_ENOMEM;
            r.st.t_submit = q.t_arrive;
            r.st.t_start = r.st.t_end = now();
            set_error("query needs " + std::to_string(Slab::bytes_for(n, E, ntiles)) +
                      " slab bytes, more than the arena holds (" + std::to_string(arena_.size) + ")");
            out.push_back(r);
            --i;
            continue;
        }
        if (rc != MF_OK) return rc;
        queue_.pop_front();
        rc = admit(i, q);
        if (rc != MF_OK) return rc;
        memset(&h_ctr_[i], 0, sizeof(Ctr));
    }
    // plan
    std::vector<Plan> plans(nslots_);
    for (int i = 0; i < nslots_; ++i) {
        SlotState& S = slots_[i];
        if (!S.live) continue;
        if (S.sched.wants_prune_decision()) {
            int rc = prune_decide(i);
            if (rc != MF_OK) return rc;
        }
        plans[i] = S.sched.plan(h_ctr_[i]);
        if (S.sched.gr_depth() > 0) {
            DeviceGraph& RG = S.run_graph ? *S.run_graph : *graphs_[S.q.graph];
            RG.depth_hint.store(S.sched.gr_depth(), std::memory_order_relaxed);
        }
        S.view.phase = S.sched.phase();
        S.view.seed_t_only = plans[i].seed_t_only ? 1 : 0;
        S.view.cert_sside = plans[i].cert_sside ? 1 : 0;
        S.view.K = S.sched.K();
        S.view.disc_cap = S.sched.disc_cap();
        S.sched.stats.steps++;
    }
    // retire finished slots
    for (int i = 0; i < nslots_; ++i) {
        SlotState& S = slots_[i];
        if (!S.live || !S.sched.done()) continue;
        Result r;
        r.tag = S.q.tag;
        r.st = S.sched.stats;
        r.st.status = S.sched.status();
        bfs_restarts += r.st.bfs_restarts;
        const Ctr& c = h_ctr_[i];
        r.st.edge_scans = r.st.vertex_scans = r.st.pushes = r.st.relabels = 0;
        for (int f = 0; f < kFamilies; ++f) {
            for (int k = 0; k < 4; ++k) {
                r.st.wk[f][k] = c.wk[f][k];
                fam_work[f][k] += c.wk[f][k];
            }
            r.st.edge_scans += c.wk[f][0];
            r.st.vertex_scans += c.wk[f][1];
            r.st.pushes += c.wk[f][2];
            r.st.relabels += c.wk[f][3];
        }
        r.st.t_end = now();
        if (r.st.status != MF_OK) set_error(S.sched.error());
        out.push_back(r);
        S.live = false;
        S.run_graph.reset();
        plans[i] = Plan();
    }
    int rc = launch_plans(plans);
    if (rc < 0) return rc;
    if (step_fn) {
        for (int i = 0; i < nslots_; ++i) {
            const SlotState& S = slots_[i];
            const Plan& p = plans[i];
            if (!S.live || p.stages == 0) continue;
            mf_step_log L;
            memset(&L, 0, sizeof(L));
            L.tag = S.q.tag;
            L.slot = i;
            const unsigned s = p.stages;
            const bool gr_bfs = (s & ST_BFS) && p.bfs_mode == BFS_GR;
            L.steps = ((s & ST_INIT) ? MF_STEP_INIT : 0u) | ((s & ST_PREFLOW) ? MF_STEP_PREFLOW : 0u) |
                      ((s & ST_GR_SEED) ? MF_STEP_SEED : 0u) | ((gr_bfs || (s & ST_LOCAL)) ? MF_STEP_BFS : 0u) |
                      (((s & ST_BFS) && !gr_bfs) || (s & (ST_BFS_SEED1 | ST_MASK)) ? MF_STEP_PRUNE : 0u) |
                      ((s & ST_DISCHARGE) ? MF_STEP_DISCHARGE : 0u) | ((s & ST_HOT) ? MF_STEP_HOT : 0u) |
                      ((s & ST_RM) ? MF_STEP_RM : 0u) | ((s & ST_CERT) ? MF_STEP_CERT : 0u) |
                      ((s & ST_ACTIVE) ? MF_STEP_COUNT : 0u);
            L.iters = S.sched.stats.iters;
            L.active = h_ctr_[i].active;
            L.active_high = h_ctr_[i].active_high;
            L.fixed_edges = (int64_t)h_ctr_[i].fixed_edges;
            L.K = S.sched.K();
            L.gr_freq = S.sched.gr_freq();
            L.path = (s & ST_DISCHARGE) && p.dis_kind == DK_PATH ? 1 : 0;
            L.cycles = (s & ST_DISCHARGE) ? std::max(1, std::min(p.cycles, std::max(1, P_.fuse))) : 0;
            for (int f = 0; f < MF_LOG_FIELDS; ++f) L.ms[f] = step_ms_[f];
            step_fn(step_user, &L);
        }
    }
    return live();
}

// Stage every live slot's view, build one launch list per stage group, issue
// the groups in stage order on the engine stream, copy the counters back and
// wait.  One host synchronisation per call.
int Engine::launch_plans(std::vector<Plan>& plans) {
    // stage the views and lists
    SlotView* hv = (SlotView*)h_stage_;
    int* hl = (int*)((char*)h_stage_ + sizeof(SlotView) * nslots_);
    for (int i = 0; i < nslots_; ++i) {
        hv[i] = slots_[i].view;
        if (!P_.bu_pick) hv[i].g.bu_pick = nullptr;
        hv[i].marks_ok = slots_[i].slab.marks_ok ? 1 : 0;
        hv[i].bfs_edges = plans[i].small_edges > 0 ? plans[i].small_edges : P_.small_edges;
    }
    size_t lpos = 0;
    struct Group { int off = 0, cnt = 0; };
    auto make_list = [&](auto pred) {
        Group g;
        g.off = (int)lpos;
        for (int i = 0; i < nslots_; ++i)
            if (slots_[i].live && pred(i)) {
                if (lpos < list_cap_) hl[lpos] = i;  // overflow is reported below
                ++lpos;
            }
        g.cnt = (int)lpos - g.off;
        return g;
    };
    // Slots whose step is only a single-workgroup BFS (a GR or pruning BFS
    // under way) run on the side stream: a few workgroups for many levels,
    // overlapped with the other slots' bandwidth-bound kernels.
    std::vector<char> side(nslots_, 0);
    {
        const unsigned bfs_only = ST_RESET_PRE | ST_GR_SEED | ST_BFS_SEED1 | ST_BFS;
        int nside = 0, nmain = 0;
        for (int i = 0; i < nslots_; ++i) {
            if (!slots_[i].live || !plans[i].stages) continue;
            const Plan& p = plans[i];
            side[i] = (p.stages & ~bfs_only) == 0 && (p.stages & ST_BFS) && p.bfs_kind == BK_SMALL;
            nside += side[i];
            nmain += !side[i];
        }
        if (nside == 0 || nmain == 0) std::fill(side.begin(), side.end(), 0);
    }
    auto has = [&](int i, unsigned bit) { return !side[i] && (plans[i].stages & bit) != 0; };
    auto hs = [&](int i, unsigned bit) { return side[i] && (plans[i].stages & bit) != 0; };
    std::vector<std::pair<unsigned, Group>> side_pre;
    Group s_seed, s_seed1[3], s_bfs[3][2];  // [mode][multi-workgroup]
    {
        std::vector<unsigned> vals;
        for (int i = 0; i < nslots_; ++i)
            if (slots_[i].live && hs(i, ST_RESET_PRE) &&
                std::find(vals.begin(), vals.end(), plans[i].reset_pre) == vals.end())
                vals.push_back(plans[i].reset_pre);
        for (unsigned w : vals)
            side_pre.push_back({w, make_list([&](int i) { return hs(i, ST_RESET_PRE) && plans[i].reset_pre == w; })});
        s_seed = make_list([&](int i) { return hs(i, ST_GR_SEED); });
        for (int m = 0; m < 3; ++m) {
            s_seed1[m] = make_list([&](int i) { return hs(i, ST_BFS_SEED1) && plans[i].seed_mode == m; });
            for (int w = 0; w < 2; ++w)
                s_bfs[m][w] = make_list([&](int i) { return hs(i, ST_BFS) && plans[i].bfs_mode == m && (plans[i].small_wgs > 1) == (w == 1); });
        }
    }
    // reset_pre grouped by distinct mask values
    std::vector<std::pair<unsigned, Group>> pre_groups;
    {
        std::vector<unsigned> vals;
        for (int i = 0; i < nslots_; ++i)
            if (slots_[i].live && has(i, ST_RESET_PRE) &&
                std::find(vals.begin(), vals.end(), plans[i].reset_pre) == vals.end())
                vals.push_back(plans[i].reset_pre);
        for (unsigned w : vals)
            pre_groups.push_back({w, make_list([&](int i) { return has(i, ST_RESET_PRE) && plans[i].reset_pre == w; })});
    }
    // a slab still holding a residual of the run graph resets marked pairs only
    // (same topology and capacity version; engines sharing a graph see each
    // other's capacity updates through the version)
    auto run_g = [&](int i) { return slots_[i].run_graph ? slots_[i].run_graph.get() : graphs_[slots_[i].q.graph].get(); };
    auto marked_ok = [&](int i) {
        const Slab& sl = slots_[i].slab;
        return sl.res_uid == run_g(i)->uid && sl.res_version == run_g(i)->cap_version.load();
    };
    Group g_init = make_list([&](int i) { return has(i, ST_INIT) && !marked_ok(i); });
    Group g_init_m = make_list([&](int i) { return has(i, ST_INIT) && marked_ok(i); });
    for (int i = 0; i < nslots_; ++i)
        if (slots_[i].live && has(i, ST_INIT)) {
            slots_[i].slab.res_uid = run_g(i)->uid;
            slots_[i].slab.res_version = run_g(i)->cap_version.load();
            slots_[i].slab.marks_ok = true;
        }
    Group g_pre = make_list([&](int i) { return has(i, ST_PREFLOW); });
    Group g_seed = make_list([&](int i) { return has(i, ST_GR_SEED); });
    Group g_seed1[3];
    for (int m = 0; m < 3; ++m) g_seed1[m] = make_list([&](int i) { return has(i, ST_BFS_SEED1) && plans[i].seed_mode == m; });
    Group g_bfs[3][3], g_bfs_multi[3];  // [kind][mode]; small-frontier slots with several workgroups apart
    for (int k = 0; k < 3; ++k)
        for (int m = 0; m < 3; ++m)
            g_bfs[k][m] = make_list([&](int i) {
                return has(i, ST_BFS) && plans[i].bfs_kind == k && plans[i].bfs_mode == m && (k != BK_SMALL || plans[i].small_wgs <= 1);
            });
    Group g_bubits[3];  // first bottom-up level of a BFS: level L's bitmap from the levels
    for (int m = 0; m < 3; ++m)
        g_bubits[m] = make_list([&](int i) {
            return has(i, ST_BFS) && plans[i].bfs_kind == BK_BU && plans[i].bfs_mode == m && plans[i].bu_bits && P_.bu_bits;
        });
    for (int m = 0; m < 3; ++m)
        g_bfs_multi[m] = make_list([&](int i) {
            return has(i, ST_BFS) && plans[i].bfs_kind == BK_SMALL && plans[i].bfs_mode == m && plans[i].small_wgs > 1;
        });
    // local relabels, one group per distinct level cap
    std::vector<std::pair<int, Group>> local_groups;
    {
        std::vector<int> caps;
        for (int i = 0; i < nslots_; ++i)
            if (slots_[i].live && has(i, ST_LOCAL) &&
                std::find(caps.begin(), caps.end(), plans[i].local_cap) == caps.end())
                caps.push_back(plans[i].local_cap);
        for (int c : caps)
            local_groups.push_back({c, make_list([&](int i) { return has(i, ST_LOCAL) && plans[i].local_cap == c; })});
    }
    Group g_mask = make_list([&](int i) { return has(i, ST_MASK); });
    Group g_dis[3];
    for (int k = 0; k < 3; ++k) g_dis[k] = make_list([&](int i) { return has(i, ST_DISCHARGE) && plans[i].dis_kind == k; });
    Group g_hot = make_list([&](int i) { return has(i, ST_HOT); });
    Group g_rm = make_list([&](int i) { return has(i, ST_RM); });
    std::vector<std::pair<unsigned, Group>> act_groups;
    {
        std::vector<unsigned> vals;
        for (int i = 0; i < nslots_; ++i)
            if (slots_[i].live && has(i, ST_RESET_ACT) &&
                std::find(vals.begin(), vals.end(), plans[i].reset_act) == vals.end())
                vals.push_back(plans[i].reset_act);
        for (unsigned w : vals)
            act_groups.push_back({w, make_list([&](int i) { return has(i, ST_RESET_ACT) && plans[i].reset_act == w; })});
    }
    Group g_act = make_list([&](int i) { return has(i, ST_ACTIVE) && !plans[i].sparse; });
    Group g_act_sp = make_list([&](int i) { return has(i, ST_ACTIVE) && plans[i].sparse; });
    // the sparse certificate where the slab's marks cover its residual, the
    // dense fingerprint elsewhere
    const bool cert_sp = P_.cert == 3;
    Group g_sink = make_list([&](int i) { return has(i, ST_SINK); });
    Group g_src = make_list([&](int i) { return has(i, ST_SOURCE); });
    Group g_cert = make_list([&](int i) { return has(i, ST_CERT) && !(cert_sp && slots_[i].slab.marks_ok); });
    Group g_cert_sp = make_list([&](int i) { return has(i, ST_CERT) && cert_sp && slots_[i].slab.marks_ok; });
    // slots of fused steps (see the fused-cycle launches below)
    int max_cycles = 1;
    for (int i = 0; i < nslots_; ++i)
        if (slots_[i].live && has(i, ST_DISCHARGE)) max_cycles = std::max(max_cycles, std::min(plans[i].cycles, std::max(1, P_.fuse)));
    max_cycles = std::min(max_cycles, 64);
    // the list discharge's device work queue {next = 0, total, prefix of the
    // slots' blocks}, uploaded with the lists: the blocks are the ones the
    // slots' last counts give (Plan::dis_grid)
    size_t dq_off = 0;
    int64_t dq_total = 0;
    if (P_.dis_queue && g_dis[DK_LIST].cnt) {
        dq_off = lpos;
        const int cnt = g_dis[DK_LIST].cnt;
        if (lpos + 3 + (size_t)cnt <= list_cap_) {
            for (int q = 0; q < cnt; ++q) {
                hl[lpos + 2 + q] = (int)dq_total;
                dq_total += plans[hl[g_dis[DK_LIST].off + q]].dis_grid;
            }
            hl[lpos] = 0;
            hl[lpos + 1] = (int)dq_total;
            hl[lpos + 2 + cnt] = (int)dq_total;
        }
        lpos += 3 + (size_t)cnt;
    }
    if (lpos > list_cap_) { set_error("launch list overflow"); return MF_EINVAL; }
    const int nlive = live();
    if (nlive == 0) return 0;
    HIPCHK(hipMemcpyAsync(d_views_, h_stage_, sizeof(SlotView) * nslots_ + sizeof(int) * lpos,
                          hipMemcpyHostToDevice, stream_));
    // the launch lists sit right after the views
    int* d_list_base = (int*)((char*)d_views_ + sizeof(SlotView) * nslots_);
    auto L = [&](const Group& g) { return d_list_base + g.off; };
    auto maxof = [&](const Group& g, auto f) {
        int64_t m = 1;
        for (int k = 0; k < g.cnt; ++k) m = std::max<int64_t>(m, f(hl[g.off + k]));
        return (int)std::min<int64_t>(m, INT32_MAX);
    };
    hipStream_t st = stream_;
    // profiling: an event pair around each kernel family's launches
    struct Timed { int fam, k, bucket; };
    std::vector<Timed> timed;  // (family, event pair index, step-log bucket)
    auto bucket_of = [](int fam) {
        switch (fam) {
        case WF_INIT: return (int)MF_LOG_INIT;
        case WF_ACTIVE: case WF_DISCHARGE: return (int)MF_LOG_PR;
        case WF_RM: return (int)MF_LOG_RM;
        case WF_CERT: return (int)MF_LOG_CERT;
        default: return (int)MF_LOG_BFS;
        }
    };
    auto tbs = [&](int fam, hipStream_t s, int bucket = -1) {
        if (!profile) return -1;
        const int k = (int)timed.size();
        if (2 * k + 1 >= (int)ev_.size()) return -1;
        (void)hipEventRecord(ev_[2 * k], s);
        timed.push_back({fam, k, bucket >= 0 ? bucket : bucket_of(fam)});
        return k;
    };
    auto tes = [&](int k, hipStream_t s) {
        if (k >= 0) (void)hipEventRecord(ev_[2 * k + 1], s);
    };
    auto tb = [&](int fam, int bucket = -1) { return tbs(fam, st, bucket); };
    auto te = [&](int k) { tes(k, st); };
    auto gridN = [&](int i) { return clampi(cdiv(slots_[i].view.g.n, 256), 1, 4096); };
    // vertex- and entry-parallel kernels are grid-stride: at most 4096
    // workgroups per slot and about 16K per launch over all its slots
    // (several times what the device holds at once), so each workgroup's
    // fixed costs (staging, block reductions, barriers) spread over several
    // chunks
    const int kLaunchBlocks = std::max(64, P_.launch_blocks), kSlotBlocks = std::max(16, P_.slot_blocks);
    auto fit = [&](const Group& g, int gx) {
        const int cap = std::min(kSlotBlocks, std::max(64, kLaunchBlocks / std::max(1, g.cnt)));
        return std::max(1, std::min(gx, cap));
    };
    auto gridE = [&](int i) { return clampi(cdiv(slots_[i].view.g.E, 1024), 1, 8192); };
    // giant rows in the group: 0 = none, 1 = discharged in whole chunks,
    // 2 = in rounds (Params::giant_rounds)
    const int giant_mode = (P_.giant_rounds == 2 ? nslots_ > 1 : P_.giant_rounds != 0) ? 2 : 1;
    auto giants = [&](const Group& g) {
        for (int k = 0; k < g.cnt; ++k)
            if (slots_[hl[g.off + k]].view.g.max_deg > kGiantDeg) return giant_mode;
        return 0;
    };
    // side stream: starts after the view / list upload on the main stream
    bool side_bfs = false;
    for (int m = 0; m < 3; ++m) side_bfs = side_bfs || s_bfs[m][0].cnt || s_bfs[m][1].cnt;
    const bool use_side = !side_pre.empty() || s_seed.cnt || side_bfs;
    if (use_side) {
        HIPCHK(hipEventRecord(ev_up_, st));
        HIPCHK(hipStreamWaitEvent(side_, ev_up_, 0));
        const int k = tbs(WF_BFS, side_);
        for (auto& pg : side_pre) launch_reset(side_, d_views_, L(pg.second), pg.second.cnt, pg.first);
        if (s_seed.cnt) launch_gr_seed(side_, d_views_, L(s_seed), s_seed.cnt, fit(s_seed, maxof(s_seed, gridN)));
        for (int m = 0; m < 3; ++m)
            if (s_seed1[m].cnt) launch_bfs_seed1(side_, d_views_, L(s_seed1[m]), s_seed1[m].cnt, fit(s_seed1[m], maxof(s_seed1[m], gridN)), m);
        for (int m = 0; m < 3; ++m)
            for (int w = 0; w < 2; ++w)
                if (s_bfs[m][w].cnt)
                    launch_bfs_small(side_, d_views_, L(s_bfs[m][w]), s_bfs[m][w].cnt, m, P_.small_levels, w ? P_.small_wgs : 1, spin_ticks_, multi_wg_cap());
        tes(k, side_);
        HIPCHK(hipEventRecord(ev_side_, side_));
    }
    for (auto& pg : pre_groups) launch_reset(st, d_views_, L(pg.second), pg.second.cnt, pg.first);
    if (g_init.cnt) {
        const int k = tb(WF_INIT);
        launch_init(st, d_views_, L(g_init), g_init.cnt, fit(g_init, maxof(g_init, [&](int i) { return std::max(gridE(i), gridN(i)); })));
        te(k);
    }
    if (g_init_m.cnt) {
        const int k = tb(WF_INIT);
        launch_init_marked(st, d_views_, L(g_init_m), g_init_m.cnt, fit(g_init_m, maxof(g_init_m, gridN)));
        te(k);
    }
    if (g_pre.cnt) { const int k = tb(WF_INIT, MF_LOG_PREFLOW); launch_preflow(st, d_views_, L(g_pre), g_pre.cnt); te(k); }
    if (g_seed.cnt) { const int k = tb(WF_BFS, MF_LOG_SEED); launch_gr_seed(st, d_views_, L(g_seed), g_seed.cnt, fit(g_seed, maxof(g_seed, gridN))); te(k); }
    if (!local_groups.empty()) {
        const int k = tb(WF_BFS);
        for (auto& lg : local_groups) {
            const Group& g = lg.second;
            const int gx = maxof(g, gridN);
            launch_local_mark(st, d_views_, L(g), g.cnt, gx, lg.first);
            launch_local_bound(st, d_views_, L(g), g.cnt, gx, lg.first);
            launch_local_bucket(st, d_views_, L(g), g.cnt, gx, lg.first);
            launch_local_bfs(st, d_views_, L(g), g.cnt, lg.first);
            launch_local_finish(st, d_views_, L(g), g.cnt, gx);
        }
        te(k);
    }
    for (int m = 0; m < 3; ++m)
        if (g_seed1[m].cnt) {
            const int k = tb(WF_BFS);
            launch_bfs_seed1(st, d_views_, L(g_seed1[m]), g_seed1[m].cnt, fit(g_seed1[m], maxof(g_seed1[m], gridN)), m);
            te(k);
        }
    for (int m = 0; m < 3; ++m) {
        if (g_bfs[BK_SMALL][m].cnt) {
            const int k = tb(WF_BFS);
            launch_bfs_small(st, d_views_, L(g_bfs[BK_SMALL][m]), g_bfs[BK_SMALL][m].cnt, m, P_.small_levels, 1, spin_ticks_, multi_wg_cap());
            te(k);
        }
        if (g_bfs_multi[m].cnt) {
            const int k = tb(WF_BFS);
            launch_bfs_small(st, d_views_, L(g_bfs_multi[m]), g_bfs_multi[m].cnt, m, P_.small_levels, P_.small_wgs, spin_ticks_, multi_wg_cap());
            te(k);
        }
        if (g_bfs[BK_TD][m].cnt) {
            const int k = tb(WF_BFS);
            launch_bfs_td(st, d_views_, L(g_bfs[BK_TD][m]), g_bfs[BK_TD][m].cnt, fit(g_bfs[BK_TD][m], maxof(g_bfs[BK_TD][m], [&](int i) { return plans[i].bfs_grid; })), m);
            te(k);
        }
        if (g_bfs[BK_BU][m].cnt) {
            const int k = tb(WF_BFS_BU);
            if (g_bubits[m].cnt)
                launch_bu_bits(st, d_views_, L(g_bubits[m]), g_bubits[m].cnt, fit(g_bubits[m], maxof(g_bubits[m], [&](int i) { return plans[i].bfs_grid; })), m);
            launch_bfs_bu(st, d_views_, L(g_bfs[BK_BU][m]), g_bfs[BK_BU][m].cnt, fit(g_bfs[BK_BU][m], maxof(g_bfs[BK_BU][m], [&](int i) { return plans[i].bfs_grid; })), m);
            te(k);
        }
    }
    if (g_mask.cnt) {
        HIPCHK(hipMemsetAsync(d_hash_, 0, sizeof(unsigned long long) * nslots_, st));
        const int k = tb(WF_BFS);
        launch_mask_combine(st, d_views_, L(g_mask), g_mask.cnt, fit(g_mask, maxof(g_mask, gridN)), d_hash_);
        te(k);
    }
    if (g_dis[DK_TILE].cnt) {
        const int k = tb(WF_DISCHARGE);
        launch_discharge_tile(st, d_views_, L(g_dis[DK_TILE]), g_dis[DK_TILE].cnt, maxof(g_dis[DK_TILE], [&](int i) { return plans[i].dis_grid; }), 0, P_.tile_rounds, giants(g_dis[DK_TILE]), P_.tile_threads, P_.tile_split);
        te(k);
    }
    if (g_dis[DK_LIST].cnt) {
        const int k = tb(WF_DISCHARGE);
        if (P_.dis_queue) {
            // one persistent launch: as many workgroups as the slots' blocks,
            // at most the per-launch cap
            const int wg = (int)std::max<int64_t>(1, std::min<int64_t>(dq_total, kLaunchBlocks));
            launch_discharge_queue(st, d_views_, L(g_dis[DK_LIST]), g_dis[DK_LIST].cnt, d_list_base + dq_off, wg, giants(g_dis[DK_LIST]));
        } else {
            launch_discharge_list(st, d_views_, L(g_dis[DK_LIST]), g_dis[DK_LIST].cnt, fit(g_dis[DK_LIST], maxof(g_dis[DK_LIST], [&](int i) { return plans[i].dis_grid; })), giants(g_dis[DK_LIST]));
        }
        te(k);
    }
    if (g_dis[DK_PATH].cnt) {
        const int k = tb(WF_DISCHARGE);
        launch_path_push(st, d_views_, L(g_dis[DK_PATH]), g_dis[DK_PATH].cnt, maxof(g_dis[DK_PATH], [&](int i) { return plans[i].dis_grid; }), P_.path_len * P_.path_attempts);
        te(k);
    }
    if (g_hot.cnt) {
        const int k = tb(WF_DISCHARGE);
        launch_hot_select(st, d_views_, L(g_hot), g_hot.cnt, P_.hot_k);
        launch_discharge_tile(st, d_views_, L(g_hot), g_hot.cnt, maxof(g_hot, [&](int i) { return slots_[i].view.g.ntiles; }), 1, P_.hot_rounds, 0, P_.tile_threads, P_.tile_split);
        te(k);
    }
    if (g_rm.cnt) {
        const int k = tb(WF_RM);
        launch_rm(st, d_views_, L(g_rm), g_rm.cnt, maxof(g_rm, [&](int i) { return slots_[i].view.g.ntiles; }), P_.rm_margin);
        te(k);
    }
    for (auto& ag : act_groups) launch_reset(st, d_views_, L(ag.second), ag.second.cnt, ag.first);
    if (g_act.cnt) { const int k = tb(WF_ACTIVE); launch_active(st, d_views_, L(g_act), g_act.cnt, fit(g_act, maxof(g_act, gridN))); te(k); }
    if (g_act_sp.cnt) { const int k = tb(WF_ACTIVE); launch_active_sparse(st, d_views_, L(g_act_sp), g_act_sp.cnt, 512); te(k); }
    // further discharge + count cycles of fused steps (deep graphs): the
    // count after each only looks at the previous list and the vertices the
    // discharge pushed into.  The slot lists sit at fixed places of cyc_d_
    // and the grids are fixed (the kernels are grid-stride), so a step whose
    // fused slots repeat replays a captured graph of all its cycles.
    if (max_cycles > 1) {
        const int ns = nslots_;
        std::vector<int32_t> key;
        key.push_back(max_cycles);
        int tiles = 1, gi = 0;
        int cnt[64][3] = {};
        for (int c = 1; c < max_cycles; ++c)
            for (int kind = 0; kind < 3; ++kind) {
                int32_t* lst = cyc_h_ + ((size_t)(c - 1) * 3 + kind) * ns;
                for (int i = 0; i < ns; ++i) {
                    if (!slots_[i].live || !has(i, ST_DISCHARGE) || plans[i].cycles <= c) continue;
                    if (kind < 2 && plans[i].dis_kind != kind) continue;
                    lst[cnt[c][kind]++] = i;
                    tiles = std::max(tiles, slots_[i].view.g.ntiles);
                    if (slots_[i].view.g.max_deg > kGiantDeg) gi = giant_mode;
                }
                key.push_back(cnt[c][kind]);
                key.insert(key.end(), lst, lst + cnt[c][kind]);
            }
        key.push_back(tiles);
        key.push_back(gi);
        HIPCHK(hipMemcpyAsync(cyc_d_, cyc_h_, sizeof(int32_t) * (size_t)(max_cycles - 1) * 3 * ns,
                              hipMemcpyHostToDevice, st));
        auto LC = [&](int c, int kind) { return cyc_d_ + ((size_t)(c - 1) * 3 + kind) * ns; };
        auto emit = [&](bool timed) {
            for (int c = 1; c < max_cycles; ++c) {
                launch_reset(st, d_views_, LC(c, 2), cnt[c][2], RST_TOUCH);
                if (cnt[c][DK_TILE]) {
                    const int k = timed ? tb(WF_DISCHARGE) : -1;
                    launch_discharge_tile(st, d_views_, LC(c, DK_TILE), cnt[c][DK_TILE], tiles, 0, P_.tile_rounds, gi, P_.tile_threads, P_.tile_split);
                    te(k);
                }
                if (cnt[c][DK_LIST]) {
                    const int k = timed ? tb(WF_DISCHARGE) : -1;
                    launch_discharge_list(st, d_views_, LC(c, DK_LIST), cnt[c][DK_LIST], 1024, gi);
                    te(k);
                }
                launch_reset(st, d_views_, LC(c, 2), cnt[c][2], RST_SPARSE);
                const int k = timed ? tb(WF_ACTIVE) : -1;
                launch_active_sparse(st, d_views_, LC(c, 2), cnt[c][2], 512);
                te(k);
            }
        };
        auto it = graph_cache_.end();
        if (P_.graphs && !profile) it = graph_cache_.find(key);
        if (it != graph_cache_.end()) {
            it->second.last = ++graph_clock_;
            HIPCHK(hipGraphLaunch(it->second.x, st));
            ++graph_replays;
        } else if (P_.graphs && !profile && graph_seen_.count(key)) {
            // second time this set of fused slots comes round: capture
            GraphEntry ge;
            HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
            emit(false);
            const hipError_t ce = hipStreamEndCapture(st, &ge.g);
            if (ce != hipSuccess) { set_error(std::string("graph capture: ") + hipGetErrorString(ce)); return MF_EHIP; }
            HIPCHK(hipGraphInstantiate(&ge.x, ge.g, nullptr, nullptr, 0));
            ge.last = ++graph_clock_;
            if (graph_cache_.size() >= 32) {
                auto lru = graph_cache_.begin();
                for (auto jt = graph_cache_.begin(); jt != graph_cache_.end(); ++jt)
                    if (jt->second.last < lru->second.last) lru = jt;
                (void)hipGraphExecDestroy(lru->second.x);
                (void)hipGraphDestroy(lru->second.g);
                graph_cache_.erase(lru);
            }
            auto ins = graph_cache_.emplace(key, ge).first;
            graph_seen_.erase(key);
            HIPCHK(hipGraphLaunch(ins->second.x, st));
            ++graph_captures;
        } else {
            if (P_.graphs && !profile) {
                if (graph_seen_.size() > 256) graph_seen_.clear();
                graph_seen_.insert(key);
            }
            emit(true);
        }
    }
    {
        if (g_sink.cnt || g_src.cnt) {
            const int k = tb(WF_CERT);
            if (g_sink.cnt) launch_cut_labels(st, d_views_, L(g_sink), g_sink.cnt, fit(g_sink, maxof(g_sink, gridN)), 0);
            if (g_src.cnt) launch_cut_labels(st, d_views_, L(g_src), g_src.cnt, fit(g_src, maxof(g_src, gridN)), 1);
            te(k);
        }
    }
    if (g_cert.cnt) { const int k = tb(WF_CERT); launch_cert(st, d_views_, L(g_cert), g_cert.cnt, fit(g_cert, maxof(g_cert, gridN)), cert_sp ? 1 : P_.cert); te(k); }
    if (g_cert_sp.cnt) { const int k = tb(WF_CERT); launch_cert(st, d_views_, L(g_cert_sp), g_cert_sp.cnt, fit(g_cert_sp, maxof(g_cert_sp, gridN)), 3); te(k); }
    HIPCHK(hipGetLastError());
    if (use_side) HIPCHK(hipStreamWaitEvent(st, ev_side_, 0));
    HIPCHK(hipMemcpyAsync(h_ctr_, d_ctrs_, sizeof(Ctr) * nslots_, hipMemcpyDeviceToHost, st));
    if (g_mask.cnt)
        HIPCHK(hipMemcpyAsync(h_hash_, d_hash_, sizeof(unsigned long long) * nslots_, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (double& x : step_ms_) x = 0.0;
    for (auto& tm : timed) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ev_[2 * tm.k], ev_[2 * tm.k + 1]) == hipSuccess) {
            fam_ms[tm.fam] += ms;
            fam_launches[tm.fam] += 1;
            step_ms_[tm.bucket] += ms;
        }
    }
    return MF_OK;
}

void Engine::prof_reset() {
    for (int f = 0; f < kFamilies; ++f) {
        fam_ms[f] = 0;
        fam_launches[f] = 0;
        for (int k = 0; k < 4; ++k) fam_work[f][k] = 0;
    }
}

int Engine::pin(int slot, const Query& q) {
    if (slot < 0 || slot >= nslots_) { set_error("bad slot"); return MF_EINVAL; }
    if (q.graph < 0 || q.graph >= (int)graphs_.size()) { set_error("bad graph index"); return MF_EINVAL; }
    if (slots_[slot].live) { set_error("slot busy"); return MF_EBUSY; }
    const int n = graphs_[q.graph]->g.n;
    if (q.s < 0 || q.t < 0 || q.s >= n || q.t >= n || q.s == q.t) { set_error("bad query"); return MF_EINVAL; }
    const DGraph& g = graphs_[q.graph]->g;
    int rc = ensure_slab(slot, g.n, g.E, g.ntiles);  // pinned queries run on the full graph
    if (rc == kNoRoom) { set_error("pin: no room in the slab arena"); return MF_ENOMEM; }
    if (rc != MF_OK) return rc;
    slots_[slot].pinned = true;
    slots_[slot].solved = false;
    slots_[slot].q = q;
    return MF_OK;
}

int Engine::solve_pinned(bool warm, std::vector<Result>& out) {
    std::vector<int> mine;
    for (int i = 0; i < nslots_; ++i) {
        SlotState& S = slots_[i];
        if (!S.pinned) continue;
        const bool w = warm && S.solved;
        int rc = admit(i, S.q, w, false);
        if (rc != MF_OK) return rc;
        memset(&h_ctr_[i], 0, sizeof(Ctr));
        mine.push_back(i);
    }
    std::vector<Result> got;
    std::vector<char> retried(nslots_, 0);
    while (true) {
        bool any = false;
        for (int i : mine) any = any || slots_[i].live;
        if (!any) break;
        const size_t before = got.size();
        int rc = step(got);
        if (rc < 0) return rc;
        // a warm answer that fails its certificate is re-solved from scratch
        for (size_t k = before; k < got.size(); ++k) {
            Result& r = got[k];
            int i = -1;
            for (int j : mine) if (slots_[j].q.tag == r.tag && !slots_[j].live) i = j;
            if (i >= 0 && r.st.status != MF_OK && warm && !retried[i]) {
                retried[i] = 1;
                int rc2 = admit(i, slots_[i].q, false, false);
                if (rc2 != MF_OK) return rc2;
                memset(&h_ctr_[i], 0, sizeof(Ctr));
                r.tag = INT64_MIN;  // drop this attempt
            }
        }
    }
    for (auto& r : got) {
        if (r.tag == INT64_MIN) continue;
        for (int j : mine)
            if (slots_[j].q.tag == r.tag) {
                slots_[j].solved = (r.st.status == MF_OK);
                if (retried[j]) r.st.drains = -1 - r.st.drains;
            }
        out.push_back(r);
    }
    return MF_OK;
}

int Engine::update_caps(int gid, const int32_t* entries, const int64_t* new_cap, int n) {
    if (gid < 0 || gid >= (int)graphs_.size()) { set_error("bad graph index"); return MF_EINVAL; }
    if (n <= 0) return MF_OK;
    DeviceGraph& G = *graphs_[gid];
    // Only forward entries (cap > 0) carry a capacity: a reverse entry would
    // change the topology the reachability classes and cached forward sets
    // were computed on.  Entries must not repeat (one scatter per entry).
    // Nothing is changed unless the whole batch is valid.
    {
        std::vector<uint64_t> seen((size_t)(G.g.E + 63) / 64, 0);
        for (int i = 0; i < n; ++i) {
            const int e = entries[i];
            if (e < 0 || e >= G.g.E) { set_error("update entry out of range"); return MF_EINVAL; }
            if (new_cap[i] < 1) { set_error("updated capacity must be >= 1"); return MF_EINVAL; }
            const uint64_t bit = 1ull << (e & 63);
            if (!G.fwd_bits.empty() && !(G.fwd_bits[(size_t)e >> 6] & bit)) {
                set_error("update entry " + std::to_string(e) + " is a reverse entry (capacity 0)");
                return MF_EINVAL;
            }
            if (seen[(size_t)e >> 6] & bit) {
                set_error("update entry " + std::to_string(e) + " repeats");
                return MF_EINVAL;
            }
            seen[(size_t)e >> 6] |= bit;
        }
    }
    int32_t* d_e = nullptr;
    int64_t *d_new = nullptr, *d_old = nullptr;
    HIPCHK(hipMalloc(&d_e, sizeof(int32_t) * n));
    HIPCHK(hipMalloc(&d_new, sizeof(int64_t) * n));
    HIPCHK(hipMalloc(&d_old, sizeof(int64_t) * n));
    HIPCHK(hipMemcpyAsync(d_e, entries, sizeof(int32_t) * n, hipMemcpyHostToDevice, stream_));
    HIPCHK(hipMemcpyAsync(d_new, new_cap, sizeof(int64_t) * n, hipMemcpyHostToDevice, stream_));
    launch_cap_scatter(stream_, G.g, G.d_cap, d_e, d_new, d_old, n);
    // adjust every pinned residual on this graph
    SlotView* hv = (SlotView*)h_stage_;
    int* hl = (int*)((char*)h_stage_ + sizeof(SlotView) * nslots_);
    int cnt = 0;
    for (int i = 0; i < nslots_; ++i) {
        hv[i] = slots_[i].view;
        if (slots_[i].pinned && slots_[i].solved && slots_[i].q.graph == gid) hl[cnt++] = i;
    }
    if (cnt) {
        HIPCHK(hipMemcpyAsync(d_views_, h_stage_, sizeof(SlotView) * nslots_ + sizeof(int) * cnt,
                              hipMemcpyHostToDevice, stream_));
        int* d_list = (int*)((char*)d_views_ + sizeof(SlotView) * nslots_);
        launch_apply_updates(stream_, d_views_, d_list, cnt, d_e, d_old, d_new, n);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(stream_));
    (void)hipFree(d_e); (void)hipFree(d_new); (void)hipFree(d_old);
    // residuals of this graph kept elsewhere predate the new capacities: the
    // next query on it (in any engine) resets every entry; pinned residuals
    // were adjusted above and stay current
    const uint64_t ver = ++G.cap_version;
    for (int i = 0; i < nslots_; ++i)
        if (slots_[i].pinned && slots_[i].solved && slots_[i].q.graph == gid && slots_[i].slab.res_uid == G.uid)
            slots_[i].slab.res_version = ver;
    // cached induced subgraphs of this graph carry old capacities
    const std::string pre = std::to_string(gid) + ":";
    for (auto it = cache_.begin(); it != cache_.end();) {
        if (it->first.compare(0, pre.size(), pre) == 0) { cache_bytes_ -= it->second.g->bytes; it = cache_.erase(it); }
        else ++it;
    }
    return MF_OK;
}

int Engine::solve_in(const Query& q, char* base, size_t cap_bytes, std::vector<Result>& out) {
    if (q.graph < 0 || q.graph >= (int)graphs_.size()) { set_error("bad graph index"); return MF_EINVAL; }
    if (live() || !queue_.empty()) { set_error("engine busy"); return MF_EBUSY; }
    for (auto& S : slots_)
        if (S.pinned) { set_error("engine has pinned slots"); return MF_EBUSY; }
    const DGraph& g = graphs_[q.graph]->g;
    const size_t need = Slab::bytes_for(g.n, g.E, g.ntiles);
    if (need > cap_bytes) {
        set_error("the graph's query slab needs " + std::to_string(need) + " bytes, more than the " +
                  std::to_string(cap_bytes) + " bytes available");
        return MF_ENOMEM;
    }
    // slot 0 runs the query on a slab laid over the caller's memory; the
    // arena block it may hold is put back afterwards
    Slab saved = slots_[0].slab;
    Slab& sl = slots_[0].slab;
    sl = Slab();
    Slab::layout(g.n, g.E, g.ntiles, base, &sl);
    sl.held = true;
    sl.bytes = need;
    const int graphs_saved = P_.graphs;
    P_.graphs = 0;  // instantiating a captured graph would allocate
    int rc = admit(0, q, false, false);
    if (rc == MF_OK) {
        memset(&h_ctr_[0], 0, sizeof(Ctr));
        while (live() > 0) {
            rc = step(out);
            if (rc < 0) break;
        }
    }
    P_.graphs = graphs_saved;
    slots_[0].live = false;
    slots_[0].run_graph.reset();
    slots_[0].slab = saved;
    return rc < 0 ? rc : MF_OK;
}

int Engine::dbg_begin(int gid, int s, int t) {
    if (gid < 0 || gid >= (int)graphs_.size()) { set_error("bad graph index"); return MF_EINVAL; }
    for (auto& S : slots_) if (S.live) { set_error("engine busy"); return MF_EBUSY; }
    Query q;
    q.graph = gid; q.s = s; q.t = t;
    const DGraph& g = graphs_[gid]->g;
    int rc = ensure_slab(0, g.n, g.E, g.ntiles);
    if (rc == kNoRoom) { set_error("no room in the slab arena"); return MF_ENOMEM; }
    if (rc != MF_OK) return rc;
    // the hooks work on the full graph: no pruning, whatever the graph's
    // reachability-class cache already holds; the reference's preflow
    rc = admit(0, q, false, false);
    if (rc != MF_OK) return rc;
    slots_[0].view.inject = 0;
    slots_[0].view.recv_ok = 0;  // hooks load residuals the marks do not follow
    memset(&h_ctr_[0], 0, sizeof(Ctr));
    return MF_OK;
}

int Engine::dbg_plan(const Plan& p) {
    std::vector<Plan> plans(nslots_);
    plans[0] = p;
    slots_[0].view.phase = 0;
    return launch_plans(plans);
}

void Engine::dbg_end() { slots_[0].live = false; slots_[0].run_graph.reset(); }

const Ctr& Engine::ctr(int i) const { return h_ctr_[i]; }
SlotView& Engine::view(int i) { return slots_[i].view; }

int Engine::serve(std::vector<Query> qs, std::vector<Result>& out, double T0_in) {
    for (size_t i = 1; i < qs.size(); ++i)
        if (qs[i].t_arrive < qs[i - 1].t_arrive) { set_error("serve: arrivals must be ascending"); return MF_EINVAL; }
    for (auto& q : qs)
        if (q.graph < 0 || q.graph >= (int)graphs_.size()) { set_error("serve: bad graph index"); return MF_EINVAL; }
    const double T0 = T0_in < 0 ? now() : T0_in;
    size_t next = 0;
    const size_t first = out.size();
    while (out.size() - first < qs.size()) {
        const double tn = now() - T0;
        while (next < qs.size() && qs[next].t_arrive <= tn) {
            Query q = qs[next++];
            q.t_arrive += T0;
            queue_.push_back(q);
        }
        if (live() == 0 && queue_.empty()) {
            if (next >= qs.size()) break;
            // idle until the next arrival
            const double wait = qs[next].t_arrive - (now() - T0);
            if (wait > 0) std::this_thread::sleep_for(std::chrono::microseconds((int64_t)(std::min(wait, 1e-3) * 1e6)));
            continue;
        }
        int rc = step(out);
        if (rc < 0) return rc;
    }
    return MF_OK;
}

int Engine::run(std::vector<Result>& out, int64_t max_steps) {
    int64_t k = 0;
    for (;;) {
        const int before = live();
        if (before == 0 && queue_.empty()) break;
        int rc = step(out);
        if (rc < 0) return rc;
        if (max_steps >= 0 && ++k >= max_steps) break;
    }
    // flush slots that finished on the last step
    return MF_OK;
}

}  // namespace mf
