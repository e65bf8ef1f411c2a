// mf_engine.h -- resident graphs, per-query slabs, the per-query scheduler
// and the step executor that drives many queries through the kernels at once.
#pragma once
#include <hip/hip_runtime_api.h>

#include <atomic>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/mscmf.h"
#include "host_graph.h"
#include "mf_device.h"

namespace mf {

// ------------------------------------------------------------------ errors
// Status codes are the ABI's (MF_OK, MF_EINVAL, ... in include/mscmf.h).
void set_error(const std::string& msg);
const std::string& last_error();

// -------------------------------------------------------------- parameters
// Defaults reproduce the reference's constants (BC:1459-1484, 1725, 1936).
struct Params {
    int K = 0;                 // cycles per vertex per discharge pass; 0 = max(4, E/V)
    int gr_freq = 128;         // GR every gr_freq iterations (BC:1466)
    int gr_freq_min = 128, gr_freq_max = 256;
    int gr_stall = 1;          // also GR after excess[t] stalls this many iterations (0 = off)
    int h_cut = 100;           // BC:1469
    int stagnate_rounds = 128; // BC:1472
    int path_switch = 8;       // BC:1480
    int path_len = 256;        // BC:1481
    int path_attempts = 4;     // BC:895
    int p_count0 = 64;         // BC:1483
    // hot-region extra sweep (BC:1899, 1936): off by default; with the
    // terminal-cut stop and the relabel's picked entries it cost 4% on R-MAT
    // 22 hub queries (batched and plain) and 3% on the 1024^2 grid
    int hot = 0, hot_k = 32, hot_min_active = 64, hot_rounds = 1;
    int rm = 1, rm_freq0 = 64, rm_freq_max = 256, rm_margin = 4;    // BC:1459, 2023, 2086
    double rm_tau = 1e-5;
    int tile_rounds = 1;       // local cycles per region pass (LDS-staged)
    double tile_frac = 1.0 / 64.0;  // region mode when active >= tile_frac * n
    // certify every answer on the GPU: 3 fingerprint over the pairs pushes
    // changed (1 where the slab's marks do not cover its residual), 1
    // fingerprint over every pair, 2 pair sums entry by entry, 0 off
    int cert = 3;
    long long iter_max = 20000000;  // BC:1476, but reported as an error
    // levels one single-workgroup BFS launch may run: bounds a step's length
    // on deep graphs so batched slots keep moving while one relabels (grid
    // 1024^2, 512 queries in 256 slots: 128 207 q/s against 166-172 at 256,
    // 201 at 96, 203 at 64 and 124 at 512)
    int small_levels = 128;
    // workgroups per slot of those multi-level BFS launches on deep graphs
    // (last relabel, or the running BFS, at least 64 levels deep); 1 = one
    // 1024-thread workgroup.  More meet at a barrier in global memory after
    // each level, so a level's scattered loads use several CUs; a shallow
    // BFS keeps one workgroup, its few levels not worth a barrier each.
    // Grid 1024^2, 128 queries in 64 slots: 4 per slot 44 q/s against 1, 2
    // and 8 at 40, 41 and 35 (every barrier's agent-scope fences and polls
    // cost more with more workgroups); one slot alone is the same at 4, 8
    // and 16.  A launch also holds at most 256 of them (launch_bfs_small).
    int small_wgs = 4;
    // a frontier whose rows sum to more entries than this leaves the
    // single-workgroup BFS for the grid-wide top-down or bottom-up levels:
    // a hub's row is spread over the device instead of one CU.  R-MAT 22
    // hub queries, 4 engines x 48 slots: 2^13 2304 q/s against 2180-2200 at
    // 2^16 and 2004 at 2^18 (plain 473 against 463)
    long long small_edges = 1 << 13;
    // ... on deep graphs (the last exact relabel, the graph's depth hint or
    // the running BFS at least 64 levels deep), where each grid-wide level
    // would be an engine step of its own: grid frontiers of up to
    // kSmallFrontier degree-8 rows stay inside
    long long small_edges_deep = 1 << 17;
    // entries a slab lists as they are first marked (at most kMarkListMax);
    // resets and the sparse certificate read the list instead of the E-byte
    // mark map while it holds every marked entry (0 = always the map)
    int mark_list = kMarkListMax;
    // direction-optimizing BFS thresholds: bottom-up once the frontier's rows
    // exceed the unvisited rows / bu_alpha, top-down again below n / bu_beta.
    // On R-MAT 22 hub queries alpha 7 beat 2, 4, 10, 14 and 28 (+3% over 14)
    // once top-down levels read reverse residuals along the row
    double bu_alpha = 7.0, bu_beta = 24.0;
    int prune = 1;             // per-query reachability pruning (BC:2678)
    double prune_ratio = 0.995;  // BC:2647
    // exact SCC routing (SCC_CONTRACT/split_queries_by_scc.cpp:49-76 made
    // exact): a query whose s and t share a strongly connected component C
    // runs on C's induced subgraph (every s-t path stays inside C); others
    // keep the reachability mask.  0 = the reference's mask for every query.
    int scc_route = 0;
    // bottom-up relabel levels skip vertices that cannot reach t's SCC when
    // that SCC is the graph's largest (DeviceGraph::d_live); 0 = scan all
    int bu_skip = 1;
    // stop as soon as excess(t) fills t's in-capacity or s's out-capacity
    // (the cut around t or s is saturated: excess(t) is the maximum flow),
    // certified with T = {t} or T = V \ {s}; 0 = only after no active
    // vertex is left
    int sink_cut = 1;
    // bottom-up relabel levels try each vertex's forward entry to its
    // longest-row head before its row (DGraph::bu_pick); 0 = rows only
    int bu_pick = 1;
    // lane and wave discharge push into lower neighbours in one pass and
    // read the row's minimum only to relabel (1); 0 = minimum pass, then
    // pushes.  Measured neutral on R-MAT 22 and 3% slower on the grid
    int push_first = 0;
    // grid-stride launches: at most slot_blocks workgroups per slot and
    // about launch_blocks per launch, so each workgroup's fixed costs
    // (staging, block sums, counter atomics) spread over several chunks.
    // On R-MAT 22 hub queries 768 per slot beat 256, 512, 1024 (-1 to -2%),
    // 2048 and 4096 (+12% batched, +21% plain over 4096); the grid +4%
    int slot_blocks = 768, launch_blocks = 16384;
    // bottom-up rows up to bu_light entries go to one lane each, longer ones
    // to a wave each (k_bfs_bu_med)
    int bu_light = 128;
    // 1: bottom-up relabel levels test rows no flow has entered or left
    // through the graph's forward mask, without reading their residuals
    int bu_recv = 1;
    // 1: the list discharge of a step runs as one persistent launch over a
    // device work queue of every slot's blocks; 0: one grid per slot, whose
    // workgroups past a slot's blocks exit at once (measured faster on
    // R-MAT 22: the queue's atomics on one word cost more than the empty
    // workgroups the hardware retires)
    int dis_queue = 0;
    // bottom-up levels test candidate parents in a frontier bitmap (n / 8
    // bytes per slot) instead of gathering their levels (0 = gather levels)
    int bu_bits = 1;
    // giant-row discharge: 1 = each chunk in rounds that stop once the row's
    // excess is gone (fewer bytes, more iterations), 0 = whole chunks, 2 =
    // rounds in engines of more than one slot.  R-MAT 22 hub queries: rounds
    // 2,453 against 2,380 q/s at 4 engines x 48 slots, 438 against 477 q/s
    // for the one-slot solver (7.8 instead of 6.3 iterations a query)
    int giant_rounds = 2;
    // shallow graphs: discharge only active vertices within disc_band
    // levels of the lowest active height (0 = every active vertex)
    int disc_band = 0;
    // also GR once the discharge has scanned gr_work x E entries since the
    // last GR (a GR costs about E scans; 0 = off).  On R-MAT 22 hub queries
    // 0.125 beat 0.0625, 0.25, 0.5, 1 and 2 (hub rows are rescanned between
    // relabels)
    double gr_work = 0.125;
    // ... on deep graphs (the last exact relabel at least 64 levels deep),
    // where path pushes scan whole paths: grid 1024^2, 512 queries in 256
    // slots 170 q/s at 0.5 against 169 at 0.25, 146 at 1 and 149 at 0.125
    // (more relabels, each about 1,700 levels)
    double gr_work_deep = 0.5;
    // discharge cycles per pass for rows scanned by a whole wave (light rows
    // use K): a second cycle pushes what the first relabel made admissible
    // without another count step (2 beat 1, 3 and 4 on R-MAT 22 by 1.6%)
    int K_heavy = 2;
    // local relabel (BC:1582-1721) before a discharge when the active height
    // band is at most local_band wide.  Off by default: on R-MAT hub queries
    // the tiles holding active vertices span most of the graph, so each local
    // relabel costs about a global one (plain solve 52.6 -> 11 queries/s at
    // R-MAT scale 22)
    int local_relabel = 0;
    int local_band = 128;
    // discharge + count cycles per engine step (one host round trip) when the
    // last global relabel was at least 64 levels deep (grids); 1 = off
    int fuse = 8;
    // replay the fused cycles of a step as a captured HIP graph when the
    // same set of slots fuses again (not while profiling)
    int graphs = 1;
    // when the discharge leaves no active vertex, certify with its labels
    // (T = {h < n}) before spending an exact global relabel on termination
    int spec_cert = 1;
    // count the active vertices after a discharge from the previous active
    // list and the vertices the discharge pushed into, instead of a pass
    // over every vertex (not after path pushes or RM, which record nothing)
    int sparse_count = 1;
    // Bounded injection (1): the preflow puts B = min(out-capacity of s,
    // in-capacity of t) on s, an upper bound on the flow, instead of
    // saturating the row of s (0, BC:173-200), and s is then an ordinary
    // vertex.  Off by default: on R-MAT 22 hub queries the excess then leaves
    // s only along shortest paths and finds the longer ones one relabel at a
    // time (6.2 global relabels per query against 2.7, 400 against 1298
    // queries/s batched).  Pinned (dynamic) queries always use 0.
    int inject = 0;
    // threads of the tile (region) discharge workgroup: 256, 512 or 1024; the
    // reference's region kernel uses 512 (BC:1725); an autotuner dimension
    int tile_threads = 512;
    // path pushes on deep graphs (the last exact relabel at least 64 levels
    // deep): up to path_switch_deep active vertices, after p_count_deep
    // iterations with that few (grids: 252 -> 178 ms per plain query, 4.2
    // global relabels instead of 7.3 at 64; grid 1024^2 with 64 slots 72 q/s
    // at 256 against 45 at 64: with about 100 active vertices the list
    // discharge moved excess one hop per iteration, 1,400 iterations a query
    // instead of 26)
    int path_switch_deep = 256;
    int p_count_deep = 1;
    // tile discharge (region mode and the hot sweep): up to this many
    // workgroups per tile, dealt by the tile's excess score as the
    // predecessor dealt blocks to regions; 1 = one workgroup per tile
    int tile_split = 1;
};

// Wall-clock ticks a multi-workgroup BFS barrier may wait before it flags
// the slot's error word (about 2 s).
long long barrier_spin_ticks();
// Workgroups one multi-workgroup BFS launch may hold: a process-wide budget
// split over the live engines' two streams each.
int multi_wg_cap();

// ---------------------------------------------------------- device memory
uint64_t next_graph_uid();

struct DeviceGraph {
    DGraph g;
    uint64_t uid = next_graph_uid();  // identifies the topology a residual belongs to
    // bumped by every capacity update; engines on other threads read it when
    // deciding whether a slab's residual is still current
    std::atomic<uint64_t> cap_version{0};
    int32_t* d_row_ptr = nullptr;
    int32_t* d_col = nullptr;
    int64_t* d_cap = nullptr;
    int32_t* d_rev = nullptr;
    int32_t* d_tile_start = nullptr;
    int32_t* d_tile_of = nullptr;
    int32_t* d_sub2old = nullptr;  // pruned run graphs: run id -> parent id
    int32_t* d_old2sub = nullptr;  // pruned run graphs: parent id -> run id (-1 = dropped)
    uint8_t* d_mask = nullptr;     // pruned run graphs: parent mask (cache check)
    size_t bytes = 0;
    int tile_max = kTileMax;
    // levels of the last exact relabel any query saw on this graph (engines
    // on several threads may share the graph)
    std::atomic<int> depth_hint{0};
    // Reachability classes for pruning: scc (over cap > 0 entries) and wcc
    // ids on the host, wcc ids on the device, and the forward-reach set of
    // each SCC once some query has needed it.
    bool classes = false;
    std::vector<int32_t> scc, wcc;
    std::vector<int32_t> scc_size;  // vertices per SCC id
    int32_t* d_wcc = nullptr;
    int32_t* d_scc = nullptr;
    // The largest SCC (id in the graph a query names) and, per vertex of
    // this graph, whether it can reach that SCC over cap > 0 entries: the
    // relabel's bottom-up levels skip the others for t in that SCC.  A
    // pruned run graph carries its parent's (giant_scc is the parent's id).
    int32_t giant_scc = -1;
    uint8_t* d_live = nullptr;
    int64_t* d_capsum = nullptr;   // certificate constants (DGraph::capsum, cert_fp)
    uint64_t* d_cert_fp = nullptr;
    int64_t* d_pcap = nullptr;     // pair capacities (DGraph::pcap)
    int32_t* d_fstart = nullptr;   // first forward entry per row (DGraph::fstart)
    int32_t* d_bu_pick = nullptr;  // DGraph::bu_pick
    uint32_t* d_fwdm = nullptr;    // DGraph::fwdm
    struct FwdSet {
        std::vector<uint8_t> mask;  // host copy
        uint8_t* d_mask = nullptr;
        int64_t count = 0;
        ~FwdSet();
    };
    // Engines on several host threads share the graph: every access to
    // fwd_cache holds fwd_mu (the sets themselves are immutable once cached).
    std::unordered_map<int32_t, std::shared_ptr<FwdSet>> fwd_cache;
    mutable std::mutex fwd_mu;
    std::shared_ptr<FwdSet> fwd_find(int32_t cls) const;
    // Inserts set unless another engine cached the class first; returns the
    // cached one.
    std::shared_ptr<FwdSet> fwd_insert(int32_t cls, std::shared_ptr<FwdSet> set);
    // forward entries (cap > 0) as a bitmap: capacity updates may only name these
    std::vector<uint64_t> fwd_bits;
    std::vector<int32_t> old2sub_host;  // pruned run graphs: host copy of d_old2sub
    ~DeviceGraph();
};

// mask8[v] = (h[v] < n), the forward-reach set left by a BFS_FWD in h.
void launch_fwd_from_h(hipStream_t st, const int32_t* h, int n, uint8_t* mask8);
// mask[v] = fwd[v] && wcc[v] == wt  (int32 0/1)
void launch_class_mask(hipStream_t st, const uint8_t* fwd, const int32_t* wcc, int32_t wt, int n,
                       int32_t* mask);
// mask[v] = ids[v] == want  (int32 0/1)
void launch_eq_mask(hipStream_t st, const int32_t* ids, int32_t want, int n, int32_t* mask);

// Tiles: each region [rstart[r], rstart[r+1]) cut into pieces of at most tile_max.
void make_tiles(const std::vector<int32_t>& rstart, int tile_max, std::vector<int32_t>& tile_start);

int upload_graph(const HostGraph& hg, int tile_max, DeviceGraph& dg);
// Allocate and compute dg's certificate constants (after its CSR is set).
int init_cert_consts(DeviceGraph& dg, hipStream_t st);

// Per-query device arrays (the reference's DynamicState, BC:47-113), carved
// from the engine's slab arena for the run graph's (n, E, ntiles).
struct Slab {
    int n_cap = 0, E_cap = 0, tiles_cap = 0;
    int64_t* res = nullptr;
    int64_t* excess = nullptr;
    int32_t* h = nullptr;
    int32_t* fa = nullptr;
    int32_t* fb = nullptr;
    int32_t* lock = nullptr;
    int32_t* aux = nullptr;
    int32_t* hl = nullptr;
    int32_t* tile_cnt = nullptr;
    int32_t* tile_hot = nullptr;
    int64_t* tile_ex = nullptr;
    int32_t* tsplit = nullptr;
    int32_t* lbkt = nullptr;
    int32_t* giant = nullptr;
    int64_t* gacc = nullptr;
    int gcap = 0;
    uint64_t* fbits = nullptr;  // 2 x ceil(n / 64) frontier bitmap words (SlotView::fbits)
    uint8_t* recv = nullptr;    // n bytes: the vertex's row left its capacities (SlotView::recv)
    uint8_t* emark = nullptr;   // E marks of changed entry pairs
    int32_t* mlist = nullptr;   // count + the marked entries (SlotView::mlist)
    int mcap = 0;
    uint64_t res_uid = 0;       // graph whose residual res holds (0 = none)
    // every entry pair whose residual differs from its capacities has a
    // marked entry (res reset from the capacities, since changed by marking
    // pushes and updates only): the sparse certificate may run
    bool marks_ok = false;
    uint64_t res_version = 0;   // ... at this capacity version
    size_t off = 0, bytes = 0;  // arena block
    bool held = false;
    bool fits(int n, int64_t E, int ntiles) const {
        return held && n <= n_cap && E <= (int64_t)E_cap && ntiles <= tiles_cap;
    }
    // Bytes of a slab for (n, E, ntiles), every array 256-byte aligned; with
    // base != nullptr the arrays of sl are carved from base.
    static size_t layout(int n, int64_t E, int ntiles, char* base, Slab* sl);
    static size_t bytes_for(int n, int64_t E, int ntiles) { return layout(n, E, ntiles, nullptr, nullptr); }
};

// One device allocation that slabs are carved from: first fit by address,
// free neighbours coalesced.
struct SlabArena {
    char* base = nullptr;
    size_t size = 0, used = 0, peak = 0;
    std::map<size_t, size_t> free_;  // offset -> bytes
    int create(size_t bytes);
    void destroy();
    bool alloc(size_t bytes, size_t* off);
    void release(size_t off, size_t bytes);
};

// --------------------------------------------------------------- scheduler
enum StageBit : unsigned {
    ST_RESET_PRE = 1u << 0,
    ST_INIT = 1u << 1,
    ST_PREFLOW = 1u << 2,
    ST_GR_SEED = 1u << 3,
    ST_BFS_SEED1 = 1u << 4,
    ST_BFS = 1u << 5,
    ST_MASK = 1u << 6,
    ST_DISCHARGE = 1u << 7,
    ST_HOT = 1u << 8,
    ST_RM = 1u << 9,
    ST_RESET_ACT = 1u << 10,
    ST_ACTIVE = 1u << 11,
    ST_CERT = 1u << 12,
    ST_LOCAL = 1u << 13,
    ST_SINK = 1u << 14,  // labels T = {t} before the certificate
    ST_SOURCE = 1u << 15,  // labels T = V \ {s} before the certificate
};

enum BfsKind : int { BK_SMALL = 0, BK_TD = 1, BK_BU = 2 };
enum DisKind : int { DK_TILE = 0, DK_LIST = 1, DK_PATH = 2 };

struct Plan {
    unsigned stages = 0;
    unsigned reset_pre = 0;
    unsigned reset_act = 0;
    int bfs_kind = BK_SMALL, bfs_mode = 0, bfs_grid = 1;
    int dis_kind = DK_LIST, dis_grid = 1;
    int seed_mode = 0;
    int local_cap = 0;        // ST_LOCAL: level cap of the local relabel
    int cycles = 1;           // ST_DISCHARGE: discharge + count cycles in this step
    bool sparse = false;      // ST_ACTIVE: count the previous list and the touched vertices only
    int small_wgs = 1;        // BK_SMALL: workgroups per slot (1 or Params::small_wgs)
    bool seed_t_only = false; // ST_GR_SEED in phase 0 after a count that saw no deficit: t is the one seed
    bool bu_bits = false;     // BK_BU: the first bottom-up level of this BFS builds level L's bitmap first
    bool cert_sside = false;  // ST_CERT: the sparse certificate scans the rows of V \ T (the smaller side)
    long long small_edges = 0; // BK_SMALL: frontier-entry bound of the multi-level launch (0 = Params::small_edges)
};

// Per-query statistics (host side).
struct QueryStats {
    int64_t flow = 0;
    int status = MF_OK;
    int n_run = 0;
    int64_t E_run = 0;
    int pruned = 0;           // ran on an induced subgraph
    int kept = 0;             // vertices kept by the pruning mask
    int64_t iters = 0;        // discharge iterations
    int64_t steps = 0;        // scheduler steps (host syncs)
    int64_t grs = 0;          // exact global relabels
    int64_t bfs_levels = 0;
    int64_t drains = 0;       // deficit drain phases
    int64_t local_relabels = 0;
    int64_t spec_fails = 0;   // certificates on discharge labels that failed
    int64_t sink_cuts = 0;    // certificates tried on the cut around t or s (excess(t) = its capacity)
    int64_t bfs_restarts = 0; // BFS runs restarted single-workgroup after a barrier time-out
    int64_t rm_fixed = 0;
    int cert_ok = 0;
    uint64_t edge_scans = 0, vertex_scans = 0, pushes = 0, relabels = 0;
    uint64_t wk[kFamilies][4] = {};  // per kernel family (see WorkFamily)
    double t_submit = 0, t_start = 0, t_end = 0;  // seconds, engine clock
};

class SlotSched {
public:
    // warm: keep the residual and excess already in the slab (dynamic mode)
    void begin(const Params* P, const DGraph& g, int s, int t, bool prune_first, bool warm = false);
    // Decide the next step given the counters of the last one.
    Plan plan(const Ctr& c);
    bool done() const { return st_ == S_DONE || st_ == S_ERROR; }
    // depth of this graph's relabels seen by earlier queries: picks the
    // multi-workgroup small-frontier BFS from a query's first relabel on
    void set_depth_hint(int d) { depth_hint_ = d; }
    int gr_depth() const { return gr_depth_; }
    // a BFS at level L on a deep graph: multi-workgroup levels and the deep
    // frontier-entry bound
    bool deep_bfs(int L) const { return gr_depth_ >= 64 || depth_hint_ >= 64 || L >= 64; }
    // height cap of the next discharge step (INT_MAX: none)
    int disc_cap() const { return disc_cap_; }
    int status() const { return status_; }
    const std::string& error() const { return err_; }
    int phase() const { return phase_; }
    int K() const { return K_; }
    int gr_freq() const { return GR_FREQ_; }
    QueryStats stats;
    // pruning hand-off: after S_MASK the engine swaps graphs and calls resume()
    bool wants_prune_decision() const { return st_ == S_PRUNE_DECIDE; }
    void resume_after_prune(const DGraph& g, int s, int t);
    // Only the forward BFS is needed (the engine has the WCC of t).
    void set_fwd_only() { fwd_only_ = true; }
    // The answer is known without a solve (t not reachable from s).
    void finish_trivial(int64_t flow);

private:
    enum St { S_PRUNE_FWD, S_PRUNE_UND, S_PRUNE_MASK, S_PRUNE_DECIDE, S_START, S_WARM_START, S_BFS,
              S_COUNTED, S_CERT, S_DONE, S_ERROR };
    Plan start_gr(const Ctr* c = nullptr);
    Plan bfs_step(const Ctr& c, int mode);
    void fail(int code, const std::string& msg);

    const Params* P_ = nullptr;
    DGraph g_;
    int s_ = 0, t_ = 0;
    St st_ = S_START;
    int status_ = MF_OK;
    std::string err_;
    int phase_ = 0;
    int bfs_mode_ = 0;
    int64_t bfs_medges_ = 0;   // edges of visited vertices (direction heuristic)
    bool bfs_bottom_up_ = false;
    bool just_gr_ = false;
    int64_t iter_ = 0, last_gr_iter_ = 0;
    int gr_depth_ = 0;  // levels of the last exact global relabel
    int depth_hint_ = 0;
    int GR_FREQ_ = 128, K_ = 4, RM_FREQ_ = 64;
    int stag_ = 0, no_stag_ = 0;
    int64_t last_pos_ = INT64_MAX;
    // the last full count's positive excess and flow into t; a sparse count
    // sees only its candidates, and without deficits the positive excess
    // off s and t is -excess(s) - excess(t), excess(s) fixed in phase 0
    int64_t pos_full_ = 0, ext_full_ = 0;
    bool counted_sparse_ = false;
    bool force_gr_ = false;
    int path_stable_ = 0, p_count_ = 64, path_rounds_ = 0;
    int64_t last_ext_ = -1;
    uint64_t scans_at_gr_ = 0;  // discharge edge scans when the last GR started
    int stall_ = 0;
    int drain_loops_ = 0;
    bool rm_off_ = false;
    bool rm_last_ = false;
    bool work_reset_ = false;
    bool fwd_only_ = false;
    bool spec_ = false;         // the certificate under way follows discharge labels, not a GR
    int disc_cap_ = 0x7fffffff;
    bool spec_failed_ = false;
    bool after_local_ = false;
    // Deficits (negative excess off s) may exist: set by RM repairs and warm
    // starts (capacity decreases), cleared only by a full count that sees
    // none.  Sparse counts do not see deficits, so they never clear it.
    bool deficits_ = false;
    // vertices the last exact relabel labelled: the size of T = {h < n} a
    // certificate on its labels (or the discharge's, which only shrink T) sees
    int64_t gr_visited_ = 0;
    // a multi-workgroup BFS barrier gave up (the device was too busy to hold
    // all its workgroups): this query's BFS runs single-workgroup from here on
    bool no_multi_ = false;
    // Restart the BFS under way after a barrier time-out.
    Plan restart_bfs(const Ctr& c);
};

// ------------------------------------------------------------------ engine
// A query submitted to the engine.
struct Query {
    int graph = 0;     // index into the engine's graph table
    int s = 0, t = 0;  // ids of that graph
    int64_t tag = 0;
    double t_arrive = 0;  // engine clock; 0 = time of submit()
};

struct Result {
    int64_t tag = 0;
    QueryStats st;
};

class Engine {
public:
    Engine() = default;
    ~Engine();
    // nslots concurrent queries; each slot's slab is sized for the largest
    // registered graph it may run.
    int init(int nslots, const Params& P);
    int add_graph(std::shared_ptr<DeviceGraph> g);  // returns graph index
    int submit(const Query& q);                       // queued
    // Run scheduler steps until the queue and all slots are empty, or until
    // max_steps.  Finished results are appended to out.
    int run(std::vector<Result>& out, int64_t max_steps = -1);
    // One step; returns number of live slots after it.
    int step(std::vector<Result>& out);
    size_t queued() const { return queue_.size(); }
    int live() const;
    const Params& params() const { return P_; }
    void set_cert_mode(int m) { P_.cert = m; }
    Params& params_mut() { return P_; }
    double now() const;
    hipStream_t stream() const { return stream_; }
    // Device pointers of slot i's slab (debug / parity hooks).
    Slab* slab(int i) { return &slots_[i].slab; }
    int nslots() const { return nslots_; }
    // Slab arena: budget in bytes (0 = nslots full slabs of the largest
    // graph).  A queued query is admitted when a slot is free and its run
    // graph's slab fits; it waits otherwise (first come, first served).
    int set_budget(size_t bytes);
    void mem_stats(size_t* budget, size_t* in_use, size_t* peak, int64_t* waits) const;
    // Serve queries arriving at offsets qs[i].t_arrive (seconds from the call,
    // ascending): each is submitted when its time comes, the engine steps
    // while anything is queued or live and idles until the next arrival.
    // T0 < 0: the call's start; else an instant on the engine clock shared by
    // engines serving side by side.
    int serve(std::vector<Query> qs, std::vector<Result>& out, double T0 = -1.0);
    // Single-slot manual control for parity hooks: admit (gid, s, t) into
    // slot 0 without scheduling, launch hand-built plans, release.
    int dbg_begin(int gid, int s, int t);
    int dbg_plan(const Plan& p);
    void dbg_end();
    // The class-path pruning mask of (s, t) (fwd(scc(s)) AND wcc == wcc(t),
    // built by the same calls as class_decide); slot 0 must hold a BFS_FWD
    // from s in h unless s's class is cached.  reachable = 0: t is not
    // reachable from s (the mask is not built).
    int dbg_class_mask(int gid, int s, int t, int32_t* mask, int64_t* kept, int* reachable);
    const Ctr& ctr(int i) const;
    SlotView& view(int i);
    DeviceGraph* graph(int gid) { return graphs_[gid].get(); }
    int ngraphs() const { return (int)graphs_.size(); }
    // Dynamic mode: pinned slots keep their residual across updates.
    int pin(int slot, const Query& q);
    int solve_pinned(bool warm, std::vector<Result>& out);
    int update_caps(int gid, const int32_t* entries, const int64_t* new_cap, int n);
    // One query on its full graph (no pruning, no captured graphs), its slab
    // carved from caller memory [base, base + cap_bytes) instead of the
    // arena: no device allocation on this path.  The engine must be idle.
    int solve_in(const Query& q, char* base, size_t cap_bytes, std::vector<Result>& out);
    int device() const { return device_; }

private:
    struct SlotState {
        bool live = false;
        bool pinned = false;
        bool solved = false;   // pinned slot holds a residual for its query
        Query q;
        SlotSched sched;
        SlotView view;
        std::shared_ptr<DeviceGraph> run_graph;  // pruned subgraph when used
        int parent_n = 0;
        Slab slab;  // kept after the query retires, reused when the next one fits
    };
    int admit(int slot, const Query& q, bool warm = false, bool prune = true);
    // (n, E, ntiles) of the graph q will run on: the cached pruned subgraph of
    // its reachability class when there is one, else its full graph.
    void run_dims(const Query& q, int* n, int64_t* E, int* ntiles) const;
    // Give slot i a slab for (n, E, ntiles): MF_OK, kNoRoom (the arena is full
    // for now) or an error.
    static constexpr int kNoRoom = 1;
    int ensure_slab(int i, int n, int64_t E, int ntiles);
    void drop_slab(int i);
    int make_arena();
    int prune_decide(int slot);
    int class_decide(int slot);
    // SCC routing: s and t share an SCC (the query runs on that SCC)
    bool intra_route(const DeviceGraph& G, const Query& q) const;
    // cache key of the subgraph q runs on: its SCC, or its reachability class
    std::string route_key(const DeviceGraph& G, const Query& q) const;
    // The forward set of SCC cls: cached, or made from d_h holding a BFS_FWD
    // from a vertex of cls (then cached).
    int fwd_set(DeviceGraph& G, int cls, const int32_t* d_h, std::shared_ptr<DeviceGraph::FwdSet>& out);
    static int64_t class_kept(const DeviceGraph& G, const DeviceGraph::FwdSet& fwd, int wt);
    int launch_plans(std::vector<Plan>& plans);

    Params P_;
    int nslots_ = 0;
    int device_ = 0;            // HIP device current at init
    bool counted_ = false;      // counted in the live-engine total
    long long spin_ticks_ = 0;  // multi-workgroup BFS barrier wait limit
    hipStream_t stream_ = nullptr;
    hipStream_t side_ = nullptr;          // single-workgroup BFS steps, overlapped with the rest
    hipEvent_t ev_up_ = nullptr, ev_side_ = nullptr;
    std::vector<std::shared_ptr<DeviceGraph>> graphs_;
    std::vector<SlotState> slots_;
    SlabArena arena_;
    // fused cycles: slot lists at fixed places (cycle, kind) so that a
    // captured graph of the cycles can be replayed; graphs keyed by the lists
    int32_t* cyc_d_ = nullptr;
    int32_t* cyc_h_ = nullptr;
    struct GraphEntry {
        hipGraph_t g = nullptr;
        hipGraphExec_t x = nullptr;
        uint64_t last = 0;
    };
    std::map<std::vector<int32_t>, GraphEntry> graph_cache_;
    std::set<std::vector<int32_t>> graph_seen_;
    uint64_t graph_clock_ = 0;
public:
    int64_t graph_replays = 0, graph_captures = 0;
    // BFS runs of retired queries restarted single-workgroup after a
    // multi-workgroup barrier gave up waiting
    int64_t bfs_restarts = 0;
    // the barrier's wait bound in wall-clock ticks (<= 0: the default, 2 s)
    void set_barrier_ticks(long long ticks) { spin_ticks_ = ticks > 0 ? ticks : barrier_spin_ticks(); }
private:
    size_t budget_ = 0;         // 0 = nslots full slabs of the largest graph
    bool arena_stale_ = false;  // default-sized arena smaller than the graphs now need
    int64_t admit_waits_ = 0;   // steps whose queue head waited for slab bytes
    std::deque<Query> queue_;
    // device copies of views and launch lists; pinned staging
    SlotView* d_views_ = nullptr;
    unsigned long long* d_hash_ = nullptr;
    void* h_stage_ = nullptr;  // pinned: views + lists
    Ctr* h_ctr_ = nullptr;     // pinned: per-slot counter snapshots
    Ctr* d_ctrs_ = nullptr;    // per-slot device counters, contiguous: one copy per step
    unsigned long long* h_hash_ = nullptr;
    size_t list_cap_ = 0;
    double t0_ = 0;
    std::vector<hipEvent_t> ev_;  // profiling event pairs
    // induced-subgraph cache: key = (graph index, mask hash, kept count)
    struct CacheEntry {
        std::shared_ptr<DeviceGraph> g;
        uint64_t last_use = 0;
        int64_t kept = 0;  // vertices of the pruning mask
        uint64_t version = 0;  // the parent's capacity version the copy was built at
    };
    std::unordered_map<std::string, CacheEntry> cache_;
    uint64_t use_clock_ = 0;
    size_t cache_bytes_ = 0;
public:
    // Per kernel family timing with HIP events on the engine stream, and the
    // algorithmic work of retired queries.
    bool profile = false;
    // per-step log (mf_engine_set_step_log): step_ms_ holds the last step's
    // stage times in the log's buckets
    using StepFn = void (*)(void*, const mf_step_log*);
    StepFn step_fn = nullptr;
    void* step_user = nullptr;
    double step_ms_[MF_LOG_FIELDS] = {};
    double fam_ms[kFamilies] = {};
    int64_t fam_launches[kFamilies] = {};
    uint64_t fam_work[kFamilies][4] = {};
    void prof_reset();
    size_t cache_budget = (size_t)64 << 30;
    int64_t cache_hits = 0, cache_misses = 0;
};

// Induced subgraph of parent on the slot's mask (aux: 0/1), built on the GPU
// (BC:2477-2551).  Vertex order, and so region/tile membership, is inherited
// from the parent.
int build_induced_gpu(const DeviceGraph& parent, const int32_t* d_mask, int kept,
                      hipStream_t st, DeviceGraph& out);

double wall_seconds();

}  // namespace mf
