// mf_device.h -- device-visible structs shared by the HIP kernels and the
// host scheduler.  Plain C++ layout (no HIP types) so host code compiled by
// any compiler can fill them.
#pragma once
#include <cstdint>

namespace mf {

// Resident topology of one region-ordered graph (device pointers).
// row_ptr[n+1], col/rev[E] (int32), cap[E] (int64): the reference's CSR with
// explicit reverse entries (BC:20-27).  Vertices are region-ordered; the
// regions are further cut into tiles of at most kTileMax vertices (a tile
// never spans two regions), tile_start[ntiles+1], tile_of[n].
struct DGraph {
    int n = 0;
    int E = 0;
    int ntiles = 0;
    int R = 0;
    int max_deg = 0;  // longest row (an upper bound for induced subgraphs)
    const int32_t* row_ptr = nullptr;
    const int32_t* col = nullptr;
    const int64_t* cap = nullptr;
    const int32_t* rev = nullptr;
    const int32_t* tile_start = nullptr;
    const int32_t* tile_of = nullptr;
    // certificate constants, kept current by capacity updates: capsum[v] =
    // sum of cap over v's row; *cert_fp = sum over entries of
    // pair_weight(pair, cert_seed) * cap (mod 2^64)
    int64_t* capsum = nullptr;
    uint64_t* cert_fp = nullptr;
    uint64_t cert_seed = 0;
    // pcap[e] = cap[e] + cap[rev e], kept current by capacity updates: the
    // pair's residuals sum to it, so res[rev e] = pcap[e] - res[e] is read
    // along e's row instead of gathered from rev e's
    int64_t* pcap = nullptr;
    // fstart[v]: the first forward entry (cap > 0) of v's row, or its row
    // start when it has none.  Bottom-up relabel levels scan a row from here
    // and wrap: a parent is reached over residual out-entries, and rows built
    // from (u, v)-sorted input lines open with the reverse entries of lower
    // ids, which carry no residual until flow crosses them.
    const int32_t* fstart = nullptr;
    // bu_pick[2v], bu_pick[2v+1]: the forward entry of v whose head has the
    // longest row, and that head (-1, -1 without forward entries).  A
    // bottom-up relabel level tries it before v's row: in skewed graphs the
    // long-row heads are few, reached early and their levels stay cached,
    // so most vertices find their parent without reading their own row.
    const int32_t* bu_pick = nullptr;
    // fwdm: bit e & 31 of fwdm[e >> 5] is set when entry e is a forward
    // entry (cap > 0; capacity updates keep forward entries forward).  A
    // bottom-up relabel level tests a row whose residuals still equal its
    // capacities through it: only its forward entries can lead to a parent.
    const uint32_t* fwdm = nullptr;
};

// Offset (in int32 words, 16-byte aligned) of the giant list's task records.
inline __host__ __device__ int64_t giant_task_offset(int64_t gcap) { return (3 * gcap + 1 + 3) / 4 * 4; }

// Per-slot device counters.  Written by kernels, copied to the host once per
// scheduler step.  Field order is part of the host/device contract.
struct Ctr {
    int32_t fsize[4];        // frontier size of level L lives in fsize[L % 4]
    int64_t fedges[4];       // row-length sum of the level-L frontier, same ring
    int32_t level;           // levels completed by the running BFS
    int32_t visited;         // vertices labelled by the running BFS
    int32_t nhl;             // entries in the heavy-row work list (front of hl)
    int32_t nhg;             // giant rows of a top-down level (back of hl)
    int32_t active;          // excess>0, h<n, not s/t
    int32_t active_high;     // ... with h > H_cut
    int32_t nlight;          // light active-list entries (front of the list)
    int32_t nheavy;          // heavy entries (back of the list)
    int32_t ngiant;          // rows longer than kGiantDeg in the giant list
    int32_t ngchunks;        // kGiantChunk-entry tasks over those rows
    // barrier of the workgroups of a multi-workgroup BFS launch: arrivals,
    // generation
    int32_t bar_count;
    uint32_t bar_gen;
    // sparse counting of fused cycles: vertices pushed into since the last
    // touch reset are listed in hl (deduplicated by stamps in aux)
    int32_t ntouched;
    int32_t epoch;           // touch epoch: stamps -(2 epoch + 1) touched, -(2 epoch + 2) counted
    int32_t prev_nlight, prev_nheavy;  // active list being replaced by a sparse count
    int64_t pos_excess;      // sum of positive excess off s,t
    int32_t neg_count;       // deficits (excess<0) off s,t
    int32_t fixed_edges;     // RM repairs in the last RM pass
    int64_t neg_sum;         // sum of deficits (negative)
    int32_t hmin, hmax;      // active height band
    int32_t cert_bad;        // certificate violations
    int32_t err;             // device-side error flags
    // level whose frontier bitmap SlotView::fbits holds (-1: none yet in
    // this BFS); a bottom-up level uses the bitmap only when it is its own
    int32_t fb_level;
    int64_t cut;             // certificate: cut(V\T,T)
    uint64_t fp;             // certificate: pair-sum fingerprint (0 when every pair sums to its capacity)
    int64_t ex_t;            // excess of t seen by the last active / certificate pass
    // units the preflow put on s (bounded injection, SlotView::inject), 0 when
    // it saturated the row of s: the excess s holds beyond its row's flow
    int64_t inject;
    // capacity into t (over t's row, pcap - cap of each entry), set by the
    // preflow: once excess(t) reaches it the cut around t is saturated
    int64_t incap_t;
    // capacity out of s (capsum of s), set by the preflow: once excess(t)
    // reaches it the cut around s is saturated
    int64_t outcap_s;
    // algorithmic work counters (roofline accounting), per kernel family:
    // wk[family][0] edge scans, [1] vertex scans, [2] pushes,
    // [3] relabels + BFS discoveries
    uint64_t wk[7][4];
};

// Kernel families of the work counters and of the engine's timing.
// WF_BFS: seeding, single-workgroup and top-down levels, local relabel;
// WF_BFS_BU: bottom-up levels.
enum WorkFamily : int {
    WF_INIT = 0, WF_BFS = 1, WF_ACTIVE = 2, WF_DISCHARGE = 3, WF_RM = 4, WF_CERT = 5, WF_BFS_BU = 6
};
constexpr int kFamilies = 7;

// Everything a kernel needs to work on one query.
struct SlotView {
    DGraph g;
    int64_t* res = nullptr;      // E residual
    int64_t* excess = nullptr;   // n
    int32_t* h = nullptr;        // n heights (GR writes BFS levels, n = unreached)
    int32_t* fa = nullptr;       // n frontier / active list A
    int32_t* fb = nullptr;       // n frontier B
    int32_t* lock = nullptr;     // n path-push ownership (-1 = free)
    int32_t* aux = nullptr;      // n scratch (prune masks, compaction)
    int32_t* hl = nullptr;       // n heavy-row work list (BFS levels, certificate)
    int32_t* tile_cnt = nullptr; // ntiles active counts
    int32_t* tile_hot = nullptr; // ntiles hot flags
    int64_t* tile_ex = nullptr;  // ntiles positive excess of the tile's active vertices (last count)
    // workgroups per tile by excess score: [0] = count, then (tile, part,
    // parts) per workgroup, at most kSplitMax per tile
    int32_t* tsplit = nullptr;
    int32_t* lbkt = nullptr;     // 2 * kLocalLevels local-relabel bucket counts and cursors
    // giant-row list (rows longer than kGiantDeg; at most E / (kGiantDeg+1)
    // of them, gcap): ids [0, gcap), per-row minimum [gcap, 2 gcap), chunk
    // prefix [2 gcap, 3 gcap], then from giant_task_offset(gcap) one int4
    // per chunk task (row, vertex, first entry, end entry; at most
    // E / kGiantChunk + gcap tasks); gacc[gcap] per-row int64 sums
    int32_t* giant = nullptr;
    int64_t* gacc = nullptr;
    int32_t gcap = 0;
    // bottom-up BFS levels: two bitmaps of ceil(n / 64) words, the vertices
    // on level L in fbits[(L & 1) * nw ...] (bit v & 63 of word v >> 6); a
    // level tests its candidate parents here (n / 8 bytes, which stay in
    // L2) instead of gathering their 4-byte levels
    uint64_t* fbits = nullptr;
    uint8_t* emark = nullptr;    // E: entry's pair changed since the residual was last reset
    // the marked entries in marking order: mlist[0] = count, entries from
    // mlist[1]; past mcap entries only the map is kept (count > mcap)
    int32_t* mlist = nullptr;
    int32_t mcap = 0;
    const int32_t* mask = nullptr;  // optional vertex mask (nullptr = all)
    // relabel BFS, bottom-up levels: only vertices with live[v] != 0 look
    // for a parent (nullptr = all).  live = can reach t's strongly connected
    // component in the graph; the others never have a residual path to t
    // (flow reaches them only from s, which the relabel never enters).
    const uint8_t* live = nullptr;
    Ctr* ctr = nullptr;
    int32_t s = 0, t = 0;
    int32_t K = 4;               // discharge cycles per light vertex per pass
    int32_t Kh = 1;              // ... per heavy (wave-scanned) vertex per pass
    int32_t H_cut = 100;         // "high" label threshold (BC:1469)
    // discharge only vertices with h <= disc_cap this step (lowest band
    // first); the others stay active and are counted as before
    int32_t disc_cap = 0x7fffffff;
    // lane and wave discharge: push first, read the row minimum only to
    // relabel (1), or the minimum first, then push (0)
    int32_t push_first = 0;
    // the slab's entry marks cover its residual (reset from the capacities,
    // changed since only by marked pushes and updates): an unmarked entry
    // holds at least its capacity
    int32_t marks_ok = 0;
    // bottom-up rows up to this length are scanned by one lane, longer ones
    // (to kGiantDeg) by a wave
    int32_t bu_light = 128;
    // multi-level single- or multi-workgroup BFS launches stop before a
    // level whose frontier rows sum to more entries than this
    int64_t bfs_edges = 1 << 16;
    int32_t phase = 0;           // 0 = normal (seeds t + deficits), 1 = drain (seeds deficits, t active)
    int32_t seed_t_only = 0;     // phase 0 with no deficit (last count): the relabel seeds t alone
    // recv[v] = 1 once flow entered v's row or left it: every push marks its
    // head (move_res), the preflow s, a repair its pusher; the residual
    // resets clear it.  A vertex never marked holds res = cap on its whole
    // row (no vertex but s and a repaired one pushes without receiving
    // first).  recv_ok: the slab's marks follow this rule (not after
    // capacity updates of a warm start, nor for residuals a debug hook
    // loaded), so bottom-up relabel levels may use them.
    uint8_t* recv = nullptr;
    int32_t recv_ok = 0;
    // the sparse certificate finds the cut entries V \ T -> T from the rows
    // of V \ T (1) or, through their pairs, from the rows of T (0): the
    // same entries, read from the smaller side's rows
    int32_t cert_sside = 0;
    // 1 = bounded injection: the preflow puts B = min(out-capacity of s,
    // in-capacity of t) on s instead of saturating its row, and s is then an
    // ordinary vertex (labelled by the relabel, active, discharged); 0 = the
    // reference's preflow (BC:173-200), s never labelled
    int32_t inject = 0;
};

constexpr int kTileMax = 4096;      // vertices staged in LDS per workgroup
constexpr int kLightDeg = 32;       // rows up to this length are scanned by one lane
constexpr int kGiantDeg = 2048;     // rows longer than this are scanned by a whole workgroup
constexpr int kSmallFrontier = 16384;  // single-workgroup BFS keeps going below this
constexpr int64_t kSmallEdges = 1 << 18; // ... and while the frontier's rows stay below this
constexpr int kLocalLevels = 256;   // level cap of the local relabel
constexpr int kGiantChunk = 2048;   // entries per workgroup task on a giant row
constexpr int kMarkListMax = 1 << 20;  // marked entries listed per slab (beyond: the map alone)
constexpr int kSplitMax = 8;        // workgroups one tile may get by excess score

}  // namespace mf
