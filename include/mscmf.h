/*
 * mscmf.h -- C ABI of the MI355X exact max-flow library (libmscmf.so).
 *
 * Plain C types only: pointers, sizes, int32/int64.  Every call returns an
 * int status (0 = ok, < 0 = error, message from mf_last_error()); nothing
 * aborts.  The caller owns host arrays (the library copies them); the
 * library owns device memory.  Handles are independent: separate engines
 * may be driven from separate threads.
 *
 * Citations are to xiyuweiguang/MScResearch,
 * Region_Aware_PR_with_Pruning_BoundaryCore.cu ("BC:line"), the path these
 * entry points replace.  The reference has no FFI of its own; its in-process
 * boundary is run_gpu_static_with_state (BC:2119-2123) and its process
 * boundary is the `<in> <out> static|dynamic` command line (BC:2561-2933),
 * which the mscmf binary (mf_cli.cpp) reproduces on top of this ABI.
 */
#ifndef MSCMF_H
#define MSCMF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define MF_OK 0
#define MF_EINVAL (-1)
#define MF_ENOMEM (-2)
#define MF_EHIP (-3)
#define MF_EIO (-4)
#define MF_EITER (-5)  /* iteration watchdog (BC:1893-1896 returned silently) */
#define MF_ECERT (-6)  /* min-cut certificate failed */
#define MF_ENODEV (-7)
#define MF_EBUSY (-8)
#define MF_ESTALE (-9) /* a binary graph cache no longer matches its text file or region count */

/* kernel families of the work counters and the profile */
#define MF_FAMILIES 7

typedef struct mf_hgraph mf_hgraph; /* host CSR + partition */
typedef struct mf_edges mf_edges;   /* generated edge list */
typedef struct mf_graph mf_graph;   /* device-resident region-ordered graph */
typedef struct mf_engine mf_engine; /* slots + scheduler + stream */

const char* mf_last_error(void);
int mf_version(void);
/* sizeof(mf_params) and sizeof(mf_result) as this library was built: a binding
 * checks its own layouts against them. */
void mf_struct_sizes(int64_t* params_bytes, int64_t* result_bytes);

/* ----------------------------------------------------------- host graph */

/* Parse graph.txt and build the CSR with reverse entries.
 * Replaces build_graph (BC:1044-1126): header "n m", up to m lines "u v cap",
 * stop at the first malformed line, skip out-of-range ids, self-loops and
 * cap <= 0; forward then reverse entry per line, rev links the pair. */
int mf_hgraph_load(const char* graph_path, mf_hgraph** out);

/* Same CSR from 0-based edge arrays (no skip rules applied beyond them). */
int mf_hgraph_from_edges(int n, int64_t cnt, const int32_t* us, const int32_t* vs,
                         const int64_t* caps, mf_hgraph** out);
int mf_hgraph_from_edgelist(const mf_edges* e, mf_hgraph** out);

/* Capacity-degree partition + region-ordered relabel, in place.
 * Replaces build_capacity_degree_partition_from_stats (BC:1130-1178),
 * partition_and_reorder (BC:1182-1260) and build_region_id (BC:1263-1276).
 * Score ties are broken by vertex id (the reference used an unstable sort). */
int mf_hgraph_partition(mf_hgraph* h, int R);

int mf_hgraph_info(const mf_hgraph* h, int* n, int64_t* E, int* R, int64_t* lines);
/* Copy arrays out; any pointer may be NULL.  Sizes: row_ptr n+1, col/cap/rev E,
 * new_id/old_id/region_id n, rstart R+1. */
int mf_hgraph_export(const mf_hgraph* h, int32_t* row_ptr, int32_t* col, int64_t* cap,
                     int32_t* rev, int32_t* new_id, int32_t* old_id, int32_t* rstart,
                     int32_t* region_id);
/* Per-vertex stats of the input lines (BC:1090-1096), in the current order. */
int mf_hgraph_stats(const mf_hgraph* h, int64_t* sum_out, int64_t* sum_in, int32_t* deg_out,
                    int32_t* deg_in);
void mf_hgraph_free(mf_hgraph* h);

/* Binary graph cache (no counterpart in the reference, which parses
 * graph.txt and partitions on every run, BC:1044-1276): the prepared graph
 * (CSR, stats, partition and id maps) as one file of 64-byte aligned
 * sections behind a 256-byte header, read back through mmap.
 * mf_hgraph_save records source_path's size and modification time when it
 * is not NULL.  mf_hgraph_open returns MF_ESTALE when source_path is given
 * and differs from the recorded one, or R >= 0 differs from the recorded
 * region count; MF_EIO on a missing, truncated or corrupted file (a 64-bit
 * hash of the sections is checked). */
int mf_hgraph_save(const mf_hgraph* h, const char* path, const char* source_path);
int mf_hgraph_open(const char* path, const char* source_path, int R, mf_hgraph** out);

/* CSR entry of each accepted input line's forward edge, in line order:
 * us/vs are the lines the graph was built from (mf_hgraph_from_edges or the
 * generator's list), in the ids they were given in; the entries are those of
 * the graph's current (region-ordered) CSR.  MF_EINVAL if the lines are not
 * the graph's.  The bulk form of the reference's per-update search for the
 * entry src->dest (OnAdd_kernel, BC:979-987), for mf_graph_update_caps. */
int mf_hgraph_line_entries(const mf_hgraph* h, int64_t cnt, const int32_t* us, const int32_t* vs,
                           int32_t* entry);

/* Reachability classes behind the per-query pruning mask (BC:1391-1445):
 * scc[n] = strongly connected component over entries with cap > 0, wcc[n] =
 * weakly connected component.  The mask of (s,t) with t reachable from s is
 * {v : v reachable from scc[s]} AND {v : wcc[v] == wcc[t]}. */
int mf_hgraph_components(const mf_hgraph* h, int32_t* scc, int32_t* wcc);

/* The reference driver's decisions for (s, t) (ids of the partitioned h),
 * reported by the CLI's log only (hazard: the reference's boundary core drops
 * interior vertices that can carry s-t flow, so this library never takes
 * it): the pruning mask keeps *pruned_n vertices (BC:1391-1445); when that
 * is below 0.995 n (BC:2690) the induced subgraph (*run_n vertices) is
 * re-partitioned and its boundary-aware core (BC:2340-2475, hop limit 2)
 * keeps *core_kept; *would_take = 1 when the reference takes that core
 * (*core_kept < 0.98 *run_n, BC:2743). */
int mf_hgraph_core_decision(const mf_hgraph* h, int s, int t, int64_t* pruned_n, int64_t* run_n,
                            int64_t* core_kept, int* would_take);

/* ----------------------------------------------- SCC contraction and routing
 * Replaces the reference's offline tools SCC_CONTRACT/scc_contract.cpp
 * (Kosaraju ids, scc_contract.cpp:44-105; contraction with capacities summed
 * per component pair, :107-134), split_queries_by_scc.cpp:49-76 and
 * convert_test_to_scc.cpp:47-60.  Component ids are the reference's, so
 * maps, contracted graphs and split files match it byte for byte (bin/
 * mscmf_scc writes them).  The contracted graph's flows are NOT exact: the
 * inside of a component is taken as uncapacitated.  mf_scc_route is the
 * exact use of the components. */
typedef struct mf_scc mf_scc;
/* From 0-based lines as the graph reader accepts them (out-of-range ids,
 * self-loops and cap <= 0 dropped; line order matters for the ids). */
int mf_scc_build(int n, int64_t m, const int32_t* us, const int32_t* vs, const int64_t* caps, mf_scc** out);
/* From graph.txt, read as mf_hgraph_load reads it (scc_contract.cpp:18-40). */
int mf_scc_load(const char* graph_path, mf_scc** out);
int mf_scc_info(const mf_scc* h, int* n, int32_t* ncomp, int64_t* m_contracted);
/* comp[n]: 0-based component id per vertex (the reference's .map is v+1, comp+1). */
int mf_scc_comp(const mf_scc* h, int32_t* comp);
/* The contracted graph, m_contracted entries sorted by (cu, cv), 0-based. */
int mf_scc_edges(const mf_scc* h, int32_t* cu, int32_t* cv, int64_t* cap);
/* Exact routing (0-based ids): route[i] = 0 when t is not reachable from s
 * (max flow 0), 1 when s and t share a component C (every s-t path stays in
 * C: solve on C's induced subgraph), 2 when t is reachable in another
 * component (solve on the original graph), -1 for an id out of range or
 * s == t.  The reference's split sends same-component queries to the
 * original graph and the rest to the contracted one. */
int mf_scc_route(const mf_scc* h, int nq, const int32_t* s, const int32_t* t, int32_t* route);
void mf_scc_free(mf_scc* h);

/* Stand-alone partition on given stats (BC:1130-1178), for parity tests. */
int mf_partition(int n, int R, const int64_t* sum_out, const int64_t* sum_in,
                 const int32_t* deg_out, const int32_t* deg_in, int32_t* part);

/* test.txt reader (BC:2603-2621): 1-based pairs until a read fails; returns
 * 0-based ids, unvalidated.  *count = pairs read (<= cap). */
int mf_read_queries(const char* path, int64_t cap, int32_t* s, int32_t* t, int64_t* count);

/* -------------------------------------------------------------- generators
 * Seeded splitmix64 -> xoshiro256** streams; identical output on any host. */
int mf_gen_uniform(int n, int64_t m, int64_t cap_max, uint64_t seed, mf_edges** out);
int mf_gen_grid(int rows, int cols, int64_t cap_max, uint64_t seed, mf_edges** out);
int mf_gen_rmat(int scale, int64_t samples, double a, double b, double c, int64_t cap_max,
                uint64_t seed, int threads, mf_edges** out);
int mf_edges_info(const mf_edges* e, int* n, int64_t* cnt);
int mf_edges_copy(const mf_edges* e, int32_t* us, int32_t* vs, int64_t* caps);
int mf_edges_write(const mf_edges* e, const char* graph_path);
void mf_edges_free(mf_edges* e);
int mf_gen_queries_uniform(int n, int q, uint64_t seed, int32_t* s, int32_t* t);
int mf_gen_queries_grid(int rows, int cols, int q, uint64_t seed, int32_t* s, int32_t* t);
int mf_gen_queries_hub(const mf_edges* e, double top_frac, int q, uint64_t seed, int32_t* s,
                       int32_t* t);
int mf_write_queries(const char* path, int q, const int32_t* s, const int32_t* t);

/* ---------------------------------------------------------- device graph */

int mf_device_count(int* count);
/* One empty kernel (k_trace_marker) on the current device, waited for: a
 * mark by which a kernel trace or a counter run is split into phases (no
 * counterpart in the reference). */
int mf_trace_marker(void);

/* Upload a (partitioned) host graph; regions are cut into tiles of at most
 * tile_max vertices (<= 4096), one workgroup per tile in region mode.
 * Replaces init_global_graph_on_device (BC:1293-1332); residency is keyed by
 * the handle, never by (V,E,R) sizes. */
int mf_graph_upload(const mf_hgraph* h, int tile_max, mf_graph** out);
int mf_graph_info(const mf_graph* g, int* n, int64_t* E, int* ntiles, int64_t* bytes);
/* Device bytes of one query slab for this graph: what an engine's slab
 * arena (mf_engine_set_budget) holds per concurrent query on it.  No
 * reference counterpart (the reference sizes its one state per run). */
int mf_graph_slab_bytes(const mf_graph* g, int64_t* bytes);
void mf_graph_free(mf_graph* g);

/* ---------------------------------------------------------------- engine */

/* Scheduler parameters; mf_params_default() gives the reference's constants
 * (BC:1459-1484, BC:1725, BC:1936, BC:2644-2648). */
typedef struct mf_params {
    int K;               /* discharge cycles per vertex per pass, 0 = max(4, E/V) */
    int gr_freq;         /* GR every gr_freq iterations */
    int gr_freq_min, gr_freq_max;
    int gr_stall;        /* extra GR after excess(t) stalls this many iterations, 0 = off */
    int h_cut;
    int stagnate_rounds;
    int path_switch, path_len, path_attempts, p_count0;
    int hot, hot_k, hot_min_active, hot_rounds;
    int rm, rm_freq0, rm_freq_max, rm_margin;
    double rm_tau;
    int tile_rounds;
    double tile_frac;    /* region mode when active >= tile_frac * n */
    int cert;            /* certify every answer on the GPU: 3 (default) fingerprint over the pairs
                            pushes changed, 1 over every pair, 2 pair sums entry by entry, 0 off */
    long long iter_max;
    double bu_alpha, bu_beta;
    int prune;           /* per-query reachability pruning + induced subgraph */
    double prune_ratio;  /* keep the subgraph when kept < ratio * n */
    double gr_work;      /* also GR after gr_work * E discharge scans since the last one, 0 = off */
    int K_heavy;         /* discharge cycles per pass for wave-scanned (long) rows */
    int local_relabel;   /* 1 = BFS relabel of active tiles between global relabels */
    int local_band;      /* ... when the active height band is at most this wide */
    int fuse;            /* discharge + count cycles per host round trip on deep graphs (1 = off) */
    int graphs;          /* 1 = replay repeated fused cycles as captured HIP graphs */
    int spec_cert;       /* 1 = certify on the discharge's labels before a terminating exact GR */
    int sparse_count;    /* 1 = count actives from the previous list and the vertices pushed into */
    int small_wgs;       /* workgroups per slot of the multi-level small-frontier BFS (1 = one) */
    long long small_edges; /* ... while the frontier's rows sum to at most this many entries */
    int mark_list;       /* marked entries listed per slab for resets and the certificate (0 = map only) */
    int inject;          /* 1 = bounded injection: the preflow puts min(out-capacity of s, in-capacity
                            of t) on s, which is then an ordinary vertex, instead of saturating the row
                            of s (BC:173-200, 0); pinned (dynamic) queries always use 0 */
    int tile_threads;    /* threads of the tile (region) discharge workgroup: 256, 512 (BC:1725) or 1024 */
    int path_switch_deep, p_count_deep; /* path_switch and p_count0 once the last exact relabel is 64+ levels deep */
    int tile_split;      /* tile discharge: up to this many workgroups per tile by excess score (1..8, 1 = off) */
    int scc_route;       /* 1: a query with s and t in one SCC runs on that SCC's induced subgraph (exact) */
    int bu_skip;         /* 1: bottom-up relabel levels skip vertices that cannot reach t's SCC when it is the largest */
    int sink_cut;        /* 1: stop once excess(t) fills t's in-capacity or s's out-capacity, certified on that cut */
    int disc_band;       /* shallow graphs: discharge only active vertices within this many levels of the lowest (0 = all) */
    int bu_pick;         /* 1: bottom-up relabel levels try each vertex's entry to its longest-row head first */
    int small_levels;    /* BFS levels one single-workgroup (or multi-workgroup) BFS launch runs at most */
    int push_first;      /* 1: lane/wave discharge pushes first and reads the row minimum only to relabel */
    int slot_blocks, launch_blocks; /* grid-stride launches: workgroups per slot / per launch at most */
    int bu_light;        /* bottom-up rows up to this length are scanned by one lane, longer ones by a wave */
    int bu_bits;         /* 1: bottom-up levels test candidate parents in a frontier bitmap, 0: in their levels */
    double gr_work_deep; /* gr_work once the last exact relabel was 64+ levels deep */
    long long small_edges_deep; /* small_edges on deep graphs (relabel, depth hint or BFS 64+ levels deep) */
    int giant_rounds;    /* giant-row discharge chunks in rounds that stop once the row is dry: 1 on, 0 off,
                            2 = on in engines of more than one slot */
    int bu_recv;         /* 1: bottom-up relabel levels test rows no flow has entered or left through the
                            graph's forward-entry mask instead of their residuals */
    int dis_queue;       /* 1: a step's list discharge runs as one persistent launch taking every slot's
                            blocks off a device work queue; 0 (default): a grid per slot */
} mf_params;

void mf_params_default(mf_params* p);

/* Engine with nslots concurrent query slots; per-query residual slabs are
 * sized for the largest graph added.  Replaces DynamicState /
 * init_state_for_st (BC:47-113, BC:1336-1372). */
int mf_engine_create(int nslots, const mf_params* p, mf_engine** out);
int mf_engine_add_graph(mf_engine* e, mf_graph* g, int* gid);
void mf_engine_free(mf_engine* e);

/* Result of one query. */
typedef struct mf_result {
    int64_t tag;
    int64_t flow;
    int status;
    int cert_ok;
    int pruned;
    int n_run;
    int64_t E_run;
    int64_t kept;
    int64_t iters, steps, grs, bfs_levels, drains, rm_fixed;
    uint64_t edge_scans, vertex_scans, pushes, relabels;
    double t_arrive, t_start, t_end; /* seconds on the engine clock */
    /* work per kernel family (init, bfs, active, discharge, rm, cert,
     * bfs_bu) x (edge scans, vertex scans, pushes, relabels + BFS
     * discoveries); bfs = seeding, single-workgroup and top-down levels,
     * bfs_bu = bottom-up levels */
    uint64_t work[MF_FAMILIES][4];
    /* 1 when the answer was certified on the cut around t as soon as
     * excess(t) reached t's in-capacity (mf_params.sink_cut), else 0 */
    int64_t sink_cut;
} mf_result;

/* Plain single query: solve (s,t) (ids of graph gid) to completion.
 * Replaces run_gpu_static_with_state (BC:2119-2205) including per-query
 * pruning (BC:2666-2698): the int64 flow is exact and certified. */
int mf_solve(mf_engine* e, int gid, int s, int t, mf_result* r);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture elsewhere in the process, tight
 * serving loops, harnesses that forbid hipMalloc).  The same plain solve as
 * mf_solve, on the full graph (no pruning, no captured cycles), with the
 * query's slab carved from a static per-device array of
 * MF_NOALLOC_SCRATCH_BYTES instead of the engine's arena; the call makes no
 * device allocation.  The graph's slab (mf_graph_slab_bytes) must fit, else
 * MF_ENOMEM and nothing runs.  The engine must be idle (nothing queued,
 * live or pinned).  Calls on one GPU share its array, which keeps the last
 * call's contents (everything a solve reads is reset by the solve itself);
 * they would race, so the library runs them one at a time per device. */
#define MF_NOALLOC_SCRATCH_BYTES (256ll << 20)
int mf_solve_noalloc(mf_engine* e, int gid, int s, int t, mf_result* r);

/* Batched / served use: queue queries, then drive the engine.  Freed slots
 * are refilled from the queue every step (continuous batching). */
int mf_engine_submit(mf_engine* e, int gid, int s, int t, int64_t tag, double t_arrive);
/* Run steps until the queue and every slot are empty (max_steps < 0) or for
 * max_steps steps; finished results are copied to out (up to cap), *nout set. */
int mf_engine_run(mf_engine* e, int64_t max_steps, mf_result* out, int cap, int* nout);
int mf_engine_live(const mf_engine* e, int* live, int64_t* queued);
/* Slab arena of the engine: per-query device arrays (residual, excess,
 * heights, lists) are carved from one HBM block of `bytes` (0 = default:
 * nslots full slabs of the largest graph).  A queued query is admitted when
 * a slot is free and its run graph's slab fits: the pruned subgraph's when
 * that is already cached, else the full graph's (the reference sized
 * DynamicState per query, BC:1336-1372).  Only while no slot is in use. */
int mf_engine_set_budget(mf_engine* e, int64_t bytes);
/* Arena size, bytes in use, peak, and steps whose queue head waited for room. */
int mf_engine_mem_stats(const mf_engine* e, int64_t* budget, int64_t* in_use, int64_t* peak,
                        int64_t* admit_waits);
/* Serving loop: query i (graph gid[i], s[i], t[i]) arrives arrive[i] seconds
 * after the call (ascending).  Arrivals are queued when their time comes,
 * admitted by slab bytes as slots free up (continuous batching) and run
 * concurrently; the call returns when every query has finished.  Results in
 * completion order, tag = i, t_arrive/t_start/t_end on the engine clock.
 * The reference answered one query at a time (BC:2628-2843). */
int mf_engine_serve(mf_engine* e, int nq, const int32_t* gid, const int32_t* s, const int32_t* t,
                    const double* arrive, mf_result* out, int* nout);
/* Serving over several engines ("lanes"), typically one per resident graph
 * class: query i goes to engines[lane[i]] (graph gid[i] of that engine).
 * Each lane runs mf_engine_serve's loop on its own host thread and HIP
 * stream, so the lanes' kernels overlap and a deep query class (grids) never
 * holds back a shallow one.  Arrival offsets share one start; results in
 * completion order, tag = i, times on the clock all engines share. */
int mf_serve_lanes(int nengines, mf_engine* const* engines, int nq, const int32_t* lane, const int32_t* gid,
                   const int32_t* s, const int32_t* t, const double* arrive, mf_result* out, int* nout);
double mf_engine_now(const mf_engine* e);
int mf_engine_cache_stats(const mf_engine* e, int64_t* hits, int64_t* misses);
/* Fused cycles replayed from captured HIP graphs, and graphs captured. */
int mf_engine_graph_stats(const mf_engine* e, int64_t* captures, int64_t* replays);
/* Per kernel family (init, bfs, active, discharge, rm, cert, bfs_bu) timing with HIP
 * events on the engine's stream while on != 0, plus the algorithmic work of
 * the queries that finished meanwhile.  Turning it on clears the totals.
 * ms[MF_FAMILIES], launches[MF_FAMILIES], work[MF_FAMILIES*4]. */
int mf_engine_profile(mf_engine* e, int on);
int mf_engine_profile_read(const mf_engine* e, double* ms, int64_t* launches, uint64_t* work);

/* Per-step log, the information behind the reference's stderr lines
 * ("[Init] kernel_1(...)=... ms, kernel_7(...)=... ms" BC:2188, the
 * per-iteration timing line BC:2106): after every engine step the callback
 * gets one record per slot that ran something.  ms[] are HIP-event times of
 * the step's stages for all slots of the step, while profiling is on
 * (mf_engine_profile), else 0. */
#define MF_STEP_INIT 1u      /* residual initialisation (BC:151-170) */
#define MF_STEP_PREFLOW 2u   /* preflow from s (BC:173-200) */
#define MF_STEP_SEED 4u      /* relabel seeding (BC:203-233) */
#define MF_STEP_BFS 8u       /* relabel BFS levels (BC:322-348) */
#define MF_STEP_DISCHARGE 16u
#define MF_STEP_HOT 32u      /* hot-region sweep (BC:1898-1984) */
#define MF_STEP_RM 64u       /* RM repair (BC:605-657) */
#define MF_STEP_CERT 128u    /* min-cut certificate */
#define MF_STEP_PRUNE 256u   /* reachability pruning BFS (BC:1391-1445) */
#define MF_STEP_COUNT 512u   /* active count */
enum { MF_LOG_INIT = 0, MF_LOG_PREFLOW, MF_LOG_SEED, MF_LOG_BFS, MF_LOG_PR, MF_LOG_RM, MF_LOG_CERT, MF_LOG_FIELDS };
typedef struct {
    int64_t tag;           /* the slot's query */
    int32_t slot;
    uint32_t steps;        /* MF_STEP_* bits of what the slot ran */
    int64_t iters;         /* discharge iterations of the query so far */
    int32_t active, active_high;  /* after the step's count (MF_STEP_COUNT) */
    int64_t fixed_edges;   /* RM repairs of the step (MF_STEP_RM) */
    int32_t K, gr_freq;    /* the scheduler's current K and GR_FREQ */
    int32_t path;          /* 1 = the discharge was a path push */
    int32_t cycles;        /* discharge + count cycles of the step */
    double ms[MF_LOG_FIELDS];  /* init, preflow, seeding, BFS (incl. pruning), push-relabel (discharge, hot
                                  sweep, count), RM, certificate */
} mf_step_log;
typedef void (*mf_step_fn)(void* user, const mf_step_log* log);
/* fn = NULL turns the log off. */
int mf_engine_set_step_log(mf_engine* e, mf_step_fn fn, void* user);

/* ------------------------------------------------------------- autotuner
 * The reference uses R = 128, K = max(4, E/V), GR_FREQ0 = 128 and a
 * 512-thread region kernel for every graph (BC:2583, BC:1465-1466,
 * BC:1725).  mf_autotune runs one coordinate-descent pass over R in {128,
 * 256, 512, 1024}, tile size in {1024, 2048, 4096} (vertices one discharge
 * workgroup stages in LDS), K in {auto, 1, 2, 4, 8}, GR_FREQ in {64, 128,
 * 256, 512}, the tile discharge workgroup size in {256, 512, 1024} and, for
 * batch > 0, slots per engine in {batch/2, batch, 2 batch}, from (R0, tile0,
 * base->K, base->gr_freq, base->tile_threads, batch, base->tile_split).  batch = 0 times the
 * plain solver (one query at a time); batch > 0 times the batched engine
 * with every sample query submitted at once.  Every configuration must
 * reproduce the first one's flows.  h: the host graph before partitioning;
 * s, t: sample queries in its ids.  cache_path (or NULL): results per graph
 * class ("n<log2 n>-d<log2 avg degree>-s<log2 skew>-k<log2 BFS depth>") and
 * mode are read from and appended to this text file. */
typedef struct mf_tune_result {
    int R, tile_max, K, gr_freq;
    int tile_threads;  /* tile discharge workgroup size */
    int slots;         /* batched tuning: slots per engine; 0 for the plain solver */
    int tile_split;    /* tile discharge workgroups per tile by excess score */
    double ms_per_query, default_ms_per_query;
    int evaluations; /* 0 = found in the cache */
    char graph_class[64];
} mf_tune_result;
int mf_autotune(const mf_hgraph* h, int nq, const int32_t* s, const int32_t* t, const mf_params* base, int R0,
                int tile0, int batch, const char* cache_path, mf_tune_result* out);
/* The graph class mf_autotune keys its cache by, into out (cap bytes). */
int mf_graph_class(const mf_hgraph* h, char* out, int cap);

/* ------------------------------------------------------- dynamic mode
 * Pinned slots keep their residual across capacity updates (the reference's
 * DynamicState reuse, BC:2230-2335).  Pinned queries run on the full graph. */

/* Pin (s,t) of graph gid to slot `slot`; it is not refilled by the queue. */
int mf_engine_pin(mf_engine* e, int slot, int gid, int s, int t, int64_t tag);
/* Solve every pinned slot, concurrently.  warm = 0: from scratch (init +
 * preflow); warm = 1: from the residual left by the last solve and updates
 * (preflow again, exact GR, drain of any deficits).  A warm answer that
 * fails its certificate is re-solved cold (counted in mf_result.drains < 0). */
int mf_engine_solve_pinned(mf_engine* e, int warm, mf_result* out, int cap, int* nout);
/* Capacity updates on graph gid.  entries are CSR entry ids of forward
 * edges (cap > 0), new_cap >= 1, no repeats; a batch breaking any of these
 * returns MF_EINVAL and changes nothing.  Engines sharing the graph must not
 * be solving on it during the call.  Every pinned residual on gid
 * is adjusted: an increase adds residual; a decrease below the edge's flow
 * sends the surplus back (excess at the tail, deficit at the head).
 * Replaces OnAdd_kernel (BC:959-1000), which handled increases only. */
int mf_graph_update_caps(mf_engine* e, int gid, int n, const int32_t* entries,
                         const int64_t* new_cap);
/* First forward entry u->v (region-ordered ids) of a host graph, or -1. */
int mf_hgraph_find_entry(const mf_hgraph* h, int u, int v, int32_t* entry);

/* ------------------------------------------------- parity hooks (tests) */

/* Residual init + source preflow on slot 0 (BC:151-200); copies res[E] and
 * excess[n] out. */
int mf_dbg_init_preflow(mf_engine* e, int gid, int s, int t, int64_t* res, int64_t* excess);
/* Exact global relabel on given residual/excess (BC:203-233, BC:322-348);
 * copies heights out (n = unreached). */
int mf_dbg_global_relabel(mf_engine* e, int gid, int s, int t, const int64_t* res,
                          const int64_t* excess, int32_t* height);
/* Pruning mask on the GPU (BC:1391-1445); mask[n] 0/1, *kept. */
int mf_dbg_prune_mask(mf_engine* e, int gid, int s, int t, int32_t* mask, int* kept);
/* The pruning mask the product builds (BC:1391-1445 by reachability class):
 * the forward set of s's SCC (from a forward BFS on the GPU, or the graph's
 * cache) AND the WCC of t, mask[n] 0/1 and *kept, when t is reachable from s
 * (*reachable = 1); *reachable = 0 means the flow is 0 and no mask is made. */
int mf_dbg_class_mask(mf_engine* e, int gid, int s, int t, int32_t* mask, int64_t* kept, int* reachable);
/* Induced subgraph on the GPU for a given mask (BC:2477-2551). */
int mf_dbg_induced(mf_engine* e, int gid, const int32_t* mask, int* n_out, int64_t* E_out,
                   int32_t* row_ptr, int32_t* col, int64_t* cap, int32_t* rev,
                   int32_t* old2sub);
/* Local relabel (BC:236-279, BC:1582-1721) on given residual, excess and
 * heights, with the tiles flagged in tile_active[ntiles] as the active
 * region and level cap lcap (1..256).  Heights are updated in place; *level
 * gets the first level the bounded BFS did not complete. */
int mf_dbg_local_relabel(mf_engine* e, int gid, int s, int t, const int64_t* res,
                         const int64_t* excess, int32_t* height, const int32_t* tile_active,
                         int lcap, int* level);
/* Min-cut certificate (mode 1 fingerprint, 2 entry by entry, 3 fingerprint
 * over the pairs whose residual left their capacities, 4 = 3 reading the cut
 * entries from the rows of V\T instead of T's) of a given residual, excess
 * and heights (T = {h < n}) on slot 0: violation count, cut(V\T, T),
 * excess(t) and whether the pair-sum fingerprint is 0. */
int mf_dbg_certify(mf_engine* e, int gid, int s, int t, const int64_t* res, const int64_t* excess,
                   const int32_t* height, int mode, int64_t* bad, int64_t* cut, int64_t* excess_t,
                   int* fp_zero);
/* Hot-region choice of the hot sweep (BC:1931-1946: the k regions with the
 * most active vertices) on given per-tile active counts (ntiles of the
 * graph): tile_hot[i] = 1 for the k tiles first in (count descending, index
 * ascending) order among tiles with a nonzero count. */
int mf_dbg_hot_select(mf_engine* e, int gid, const int32_t* tile_cnt, int k, int32_t* tile_hot);
/* Wait bound, in wall-clock ticks, of the engine's multi-workgroup BFS
 * barriers (ticks <= 0: the default of 2 s).  A barrier that waits longer
 * flags its slot and the BFS restarts single-workgroup; a tiny bound forces
 * that path from the host.  *restarts (may be NULL) gets the restarts of the
 * queries the engine has retired so far.  Call it while the engine is idle
 * (nothing live).  No counterpart in the reference, whose BFS is one launch
 * per level (BC:1538-1572). */
int mf_dbg_barrier_ticks(mf_engine* e, long long ticks, long long* restarts);
/* Solve without pruning on slot 0 and copy the final residual, excess and
 * heights out (for an independent CPU certificate). */
int mf_dbg_solve_state(mf_engine* e, int gid, int s, int t, mf_result* r, int64_t* res,
                       int64_t* excess, int32_t* height);

#ifdef __cplusplus
}
#endif
#endif
