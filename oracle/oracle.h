/*
 * oracle.h -- CPU restatement of the reference's exact max-flow path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker.
 * The product (mscresearch_amd/) never links, loads or calls anything here.
 *
 * Reference citations are to Region_Aware_PR_with_Pruning_BoundaryCore.cu
 * ("BC:line") of xiyuweiguang/MScResearch.
 *
 * Pinning (what checks each function besides retyping it):
 *   orc_parse_*          hand-written fixtures (skip/stop rules), E == 2*accepted
 *   orc_build_csr        rev[rev[e]]==e, col[rev[e]] is the row owner, fixtures
 *   orc_partition        closed-form scores on tiny graphs, greedy min-load property
 *   orc_reorder          permutation / region-contiguity / isomorphism invariants
 *   orc_prune_mask       brute force (transitive closure) on tiny graphs
 *   orc_induced          brute force edge filter on tiny graphs
 *   orc_core_mask        brute force restatement on tiny graphs
 *   orc_bfs_heights      brute-force all-pairs distances on tiny graphs
 *   orc_dinic            networkx / scipy golden fixtures, closed forms,
 *                        brute-force min cut (2^n subsets) on tiny graphs
 *   orc_certify          planted mistakes (tests corrupt a residual on the CPU)
 *   orc_scc_kosaraju     the reference tool built from its own source
 *                        (oracle/_ref/scc_contract, byte-identical .map),
 *                        scipy's strong components (same partition), ids in
 *                        a topological order of the condensation
 *   orc_scc_contract     the reference tool's graph_scc.txt, brute-force sums
 */
#ifndef MSC_ORACLE_H
#define MSC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* graph.txt header: "n m" (BC:1054-1058).  Returns 0, or -1 if unopenable,
 * -2 if the header is malformed. */
int orc_parse_header(const char* path, int* n, int64_t* m);

/* Reads up to m lines "u v cap" after the header; stops at the first failed
 * read, skips out-of-range ids, u==v and cap<=0 (BC:1066-1076).  Writes the
 * accepted edges, 0-based, into us/vs/caps (capacity m) and returns the count
 * (>=0) or a negative error. */
int64_t orc_parse_edges(const char* path, int64_t m, int32_t* us, int32_t* vs, int64_t* caps);

/* test.txt: pairs "s t" (1-based) until a read fails (BC:2609-2621).  Writes
 * 0-based ids; out-of-range entries become (-1,-1).  Returns count or -1. */
int64_t orc_parse_queries(const char* path, int64_t cap_q, int32_t* s, int32_t* t);

/* CSR with explicit reverse entries, one forward (cap c) and one reverse
 * (cap 0) per accepted line, in line order (BC:1078-1125).  row_ptr[n+1],
 * col/cap/rev[2*cnt]; stats arrays [n] (may be NULL). */
void orc_build_csr(int n, int64_t cnt, const int32_t* us, const int32_t* vs, const int64_t* caps,
                   int32_t* row_ptr, int32_t* col, int64_t* cap, int32_t* rev,
                   int64_t* sum_out, int64_t* sum_in, int32_t* deg_out, int32_t* deg_in);

/* Forward CSR entry (of orc_build_csr's layout, row_ptr given) of every
 * accepted line, in line order. */
void orc_line_entries(int n, int64_t cnt, const int32_t* us, const int32_t* vs, const int32_t* row_ptr,
                      int32_t* entry);

/* Per-vertex stats recomputed from a CSR, counting entries with cap>0
 * (the per-query variant, BC:2707-2720). */
void orc_csr_stats(int n, const int32_t* row_ptr, const int32_t* col, const int64_t* cap,
                   int64_t* sum_out, int64_t* sum_in, int32_t* deg_out, int32_t* deg_in);

/* Capacity-degree score and greedy least-loaded assignment (BC:1130-1178).
 * Ties in the descending score sort are broken by ascending vertex id. */
void orc_partition(int n, int R, const int64_t* sum_out, const int64_t* sum_in,
                   const int32_t* deg_out, const int32_t* deg_in, double* score, int32_t* part);

/* Region-contiguous renumbering (BC:1182-1260) and region_id (BC:1263-1276).
 * Outputs: new_id/old_id[n], rstart/rend[R], region_id[n], and the reordered
 * CSR (row_ptr2[n+1], col2/cap2/rev2[E]). */
void orc_reorder(int n, int R, const int32_t* part,
                 const int32_t* row_ptr, const int32_t* col, const int64_t* cap, const int32_t* rev,
                 int32_t* new_id, int32_t* old_id, int32_t* rstart, int32_t* rend, int32_t* region_id,
                 int32_t* row_ptr2, int32_t* col2, int64_t* cap2, int32_t* rev2);

/* mask = fwd(s; cap>0) AND undirected-component(t), plus {s,t} (BC:1391-1445). */
void orc_prune_mask(int n, const int32_t* row_ptr, const int32_t* col, const int64_t* cap,
                    int s, int t, int32_t* mask);

/* Induced subgraph on mask (BC:2477-2551).  old2sub[n], sub2old[n'] (capacity
 * n).  Returns n' and writes *m_out; row_ptr2 capacity n+1, edge arrays
 * capacity E. */
int orc_induced(int n, const int32_t* row_ptr, const int32_t* col, const int64_t* cap,
                const int32_t* rev, const int32_t* mask, int32_t* old2sub, int32_t* sub2old,
                int32_t* row_ptr2, int32_t* col2, int64_t* cap2, int32_t* rev2, int64_t* m_out);

/* Boundary-aware core mask (BC:2340-2475); the reference's non-conservative
 * shrink, restated for the drop-in's diagnostic log only. */
void orc_core_mask(int n, int R, const int32_t* row_ptr, const int32_t* col,
                   const int32_t* region_id, int s, int t, int hop_limit, int32_t* mask,
                   int64_t* kept_regions);

/* Residual init (BC:151-170): res = cap, excess = 0. */
void orc_init_residual(int n, int64_t E, const int64_t* cap, int64_t* res, int64_t* excess);

/* Source preflow (BC:173-200): saturate every s entry with res>0. */
void orc_preflow(int s, const int32_t* row_ptr, const int32_t* col, const int32_t* rev,
                 int64_t* res, int64_t* excess);

/* Global-relabel BFS heights (BC:203-233 seeding, BC:322-348 expansion):
 * h = n everywhere, level-0 seeds are t and every v!=s with excess<0 whose
 * mask is set (mask may be NULL = all ones); BFS walks residual edges
 * backwards (w joins when res(w->v)>0).  s is never labelled (its height
 * stays n).  Writes height[n]; returns the number of labelled vertices. */
int64_t orc_bfs_heights(int n, const int32_t* row_ptr, const int32_t* col, const int32_t* rev,
                        const int64_t* res, const int64_t* excess, const int32_t* mask,
                        int s, int t, int32_t* height);

/* Exact int64 max flow (Dinic).  If res_out is non-NULL the final residual
 * (pair-consistent with cap) is written there.  Returns the flow value. */
int64_t orc_dinic(int n, const int32_t* row_ptr, const int32_t* col, const int64_t* cap,
                  const int32_t* rev, int s, int t, int64_t* res_out);

/* Min-cut certificate on a final residual.  Checks: 0<=res, pair sums equal
 * cap pair sums, excess recomputed from res is >=0 off s, the set T that
 * reaches t in the residual excludes s and holds no excess besides t, and
 * excess[t] == cut(V\T,T) == flow.  Returns 0 if all hold, else a code
 * (1 capacity, 2 pair, 3 negative excess, 4 s reaches t, 5 excess in T,
 * 6 cut != excess[t], 7 excess[t] != flow, 8 excess mismatch). */
int orc_certify(int n, const int32_t* row_ptr, const int32_t* col, const int64_t* cap,
                const int32_t* rev, const int64_t* res, const int64_t* excess_or_null,
                int s, int t, int64_t flow);

/* Multi-threaded Dinic over a list of queries on one graph (one query per
 * thread).  Returns wall seconds; flows[q] written. */
double orc_dinic_many(int n, const int32_t* row_ptr, const int32_t* col, const int64_t* cap,
                      const int32_t* rev, int nq, const int32_t* s, const int32_t* t,
                      int threads, int64_t* flows);
/* The same; prune != 0 first restricts each query to the pruning mask of
 * orc_prune_mask (BC:1391-1445), answering 0 when t is not reachable from s.
 * Each thread restores only the residual entries its last query changed. */
double orc_dinic_many_ex(int n, const int32_t* row_ptr, const int32_t* col, const int64_t* cap,
                         const int32_t* rev, int nq, const int32_t* s, const int32_t* t,
                         int threads, int prune, int64_t* flows);

/* SCC tools (SCC_CONTRACT/scc_contract.cpp).  Kosaraju component ids of the
 * accepted lines (line order kept; it fixes the ids); returns the count. */
int orc_scc_kosaraju(int n, int64_t cnt, const int32_t* us, const int32_t* vs, int32_t* comp);
/* Lines between different components summed per (comp[u], comp[v]), sorted
 * by that pair; cu/cv/ccap hold up to cnt entries; returns the entry count. */
int64_t orc_scc_contract(int64_t cnt, const int32_t* us, const int32_t* vs, const int64_t* caps,
                         const int32_t* comp, int32_t* cu, int32_t* cv, int64_t* ccap);

#ifdef __cplusplus
}
#endif
#endif
