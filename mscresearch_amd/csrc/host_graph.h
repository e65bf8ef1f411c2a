// host_graph.h -- host-side graph preparation for the MI355X max-flow path.
//
// Parsing, CSR-with-reverse-entries construction, the capacity-degree region
// partition and the region-ordered relabelling.  Everything here is exact
// integer work except the partition score (double, used only to order
// vertices).  Reference: Region_Aware_PR_with_Pruning_BoundaryCore.cu
// (BC:1044-1276).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace mf {

struct HostGraph {
    int n = 0;                 // vertices
    int64_t lines = 0;         // accepted input lines
    std::vector<int32_t> row_ptr;  // n+1
    std::vector<int32_t> col;      // E
    std::vector<int64_t> cap;      // E
    std::vector<int32_t> rev;      // E
    // per-vertex stats over original forward edges (BC:1090-1096)
    std::vector<int64_t> sum_out, sum_in;
    std::vector<int32_t> deg_out, deg_in;
    // partition state; empty until partition_and_reorder() runs
    int R = 0;
    std::vector<int32_t> new_id, old_id;   // original <-> region-ordered ids
    std::vector<int32_t> rstart;           // R+1 region boundaries (region-ordered ids)
    std::vector<int32_t> region_id;        // n
    int64_t E() const { return (int64_t)col.size(); }
};

// Errors are reported as negative codes plus a message in err.
// graph.txt reader (BC:1044-1076): header "n m", then up to m lines
// "u v cap"; stops at the first malformed line; skips out-of-range ids,
// self-loops and cap<=0.
int parse_graph_file(const std::string& path, std::vector<int32_t>& us, std::vector<int32_t>& vs,
                     std::vector<int64_t>& caps, int& n, std::string& err);

// test.txt reader (BC:2603-2621): 1-based pairs until a read fails; returns
// 0-based ids (unmapped, may be out of range).
int parse_query_file(const std::string& path, std::vector<int32_t>& s, std::vector<int32_t>& t,
                     std::string& err);

// CSR with one forward (cap c) and one reverse (cap 0) entry per edge, in
// edge order (BC:1078-1125).
void build_csr(int n, const std::vector<int32_t>& us, const std::vector<int32_t>& vs,
               const std::vector<int64_t>& caps, HostGraph& g);

// Stats recomputed from CSR entries with cap>0 (BC:2707-2720).
void csr_stats(const HostGraph& g, std::vector<int64_t>& so, std::vector<int64_t>& si,
               std::vector<int32_t>& dout, std::vector<int32_t>& din);

// Capacity-degree partition (BC:1130-1178); ties in score broken by id.
void capacity_degree_partition(int n, int R, const std::vector<int64_t>& so,
                               const std::vector<int64_t>& si, const std::vector<int32_t>& dout,
                               const std::vector<int32_t>& din, std::vector<int32_t>& part);

// Region-contiguous relabel (BC:1182-1276).  Rewrites g in place to the
// region-ordered CSR and fills new_id/old_id/rstart/region_id.
void partition_and_reorder(HostGraph& g, int R, const std::vector<int32_t>& part);

// Strongly connected components over entries with cap > 0 (iterative
// Tarjan) and weakly connected components over all entries (union-find).
// Two sources in one SCC reach the same vertex set, and the undirected
// component of t is its WCC, so the pruning mask of BC:1391-1445 is a
// function of (scc[s], wcc[t]).
void strong_components(const HostGraph& g, std::vector<int32_t>& comp);
void weak_components(const HostGraph& g, std::vector<int32_t>& comp);

// What the reference's driver would do with (s, t) (BC:2666-2753), for a
// diagnostic only: its pruning mask (forward reach of s over cap > 0 AND the
// undirected component of t, plus s and t, BC:1391-1445) keeps pruned_n
// vertices; when that is below 0.995 n it builds the induced subgraph
// (BC:2477-2551, run_n vertices), re-partitions it into min(R, run_n) regions
// (BC:2705-2731) and computes its boundary-aware core (regions on an s-t
// region path, their boundary vertices and the regions of s and t, grown by
// hop_limit in-region hops, BC:2340-2475), which it takes when core_kept <
// 0.98 run_n.  The build never takes that core: it drops interior vertices
// that can carry s-t flow.  s, t: ids of g (partitioned).
struct CoreDecision {
    int64_t pruned_n = 0, run_n = 0, core_kept = 0;
    bool pruned = false, would_take = false;
};
CoreDecision reference_core_decision(const HostGraph& g, int s, int t, int hop_limit = 2);

// ------------------------------------------------------------ generators
// Seeded, platform-independent generators (splitmix64 -> xoshiro256**).
// Each writes an edge list with 0-based ids in emission order.
void gen_uniform(int n, int64_t m, int64_t cap_max, uint64_t seed, std::vector<int32_t>& us,
                 std::vector<int32_t>& vs, std::vector<int64_t>& caps);
void gen_grid(int rows, int cols, int64_t cap_max, uint64_t seed, std::vector<int32_t>& us,
              std::vector<int32_t>& vs, std::vector<int64_t>& caps);
// R-MAT (a,b,c,d) with 16*2^scale samples, self-loops dropped, duplicates
// merged by summing log-uniform [1,cap_max] capacities; output sorted by (u,v).
void gen_rmat(int scale, int64_t samples, double a, double b, double c, int64_t cap_max,
              uint64_t seed, int threads, std::vector<int32_t>& us, std::vector<int32_t>& vs,
              std::vector<int64_t>& caps);

// Query generators (0-based ids).
void gen_queries_uniform(int n, int q, uint64_t seed, std::vector<int32_t>& s,
                         std::vector<int32_t>& t);
void gen_queries_grid(int rows, int cols, int q, uint64_t seed, std::vector<int32_t>& s,
                      std::vector<int32_t>& t);
// s from the top `top_frac` of vertices by out-degree (ties by id), t uniform
// over vertices with in-degree > 0, s != t.
void gen_queries_hub(int n, const std::vector<int32_t>& us, const std::vector<int32_t>& vs,
                     double top_frac, int q, uint64_t seed, std::vector<int32_t>& s,
                     std::vector<int32_t>& t);

// ------------------------------------------------------- binary graph cache
// The prepared graph (CSR, stats, partition) as one file: a 256-byte header,
// then 64-byte aligned sections read back through mmap.  The header records
// the text file it was built from (size, modification time in ns) and a
// 64-bit hash of the sections.  save_graph_binary: 0 or -1 (err).
// load_graph_binary: 0; 1 when `source` is given and its size or time differ,
// or R >= 0 differs from the recorded region count (stale); -1 on a missing,
// truncated or corrupted file (err).
int save_graph_binary(const HostGraph& g, const std::string& path, const std::string& source,
                      std::string& err);
int load_graph_binary(const std::string& path, const std::string& source, int R, HostGraph& g,
                      std::string& err);

int write_graph_file(const std::string& path, int n, const std::vector<int32_t>& us,
                     const std::vector<int32_t>& vs, const std::vector<int64_t>& caps);
int write_query_file(const std::string& path, const std::vector<int32_t>& s,
                     const std::vector<int32_t>& t);

}  // namespace mf
