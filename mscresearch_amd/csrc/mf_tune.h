// mf_tune.h -- autotuner of R, tile size, K, GR_FREQ, the tile discharge
// workgroup size, workgroups per tile by excess score and slots per engine,
// per graph class.
#pragma once
#include <string>
#include <vector>

#include "host_graph.h"
#include "mf_engine.h"

namespace mf {

struct TuneResult {
    int R = 0, tile_max = 0, K = 0, gr_freq = 0;
    int tile_threads = 512;         // tile discharge workgroup size
    int slots = 0;                  // batched: engine slots (0 = tuned for the plain solver)
    int tile_split = 1;             // tile discharge workgroups per tile by excess score
    double ms = 0, ms_default = 0;  // ms per sample query: tuned, defaults
    int evaluations = 0;            // configurations timed (0: taken from the cache)
    std::string cls;
};

// "n<log2 n>-d<log2 avg degree>-s<log2 max/avg degree>-k<log2 BFS depth>"
std::string graph_class(const HostGraph& g);

// g: the graph before partitioning; s, t: sample queries (its ids).  One
// coordinate-descent pass over R, tile size, K, GR_FREQ, tile workgroup size,
// workgroups per tile by excess score and (batch > 0) slots in {batch/2,
// batch, 2 batch} from (R0, tile0, base.K, base.gr_freq, base.tile_threads,
// base.tile_split, batch); batch = 0 times the plain solver.  Every
// configuration must give the first one's flows.  cache_path (may be empty):
// "class:mode R tile K gr_freq threads slots ms ms_default [split]" lines,
// read first and appended to (a line without split means 1).
int autotune(const HostGraph& g, const std::vector<int32_t>& s, const std::vector<int32_t>& t, const Params& base,
             int R0, int tile0, int batch, const std::string& cache_path, TuneResult& out);

}  // namespace mf
