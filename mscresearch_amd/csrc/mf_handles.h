// mf_handles.h -- the opaque handles of include/mscmf.h as this library
// defines them, shared by the translation units of the C ABI.
#pragma once
#include <memory>
#include <vector>

#include "../../include/mscmf.h"
#include "host_graph.h"
#include "mf_engine.h"
#include "scc.h"

struct mf_hgraph {
    mf::HostGraph g;
};
struct mf_edges {
    int n = 0;
    std::vector<int32_t> us, vs;
    std::vector<int64_t> caps;
};
struct mf_graph {
    std::shared_ptr<mf::DeviceGraph> dg;
};
struct mf_scc {
    mf::SccContract c;
};
struct mf_engine {
    mf::Engine eng;
    std::vector<mf::Result> done;  // results not yet handed out
};

namespace mf {
// Engine result -> ABI result.
void result_to_c(const Result& r, mf_result* o);
}  // namespace mf
