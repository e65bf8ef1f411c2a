// scc.h -- strongly connected components of the input lines, their
// contraction, and exact query routing by component.
//
// The reference's offline tools (SCC_CONTRACT/scc_contract.cpp,
// split_queries_by_scc.cpp, convert_test_to_scc.cpp) contract every SCC of
// graph.txt to one vertex and send cross-component queries to the
// contracted graph.  Component ids here are the reference's (Kosaraju,
// scc_contract.cpp:44-105), so its .map, contracted graph and split files
// come out byte for byte.  Flows on the contracted graph are NOT exact (the
// inside of a component is taken as uncapacitated); the exact use of the
// components is routing: t unreachable from s (max flow 0), s and t in one
// component C (every s-t path stays in C), or across components.
#pragma once
#include <cstdint>
#include <vector>

namespace mf {

struct SccContract {
    int n = 0;
    int32_t ncomp = 0;
    std::vector<int32_t> comp;     // vertex -> component id (0-based)
    // contracted graph: one entry per ordered component pair with an input
    // line between them, capacities summed, sorted by (cu, cv)
    std::vector<int32_t> cu, cv;
    std::vector<int64_t> ccap;
    std::vector<int64_t> crow;     // ncomp+1 offsets of cu's entries
};

// Lines as the graph reader accepts them (0-based, out-of-range ids,
// self-loops and cap <= 0 already dropped; line order kept: the depth-first
// search scans each vertex's lines in that order, which fixes the ids).
void scc_contract(int n, int64_t m, const int32_t* u, const int32_t* v, const int64_t* cap, SccContract& out);

// Per query: 0 = t not reachable from s (the max flow is 0), 1 = s and t in
// one component, 2 = t reachable from s in another component, -1 = an id out
// of [0, n) or s == t.  Component ids are a topological order of the
// condensation (every contracted entry has cu < cv), so a search from s's
// component stops at t's id; queries sharing a source component share it.
void scc_route(const SccContract& c, int nq, const int32_t* s, const int32_t* t, int32_t* route);

}  // namespace mf
