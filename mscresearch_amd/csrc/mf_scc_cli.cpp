// mf_scc_cli.cpp -- bin/mscmf_scc: the reference's SCC preprocessing tools
// behind one command, plus the exact routing they lack.
//
//   mscmf_scc contract <graph.txt> <graph_scc.txt>      (SCC_CONTRACT/scc_contract.cpp)
//   mscmf_scc convert  <test.txt> <map> <test_scc.txt>  (convert_test_to_scc.cpp)
//   mscmf_scc split    <test.txt> <map> <inter> <intra> (split_queries_by_scc.cpp)
//   mscmf_scc route    <graph.txt> <test.txt> <routes.txt>
//
// contract, convert and split write the reference's files byte for byte
// (graph_scc.txt and graph_scc.txt.map; the converted or split queries and
// <inter>.index) and its stderr lines.  Flows on the contracted graph are not
// exact.  route writes one line "<query> <zero|intra|inter|invalid>" per
// query: zero = t is not reachable from s (the max flow is 0), intra = s and
// t share a component (solve on it alone), inter = solve on the graph.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mscmf.h"

namespace {

int usage(const char* argv0) {
    std::fprintf(stderr,
                 "Usage: %s contract <graph.txt> <graph_scc.txt>\n"
                 "       %s convert <test.txt> <graph_scc.txt.map> <test_scc.txt>\n"
                 "       %s split <test.txt> <graph_scc.txt.map> <test_inter.txt> <test_intra.txt>\n"
                 "       %s route <graph.txt> <test.txt> <routes.txt>\n",
                 argv0, argv0, argv0, argv0);
    return 1;
}

struct File {
    FILE* f = nullptr;
    explicit File(FILE* x) : f(x) {}
    ~File() { if (f) std::fclose(f); }
};

int contract(const char* in, const char* out_path) {
    // the reference checks the header before reading lines
    {
        File f(std::fopen(in, "r"));
        if (!f.f) { std::fprintf(stderr, "Cannot open %s\n", in); return 1; }
        int n = 0, m = 0;
        if (std::fscanf(f.f, "%d %d", &n, &m) != 2) { std::fprintf(stderr, "Invalid header in %s\n", in); return 1; }
    }
    mf_scc* h = nullptr;
    if (mf_scc_load(in, &h) != MF_OK) { std::fprintf(stderr, "%s\n", mf_last_error()); return 1; }
    int n = 0;
    int32_t C = 0;
    int64_t mc = 0;
    mf_scc_info(h, &n, &C, &mc);
    std::fprintf(stderr, "[SCC] Graph loaded: n=%d\n", n);
    std::fprintf(stderr, "[SCC] found components: C=%d\n", C);
    std::vector<int32_t> cu((size_t)mc), cv((size_t)mc), comp((size_t)n);
    std::vector<int64_t> cap((size_t)mc);
    mf_scc_edges(h, cu.data(), cv.data(), cap.data());
    mf_scc_comp(h, comp.data());
    mf_scc_free(h);
    std::fprintf(stderr, "[SCC] contracted graph: n'=%d, m'=%d\n", C, (int)mc);
    File out(std::fopen(out_path, "w"));
    if (!out.f) { std::fprintf(stderr, "Cannot open %s for write\n", out_path); return 1; }
    std::fprintf(out.f, "%d %d\n", C, (int)mc);
    for (int64_t i = 0; i < mc; ++i) std::fprintf(out.f, "%d %d %lld\n", cu[i] + 1, cv[i] + 1, (long long)cap[i]);
    const std::string map_path = std::string(out_path) + ".map";
    File mp(std::fopen(map_path.c_str(), "w"));
    if (mp.f) {
        for (int v = 0; v < n; ++v) std::fprintf(mp.f, "%d %d\n", v + 1, comp[v] + 1);
        std::fprintf(stderr, "[SCC] mapping written to %s\n", map_path.c_str());
    }
    std::fprintf(stderr, "[SCC] Done.\n");
    return 0;
}

// "<vertex> <component>" pairs (1-based) until a read fails; comp_of[v] = -1
// for a vertex the map does not name; *max_v = the largest vertex named.
bool read_map(const char* path, std::vector<int>& comp_of, int* max_v) {
    File f(std::fopen(path, "r"));
    if (!f.f) return false;
    std::vector<std::pair<int, int>> kv;
    int v = 0, c = 0;
    *max_v = 0;
    while (std::fscanf(f.f, "%d %d", &v, &c) == 2) {
        kv.emplace_back(v, c);
        if (v > *max_v) *max_v = v;
    }
    comp_of.assign((size_t)*max_v + 1, -1);
    for (auto& p : kv)
        if (p.first >= 0) comp_of[(size_t)p.first] = p.second;
    return true;
}

int convert(const char* test, const char* map, const char* out_path) {
    std::vector<int> comp_of;
    int max_v = 0;
    if (!read_map(map, comp_of, &max_v)) { std::fprintf(stderr, "Cannot open %s\n", map); return 1; }
    File tf(std::fopen(test, "r"));
    if (!tf.f) { std::fprintf(stderr, "Cannot open %s\n", test); return 1; }
    File of(std::fopen(out_path, "w"));
    if (!of.f) { std::fprintf(stderr, "Cannot open %s for write\n", out_path); return 1; }
    int u = 0, v = 0;
    while (std::fscanf(tf.f, "%d %d", &u, &v) == 2) {
        const bool in = u > 0 && u <= max_v && v > 0 && v <= max_v;
        const int cu = in ? comp_of[(size_t)u] : 0, cv = in ? comp_of[(size_t)v] : 0;
        if (!in || cu <= 0 || cv <= 0) std::fprintf(of.f, "0 0\n");
        else std::fprintf(of.f, "%d %d\n", cu, cv);
    }
    std::fprintf(stderr, "[Convert] done, written to %s\n", out_path);
    return 0;
}

int split(const char* test, const char* map, const char* inter, const char* intra) {
    std::vector<int> comp_of;
    int max_v = 0;
    if (!read_map(map, comp_of, &max_v)) { std::fprintf(stderr, "Cannot open %s\n", map); return 1; }
    File tf(std::fopen(test, "r"));
    if (!tf.f) { std::fprintf(stderr, "Cannot open %s\n", test); return 1; }
    File fi(std::fopen(inter, "w")), fa(std::fopen(intra, "w"));
    if (!fi.f || !fa.f) { std::fprintf(stderr, "Cannot open output test files\n"); return 1; }
    File idx(std::fopen((std::string(inter) + ".index").c_str(), "w"));
    int u = 0, v = 0, q = 0, ni = 0, na = 0;
    while (std::fscanf(tf.f, "%d %d", &u, &v) == 2) {
        ++q;
        const bool in = u > 0 && u <= max_v && v > 0 && v <= max_v;
        const int cu = in ? comp_of[(size_t)u] : 0, cv = in ? comp_of[(size_t)v] : 0;
        if (in && cu > 0 && cv > 0 && cu == cv) {
            // one component: the original graph, original ids
            std::fprintf(fa.f, "%d %d\n", u, v);
            if (idx.f) std::fprintf(idx.f, "%d intra %d\n", q, ++na);
        } else {
            // across components: component ids; unknown ids pass through
            if (in && cu > 0 && cv > 0) std::fprintf(fi.f, "%d %d\n", cu, cv);
            else std::fprintf(fi.f, "%d %d\n", u, v);
            if (idx.f) std::fprintf(idx.f, "%d inter %d\n", q, ++ni);
        }
    }
    std::fprintf(stderr, "[Split] total queries=%d, inter=%d, intra=%d\n", q, ni, na);
    return 0;
}

int route(const char* graph, const char* test, const char* out_path) {
    mf_scc* h = nullptr;
    if (mf_scc_load(graph, &h) != MF_OK) { std::fprintf(stderr, "%s\n", mf_last_error()); return 1; }
    std::vector<int32_t> s(1 << 16), t(1 << 16);
    int64_t nq = 0;
    for (;;) {
        if (mf_read_queries(test, (int64_t)s.size(), s.data(), t.data(), &nq) != MF_OK) {
            std::fprintf(stderr, "%s\n", mf_last_error());
            mf_scc_free(h);
            return 1;
        }
        if (nq < (int64_t)s.size()) break;
        s.resize(s.size() * 2);
        t.resize(t.size() * 2);
    }
    std::vector<int32_t> r((size_t)nq);
    mf_scc_route(h, (int)nq, s.data(), t.data(), r.data());
    mf_scc_free(h);
    File of(std::fopen(out_path, "w"));
    if (!of.f) { std::fprintf(stderr, "Cannot open %s for write\n", out_path); return 1; }
    static const char* names[] = {"invalid", "zero", "intra", "inter"};
    int cnt[4] = {0, 0, 0, 0};
    for (int64_t i = 0; i < nq; ++i) {
        const int k = r[(size_t)i] + 1;
        ++cnt[k];
        std::fprintf(of.f, "%lld %s\n", (long long)(i + 1), names[k]);
    }
    std::fprintf(stderr, "[Route] total queries=%lld, zero=%d, intra=%d, inter=%d, invalid=%d\n", (long long)nq,
                 cnt[1], cnt[2], cnt[3], cnt[0]);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) return usage(argv[0]);
    const std::string cmd = argv[1];
    if (cmd == "contract" && argc >= 4) return contract(argv[2], argv[3]);
    if (cmd == "convert" && argc >= 5) return convert(argv[2], argv[3], argv[4]);
    if (cmd == "split" && argc >= 6) return split(argv[2], argv[3], argv[4], argv[5]);
    if (cmd == "route" && argc >= 5) return route(argv[2], argv[3], argv[4]);
    return usage(argv[0]);
}
