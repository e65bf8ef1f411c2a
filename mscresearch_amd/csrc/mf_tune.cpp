// mf_tune.cpp -- autotuner: region count R, tile size (the vertex range one
// discharge workgroup stages in LDS), local-cycle bound K, GR_FREQ, the
// tile discharge workgroup size and, for the batched engine, slots per
// engine, per graph class.
//
// The reference fixes R = 128 (BC:2583), K = max(4, E/V) (BC:1465),
// GR_FREQ0 = 128 (BC:1466) and 512 threads for its region kernel (BC:1725)
// for every graph.  Here one coordinate-descent pass over the knobs times
// the plain solver (one query at a time) or the batched engine (all sample
// queries at once) on sample queries of the graph; every candidate must
// reproduce the flows of the first (default) configuration, so tuning never
// changes an answer.  Results are kept per graph class (sizes, degree skew
// and BFS depth, bucketed) and mode in a text file.
#include "mf_tune.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <map>
#include <sstream>

namespace mf {

namespace {

double seconds_now() {
    using namespace std::chrono;
    return duration_cast<duration<double>>(steady_clock::now().time_since_epoch()).count();
}

int ilog2(double x) { return x <= 1.0 ? 0 : (int)std::floor(std::log2(x)); }

// Levels of an undirected BFS from the vertex of largest degree: the depth
// scale that decides how many discharge cycles an excess needs.
int bfs_depth(const HostGraph& g) {
    const int n = g.n;
    if (n == 0) return 0;
    int src = 0;
    for (int v = 1; v < n; ++v)
        if (g.row_ptr[v + 1] - g.row_ptr[v] > g.row_ptr[src + 1] - g.row_ptr[src]) src = v;
    std::vector<int32_t> lvl(n, -1), cur{src}, nxt;
    lvl[src] = 0;
    int L = 0;
    while (!cur.empty()) {
        nxt.clear();
        for (int v : cur)
            for (int e = g.row_ptr[v]; e < g.row_ptr[v + 1]; ++e) {
                const int w = g.col[e];
                if (lvl[w] < 0) { lvl[w] = L + 1; nxt.push_back(w); }
            }
        if (!nxt.empty()) ++L;
        cur.swap(nxt);
    }
    return L;
}

}  // namespace

std::string graph_class(const HostGraph& g) {
    const double n = std::max(1, g.n);
    const double avg = (double)g.E() / n;
    int maxdeg = 0;
    for (int v = 0; v < g.n; ++v) maxdeg = std::max(maxdeg, g.row_ptr[v + 1] - g.row_ptr[v]);
    std::ostringstream o;
    o << "n" << ilog2(n) << "-d" << ilog2(avg + 1) << "-s" << ilog2(maxdeg / std::max(1.0, avg) + 1) << "-k"
      << ilog2(bfs_depth(g) + 1);
    return o.str();
}

namespace {

struct Setup {
    int R = 0, tile = 0;
    std::shared_ptr<DeviceGraph> dg;
    std::vector<int32_t> s, t;  // queries in the partitioned ids
};

int prepare(const HostGraph& g0, int R, int tile, const std::vector<int32_t>& s0, const std::vector<int32_t>& t0,
            Setup& out) {
    HostGraph g = g0;
    std::vector<int32_t> part;
    capacity_degree_partition(g.n, R, g.sum_out, g.sum_in, g.deg_out, g.deg_in, part);
    partition_and_reorder(g, R, part);
    auto dg = std::make_shared<DeviceGraph>();
    const int rc = upload_graph(g, tile, *dg);
    if (rc != MF_OK) return rc;
    out.R = R;
    out.tile = tile;
    out.dg = dg;
    out.s.resize(s0.size());
    out.t.resize(t0.size());
    for (size_t i = 0; i < s0.size(); ++i) {
        out.s[i] = g.new_id[s0[i]];
        out.t[i] = g.new_id[t0[i]];
    }
    return MF_OK;
}

// ms per sample query, flows into `flows` (by sample index).  slots == 0:
// the plain solver, one query at a time on one slot; slots >= 1: every
// sample query submitted at once to an engine of that many slots (the
// batched engine's continuous refill), wall time / queries.  One untimed
// round first warms the graph's reachability caches and the slabs.
int measure(const Setup& su, const Params& P, int slots, std::vector<int64_t>& flows, double& ms) {
    Engine eng;
    int rc = eng.init(std::max(1, slots), P);
    if (rc != MF_OK) return rc;
    const int gid = eng.add_graph(su.dg);
    const int nq = (int)su.s.size();
    std::vector<Result> out;
    auto run = [&](int lo, int hi) -> int {
        out.clear();
        for (int i = lo; i < hi; ++i) {
            Query q;
            q.graph = gid;
            q.s = su.s[i];
            q.t = su.t[i];
            q.tag = i;
            const int r = eng.submit(q);
            if (r != MF_OK) return r;
        }
        const int r = eng.run(out);
        if (r != MF_OK) return r;
        for (auto& x : out)
            if (x.st.status != MF_OK) {
                set_error("autotune: a sample query failed: " + last_error());
                return x.st.status;
            }
        if ((int)out.size() != hi - lo) { set_error("autotune: a sample query gave no result"); return MF_EHIP; }
        for (auto& x : out) flows[x.tag] = x.st.flow;
        return MF_OK;
    };
    flows.assign(nq, 0);
    // warm-up: plain, the first query; batched, a round of min(slots, nq) queries
    rc = run(0, slots == 0 ? std::min(1, nq) : std::min(slots, nq));
    if (rc != MF_OK) return rc;
    const double t0 = seconds_now();
    if (slots == 0) {
        for (int i = 0; i < nq; ++i)
            if ((rc = run(i, i + 1)) != MF_OK) return rc;
    } else if ((rc = run(0, nq)) != MF_OK) {
        return rc;
    }
    ms = (seconds_now() - t0) * 1e3 / std::max(1, nq);
    return MF_OK;
}

std::string cache_key(const std::string& cls, int batch) { return cls + (batch > 0 ? ":b" + std::to_string(batch) : ":plain"); }

bool read_cache(const std::string& path, const std::string& key, TuneResult& r) {
    std::ifstream f(path);
    std::string line;
    bool found = false;
    while (std::getline(f, line)) {
        std::istringstream in(line);
        std::string c;
        TuneResult x;
        if (!(in >> c >> x.R >> x.tile_max >> x.K >> x.gr_freq >> x.tile_threads >> x.slots >> x.ms >> x.ms_default))
            continue;
        if (!(in >> x.tile_split)) x.tile_split = 1;
        if (c == key) { r = x; found = true; }  // the last entry of a key wins
    }
    return found;
}

}  // namespace

int autotune(const HostGraph& g0, const std::vector<int32_t>& s, const std::vector<int32_t>& t, const Params& base,
             int R0, int tile0, int batch, const std::string& cache_path, TuneResult& out) {
    if (!g0.rstart.empty()) { set_error("autotune needs the graph before partitioning"); return MF_EINVAL; }
    if (s.empty() || s.size() != t.size()) { set_error("autotune needs sample queries"); return MF_EINVAL; }
    if (batch < 0) { set_error("autotune: batch must be >= 0"); return MF_EINVAL; }
    for (size_t i = 0; i < s.size(); ++i)
        if (s[i] < 0 || t[i] < 0 || s[i] >= g0.n || t[i] >= g0.n || s[i] == t[i]) {
            set_error("autotune: bad sample query");
            return MF_EINVAL;
        }
    const std::string cls = graph_class(g0);
    const std::string key = cache_key(cls, batch);
    if (!cache_path.empty()) {
        TuneResult c;
        if (read_cache(cache_path, key, c)) {
            out = c;
            out.cls = cls;
            out.evaluations = 0;
            return MF_OK;
        }
    }
    struct Cand { int R, tile, K, gr, threads, slots, split; };
    Cand best{R0, tile0, base.K, base.gr_freq, base.tile_threads, batch, std::max(1, base.tile_split)};
    std::map<std::pair<int, int>, Setup> setups;
    auto setup_for = [&](int R, int tile, Setup** su) -> int {
        auto key2 = std::make_pair(R, tile);
        auto it = setups.find(key2);
        if (it == setups.end()) {
            // keep at most the incumbent's setup and this one resident
            for (auto jt = setups.begin(); jt != setups.end();)
                if (jt->first != std::make_pair(best.R, best.tile)) jt = setups.erase(jt); else ++jt;
            Setup nsu;
            const int rc = prepare(g0, R, tile, s, t, nsu);
            if (rc != MF_OK) return rc;
            it = setups.emplace(key2, std::move(nsu)).first;
        }
        *su = &it->second;
        return MF_OK;
    };
    std::vector<int64_t> ref;
    int evals = 0;
    auto score = [&](const Cand& c, double& ms) -> int {
        Setup* su = nullptr;
        int rc = setup_for(c.R, c.tile, &su);
        if (rc != MF_OK) return rc;
        Params P = base;
        P.K = c.K;
        P.gr_freq = c.gr;
        P.gr_freq_min = std::min(c.gr, base.gr_freq_min);
        P.gr_freq_max = std::max(2 * c.gr, base.gr_freq_max);
        P.tile_threads = c.threads;
        P.tile_split = c.split;
        std::vector<int64_t> flows;
        rc = measure(*su, P, c.slots, flows, ms);
        if (rc != MF_OK) return rc;
        ++evals;
        if (ref.empty()) {
            ref = flows;
        } else if (flows != ref) {
            set_error("autotune: a configuration changed a flow");
            return MF_ECERT;
        }
        return MF_OK;
    };
    double best_ms = 0;
    int rc = score(best, best_ms);
    if (rc != MF_OK) return rc;
    const double default_ms = best_ms;
    const std::vector<int> Rs = {128, 256, 512, 1024};
    const std::vector<int> tiles = {1024, 2048, 4096};
    const std::vector<int> Ks = {0, 1, 2, 4, 8};
    const std::vector<int> grs = {64, 128, 256, 512};
    const std::vector<int> threads = {256, 512, 1024};
    const std::vector<int> splits = {1, 2, 4, 8};
    std::vector<int> slots;
    if (batch > 0)
        for (int k : {batch / 2, batch, 2 * batch})
            if (k >= 1 && std::find(slots.begin(), slots.end(), k) == slots.end()) slots.push_back(k);
    auto sweep = [&](auto field, const std::vector<int>& vals) -> int {
        for (int v : vals) {
            Cand c = best;
            if (field(c) == v) continue;
            field(c) = v;
            double ms = 0;
            const int r = score(c, ms);
            if (r != MF_OK) return r;
            if (ms < best_ms) { best_ms = ms; best = c; }
        }
        return MF_OK;
    };
    if ((rc = sweep([](Cand& c) -> int& { return c.R; }, Rs)) != MF_OK) return rc;
    if ((rc = sweep([](Cand& c) -> int& { return c.tile; }, tiles)) != MF_OK) return rc;
    if ((rc = sweep([](Cand& c) -> int& { return c.K; }, Ks)) != MF_OK) return rc;
    if ((rc = sweep([](Cand& c) -> int& { return c.gr; }, grs)) != MF_OK) return rc;
    if ((rc = sweep([](Cand& c) -> int& { return c.threads; }, threads)) != MF_OK) return rc;
    if ((rc = sweep([](Cand& c) -> int& { return c.split; }, splits)) != MF_OK) return rc;
    if ((rc = sweep([](Cand& c) -> int& { return c.slots; }, slots)) != MF_OK) return rc;
    out.R = best.R;
    out.tile_max = best.tile;
    out.K = best.K;
    out.gr_freq = best.gr;
    out.tile_threads = best.threads;
    out.slots = best.slots;
    out.tile_split = best.split;
    out.ms = best_ms;
    out.ms_default = default_ms;
    out.evaluations = evals;
    out.cls = cls;
    if (!cache_path.empty()) {
        std::ofstream f(cache_path, std::ios::app);
        f << key << " " << out.R << " " << out.tile_max << " " << out.K << " " << out.gr_freq << " " << out.tile_threads
          << " " << out.slots << " " << out.ms << " " << out.ms_default << " " << out.tile_split << "\n";
    }
    return MF_OK;
}

}  // namespace mf
