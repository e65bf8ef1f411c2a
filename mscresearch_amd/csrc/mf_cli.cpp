// mf_cli.cpp -- drop-in command line: mscmf <input_dir> <output_dir> [static|dynamic]
//
// Same invocation, inputs and outputs as the reference's main()
// (Region_Aware_PR_with_Pruning_BoundaryCore.cu, BC:2561-2933):
//   <in>/graph.txt, <in>/test.txt (1-based) -> <out>/res_mt.txt (static) or
//   <out>/res_dynamic.txt (dynamic), one "%lld %lld\n" = flow, integer ms per
//   query in input order; invalid or s==t queries give "0 0".
// Additions: <out>/metrics.json (microsecond latencies, QPS, p50/p99), and in
// dynamic mode an optional <in>/updates.txt ("round u v new_cap" per line,
// 1-based ids) whose rounds are applied and re-solved warm, results in
// <out>/res_dynamic_rounds.txt ("round query flow time_ms").
// Environment: MSCMF_R (regions, default 128 as BC:2583), MSCMF_SLOTS
// (concurrent queries, default 1 = one query at a time like the reference),
// MSCMF_TILE (tile size), MSCMF_PRUNE (0/1), MSCMF_GRAPH_CACHE (path of a
// binary cache of the prepared graph: loaded while it matches graph.txt and
// MSCMF_R, else rebuilt and written), MSCMF_LOG=1 (static mode with one
// slot: the reference's per-query and per-iteration stderr lines, timed).
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <string>
#include <vector>

#include "../../include/mscmf.h"

namespace {

int env_int(const char* k, int dflt) {
    const char* v = getenv(k);
    return (v && *v) ? atoi(v) : dflt;
}

double pct(std::vector<double> v, double p) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    // exact order statistic: smallest value with at least p of the mass at or below it
    size_t k = (size_t)std::ceil(p * (double)v.size());
    if (k < 1) k = 1;
    if (k > v.size()) k = v.size();
    return v[k - 1];
}

void write_metrics(const std::string& path, const char* mode, const std::vector<double>& lat_us,
                   double wall_s, int64_t edges_total, int nvalid) {
    FILE* f = fopen(path.c_str(), "w");
    if (!f) return;
    fprintf(f, "{\"mode\": \"%s\", \"queries\": %d, \"wall_s\": %.6f, \"qps\": %.3f, "
               "\"residual_edges_per_s\": %.3f, \"p50_us\": %.1f, \"p99_us\": %.1f, \"latency_us\": [",
            mode, nvalid, wall_s, wall_s > 0 ? nvalid / wall_s : 0.0,
            wall_s > 0 ? (double)edges_total / wall_s : 0.0, pct(lat_us, 0.5), pct(lat_us, 0.99));
    for (size_t i = 0; i < lat_us.size(); ++i) fprintf(f, "%s%.1f", i ? ", " : "", lat_us[i]);
    fprintf(f, "]}\n");
    fclose(f);
}

// MSCMF_LOG=1 (static mode, one slot): the reference's per-query stderr
// lines, in the format its plotting scripts parse ("[Init] kernel_1(...)"
// BC:2188, the per-iteration timing line of the predecessor solver
// Region_Aware_PR_without_Pruning.cu:1802-1807, "[QueryPrune+RegionBoundary]"
// BC:2815-2819), from the engine's step log with HIP-event timing.  One
// "Iteration" line per discharge step: the relabel and count time since the
// previous one goes into it, as the reference's iteration starts with its
// global relabel.
struct RefLog {
    int path_len = 256;
    int64_t iteration = 0;
    double k1 = 0, k7 = 0, init_hl = 0, bfs = 0, pr = 0, prune = 0;
    bool init_printed = false;
    void start() { *this = RefLog{path_len}; }
};

void ref_step_log(void* user, const mf_step_log* L) {
    RefLog& R = *(RefLog*)user;
    const uint32_t s = L->steps;
    R.k1 += L->ms[MF_LOG_INIT];
    R.k7 += L->ms[MF_LOG_PREFLOW];
    R.init_hl += L->ms[MF_LOG_SEED];
    if (s & MF_STEP_PRUNE) R.prune += L->ms[MF_LOG_BFS];
    else R.bfs += L->ms[MF_LOG_BFS];
    R.pr += L->ms[MF_LOG_PR];
    if ((s & MF_STEP_PREFLOW) && !R.init_printed) {
        fprintf(stderr, "[Init] kernel_1(init residual/excess)=%.3f ms, kernel_7(preflow)=%.3f ms\n", R.k1, R.k7);
        R.init_printed = true;
    }
    if (s & MF_STEP_DISCHARGE) {
        fprintf(stderr,
                "Iteration %lld: init_hl=%.3f ms, bfs=%.3f ms, push_relabel=%.3f ms (K=%d, mode=%s, deeppath=%d), "
                "remove_invalid=%.3f ms, active=%d, active_high=%d, fixed_edges=%lld, GR_FREQ0=%d\n\n",
                (long long)R.iteration, R.init_hl, R.bfs, R.pr, L->K, L->path ? "active-aware" : "region-aware",
                R.path_len, L->ms[MF_LOG_RM], L->active, L->active_high,
                (long long)((s & MF_STEP_RM) ? L->fixed_edges : 0), L->gr_freq);
        R.iteration += 1;
        R.init_hl = R.bfs = R.pr = 0;
    }
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        fprintf(stderr, "Usage: %s <input_dir> <output_dir> [mode]\n", argv[0]);
        fprintf(stderr, "  mode = static (default) | dynamic\n");
        return 1;
    }
    const std::string in = argv[1], out = argv[2];
    const std::string mode = argc >= 4 ? argv[3] : "static";
    mkdir(out.c_str(), 0755);  // errors ignored, as BC:2555-2557

    mf_hgraph* h = nullptr;
    const int R = env_int("MSCMF_R", 128);
    const std::string gpath = in + "/graph.txt";
    // MSCMF_GRAPH_CACHE: binary cache of the prepared graph, used while it
    // matches graph.txt (size, modification time) and MSCMF_R, else rebuilt
    const char* cache = getenv("MSCMF_GRAPH_CACHE");
    if (cache && *cache) {
        const auto t0 = std::chrono::steady_clock::now();
        if (mf_hgraph_open(cache, gpath.c_str(), R, &h) == MF_OK)
            fprintf(stderr, "[GraphCache] loaded %s in %.0f ms\n", cache,
                    std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
        else
            fprintf(stderr, "[GraphCache] %s; rebuilding from %s\n", mf_last_error(), gpath.c_str());
    }
    if (!h) {
        if (mf_hgraph_load(gpath.c_str(), &h) != MF_OK) {
            fprintf(stderr, "%s\n", mf_last_error());
            exit(1);
        }
        mf_hgraph_partition(h, R);
        if (cache && *cache) {
            if (mf_hgraph_save(h, cache, gpath.c_str()) == MF_OK) fprintf(stderr, "[GraphCache] wrote %s\n", cache);
            else fprintf(stderr, "[GraphCache] not written: %s\n", mf_last_error());
        }
    }
    int n = 0, Rg = 0;
    int64_t E = 0, lines = 0;
    mf_hgraph_info(h, &n, &E, &Rg, &lines);
    std::vector<int32_t> new_id(n), old_id(n);
    mf_hgraph_export(h, nullptr, nullptr, nullptr, nullptr, new_id.data(), old_id.data(), nullptr, nullptr);

    int64_t nq = 0;
    if (mf_read_queries((in + "/test.txt").c_str(), 0, nullptr, nullptr, &nq) != MF_OK) {
        fprintf(stderr, "Cannot open %s\n", (in + "/test.txt").c_str());
        return 1;
    }
    std::vector<int32_t> qs(nq), qt(nq);
    mf_read_queries((in + "/test.txt").c_str(), nq, qs.data(), qt.data(), &nq);
    // map to region-ordered ids; out of range -> (-1,-1) (BC:2609-2620)
    for (int64_t i = 0; i < nq; ++i) {
        if (qs[i] >= 0 && qs[i] < n && qt[i] >= 0 && qt[i] < n) {
            qs[i] = new_id[qs[i]];
            qt[i] = new_id[qt[i]];
        } else {
            qs[i] = qt[i] = -1;
        }
    }
    fprintf(stderr, "Graph (reordered by Region): n=%d, m=%lld, Regions=%d, Queries=%lld\n", n, (long long)E, R,
            (long long)nq);
    if (mode != "static" && mode != "dynamic") {
        fprintf(stderr, "Unknown mode: %s (use static or dynamic)\n", mode.c_str());
        mf_hgraph_free(h);
        return 1;
    }

    mf_graph* g = nullptr;
    if (mf_graph_upload(h, env_int("MSCMF_TILE", 4096), &g) != MF_OK) {
        fprintf(stderr, "%s\n", mf_last_error());
        return 2;
    }
    mf_params P;
    mf_params_default(&P);
    P.prune = env_int("MSCMF_PRUNE", 1);
    const int slots = std::max(1, env_int("MSCMF_SLOTS", 1));
    mf_engine* eng = nullptr;
    if (mf_engine_create(std::max<int>(slots, mode == "dynamic" ? 10 : 1), &P, &eng) != MF_OK) {
        fprintf(stderr, "%s\n", mf_last_error());
        return 2;
    }
    int gid = 0;
    if (mf_engine_add_graph(eng, g, &gid) != MF_OK) {
        fprintf(stderr, "%s\n", mf_last_error());
        return 2;
    }

    RefLog reflog;
    reflog.path_len = P.path_len;
    const bool ref_log = mode == "static" && slots == 1 && env_int("MSCMF_LOG", 0) != 0;
    if (ref_log) {
        mf_engine_profile(eng, 1);
        mf_engine_set_step_log(eng, ref_step_log, &reflog);
    }
    std::vector<long long> flows(nq, 0), times(nq, 0);
    std::vector<double> lat_us;
    int64_t edges_total = 0;
    int rc_final = 0;
    const auto t_wall0 = std::chrono::steady_clock::now();
    auto wall_since = [&]() {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t_wall0).count();
    };

    if (mode == "static") {
        for (int64_t i = 0; i < nq; ++i) {
            const int s = qs[i], t = qt[i];
            if (ref_log) {
                const int os = (s >= 0) ? old_id[s] : -1, ot = (t >= 0) ? old_id[t] : -1;
                fprintf(stderr, "[Static] Query %lld: s=%d, t=%d... ", (long long)i + 1, os + 1, ot + 1);
                if (s < 0 || t < 0 || s == t) fprintf(stderr, "Invalid.\n");
                reflog.start();
            }
            if (s < 0 || t < 0 || s == t) continue;
            if (mf_engine_submit(eng, gid, s, t, i, 0.0) != MF_OK) {
                fprintf(stderr, "%s\n", mf_last_error());
                return 2;
            }
            if (slots == 1) {
                // one query at a time, like the reference
                std::vector<mf_result> rs(4);
                int k = 0;
                mf_engine_run(eng, -1, rs.data(), (int)rs.size(), &k);
                for (int j = 0; j < k; ++j) {
                    const mf_result& r = rs[j];
                    if (r.status != MF_OK) {
                        fprintf(stderr, "Query %lld failed: %s\n", (long long)r.tag + 1, mf_last_error());
                        rc_final = 3;
                    }
                    flows[r.tag] = r.flow;
                    const double ms = (r.t_end - r.t_arrive) * 1e3;
                    times[r.tag] = (long long)ms;
                    lat_us.push_back(ms * 1e3);
                    edges_total += r.E_run;
                    if (ref_log) {
                        if (r.pruned)
                            fprintf(stderr,
                                    "[QueryPrune+RegionBoundary] prune_time=%lld ms, pruned_n=%lld/%d, pruned_m=%lld, "
                                    "run_n=%d, run_m=%lld, Flow=%lld, Time=%lld ms\n",
                                    (long long)reflog.prune, (long long)r.kept, n, (long long)r.E_run, r.n_run,
                                    (long long)r.E_run, (long long)r.flow, times[r.tag]);
                        else
                            fprintf(stderr,
                                    "[QueryPrune+RegionBoundary] prune_time=%lld ms, pruned_n=%lld/%d, pruned_m=%lld, "
                                    "fallback=full, run_n=%d, run_m=%lld, Flow=%lld, Time=%lld ms\n",
                                    (long long)reflog.prune, (long long)r.kept, n, (long long)r.E_run, r.n_run,
                                    (long long)r.E_run, (long long)r.flow, times[r.tag]);
                        // the reference would go on to its boundary-aware core
                        // (BC:2733-2753), which can cut s-t paths; say so
                        int64_t pn = 0, rn = 0, ck = 0;
                        int take = 0;
                        if (mf_hgraph_core_decision(h, s, t, &pn, &rn, &ck, &take) == MF_OK && take)
                            fprintf(stderr,
                                    "[RegionBoundary] the reference would run this query on its boundary core: "
                                    "core_n=%lld of run_n=%lld (< 0.98); not taken here, it can drop vertices that "
                                    "carry s-t flow\n",
                                    (long long)ck, (long long)rn);
                    }
                }
            }
        }
        if (slots > 1) {
            std::vector<mf_result> rs(nq + 1);
            int k = 0;
            mf_engine_run(eng, -1, rs.data(), (int)rs.size(), &k);
            for (int j = 0; j < k; ++j) {
                const mf_result& r = rs[j];
                if (r.status != MF_OK) rc_final = 3;
                flows[r.tag] = r.flow;
                const double ms = (r.t_end - r.t_arrive) * 1e3;
                times[r.tag] = (long long)ms;
                lat_us.push_back(ms * 1e3);
                edges_total += r.E_run;
            }
        }
        long long total = 0;
        for (int64_t i = 0; i < nq; ++i) {
            const int s = qs[i], t = qt[i];
            const int os = (s >= 0) ? old_id[s] : -1, ot = (t >= 0) ? old_id[t] : -1;
            if (ref_log) {  // printed per query above
                if (s >= 0 && t >= 0 && s != t) total += times[i];
                continue;
            }
            if (s < 0 || t < 0 || s == t) {
                fprintf(stderr, "[Static] Query %lld: s=%d, t=%d... Invalid.\n", (long long)i + 1, os + 1, ot + 1);
                continue;
            }
            fprintf(stderr, "[Static] Query %lld: s=%d, t=%d... Flow=%lld, Time=%lld ms\n", (long long)i + 1, os + 1,
                    ot + 1, flows[i], times[i]);
            total += times[i];
        }
        const double wall = wall_since();
        FILE* f = fopen((out + "/res_mt.txt").c_str(), "w");
        if (!f) {
            fprintf(stderr, "Cannot open output file %s\n", (out + "/res_mt.txt").c_str());
            return 1;
        }
        for (int64_t i = 0; i < nq; ++i) fprintf(f, "%lld %lld\n", flows[i], times[i]);
        fclose(f);
        write_metrics(out + "/metrics.json", "static", lat_us, wall, edges_total, (int)lat_us.size());
        fprintf(stderr, "Total time: %lld ms\n", total);
    } else if (nq == 0) {
        // BC:2845-2847: no queries, no res_dynamic.txt
        fprintf(stderr, "[Dynamic] No queries found.\n");
    } else {
        // rounds of capacity updates, if any
        std::map<int, std::vector<std::pair<int32_t, int64_t>>> rounds;
        {
            FILE* uf = fopen((in + "/updates.txt").c_str(), "r");
            if (uf) {
                int r, u, v;
                long long c;
                while (fscanf(uf, "%d %d %d %lld", &r, &u, &v, &c) == 4) {
                    --u; --v;
                    if (u < 0 || u >= n || v < 0 || v >= n || c < 1) continue;
                    int32_t e = -1;
                    mf_hgraph_find_entry(h, new_id[u], new_id[v], &e);
                    if (e >= 0) rounds[r].push_back({e, (int64_t)c});
                }
                fclose(uf);
            }
        }
        FILE* dout = fopen((out + "/res_dynamic.txt").c_str(), "w");
        if (!dout) {
            fprintf(stderr, "Cannot open dynamic output file %s\n", (out + "/res_dynamic.txt").c_str());
            return 1;
        }
        FILE* rout = rounds.empty() ? nullptr : fopen((out + "/res_dynamic_rounds.txt").c_str(), "w");
        const int W = std::max(10, slots);  // window (BC:2852 uses 10)
        long long total = 0;
        for (int64_t w0 = 0; w0 < nq; w0 += W) {
            const int64_t w1 = std::min<int64_t>(nq, w0 + W);
            fprintf(stderr, "\n[Dynamic] ===== Processing queries [%lld, %lld) =====\n", (long long)w0, (long long)w1);
            mf_engine_free(eng);
            eng = nullptr;
            if (!rounds.empty() && w0 > 0) {
                // every window starts from the capacities of graph.txt
                mf_graph_free(g);
                g = nullptr;
                if (mf_graph_upload(h, env_int("MSCMF_TILE", 4096), &g) != MF_OK) {
                    fprintf(stderr, "%s\n", mf_last_error());
                    return 2;
                }
            }
            if (mf_engine_create(W, &P, &eng) != MF_OK || mf_engine_add_graph(eng, g, &gid) != MF_OK) {
                fprintf(stderr, "%s\n", mf_last_error());
                return 2;
            }
            int slot = 0;
            for (int64_t i = w0; i < w1; ++i)
                if (qs[i] >= 0 && qt[i] >= 0 && qs[i] != qt[i]) mf_engine_pin(eng, slot++, gid, qs[i], qt[i], i);
            std::vector<mf_result> rs(W);
            int k = 0;
            if (slot) {
                if (mf_engine_solve_pinned(eng, 0, rs.data(), W, &k) != MF_OK) rc_final = 3;
            }
            std::vector<long long> f0(w1 - w0, 0), t0(w1 - w0, 0);
            for (int j = 0; j < k; ++j) {
                const int64_t i = rs[j].tag;
                f0[i - w0] = rs[j].flow;
                t0[i - w0] = (long long)((rs[j].t_end - rs[j].t_start) * 1e3);
                lat_us.push_back((rs[j].t_end - rs[j].t_start) * 1e6);
                fprintf(stderr, "[Dynamic] Query %lld Init: flow=%lld, time=%lld ms\n", (long long)i + 1, f0[i - w0],
                        t0[i - w0]);
            }
            for (int64_t i = w0; i < w1; ++i) {
                fprintf(dout, "%lld %lld\n", f0[i - w0], t0[i - w0]);
                total += t0[i - w0];
            }
            for (auto& rd : rounds) {
                std::vector<int32_t> ent;
                std::vector<int64_t> cap;
                std::map<int32_t, int64_t> last;
                for (auto& pr : rd.second) last[pr.first] = pr.second;
                for (auto& pr : last) { ent.push_back(pr.first); cap.push_back(pr.second); }
                if (mf_graph_update_caps(eng, gid, (int)ent.size(), ent.data(), cap.data()) != MF_OK) {
                    fprintf(stderr, "%s\n", mf_last_error());
                    rc_final = 3;
                    break;
                }
                if (slot && mf_engine_solve_pinned(eng, 1, rs.data(), W, &k) != MF_OK) rc_final = 3;
                for (int j = 0; j < k && rout; ++j)
                    fprintf(rout, "%d %lld %lld %lld\n", rd.first, (long long)rs[j].tag + 1, (long long)rs[j].flow,
                            (long long)((rs[j].t_end - rs[j].t_start) * 1e3));
            }
        }
        fclose(dout);
        if (rout) fclose(rout);
        const double wall = wall_since();
        write_metrics(out + "/metrics.json", "dynamic", lat_us, wall, edges_total, (int)lat_us.size());
        fprintf(stderr, "Total time: %lld ms\n", total);
        fprintf(stderr, "[Dynamic] Summary results written to %s\n", (out + "/res_dynamic.txt").c_str());
    }
    mf_engine_free(eng);
    mf_graph_free(g);
    mf_hgraph_free(h);
    return rc_final;
}
