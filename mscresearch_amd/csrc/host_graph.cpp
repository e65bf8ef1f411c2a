// host_graph.cpp -- parsing, CSR, capacity-degree partition, region reorder,
// seeded generators.  See host_graph.h.
#include "host_graph.h"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <queue>
#include <thread>
#include <type_traits>

namespace mf {

// ------------------------------------------------------------------ scanner
// A reader with the token rules of scanf's %d / %lld: leading white space is
// skipped, an optional sign, then at least one digit.  Out-of-range values
// saturate like strtol/strtoll and %d keeps the low 32 bits of the long.
namespace {

struct MappedFile {
    const char* p = nullptr;
    size_t n = 0;
    int fd = -1;
    void* base = nullptr;
    bool open(const std::string& path) {
        fd = ::open(path.c_str(), O_RDONLY);
        if (fd < 0) return false;
        struct stat st;
        if (fstat(fd, &st) != 0) return false;
        n = (size_t)st.st_size;
        if (n == 0) { p = ""; return true; }
        base = mmap(nullptr, n, PROT_READ, MAP_PRIVATE, fd, 0);
        if (base == MAP_FAILED) { base = nullptr; return false; }
        madvise(base, n, MADV_SEQUENTIAL);
        p = (const char*)base;
        return true;
    }
    ~MappedFile() {
        if (base) munmap(base, n);
        if (fd >= 0) ::close(fd);
    }
};

inline bool is_space(char c) {
    return c == ' ' || c == '\n' || c == '\t' || c == '\r' || c == '\v' || c == '\f';
}

struct Scanner {
    const char* p;
    const char* e;
    // Reads one signed integer with strtoll saturation; false on failure.
    bool read_ll(long long& out) {
        while (p < e && is_space(*p)) ++p;
        if (p >= e) return false;
        bool neg = false;
        const char* q = p;
        if (*q == '+' || *q == '-') { neg = (*q == '-'); ++q; }
        if (q >= e || *q < '0' || *q > '9') return false;
        unsigned long long acc = 0;
        bool over = false;
        const unsigned long long lim = neg ? (unsigned long long)LLONG_MAX + 1ULL : (unsigned long long)LLONG_MAX;
        while (q < e && *q >= '0' && *q <= '9') {
            unsigned d = (unsigned)(*q - '0');
            if (!over) {
                if (acc > (lim - d) / 10ULL) over = true;
                else acc = acc * 10ULL + d;
            }
            ++q;
        }
        p = q;
        if (over) out = neg ? LLONG_MIN : LLONG_MAX;
        else out = neg ? (long long)(0ULL - acc) : (long long)acc;
        return true;
    }
    bool read_int(int& out) {
        long long v;
        if (!read_ll(v)) return false;
        out = (int)(long)v;  // %d stores the conversion result into an int
        return true;
    }
};

}  // namespace

int parse_graph_file(const std::string& path, std::vector<int32_t>& us, std::vector<int32_t>& vs,
                     std::vector<int64_t>& caps, int& n, std::string& err) {
    MappedFile f;
    if (!f.open(path)) { err = "Cannot open " + path; return -1; }
    Scanner sc{f.p, f.p + f.n};
    int nn, mm;
    if (!sc.read_int(nn) || !sc.read_int(mm)) { err = "Invalid graph header in " + path; return -2; }
    if (nn < 0) { err = "Invalid graph header in " + path; return -2; }
    n = nn;
    us.clear(); vs.clear(); caps.clear();
    if (mm > 0) { us.reserve(mm); vs.reserve(mm); caps.reserve(mm); }
    for (int i = 0; i < mm; ++i) {
        int u, v;
        long long c;
        if (!sc.read_int(u) || !sc.read_int(v) || !sc.read_ll(c)) break;
        --u; --v;
        if (u < 0 || u >= n || v < 0 || v >= n) continue;
        if (u == v) continue;
        if (c <= 0) continue;
        us.push_back(u); vs.push_back(v); caps.push_back((int64_t)c);
    }
    return 0;
}

int parse_query_file(const std::string& path, std::vector<int32_t>& s, std::vector<int32_t>& t,
                     std::string& err) {
    MappedFile f;
    if (!f.open(path)) { err = "Cannot open " + path; return -1; }
    Scanner sc{f.p, f.p + f.n};
    s.clear(); t.clear();
    for (;;) {
        int u, v;
        if (!sc.read_int(u) || !sc.read_int(v)) break;
        s.push_back(u - 1);
        t.push_back(v - 1);
    }
    return 0;
}

// --------------------------------------------------------------------- CSR

void build_csr(int n, const std::vector<int32_t>& us, const std::vector<int32_t>& vs,
               const std::vector<int64_t>& caps, HostGraph& g) {
    const int64_t cnt = (int64_t)us.size();
    g = HostGraph();
    g.n = n;
    g.lines = cnt;
    g.row_ptr.assign((size_t)n + 1, 0);
    for (int64_t i = 0; i < cnt; ++i) { g.row_ptr[us[i] + 1]++; g.row_ptr[vs[i] + 1]++; }
    for (int v = 0; v < n; ++v) g.row_ptr[v + 1] += g.row_ptr[v];
    const int64_t E = 2 * cnt;
    g.col.resize(E); g.cap.resize(E); g.rev.resize(E);
    std::vector<int32_t> cur(g.row_ptr.begin(), g.row_ptr.end() - 1);
    g.sum_out.assign(n, 0); g.sum_in.assign(n, 0);
    g.deg_out.assign(n, 0); g.deg_in.assign(n, 0);
    for (int64_t i = 0; i < cnt; ++i) {
        const int u = us[i], v = vs[i];
        const int pu = cur[u]++;
        const int pv = cur[v]++;
        g.col[pu] = v; g.cap[pu] = caps[i]; g.rev[pu] = pv;
        g.col[pv] = u; g.cap[pv] = 0;       g.rev[pv] = pu;
        g.sum_out[u] += caps[i]; g.sum_in[v] += caps[i];
        g.deg_out[u]++; g.deg_in[v]++;
    }
}

void csr_stats(const HostGraph& g, std::vector<int64_t>& so, std::vector<int64_t>& si,
               std::vector<int32_t>& dout, std::vector<int32_t>& din) {
    so.assign(g.n, 0); si.assign(g.n, 0); dout.assign(g.n, 0); din.assign(g.n, 0);
    for (int v = 0; v < g.n; ++v)
        for (int e = g.row_ptr[v]; e < g.row_ptr[v + 1]; ++e)
            if (g.cap[e] > 0) {
                const int u = g.col[e];
                so[v] += g.cap[e]; si[u] += g.cap[e];
                dout[v]++; din[u]++;
            }
}

// ---------------------------------------------------------------- partition

void capacity_degree_partition(int n, int R, const std::vector<int64_t>& so,
                               const std::vector<int64_t>& si, const std::vector<int32_t>& dout,
                               const std::vector<int32_t>& din, std::vector<int32_t>& part) {
    // Same arithmetic, in the same order, as BC:1141-1160.
    std::vector<double> score(n);
    for (int u = 0; u < n; ++u) {
        const double s_o = (double)so[u], s_i = (double)si[u];
        const double d_o = (double)dout[u], d_i = (double)din[u];
        const double num = std::log(s_o * d_o + 1.0) + std::log(s_i * d_i + 1.0);
        const double prod = d_o * d_i;
        const double den = std::log(std::sqrt(prod) + 1.0);
        score[u] = (den > 0.0) ? num / den : 1.0;
    }
    std::vector<int32_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int a, int b) {
        if (score[a] != score[b]) return score[a] > score[b];
        return a < b;
    });
    // Least-loaded region with the lowest index on ties == the lexicographic
    // minimum of (load, index); a heap gives it in O(log R) per vertex.
    using Item = std::pair<double, int>;
    std::priority_queue<Item, std::vector<Item>, std::greater<Item>> heap;
    for (int r = 0; r < R; ++r) heap.push({0.0, r});
    part.assign(n, 0);
    for (int i = 0; i < n; ++i) {
        Item it = heap.top();
        heap.pop();
        const int v = order[i];
        part[v] = it.second;
        it.first += score[v];
        heap.push(it);
    }
}

void partition_and_reorder(HostGraph& g, int R, const std::vector<int32_t>& part) {
    const int n = g.n;
    std::vector<int32_t> cnt(R + 1, 0);
    for (int v = 0; v < n; ++v) {
        int r = part[v];
        if (r < 0 || r >= R) r = 0;
        cnt[r]++;
    }
    g.R = R;
    g.rstart.assign(R + 1, 0);
    for (int r = 0; r < R; ++r) g.rstart[r + 1] = g.rstart[r] + cnt[r];
    std::vector<int32_t> fill(g.rstart.begin(), g.rstart.end() - 1);
    g.new_id.assign(n, -1);
    g.old_id.assign(n, -1);
    for (int v = 0; v < n; ++v) {
        int r = part[v];
        if (r < 0 || r >= R) r = 0;
        const int id = fill[r]++;
        g.new_id[v] = id;
        g.old_id[id] = v;
    }
    g.region_id.assign(n, 0);
    for (int r = 0; r < R; ++r)
        for (int v = g.rstart[r]; v < g.rstart[r + 1]; ++v) g.region_id[v] = r;
    std::vector<int32_t> rp2((size_t)n + 1, 0);
    for (int u = 0; u < n; ++u) {
        const int o = g.old_id[u];
        rp2[u + 1] = rp2[u] + (g.row_ptr[o + 1] - g.row_ptr[o]);
    }
    const int64_t E = g.E();
    std::vector<int32_t> col2(E), rev2(E);
    std::vector<int64_t> cap2(E);
    for (int o = 0; o < n; ++o) {
        const int base = rp2[g.new_id[o]];
        for (int e = g.row_ptr[o]; e < g.row_ptr[o + 1]; ++e) {
            const int ne = base + (e - g.row_ptr[o]);
            const int w = g.col[e];
            col2[ne] = g.new_id[w];
            cap2[ne] = g.cap[e];
            rev2[ne] = rp2[g.new_id[w]] + (g.rev[e] - g.row_ptr[w]);
        }
    }
    // Stats follow the vertices.
    std::vector<int64_t> so(n), si(n);
    std::vector<int32_t> dout(n), din(n);
    if ((int)g.sum_out.size() == n) {
        for (int u = 0; u < n; ++u) {
            const int o = g.old_id[u];
            so[u] = g.sum_out[o]; si[u] = g.sum_in[o]; dout[u] = g.deg_out[o]; din[u] = g.deg_in[o];
        }
        g.sum_out.swap(so); g.sum_in.swap(si); g.deg_out.swap(dout); g.deg_in.swap(din);
    }
    g.row_ptr.swap(rp2);
    g.col.swap(col2);
    g.cap.swap(cap2);
    g.rev.swap(rev2);
}

// ------------------------------------------------------------- components

void strong_components(const HostGraph& g, std::vector<int32_t>& comp) {
    const int n = g.n;
    comp.assign(n, -1);
    std::vector<int32_t> index(n, -1), low(n, 0), stk;
    std::vector<char> on(n, 0);
    std::vector<std::pair<int32_t, int32_t>> call;  // (vertex, next entry)
    int32_t idx = 0, nc = 0;
    stk.reserve(1024);
    call.reserve(1024);
    for (int r = 0; r < n; ++r) {
        if (index[r] != -1) continue;
        index[r] = low[r] = idx++;
        stk.push_back(r);
        on[r] = 1;
        call.push_back({r, g.row_ptr[r]});
        while (!call.empty()) {
            const int v = call.back().first;
            int e = call.back().second;
            const int end = g.row_ptr[v + 1];
            bool descended = false;
            for (; e < end; ++e) {
                if (g.cap[e] <= 0) continue;
                const int w = g.col[e];
                if (index[w] == -1) {
                    call.back().second = e + 1;
                    index[w] = low[w] = idx++;
                    stk.push_back(w);
                    on[w] = 1;
                    call.push_back({w, g.row_ptr[w]});
                    descended = true;
                    break;
                }
                if (on[w] && index[w] < low[v]) low[v] = index[w];
            }
            if (descended) continue;
            if (low[v] == index[v]) {
                for (;;) {
                    const int x = stk.back();
                    stk.pop_back();
                    on[x] = 0;
                    comp[x] = nc;
                    if (x == v) break;
                }
                ++nc;
            }
            call.pop_back();
            if (!call.empty()) {
                const int u = call.back().first;
                if (low[v] < low[u]) low[u] = low[v];
            }
        }
    }
}

void weak_components(const HostGraph& g, std::vector<int32_t>& comp) {
    const int n = g.n;
    std::vector<int32_t> par(n);
    std::iota(par.begin(), par.end(), 0);
    auto find = [&](int x) {
        while (par[x] != x) {
            par[x] = par[par[x]];
            x = par[x];
        }
        return x;
    };
    for (int v = 0; v < n; ++v)
        for (int e = g.row_ptr[v]; e < g.row_ptr[v + 1]; ++e) {
            int a = find(v), b = find(g.col[e]);
            if (a != b) { if (a < b) par[b] = a; else par[a] = b; }
        }
    comp.assign(n, 0);
    for (int v = 0; v < n; ++v) comp[v] = find(v);
}

// ------------------------------------------- the reference's core decision

CoreDecision reference_core_decision(const HostGraph& g, int s, int t, int hop_limit) {
    CoreDecision d;
    const int n = g.n;
    if (s < 0 || t < 0 || s >= n || t >= n || g.region_id.size() != (size_t)n) return d;
    // pruning mask (BC:1391-1445)
    std::vector<char> fwd(n, 0), und(n, 0);
    std::vector<int32_t> q;
    q.reserve(n);
    fwd[s] = 1;
    q.push_back(s);
    for (size_t h = 0; h < q.size(); ++h) {
        const int v = q[h];
        for (int e = g.row_ptr[v]; e < g.row_ptr[v + 1]; ++e)
            if (g.cap[e] > 0 && !fwd[g.col[e]]) { fwd[g.col[e]] = 1; q.push_back(g.col[e]); }
    }
    q.clear();
    und[t] = 1;
    q.push_back(t);
    for (size_t h = 0; h < q.size(); ++h) {
        const int v = q[h];
        for (int e = g.row_ptr[v]; e < g.row_ptr[v + 1]; ++e)
            if (!und[g.col[e]]) { und[g.col[e]] = 1; q.push_back(g.col[e]); }
    }
    std::vector<int32_t> old2sub(n, -1), sub2old;
    for (int v = 0; v < n; ++v)
        if ((fwd[v] && und[v]) || v == s || v == t) {
            old2sub[v] = (int32_t)sub2old.size();
            sub2old.push_back(v);
        }
    d.pruned_n = (int64_t)sub2old.size();
    d.run_n = n;
    if (!(d.pruned_n > 0 && (double)d.pruned_n < 0.995 * (double)n)) return d;
    // induced subgraph (BC:2477-2551): kept entries in row order, rev remapped
    HostGraph sub;
    const int k = (int)sub2old.size();
    sub.n = k;
    sub.row_ptr.assign((size_t)k + 1, 0);
    std::vector<int32_t> e2s(g.E(), -1);
    for (int i = 0; i < k; ++i) {
        const int v = sub2old[i];
        int cnt = 0;
        for (int e = g.row_ptr[v]; e < g.row_ptr[v + 1]; ++e)
            if (old2sub[g.col[e]] >= 0) e2s[e] = sub.row_ptr[i] + cnt++;
        sub.row_ptr[i + 1] = sub.row_ptr[i] + cnt;
    }
    sub.col.resize(sub.row_ptr[k]);
    sub.cap.resize(sub.row_ptr[k]);
    sub.rev.resize(sub.row_ptr[k]);
    for (int i = 0; i < k; ++i) {
        const int v = sub2old[i];
        for (int e = g.row_ptr[v]; e < g.row_ptr[v + 1]; ++e)
            if (e2s[e] >= 0) {
                sub.col[e2s[e]] = old2sub[g.col[e]];
                sub.cap[e2s[e]] = g.cap[e];
                sub.rev[e2s[e]] = e2s[g.rev[e]];
            }
    }
    d.pruned = true;
    d.run_n = k;
    // re-partition on stats over cap > 0 entries (BC:2705-2731)
    const int runR = std::min(std::max(1, g.R), std::max(1, k));
    std::vector<int64_t> so, si;
    std::vector<int32_t> dout, din, part;
    csr_stats(sub, so, si, dout, din);
    capacity_degree_partition(k, runR, so, si, dout, din, part);
    partition_and_reorder(sub, runR, part);
    const int rs_v = sub.new_id[old2sub[s]], rt_v = sub.new_id[old2sub[t]];
    // boundary-aware core (BC:2340-2475)
    const int R = runR;
    std::vector<std::vector<int32_t>> adj(R), radj(R);
    std::vector<char> bnd(k, 0);
    for (int v = 0; v < k; ++v) {
        const int rv = sub.region_id[v];
        for (int e = sub.row_ptr[v]; e < sub.row_ptr[v + 1]; ++e) {
            const int ru = sub.region_id[sub.col[e]];
            if (ru != rv) {
                adj[rv].push_back(ru);
                radj[ru].push_back(rv);
                bnd[v] = 1;
            }
        }
    }
    std::vector<char> from_s(R, 0), to_t(R, 0);
    std::vector<int32_t> rq;
    from_s[sub.region_id[rs_v]] = 1;
    rq.push_back(sub.region_id[rs_v]);
    for (size_t h = 0; h < rq.size(); ++h)
        for (int x : adj[rq[h]])
            if (!from_s[x]) { from_s[x] = 1; rq.push_back(x); }
    rq.clear();
    to_t[sub.region_id[rt_v]] = 1;
    rq.push_back(sub.region_id[rt_v]);
    for (size_t h = 0; h < rq.size(); ++h)
        for (int x : radj[rq[h]])
            if (!to_t[x]) { to_t[x] = 1; rq.push_back(x); }
    const int rs = sub.region_id[rs_v], rt = sub.region_id[rt_v];
    std::vector<int32_t> dist(k, -1);
    std::vector<char> mask(k, 0);
    q.clear();
    for (int v = 0; v < k; ++v) {
        const int r = sub.region_id[v];
        if (!(from_s[r] && to_t[r])) continue;
        if (r == rs || r == rt || bnd[v]) {
            mask[v] = 1;
            dist[v] = 0;
            q.push_back(v);
        }
    }
    for (size_t h = 0; h < q.size(); ++h) {
        const int v = q[h];
        if (dist[v] >= hop_limit) continue;
        const int rv = sub.region_id[v];
        for (int e = sub.row_ptr[v]; e < sub.row_ptr[v + 1]; ++e) {
            const int u = sub.col[e];
            if (sub.region_id[u] != rv || dist[u] != -1) continue;
            dist[u] = dist[v] + 1;
            mask[u] = 1;
            q.push_back(u);
        }
    }
    mask[rs_v] = 1;
    mask[rt_v] = 1;
    int64_t kept = 0;
    for (int v = 0; v < k; ++v) kept += mask[v];
    d.core_kept = kept;
    d.would_take = kept > 0 && (double)kept < 0.98 * (double)k;
    return d;
}

// --------------------------------------------------------------- generators

namespace {

inline uint64_t splitmix64(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (int i = 0; i < 4; ++i) s[i] = splitmix64(x);
    }
    static inline uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
    inline uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    // [0, k) by the high half of a 64x64 product.
    inline uint64_t below(uint64_t k) { return (uint64_t)(((unsigned __int128)next() * k) >> 64); }
    // [0, 1) with 53 bits.
    inline double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

}  // namespace

void gen_uniform(int n, int64_t m, int64_t cap_max, uint64_t seed, std::vector<int32_t>& us,
                 std::vector<int32_t>& vs, std::vector<int64_t>& caps) {
    Rng rng(seed);
    us.clear(); vs.clear(); caps.clear();
    for (int64_t i = 0; i < m; ++i) {
        const int u = (int)rng.below(n);
        const int v = (int)rng.below(n);
        const int64_t c = 1 + (int64_t)rng.below((uint64_t)cap_max);
        if (u == v) continue;
        us.push_back(u); vs.push_back(v); caps.push_back(c);
    }
}

void gen_grid(int rows, int cols, int64_t cap_max, uint64_t seed, std::vector<int32_t>& us,
              std::vector<int32_t>& vs, std::vector<int64_t>& caps) {
    Rng rng(seed);
    us.clear(); vs.clear(); caps.clear();
    auto emit = [&](int a, int b) {
        us.push_back(a); vs.push_back(b); caps.push_back(1 + (int64_t)rng.below((uint64_t)cap_max));
        us.push_back(b); vs.push_back(a); caps.push_back(1 + (int64_t)rng.below((uint64_t)cap_max));
    };
    for (int r = 0; r < rows; ++r)
        for (int c = 0; c < cols; ++c) {
            const int v = r * cols + c;
            if (c + 1 < cols) emit(v, v + 1);
            if (r + 1 < rows) emit(v, v + cols);
        }
}

void gen_rmat(int scale, int64_t samples, double a, double b, double c, int64_t cap_max,
              uint64_t seed, int threads, std::vector<int32_t>& us, std::vector<int32_t>& vs,
              std::vector<int64_t>& caps) {
    // Samples are drawn in fixed blocks, block k from its own stream seeded by
    // (seed, k), so the output does not depend on the thread count.
    const int64_t BLOCK = 1 << 20;
    const int64_t nblk = (samples + BLOCK - 1) / BLOCK;
    std::vector<uint64_t> keys((size_t)samples);
    std::vector<int64_t> cvals((size_t)samples);
    const double ab = a + b, abc = a + b + c;
    const double lnc = std::log((double)cap_max);
    auto work = [&](int64_t k0, int64_t k1) {
        for (int64_t k = k0; k < k1; ++k) {
            uint64_t sm = seed ^ (0xD1B54A32D192ED03ULL * (uint64_t)(k + 1));
            Rng rng(splitmix64(sm));
            const int64_t i0 = k * BLOCK, i1 = std::min(samples, i0 + BLOCK);
            for (int64_t i = i0; i < i1; ++i) {
                uint32_t u = 0, v = 0;
                for (int l = 0; l < scale; ++l) {
                    const double p = rng.unit();
                    u <<= 1; v <<= 1;
                    if (p < a) {
                    } else if (p < ab) {
                        v |= 1;
                    } else if (p < abc) {
                        u |= 1;
                    } else {
                        u |= 1; v |= 1;
                    }
                }
                int64_t cv = (int64_t)std::floor(std::exp(rng.unit() * lnc));
                if (cv < 1) cv = 1;
                keys[i] = ((uint64_t)u << 32) | v;
                cvals[i] = (u == v) ? 0 : cv;
            }
        }
    };
    if (threads < 1) threads = 1;
    std::vector<std::thread> th;
    for (int w = 0; w < threads; ++w) {
        const int64_t k0 = nblk * w / threads, k1 = nblk * (w + 1) / threads;
        th.emplace_back(work, k0, k1);
    }
    for (auto& x : th) x.join();
    // sort (key, cap) pairs by key with an LSD radix sort on an index array
    std::vector<uint32_t> idx((size_t)samples), tmp((size_t)samples);
    std::iota(idx.begin(), idx.end(), 0u);
    for (int pass = 0; pass < 4; ++pass) {
        const int shift = pass * 16;
        std::vector<int64_t> count(65537, 0);
        for (int64_t i = 0; i < samples; ++i) count[((keys[idx[i]] >> shift) & 0xFFFF) + 1]++;
        for (int d = 0; d < 65536; ++d) count[d + 1] += count[d];
        for (int64_t i = 0; i < samples; ++i) tmp[count[(keys[idx[i]] >> shift) & 0xFFFF]++] = idx[i];
        idx.swap(tmp);
    }
    us.clear(); vs.clear(); caps.clear();
    us.reserve(samples); vs.reserve(samples); caps.reserve(samples);
    int64_t i = 0;
    while (i < samples) {
        const uint64_t k = keys[idx[i]];
        int64_t sum = 0;
        while (i < samples && keys[idx[i]] == k) { sum += cvals[idx[i]]; ++i; }
        const int u = (int)(k >> 32), v = (int)(k & 0xFFFFFFFFu);
        if (u == v) continue;  // self-loop dropped
        us.push_back(u); vs.push_back(v); caps.push_back(sum);
    }
}

void gen_queries_uniform(int n, int q, uint64_t seed, std::vector<int32_t>& s,
                         std::vector<int32_t>& t) {
    Rng rng(seed);
    s.clear(); t.clear();
    while ((int)s.size() < q) {
        const int a = (int)rng.below(n), b = (int)rng.below(n);
        if (a == b) continue;
        s.push_back(a); t.push_back(b);
    }
}

void gen_queries_grid(int rows, int cols, int q, uint64_t seed, std::vector<int32_t>& s,
                      std::vector<int32_t>& t) {
    Rng rng(seed);
    s.clear(); t.clear();
    for (int i = 0; i < q; ++i) {
        const int rs = (int)rng.below(rows), rt = (int)rng.below(rows);
        s.push_back(rs * cols + 0);
        t.push_back(rt * cols + (cols - 1));
    }
}

void gen_queries_hub(int n, const std::vector<int32_t>& us, const std::vector<int32_t>& vs,
                     double top_frac, int q, uint64_t seed, std::vector<int32_t>& s,
                     std::vector<int32_t>& t) {
    std::vector<int64_t> dout(n, 0), din(n, 0);
    for (size_t i = 0; i < us.size(); ++i) { dout[us[i]]++; din[vs[i]]++; }
    std::vector<int32_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return dout[a] > dout[b]; });
    int64_t top = (int64_t)std::floor(top_frac * (double)n);
    if (top < 1) top = 1;
    std::vector<int32_t> tin;
    for (int v = 0; v < n; ++v) if (din[v] > 0) tin.push_back(v);
    Rng rng(seed);
    s.clear(); t.clear();
    if (tin.empty()) return;
    while ((int)s.size() < q) {
        const int a = order[rng.below((uint64_t)top)];
        const int b = tin[rng.below(tin.size())];
        if (a == b) continue;
        s.push_back(a); t.push_back(b);
    }
}

int write_graph_file(const std::string& path, int n, const std::vector<int32_t>& us,
                     const std::vector<int32_t>& vs, const std::vector<int64_t>& caps) {
    FILE* f = fopen(path.c_str(), "w");
    if (!f) return -1;
    std::vector<char> buf(1 << 22);
    setvbuf(f, buf.data(), _IOFBF, buf.size());
    fprintf(f, "%d %zu\n", n, us.size());
    for (size_t i = 0; i < us.size(); ++i)
        fprintf(f, "%d %d %lld\n", us[i] + 1, vs[i] + 1, (long long)caps[i]);
    fclose(f);
    return 0;
}

int write_query_file(const std::string& path, const std::vector<int32_t>& s,
                     const std::vector<int32_t>& t) {
    FILE* f = fopen(path.c_str(), "w");
    if (!f) return -1;
    for (size_t i = 0; i < s.size(); ++i) fprintf(f, "%d %d\n", s[i] + 1, t[i] + 1);
    fclose(f);
    return 0;
}

// ------------------------------------------------------- binary graph cache

namespace {

constexpr char kBinMagic[8] = {'M', 'S', 'C', 'M', 'F', 'G', 'R', '1'};
constexpr int kBinSections = 12;

struct BinHeader {
    char magic[8];
    uint32_t version;
    uint32_t header_bytes;
    int32_t n, R;
    int64_t E, lines;
    int64_t src_size, src_mtime_ns;  // -1 when no source file was named
    uint64_t hash;                   // of every section's bytes, in order
    int64_t off[kBinSections], len[kBinSections];  // bytes
};
static_assert(sizeof(BinHeader) == 256, "binary graph header is 256 bytes");

// Four independent multiply-xorshift lanes over 8-byte words, then a final
// mix; the tail shorter than 32 bytes goes byte by byte.  Chained across
// sections through the seed.
uint64_t hash_bytes(const void* p, size_t nbytes, uint64_t seed) {
    const uint64_t K = 0x9fb21c651e98df25ULL;
    uint64_t a[4] = {seed ^ 0x243f6a8885a308d3ULL, seed ^ 0x13198a2e03707344ULL, seed ^ 0xa4093822299f31d0ULL,
                     seed ^ 0x082efa98ec4e6c89ULL};
    const unsigned char* b = (const unsigned char*)p;
    size_t i = 0;
    for (; i + 32 <= nbytes; i += 32)
        for (int k = 0; k < 4; ++k) {
            uint64_t w;
            memcpy(&w, b + i + 8 * k, 8);
            a[k] = (a[k] ^ w) * K;
            a[k] ^= a[k] >> 29;
        }
    uint64_t h = a[0] ^ (a[1] * 3) ^ (a[2] * 5) ^ (a[3] * 7) ^ (uint64_t)nbytes;
    for (; i < nbytes; ++i) h = (h ^ b[i]) * 0x100000001b3ULL;
    h ^= h >> 33;
    h *= 0xff51afd7ed558ccdULL;
    h ^= h >> 33;
    return h;
}

bool source_identity(const std::string& src, int64_t& size, int64_t& mtime_ns) {
    struct stat st;
    if (src.empty() || stat(src.c_str(), &st) != 0) return false;
    size = (int64_t)st.st_size;
    mtime_ns = (int64_t)st.st_mtim.tv_sec * 1000000000LL + (int64_t)st.st_mtim.tv_nsec;
    return true;
}

template <typename T>
void section(std::vector<std::pair<const void*, size_t>>& out, const std::vector<T>& v) {
    out.push_back({v.data(), v.size() * sizeof(T)});
}

}  // namespace

int save_graph_binary(const HostGraph& g, const std::string& path, const std::string& source,
                      std::string& err) {
    std::vector<std::pair<const void*, size_t>> secs;
    section(secs, g.row_ptr);
    section(secs, g.col);
    section(secs, g.cap);
    section(secs, g.rev);
    section(secs, g.sum_out);
    section(secs, g.sum_in);
    section(secs, g.deg_out);
    section(secs, g.deg_in);
    section(secs, g.new_id);
    section(secs, g.old_id);
    section(secs, g.rstart);
    section(secs, g.region_id);
    BinHeader hd;
    memset(&hd, 0, sizeof(hd));
    memcpy(hd.magic, kBinMagic, 8);
    hd.version = 1;
    hd.header_bytes = sizeof(BinHeader);
    hd.n = g.n;
    hd.R = g.R;
    hd.E = g.E();
    hd.lines = g.lines;
    hd.src_size = hd.src_mtime_ns = -1;
    if (!source.empty() && !source_identity(source, hd.src_size, hd.src_mtime_ns)) {
        err = "cannot stat " + source;
        return -1;
    }
    int64_t off = sizeof(BinHeader);
    uint64_t h = 0;
    for (int k = 0; k < kBinSections; ++k) {
        hd.off[k] = off;
        hd.len[k] = (int64_t)secs[k].second;
        off += (hd.len[k] + 63) & ~(int64_t)63;
        h = hash_bytes(secs[k].first, secs[k].second, h);
    }
    hd.hash = h;
    // write beside the target, then rename: a reader never sees half a file
    const std::string tmp = path + ".tmp" + std::to_string((long long)getpid());
    FILE* f = fopen(tmp.c_str(), "wb");
    if (!f) {
        err = "cannot write " + tmp;
        return -1;
    }
    static const char zeros[64] = {0};
    bool ok = fwrite(&hd, sizeof(hd), 1, f) == 1;
    for (int k = 0; k < kBinSections && ok; ++k) {
        if (secs[k].second) ok = fwrite(secs[k].first, 1, secs[k].second, f) == secs[k].second;
        const size_t padb = (size_t)(((hd.len[k] + 63) & ~(int64_t)63) - hd.len[k]);
        if (ok && padb) ok = fwrite(zeros, 1, padb, f) == padb;
    }
    ok = (fclose(f) == 0) && ok;
    if (!ok || rename(tmp.c_str(), path.c_str()) != 0) {
        unlink(tmp.c_str());
        err = "cannot write " + path;
        return -1;
    }
    return 0;
}

int load_graph_binary(const std::string& path, const std::string& source, int R, HostGraph& g,
                      std::string& err) {
    const int fd = open(path.c_str(), O_RDONLY);
    if (fd < 0) {
        err = "cannot open " + path;
        return -1;
    }
    struct stat st;
    if (fstat(fd, &st) != 0 || st.st_size < (off_t)sizeof(BinHeader)) {
        close(fd);
        err = "truncated graph cache " + path;
        return -1;
    }
    const size_t fsize = (size_t)st.st_size;
    void* base = mmap(nullptr, fsize, PROT_READ, MAP_PRIVATE, fd, 0);
    close(fd);
    if (base == MAP_FAILED) {
        err = "cannot map " + path;
        return -1;
    }
    struct Unmap {
        void* p;
        size_t n;
        ~Unmap() { munmap(p, n); }
    } um{base, fsize};
    BinHeader hd;
    memcpy(&hd, base, sizeof(hd));
    if (memcmp(hd.magic, kBinMagic, 8) != 0 || hd.version != 1 || hd.header_bytes != sizeof(BinHeader) ||
        hd.n < 0 || hd.E < 0 || hd.R < 0) {
        err = "not a graph cache: " + path;
        return -1;
    }
    const int64_t n = hd.n, E = hd.E, Rg = hd.R;
    const int64_t want[kBinSections] = {
        4 * (n + 1), 4 * E, 8 * E, 4 * E, 8 * n, 8 * n, 4 * n, 4 * n,
        Rg ? 4 * n : 0, Rg ? 4 * n : 0, Rg ? 4 * (Rg + 1) : 0, Rg ? 4 * n : 0};
    for (int k = 0; k < kBinSections; ++k)
        if ((hd.len[k] != want[k] && !(k >= 4 && k <= 7 && hd.len[k] == 0)) ||  // stats may be absent
            hd.off[k] < (int64_t)sizeof(BinHeader) || hd.off[k] % 64 != 0 ||
            hd.off[k] + hd.len[k] > (int64_t)fsize) {
            err = "truncated or inconsistent graph cache " + path;
            return -1;
        }
    if (!source.empty()) {
        int64_t sz = 0, mt = 0;
        if (!source_identity(source, sz, mt) || sz != hd.src_size || mt != hd.src_mtime_ns) return 1;
    }
    if (R >= 0 && R != hd.R) return 1;
    const char* b = (const char*)base;
    uint64_t h = 0;
    for (int k = 0; k < kBinSections; ++k) h = hash_bytes(b + hd.off[k], (size_t)hd.len[k], h);
    if (h != hd.hash) {
        err = "graph cache hash mismatch: " + path;
        return -1;
    }
    auto take = [&](int k, auto& v) {
        using T = typename std::decay_t<decltype(v)>::value_type;
        const T* p = (const T*)(b + hd.off[k]);
        v.assign(p, p + hd.len[k] / (int64_t)sizeof(T));
    };
    HostGraph out;
    out.n = hd.n;
    out.lines = hd.lines;
    out.R = hd.R;
    take(0, out.row_ptr);
    take(1, out.col);
    take(2, out.cap);
    take(3, out.rev);
    take(4, out.sum_out);
    take(5, out.sum_in);
    take(6, out.deg_out);
    take(7, out.deg_in);
    take(8, out.new_id);
    take(9, out.old_id);
    take(10, out.rstart);
    take(11, out.region_id);
    if (out.row_ptr.empty() || out.row_ptr[0] != 0 || out.row_ptr[n] != (int32_t)E) {
        err = "inconsistent graph cache " + path;
        return -1;
    }
    g = std::move(out);
    return 0;
}

}  // namespace mf
