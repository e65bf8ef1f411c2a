// scc.cpp -- SCC ids in the reference's numbering, the contracted graph and
// exact query routing by component.  Host code: the ids come from a
// sequential depth-first order (SCC_CONTRACT/scc_contract.cpp:44-105), which
// the maps and split files the reference writes depend on.
#include "scc.h"

#include <algorithm>
#include <utility>

namespace mf {

namespace {

// Lines grouped by key[i] (a CSR over n rows), line order kept within a row.
void rows_by(int n, int64_t m, const int32_t* key, const int32_t* other, std::vector<int64_t>& ptr,
             std::vector<int32_t>& nb) {
    ptr.assign((size_t)n + 1, 0);
    for (int64_t i = 0; i < m; ++i) ++ptr[(size_t)key[i] + 1];
    for (int x = 0; x < n; ++x) ptr[x + 1] += ptr[x];
    nb.resize((size_t)m);
    std::vector<int64_t> pos(ptr.begin(), ptr.end() - 1);
    for (int64_t i = 0; i < m; ++i) nb[(size_t)pos[key[i]]++] = other[i];
}

}  // namespace

void scc_contract(int n, int64_t m, const int32_t* u, const int32_t* v, const int64_t* cap, SccContract& out) {
    out = SccContract();
    out.n = n;
    std::vector<int64_t> optr, iptr;
    std::vector<int32_t> onb, inb;
    rows_by(n, m, u, v, optr, onb);
    rows_by(n, m, v, u, iptr, inb);

    // Pass 1 (scc_contract.cpp:44-81): roots 0..n-1 in turn; a vertex is
    // recorded once its last out-line has been scanned.
    std::vector<int32_t> order;
    order.reserve((size_t)n);
    std::vector<uint8_t> seen((size_t)n, 0);
    std::vector<int32_t> stk;
    std::vector<int64_t> next;  // per stack entry: its next out-line
    for (int r = 0; r < n; ++r) {
        if (seen[r]) continue;
        seen[r] = 1;
        stk.push_back(r);
        next.push_back(optr[r]);
        while (!stk.empty()) {
            const int x = stk.back();
            const int64_t k = next.back();
            if (k < optr[(size_t)x + 1]) {
                next.back() = k + 1;
                const int y = onb[(size_t)k];
                if (!seen[y]) {
                    seen[y] = 1;
                    stk.push_back(y);
                    next.push_back(optr[y]);
                }
            } else {
                order.push_back(x);
                stk.pop_back();
                next.pop_back();
            }
        }
    }

    // Pass 2 (scc_contract.cpp:83-103): in reverse recording order, each
    // vertex not yet labelled opens the next id, which spreads over in-lines.
    out.comp.assign((size_t)n, -1);
    int32_t C = 0;
    for (int64_t i = (int64_t)order.size() - 1; i >= 0; --i) {
        const int r = order[(size_t)i];
        if (out.comp[r] >= 0) continue;
        out.comp[r] = C;
        stk.clear();
        stk.push_back(r);
        while (!stk.empty()) {
            const int x = stk.back();
            stk.pop_back();
            for (int64_t k = iptr[x]; k < iptr[(size_t)x + 1]; ++k) {
                const int y = inb[(size_t)k];
                if (out.comp[y] < 0) {
                    out.comp[y] = C;
                    stk.push_back(y);
                }
            }
        }
        ++C;
    }
    out.ncomp = C;

    // Contraction (scc_contract.cpp:107-125): lines between components,
    // bucketed by tail component, sorted by head component, summed.
    std::vector<int64_t> bptr((size_t)C + 1, 0);
    for (int64_t i = 0; i < m; ++i) {
        const int32_t a = out.comp[u[i]], b = out.comp[v[i]];
        if (a != b) ++bptr[(size_t)a + 1];
    }
    for (int32_t c = 0; c < C; ++c) bptr[c + 1] += bptr[c];
    std::vector<std::pair<int32_t, int64_t>> ent((size_t)bptr[C]);
    {
        std::vector<int64_t> pos(bptr.begin(), bptr.end() - 1);
        for (int64_t i = 0; i < m; ++i) {
            const int32_t a = out.comp[u[i]], b = out.comp[v[i]];
            if (a != b) ent[(size_t)pos[a]++] = {b, cap[i]};
        }
    }
    out.crow.assign((size_t)C + 1, 0);
    for (int32_t c = 0; c < C; ++c) {
        auto lo = ent.begin() + bptr[c], hi = ent.begin() + bptr[(size_t)c + 1];
        std::sort(lo, hi, [](const std::pair<int32_t, int64_t>& x, const std::pair<int32_t, int64_t>& y) {
            return x.first < y.first;
        });
        for (auto it = lo; it != hi;) {
            const int32_t b = it->first;
            int64_t sum = 0;
            for (; it != hi && it->first == b; ++it) sum += it->second;
            out.cu.push_back(c);
            out.cv.push_back(b);
            out.ccap.push_back(sum);
        }
        out.crow[(size_t)c + 1] = (int64_t)out.cu.size();
    }
}

void scc_route(const SccContract& c, int nq, const int32_t* s, const int32_t* t, int32_t* route) {
    std::vector<int> pend;
    for (int q = 0; q < nq; ++q) {
        if (s[q] < 0 || t[q] < 0 || s[q] >= c.n || t[q] >= c.n || s[q] == t[q]) { route[q] = -1; continue; }
        const int32_t a = c.comp[s[q]], b = c.comp[t[q]];
        if (a == b) route[q] = 1;
        else if (a > b) route[q] = 0;  // ids are topological: nothing leads back
        else pend.push_back(q);
    }
    std::sort(pend.begin(), pend.end(), [&](int x, int y) { return c.comp[s[x]] < c.comp[s[y]]; });
    std::vector<int32_t> stamp((size_t)c.ncomp, -1);
    std::vector<int32_t> queue;
    int32_t group = 0;
    for (size_t i = 0; i < pend.size();) {
        const int32_t a = c.comp[s[pend[i]]];
        size_t j = i;
        int32_t lim = a;
        for (; j < pend.size() && c.comp[s[pend[j]]] == a; ++j) lim = std::max(lim, c.comp[t[pend[j]]]);
        // components reachable from a, up to the largest target id
        queue.clear();
        queue.push_back(a);
        stamp[a] = group;
        for (size_t h = 0; h < queue.size(); ++h) {
            const int32_t x = queue[h];
            for (int64_t k = c.crow[x]; k < c.crow[(size_t)x + 1]; ++k) {
                const int32_t y = c.cv[(size_t)k];
                if (y <= lim && stamp[y] != group) {
                    stamp[y] = group;
                    queue.push_back(y);
                }
            }
        }
        for (size_t k = i; k < j; ++k) route[pend[k]] = stamp[c.comp[t[pend[k]]]] == group ? 2 : 0;
        ++group;
        i = j;
    }
}

}  // namespace mf
