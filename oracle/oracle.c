/*
 * oracle.c -- CPU restatement of the reference's max-flow path (test infrastructure).
 *
 * This file is the checker.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product never does.
 * Citations "BC:n" refer to Region_Aware_PR_with_Pruning_BoundaryCore.cu.
 * Every function is pinned by an independent check listed in oracle.h.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ---------------------------------------------------------------- parsing */

/* BC:1052-1058: header "n m" read with %d %d. */
int orc_parse_header(const char* path, int* n, int64_t* m) {
    FILE* f = fopen(path, "r");
    if (!f) return -1;
    int nn, mm;
    int ok = fscanf(f, "%d %d", &nn, &mm);
    fclose(f);
    if (ok != 2) return -2;
    *n = nn;
    *m = mm;
    return 0;
}

/* BC:1066-1076: up to m reads of "%d %d %lld"; stop on the first failed read,
 * skip ids out of [1,n], self-loops and cap<=0. */
int64_t orc_parse_edges(const char* path, int64_t m, int32_t* us, int32_t* vs, int64_t* caps) {
    FILE* f = fopen(path, "r");
    if (!f) return -1;
    int n, mm;
    if (fscanf(f, "%d %d", &n, &mm) != 2) { fclose(f); return -2; }
    int64_t cnt = 0;
    for (int64_t i = 0; i < m; ++i) {
        int u, v;
        long long c;
        if (fscanf(f, "%d %d %lld", &u, &v, &c) != 3) break;
        --u; --v;
        if (u < 0 || u >= n || v < 0 || v >= n) continue;
        if (u == v) continue;
        if (c <= 0) continue;
        us[cnt] = u; vs[cnt] = v; caps[cnt] = (int64_t)c;
        ++cnt;
    }
    fclose(f);
    return cnt;
}

/* BC:2603-2621: read pairs until a read fails; 1-based to 0-based; any id out
 * of range makes the query (-1,-1).  The caller maps ids afterwards. */
int64_t orc_parse_queries(const char* path, int64_t cap_q, int32_t* s, int32_t* t) {
    FILE* f = fopen(path, "r");
    if (!f) return -1;
    int64_t q = 0;
    int u, v;
    while (q < cap_q && fscanf(f, "%d %d", &u, &v) == 2) {
        s[q] = u - 1;
        t[q] = v - 1;
        ++q;
    }
    fclose(f);
    return q;
}

/* -------------------------------------------------------------------- CSR */

/* BC:1078-1125: forward entry appended to u's list, then the reverse entry to
 * v's list, line by line; rev links the two.  Stats as BC:1090-1096. */
void orc_build_csr(int n, int64_t cnt, const int32_t* us, const int32_t* vs, const int64_t* caps,
                   int32_t* row_ptr, int32_t* col, int64_t* cap, int32_t* rev,
                   int64_t* sum_out, int64_t* sum_in, int32_t* deg_out, int32_t* deg_in) {
    int32_t* deg = (int32_t*)calloc((size_t)n + 1, sizeof(int32_t));
    for (int64_t i = 0; i < cnt; ++i) { deg[us[i]]++; deg[vs[i]]++; }
    row_ptr[0] = 0;
    for (int v = 0; v < n; ++v) row_ptr[v + 1] = row_ptr[v] + deg[v];
    int32_t* cur = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
    for (int v = 0; v < n; ++v) cur[v] = row_ptr[v];
    if (sum_out) for (int v = 0; v < n; ++v) { sum_out[v] = 0; sum_in[v] = 0; deg_out[v] = 0; deg_in[v] = 0; }
    for (int64_t i = 0; i < cnt; ++i) {
        int u = us[i], v = vs[i];
        int pu = cur[u]++;
        int pv = cur[v]++;
        col[pu] = v; cap[pu] = caps[i]; rev[pu] = pv;
        col[pv] = u; cap[pv] = 0;       rev[pv] = pu;
        if (sum_out) { sum_out[u] += caps[i]; sum_in[v] += caps[i]; deg_out[u]++; deg_in[v]++; }
    }
    free(cur);
    free(deg);
}

/* Forward entry of each accepted line in the CSR orc_build_csr makes: line i
 * takes the next free place of row us[i] (BC:1078-1125), so replaying the
 * row cursors gives its entry. */
void orc_line_entries(int n, int64_t cnt, const int32_t* us, const int32_t* vs, const int32_t* row_ptr,
                      int32_t* entry) {
    int32_t* cur = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
    for (int v = 0; v < n; ++v) cur[v] = row_ptr[v];
    for (int64_t i = 0; i < cnt; ++i) {
        entry[i] = cur[us[i]]++;
        cur[vs[i]]++;
    }
    free(cur);
}

/* BC:2707-2720: stats over CSR entries with cap>0. */
void orc_csr_stats(int n, const int32_t* row_ptr, const int32_t* col, const int64_t* cap,
                   int64_t* sum_out, int64_t* sum_in, int32_t* deg_out, int32_t* deg_in) {
    for (int v = 0; v < n; ++v) { sum_out[v] = 0; sum_in[v] = 0; deg_out[v] = 0; deg_in[v] = 0; }
    for (int v = 0; v < n; ++v)
        for (int e = row_ptr[v]; e < row_ptr[v + 1]; ++e)
            if (cap[e] > 0) {
                int u = col[e];
                sum_out[v] += cap[e]; sum_in[u] += cap[e];
                deg_out[v]++; deg_in[u]++;
            }
}

/* -------------------------------------------------------------- partition */

typedef struct { int v; double score; } vinfo_t;

static int cmp_vinfo(const void* a, const void* b) {
    const vinfo_t* x = (const vinfo_t*)a;
    const vinfo_t* y = (const vinfo_t*)b;
    if (x->score > y->score) return -1;
    if (x->score < y->score) return 1;
    return (x->v < y->v) ? -1 : (x->v > y->v);
}

/* BC:1141-1177.  score = [ln(so*do+1)+ln(si*di+1)] / ln(sqrt(do*di)+1), or 1
 * when the denominator is not positive; descending sort (ties: lower id
 * first); each vertex goes to the region with the least load, lowest index on
 * ties, and adds its score to that load. */
void orc_partition(int n, int R, const int64_t* sum_out, const int64_t* sum_in,
                   const int32_t* deg_out, const int32_t* deg_in, double* score, int32_t* part) {
    vinfo_t* vi = (vinfo_t*)malloc(sizeof(vinfo_t) * (size_t)(n > 0 ? n : 1));
    for (int u = 0; u < n; ++u) {
        double so = (double)sum_out[u], si = (double)sum_in[u];
        double dout = (double)deg_out[u], din = (double)deg_in[u];
        double num = log(so * dout + 1.0) + log(si * din + 1.0);
        double prod = dout * din;
        double den = log(sqrt(prod) + 1.0);
        double sc = (den > 0.0) ? num / den : 1.0;
        vi[u].v = u;
        vi[u].score = sc;
        if (score) score[u] = sc;
    }
    qsort(vi, (size_t)n, sizeof(vinfo_t), cmp_vinfo);
    double* load = (double*)calloc((size_t)R, sizeof(double));
    for (int i = 0; i < n; ++i) {
        int best = 0;
        double bl = load[0];
        for (int r = 1; r < R; ++r)
            if (load[r] < bl) { bl = load[r]; best = r; }
        part[vi[i].v] = best;
        load[best] += vi[i].score;
    }
    free(load);
    free(vi);
}

/* BC:1182-1260 and BC:1263-1276.  Regions laid out in index order, vertices
 * inside a region in ascending old id; each row keeps its entry order. */
void orc_reorder(int n, int R, const int32_t* part,
                 const int32_t* row_ptr, const int32_t* col, const int64_t* cap, const int32_t* rev,
                 int32_t* new_id, int32_t* old_id, int32_t* rstart, int32_t* rend, int32_t* region_id,
                 int32_t* row_ptr2, int32_t* col2, int64_t* cap2, int32_t* rev2) {
    int32_t* cnt = (int32_t*)calloc((size_t)R + 1, sizeof(int32_t));
    for (int v = 0; v < n; ++v) {
        int r = part[v];
        if (r < 0 || r >= R) r = 0;
        cnt[r]++;
    }
    int cur = 0;
    for (int r = 0; r < R; ++r) { rstart[r] = cur; cur += cnt[r]; rend[r] = cur; }
    int32_t* fill = (int32_t*)malloc(sizeof(int32_t) * ((size_t)R + 1));
    for (int r = 0; r < R; ++r) fill[r] = rstart[r];
    for (int v = 0; v < n; ++v) {
        int r = part[v];
        if (r < 0 || r >= R) r = 0;
        int id = fill[r]++;
        new_id[v] = id;
        old_id[id] = v;
    }
    for (int r = 0; r < R; ++r)
        for (int v = rstart[r]; v < rend[r]; ++v) region_id[v] = r;
    row_ptr2[0] = 0;
    for (int u = 0; u < n; ++u) {
        int o = old_id[u];
        row_ptr2[u + 1] = row_ptr2[u] + (row_ptr[o + 1] - row_ptr[o]);
    }
    /* old entry e of old vertex o maps to row_ptr2[new_id[o]] + (e - row_ptr[o]) */
    for (int o = 0; o < n; ++o) {
        int base = row_ptr2[new_id[o]];
        for (int e = row_ptr[o]; e < row_ptr[o + 1]; ++e) {
            int ne = base + (e - row_ptr[o]);
            col2[ne] = new_id[col[e]];
            cap2[ne] = cap[e];
            int re = rev[e];
            int ro = col[e]; /* rev[e] lives in the row of col[e] */
            rev2[ne] = row_ptr2[new_id[ro]] + (re - row_ptr[ro]);
        }
    }
    free(fill);
    free(cnt);
}

/* ---------------------------------------------------------------- pruning */

/* BC:1391-1445. */
void orc_prune_mask(int n, const int32_t* row_ptr, const int32_t* col, const int64_t* cap,
                    int s, int t, int32_t* mask) {
    for (int v = 0; v < n; ++v) mask[v] = 0;
    if (s < 0 || s >= n || t < 0 || t >= n) return;
    char* fwd = (char*)calloc((size_t)n, 1);
    char* und = (char*)calloc((size_t)n, 1);
    int32_t* q = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int qh = 0, qt = 0;
    fwd[s] = 1; q[qt++] = s;
    while (qh < qt) {
        int v = q[qh++];
        for (int e = row_ptr[v]; e < row_ptr[v + 1]; ++e) {
            int u = col[e];
            if (cap[e] > 0 && !fwd[u]) { fwd[u] = 1; q[qt++] = u; }
        }
    }
    qh = qt = 0;
    und[t] = 1; q[qt++] = t;
    while (qh < qt) {
        int v = q[qh++];
        for (int e = row_ptr[v]; e < row_ptr[v + 1]; ++e) {
            int u = col[e];
            if (!und[u]) { und[u] = 1; q[qt++] = u; }
        }
    }
    for (int v = 0; v < n; ++v) mask[v] = (fwd[v] && und[v]) ? 1 : 0;
    mask[s] = 1;
    mask[t] = 1;
    free(q); free(und); free(fwd);
}

/* BC:2477-2551: kept vertices in ascending id, kept entries in row order. */
int orc_induced(int n, const int32_t* row_ptr, const int32_t* col, const int64_t* cap,
                const int32_t* rev, const int32_t* mask, int32_t* old2sub, int32_t* sub2old,
                int32_t* row_ptr2, int32_t* col2, int64_t* cap2, int32_t* rev2, int64_t* m_out) {
    int sn = 0;
    for (int v = 0; v < n; ++v) {
        if (mask[v]) { old2sub[v] = sn; sub2old[sn] = v; ++sn; }
        else old2sub[v] = -1;
    }
    int64_t E = row_ptr[n];
    int32_t* e2s = (int32_t*)malloc(sizeof(int32_t) * (size_t)(E > 0 ? E : 1));
    for (int64_t e = 0; e < E; ++e) e2s[e] = -1;
    row_ptr2[0] = 0;
    int pos = 0;
    for (int i = 0; i < sn; ++i) {
        int vo = sub2old[i];
        for (int e = row_ptr[vo]; e < row_ptr[vo + 1]; ++e) {
            int us = old2sub[col[e]];
            if (us != -1) { col2[pos] = us; cap2[pos] = cap[e]; e2s[e] = pos; ++pos; }
        }
        row_ptr2[i + 1] = pos;
    }
    for (int i = 0; i < sn; ++i) {
        int vo = sub2old[i];
        for (int e = row_ptr[vo]; e < row_ptr[vo + 1]; ++e) {
            int se = e2s[e];
            if (se == -1) continue;
            int re = rev[e];
            rev2[se] = (re >= 0 && re < E) ? e2s[re] : -1;
        }
    }
    free(e2s);
    *m_out = pos;
    return sn;
}

/* BC:2340-2475 (coarse region graph, boundary vertices, regions on any
 * region path s->t, hop-limited in-region growth from boundary vertices). */
static int cmp_int(const void* a, const void* b) {
    int x = *(const int*)a, y = *(const int*)b;
    return (x > y) - (x < y);
}

void orc_core_mask(int n, int R, const int32_t* row_ptr, const int32_t* col,
                   const int32_t* region_id, int s, int t, int hop_limit, int32_t* mask,
                   int64_t* kept_regions) {
    for (int v = 0; v < n; ++v) mask[v] = 0;
    *kept_regions = 0;
    if (s < 0 || s >= n || t < 0 || t >= n) return;
    /* dense R x R adjacency is fine for the oracle's sizes */
    char* adj = (char*)calloc((size_t)R * (size_t)R, 1);
    char* bnd = (char*)calloc((size_t)n, 1);
    for (int v = 0; v < n; ++v) {
        int rv = region_id[v];
        for (int e = row_ptr[v]; e < row_ptr[v + 1]; ++e) {
            int ru = region_id[col[e]];
            if (rv != ru) { adj[(size_t)rv * R + ru] = 1; bnd[v] = 1; }
        }
    }
    char* fs = (char*)calloc((size_t)R, 1);
    char* tt = (char*)calloc((size_t)R, 1);
    int* q = (int*)malloc(sizeof(int) * (size_t)(n > R ? n : R));
    int qh = 0, qt = 0;
    int rs = region_id[s], rt = region_id[t];
    fs[rs] = 1; q[qt++] = rs;
    while (qh < qt) {
        int r = q[qh++];
        for (int x = 0; x < R; ++x) if (adj[(size_t)r * R + x] && !fs[x]) { fs[x] = 1; q[qt++] = x; }
    }
    qh = qt = 0;
    tt[rt] = 1; q[qt++] = rt;
    while (qh < qt) {
        int r = q[qh++];
        for (int x = 0; x < R; ++x) if (adj[(size_t)x * R + r] && !tt[x]) { tt[x] = 1; q[qt++] = x; }
    }
    char* keep = (char*)calloc((size_t)R, 1);
    int64_t kept = 0;
    for (int r = 0; r < R; ++r) if (fs[r] && tt[r]) { keep[r] = 1; ++kept; }
    *kept_regions = kept;
    int* dist = (int*)malloc(sizeof(int) * (size_t)n);
    qh = qt = 0;
    for (int v = 0; v < n; ++v) {
        dist[v] = -1;
        int r = region_id[v];
        if (!keep[r]) continue;
        if (r == rs || r == rt || bnd[v]) { mask[v] = 1; dist[v] = 0; q[qt++] = v; }
    }
    while (qh < qt) {
        int v = q[qh++];
        if (dist[v] >= hop_limit) continue;
        int rv = region_id[v];
        for (int e = row_ptr[v]; e < row_ptr[v + 1]; ++e) {
            int u = col[e];
            if (region_id[u] != rv) continue;
            if (!keep[rv]) continue;
            if (dist[u] == -1) { dist[u] = dist[v] + 1; mask[u] = 1; q[qt++] = u; }
        }
    }
    mask[s] = 1;
    mask[t] = 1;
    (void)cmp_int;
    free(dist); free(keep); free(q); free(tt); free(fs); free(bnd); free(adj);
}

/* ------------------------------------------------------- push-relabel bits */

/* BC:151-170 */
void orc_init_residual(int n, int64_t E, const int64_t* cap, int64_t* res, int64_t* excess) {
    for (int64_t e = 0; e < E; ++e) res[e] = cap[e];
    for (int v = 0; v < n; ++v) excess[v] = 0;
}

/* BC:173-200 (without the mirrored rev_res array) */
void orc_preflow(int s, const int32_t* row_ptr, const int32_t* col, const int32_t* rev,
                 int64_t* res, int64_t* excess) {
    for (int e = row_ptr[s]; e < row_ptr[s + 1]; ++e) {
        int64_t d = res[e];
        if (d > 0) {
            excess[col[e]] += d;
            excess[s] -= d;
            res[e] -= d;
            res[rev[e]] += d;
        }
    }
}

/* BC:203-233 + BC:322-348, with s never labelled. */
int64_t orc_bfs_heights(int n, const int32_t* row_ptr, const int32_t* col, const int32_t* rev,
                        const int64_t* res, const int64_t* excess, const int32_t* mask,
                        int s, int t, int32_t* height) {
    int32_t* q = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int qh = 0, qt = 0;
    for (int v = 0; v < n; ++v) height[v] = n;
    for (int v = 0; v < n; ++v) {
        if (v == s) continue;
        if (v != t && mask && mask[v] == 0) continue;
        if (v == t || (excess && excess[v] < 0)) { height[v] = 0; q[qt++] = v; }
    }
    while (qh < qt) {
        int v = q[qh++];
        for (int e = row_ptr[v]; e < row_ptr[v + 1]; ++e) {
            int w = col[e];
            if (w == s) continue;
            if (res[rev[e]] > 0 && height[w] == n) { height[w] = height[v] + 1; q[qt++] = w; }
        }
    }
    free(q);
    return qt;
}

/* ------------------------------------------------------------------ Dinic */

typedef struct {
    int n;
    const int32_t *row_ptr, *col, *rev;
    int64_t* res;
    int32_t *level, *it, *q, *path;
    /* optional: vertices outside the mask are never entered (level -2) */
    const uint8_t* mask;
    /* optional: entries whose residual changed, so that a caller holding
     * res == cap between queries restores only those (touched[e] = 1) */
    uint8_t* touched;
    int32_t* tlist;
    int64_t ntouched;
} dinic_ws;

static int dinic_bfs(dinic_ws* w, int s, int t) {
    int n = w->n;
    if (w->mask)
        for (int v = 0; v < n; ++v) w->level[v] = w->mask[v] ? -1 : -2;
    else
        for (int v = 0; v < n; ++v) w->level[v] = -1;
    int qh = 0, qt = 0;
    w->level[s] = 0; w->q[qt++] = s;
    int tl = -1;
    while (qh < qt) {
        int v = w->q[qh++];
        if (tl >= 0 && w->level[v] >= tl) break;
        for (int e = w->row_ptr[v]; e < w->row_ptr[v + 1]; ++e) {
            int u = w->col[e];
            if (w->res[e] > 0 && w->level[u] == -1) {
                w->level[u] = w->level[v] + 1;
                if (u == t) tl = w->level[u];
                w->q[qt++] = u;
            }
        }
    }
    return w->level[t] >= 0;
}

static int64_t dinic_run(dinic_ws* w, int s, int t) {
    int64_t total = 0;
    if (s == t) return 0;
    while (dinic_bfs(w, s, t)) {
        for (int v = 0; v < w->n; ++v) w->it[v] = w->row_ptr[v];
        int top = 0, v = s;
        for (;;) {
            if (v == t) {
                int64_t b = INT64_MAX;
                for (int i = 0; i < top; ++i) if (w->res[w->path[i]] < b) b = w->res[w->path[i]];
                for (int i = 0; i < top; ++i) {
                    int e = w->path[i];
                    w->res[e] -= b;
                    w->res[w->rev[e]] += b;
                    if (w->touched && !w->touched[e]) {
                        w->touched[e] = 1;
                        w->tlist[w->ntouched++] = e;
                    }
                }
                total += b;
                int k = 0;
                while (k < top && w->res[w->path[k]] != 0) ++k;
                top = k;
                v = (top == 0) ? s : w->col[w->path[top - 1]];
                continue;
            }
            int e = w->it[v], end = w->row_ptr[v + 1];
            int lv = w->level[v] + 1;
            while (e < end && !(w->res[e] > 0 && w->level[w->col[e]] == lv)) ++e;
            w->it[v] = e;
            if (e < end) {
                w->path[top++] = e;
                v = w->col[e];
            } else {
                if (v == s) break;
                w->level[v] = -1;
                --top;
                v = (top == 0) ? s : w->col[w->path[top - 1]];
            }
        }
    }
    return total;
}

int64_t orc_dinic(int n, const int32_t* row_ptr, const int32_t* col, const int64_t* cap,
                  const int32_t* rev, int s, int t, int64_t* res_out) {
    if (s < 0 || t < 0 || s >= n || t >= n || s == t) return 0;
    int64_t E = row_ptr[n];
    dinic_ws w;
    memset(&w, 0, sizeof(w));
    w.n = n; w.row_ptr = row_ptr; w.col = col; w.rev = rev;
    w.res = res_out ? res_out : (int64_t*)malloc(sizeof(int64_t) * (size_t)(E > 0 ? E : 1));
    for (int64_t e = 0; e < E; ++e) w.res[e] = cap[e];
    w.level = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    w.it = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    w.q = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    w.path = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
    int64_t f = dinic_run(&w, s, t);
    free(w.path); free(w.q); free(w.it); free(w.level);
    if (!res_out) free(w.res);
    return f;
}

/* ------------------------------------------------------------ certificate */

int orc_certify(int n, const int32_t* row_ptr, const int32_t* col, const int64_t* cap,
                const int32_t* rev, const int64_t* res, const int64_t* excess_or_null,
                int s, int t, int64_t flow) {
    int64_t E = row_ptr[n];
    for (int64_t e = 0; e < E; ++e) {
        if (res[e] < 0) return 1;
        if (res[e] + res[rev[e]] != cap[e] + cap[rev[e]]) return 2;
    }
    int64_t* ex = (int64_t*)calloc((size_t)n, sizeof(int64_t));
    for (int v = 0; v < n; ++v)
        for (int e = row_ptr[v]; e < row_ptr[v + 1]; ++e) ex[v] += res[e] - cap[e];
    int rc = 0;
    if (excess_or_null)
        for (int v = 0; v < n && !rc; ++v) if (ex[v] != excess_or_null[v]) rc = 8;
    for (int v = 0; v < n && !rc; ++v) if (v != s && ex[v] < 0) rc = 3;
    if (rc) { free(ex); return rc; }
    char* inT = (char*)calloc((size_t)n, 1);
    int32_t* q = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int qh = 0, qt = 0;
    inT[t] = 1; q[qt++] = t;
    while (qh < qt) {
        int v = q[qh++];
        for (int e = row_ptr[v]; e < row_ptr[v + 1]; ++e) {
            int w = col[e];
            if (!inT[w] && res[rev[e]] > 0) { inT[w] = 1; q[qt++] = w; }
        }
    }
    if (inT[s]) rc = 4;
    for (int v = 0; v < n && !rc; ++v) if (inT[v] && v != t && ex[v] != 0) rc = 5;
    if (!rc) {
        int64_t cut = 0;
        for (int u = 0; u < n; ++u) {
            if (inT[u]) continue;
            for (int e = row_ptr[u]; e < row_ptr[u + 1]; ++e) if (inT[col[e]]) cut += cap[e];
        }
        if (cut != ex[t]) rc = 6;
        else if (ex[t] != flow) rc = 7;
    }
    free(q); free(inT); free(ex);
    return rc;
}

/* --------------------------------------------------------- threaded Dinic */

typedef struct {
    int n;
    const int32_t *row_ptr, *col, *rev;
    const int64_t* cap;
    int nq;
    const int32_t *s, *t;
    int64_t* flows;
    int prune;
    atomic_int next;
} many_ctx;

/* The pruning mask of BC:1391-1445 (fwd(s) over cap > 0, AND the undirected
 * component of t, plus s and t) into mask8, using q as the BFS queue.
 * Returns 0 when t is not reachable from s over cap > 0 (the flow is 0). */
static int prune_mask8(int n, const int32_t* row_ptr, const int32_t* col, const int64_t* cap, int s, int t,
                       uint8_t* fwd, uint8_t* mask8, int32_t* q) {
    memset(fwd, 0, (size_t)n);
    memset(mask8, 0, (size_t)n);
    int qh = 0, qt = 0;
    fwd[s] = 1; q[qt++] = s;
    while (qh < qt) {
        int v = q[qh++];
        for (int e = row_ptr[v]; e < row_ptr[v + 1]; ++e)
            if (cap[e] > 0 && !fwd[col[e]]) { fwd[col[e]] = 1; q[qt++] = col[e]; }
    }
    if (!fwd[t]) return 0;
    qh = qt = 0;
    mask8[t] = 1; q[qt++] = t;
    while (qh < qt) {
        int v = q[qh++];
        for (int e = row_ptr[v]; e < row_ptr[v + 1]; ++e)
            if (!mask8[col[e]]) { mask8[col[e]] = 1; q[qt++] = col[e]; }
    }
    for (int v = 0; v < n; ++v) mask8[v] = (uint8_t)(mask8[v] & fwd[v]);
    mask8[s] = mask8[t] = 1;
    return 1;
}

/* One thread: the residual is copied from cap once and, after each query,
 * only the entries Dinic changed are restored. */
static void* many_worker(void* arg) {
    many_ctx* c = (many_ctx*)arg;
    int64_t E = c->row_ptr[c->n];
    const size_t ee = (size_t)(E > 0 ? E : 1), nn = (size_t)(c->n > 0 ? c->n : 1);
    dinic_ws w;
    memset(&w, 0, sizeof(w));
    w.n = c->n; w.row_ptr = c->row_ptr; w.col = c->col; w.rev = c->rev;
    w.res = (int64_t*)malloc(sizeof(int64_t) * ee);
    w.level = (int32_t*)malloc(sizeof(int32_t) * nn);
    w.it = (int32_t*)malloc(sizeof(int32_t) * nn);
    w.q = (int32_t*)malloc(sizeof(int32_t) * nn);
    w.path = (int32_t*)malloc(sizeof(int32_t) * (nn + 1));
    w.touched = (uint8_t*)calloc(ee, 1);
    w.tlist = (int32_t*)malloc(sizeof(int32_t) * ee);
    uint8_t* fwd = c->prune ? (uint8_t*)malloc(nn) : NULL;
    uint8_t* mask8 = c->prune ? (uint8_t*)malloc(nn) : NULL;
    memcpy(w.res, c->cap, sizeof(int64_t) * (size_t)E);
    for (;;) {
        int q = atomic_fetch_add(&c->next, 1);
        if (q >= c->nq) break;
        int s = c->s[q], t = c->t[q];
        if (s < 0 || t < 0 || s >= c->n || t >= c->n || s == t) { c->flows[q] = 0; continue; }
        w.mask = NULL;
        if (c->prune) {
            if (!prune_mask8(c->n, c->row_ptr, c->col, c->cap, s, t, fwd, mask8, w.q)) { c->flows[q] = 0; continue; }
            w.mask = mask8;
        }
        w.ntouched = 0;
        c->flows[q] = dinic_run(&w, s, t);
        for (int64_t k = 0; k < w.ntouched; ++k) {
            const int e = w.tlist[k];
            w.res[e] = c->cap[e];
            w.res[c->rev[e]] = c->cap[c->rev[e]];
            w.touched[e] = 0;
        }
    }
    free(mask8); free(fwd);
    free(w.tlist); free(w.touched);
    free(w.path); free(w.q); free(w.it); free(w.level); free(w.res);
    return NULL;
}

double orc_dinic_many_ex(int n, const int32_t* row_ptr, const int32_t* col, const int64_t* cap,
                         const int32_t* rev, int nq, const int32_t* s, const int32_t* t,
                         int threads, int prune, int64_t* flows) {
    many_ctx c;
    c.n = n; c.row_ptr = row_ptr; c.col = col; c.rev = rev; c.cap = cap;
    c.nq = nq; c.s = s; c.t = t; c.flows = flows; c.prune = prune;
    atomic_init(&c.next, 0);
    if (threads < 1) threads = 1;
    struct timespec a, b;
    clock_gettime(CLOCK_MONOTONIC, &a);
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
    for (int i = 0; i < threads; ++i) pthread_create(&th[i], NULL, many_worker, &c);
    for (int i = 0; i < threads; ++i) pthread_join(th[i], NULL);
    clock_gettime(CLOCK_MONOTONIC, &b);
    free(th);
    return (double)(b.tv_sec - a.tv_sec) + 1e-9 * (double)(b.tv_nsec - a.tv_nsec);
}

double orc_dinic_many(int n, const int32_t* row_ptr, const int32_t* col, const int64_t* cap,
                      const int32_t* rev, int nq, const int32_t* s, const int32_t* t,
                      int threads, int64_t* flows) {
    return orc_dinic_many_ex(n, row_ptr, col, cap, rev, nq, s, t, threads, 0, flows);
}

/* ------------------------------------------------------- SCC contraction */

/* SCC_CONTRACT/scc_contract.cpp:31-105 restated (Kosaraju).  Adjacency in
 * line order.  Pass 1 (:44-81): for r = 0..n-1 not yet seen, a depth-first
 * search over out-lines; a vertex is appended to `order` once every one of its
 * out-lines has been scanned.  Pass 2 (:83-103): walking `order` from its last
 * vertex to its first, a vertex without an id opens the next id, which then
 * spreads over in-lines to every vertex without one. */
int orc_scc_kosaraju(int n, int64_t cnt, const int32_t* us, const int32_t* vs, int32_t* comp) {
    int64_t* op = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
    int64_t* ip = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
    int32_t* oa = (int32_t*)malloc(sizeof(int32_t) * (size_t)(cnt > 0 ? cnt : 1));
    int32_t* ia = (int32_t*)malloc(sizeof(int32_t) * (size_t)(cnt > 0 ? cnt : 1));
    for (int64_t i = 0; i < cnt; ++i) { op[us[i] + 1]++; ip[vs[i] + 1]++; }
    for (int v = 0; v < n; ++v) { op[v + 1] += op[v]; ip[v + 1] += ip[v]; }
    {
        int64_t* fo = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
        int64_t* fi = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
        for (int v = 0; v < n; ++v) { fo[v] = op[v]; fi[v] = ip[v]; }
        for (int64_t i = 0; i < cnt; ++i) { oa[fo[us[i]]++] = vs[i]; ia[fi[vs[i]]++] = us[i]; }
        free(fo);
        free(fi);
    }
    char* seen = (char*)calloc((size_t)n + 1, 1);
    int32_t* order = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int32_t* sv = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int64_t* sk = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    int no = 0;
    for (int r = 0; r < n; ++r) {
        if (seen[r]) continue;
        int top = 0;
        seen[r] = 1;
        sv[0] = r;
        sk[0] = op[r];
        while (top >= 0) {
            int x = sv[top];
            if (sk[top] < op[x + 1]) {
                int y = oa[sk[top]++];
                if (!seen[y]) {
                    seen[y] = 1;
                    ++top;
                    sv[top] = y;
                    sk[top] = op[y];
                }
            } else {
                order[no++] = x;
                --top;
            }
        }
    }
    for (int v = 0; v < n; ++v) comp[v] = -1;
    int C = 0;
    for (int i = no - 1; i >= 0; --i) {
        int r = order[i];
        if (comp[r] != -1) continue;
        int sp = 0;
        comp[r] = C;
        sv[sp++] = r;  /* each vertex is pushed once, when it gets its id */
        while (sp > 0) {
            int x = sv[--sp];
            for (int64_t k = ip[x]; k < ip[x + 1]; ++k) {
                int y = ia[k];
                if (comp[y] == -1) {
                    comp[y] = C;
                    sv[sp++] = y;
                }
            }
        }
        ++C;
    }
    free(op); free(ip); free(oa); free(ia); free(seen); free(order); free(sv); free(sk);
    return C;
}

typedef struct { int32_t a, b; int64_t c; } scc_trip;

static int trip_cmp(const void* x, const void* y) {
    const scc_trip* p = (const scc_trip*)x;
    const scc_trip* q = (const scc_trip*)y;
    if (p->a != q->a) return p->a < q->a ? -1 : 1;
    if (p->b != q->b) return p->b < q->b ? -1 : 1;
    return 0;
}

/* scc_contract.cpp:107-125 restated: every line between two components adds
 * its capacity to the entry (comp[u], comp[v]); entries in (cu, cv) order. */
int64_t orc_scc_contract(int64_t cnt, const int32_t* us, const int32_t* vs, const int64_t* caps,
                         const int32_t* comp, int32_t* cu, int32_t* cv, int64_t* ccap) {
    scc_trip* tr = (scc_trip*)malloc(sizeof(scc_trip) * (size_t)(cnt > 0 ? cnt : 1));
    int64_t k = 0;
    for (int64_t i = 0; i < cnt; ++i) {
        if (comp[us[i]] == comp[vs[i]]) continue;
        tr[k].a = comp[us[i]];
        tr[k].b = comp[vs[i]];
        tr[k].c = caps[i];
        ++k;
    }
    qsort(tr, (size_t)k, sizeof(scc_trip), trip_cmp);
    int64_t m = 0;
    for (int64_t i = 0; i < k; ++i) {
        if (m > 0 && cu[m - 1] == tr[i].a && cv[m - 1] == tr[i].b) {
            ccap[m - 1] += tr[i].c;
        } else {
            cu[m] = tr[i].a;
            cv[m] = tr[i].b;
            ccap[m] = tr[i].c;
            ++m;
        }
    }
    free(tr);
    return m;
}
