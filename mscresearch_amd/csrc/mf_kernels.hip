// mf_kernels.hip -- CDNA4 (gfx950) kernels for exact integer push-relabel.
//
// Design notes (the reference kernels are cited as BC:line of
// Region_Aware_PR_with_Pruning_BoundaryCore.cu):
//  * Wave64 everywhere: frontier / active-list compaction by ballot + LDS
//    staging (one global atomic per workgroup), rows longer than kLightDeg
//    are scanned by a whole wave with coalesced col/res loads.
//  * All flow arithmetic is int64 with native global 64-bit atomics.  Only
//    the owner of a vertex (one lane, one wave, or one path-push wave holding
//    its lock) takes excess or residual out of it; everybody else only adds.
//    The owner reads excess and residuals with relaxed agent-scope atomic
//    loads, so what it sees is a lower bound and a push never overdraws.
//  * Exactness does not rest on heights: termination is decided by an exact
//    BFS (global relabel) and every answer can be certified by launch_cert.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdint>

#include "mf_kernels.h"

namespace mf {

namespace {

constexpr int WAVE = 64;

template <typename P>
__device__ __forceinline__ int64_t ld64(P p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Residual read in a step that pushes nothing on the slot (BFS levels,
// local relabel, certificate): the last writer was an earlier kernel, so a
// plain load sees its value and may hit in L2.  The agent-scope ld64 is
// for discharge kernels, where other workgroups of the same launch add to
// residuals.
template <typename P>
__device__ __forceinline__ int64_t ld_res(P p) {
    return *p;
}
template <typename P>
__device__ __forceinline__ int64_t add64(P p, int64_t v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int64_t add64_lds(int64_t* p, int64_t v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <typename P>
__device__ __forceinline__ int32_t add32(P p, int32_t v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename P>
__device__ __forceinline__ uint64_t addu64(P p, uint64_t v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename P>
__device__ __forceinline__ bool cas32(P p, int32_t expect, int32_t want) {
    return __hip_atomic_compare_exchange_strong(p, &expect, want, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
}
template <typename P>
__device__ __forceinline__ void min32(P p, int32_t v) {
    __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename P>
__device__ __forceinline__ void max32(P p, int32_t v) {
    __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (WAVE - 1)); }

__device__ __forceinline__ int wave_min_i32(int x) {
    for (int o = WAVE / 2; o > 0; o >>= 1) x = min(x, __shfl_xor(x, o, WAVE));
    return x;
}
__device__ __forceinline__ int wave_max_i32(int x) {
    for (int o = WAVE / 2; o > 0; o >>= 1) x = max(x, __shfl_xor(x, o, WAVE));
    return x;
}
__device__ __forceinline__ int wave_sum_i32(int x) {
    for (int o = WAVE / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, WAVE);
    return x;
}
__device__ __forceinline__ int64_t wave_sum_i64(int64_t x) {
    for (int o = WAVE / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, WAVE);
    return x;
}
__device__ __forceinline__ int64_t wave_incl_scan_i64(int64_t x) {
    const int lane = lane_id();
    for (int o = 1; o < WAVE; o <<= 1) {
        const int64_t y = __shfl_up(x, o, WAVE);
        if (lane >= o) x += y;
    }
    return x;
}

// Block-wide sum of a per-thread value, result valid in thread 0.  `red`
// must hold blockDim.x / WAVE entries.
template <typename T>
__device__ __forceinline__ T block_sum(T v, T* red) {
    for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    const int w = threadIdx.x / WAVE;
    if (lane_id() == 0) red[w] = v;
    __syncthreads();
    T s = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x / WAVE); ++i) s += red[i];
    __syncthreads();
    return s;
}

// Every pointer of a slot is device global memory.  Kernels see the slot
// through GSlot, whose pointers are typed address_space(1): the compiler then
// emits global (not flat) loads, stores and atomics, without the lgkmcnt wait
// each flat access carries in an inner loop.
#define GAS __attribute__((address_space(1)))

struct GGraph {
    int n, E, ntiles, R;
    const GAS int32_t* row_ptr;
    const GAS int32_t* col;
    const GAS int64_t* cap;
    const GAS int32_t* rev;
    const GAS int32_t* tile_start;
    const GAS int32_t* tile_of;
    GAS int64_t* capsum;
    GAS uint64_t* cert_fp;
    uint64_t cert_seed;
    const GAS int64_t* pcap;
    const GAS int32_t* fstart;
    const GAS int32_t* bu_pick;
    const GAS uint32_t* fwdm;
};

struct GSlot {
    GGraph g;
    GAS int64_t* res;
    GAS int64_t* excess;
    GAS int32_t* h;
    GAS int32_t* fa;
    GAS int32_t* fb;
    GAS int32_t* lock;
    GAS int32_t* aux;
    GAS int32_t* hl;
    GAS int32_t* tile_cnt;
    GAS int32_t* tile_hot;
    GAS int64_t* tile_ex;
    GAS int32_t* tsplit;
    GAS int32_t* lbkt;
    GAS int32_t* giant;
    GAS int64_t* gacc;
    int32_t gcap;
    GAS uint64_t* fbits;
    GAS uint8_t* recv;
    int32_t recv_ok, cert_sside;
    GAS uint8_t* emark;
    GAS int32_t* mlist;
    int32_t mcap;
    const GAS int32_t* mask;
    const GAS uint8_t* live;
    GAS Ctr* ctr;
    int32_t s, t, K, Kh, H_cut, phase, seed_t_only, inj, disc_cap, push_first, marks_ok, bu_light;
    int64_t bfs_edges;
    int32_t tstamp;  // this cycle's touch stamp (negative)
};

template <typename T>
__device__ __forceinline__ GAS T* as_global(T* p) {
    return (GAS T*)p;
}

__device__ __forceinline__ GSlot load_slot_at(const SlotView* slots, int slot);

__device__ __forceinline__ GSlot load_slot(const SlotView* slots, const int* list) {
    return load_slot_at(slots, list[blockIdx.y]);
}

__device__ __forceinline__ GSlot load_slot_at(const SlotView* slots, int slot) {
    const SlotView v = slots[slot];
    GSlot S;
    S.g.n = v.g.n;
    S.g.E = v.g.E;
    S.g.ntiles = v.g.ntiles;
    S.g.R = v.g.R;
    S.g.row_ptr = as_global(v.g.row_ptr);
    S.g.col = as_global(v.g.col);
    S.g.cap = as_global(v.g.cap);
    S.g.rev = as_global(v.g.rev);
    S.g.tile_start = as_global(v.g.tile_start);
    S.g.tile_of = as_global(v.g.tile_of);
    S.g.capsum = as_global(v.g.capsum);
    S.g.cert_fp = as_global(v.g.cert_fp);
    S.g.cert_seed = v.g.cert_seed;
    S.g.pcap = as_global(v.g.pcap);
    S.g.fstart = as_global(v.g.fstart);
    S.g.bu_pick = as_global(v.g.bu_pick);
    S.g.fwdm = as_global(v.g.fwdm);
    S.res = as_global(v.res);
    S.excess = as_global(v.excess);
    S.h = as_global(v.h);
    S.fa = as_global(v.fa);
    S.fb = as_global(v.fb);
    S.lock = as_global(v.lock);
    S.aux = as_global(v.aux);
    S.hl = as_global(v.hl);
    S.tile_cnt = as_global(v.tile_cnt);
    S.tile_hot = as_global(v.tile_hot);
    S.tile_ex = as_global(v.tile_ex);
    S.tsplit = as_global(v.tsplit);
    S.lbkt = as_global(v.lbkt);
    S.giant = as_global(v.giant);
    S.gacc = as_global(v.gacc);
    S.gcap = v.gcap;
    S.fbits = as_global(v.fbits);
    S.recv = as_global(v.recv);
    S.recv_ok = v.recv_ok;
    S.cert_sside = v.cert_sside;
    S.emark = as_global(v.emark);
    S.mlist = as_global(v.mlist);
    S.mcap = v.mcap;
    S.mask = as_global(v.mask);
    S.live = as_global(v.live);
    S.ctr = as_global(v.ctr);
    S.s = v.s;
    S.t = v.t;
    S.K = v.K;
    S.Kh = v.Kh;
    S.H_cut = v.H_cut;
    S.disc_cap = v.disc_cap;
    S.push_first = v.push_first;
    S.marks_ok = v.marks_ok;
    S.bu_light = v.bu_light;
    S.bfs_edges = v.bfs_edges;
    S.phase = v.phase;
    S.seed_t_only = v.seed_t_only;
    S.inj = v.inject;
    S.tstamp = -(2 * S.ctr->epoch + 1);
    return S;
}

// A place in a global list for each active lane of the wave: one atomic on
// the list's count per wave instead of one per lane (the count is one hot
// word).  Every lane that reaches the call takes part.
__device__ __forceinline__ int wave_slot(GAS int32_t* cnt) {
    const uint64_t m = __ballot(true);
    const int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane_id() == leader) base = add32(cnt, __popcll(m));
    base = __shfl(base, leader, WAVE);
    return base + __popcll(m & ((1ULL << lane_id()) - 1));
}

// LDS-staged append buffer: entries are collected with LDS atomics and
// flushed with one global atomic per workgroup.  Past CAP entries go
// straight to global memory, their places taken with one global atomic per
// wave (the count is one hot word per list); the row-length sum of every
// entry stays in LDS until the flush.
template <int CAP>
struct Stage {
    int32_t buf[CAP];
    int64_t deg_sum;
    int cnt;
    int base;
};

template <int CAP>
__device__ __forceinline__ void stage_init(Stage<CAP>& st) {
    if (threadIdx.x == 0) { st.cnt = 0; st.deg_sum = 0; }
}

template <int CAP>
__device__ __forceinline__ void stage_push(Stage<CAP>& st, int32_t item, int deg, GAS int32_t* gbuf,
                                           GAS int32_t* gcnt, GAS int64_t* gdeg) {
    const int pos = atomicAdd(&st.cnt, 1);
    add64_lds(&st.deg_sum, (int64_t)deg);
    if (pos < CAP) {
        st.buf[pos] = item;
    } else {
        gbuf[wave_slot(gcnt)] = item;  // the lanes of this call that overflow
    }
}

template <int CAP>
__device__ __forceinline__ void stage_flush(Stage<CAP>& st, GAS int32_t* gbuf, GAS int32_t* gcnt,
                                            GAS int64_t* gdeg) {
    __syncthreads();
    const int n = min(st.cnt, CAP);
    if (threadIdx.x == 0) {
        st.base = n > 0 ? add32(gcnt, n) : 0;
        if (gdeg && st.deg_sum) add64(gdeg, st.deg_sum);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) gbuf[st.base + i] = st.buf[i];
}

struct Work {
    uint64_t edges = 0, verts = 0, pushes = 0, relabels = 0;
};

// Mark entry e (its pair changed since the residual was last reset) and,
// the first time, append it to the slab's list, so resets and certificates
// visit the marked entries only.  The mark byte is set through an atomic OR
// on its 32-bit word, whose old value says who marked it first.  The
// lanes of a wave that mark first take their list places with one atomic
// add (the count is one hot word per slab); past mcap entries only the map
// is kept (the count goes on past mcap).  Every lane that reaches the call
// takes part in the ballot.
__device__ __forceinline__ void mark_entry(const GSlot& S, int e) {
    bool first = false;
    if (!S.emark[e]) {
        GAS uint32_t* w = (GAS uint32_t*)(S.emark + (e & ~3));  // slab arrays are 256-byte aligned
        const uint32_t bit = 1u << (8 * (e & 3));
        first = !(__hip_atomic_fetch_or(w, bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & bit);
    }
    const uint64_t m = __ballot(first);
    if (m == 0) return;
    const int lane = lane_id();
    const int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == leader) base = add32(&S.mlist[0], __popcll(m));
    base = __shfl(base, leader, WAVE);
    if (first) {
        const int k = base + __popcll(m & ((1ULL << lane) - 1));
        if (k < S.mcap) S.mlist[1 + k] = e;
    }
}

// Move d units of residual across entry e (its pair rev[e] gains them) and
// mark the entry, so the next query on this slab resets only marked pairs.
__device__ __forceinline__ void move_res(const GSlot& S, int e, int64_t d) {
    __hip_atomic_fetch_add(&S.res[e], -d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_add(&S.res[S.g.rev[e]], d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (S.recv) S.recv[S.g.col[e]] = 1;  // flow entered the head's row
    mark_entry(S, e);
}

__device__ __forceinline__ void work_flush(const GSlot& S, Work w, uint64_t* red, int fam) {
    // The four counters in one block reduction (one barrier pair, the
    // workgroup's fixed cost in a grid-stride kernel), one atomic each from
    // thread 0.  Up to 1024 threads.
    (void)red;
    __shared__ uint64_t wred[4][1024 / WAVE];
    uint64_t x[4] = {w.edges, w.verts, w.pushes, w.relabels};
#pragma unroll
    for (int k = 0; k < 4; ++k)
        for (int o = WAVE / 2; o > 0; o >>= 1) x[k] += __shfl_xor(x[k], o, WAVE);
    const int wv = threadIdx.x / WAVE;
    if (lane_id() == 0)
#pragma unroll
        for (int k = 0; k < 4; ++k) wred[k][wv] = x[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t[4] = {0, 0, 0, 0};
        for (int i = 0; i < (int)(blockDim.x / WAVE); ++i)
#pragma unroll
            for (int k = 0; k < 4; ++k) t[k] += wred[k][i];
        GAS uint64_t* c = S.ctr->wk[fam];
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (t[k]) addu64(&c[k], t[k]);
    }
    __syncthreads();
}

}  // namespace

// ------------------------------------------------------------------ resets

__global__ void k_reset(const SlotView* __restrict__ slots, const int* __restrict__ list,
                        unsigned what) {
    const GSlot S = load_slot(slots, list);
    GAS Ctr* c = S.ctr;
    if (threadIdx.x == 0) {
        if (what & RST_BFS) {
            for (int i = 0; i < 4; ++i) { c->fsize[i] = 0; c->fedges[i] = 0; }
            c->level = 0;
            c->nhl = 0;
            c->nhg = 0;
            c->ngiant = 0;
            c->bar_count = 0;
            c->visited = 0;
            c->fb_level = -1;
        }
        if (what & (RST_ACTIVE | RST_SPARSE)) {
            c->prev_nlight = c->nlight;
            c->prev_nheavy = c->nheavy;
        }
        if (what & (RST_ACTIVE | RST_TOUCH)) {
            c->ntouched = 0;
            c->epoch = c->epoch + 1;
        }
        if (what & (RST_ACTIVE | RST_SPARSE)) {
            c->active = 0; c->active_high = 0; c->nlight = 0; c->nheavy = 0; c->ngiant = 0;
            c->pos_excess = 0; c->neg_count = 0; c->neg_sum = 0;
            c->hmin = INT_MAX; c->hmax = -1;
        }
        if (what & RST_CERT) { c->cert_bad = 0; c->cut = 0; c->fp = 0; c->nhl = 0; c->ngiant = 0; }
        if (what & RST_RM) c->fixed_edges = 0;
        if (what & RST_WORK)
            for (int f = 0; f < kFamilies; ++f)
                for (int k = 0; k < 4; ++k) c->wk[f][k] = 0;
        if (what & RST_ERR) c->err = 0;
    }
    if (what & (RST_ACTIVE | RST_SPARSE))
        for (int i = threadIdx.x; i < S.g.ntiles; i += blockDim.x) {
            S.tile_cnt[i] = 0;
            S.tile_hot[i] = 0;
            S.tile_ex[i] = 0;
        }
}

void launch_reset(hipStream_t st, const SlotView* slots, const int* list, int nlist, unsigned what) {
    hipLaunchKernelGGL(k_reset, dim3(1, nlist), dim3(256), 0, st, slots, list, what);
}

// -------------------------------------------------------- residual init (a10)
// BC:151-170: res = cap, excess = 0.  Edge-parallel copy; locks freed and
// touch stamps cleared (the slab's memory may have served another slot,
// whose touch epochs differ from this one's).

__global__ __launch_bounds__(256) void k_init(const SlotView* __restrict__ slots,
                                              const int* __restrict__ list) {
    const GSlot S = load_slot(slots, list);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (int64_t e = i0; e < S.g.E; e += stride) {
        S.res[e] = S.g.cap[e];
        S.emark[e] = 0;
    }
    for (int64_t v = i0; v < S.g.n; v += stride) {
        S.excess[v] = 0;
        S.h[v] = 0;
        S.lock[v] = -1;
        S.aux[v] = 0;  // no touch stamp of an earlier owner of this memory survives
        if (S.recv) S.recv[v] = 0;
    }
    if (i0 == 0) {
        S.mlist[0] = 0;  // the marks are all cleared above
        addu64(&S.ctr->wk[WF_INIT][0], (uint64_t)S.g.E);
        addu64(&S.ctr->wk[WF_INIT][1], (uint64_t)S.g.n);
    }
}

void launch_init(hipStream_t st, const SlotView* slots, const int* list, int nlist, int gridx) {
    hipLaunchKernelGGL(k_init, dim3(gridx, nlist), dim3(256), 0, st, slots, list);
}

// Reset of a slab that last held a residual of the same graph: only entry
// pairs marked by a push (or an update) since go back to their capacities.
// The marked entries come from the slab's list, or, when it overflowed,
// from the mark map read 8 entries per load.  The list count is cleared by
// k_marks_clear after every workgroup has read it.
__global__ __launch_bounds__(256) void k_init_marked(const SlotView* __restrict__ slots,
                                                     const int* __restrict__ list) {
    const GSlot S = load_slot(slots, list);
    __shared__ uint64_t red[256 / WAVE];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t E = S.g.E;
    const int cnt = S.mlist[0];
    const int64_t nw = cnt <= S.mcap ? 0 : E / 8;
    const GAS uint64_t* m8 = (const GAS uint64_t*)S.emark;  // slab arrays are 256-byte aligned
    Work wk;
    if (cnt <= S.mcap)
        for (int64_t k = i0; k < cnt; k += stride) {
            const int e = S.mlist[1 + k];
            const int r = S.g.rev[e];
            S.res[e] = S.g.cap[e];
            S.res[r] = S.g.cap[r];
            S.emark[e] = 0;
            wk.edges++;
        }
    for (int64_t w = i0; w < nw; w += stride) {
        const uint64_t m = m8[w];
        if (m == 0) continue;
        for (int b = 0; b < 8; ++b)
            if ((m >> (8 * b)) & 0xff) {
                const int e = (int)(w * 8 + b);
                const int r = S.g.rev[e];
                S.res[e] = S.g.cap[e];
                S.res[r] = S.g.cap[r];
                S.emark[e] = 0;
                wk.edges++;
            }
    }
    if (cnt > S.mcap)
        for (int64_t e = nw * 8 + i0; e < E; e += stride)  // tail
            if (S.emark[e]) {
                const int r = S.g.rev[(int)e];
                S.res[e] = S.g.cap[e];
                S.res[r] = S.g.cap[r];
                S.emark[e] = 0;
                wk.edges++;
            }
    for (int64_t v = i0; v < S.g.n; v += stride) {
        S.excess[v] = 0;
        S.h[v] = 0;
        S.lock[v] = -1;
        S.aux[v] = 0;
        if (S.recv) S.recv[v] = 0;
    }
    if (i0 == 0) wk.verts += (uint64_t)S.g.n;
    work_flush(S, wk, red, WF_INIT);
}

// After k_init_marked: the marks are all cleared, so the list starts empty.
__global__ void k_marks_clear(const SlotView* __restrict__ slots, const int* __restrict__ list) {
    const GSlot S = load_slot(slots, list);
    if (threadIdx.x == 0) S.mlist[0] = 0;
}

void launch_init_marked(hipStream_t st, const SlotView* slots, const int* list, int nlist, int gridx) {
    hipLaunchKernelGGL(k_init_marked, dim3(gridx, nlist), dim3(256), 0, st, slots, list);
    hipLaunchKernelGGL(k_marks_clear, dim3(1, nlist), dim3(64), 0, st, slots, list);
}

// ---------------------------------------------------------------- preflow (a11)
// BC:173-200: saturate every entry of s with positive residual.  One
// workgroup strides the row of s instead of one thread.

__global__ __launch_bounds__(256) void k_preflow(const SlotView* __restrict__ slots,
                                                 const int* __restrict__ list) {
    const GSlot S = load_slot(slots, list);
    __shared__ int64_t red[256 / WAVE];
    const int s = S.s;
    // workgroup 0: the counters; every workgroup: its share of s's row
    int64_t in_t = 0;
    if (blockIdx.x == 0) {
        // In-capacity of t: over t's row, pcap - cap is the capacity of each
        // entry's pair into t.
        const int tb = S.g.row_ptr[S.t], te = S.g.row_ptr[S.t + 1];
        for (int e = tb + threadIdx.x; e < te; e += blockDim.x) in_t += S.g.pcap[e] - S.g.cap[e];
        in_t = block_sum<int64_t>(in_t, red);
        if (threadIdx.x == 0) {
            S.ctr->incap_t = in_t;
            S.ctr->outcap_s = S.g.capsum[s];
            if (S.recv) S.recv[s] = 1;  // s pushes without receiving
        }
    }
    if (S.inj) {
        if (blockIdx.x != 0) return;
        // Bounded injection: no flow can exceed the out-capacity of s or the
        // in-capacity of t (each an s-t cut), so s starts with B units of
        // excess and an unsaturated row.
        if (threadIdx.x == 0) {
            const int64_t out_s = S.g.capsum[s];
            const int64_t B = out_s < in_t ? out_s : in_t;
            S.excess[s] = B;
            S.ctr->inject = B;
        }
        return;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) S.ctr->inject = 0;
    const int rb = S.g.row_ptr[s], re = S.g.row_ptr[s + 1];
    int64_t sum = 0;
    for (int e = rb + (int)(blockIdx.x * blockDim.x + threadIdx.x); e < re; e += (int)(gridDim.x * blockDim.x)) {
        const int64_t d = ld64(&S.res[e]);
        if (d > 0) {
            move_res(S, e, d);
            add64(&S.excess[S.g.col[e]], d);
            sum += d;
        }
    }
    const int64_t tot = block_sum<int64_t>(sum, red);
    if (threadIdx.x == 0 && tot) add64(&S.excess[s], -tot);
}

void launch_preflow(hipStream_t st, const SlotView* slots, const int* list, int nlist) {
    // s is often a hub: its row over 16 workgroups per slot
    hipLaunchKernelGGL(k_preflow, dim3(16, nlist), dim3(256), 0, st, slots, list);
}

// ------------------------------------------------------------- BFS (a12/a13)
// Level-L frontier lives in fa when L is even, fb when odd; its size in
// ctr->fsize[L % 4].  Level L writes level L+1 and clears ring slot L+2.
//
//   BFS_GR : backwards over residual edges (w joins via entry v->w when
//            res[rev] > 0), visited marker h == n, s never joins, optional
//            vertex mask (BC:203-233, BC:322-348).
//   BFS_FWD: forwards over entries with cap > 0 (pruning, BC:1402-1416);
//            visited marker h == n.
//   BFS_UND: over every entry (undirected component, BC:1419-1433);
//            visited marker aux == n.

template <int MODE>
__device__ __forceinline__ GAS int32_t* bfs_lvl(const GSlot& S) {
    return MODE == BFS_UND ? S.aux : S.h;
}

// Bottom-up: may unvisited v join level L+1 through entry e: v -> x where x
// is on level L?  GR: v can push to x (res[e] > 0).  FWD: x -> v carries
// capacity, i.e. e is a reverse entry (cap[e] == 0; forward entries always
// have cap > 0 and reverse entries cap 0).  UND: any entry.
template <int MODE>
__device__ __forceinline__ bool bfs_bu_edge_ok(const GSlot& S, int e) {
    if (MODE == BFS_GR) return ld_res(&S.res[e]) > 0;
    if (MODE == BFS_FWD) return S.g.cap[e] == 0;
    return true;
}

// Frontier bitmaps of the bottom-up levels (SlotView::fbits): words per
// bitmap, level L's bitmap, and "is x on level L" from it (a 4-byte load
// from n / 8 bytes) or, without bitmaps, from x's level.
__device__ __forceinline__ int fb_words(int n) { return (n + 63) >> 6; }
__device__ __forceinline__ GAS uint64_t* fb_level(const GSlot& S, int L) {
    return S.fbits + (size_t)(L & 1) * (size_t)fb_words(S.g.n);
}
// Level L's bitmap as 32-bit words when it is complete (Ctr::fb_level),
// else nullptr: candidate parents are then tested by their levels.
__device__ __forceinline__ const GAS uint32_t* fb_valid(const GSlot& S, int L) {
    return (S.fbits && S.ctr->fb_level == L) ? (const GAS uint32_t*)fb_level(S, L) : nullptr;
}
__device__ __forceinline__ bool on_level(const GAS uint32_t* bits, const GAS int32_t* lvl, int x, int L) {
    if (bits) return (bits[x >> 5] >> (x & 31)) & 1u;
    return lvl[x] == L;
}
// Vertex x joined level L + 1 outside the wave that owns its bitmap word.
__device__ __forceinline__ void fb_set_next(const GSlot& S, int x, int L) {
    if (S.fbits)
        __hip_atomic_fetch_or(fb_level(S, L + 1) + (x >> 6), 1ull << (x & 63), __ATOMIC_RELAXED,
                              __HIP_MEMORY_SCOPE_AGENT);
}

template <int MODE>
__device__ __forceinline__ bool bfs_vertex_ok(const GSlot& S, int v) {
    if (MODE == BFS_GR) {
        if (v == S.s && !S.inj) return false;
        if (S.mask && S.mask[v] == 0) return false;
    }
    return true;
}

// May frontier vertex v (entry e: v -> w) discover w?  GR: w != s, w in
// the mask, and w can push to v (res[rev e] > 0); FWD: e carries capacity;
// UND: always.  Split in two: aux = edge_aux(e), loaded with the column,
// then edge_ok_aux(aux, w).  GR with ROW: res[rev e] = pcap[e] - res[e] is
// read along e's row (16 coalesced bytes), so w's level is gathered only
// for entries that pass: the wide frontiers of the grid-wide top-down
// levels.  GR without ROW: aux = rev[e] and res[rev e] is gathered only for
// unvisited w: the single-workgroup levels, whose narrow frontiers meet
// mostly visited neighbours and wait on each load's latency.
template <int MODE, bool ROW>
__device__ __forceinline__ int edge_aux(const GSlot& S, int e) {
    if (MODE == BFS_GR) return ROW ? (S.g.pcap[e] - ld_res(&S.res[e]) > 0 ? 1 : 0) : S.g.rev[e];
    if (MODE == BFS_FWD) return S.g.cap[e] > 0 ? 1 : 0;
    return 1;
}

// GR without ROW, a fresh row (td_entries): the pair's residual is known
// from the forward mask, carried in aux as kFreshYes / kFreshNo instead of
// the pair's entry index.
constexpr int kFreshNo = -1, kFreshYes = -2;

template <int MODE, bool ROW>
__device__ __forceinline__ bool edge_ok_aux(const GSlot& S, int aux, int w) {
    if (MODE == BFS_GR) {
        if (w == S.s && !S.inj) return false;
        if (S.mask && S.mask[w] == 0) return false;
        if (!ROW) return aux == kFreshYes || (aux >= 0 && ld_res(&S.res[aux]) > 0);
    }
    return aux != 0;
}

// Whether frontier vertex v's row still holds its capacities (GR only).
template <int MODE>
__device__ __forceinline__ bool td_fresh(const GSlot& S, int v) {
    return MODE == BFS_GR && S.recv_ok && S.g.fwdm && S.recv[v] == 0;
}

// Top-down expansion of entries e0, e0 + step, ... below e1 by one thread,
// U entries at a time: their column and entry-side loads, then the level
// gathers of the entries that pass, each issued together (one memory round
// trip per stage instead of one per entry and stage).  claim(w) runs for
// each vertex this thread moved from n to level L + 1.  fresh (GR, ROW): no
// flow has entered or left the row's vertex (SlotView::recv), so its
// residuals are its capacities and w can push into it exactly when the
// entry is a reverse one (its pair, in w's row, is forward with cap > 0):
// the forward mask answers instead of the residuals.
template <int MODE, int U, bool ROW = true, typename Claim>
__device__ __forceinline__ void td_entries(const GSlot& S, int e0, int e1, int step, int L,
                                           GAS int32_t* lvl, Claim claim, bool fresh = false) {
    const int n = S.g.n;
    for (int e = e0; e < e1; e += U * step) {
        int w[U], ax[U], l[U];
        bool ok[U];
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const int ek = e + k * step;
            w[k] = -1;
            ax[k] = 0;
            if (ek < e1) {
                w[k] = S.g.col[ek];
                if (MODE == BFS_GR && fresh) {
                    const bool into = !((S.g.fwdm[ek >> 5] >> (ek & 31)) & 1u);  // a reverse entry
                    ax[k] = ROW ? (into ? 1 : 0) : (into ? kFreshYes : kFreshNo);
                } else {
                    ax[k] = edge_aux<MODE, ROW>(S, ek);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < U; ++k) {
            // without ROW a GR aux is an entry index (or a fresh answer), not a flag
            const bool pre = (MODE == BFS_GR && !ROW) ? ax[k] != kFreshNo : ax[k] != 0;
            l[k] = (w[k] >= 0 && pre) ? lvl[w[k]] : -1;
        }
#pragma unroll
        for (int k = 0; k < U; ++k) ok[k] = l[k] == n && edge_ok_aux<MODE, ROW>(S, ax[k], w[k]);
#pragma unroll
        for (int k = 0; k < U; ++k)
            if (ok[k] && cas32(&lvl[w[k]], n, L + 1)) claim(w[k]);
    }
}

// Top-down expansion for deep, narrow BFS levels (grids), where every level
// is a chain of dependent loads and the frontier's neighbours are mostly
// visited already: the column and pair entry of U entries, then, all issued
// together, each neighbour's level, its pair residual (GR) and its row
// bounds (the degree its frontier entry carries), so a level waits on two
// gathers instead of four; claim(w, deg) for each vertex moved from n to
// level L + 1.
template <int MODE, int U, typename Claim>
__device__ __forceinline__ void td_entries_deep(const GSlot& S, int e0, int e1, int step, int L,
                                                GAS int32_t* lvl, Claim claim, bool fresh = false) {
    const int n = S.g.n;
    for (int e = e0; e < e1; e += U * step) {
        int w[U], ax[U], l[U], d[U];
        bool okx[U];
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const int ek = e + k * step;
            w[k] = -1;
            ax[k] = 0;
            if (ek < e1) {
                w[k] = S.g.col[ek];
                if (MODE == BFS_GR && fresh)  // a reverse entry of a fresh row: its pair holds its capacity
                    ax[k] = ((S.g.fwdm[ek >> 5] >> (ek & 31)) & 1u) ? kFreshNo : kFreshYes;
                else
                    ax[k] = MODE == BFS_GR ? S.g.rev[ek] : edge_aux<MODE, true>(S, ek);
            }
        }
#pragma unroll
        for (int k = 0; k < U; ++k) {
            l[k] = -1;
            okx[k] = false;
            d[k] = 0;
            if (w[k] >= 0 && !(MODE == BFS_GR && ax[k] == kFreshNo)) {
                l[k] = lvl[w[k]];
                if (MODE == BFS_GR) okx[k] = ax[k] == kFreshYes || (ax[k] >= 0 && ld_res(&S.res[ax[k]]) > 0);
                else okx[k] = ax[k] != 0;
                d[k] = S.g.row_ptr[w[k] + 1] - S.g.row_ptr[w[k]];
            }
        }
#pragma unroll
        for (int k = 0; k < U; ++k) {
            bool ok = l[k] == n && okx[k];
            if (ok && MODE == BFS_GR) ok = (w[k] != S.s || S.inj) && (S.mask == nullptr || S.mask[w[k]] != 0);
            if (ok && cas32(&lvl[w[k]], n, L + 1)) claim(w[k], d[k]);
        }
    }
}

// Bottom-up test of entries [rb, re) by one thread, U entries at a time
// (their residual/capacity and column loads together, then their level
// gathers): does an entry lead to level L?  `scanned` = entries up to the end
// of the group holding the first such entry.
template <int MODE, int U>
__device__ __forceinline__ bool bu_range_lane(const GSlot& S, int rb, int re, int L, const GAS int32_t* lvl,
                                              const GAS uint32_t* bits, int& scanned) {
    for (int e = rb; e < re; e += U) {
        bool ok[U];
        int c[U];
#pragma unroll
        for (int k = 0; k < U; ++k) {
            const int ek = e + k;
            ok[k] = false;
            c[k] = 0;
            if (ek < re) {
                ok[k] = bfs_bu_edge_ok<MODE>(S, ek);
                c[k] = S.g.col[ek];
            }
        }
        bool on[U];
#pragma unroll
        for (int k = 0; k < U; ++k) on[k] = ok[k] && on_level(bits, lvl, c[k], L);
        bool hit = false;
#pragma unroll
        for (int k = 0; k < U; ++k) hit |= on[k];
        if (hit) {
            scanned = min(e + U, re) - rb;
            return true;
        }
    }
    scanned = re - rb;
    return false;
}

// First entry of v's row a bottom-up level looks at: the relabel (GR) reaches
// a parent over a residual out-entry, so it starts at the first forward
// entry (DGraph::fstart) and wraps; the pruning BFS modes start at the row
// start.
template <int MODE>
__device__ __forceinline__ int bu_first(const GSlot& S, int v, int rb) {
    return (MODE == BFS_GR && S.g.fstart) ? S.g.fstart[v] : rb;
}

// Bottom-up test of v's row [rb, re) from fs = bu_first(v), wrapping to
// [rb, fs).
template <int MODE, int U>
__device__ __forceinline__ bool bu_row_lane(const GSlot& S, int rb, int fs, int re, int L, const GAS int32_t* lvl,
                                            const GAS uint32_t* bits, int& scanned) {
    int a = 0;
    if (bu_range_lane<MODE, U>(S, fs, re, L, lvl, bits, a)) {
        scanned = a;
        return true;
    }
    int b = 0;
    const bool hit = fs > rb && bu_range_lane<MODE, U>(S, rb, fs, L, lvl, bits, b);
    scanned = a + b;
    return hit;
}

// Bottom-up test of a row whose residuals still equal its capacities
// (SlotView::recv): a parent is a forward entry whose head is on level L, so
// only the forward entries of [fs, re) are read (the graph's forward mask,
// 32 entries per load; [rb, fs) holds reverse entries only), their columns
// and level tests U at a time.  `scanned` = entries covered up to the group
// holding the parent.
template <int U>
__device__ __forceinline__ bool bu_row_fresh(const GSlot& S, int fs, int re, int L, const GAS int32_t* lvl,
                                             const GAS uint32_t* bits, int& scanned) {
    for (int e = fs; e < re;) {
        const int lo = e & 31;
        const int span = min(32 - lo, re - e);
        uint32_t m = S.g.fwdm[e >> 5] >> lo;
        if (span < 32) m &= (1u << span) - 1u;
        while (m) {
            int ek[U], c[U];
#pragma unroll
            for (int k = 0; k < U; ++k) {
                ek[k] = -1;
                if (m) {
                    ek[k] = e + __builtin_ctz(m);
                    m &= m - 1u;
                }
            }
#pragma unroll
            for (int k = 0; k < U; ++k) c[k] = ek[k] >= 0 ? S.g.col[ek[k]] : -1;
            bool hit = false;
            int last = e;
#pragma unroll
            for (int k = 0; k < U; ++k) {
                hit |= c[k] >= 0 && on_level(bits, lvl, c[k], L);
                last = max(last, ek[k]);
            }
            if (hit) {
                scanned = last + 1 - fs;
                return true;
            }
        }
        e += span;
    }
    scanned = re - fs;
    return false;
}

// GR seeding (a12): h = n everywhere, level 0 = t plus deficits (phase 0) or
// deficits only (phase 1, draining from t), under the mask; s stays n.
__global__ __launch_bounds__(256) void k_gr_seed(const SlotView* __restrict__ slots,
                                                 const int* __restrict__ list) {
    const GSlot S = load_slot(slots, list);
    __shared__ Stage<1024> st;
    __shared__ uint64_t red[256 / WAVE];
    stage_init(st);
    __syncthreads();
    const int n = S.g.n;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    Work wk;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int v = (int)i;
        bool seed = false;
        if (v != S.s && (S.mask == nullptr || S.mask[v] != 0 || v == S.t)) {
            if (S.phase == 0) seed = (v == S.t) || (!S.seed_t_only && S.excess[v] < 0);
            else seed = (v != S.t) && (S.excess[v] < 0);
        }
        S.h[v] = seed ? 0 : n;
        wk.verts++;
        if (seed) {
            const int deg = S.g.row_ptr[v + 1] - S.g.row_ptr[v];
            stage_push(st, v, deg, S.fa, &S.ctr->fsize[0], &S.ctr->fedges[0]);
            wk.relabels++;
        }
    }
    stage_flush(st, S.fa, &S.ctr->fsize[0], &S.ctr->fedges[0]);
    const uint64_t nv = block_sum<uint64_t>(wk.relabels, red);
    if (threadIdx.x == 0 && nv) add32(&S.ctr->visited, (int32_t)nv);
    work_flush(S, wk, red, WF_BFS);
}

void launch_gr_seed(hipStream_t st, const SlotView* slots, const int* list, int nlist, int gridx) {
    hipLaunchKernelGGL(k_gr_seed, dim3(gridx, nlist), dim3(256), 0, st, slots, list);
}

// Single-source seeding for the pruning BFS: lvl = n, lvl[src] = 0.
// FWD starts at s, UND at t.
__global__ __launch_bounds__(256) void k_bfs_seed1(const SlotView* __restrict__ slots,
                                                   const int* __restrict__ list, int mode) {
    const GSlot S = load_slot(slots, list);
    GAS int32_t* lvl = (mode == BFS_UND) ? S.aux : S.h;
    const int src = (mode == BFS_UND) ? S.t : S.s;
    const int n = S.g.n;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        lvl[i] = (i == src) ? 0 : n;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        S.fa[0] = src;
        S.ctr->fsize[0] = 1;
        S.ctr->fedges[0] = S.g.row_ptr[src + 1] - S.g.row_ptr[src];
        S.ctr->visited = 1;
    }
}

void launch_bfs_seed1(hipStream_t st, const SlotView* slots, const int* list, int nlist, int gridx,
                      int mode) {
    hipLaunchKernelGGL(k_bfs_seed1, dim3(gridx, nlist), dim3(256), 0, st, slots, list, mode);
}

// Top-down expansion of one frontier vertex by one lane.
template <int MODE, int CAP>
__device__ __forceinline__ void td_lane(const GSlot& S, int v, int L, GAS int32_t* lvl,
                                        Stage<CAP>& st, GAS int32_t* out, GAS int32_t* ocnt,
                                        GAS int64_t* odeg, Work& wk) {
    const int rb = S.g.row_ptr[v], re = S.g.row_ptr[v + 1];
    wk.edges += (uint64_t)(re - rb);
    td_entries<MODE, 4>(S, rb, re, 1, L, lvl, [&](int w) {
        stage_push(st, w, S.g.row_ptr[w + 1] - S.g.row_ptr[w], out, ocnt, odeg);
        wk.relabels++;
    }, td_fresh<MODE>(S, v));
}

// Top-down expansion of one (heavy) frontier vertex by the whole workgroup.
template <int MODE, int CAP>
__device__ __forceinline__ void td_block(const GSlot& S, int v, int L, GAS int32_t* lvl,
                                         Stage<CAP>& st, GAS int32_t* out, GAS int32_t* ocnt,
                                         GAS int64_t* odeg, Work& wk) {
    const int rb = S.g.row_ptr[v], re = S.g.row_ptr[v + 1];
    if (threadIdx.x == 0) wk.edges += (uint64_t)(re - rb);
    td_entries<MODE, 2>(S, rb + (int)threadIdx.x, re, (int)blockDim.x, L, lvl, [&](int w) {
        stage_push(st, w, S.g.row_ptr[w + 1] - S.g.row_ptr[w], out, ocnt, odeg);
        wk.relabels++;
    }, td_fresh<MODE>(S, v));
}

template <int MODE>
__global__ __launch_bounds__(256) void k_bfs_td(const SlotView* __restrict__ slots,
                                                const int* __restrict__ list) {
    const GSlot S = load_slot(slots, list);
    __shared__ Stage<2048> st;
    __shared__ uint64_t red[256 / WAVE];
    const int L = S.ctr->level;
    const int fsz = S.ctr->fsize[L & 3];
    if ((int)(blockIdx.x * blockDim.x) >= fsz) return;  // uniform per block
    GAS int32_t* lvl = bfs_lvl<MODE>(S);
    const GAS int32_t* in = (L & 1) ? S.fb : S.fa;
    GAS int32_t* out = (L & 1) ? S.fa : S.fb;
    GAS int32_t* ocnt = &S.ctr->fsize[(L + 1) & 3];
    GAS int64_t* odeg = &S.ctr->fedges[(L + 1) & 3];
    stage_init(st);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        S.ctr->fsize[(L + 2) & 3] = 0;
        S.ctr->fedges[(L + 2) & 3] = 0;
    }
    __syncthreads();
    Work wk;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < fsz; i += gridDim.x * blockDim.x) {
        const int v = in[i];
        wk.verts++;
        const int deg = S.g.row_ptr[v + 1] - S.g.row_ptr[v];
        // longer rows go over the grid: medium rows to hl (k_bfs_td_heavy,
        // a wave each), giant rows to the giant list (k_bfs_td_giant, in
        // kGiantChunk-entry tasks)
        if (deg <= kLightDeg) td_lane<MODE>(S, v, L, lvl, st, out, ocnt, odeg, wk);
        else if (deg <= kGiantDeg) S.hl[wave_slot(&S.ctr->nhl)] = v;
        else S.giant[wave_slot(&S.ctr->ngiant)] = v;
    }
    stage_flush(st, out, ocnt, odeg);
    const uint64_t nv = block_sum<uint64_t>(wk.relabels, red);
    if (threadIdx.x == 0 && nv) add32(&S.ctr->visited, (int32_t)nv);
    work_flush(S, wk, red, WF_BFS);
}

// Top-down expansion of one (medium) frontier vertex by one wave.
template <int MODE, int CAP>
__device__ __forceinline__ void td_wave(const GSlot& S, int v, int L, GAS int32_t* lvl,
                                        Stage<CAP>& st, GAS int32_t* out, GAS int32_t* ocnt,
                                        GAS int64_t* odeg, Work& wk) {
    const int rb = S.g.row_ptr[v], re = S.g.row_ptr[v + 1];
    if (lane_id() == 0) wk.edges += (uint64_t)(re - rb);
    td_entries<MODE, 2>(S, rb + lane_id(), re, WAVE, L, lvl, [&](int w) {
        stage_push(st, w, S.g.row_ptr[w + 1] - S.g.row_ptr[w], out, ocnt, odeg);
        wk.relabels++;
    }, td_fresh<MODE>(S, v));
}

// Medium rows of a top-down level, a wave each, taken grid-stride from the
// level's list.  Hubs sit together at region starts, so handing them out
// across the grid is what balances the level.
template <int MODE>
__global__ __launch_bounds__(256) void k_bfs_td_heavy(const SlotView* __restrict__ slots,
                                                      const int* __restrict__ list) {
    const GSlot S = load_slot(slots, list);
    __shared__ Stage<2048> st;
    __shared__ uint64_t red[256 / WAVE];
    const int nm = S.ctr->nhl;
    const int wpb = (int)(blockDim.x / WAVE);
    if ((int)blockIdx.x * wpb >= nm) return;  // uniform per block
    const int L = S.ctr->level;
    GAS int32_t* lvl = bfs_lvl<MODE>(S);
    GAS int32_t* out = (L & 1) ? S.fa : S.fb;
    GAS int32_t* ocnt = &S.ctr->fsize[(L + 1) & 3];
    GAS int64_t* odeg = &S.ctr->fedges[(L + 1) & 3];
    stage_init(st);
    __syncthreads();
    Work wk;
    const int gw = (int)blockIdx.x * wpb + (int)(threadIdx.x / WAVE);
    for (int k = gw; k < nm; k += (int)gridDim.x * wpb) td_wave<MODE>(S, S.hl[k], L, lvl, st, out, ocnt, odeg, wk);
    stage_flush(st, out, ocnt, odeg);
    const uint64_t nv = block_sum<uint64_t>(wk.relabels, red);
    if (threadIdx.x == 0 && nv) add32(&S.ctr->visited, (int32_t)nv);
    work_flush(S, wk, red, WF_BFS);
}

// One thread per vertex, no workgroup barrier inside the loop (a wave that
// meets a long scan holds back only itself): light rows are scanned by
// their lane; medium rows go to hl (k_bfs_bu_med, a wave each), giant rows
// to the giant list (k_bfs_bu_heavy, a workgroup each).
template <int MODE>
__global__ __launch_bounds__(256) void k_bfs_bu(const SlotView* __restrict__ slots,
                                                const int* __restrict__ list) {
    const GSlot S = load_slot(slots, list);
    __shared__ Stage<2048> st;
    __shared__ uint64_t red[256 / WAVE];
    const int L = S.ctr->level;
    GAS int32_t* lvl = bfs_lvl<MODE>(S);
    GAS int32_t* out = (L & 1) ? S.fa : S.fb;
    GAS int32_t* ocnt = &S.ctr->fsize[(L + 1) & 3];
    GAS int64_t* odeg = &S.ctr->fedges[(L + 1) & 3];
    stage_init(st);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        S.ctr->fsize[(L + 2) & 3] = 0;
        S.ctr->fedges[(L + 2) & 3] = 0;
    }
    __syncthreads();
    const int n = S.g.n;
    // with bitmaps, level L + 1's bitmap is written here word by word: a
    // wave's 64 vertices are one word (every word, so no clearing), and the
    // medium and giant rows listed for the kernels after this one are OR-ed
    // in by them
    GAS uint64_t* nbits = S.fbits ? fb_level(S, L + 1) : nullptr;
    const GAS uint32_t* bits = fb_valid(S, L);
    Work wk;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(WAVE - 1)); base < n; base += stride) {
        const int v = (int)base + lane_id();
        bool joined = false;
        if (v < n && lvl[v] == n && bfs_vertex_ok<MODE>(S, v) && (MODE != BFS_GR || S.live == nullptr || S.live[v])) {
            wk.verts++;
            const int rb = S.g.row_ptr[v], re = S.g.row_ptr[v + 1];
            // no flow entered or left v's row: its residuals are its capacities
            const bool fresh = MODE == BFS_GR && S.recv_ok && S.g.fwdm && S.recv[v] == 0;
            bool picked = false;
            if (MODE == BFS_GR && S.g.bu_pick) {
                // the picked entry first (its head's level is likely cached);
                // a forward entry no push has marked still holds its
                // capacity, so its mark byte (128 entries a line) answers
                // for its residual (16 a line), and a fresh row needs neither
                const int pe = S.g.bu_pick[2 * v], pw = S.g.bu_pick[2 * v + 1];
                picked = pe >= 0 && on_level(bits, lvl, pw, L) &&
                         (fresh || (S.marks_ok && S.emark[pe] == 0) || ld_res(&S.res[pe]) > 0);
                if (picked) {
                    wk.edges += 1;
                    joined = true;
                }
            }
            if (picked) {
            } else if (re - rb > kGiantDeg) {
                S.giant[wave_slot(&S.ctr->ngiant)] = v;
            } else if (re - rb > S.bu_light) {
                S.hl[wave_slot(&S.ctr->nhl)] = v;
            } else {
                // the parent test only for entries that pass the residual
                // test (most reverse entries carry no flow)
                int scanned = 0;
                if (fresh)
                    joined = bu_row_fresh<4>(S, bu_first<MODE>(S, v, rb), re, L, lvl, bits, scanned);
                else
                    joined = bu_row_lane<MODE, 2>(S, rb, bu_first<MODE>(S, v, rb), re, L, lvl, bits, scanned);
                wk.edges += (uint64_t)scanned;
            }
            if (joined) {
                lvl[v] = L + 1;  // only this lane writes v
                stage_push(st, v, re - rb, out, ocnt, odeg);
                wk.relabels++;
            }
        }
        if (nbits) {
            const uint64_t m = __ballot(joined);
            if (lane_id() == 0) nbits[base >> 6] = m;
        }
    }
    stage_flush(st, out, ocnt, odeg);
    const uint64_t nv = block_sum<uint64_t>(wk.relabels, red);
    if (threadIdx.x == 0 && nv) add32(&S.ctr->visited, (int32_t)nv);
    work_flush(S, wk, red, WF_BFS_BU);
}

// Medium unvisited rows of a bottom-up level (SlotView::bu_light < length <=
// kGiantDeg), listed in hl by k_bfs_bu: a wave each, grid-stride, 64
// entries per step (ballot), stopping at the first step holding a parent.
template <int MODE>
__global__ __launch_bounds__(256) void k_bfs_bu_med(const SlotView* __restrict__ slots,
                                                    const int* __restrict__ list) {
    const GSlot S = load_slot(slots, list);
    __shared__ Stage<256> st;
    __shared__ uint64_t red[256 / WAVE];
    const int nm = S.ctr->nhl;
    const int wpb = (int)(blockDim.x / WAVE);
    if ((int)blockIdx.x * wpb >= nm) return;  // uniform per block
    const int L = S.ctr->level;
    GAS int32_t* lvl = bfs_lvl<MODE>(S);
    const GAS uint32_t* bits = fb_valid(S, L);
    GAS int32_t* out = (L & 1) ? S.fa : S.fb;
    GAS int32_t* ocnt = &S.ctr->fsize[(L + 1) & 3];
    GAS int64_t* odeg = &S.ctr->fedges[(L + 1) & 3];
    stage_init(st);
    __syncthreads();
    Work wk;
    const int lane = lane_id();
    for (int k = (int)blockIdx.x * wpb + (int)(threadIdx.x / WAVE); k < nm; k += (int)gridDim.x * wpb) {
        const int u = S.hl[k];
        const int rb = S.g.row_ptr[u], re = S.g.row_ptr[u + 1];
        const int fs = bu_first<MODE>(S, u, rb);
        const int d = re - rb;
        bool hit = false;
        int scanned = 0;
        // positions fs, fs + 1, ... wrapping from re to rb
        for (int k0 = 0; k0 < d; k0 += WAVE) {
            int e = fs + k0 + lane;
            if (e >= re) e -= d;
            bool ok = false;
            if (k0 + lane < d) {
                const bool r = bfs_bu_edge_ok<MODE>(S, e);
                const int c = S.g.col[e];  // issued with the residual load
                ok = r && on_level(bits, lvl, c, L);
            }
            scanned += min(WAVE, d - k0);
            if (__ballot(ok)) { hit = true; break; }
        }
        if (lane == 0) {
            wk.edges += (uint64_t)scanned;
            if (hit) {
                lvl[u] = L + 1;  // only this wave writes u
                fb_set_next(S, u, L);
                stage_push(st, u, re - rb, out, ocnt, odeg);
                wk.relabels++;
            }
        }
    }
    stage_flush(st, out, ocnt, odeg);
    const uint64_t nv = block_sum<uint64_t>(wk.relabels, red);
    if (threadIdx.x == 0 && nv) add32(&S.ctr->visited, (int32_t)nv);
    work_flush(S, wk, red, WF_BFS_BU);
}

// Giant unvisited rows of a bottom-up level: a workgroup scans the row in
// chunks and stops at the first chunk holding a parent.
template <int MODE>
__global__ __launch_bounds__(256) void k_bfs_bu_heavy(const SlotView* __restrict__ slots,
                                                      const int* __restrict__ list) {
    const GSlot S = load_slot(slots, list);
    __shared__ Stage<256> st;
    __shared__ uint64_t red[256 / WAVE];
    __shared__ int found;
    const int nh = S.ctr->ngiant;
    if ((int)blockIdx.x >= nh) return;  // uniform per block
    const int L = S.ctr->level;
    GAS int32_t* lvl = bfs_lvl<MODE>(S);
    const GAS uint32_t* bits = fb_valid(S, L);
    GAS int32_t* out = (L & 1) ? S.fa : S.fb;
    GAS int32_t* ocnt = &S.ctr->fsize[(L + 1) & 3];
    GAS int64_t* odeg = &S.ctr->fedges[(L + 1) & 3];
    stage_init(st);
    Work wk;
    for (int k = blockIdx.x; k < nh; k += gridDim.x) {
        const int v = S.giant[k];
        const int rb = S.g.row_ptr[v], re = S.g.row_ptr[v + 1];
        const int fs = bu_first<MODE>(S, v, rb);
        const int d = re - rb;
        if (threadIdx.x == 0) found = 0;
        __syncthreads();
        int k0 = 0;
        // positions fs, fs + 1, ... wrapping from re to rb
        for (; k0 < d; k0 += blockDim.x) {
            const int kk = k0 + (int)threadIdx.x;
            if (kk < d) {
                const int e = fs + kk < re ? fs + kk : fs + kk - d;
                const bool r = bfs_bu_edge_ok<MODE>(S, e);
                const int c = S.g.col[e];  // issued with the residual load
                if (r && on_level(bits, lvl, c, L)) found = 1;
            }
            __syncthreads();
            const int f = found;
            __syncthreads();
            if (f) break;
        }
        if (threadIdx.x == 0) {
            wk.edges += (uint64_t)min(k0 + (int)blockDim.x, d);
            if (found) {
                lvl[v] = L + 1;
                fb_set_next(S, v, L);
                stage_push(st, v, re - rb, out, ocnt, odeg);
                wk.relabels++;
            }
        }
        __syncthreads();
    }
    stage_flush(st, out, ocnt, odeg);
    const uint64_t nv = block_sum<uint64_t>(wk.relabels, red);
    if (threadIdx.x == 0 && nv) add32(&S.ctr->visited, (int32_t)nv);
    work_flush(S, wk, red, WF_BFS_BU);
}

// Level bookkeeping after a multi-workgroup level: level += 1; after a
// bottom-up level (bu = 1) the new level's bitmap is complete.
__global__ void k_bfs_advance(const SlotView* __restrict__ slots, const int* __restrict__ list, int bu) {
    const GSlot S = load_slot(slots, list);
    if (threadIdx.x == 0) {
        if (bu && S.fbits) S.ctr->fb_level = S.ctr->level + 1;
        S.ctr->level = S.ctr->level + 1;
        S.ctr->nhl = 0;
        S.ctr->nhg = 0;
        S.ctr->ngiant = 0;
    }
}

// Workgroups per slot for a grid-stride kernel that would like `want`: at
// most about 16K over the nlist slots of the launch (fewer, fuller
// workgroups; every one of them is dispatched even when it has no work).
static int per_slot_grid(int want, int nlist) {
    const int cap = 16384 / (nlist > 0 ? nlist : 1);
    return want < cap ? want : (cap > 16 ? cap : 16);
}

// Level L's bitmap from the levels (a wave per 64-vertex word), before the
// first bottom-up level of a BFS; later bottom-up levels write their
// successor's bitmap themselves.
template <int MODE>
__global__ __launch_bounds__(256) void k_bu_bits(const SlotView* __restrict__ slots,
                                                 const int* __restrict__ list) {
    const GSlot S = load_slot(slots, list);
    if (!S.fbits) return;
    const int L = S.ctr->level;
    const GAS int32_t* lvl = bfs_lvl<MODE>(S);
    GAS uint64_t* bits = fb_level(S, L);
    const int n = S.g.n;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(WAVE - 1)); base < n; base += stride) {
        const int v = (int)base + lane_id();
        const uint64_t m = __ballot(v < n && lvl[v] == L);
        if (lane_id() == 0) bits[base >> 6] = m;
    }
    // read by the next kernel, once every workgroup of this one is done
    if (blockIdx.x == 0 && threadIdx.x == 0) S.ctr->fb_level = L;
}

void launch_bu_bits(hipStream_t st, const SlotView* slots, const int* list, int nlist, int gridx, int mode) {
    if (mode == BFS_GR) hipLaunchKernelGGL(k_bu_bits<BFS_GR>, dim3(gridx, nlist), dim3(256), 0, st, slots, list);
    else if (mode == BFS_FWD) hipLaunchKernelGGL(k_bu_bits<BFS_FWD>, dim3(gridx, nlist), dim3(256), 0, st, slots, list);
    else hipLaunchKernelGGL(k_bu_bits<BFS_UND>, dim3(gridx, nlist), dim3(256), 0, st, slots, list);
}

void launch_bfs_bu(hipStream_t st, const SlotView* slots, const int* list, int nlist, int gridx,
                   int mode) {
    const int gh = per_slot_grid(gridx < 2048 ? gridx : 2048, nlist);
    if (mode == BFS_GR) {
        hipLaunchKernelGGL(k_bfs_bu<BFS_GR>, dim3(gridx, nlist), dim3(256), 0, st, slots, list);
        hipLaunchKernelGGL(k_bfs_bu_med<BFS_GR>, dim3(gh, nlist), dim3(256), 0, st, slots, list);
        hipLaunchKernelGGL(k_bfs_bu_heavy<BFS_GR>, dim3(gh, nlist), dim3(256), 0, st, slots, list);
    } else if (mode == BFS_FWD) {
        hipLaunchKernelGGL(k_bfs_bu<BFS_FWD>, dim3(gridx, nlist), dim3(256), 0, st, slots, list);
        hipLaunchKernelGGL(k_bfs_bu_med<BFS_FWD>, dim3(gh, nlist), dim3(256), 0, st, slots, list);
        hipLaunchKernelGGL(k_bfs_bu_heavy<BFS_FWD>, dim3(gh, nlist), dim3(256), 0, st, slots, list);
    } else {
        hipLaunchKernelGGL(k_bfs_bu<BFS_UND>, dim3(gridx, nlist), dim3(256), 0, st, slots, list);
        hipLaunchKernelGGL(k_bfs_bu_med<BFS_UND>, dim3(gh, nlist), dim3(256), 0, st, slots, list);
        hipLaunchKernelGGL(k_bfs_bu_heavy<BFS_UND>, dim3(gh, nlist), dim3(256), 0, st, slots, list);
    }
    hipLaunchKernelGGL(k_bfs_advance, dim3(1, nlist), dim3(64), 0, st, slots, list, 1);
}

// Single-workgroup BFS: runs levels back to back with only workgroup
// barriers while the frontier holds at most kSmallFrontier vertices and at
// most max_levels levels.  This is the deep-and-narrow case (grids), where a
// launch per level would cost more than the level itself.
template <int MODE>
__global__ __launch_bounds__(1024) void k_bfs_small(const SlotView* __restrict__ slots,
                                                    const int* __restrict__ list, int max_levels) {
    const GSlot S = load_slot(slots, list);
    __shared__ int heavy[1024];
    __shared__ int nheavy;
    __shared__ int next_cnt;
    __shared__ int64_t next_deg;
    __shared__ uint64_t red[1024 / WAVE];
    GAS int32_t* lvl = bfs_lvl<MODE>(S);
    int L = S.ctr->level;
    int fsz = S.ctr->fsize[L & 3];
    Work wk;
    int done_levels = 0;
    int visited = 0;
    int64_t fdeg = S.ctr->fedges[L & 3];
    while (fsz > 0 && fsz <= kSmallFrontier && fdeg <= S.bfs_edges && done_levels < max_levels) {
        const GAS int32_t* in = (L & 1) ? S.fb : S.fa;
        GAS int32_t* out = (L & 1) ? S.fa : S.fb;
        if (threadIdx.x == 0) { nheavy = 0; next_cnt = 0; next_deg = 0; }
        __syncthreads();
        for (int i = threadIdx.x; i < fsz; i += blockDim.x) {
            const int v = in[i];
            wk.verts++;
            const int rb = S.g.row_ptr[v], re = S.g.row_ptr[v + 1];
            if (re - rb > kLightDeg) {
                const int k = atomicAdd(&nheavy, 1);
                if (k < 1024) { heavy[k] = v; continue; }
            }
            wk.edges += (uint64_t)(re - rb);
            td_entries<MODE, 4, false>(S, rb, re, 1, L, lvl, [&](int w) {
                const int p = atomicAdd(&next_cnt, 1);
                out[p] = w;
                add64_lds(&next_deg, (int64_t)(S.g.row_ptr[w + 1] - S.g.row_ptr[w]));
                wk.relabels++;
            }, td_fresh<MODE>(S, v));
        }
        __syncthreads();
        const int nh = min(nheavy, 1024);
        for (int k = 0; k < nh; ++k) {
            const int v = heavy[k];
            const int rb = S.g.row_ptr[v], re = S.g.row_ptr[v + 1];
            if (threadIdx.x == 0) wk.edges += (uint64_t)(re - rb);
            td_entries<MODE, 2, false>(S, rb + (int)threadIdx.x, re, (int)blockDim.x, L, lvl, [&](int w) {
                const int p = atomicAdd(&next_cnt, 1);
                out[p] = w;
                add64_lds(&next_deg, (int64_t)(S.g.row_ptr[w + 1] - S.g.row_ptr[w]));
                wk.relabels++;
            }, td_fresh<MODE>(S, v));
        }
        __syncthreads();
        fsz = next_cnt;
        fdeg = next_deg;
        ++L;
        ++done_levels;
        visited += fsz;
        __syncthreads();  // next_cnt / next_deg are reset by the next level
    }
    if (threadIdx.x == 0 && done_levels > 0) {
        // the ring as a multi-workgroup continuation expects it: level L's
        // frontier in slot L, slot L+1 empty for the next level's appends
        S.ctr->fsize[L & 3] = fsz;
        S.ctr->fedges[L & 3] = fdeg;
        S.ctr->fsize[(L + 1) & 3] = 0;
        S.ctr->fedges[(L + 1) & 3] = 0;
        S.ctr->visited = S.ctr->visited + visited;
        S.ctr->level = L;
    }
    work_flush(S, wk, red, WF_BFS);
}

// Barrier of the gridDim.x workgroups of one slot in a multi-workgroup BFS.
// Every thread publishes its writes (agent-scope release) before the
// workgroup arrives, and takes the others' (acquire) after the generation
// moves; the last arrival resets the count before it moves the generation,
// so no workgroup can arrive at the next barrier early.  The wait is bounded:
// past spin_ticks of the wall clock it gives up, flags the slot's error word
// (bit 4) and returns false; the scheduler then runs that BFS again with a
// single workgroup.  The multi-workgroup launches of all engines of the
// process together hold at most 1024 workgroups of 256 threads (each launch
// at most its share, launch_bfs_small's cap_total), half of what the device
// holds, and no other kernel waits on them, so a missing workgroup is at
// worst still being dispatched.
__device__ __forceinline__ bool slot_barrier(GAS Ctr* c, uint32_t& gen, long long spin_ticks, int* ok_s) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0) {
        int ok = 1;
        const int G = (int)gridDim.x;
        // acq_rel: the last arrival's release of the generation then carries
        // every workgroup's writes (each released them before arriving)
        const int arrived = __hip_atomic_fetch_add(&c->bar_count, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) + 1;
        if (arrived == G) {
            __hip_atomic_store(&c->bar_count, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&c->bar_gen, gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            const long long t0 = wall_clock64();
            // Relaxed polls: the acquire fence after the barrier pairs with
            // the release store of the generation.  An acquire load here
            // would invalidate this XCD's L2 on every poll, under the feet of
            // every other workgroup on the XCD.
            while (__hip_atomic_load(&c->bar_gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) {
                __builtin_amdgcn_s_sleep(1);
                if (wall_clock64() - t0 > spin_ticks) {
                    __hip_atomic_fetch_or(&c->err, 4, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ok = 0;
                    break;
                }
            }
        }
        *ok_s = ok;
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    gen += 1u;
    return *ok_s != 0;
}

// Multi-workgroup version of k_bfs_small: gridDim.x workgroups of 256
// threads per slot share each level's frontier (grid-stride) and append the
// next one through LDS staging; the level counts are read after the barrier,
// so every workgroup takes the same decision to go on.  The stopping rules
// and the ring left behind are k_bfs_small's.
template <int MODE>
__global__ __launch_bounds__(256) void k_bfs_multi(const SlotView* __restrict__ slots,
                                                   const int* __restrict__ list, int max_levels,
                                                   long long spin_ticks) {
    const GSlot S = load_slot(slots, list);
    __shared__ Stage<1024> st;
    __shared__ int heavy[256];
    __shared__ int nheavy;
    __shared__ int ok_s;
    __shared__ int next_fsz;
    __shared__ int64_t next_fdeg;
    __shared__ uint64_t red[256 / WAVE];
    GAS int32_t* lvl = bfs_lvl<MODE>(S);
    const int G = (int)gridDim.x;
    int L = S.ctr->level;
    int fsz = S.ctr->fsize[L & 3];
    int64_t fdeg = S.ctr->fedges[L & 3];
    // every workgroup reads the same generation: it only moves once all arrived
    uint32_t gen = __hip_atomic_load(&S.ctr->bar_gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    Work wk;
    int done_levels = 0;
    int visited = 0;
    while (fsz > 0 && fsz <= kSmallFrontier && fdeg <= S.bfs_edges && done_levels < max_levels) {
        const GAS int32_t* in = (L & 1) ? S.fb : S.fa;
        GAS int32_t* out = (L & 1) ? S.fa : S.fb;
        GAS int32_t* ocnt = &S.ctr->fsize[(L + 1) & 3];
        GAS int64_t* odeg = &S.ctr->fedges[(L + 1) & 3];
        stage_init(st);
        if (threadIdx.x == 0) nheavy = 0;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            // level L + 2's counts start at zero (nobody reads or appends them
            // before the next barrier)
            S.ctr->fsize[(L + 2) & 3] = 0;
            S.ctr->fedges[(L + 2) & 3] = 0;
        }
        __syncthreads();
        for (int i = (int)(blockIdx.x * blockDim.x + threadIdx.x); i < fsz; i += G * (int)blockDim.x) {
            const int v = in[i];
            wk.verts++;
            const int rb = S.g.row_ptr[v], re = S.g.row_ptr[v + 1];
            if (re - rb > kLightDeg) {
                const int k = atomicAdd(&nheavy, 1);
                if (k < 256) { heavy[k] = v; continue; }
            }
            wk.edges += (uint64_t)(re - rb);
            td_entries_deep<MODE, 4>(S, rb, re, 1, L, lvl, [&](int w, int deg) {
                stage_push(st, w, deg, out, ocnt, odeg);
                wk.relabels++;
            }, td_fresh<MODE>(S, v));
        }
        __syncthreads();
        const int nh = min(nheavy, 256);
        for (int k = 0; k < nh; ++k) {
            const int v = heavy[k];
            const int rb = S.g.row_ptr[v], re = S.g.row_ptr[v + 1];
            if (threadIdx.x == 0) wk.edges += (uint64_t)(re - rb);
            td_entries_deep<MODE, 2>(S, rb + (int)threadIdx.x, re, (int)blockDim.x, L, lvl, [&](int w, int deg) {
                stage_push(st, w, deg, out, ocnt, odeg);
                wk.relabels++;
            }, td_fresh<MODE>(S, v));
        }
        stage_flush(st, out, ocnt, odeg);
        if (!slot_barrier(S.ctr, gen, spin_ticks, &ok_s)) break;
        // the next level's counts: nobody changes them from here on
        if (threadIdx.x == 0) {
            next_fsz = __hip_atomic_load(ocnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            next_fdeg = __hip_atomic_load(odeg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
        fsz = next_fsz;
        fdeg = next_fdeg;
        ++L;
        ++done_levels;
        visited += fsz;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && done_levels > 0) {
        // the ring as a multi-workgroup continuation expects it: level L's
        // frontier in slot L (appended there), slot L+1 zeroed by the last level
        S.ctr->visited = S.ctr->visited + visited;
        S.ctr->level = L;
    }
    work_flush(S, wk, red, WF_BFS);
}

void launch_bfs_small(hipStream_t st, const SlotView* slots, const int* list, int nlist, int mode,
                      int max_levels, int wgs, long long spin_ticks, int cap_total) {
    if (wgs <= 1) {
        if (mode == BFS_GR) hipLaunchKernelGGL(k_bfs_small<BFS_GR>, dim3(1, nlist), dim3(1024), 0, st, slots, list, max_levels);
        else if (mode == BFS_FWD) hipLaunchKernelGGL(k_bfs_small<BFS_FWD>, dim3(1, nlist), dim3(1024), 0, st, slots, list, max_levels);
        else hipLaunchKernelGGL(k_bfs_small<BFS_UND>, dim3(1, nlist), dim3(1024), 0, st, slots, list, max_levels);
        return;
    }
    // at most cap_total workgroups per launch (at most 256): every multi-
    // workgroup launch of every engine fits on the device at once, and many
    // slots share the barriers' cost (grid 1024^2 with 128 slots: 2 per slot
    // 57 q/s against 54 at 4)
    const long long cap = std::min(256, std::max(1, cap_total));
    int g = wgs;
    while (g > 1 && (long long)g * nlist > cap) g >>= 1;
    if (mode == BFS_GR) hipLaunchKernelGGL(k_bfs_multi<BFS_GR>, dim3(g, nlist), dim3(256), 0, st, slots, list, max_levels, spin_ticks);
    else if (mode == BFS_FWD) hipLaunchKernelGGL(k_bfs_multi<BFS_FWD>, dim3(g, nlist), dim3(256), 0, st, slots, list, max_levels, spin_ticks);
    else hipLaunchKernelGGL(k_bfs_multi<BFS_UND>, dim3(g, nlist), dim3(256), 0, st, slots, list, max_levels, spin_ticks);
}

// ------------------------------------------------------- local relabel (a14)
// BC:236-279 and BC:1582-1721 relabel the tiles that hold active vertices by
// a BFS from t and the deficits inside them, for at most band + 2 levels, and
// overwrite the heights the BFS reaches (which may lower them).  Here the
// same bounded BFS over the active tiles only ever raises heights: the
// heights around the relabelled set R' enter as sources, so the result is the
// largest valid labelling of R' below lcap and RM never has invalid entries
// to repair.  Vertices of R' are marked n + 1 while the BFS runs.

__device__ __forceinline__ bool local_in_region(const GSlot& S, int v) {
    return S.tile_cnt[S.g.tile_of[v]] > 0;
}

__global__ __launch_bounds__(256) void k_local_mark(const SlotView* __restrict__ slots,
                                                    const int* __restrict__ list, int lcap) {
    const GSlot S = load_slot(slots, list);
    __shared__ Stage<1024> st;
    __shared__ uint64_t red[256 / WAVE];
    stage_init(st);
    if (blockIdx.x == 0)
        for (int i = threadIdx.x; i < 2 * kLocalLevels; i += blockDim.x) S.lbkt[i] = 0;
    __syncthreads();
    const int n = S.g.n;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    Work wk;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int v = (int)i;
        if (!local_in_region(S, v)) continue;
        const int hv = S.h[v];
        S.aux[v] = hv;
        wk.verts++;
        if (hv >= lcap || v == S.s || (S.mask != nullptr && S.mask[v] == 0 && v != S.t)) continue;
        bool seed = false;
        if (hv == 0) {
            if (S.phase == 0) seed = (v == S.t) || (S.excess[v] < 0);
            else seed = (v != S.t) && (S.excess[v] < 0);
        }
        if (seed) {
            stage_push(st, v, S.g.row_ptr[v + 1] - S.g.row_ptr[v], S.fa, &S.ctr->fsize[0], &S.ctr->fedges[0]);
        } else {
            S.h[v] = n + 1;
        }
    }
    stage_flush(st, S.fa, &S.ctr->fsize[0], &S.ctr->fedges[0]);
    work_flush(S, wk, red, WF_BFS);
}

void launch_local_mark(hipStream_t st, const SlotView* slots, const int* list, int nlist, int gridx,
                       int lcap) {
    hipLaunchKernelGGL(k_local_mark, dim3(gridx, nlist), dim3(256), 0, st, slots, list, lcap);
}

// b(v) for marked v into fb[v]; counts per b below lcap into lbkt[0..lcap).
// Rows up to kLightDeg by one lane, longer rows by a wave.
__global__ __launch_bounds__(256) void k_local_bound(const SlotView* __restrict__ slots,
                                                     const int* __restrict__ list, int lcap) {
    const GSlot S = load_slot(slots, list);
    __shared__ int heavy[256];
    __shared__ int nheavy;
    __shared__ uint64_t red[256 / WAVE];
    const int n = S.g.n;
    const int mark = n + 1;
    const int lane = lane_id();
    const int wave = threadIdx.x / WAVE;
    Work wk;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n; base += (int64_t)gridDim.x * blockDim.x) {
        if (threadIdx.x == 0) nheavy = 0;
        __syncthreads();
        const int v = (int)(base + threadIdx.x);
        if (v < n && S.h[v] == mark) {
            const int rb = S.g.row_ptr[v], re = S.g.row_ptr[v + 1];
            if (re - rb > kLightDeg) {
                heavy[atomicAdd(&nheavy, 1)] = v;
            } else {
                int b = INT_MAX;
                wk.edges += (uint64_t)(re - rb);
                for (int e = rb; e < re; ++e) {
                    if (ld_res(&S.res[e]) <= 0) continue;
                    const int hx = S.h[S.g.col[e]];
                    if (hx != mark) b = min(b, hx + 1);
                }
                const int bb = b < lcap ? b : lcap;
                S.fb[v] = bb;
                if (bb < lcap) add32(&S.lbkt[bb], 1);
            }
        }
        __syncthreads();
        const int nh = nheavy;
        for (int k = wave; k < nh; k += (int)(blockDim.x / WAVE)) {
            const int u = heavy[k];
            const int rb = S.g.row_ptr[u], re = S.g.row_ptr[u + 1];
            if (lane == 0) wk.edges += (uint64_t)(re - rb);
            int b = INT_MAX;
            for (int e = rb + lane; e < re; e += WAVE) {
                if (ld_res(&S.res[e]) <= 0) continue;
                const int hx = S.h[S.g.col[e]];
                if (hx != mark) b = min(b, hx + 1);
            }
            b = wave_min_i32(b);
            if (lane == 0) {
                const int bb = b < lcap ? b : lcap;
                S.fb[u] = bb;
                if (bb < lcap) add32(&S.lbkt[bb], 1);
            }
        }
        __syncthreads();
    }
    work_flush(S, wk, red, WF_BFS);
}

void launch_local_bound(hipStream_t st, const SlotView* slots, const int* list, int nlist, int gridx,
                        int lcap) {
    hipLaunchKernelGGL(k_local_bound, dim3(gridx, nlist), dim3(256), 0, st, slots, list, lcap);
}

// Exclusive offsets of the lcap bucket counts into off[0..lcap], by one thread.
__device__ __forceinline__ void local_offsets(const GSlot& S, int lcap, int* off) {
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int b = 0; b < lcap; ++b) { off[b] = acc; acc += S.lbkt[b]; }
        off[lcap] = acc;
    }
    __syncthreads();
}

__global__ __launch_bounds__(256) void k_local_bucket(const SlotView* __restrict__ slots,
                                                      const int* __restrict__ list, int lcap) {
    const GSlot S = load_slot(slots, list);
    __shared__ int off[kLocalLevels + 1];
    local_offsets(S, lcap, off);
    const int n = S.g.n;
    const int mark = n + 1;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int v = (int)i;
        if (S.h[v] != mark) continue;
        const int b = S.fb[v];
        if (b >= lcap) continue;
        const int pos = off[b] + add32(&S.lbkt[kLocalLevels + b], 1);
        S.hl[pos] = v;
    }
}

void launch_local_bucket(hipStream_t st, const SlotView* slots, const int* list, int nlist, int gridx,
                         int lcap) {
    hipLaunchKernelGGL(k_local_bucket, dim3(gridx, nlist), dim3(256), 0, st, slots, list, lcap);
}

// One workgroup per slot.  Level L: bucket L joins the frontier, then the
// frontier is expanded backwards over residual entries into marked vertices.
// Stops early (ctr->level = L + 1) when the frontier outgrows one workgroup;
// otherwise ctr->level = lcap.
__global__ __launch_bounds__(1024) void k_local_bfs(const SlotView* __restrict__ slots,
                                                    const int* __restrict__ list, int lcap) {
    const GSlot S = load_slot(slots, list);
    __shared__ int off[kLocalLevels + 1];
    __shared__ int heavy[1024];
    __shared__ int nheavy;
    __shared__ int next_cnt;
    __shared__ int64_t next_deg;
    __shared__ uint64_t red[1024 / WAVE];
    local_offsets(S, lcap, off);
    const int n = S.g.n;
    const int mark = n + 1;
    int fsz = S.ctr->fsize[0];
    int64_t fdeg = S.ctr->fedges[0];
    GAS int32_t* in = S.fa;
    GAS int32_t* out = S.fb;
    Work wk;
    int Lf = lcap;
    for (int L = 0; L < lcap; ++L) {
        // bucket L joins level L
        if (threadIdx.x == 0) { next_cnt = 0; next_deg = 0; }
        __syncthreads();
        for (int i = off[L] + (int)threadIdx.x; i < off[L + 1]; i += blockDim.x) {
            const int u = S.hl[i];
            if (cas32(&S.h[u], mark, L)) {
                in[fsz + atomicAdd(&next_cnt, 1)] = u;
                add64_lds(&next_deg, (int64_t)(S.g.row_ptr[u + 1] - S.g.row_ptr[u]));
                wk.relabels++;
            }
        }
        __syncthreads();
        fsz += next_cnt;
        fdeg += next_deg;
        __syncthreads();  // everybody has read next_cnt before it is reset
        if (fsz > kSmallFrontier || fdeg > kSmallEdges) { Lf = L + 1; break; }
        if (L + 1 >= lcap) break;
        if (fsz == 0) continue;
        if (threadIdx.x == 0) { nheavy = 0; next_cnt = 0; next_deg = 0; }
        __syncthreads();
        for (int i = threadIdx.x; i < fsz; i += blockDim.x) {
            const int v = in[i];
            wk.verts++;
            const int rb = S.g.row_ptr[v], re = S.g.row_ptr[v + 1];
            if (re - rb > kLightDeg) {
                const int k = atomicAdd(&nheavy, 1);
                if (k < 1024) { heavy[k] = v; continue; }
            }
            wk.edges += (uint64_t)(re - rb);
            for (int e = rb; e < re; ++e) {
                const int w = S.g.col[e];
                if (S.h[w] != mark) continue;
                if (S.g.pcap[e] - ld_res(&S.res[e]) <= 0) continue;  // res[rev e]
                if (cas32(&S.h[w], mark, L + 1)) {
                    out[atomicAdd(&next_cnt, 1)] = w;
                    add64_lds(&next_deg, (int64_t)(S.g.row_ptr[w + 1] - S.g.row_ptr[w]));
                    wk.relabels++;
                }
            }
        }
        __syncthreads();
        const int nh = min(nheavy, 1024);
        for (int k = 0; k < nh; ++k) {
            const int v = heavy[k];
            const int rb = S.g.row_ptr[v], re = S.g.row_ptr[v + 1];
            if (threadIdx.x == 0) wk.edges += (uint64_t)(re - rb);
            for (int e = rb + (int)threadIdx.x; e < re; e += blockDim.x) {
                const int w = S.g.col[e];
                if (S.h[w] != mark) continue;
                if (S.g.pcap[e] - ld_res(&S.res[e]) <= 0) continue;  // res[rev e]
                if (cas32(&S.h[w], mark, L + 1)) {
                    out[atomicAdd(&next_cnt, 1)] = w;
                    add64_lds(&next_deg, (int64_t)(S.g.row_ptr[w + 1] - S.g.row_ptr[w]));
                    wk.relabels++;
                }
            }
        }
        __syncthreads();
        fsz = next_cnt;
        fdeg = next_deg;
        GAS int32_t* tmp = in;
        in = out;
        out = tmp;
        __syncthreads();
    }
    if (threadIdx.x == 0) S.ctr->level = Lf;
    work_flush(S, wk, red, WF_BFS);
}

void launch_local_bfs(hipStream_t st, const SlotView* slots, const int* list, int nlist, int lcap) {
    hipLaunchKernelGGL(k_local_bfs, dim3(1, nlist), dim3(1024), 0, st, slots, list, lcap);
}

__global__ __launch_bounds__(256) void k_local_finish(const SlotView* __restrict__ slots,
                                                      const int* __restrict__ list) {
    const GSlot S = load_slot(slots, list);
    const int n = S.g.n;
    const int mark = n + 1;
    const int Lf = S.ctr->level;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int v = (int)i;
        if (!local_in_region(S, v)) continue;
        const int hv = S.h[v];
        const int ho = S.aux[v];
        S.h[v] = max(ho, hv == mark ? Lf : hv);
    }
}

void launch_local_finish(hipStream_t st, const SlotView* slots, const int* list, int nlist, int gridx) {
    hipLaunchKernelGGL(k_local_finish, dim3(gridx, nlist), dim3(256), 0, st, slots, list);
}

// ------------------------------------------------ active list + counters (a15, a18)
// BC:664-695 (active list, counts), BC:559-573 (positive excess sum),
// BC:283-318 (per-region active counts), BC:519-536 (active height band).
// Active: excess > 0, h < n, v != s and (phase 0) v != t.  Light rows go to
// fa[0..nlight), heavy rows to fa[n-1], fa[n-2], ...

__global__ __launch_bounds__(256) void k_active(const SlotView* __restrict__ slots,
                                                const int* __restrict__ list) {
    const GSlot S = load_slot(slots, list);
    __shared__ Stage<1024> st;
    __shared__ int64_t red64[256 / WAVE];
    __shared__ uint64_t redu[256 / WAVE];
    __shared__ int bmin, bmax;
    stage_init(st);
    if (threadIdx.x == 0) { bmin = INT_MAX; bmax = -1; }
    __syncthreads();
    const int n = S.g.n;
    int64_t pos = 0, negs = 0;
    uint64_t act = 0, high = 0, negc = 0, nv = 0;
    int hmin = INT_MAX, hmax = -1;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int v = (int)i;
        ++nv;
        if (v == S.s && !S.inj) continue;
        // no flow has entered or left v (SlotView::recv): its excess is 0
        if (S.recv_ok && v != S.t && S.recv[v] == 0) continue;
        const int64_t ex = S.excess[v];
        if (v == S.t) S.ctr->ex_t = ex;
        if (v == S.t && S.phase == 0) continue;
        if (ex > 0) pos += ex;
        if (ex < 0) { negs += ex; ++negc; }
        if (ex <= 0) continue;
        const int hv = S.h[v];
        if (hv >= n) continue;
        ++act;
        if (hv > S.H_cut) ++high;
        hmin = min(hmin, hv);
        hmax = max(hmax, hv);
        {
            const int tl = S.g.tile_of[v];
            add32(&S.tile_cnt[tl], 1);
            add64(&S.tile_ex[tl], ex);
        }
        const int deg = S.g.row_ptr[v + 1] - S.g.row_ptr[v];
        if (deg <= kLightDeg) {
            stage_push(st, v, 0, S.fa, &S.ctr->nlight, (GAS int64_t*)nullptr);
        } else {
            const int k = add32(&S.ctr->nheavy, 1);
            S.fa[n - 1 - k] = v;
            if (deg > kGiantDeg) {  // also to the giant list: row minimum slot starts empty
                const int g = add32(&S.ctr->ngiant, 1);
                S.giant[g] = v;
                S.giant[S.gcap + g] = INT_MAX;
            }
        }
    }
    stage_flush(st, S.fa, &S.ctr->nlight, (GAS int64_t*)nullptr);
    hmin = wave_min_i32(hmin);
    hmax = wave_max_i32(hmax);
    if (lane_id() == 0) { atomicMin(&bmin, hmin); atomicMax(&bmax, hmax); }
    const int64_t bpos = block_sum<int64_t>(pos, red64);
    const int64_t bneg = block_sum<int64_t>(negs, red64);
    const uint64_t bact = block_sum<uint64_t>(act, redu);
    const uint64_t bhigh = block_sum<uint64_t>(high, redu);
    const uint64_t bnegc = block_sum<uint64_t>(negc, redu);
    const uint64_t bnv = block_sum<uint64_t>(nv, redu);
    if (threadIdx.x == 0) {
        if (bpos) add64(&S.ctr->pos_excess, bpos);
        if (bneg) add64(&S.ctr->neg_sum, bneg);
        if (bact) add32(&S.ctr->active, (int32_t)bact);
        if (bhigh) add32(&S.ctr->active_high, (int32_t)bhigh);
        if (bnegc) add32(&S.ctr->neg_count, (int32_t)bnegc);
        if (bnv) addu64(&S.ctr->wk[WF_ACTIVE][1], bnv);
        if (bmax >= 0) { min32(&S.ctr->hmin, bmin); max32(&S.ctr->hmax, bmax); }
    }
}

void launch_active(hipStream_t st, const SlotView* slots, const int* list, int nlist, int gridx) {
    hipLaunchKernelGGL(k_active, dim3(gridx, nlist), dim3(256), 0, st, slots, list);
}

// Sparse count of a fused cycle: the candidates are the active list being
// replaced (fa, prev_nlight / prev_nheavy) and the vertices pushed into this
// cycle (hl, ntouched); each is counted once (exchange on its aux stamp).  The
// new list goes to fb (light at the front, heavy at the back) and
// k_list_copy moves it to fa.  Same counters as k_active, except that excess
// nobody pushed to or discharged is not re-summed (deficits are counted by the
// full pass that follows every global relabel).
__global__ __launch_bounds__(256) void k_active_sparse(const SlotView* __restrict__ slots,
                                                       const int* __restrict__ list) {
    const GSlot S = load_slot(slots, list);
    __shared__ Stage<1024> st;
    __shared__ int64_t red64[256 / WAVE];
    __shared__ uint64_t redu[256 / WAVE];
    __shared__ int bmin, bmax;
    stage_init(st);
    if (threadIdx.x == 0) { bmin = INT_MAX; bmax = -1; }
    __syncthreads();
    const int n = S.g.n;
    const int pl = S.ctr->prev_nlight, ph = S.ctr->prev_nheavy, nt = S.ctr->ntouched;
    const int cstamp = S.tstamp - 1;  // -(2 epoch + 2)
    const int total = pl + ph + nt;
    int64_t pos = 0;
    uint64_t act = 0, high = 0, nv = 0;
    int hmin = INT_MAX, hmax = -1;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        const int v = i < pl ? S.fa[i] : (i < pl + ph ? S.fa[n - 1 - (i - pl)] : S.hl[i - pl - ph]);
        if (__hip_atomic_exchange(&S.aux[v], cstamp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == cstamp) continue;
        ++nv;
        if (v == S.s && !S.inj) continue;
        const int64_t ex = ld64(&S.excess[v]);
        if (v == S.t) S.ctr->ex_t = ex;
        if (v == S.t && S.phase == 0) continue;
        if (ex <= 0) continue;
        pos += ex;
        const int hv = S.h[v];
        if (hv >= n) continue;
        ++act;
        if (hv > S.H_cut) ++high;
        hmin = min(hmin, hv);
        hmax = max(hmax, hv);
        {
            const int tl = S.g.tile_of[v];
            add32(&S.tile_cnt[tl], 1);
            add64(&S.tile_ex[tl], ex);
        }
        const int deg = S.g.row_ptr[v + 1] - S.g.row_ptr[v];
        if (deg <= kLightDeg) {
            stage_push(st, v, 0, S.fb, &S.ctr->nlight, (GAS int64_t*)nullptr);
        } else {
            const int k = add32(&S.ctr->nheavy, 1);
            S.fb[n - 1 - k] = v;
            if (deg > kGiantDeg) {
                const int g = add32(&S.ctr->ngiant, 1);
                S.giant[g] = v;
                S.giant[S.gcap + g] = INT_MAX;
            }
        }
    }
    stage_flush(st, S.fb, &S.ctr->nlight, (GAS int64_t*)nullptr);
    hmin = wave_min_i32(hmin);
    hmax = wave_max_i32(hmax);
    if (lane_id() == 0) { atomicMin(&bmin, hmin); atomicMax(&bmax, hmax); }
    const int64_t bpos = block_sum<int64_t>(pos, red64);
    const uint64_t bact = block_sum<uint64_t>(act, redu);
    const uint64_t bhigh = block_sum<uint64_t>(high, redu);
    const uint64_t bnv = block_sum<uint64_t>(nv, redu);
    if (threadIdx.x == 0) {
        if (bpos) add64(&S.ctr->pos_excess, bpos);
        if (bact) add32(&S.ctr->active, (int32_t)bact);
        if (bhigh) add32(&S.ctr->active_high, (int32_t)bhigh);
        if (bnv) addu64(&S.ctr->wk[WF_ACTIVE][1], bnv);
        if (bmax >= 0) { min32(&S.ctr->hmin, bmin); max32(&S.ctr->hmax, bmax); }
    }
}

__global__ __launch_bounds__(256) void k_list_copy(const SlotView* __restrict__ slots,
                                                   const int* __restrict__ list) {
    const GSlot S = load_slot(slots, list);
    const int n = S.g.n;
    const int nl = S.ctr->nlight, nh = S.ctr->nheavy;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nl + nh; i += gridDim.x * blockDim.x) {
        const int j = i < nl ? i : n - 1 - (i - nl);
        S.fa[j] = S.fb[j];
    }
}

void launch_active_sparse(hipStream_t st, const SlotView* slots, const int* list, int nlist, int gridx) {
    hipLaunchKernelGGL(k_active_sparse, dim3(gridx, nlist), dim3(256), 0, st, slots, list);
    hipLaunchKernelGGL(k_list_copy, dim3(gridx, nlist), dim3(256), 0, st, slots, list);
}

// ------------------------------------------------------------- discharge (a16)
// One discharge cycle (the reference's cycle is BC:467-511): find the lowest
// neighbour over residual entries; if v is not above it, relabel v to lowest
// + 1 (the reference's relabel); then push min(excess, res) to every entry at
// that lowest height in row order until the excess is gone.  A cycle is a
// relabel followed by the reference's single pushes, done in two row scans,
// so excess moves on the cycle that relabels instead of the next one.

struct TileCtx {
    int a = 0, b = 0;          // staged tile [a, b); b == a means list mode
    int32_t* lh = nullptr;     // LDS heights of the tile
    int64_t* ldx = nullptr;    // LDS pending excess deltas of the tile
};

__device__ __forceinline__ int nbr_h(const GSlot& S, const TileCtx& tc, int w) {
    return (w >= tc.a && w < tc.b) ? tc.lh[w - tc.a] : S.h[w];
}

// w received excess this cycle: listed once per touch epoch for a sparse count.
__device__ __forceinline__ void touch(const GSlot& S, int w) {
    if (__hip_atomic_exchange(&S.aux[w], S.tstamp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != S.tstamp)
        S.hl[wave_slot(&S.ctr->ntouched)] = w;
}

__device__ __forceinline__ void push_into(const GSlot& S, const TileCtx& tc, int w, int64_t d) {
    if (w >= tc.a && w < tc.b) {
        add64_lds(&tc.ldx[w - tc.a], d);  // touched when the tile flushes
    } else {
        add64(&S.excess[w], d);
        touch(S, w);
    }
}

// Owner view of v's excess.
__device__ __forceinline__ int64_t own_excess(const GSlot& S, const TileCtx& tc, int v) {
    int64_t ex = ld64(&S.excess[v]);
    if (v >= tc.a && v < tc.b) ex += tc.ldx[v - tc.a];
    return ex;
}

// Take tot out of v; returns the owner's new view.
__device__ __forceinline__ int64_t take_excess(const GSlot& S, const TileCtx& tc, int v,
                                               int64_t ex_view, int64_t tot) {
    if (v >= tc.a && v < tc.b) {
        add64_lds(&tc.ldx[v - tc.a], -tot);
        return ex_view - tot;
    }
    return add64(&S.excess[v], -tot) - tot;
}

__device__ __forceinline__ int own_h(const GSlot& S, const TileCtx& tc, int v) {
    return (v >= tc.a && v < tc.b) ? tc.lh[v - tc.a] : S.h[v];
}

__device__ __forceinline__ void set_h(const GSlot& S, const TileCtx& tc, int v, int hv) {
    if (v >= tc.a && v < tc.b) tc.lh[v - tc.a] = hv;
    else S.h[v] = hv;
}

// Light row: one lane, `cycles` discharge cycles.
__device__ void lane_discharge_minfirst(const GSlot& S, const TileCtx& tc, int v, int cycles, Work& wk) {
    const int n = S.g.n;
    const int rb = S.g.row_ptr[v], re = S.g.row_ptr[v + 1];
    int hv = own_h(S, tc, v);
    int64_t ex = own_excess(S, tc, v);
    if (hv > S.disc_cap) return;  // above this step's band
    for (int c = 0; c < cycles; ++c) {
        if (ex <= 0 || hv >= n) break;
        int hmin = INT_MAX;
        for (int e = rb; e < re; ++e) {
            if (ld64(&S.res[e]) > 0) hmin = min(hmin, nbr_h(S, tc, S.g.col[e]));
        }
        wk.edges += (uint64_t)(re - rb);
        if (hmin >= n) {  // nothing below n is reachable from v
            hv = n;
            set_h(S, tc, v, hv);
            wk.relabels++;
            break;
        }
        if (hv <= hmin) {  // relabel (BC:504-507), then push at the new height
            hv = hmin + 1;
            set_h(S, tc, v, hv);
            wk.relabels++;
        }
        {
            int64_t rem = ex;
            for (int e = rb; e < re && rem > 0; ++e) {
                const int64_t r = ld64(&S.res[e]);
                if (r <= 0) continue;
                const int w = S.g.col[e];
                if (nbr_h(S, tc, w) != hmin) continue;
                const int64_t d = rem < r ? rem : r;
                move_res(S, e, d);
                push_into(S, tc, w, d);
                rem -= d;
                wk.pushes++;
            }
            const int64_t tot = ex - rem;
            if (tot > 0) ex = take_excess(S, tc, v, ex, tot);
        }
    }
}

// Heavy row: one wave, `cycles` discharge cycles; wave-uniform control.
__device__ void wave_discharge_minfirst(const GSlot& S, const TileCtx& tc, int v, int cycles, Work& wk) {
    const int n = S.g.n;
    const int lane = lane_id();
    const int rb = S.g.row_ptr[v], re = S.g.row_ptr[v + 1];
    int hv = __shfl(own_h(S, tc, v), 0, WAVE);
    int64_t ex = __shfl(own_excess(S, tc, v), 0, WAVE);
    if (hv > S.disc_cap) return;  // above this step's band (wave-uniform)
    for (int c = 0; c < cycles; ++c) {
        if (ex <= 0 || hv >= n) break;
        int hmin = INT_MAX;
        for (int e = rb + lane; e < re; e += WAVE)
            if (ld64(&S.res[e]) > 0) hmin = min(hmin, nbr_h(S, tc, S.g.col[e]));
        hmin = wave_min_i32(hmin);
        if (lane == 0) wk.edges += (uint64_t)(re - rb);
        if (hmin >= n) {
            hv = n;
            if (lane == 0) { set_h(S, tc, v, hv); wk.relabels++; }
            break;
        }
        if (hv <= hmin) {  // relabel, then push at the new height
            hv = hmin + 1;
            if (lane == 0) { set_h(S, tc, v, hv); wk.relabels++; }
        }
        {
            int64_t rem = ex;
            for (int base = rb; base < re && rem > 0; base += WAVE) {
                const int e = base + lane;
                int64_t r = 0;
                if (e < re) {
                    r = ld64(&S.res[e]);
                    if (r > 0 && nbr_h(S, tc, S.g.col[e]) != hmin) r = 0;
                    if (r < 0) r = 0;
                }
                const int64_t incl = wave_incl_scan_i64(r);
                const int64_t excl = incl - r;
                const int64_t d = (r > 0 && rem > excl) ? ((rem - excl) < r ? (rem - excl) : r) : 0;
                if (d > 0) {
                    move_res(S, e, d);
                    push_into(S, tc, S.g.col[e], d);
                    wk.pushes++;
                }
                const int64_t chunk = __shfl(incl, WAVE - 1, WAVE);
                rem -= (chunk < rem ? chunk : rem);
            }
            const int64_t tot = ex - rem;
            if (tot > 0) {
                int64_t nex = 0;
                if (lane == 0) nex = take_excess(S, tc, v, ex, tot);
                ex = __shfl(nex, 0, WAVE);
            }
        }
    }
}

// Push-first discharge (mf_params.push_first): one pass pushes the excess into lower
// neighbours (h < hv) as it meets them and stops once the excess is gone,
// keeping the lowest height among the entries that still have residual; only
// a vertex whose whole row was scanned with excess left is relabelled (to
// that height + 1, n when nothing below n - 1 remains) and pushes again at
// once.  Under valid labels the lower neighbours are exactly those at
// hv - 1, so the pushes are the minimum-first ones; the full-row minimum is
// read only when a relabel needs it.
__device__ void lane_discharge(const GSlot& S, const TileCtx& tc, int v, int cycles, Work& wk) {
    if (!S.push_first) { lane_discharge_minfirst(S, tc, v, cycles, wk); return; }
    const int n = S.g.n;
    const int rb = S.g.row_ptr[v], re = S.g.row_ptr[v + 1];
    int hv = own_h(S, tc, v);
    int64_t ex = own_excess(S, tc, v);
    if (hv > S.disc_cap) return;  // above this step's band
    for (int c = 0; c < cycles; ++c) {
        if (ex <= 0 || hv >= n) break;
        for (int pass = 0; pass < 2; ++pass) {
            int64_t rem = ex;
            int m = INT_MAX;
            int e = rb;
            for (; e < re && rem > 0; ++e) {
                const int64_t r = ld64(&S.res[e]);
                if (r <= 0) continue;
                const int w = S.g.col[e];
                const int hw = nbr_h(S, tc, w);
                if (hw < hv) {
                    const int64_t d = rem < r ? rem : r;
                    move_res(S, e, d);
                    push_into(S, tc, w, d);
                    rem -= d;
                    wk.pushes++;
                    if (d < r) m = min(m, hw);
                } else {
                    m = min(m, hw);
                }
            }
            wk.edges += (uint64_t)(e - rb);
            const int64_t tot = ex - rem;
            if (tot > 0) ex = take_excess(S, tc, v, ex, tot);
            if (rem <= 0 || pass == 1) break;
            // the whole row is scanned and excess is left: relabel
            hv = m >= n - 1 ? n : m + 1;
            set_h(S, tc, v, hv);
            wk.relabels++;
            if (hv >= n || ex <= 0) break;
        }
    }
}

__device__ void wave_discharge(const GSlot& S, const TileCtx& tc, int v, int cycles, Work& wk) {
    if (!S.push_first) { wave_discharge_minfirst(S, tc, v, cycles, wk); return; }
    const int n = S.g.n;
    const int lane = lane_id();
    const int rb = S.g.row_ptr[v], re = S.g.row_ptr[v + 1];
    int hv = __shfl(own_h(S, tc, v), 0, WAVE);
    int64_t ex = __shfl(own_excess(S, tc, v), 0, WAVE);
    if (hv > S.disc_cap) return;  // above this step's band (wave-uniform)
    for (int c = 0; c < cycles; ++c) {
        if (ex <= 0 || hv >= n) break;
        for (int pass = 0; pass < 2; ++pass) {
            int64_t rem = ex;
            int m = INT_MAX;
            int base = rb;
            for (; base < re && rem > 0; base += WAVE) {
                const int e = base + lane;
                int64_t r = 0;
                int hw = INT_MAX, w = 0;
                if (e < re) {
                    r = ld64(&S.res[e]);
                    if (r > 0) {
                        w = S.g.col[e];
                        hw = nbr_h(S, tc, w);
                    } else {
                        r = 0;
                    }
                }
                const int64_t a = (r > 0 && hw < hv) ? r : 0;  // into a lower neighbour
                const int64_t incl = wave_incl_scan_i64(a);
                const int64_t excl = incl - a;
                const int64_t d = (a > 0 && rem > excl) ? ((rem - excl) < a ? (rem - excl) : a) : 0;
                if (d > 0) {
                    move_res(S, e, d);
                    push_into(S, tc, w, d);
                    wk.pushes++;
                }
                if (r > d) m = min(m, hw);
                const int64_t chunk = __shfl(incl, WAVE - 1, WAVE);
                rem -= (chunk < rem ? chunk : rem);
            }
            if (lane == 0) wk.edges += (uint64_t)(min(base, re) - rb);
            const int64_t tot = ex - rem;
            if (tot > 0) {
                int64_t nex = 0;
                if (lane == 0) nex = take_excess(S, tc, v, ex, tot);
                ex = __shfl(nex, 0, WAVE);
            }
            if (rem <= 0 || pass == 1) break;
            // the whole row is scanned and excess is left: relabel
            m = wave_min_i32(m);
            hv = m >= n - 1 ? n : m + 1;
            if (lane == 0) { set_h(S, tc, v, hv); wk.relabels++; }
            if (hv >= n || ex <= 0) break;
        }
    }
}

// Block-wide helpers (blockDim a multiple of 64, scratch holds blockDim/64
// entries).  Results are valid in every thread.
__device__ __forceinline__ int block_min_i32(int x, int* scratch) {
    x = wave_min_i32(x);
    const int w = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
    if (lane_id() == 0) scratch[w] = x;
    __syncthreads();
    int m = scratch[0];
    for (int j = 1; j < nw; ++j) m = min(m, scratch[j]);
    __syncthreads();
    return m;
}

// Inclusive scan of x over the workgroup; *total gets the block sum.
__device__ __forceinline__ int64_t block_incl_scan_i64(int64_t x, int64_t* scratch, int64_t* total) {
    const int w = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
    const int64_t incl = wave_incl_scan_i64(x);
    if (lane_id() == WAVE - 1) scratch[w] = incl;
    __syncthreads();
    int64_t before = 0, all = 0;
    for (int j = 0; j < nw; ++j) {
        const int64_t t = scratch[j];
        if (j < w) before += t;
        all += t;
    }
    __syncthreads();
    *total = all;
    return incl + before;
}

// Active-list discharge: light entries one lane each, heavy entries one
// wave each.  Blocks [0, light_blocks) take light entries.
__global__ __launch_bounds__(256) void k_discharge_list(const SlotView* __restrict__ slots,
                                                        const int* __restrict__ list) {
    const GSlot S = load_slot(slots, list);
    __shared__ uint64_t red[256 / WAVE];
    const int nl = S.ctr->nlight, nh = S.ctr->nheavy;
    const int lb = (nl + 255) / 256;
    const int hb = (nh + 3) / 4;
    if ((int)blockIdx.x >= lb + hb) return;
    TileCtx tc;
    Work wk;
    // work blocks [0, lb) light, [lb, lb + hb) heavy, dealt grid-stride so any
    // grid covers the whole list (fused cycles launch before the count is known)
    for (int b = blockIdx.x; b < lb + hb; b += gridDim.x) {
        if (b < lb) {
            const int i = b * 256 + threadIdx.x;
            if (i < nl) lane_discharge(S, tc, S.fa[i], S.K, wk);
        } else {
            const int k = (b - lb) * 4 + (int)(threadIdx.x / WAVE);
            if (k < nh) {
                const int v = S.fa[S.g.n - 1 - k];
                if (S.g.row_ptr[v + 1] - S.g.row_ptr[v] <= kGiantDeg) wave_discharge(S, tc, v, S.Kh, wk);
            }
        }
    }
    work_flush(S, wk, red, WF_DISCHARGE);
}

// One persistent launch for the list discharge of every slot of a step, fed
// by a device work queue: dq = {next, total, prefix[0..nlist]}, the prefix of
// the slots' blocks (light rows 256 to a block, heavy rows 4) over the
// launch list, written by the host from the counts it read and uploaded with
// the list (so `next` starts at 0 every step).  The workgroups take blocks
// off it with one atomic each until it is empty, so slots with uneven active
// lists share one grid instead of each getting the largest one's.  A block is
// k_discharge_list's block b of its slot; each wave adds its work counters
// to its block's slot.
__global__ __launch_bounds__(256) void k_discharge_queue(const SlotView* __restrict__ slots,
                                                         const int* __restrict__ list, int nlist,
                                                         int32_t* __restrict__ dq) {
    __shared__ int s_item, s_j;
    const int total = dq[1];
    TileCtx tc;
    for (;;) {
        if (threadIdx.x == 0) {
            const int item = atomicAdd(&dq[0], 1);
            int j = 0;
            if (item < total) {  // the last list entry whose prefix is <= item
                int lo = 0, hi = nlist - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (dq[2 + mid] <= item) lo = mid;
                    else hi = mid - 1;
                }
                j = lo;
            }
            s_item = item;
            s_j = j;
        }
        __syncthreads();
        const int item = s_item, j = s_j;
        __syncthreads();  // s_item / s_j are rewritten by the next take
        if (item >= total) break;  // uniform per block
        const GSlot S = load_slot_at(slots, list[j]);
        const int nl = S.ctr->nlight, nh = S.ctr->nheavy;
        const int lb = (nl + 255) / 256;
        const int b = item - dq[2 + j];
        Work wk;
        if (b < lb) {
            const int i = b * 256 + threadIdx.x;
            if (i < nl) lane_discharge(S, tc, S.fa[i], S.K, wk);
        } else {
            const int k = (b - lb) * 4 + (int)(threadIdx.x / WAVE);
            if (k < nh) {
                const int v = S.fa[S.g.n - 1 - k];
                if (S.g.row_ptr[v + 1] - S.g.row_ptr[v] <= kGiantDeg) wave_discharge(S, tc, v, S.Kh, wk);
            }
        }
        uint64_t x[4] = {wk.edges, wk.verts, wk.pushes, wk.relabels};
        for (int q = 0; q < 4; ++q) {
            for (int o = WAVE / 2; o > 0; o >>= 1) x[q] += __shfl_xor(x[q], o, WAVE);
            if (lane_id() == 0 && x[q]) addu64(&S.ctr->wk[WF_DISCHARGE][q], x[q]);
        }
    }
}

void launch_giant(hipStream_t st, const SlotView* slots, const int* list, int nlist, int giants);

void launch_discharge_queue(hipStream_t st, const SlotView* slots, const int* list, int nlist, int32_t* dq,
                            int workgroups, int giants) {
    hipLaunchKernelGGL(k_discharge_queue, dim3(workgroups > 0 ? workgroups : 1), dim3(256), 0, st, slots, list,
                       nlist, dq);
    if (giants) launch_giant(st, slots, list, nlist, giants);
}

// Giant rows (longer than kGiantDeg), spread over the whole grid in tasks of
// kGiantChunk entries.  k_giant_prep turns the slot's giant list into a
// prefix of task counts (one workgroup per slot); a workgroup maps a task to
// (row, chunk) by binary search.
// 16-byte task record of the giant list (a vector type keeps the load one
// instruction): row in the list, vertex, first entry, end entry.
typedef int TaskRec __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(1024) void k_giant_prep(const SlotView* __restrict__ slots,
                                                     const int* __restrict__ list) {
    const GSlot S = load_slot(slots, list);
    __shared__ int wsum[1024 / WAVE];
    __shared__ int carry;
    const int ng = S.ctr->ngiant;
    GAS int32_t* pre = S.giant + 2 * S.gcap;
    GAS TaskRec* rec = (GAS TaskRec*)(S.giant + giant_task_offset(S.gcap));
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < ng; base += blockDim.x) {
        const int g = base + (int)threadIdx.x;
        int c = 0, v = 0, rb = 0, re = 0;
        if (g < ng) {
            v = S.giant[g];
            rb = S.g.row_ptr[v];
            re = S.g.row_ptr[v + 1];
            c = (re - rb + kGiantChunk - 1) / kGiantChunk;
            S.gacc[g] = 0;
        }
        // block exclusive scan of c
        int x = c;
        const int lane = lane_id();
        for (int o = 1; o < WAVE; o <<= 1) {
            const int y = __shfl_up(x, o, WAVE);
            if (lane >= o) x += y;
        }
        const int w = threadIdx.x / WAVE;
        if (lane == WAVE - 1) wsum[w] = x;
        __syncthreads();
        int before = carry;
        int all = 0;
        for (int j = 0; j < (int)(blockDim.x / WAVE); ++j) {
            if (j < w) before += wsum[j];
            all += wsum[j];
        }
        if (g < ng) {
            const int p = before + x - c;
            pre[g] = p;
            for (int j = 0; j < c; ++j) {  // one record per task: row, vertex, entry range
                const int e0 = rb + j * kGiantChunk;
                rec[p + j] = TaskRec{g, v, e0, min(re, e0 + kGiantChunk)};
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) carry += all;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        pre[ng] = carry;
        S.ctr->ngchunks = carry;
    }
}

// Task -> (row k in the giant list, vertex v, entries [e0, e1)), the record
// k_giant_prep wrote: every thread reads the same 16 bytes, so no barrier is
// needed.
struct GTask { int k, v, e0, e1; };
__device__ __forceinline__ GTask giant_task(const GSlot& S, int task) {
    const TaskRec r = ((const GAS TaskRec*)(S.giant + giant_task_offset(S.gcap)))[task];
    return GTask{r.x, r.y, r.z, r.w};
}

// k_giant_push: one pass over each chunk of a giant row, in rounds of one
// entry per thread.  A round claims min(excess, residual into its lower
// entries) from the row's vertex with a compare-and-swap that never
// overdraws and pushes it to those entries in prefix-sum order; the chunk
// stops at the round that finds less excess than it could take.  The
// lowest height over the entries that keep residual goes to the row's
// minimum slot, and gacc counts the chunks that reported; k_giant_relabel
// raises a vertex that still has excess to that minimum + 1 (n when it is
// n - 1 or more) once every chunk of its row reported.  Only the row's
// chunks take excess or residual out of the row, so the lower-bound reads of
// the single-owner scheme still hold.
template <typename P>
__device__ __forceinline__ bool cas64(P p, int64_t& expect, int64_t want) {
    return __hip_atomic_compare_exchange_strong(p, &expect, want, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
}

// phase 0: every chunk; 1: each row's lead chunk only (the one holding its
// first forward entry, where a hub's out-entries with residual start);
// 2: the other chunks, after the lead chunks have run.  A hub's excess
// (one preflow entry's capacity, or what its neighbours pushed) usually
// fits into the lower entries of its lead chunk: the other chunks then find
// the row dry and read none of their entries.
template <int PHASE, bool ROUNDS>
__global__ __launch_bounds__(256) void k_giant_push(const SlotView* __restrict__ slots,
                                                    const int* __restrict__ list) {
    constexpr int phase = PHASE;
    const GSlot S = load_slot(slots, list);
    __shared__ int64_t lscr[256 / WAVE];
    __shared__ int iscr[256 / WAVE];
    __shared__ int64_t bcast;
    __shared__ uint64_t red[256 / WAVE];
    __shared__ int sh[4];
    constexpr int PER = kGiantChunk / 256;
    const int T = S.ctr->ngchunks;
    if ((int)blockIdx.x >= T) return;  // uniform per block
    const int n = S.g.n;
    Work wk;
    for (int task = blockIdx.x; task < T; task += gridDim.x) {
        const GTask gt = giant_task(S, task);
        const int k = gt.k, v = gt.v;
        const int e0 = gt.e0, e1 = gt.e1;
        if (phase != 0) {
            const int fs = S.g.fstart ? S.g.fstart[v] : S.g.row_ptr[v];
            const bool lead = fs >= e0 && fs < e1;
            if (lead != (phase == 1)) continue;  // uniform per block
        }
        // the row's excess and height first: a chunk that finds its row dry
        // (the row's other chunks took the excess) reads none of its entries
        if (threadIdx.x == 0) {
            bcast = ld64(&S.excess[v]);
            sh[3] = S.h[v];
        }
        __syncthreads();
        const int64_t ex0 = bcast;
        const int hv = sh[3];
        __syncthreads();
        // dry, cut off or above this step's band: no report, no relabel
        if (ex0 <= 0 || hv >= n || hv > S.disc_cap) continue;
        int m = INT_MAX - 1;
        if constexpr (ROUNDS) {
            // the chunk in rounds of blockDim.x entries, each claiming from
            // the row what its lower entries can take: once a round claims
            // less than that, the row is dry and the rest of the chunk is
            // neither read nor gathered (a hub's excess usually fits into its
            // first lower entries); a chunk that read every round reports its
            // minimum.  Fewer bytes per discharge, but the excess lands on
            // the first lower entries of a chunk instead of all of them, which
            // costs iterations: engines of many slots (bandwidth) use it, a
            // one-slot solver (latency) does not
            bool dry = false;
            int rounds = 0;
            for (int j = 0; j < PER; ++j) {
                const int e = e0 + j * (int)blockDim.x + (int)threadIdx.x;
                int64_t r = 0;
                int cw = 0, hw = INT_MAX;
                if (e < e1) {
                    r = ld64(&S.res[e]);
                    cw = S.g.col[e];
                }
                int64_t a = 0;
                if (r > 0) {
                    hw = S.h[cw];
                    if (hw < hv) a = r;
                } else {
                    r = 0;
                }
                int64_t tot = 0;
                const int64_t incl = block_incl_scan_i64(a, lscr, &tot);
                ++rounds;
                int64_t got = 0;
                if (tot > 0) {  // block-uniform
                    if (threadIdx.x == 0) {
                        int64_t cur = ld64(&S.excess[v]);
                        int64_t take = 0;
                        while (cur > 0) {
                            take = cur < tot ? cur : tot;
                            if (cas64(&S.excess[v], cur, cur - take)) break;
                        }
                        bcast = cur > 0 ? take : 0;
                    }
                    __syncthreads();
                    got = bcast;
                    __syncthreads();
                }
                const int64_t rem = got - (incl - a);
                if (a > 0 && rem > 0) {
                    const int64_t d = rem < a ? rem : a;
                    move_res(S, e, d);
                    add64(&S.excess[cw], d);
                    touch(S, cw);
                    r -= d;
                    wk.pushes++;
                }
                if (r > 0) m = min(m, hw);
                if (got < tot) {  // block-uniform: the row's excess ran out in this round
                    dry = true;
                    break;
                }
                if (e0 + (j + 1) * (int)blockDim.x >= e1) break;
            }
            if (threadIdx.x == 0) wk.edges += (uint64_t)min(e1 - e0, rounds * (int)blockDim.x);
            if (dry) continue;  // block-uniform
        } else {
            // the whole chunk into registers, one claim, pushes in prefix order
            int64_t r[PER];
            int hw[PER], cw[PER];
#pragma unroll
            for (int j = 0; j < PER; ++j) {
                const int e = e0 + j * (int)blockDim.x + (int)threadIdx.x;
                r[j] = 0;
                cw[j] = 0;
                if (e < e1) {
                    r[j] = ld64(&S.res[e]);
                    cw[j] = S.g.col[e];
                }
            }
            int64_t tsum = 0;
#pragma unroll
            for (int j = 0; j < PER; ++j) {
                hw[j] = INT_MAX;
                if (r[j] > 0) {
                    hw[j] = S.h[cw[j]];
                    if (hw[j] < hv) tsum += r[j];
                } else {
                    r[j] = 0;
                }
            }
            int64_t tot = 0;
            const int64_t incl = block_incl_scan_i64(tsum, lscr, &tot);
            if (threadIdx.x == 0) wk.edges += (uint64_t)(e1 - e0);
            int64_t got = 0;
            if (tot > 0) {  // block-uniform
                if (threadIdx.x == 0) {
                    int64_t cur = ld64(&S.excess[v]);
                    int64_t a = 0;
                    while (cur > 0) {
                        a = cur < tot ? cur : tot;
                        if (cas64(&S.excess[v], cur, cur - a)) break;
                    }
                    bcast = cur > 0 ? a : 0;
                }
                __syncthreads();
                got = bcast;
                __syncthreads();
            }
            int64_t rem = got - (incl - tsum);
#pragma unroll
            for (int j = 0; j < PER; ++j) {
                if (r[j] <= 0) continue;
                if (hw[j] < hv && rem > 0) {
                    const int e = e0 + j * (int)blockDim.x + (int)threadIdx.x;
                    const int64_t d = rem < r[j] ? rem : r[j];
                    move_res(S, e, d);
                    add64(&S.excess[cw[j]], d);
                    touch(S, cw[j]);
                    rem -= d;
                    r[j] -= d;
                    wk.pushes++;
                }
                if (r[j] > 0) m = min(m, hw[j]);
            }
        }
        m = block_min_i32(m, iscr);
        if (threadIdx.x == 0) {
            min32(&S.giant[S.gcap + k], m);
            add64(&S.gacc[k], 1);
        }
    }
    work_flush(S, wk, red, WF_DISCHARGE);
}

__global__ __launch_bounds__(256) void k_giant_relabel(const SlotView* __restrict__ slots,
                                                       const int* __restrict__ list) {
    const GSlot S = load_slot(slots, list);
    __shared__ uint64_t red[256 / WAVE];
    const int ng = S.ctr->ngiant;
    const int n = S.g.n;
    Work wk;
    for (int k = threadIdx.x; k < ng; k += blockDim.x) {
        const int v = S.giant[k];
        const int rb = S.g.row_ptr[v], re = S.g.row_ptr[v + 1];
        const int nch = (re - rb + kGiantChunk - 1) / kGiantChunk;
        if (S.gacc[k] != (int64_t)nch) continue;  // a chunk found the vertex dry
        if (ld64(&S.excess[v]) <= 0) continue;
        const int hv = S.h[v];
        const int m = S.giant[S.gcap + k];
        const int nh = m >= n - 1 ? n : m + 1;
        if (nh > hv) {
            S.h[v] = nh;
            wk.relabels++;
        }
    }
    work_flush(S, wk, red, WF_DISCHARGE);
}

// Giant rows of a top-down level in kGiantChunk-entry tasks over the grid
// (the task prefix comes from k_giant_prep over the level's giant list).
template <int MODE>
__global__ __launch_bounds__(256) void k_bfs_td_giant(const SlotView* __restrict__ slots,
                                                      const int* __restrict__ list) {
    const GSlot S = load_slot(slots, list);
    __shared__ Stage<2048> st;
    __shared__ uint64_t red[256 / WAVE];
    const int T = S.ctr->ngchunks;
    if ((int)blockIdx.x >= T) return;  // uniform per block
    const int L = S.ctr->level;
    GAS int32_t* lvl = bfs_lvl<MODE>(S);
    GAS int32_t* out = (L & 1) ? S.fa : S.fb;
    GAS int32_t* ocnt = &S.ctr->fsize[(L + 1) & 3];
    GAS int64_t* odeg = &S.ctr->fedges[(L + 1) & 3];
    stage_init(st);
    Work wk;
    for (int task = blockIdx.x; task < T; task += gridDim.x) {
        const GTask gt = giant_task(S, task);
        const int e0 = gt.e0, e1 = gt.e1;
        if (threadIdx.x == 0) wk.edges += (uint64_t)(e1 - e0);
        td_entries<MODE, 4>(S, e0 + (int)threadIdx.x, e1, (int)blockDim.x, L, lvl, [&](int w) {
            stage_push(st, w, S.g.row_ptr[w + 1] - S.g.row_ptr[w], out, ocnt, odeg);
            wk.relabels++;
        }, td_fresh<MODE>(S, gt.v));
    }
    stage_flush(st, out, ocnt, odeg);
    const uint64_t nv = block_sum<uint64_t>(wk.relabels, red);
    if (threadIdx.x == 0 && nv) add32(&S.ctr->visited, (int32_t)nv);
    work_flush(S, wk, red, WF_BFS);
}

void launch_bfs_td(hipStream_t st, const SlotView* slots, const int* list, int nlist, int gridx,
                   int mode) {
    // gridx: up to frontier / 256 workgroups (k_bfs_td is grid-stride);
    // medium rows get up to 2048 workgroups of 4 waves, giant rows 1024
    // workgroups of chunk tasks
    const int gh = per_slot_grid(gridx * 64 < 2048 ? gridx * 64 : 2048, nlist);
    const int gg = per_slot_grid(1024, nlist);
    if (mode == BFS_GR) {
        hipLaunchKernelGGL(k_bfs_td<BFS_GR>, dim3(gridx, nlist), dim3(256), 0, st, slots, list);
        hipLaunchKernelGGL(k_giant_prep, dim3(1, nlist), dim3(1024), 0, st, slots, list);
        hipLaunchKernelGGL(k_bfs_td_giant<BFS_GR>, dim3(gg, nlist), dim3(256), 0, st, slots, list);
        hipLaunchKernelGGL(k_bfs_td_heavy<BFS_GR>, dim3(gh, nlist), dim3(256), 0, st, slots, list);
    } else if (mode == BFS_FWD) {
        hipLaunchKernelGGL(k_bfs_td<BFS_FWD>, dim3(gridx, nlist), dim3(256), 0, st, slots, list);
        hipLaunchKernelGGL(k_giant_prep, dim3(1, nlist), dim3(1024), 0, st, slots, list);
        hipLaunchKernelGGL(k_bfs_td_giant<BFS_FWD>, dim3(gg, nlist), dim3(256), 0, st, slots, list);
        hipLaunchKernelGGL(k_bfs_td_heavy<BFS_FWD>, dim3(gh, nlist), dim3(256), 0, st, slots, list);
    } else {
        hipLaunchKernelGGL(k_bfs_td<BFS_UND>, dim3(gridx, nlist), dim3(256), 0, st, slots, list);
        hipLaunchKernelGGL(k_giant_prep, dim3(1, nlist), dim3(1024), 0, st, slots, list);
        hipLaunchKernelGGL(k_bfs_td_giant<BFS_UND>, dim3(gg, nlist), dim3(256), 0, st, slots, list);
        hipLaunchKernelGGL(k_bfs_td_heavy<BFS_UND>, dim3(gh, nlist), dim3(256), 0, st, slots, list);
    }
    hipLaunchKernelGGL(k_bfs_advance, dim3(1, nlist), dim3(64), 0, st, slots, list, 0);
}

// giants: 1 = whole chunks, 2 = chunks in rounds (k_giant_push)
void launch_giant(hipStream_t st, const SlotView* slots, const int* list, int nlist, int giants) {
    hipLaunchKernelGGL(k_giant_prep, dim3(1, nlist), dim3(1024), 0, st, slots, list);
    const dim3 grid(per_slot_grid(512, nlist), nlist);
    if (giants == 2) {
        hipLaunchKernelGGL((k_giant_push<1, true>), grid, dim3(256), 0, st, slots, list);
        hipLaunchKernelGGL((k_giant_push<2, true>), grid, dim3(256), 0, st, slots, list);
    } else {
        hipLaunchKernelGGL((k_giant_push<1, false>), grid, dim3(256), 0, st, slots, list);
        hipLaunchKernelGGL((k_giant_push<2, false>), grid, dim3(256), 0, st, slots, list);
    }
    hipLaunchKernelGGL(k_giant_relabel, dim3(1, nlist), dim3(256), 0, st, slots, list);
}

void launch_discharge_list(hipStream_t st, const SlotView* slots, const int* list, int nlist,
                           int gridx, int giants) {
    hipLaunchKernelGGL(k_discharge_list, dim3(gridx, nlist), dim3(256), 0, st, slots, list);
    if (giants) launch_giant(st, slots, list, nlist, giants);
}

// Region (tile) discharge (BC:432-515, one block per region): the tile's
// heights and pending excess live in LDS for the whole launch; pushes to
// vertices of the same tile are LDS atomics, other pushes global atomics.
// `rounds` rebuilds of the tile's active list, one cycle per vertex each.
template <int NT>
__global__ __launch_bounds__(NT) void k_discharge_tile(const SlotView* __restrict__ slots,
                                                       const int* __restrict__ list, int hot_only,
                                                       int rounds, int split) {
    const GSlot S = load_slot(slots, list);
    // split: the workgroup takes part `part` of `parts` of its tile's vertex
    // range, as k_tile_alloc dealt them out by excess score; else one
    // workgroup per tile
    int tile = blockIdx.x, part = 0, parts = 1;
    if (split) {
        if ((int)blockIdx.x >= S.tsplit[0]) return;
        tile = S.tsplit[1 + 3 * blockIdx.x];
        part = S.tsplit[2 + 3 * blockIdx.x];
        parts = S.tsplit[3 + 3 * blockIdx.x];
    }
    if (tile >= S.g.ntiles) return;
    if (hot_only && S.tile_hot[tile] == 0) return;
    // no active vertex at the last count: nothing to do (vertices pushed into
    // since are counted next cycle)
    if (S.tile_cnt[tile] == 0) return;
    __shared__ int32_t lh[kTileMax];
    __shared__ int64_t ldx[kTileMax];
    __shared__ int32_t alist[kTileMax];
    __shared__ int nlight, nheavy;
    __shared__ uint64_t red[NT / WAVE];
    const int ta = S.g.tile_start[tile], tsz = S.g.tile_start[tile + 1] - ta;
    const int a = ta + (int)((int64_t)tsz * part / parts), b = ta + (int)((int64_t)tsz * (part + 1) / parts);
    const int size = b - a;  // <= kTileMax by construction of tile_start
    const int n = S.g.n;
    for (int i = threadIdx.x; i < size; i += blockDim.x) { lh[i] = S.h[a + i]; ldx[i] = 0; }
    TileCtx tc;
    tc.a = a; tc.b = b; tc.lh = lh; tc.ldx = ldx;
    Work wk;
    for (int r = 0; r < rounds; ++r) {
        if (threadIdx.x == 0) { nlight = 0; nheavy = 0; }
        __syncthreads();
        for (int i = threadIdx.x; i < size; i += blockDim.x) {
            const int v = a + i;
            if (v == S.s && !S.inj) continue;
            if (v == S.t && S.phase == 0) continue;
            if (lh[i] >= n || lh[i] > S.disc_cap) continue;
            const int64_t ex = ld64(&S.excess[v]) + ldx[i];
            if (ex <= 0) continue;
            const int deg = S.g.row_ptr[v + 1] - S.g.row_ptr[v];
            if (deg <= kLightDeg) alist[atomicAdd(&nlight, 1)] = i;
            else if (deg <= kGiantDeg) alist[size - 1 - atomicAdd(&nheavy, 1)] = i;
            // giant rows: k_giant_push / k_giant_relabel after the tile pass
        }
        if (threadIdx.x == 0) wk.verts += (uint64_t)size;
        __syncthreads();
        const int nl = nlight, nh = nheavy;
        if (nl + nh == 0) break;
        for (int k = threadIdx.x; k < nl; k += blockDim.x) lane_discharge(S, tc, a + alist[k], 1, wk);
        for (int k = threadIdx.x / WAVE; k < nh; k += blockDim.x / WAVE)
            wave_discharge(S, tc, a + alist[size - 1 - k], 1, wk);
        __syncthreads();
    }
    __syncthreads();
    for (int i = threadIdx.x; i < size; i += blockDim.x) {
        S.h[a + i] = lh[i];
        if (ldx[i] != 0) {
            add64(&S.excess[a + i], ldx[i]);
            if (ldx[i] > 0) touch(S, a + i);
        }
    }
    work_flush(S, wk, red, WF_DISCHARGE);
}

// Workgroups per tile by excess score (the predecessor's region blocks,
// Region_Aware_PR_without_Pruning.cu:173-215 and :247-338): a tile with
// active vertices (hot ones only, for the hot sweep) scores its active
// vertices' positive excess plus (its active count / n) x the slot's
// positive excess, and gets 1 + floor((kmax - 1) x score / top score)
// workgroups, each a contiguous part of the tile's vertex range.  One
// workgroup per slot writes the (tile, part, parts) records in tile order.
__global__ __launch_bounds__(1024) void k_tile_alloc(const SlotView* __restrict__ slots,
                                                     const int* __restrict__ list, int kmax, int hot_only) {
    const GSlot S = load_slot(slots, list);
    __shared__ uint64_t redu[1024 / WAVE];
    __shared__ int wsum[1024 / WAVE];
    __shared__ int carry;
    const int T = S.g.ntiles;
    const int n = S.g.n;
    const int64_t pos = S.ctr->pos_excess > 0 ? S.ctr->pos_excess : 0;
    auto score = [&](int i) -> int64_t {
        const int c = S.tile_cnt[i];
        if (c <= 0 || (hot_only && S.tile_hot[i] == 0)) return -1;
        const double ratio = (double)c / (double)(n > 0 ? n : 1);
        return S.tile_ex[i] + (int64_t)(ratio * (double)pos);
    };
    int64_t mx = 0;
    for (int i = threadIdx.x; i < T; i += blockDim.x) mx = max(mx, score(i));
    // block max through the unsigned sum's reduction slots
    {
        const int lane = lane_id();
        for (int o = WAVE / 2; o > 0; o >>= 1) mx = max(mx, (int64_t)__shfl_down((long long)mx, o, WAVE));
        if (lane == 0) redu[threadIdx.x / WAVE] = (uint64_t)mx;
        __syncthreads();
        if (threadIdx.x == 0) {
            int64_t m = 0;
            for (int w = 0; w < (int)(blockDim.x / WAVE); ++w) m = max(m, (int64_t)redu[w]);
            redu[0] = (uint64_t)m;
            carry = 0;
        }
        __syncthreads();
        mx = (int64_t)redu[0];
    }
    for (int base = 0; base < T; base += blockDim.x) {
        const int i = base + (int)threadIdx.x;
        int b = 0;
        if (i < T) {
            const int64_t sc = score(i);
            if (sc >= 0) {
                const double f = mx > 0 ? (double)sc / (double)mx : 0.0;
                b = 1 + (int)((double)(kmax - 1) * f);
                b = min(b, kmax);
            }
        }
        // block exclusive scan of b
        int x = b;
        const int lane = lane_id();
        for (int o = 1; o < WAVE; o <<= 1) {
            const int y = __shfl_up(x, o, WAVE);
            if (lane >= o) x += y;
        }
        const int w = threadIdx.x / WAVE;
        if (lane == WAVE - 1) wsum[w] = x;
        __syncthreads();
        int before = carry, all = 0;
        for (int j = 0; j < (int)(blockDim.x / WAVE); ++j) {
            if (j < w) before += wsum[j];
            all += wsum[j];
        }
        const int off = before + x - b;
        for (int j = 0; j < b; ++j) {
            S.tsplit[1 + 3 * (off + j)] = i;
            S.tsplit[2 + 3 * (off + j)] = j;
            S.tsplit[3 + 3 * (off + j)] = b;
        }
        __syncthreads();
        if (threadIdx.x == 0) carry += all;
        __syncthreads();
    }
    if (threadIdx.x == 0) S.tsplit[0] = carry;
}

// Labels of a saturated terminal cut, for the certificate (T = {h < n}):
// side 0, the cut around t: T = {t} (h = 0 at t, n elsewhere); side 1, the
// cut around s: T = V \ {s} (h = n at s, 0 elsewhere).
__global__ __launch_bounds__(256) void k_cut_labels(const SlotView* __restrict__ slots,
                                                    const int* __restrict__ list, int side) {
    const GSlot S = load_slot(slots, list);
    const int n = S.g.n;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x)
        S.h[v] = side == 0 ? (v == S.t ? 0 : n) : (v == S.s ? n : 0);
}

void launch_cut_labels(hipStream_t st, const SlotView* slots, const int* list, int nlist, int gridx, int side) {
    hipLaunchKernelGGL(k_cut_labels, dim3(gridx, nlist), dim3(256), 0, st, slots, list, side);
}

__global__ void k_trace_marker() {}

void launch_trace_marker(hipStream_t st) { hipLaunchKernelGGL(k_trace_marker, dim3(1), dim3(64), 0, st); }

void launch_discharge_tile(hipStream_t st, const SlotView* slots, const int* list, int nlist,
                           int gridx, int hot_only, int rounds, int giants, int threads, int split) {
    int sp = 0;
    if (split > 1) {
        hipLaunchKernelGGL(k_tile_alloc, dim3(1, nlist), dim3(1024), 0, st, slots, list, std::min(split, kSplitMax),
                           hot_only);
        gridx *= std::min(split, kSplitMax);
        sp = 1;
    }
    if (threads <= 256)
        hipLaunchKernelGGL(k_discharge_tile<256>, dim3(gridx, nlist), dim3(256), 0, st, slots, list, hot_only, rounds, sp);
    else if (threads <= 512)
        hipLaunchKernelGGL(k_discharge_tile<512>, dim3(gridx, nlist), dim3(512), 0, st, slots, list, hot_only, rounds, sp);
    else
        hipLaunchKernelGGL(k_discharge_tile<1024>, dim3(gridx, nlist), dim3(1024), 0, st, slots, list, hot_only,
                           rounds, sp);
    if (!hot_only && giants) launch_giant(st, slots, list, nlist, giants);
}

// ------------------------------------------------------------ path push (a17)
// BC:773-955, made owner-exclusive: a wave walks from an active vertex down
// admissible entries (h[w] == h[cur] - 1), holding lock[cur] while it takes
// excess out of cur, and hands the lock over hop by hop.  A dead end
// relabels cur to lowest residual neighbour + 1 (n when none is below n-1).

__global__ __launch_bounds__(256) void k_path_push(const SlotView* __restrict__ slots,
                                                   const int* __restrict__ list, int max_hops) {
    const GSlot S = load_slot(slots, list);
    __shared__ uint64_t red[256 / WAVE];
    const int nl = S.ctr->nlight, nh = S.ctr->nheavy;
    const int k = blockIdx.x * 4 + (int)(threadIdx.x / WAVE);
    const int lane = lane_id();
    const int n = S.g.n;
    Work wk;
    if (k < nl + nh) {
        const int v = (k < nl) ? S.fa[k] : S.fa[n - 1 - (k - nl)];
        const int tag = v;
        int got = 0;
        if (lane == 0) got = cas32(&S.lock[v], -1, tag) ? 1 : 0;
        got = __shfl(got, 0, WAVE);
        if (got) {
            int cur = v;
            int hknown = -1;  // height this wave last wrote for cur (lane 0 stores it)
            for (int hop = 0; hop < max_hops; ++hop) {
                int64_t ex = 0;
                if (lane == 0) ex = ld64(&S.excess[cur]);
                ex = __shfl(ex, 0, WAVE);
                if (ex <= 0) break;
                if (cur == S.t && S.phase == 0) break;
                int hc = hknown;
                if (hc < 0) hc = __shfl(S.h[cur], 0, WAVE);
                if (hc >= n) break;
                const int rb = S.g.row_ptr[cur], re = S.g.row_ptr[cur + 1];
                if (lane == 0) wk.edges += (uint64_t)(re - rb);
                int best = -1;
                for (int base = rb; base < re; base += WAVE) {
                    const int e = base + lane;
                    bool ok = false;
                    if (e < re) ok = (ld64(&S.res[e]) > 0) && (S.h[S.g.col[e]] == hc - 1);
                    const unsigned long long m = __ballot(ok);
                    if (m) { best = base + (int)__ffsll((long long)m) - 1; break; }
                }
                if (best >= 0) {
                    const int w = S.g.col[best];
                    int moved = 0;
                    if (lane == 0) {
                        const int64_t r = ld64(&S.res[best]);
                        const int64_t d = ex < r ? ex : r;
                        if (d > 0) {
                            move_res(S, best, d);
                            add64(&S.excess[w], d);
                            add64(&S.excess[cur], -d);
                            wk.pushes++;
                        }
                        // hand the lock to w unless w is a terminal or taken
                        if (w != S.s && !(w == S.t && S.phase == 0) && cas32(&S.lock[w], -1, tag)) {
                            __hip_atomic_store(&S.lock[cur], -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            moved = 1;
                        }
                    }
                    moved = __shfl(moved, 0, WAVE);
                    if (!moved) break;
                    cur = w;
                    hknown = -1;
                } else {
                    int hmin = INT_MAX;
                    for (int e = rb + lane; e < re; e += WAVE)
                        if (ld64(&S.res[e]) > 0) hmin = min(hmin, S.h[S.g.col[e]]);
                    hmin = wave_min_i32(hmin);
                    const int hnew = (hmin == INT_MAX || hmin >= n - 1) ? n : hmin + 1;
                    if (lane == 0) {
                        S.h[cur] = hnew;
                        wk.relabels++;
                    }
                    hknown = hnew;
                    if (hnew >= n) break;
                }
            }
            if (lane == 0) __hip_atomic_store(&S.lock[cur], -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    work_flush(S, wk, red, WF_DISCHARGE);
}

void launch_path_push(hipStream_t st, const SlotView* slots, const int* list, int nlist, int gridx,
                      int max_hops) {
    hipLaunchKernelGGL(k_path_push, dim3(gridx, nlist), dim3(256), 0, st, slots, list, max_hops);
}

// ------------------------------------------------------------- hot tiles (a19)
// BC:1899-1946: the top-k regions by active count (ties by lower index, only
// regions with at least one active vertex).  One workgroup per slot: each
// tile's rank is the number of tiles that beat it.

__global__ __launch_bounds__(1024) void k_hot_select(const SlotView* __restrict__ slots,
                                                     const int* __restrict__ list, int k) {
    constexpr int kLds = 8192;  // tile counts kept in LDS up to this many tiles
    const GSlot S = load_slot(slots, list);
    const int T = S.g.ntiles;
    __shared__ int lc[kLds];
    __shared__ int cnt_gt, maxc;
    __shared__ int lo_s, hi_s;
    const bool in_lds = T <= kLds;
    auto tc = [&](int i) { return in_lds ? lc[i] : S.tile_cnt[i]; };
    if (threadIdx.x == 0) maxc = 0;
    __syncthreads();
    int m = 0;
    for (int i = threadIdx.x; i < T; i += blockDim.x) {
        const int c = S.tile_cnt[i];
        if (in_lds) lc[i] = c;
        m = max(m, c);
    }
    m = wave_max_i32(m);
    if (lane_id() == 0) atomicMax(&maxc, m);
    __syncthreads();
    if (threadIdx.x == 0) { lo_s = 0; hi_s = maxc; }
    __syncthreads();
    // Threshold search: smallest x with |{cnt > x}| < k (at most the largest
    // count) by bisection -> tiles with cnt > x are hot, ties at cnt == x
    // filled by lowest index.
    while (true) {
        const int lo = lo_s, hi = hi_s;
        if (lo >= hi) break;
        const int mid = lo + (hi - lo) / 2;
        if (threadIdx.x == 0) cnt_gt = 0;
        __syncthreads();
        int c = 0;
        for (int i = threadIdx.x; i < T; i += blockDim.x) c += (tc(i) > mid) ? 1 : 0;
        c = wave_sum_i32(c);
        if (lane_id() == 0 && c) atomicAdd(&cnt_gt, c);
        __syncthreads();
        if (threadIdx.x == 0) {
            if (cnt_gt < k) hi_s = mid; else lo_s = mid + 1;
        }
        __syncthreads();
    }
    const int x = lo_s;
    __shared__ int above, ties;
    if (threadIdx.x == 0) { above = 0; ties = 0; }
    __syncthreads();
    int c = 0;
    for (int i = threadIdx.x; i < T; i += blockDim.x) c += (tc(i) > x) ? 1 : 0;
    c = wave_sum_i32(c);
    if (lane_id() == 0 && c) atomicAdd(&above, c);
    __syncthreads();
    const int room = k - above;
    // Ties at cnt == x (x >= 1): lowest indices first; serial over chunks keeps order.
    for (int base = 0; base < T; base += blockDim.x) {
        const int i = base + threadIdx.x;
        int hot = 0;
        bool tie = false;
        if (i < T) {
            const int ci = tc(i);
            hot = (ci > x) ? 1 : 0;
            tie = (ci == x && x >= 1);
        }
        // rank ties within the chunk in index order
        __shared__ int tie_base;
        __shared__ int wave_ties[1024 / WAVE];
        const unsigned long long mk = __ballot(tie);
        const int w = threadIdx.x / WAVE;
        if (lane_id() == 0) wave_ties[w] = __popcll(mk);
        if (threadIdx.x == 0) tie_base = ties;
        __syncthreads();
        int before = tie_base;
        for (int j = 0; j < w; ++j) before += wave_ties[j];
        before += __popcll(mk & ((1ull << lane_id()) - 1ull));
        if (tie && before < room) hot = 1;
        if (i < T) S.tile_hot[i] = hot;
        __syncthreads();
        if (threadIdx.x == 0) {
            int tot = 0;
            for (int j = 0; j < (int)(blockDim.x / WAVE); ++j) tot += wave_ties[j];
            ties = tie_base + tot;
        }
        __syncthreads();
    }
}

void launch_hot_select(hipStream_t st, const SlotView* slots, const int* list, int nlist, int k) {
    hipLaunchKernelGGL(k_hot_select, dim3(1, nlist), dim3(1024), 0, st, slots, list, k);
}

// ---------------------------------------------------------------- RM (a20)
// BC:605-657: for v in a tile holding active vertices, with h(v) in the band
// and v not a terminal, every entry v->w with h(w) > h(v)+1 and excess(w) <= 0
// has its residual w->v pushed from w to v (w may go into deficit).  The
// condition is exclusive per entry pair, so each pair has one writer.

__global__ __launch_bounds__(256) void k_rm(const SlotView* __restrict__ slots,
                                            const int* __restrict__ list, int margin) {
    const GSlot S = load_slot(slots, list);
    __shared__ uint64_t red[256 / WAVE];
    const int tile = blockIdx.x;
    if (tile >= S.g.ntiles) return;
    if (S.tile_cnt[tile] == 0) return;
    const int n = S.g.n;
    const int hmin = S.ctr->hmin, hmax = S.ctr->hmax;
    int lo = 0, hi = n - 1;
    if (hmax >= 0 && hmin <= hmax) { lo = max(0, hmin - margin); hi = min(n - 1, hmax + margin); }
    const int a = S.g.tile_start[tile], b = S.g.tile_start[tile + 1];
    uint64_t fixed = 0;
    Work wk;
    for (int v = a + threadIdx.x; v < b; v += blockDim.x) {
        const int hv = S.h[v];
        if (hv < lo || hv > hi) continue;
        if (v == S.s || v == S.t) continue;
        const int rb = S.g.row_ptr[v], re = S.g.row_ptr[v + 1];
        wk.edges += (uint64_t)(re - rb);
        int64_t gained = 0;
        for (int e = rb; e < re; ++e) {
            const int w = S.g.col[e];
            if (S.h[w] <= hv + 1) continue;
            if (ld64(&S.excess[w]) > 0) continue;
            const int re_ = S.g.rev[e];
            const int64_t d = ld64(&S.res[re_]);
            if (d <= 0) continue;
            move_res(S, re_, d);
            if (S.recv) S.recv[w] = 1;  // w pushed without receiving
            add64(&S.excess[w], -d);
            gained += d;
            ++fixed;
            wk.pushes++;
        }
        if (gained) add64(&S.excess[v], gained);
    }
    const uint64_t bf = block_sum<uint64_t>(fixed, red);
    if (threadIdx.x == 0 && bf) add32(&S.ctr->fixed_edges, (int32_t)bf);
    work_flush(S, wk, red, WF_RM);
}

void launch_rm(hipStream_t st, const SlotView* slots, const int* list, int nlist, int gridx,
               int margin) {
    hipLaunchKernelGGL(k_rm, dim3(gridx, nlist), dim3(256), 0, st, slots, list, margin);
}

// ------------------------------------------------------------ certificate
// T = {v : h(v) < n} from the final exact BFS.  Violations counted: negative
// residual, pair sum != cap pair sum, stored excess != excess recomputed from
// residuals, deficit off s, excess inside T off t, t outside T, s inside T,
// a residual entry from outside T into T.  cut = sum of cap over entries
// from outside T into T.

// Every entry pair has exactly one entry with capacity (forward) and one
// without (reverse).  Three modes:
//  FULL (k_cert, k_cert_chunks, k_cert_fin): pair sums from the forward
//        side, res[e] + res[rev e] == cap[e] (a gather), and each stored
//        excess against its row's recomputed sum, exactly;
//  FP   (k_cert_fp): pair sums and stored excesses through one 64-bit
//        fingerprint of odd weights (w per pair, u per vertex); the cap
//        terms are graph constants (g.cert_fp, g.capsum), so an entry costs
//        res + rev, both coalesced.
//  SPARSE (k_cert_sp_pairs, k_cert_sp_rows): the flow certified is the one
//        the marked pairs hold; every other pair carries none (res = cap).
//        The same fingerprint over the marked pairs' entries only, with
//        res - cap in place of res, so an unmarked pair adds nothing; the
//        mark map is read 8 entries per load.  A valid flow of value
//        excess(t) beside a closed cut of that capacity proves the value
//        maximal whatever unmarked residuals hold.  Used where every pair
//        whose residual left its capacities has a marked entry (residual
//        reset from the capacities, then changed by pushes and capacity
//        updates only, both of which mark).
// Rows of T also check closedness and sum the cut from T's side: entry
// e = w -> u with u outside T closes pair u -> w.
enum CertMode : int { CERT_FP = 1, CERT_FULL = 2, CERT_SPARSE = 3 };

__device__ __forceinline__ uint64_t pair_weight(uint64_t pair, uint64_t seed) {
    uint64_t x = pair + seed + 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return (x ^ (x >> 31)) | 1ULL;
}

__device__ __forceinline__ void cert_entry(const GSlot& S, int e, bool inT, int n, int64_t& ex,
                                           uint64_t& bad, int64_t& cut) {
    const int64_t r = S.res[e];
    if (r < 0) ++bad;
    const int64_t c = S.g.cap[e];
    ex += r - c;
    if (c > 0) {  // forward entry; its reverse entry has capacity 0 by construction
        const int rv = S.g.rev[e];
        if (r + S.res[rv] != c) ++bad;
    }
    if (inT) {
        if (S.h[S.g.col[e]] >= n) {  // col[e] outside T: its pair entry leaves V\T into T
            const int rv = S.g.rev[e];
            cut += S.g.cap[rv];
            if (S.res[rv] > 0) ++bad;
        }
    }
}

__device__ __forceinline__ void cert_vertex(const GSlot& S, int v, int64_t ex, bool inT, uint64_t& bad) {
    // s holds the injected units beyond its row's flow
    if (ex != S.excess[v] - (v == S.s ? S.ctr->inject : 0)) ++bad;
    if (v != S.s && ex < 0) ++bad;
    if (inT && v != S.t && ex != 0) ++bad;
    if (v == S.t) S.ctr->ex_t = S.excess[v];
    if (v == S.t && !inT) ++bad;
    if (v == S.s && inT) ++bad;
}

// Light rows: one lane each; medium rows: one wave each; giant rows go to
// the slot's giant list, whose rows k_cert_chunks cuts into kGiantChunk-entry
// tasks over the whole grid (hub rows sit together at region starts, so they
// are spread over the grid instead of the block that met them).
__device__ __forceinline__ void cert_flush(const GSlot& S, int64_t cut, uint64_t bad, uint64_t edges,
                                           uint64_t fp, int64_t* red64, uint64_t* redu) {
    const int64_t bc = block_sum<int64_t>(cut, red64);
    const uint64_t bb = block_sum<uint64_t>(bad, redu);
    const uint64_t be = block_sum<uint64_t>(edges, redu);
    const uint64_t bf = block_sum<uint64_t>(fp, redu);  // wraps mod 2^64, order-free
    if (threadIdx.x == 0) {
        if (bf) addu64(&S.ctr->fp, bf);
        if (bc) add64(&S.ctr->cut, bc);
        if (bb) add32(&S.ctr->cert_bad, (int32_t)bb);
        if (be) addu64(&S.ctr->wk[WF_CERT][0], be);
    }
}

__global__ __launch_bounds__(256) void k_cert(const SlotView* __restrict__ slots,
                                              const int* __restrict__ list) {
    const GSlot S = load_slot(slots, list);
    __shared__ int64_t red64[256 / WAVE];
    __shared__ uint64_t redu[256 / WAVE];
    __shared__ int med[256];
    __shared__ int nmed;
    const int n = S.g.n;
    int64_t cut = 0;
    uint64_t bad = 0, edges = 0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n; base += stride) {
        if (threadIdx.x == 0) nmed = 0;
        __syncthreads();
        const int64_t vi = base + threadIdx.x;
        if (vi < n) {
            const int v = (int)vi;
            const int rb = S.g.row_ptr[v], re = S.g.row_ptr[v + 1];
            if (re - rb <= kLightDeg) {  // one lane
                const bool inT = S.h[v] < n;
                int64_t ex = 0;
                for (int e = rb; e < re; ++e) cert_entry(S, e, inT, n, ex, bad, cut);
                edges += (uint64_t)(re - rb);
                cert_vertex(S, v, ex, inT, bad);
            } else if (re - rb <= kGiantDeg) {  // one wave, below
                med[atomicAdd(&nmed, 1)] = v;
            } else {  // giant list: k_cert_chunks, k_cert_fin
                S.giant[add32(&S.ctr->ngiant, 1)] = v;
            }
        }
        __syncthreads();
        const int nm = nmed;
        for (int k = threadIdx.x / WAVE; k < nm; k += blockDim.x / WAVE) {
            const int v = med[k];
            const int rb = S.g.row_ptr[v], re = S.g.row_ptr[v + 1];
            const bool inT = S.h[v] < n;
            int64_t ex = 0;
            for (int e = rb + lane_id(); e < re; e += WAVE) cert_entry(S, e, inT, n, ex, bad, cut);
            ex = wave_sum_i64(ex);
            if (lane_id() == 0) {
                edges += (uint64_t)(re - rb);
                cert_vertex(S, v, ex, inT, bad);
            }
        }
        __syncthreads();
    }
    cert_flush(S, cut, bad, edges, 0, red64, redu);
}

// Fingerprint mode, entry-parallel, one kernel.  A workgroup takes 256
// consecutive vertices: their row starts, heights and vertex weights go to
// LDS and the entries of their rows are read coalesced, four per thread in
// flight, the row of an entry found by binary lifting over the LDS row
// starts.  Each entry adds (w(pair) + u(row)) * res to the fingerprint and
// each vertex adds -u(v) * (capsum(v) + stored excess(v)), u an odd hash of
// the vertex id under a second seed.  The u terms sum to zero exactly when
// every stored excess equals its row's residual sum minus capacity sum, so
// one 64-bit sum checks pair sums and stored excesses together (a single
// wrong term can never give zero, several cancel with chance about 2^-64);
// the vertex checks then read the stored excess.  A hub row is read by its
// vertex's workgroup like any other: at most ~10^5 coalesced entries.
__device__ __forceinline__ uint64_t vertex_weight(int v, uint64_t seed) {
    return pair_weight((uint64_t)(uint32_t)v, seed ^ 0x6a09e667f3bcc909ULL);
}

__global__ __launch_bounds__(256) void k_cert_fp(const SlotView* __restrict__ slots,
                                                 const int* __restrict__ list) {
    constexpr int U = 4;  // entries per thread in flight
    const GSlot S = load_slot(slots, list);
    __shared__ int64_t red64[256 / WAVE];
    __shared__ uint64_t redu[256 / WAVE];
    __shared__ int rp[257];
    __shared__ int hh[256];
    __shared__ uint64_t uw[256];
    const int n = S.g.n;
    const uint64_t seed = S.g.cert_seed;
    int64_t cut = 0;
    uint64_t bad = 0, edges = 0, fp = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) fp = 0ULL - *S.g.cert_fp;  // the pairs' capacity terms
    const int nblk = (n + 255) / 256;
    const int i = (int)threadIdx.x;
    for (int b = blockIdx.x; b < nblk; b += gridDim.x) {
        const int v0 = b * 256;
        const int nv = min(256, n - v0);
        __syncthreads();  // the previous block's LDS reads are done
        if (i < nv) {
            const int v = v0 + i;
            const int rb = S.g.row_ptr[v], re = S.g.row_ptr[v + 1];
            rp[i] = rb;
            if (i == nv - 1) rp[nv] = re;
            const int hv = S.h[v];
            hh[i] = hv;
            const uint64_t u = vertex_weight(v, seed);
            uw[i] = u;
            const bool inT = hv < n;
            const int64_t x = S.excess[v];
            fp -= u * (uint64_t)(S.g.capsum[v] + x - (v == S.s ? S.ctr->inject : 0));
            if (v != S.s && x < 0) ++bad;
            if (inT && v != S.t && x != 0) ++bad;
            if (v == S.t) {
                S.ctr->ex_t = x;
                if (!inT) ++bad;
            }
            if (v == S.s && inT) ++bad;
            edges += (uint64_t)(re - rb);
        }
        __syncthreads();
        const int E1 = rp[nv];
        int lo = 0;  // row of this thread's previous entry: rows only grow along e
        for (int eb = rp[0] + i; eb < E1; eb += U * 256) {
            int r[U], rv[U];
            int64_t x[U];
#pragma unroll
            for (int j = 0; j < U; ++j) {
                const int e = eb + j * 256;
                // last row with rp[row] <= e, within lo + 255: binary lifting
                // without branches (a probe past the last row reads the last
                // row, which is then the answer exactly when it starts at or
                // before e; row starts never decrease)
                int rr = lo;
#pragma unroll
                for (int step = 128; step > 0; step >>= 1) {
                    const int c = min(rr + step, nv - 1);
                    rr = rp[c] <= e ? c : rr;
                }
                r[j] = rr;
                x[j] = 0;
                rv[j] = 0;
                if (e < E1) {
                    x[j] = S.res[e];
                    rv[j] = S.g.rev[e];
                }
            }
            lo = r[U - 1];
#pragma unroll
            for (int j = 0; j < U; ++j) {
                const int e = eb + j * 256;
                if (e >= E1) break;
                if (x[j] < 0) ++bad;
                fp += (pair_weight((uint64_t)min(e, rv[j]), seed) + uw[r[j]]) * (uint64_t)x[j];
                if (hh[r[j]] < n && S.h[S.g.col[e]] >= n) {  // pair col[e] -> row enters T
                    cut += S.g.cap[rv[j]];
                    if (S.res[rv[j]] > 0) ++bad;
                }
            }
        }
    }
    cert_flush(S, cut, bad, edges, fp, red64, redu);
}

// Sparse mode, pair side: each marked pair once (from its lower marked
// entry), both entries' terms (w(pair) + u(row)) * (res - cap); the row of
// an entry is the column of its pair entry.  The marked entries come from
// the slab's list, or from the mark map when the list overflowed; the
// bytes read for that (4 per listed entry, or the map's E) count as
// 16-byte vertex-scan units in the work counters.
__device__ __forceinline__ void cert_sp_pair(const GSlot& S, int e, uint64_t seed, uint64_t& bad,
                                             uint64_t& fp, uint64_t& pairs) {
    const int r = S.g.rev[e];
    if (r < e && S.emark[r]) return;  // taken from r
    const int64_t xe = S.res[e], xr = S.res[r];
    const int64_t ce = S.g.cap[e], cr = S.g.cap[r];
    if (xe < 0) ++bad;
    if (xr < 0) ++bad;
    const uint64_t wp = pair_weight((uint64_t)min(e, r), seed);
    fp += (wp + vertex_weight(S.g.col[r], seed)) * (uint64_t)(xe - ce) +
          (wp + vertex_weight(S.g.col[e], seed)) * (uint64_t)(xr - cr);
    ++pairs;
}

__global__ __launch_bounds__(256) void k_cert_sp_pairs(const SlotView* __restrict__ slots,
                                                       const int* __restrict__ list) {
    const GSlot S = load_slot(slots, list);
    __shared__ int64_t red64[256 / WAVE];
    __shared__ uint64_t redu[256 / WAVE];
    const uint64_t seed = S.g.cert_seed;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t E = S.g.E;
    const int cnt = S.mlist[0];
    const bool listed = cnt <= S.mcap;
    uint64_t bad = 0, fp = 0, pairs = 0;
    if (listed) {
        for (int64_t k = i0; k < cnt; k += stride) cert_sp_pair(S, S.mlist[1 + k], seed, bad, fp, pairs);
    } else {
        const int64_t nw = E / 8;
        const GAS uint64_t* m8 = (const GAS uint64_t*)S.emark;  // slab arrays are 256-byte aligned
        for (int64_t w = i0; w < nw; w += stride) {
            uint64_t m = m8[w];
            while (m) {
                const int b = __builtin_ctzll(m) >> 3;
                m &= ~(0xffULL << (8 * b));
                cert_sp_pair(S, (int)(w * 8 + b), seed, bad, fp, pairs);
            }
        }
        for (int64_t e = nw * 8 + i0; e < E; e += stride)  // tail
            if (S.emark[e]) cert_sp_pair(S, (int)e, seed, bad, fp, pairs);
    }
    cert_flush(S, 0, bad, 2 * pairs, fp, red64, redu);
    // list or map bytes read, in 16-byte vertex-scan units
    const uint64_t units = block_sum<uint64_t>(i0 == 0 ? (uint64_t)(listed ? cnt / 4 : E / 16) : 0, redu);
    if (threadIdx.x == 0 && units) addu64(&S.ctr->wk[WF_CERT][1], units);
}

// Sparse mode, vertex side: -u(v) * stored excess(v) and the vertex checks
// for every vertex; then the cut entries V\T -> T for the cut and
// closedness, the residual of an unmarked pair being its capacity: found in
// the rows of T through their pairs (workgroups holding a vertex of T), or,
// with cert_sside, in the rows of V\T themselves (workgroups holding one),
// whichever side the scheduler expects to be smaller.
__global__ __launch_bounds__(256) void k_cert_sp_rows(const SlotView* __restrict__ slots,
                                                      const int* __restrict__ list) {
    const GSlot S = load_slot(slots, list);
    __shared__ int64_t red64[256 / WAVE];
    __shared__ uint64_t redu[256 / WAVE];
    __shared__ int rp[257];
    __shared__ int tin[256];
    const int n = S.g.n;
    const uint64_t seed = S.g.cert_seed;
    int64_t cut = 0;
    uint64_t bad = 0, edges = 0, fp = 0, verts = 0;
    const int nblk = (n + 255) / 256;
    const int i = (int)threadIdx.x;
    for (int b = blockIdx.x; b < nblk; b += gridDim.x) {
        const int v0 = b * 256;
        const int nv = min(256, n - v0);
        __syncthreads();  // the previous block's LDS reads are done
        bool inT = false;
        if (i < nv) {
            const int v = v0 + i;
            inT = S.h[v] < n;
            const int64_t x = S.excess[v];
            fp -= vertex_weight(v, seed) * (uint64_t)(x - (v == S.s ? S.ctr->inject : 0));
            if (v != S.s && x < 0) ++bad;
            if (inT && v != S.t && x != 0) ++bad;
            if (v == S.t) {
                S.ctr->ex_t = x;
                if (!inT) ++bad;
            }
            if (v == S.s && inT) ++bad;
            ++verts;
            rp[i] = S.g.row_ptr[v];
            if (i == nv - 1) rp[nv] = S.g.row_ptr[v + 1];
        }
        const bool scan = i < nv && (S.cert_sside ? !inT : inT);
        if (i < nv) tin[i] = scan ? 1 : 0;
        if (!__syncthreads_or(scan)) continue;  // uniform per block
        const int E1 = rp[nv];
        int lo = 0;  // rows only grow along e
        for (int e = rp[0] + i; e < E1; e += 256) {
            int rr = lo;
#pragma unroll
            for (int step = 128; step > 0; step >>= 1) {
                const int c = min(rr + step, nv - 1);
                rr = rp[c] <= e ? c : rr;
            }
            lo = rr;
            if (!tin[rr]) continue;
            ++edges;
            const int w = S.g.col[e];
            if (S.cert_sside) {
                if (S.h[w] < n) {  // the row's entry row -> col[e] enters T
                    const int rv = S.g.rev[e];
                    const int64_t c = S.g.cap[e];
                    cut += c;
                    const int64_t x = (S.emark[e] | S.emark[rv]) ? S.res[e] : c;
                    if (x > 0) ++bad;
                }
            } else if (S.h[w] >= n) {  // pair col[e] -> row enters T
                const int rv = S.g.rev[e];
                const int64_t c = S.g.cap[rv];
                cut += c;
                const int64_t x = (S.emark[e] | S.emark[rv]) ? S.res[rv] : c;
                if (x > 0) ++bad;
            }
        }
    }
    cert_flush(S, cut, bad, edges, fp, red64, redu);
    const uint64_t bv = block_sum<uint64_t>(verts, redu);
    if (threadIdx.x == 0 && bv) addu64(&S.ctr->wk[WF_CERT][1], bv);
}

// Giant rows of the certificate in kGiantChunk-entry tasks over the grid:
// each chunk adds its part of the row's recomputed excess into gacc[row];
// k_cert_fin then checks each giant vertex.
__global__ __launch_bounds__(256) void k_cert_chunks(const SlotView* __restrict__ slots,
                                                     const int* __restrict__ list) {
    const GSlot S = load_slot(slots, list);
    __shared__ int64_t red64[256 / WAVE];
    __shared__ uint64_t redu[256 / WAVE];
    const int T = S.ctr->ngchunks;
    if ((int)blockIdx.x >= T) return;  // uniform per block
    const int n = S.g.n;
    int64_t cut = 0;
    uint64_t bad = 0, edges = 0;
    for (int task = blockIdx.x; task < T; task += gridDim.x) {
        const GTask gt = giant_task(S, task);
        const int k = gt.k, v = gt.v;
        const int e0 = gt.e0, e1 = gt.e1;
        const bool inT = S.h[v] < n;
        int64_t ex = 0;
        for (int e = e0 + (int)threadIdx.x; e < e1; e += blockDim.x) cert_entry(S, e, inT, n, ex, bad, cut);
        const int64_t tot = block_sum<int64_t>(ex, red64);
        if (threadIdx.x == 0) {
            if (tot) add64(&S.gacc[k], tot);
            edges += (uint64_t)(e1 - e0);
        }
    }
    cert_flush(S, cut, bad, edges, 0, red64, redu);
}

__global__ __launch_bounds__(256) void k_cert_fin(const SlotView* __restrict__ slots,
                                                  const int* __restrict__ list) {
    const GSlot S = load_slot(slots, list);
    __shared__ int64_t red64[256 / WAVE];
    __shared__ uint64_t redu[256 / WAVE];
    const int ng = S.ctr->ngiant;
    const int n = S.g.n;
    uint64_t bad = 0;
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < ng; g += gridDim.x * blockDim.x) {
        const int v = S.giant[g];
        cert_vertex(S, v, ld64(&S.gacc[g]), S.h[v] < n, bad);
    }
    // every k_cert / k_cert_chunks block has added its part by now (stream order)
    cert_flush(S, 0, bad, 0, 0, red64, redu);
}

void launch_cert(hipStream_t st, const SlotView* slots, const int* list, int nlist, int gridx, int mode) {
    if (mode == CERT_SPARSE) {
        hipLaunchKernelGGL(k_cert_sp_pairs, dim3(gridx, nlist), dim3(256), 0, st, slots, list);
        hipLaunchKernelGGL(k_cert_sp_rows, dim3(gridx, nlist), dim3(256), 0, st, slots, list);
        return;
    }
    if (mode != CERT_FULL) {
        hipLaunchKernelGGL(k_cert_fp, dim3(gridx, nlist), dim3(256), 0, st, slots, list);
        return;
    }
    hipLaunchKernelGGL(k_cert, dim3(gridx, nlist), dim3(256), 0, st, slots, list);
    hipLaunchKernelGGL(k_giant_prep, dim3(1, nlist), dim3(1024), 0, st, slots, list);
    hipLaunchKernelGGL(k_cert_chunks, dim3(per_slot_grid(1024, nlist), nlist), dim3(256), 0, st, slots, list);
    hipLaunchKernelGGL(k_cert_fin, dim3(4, nlist), dim3(256), 0, st, slots, list);
}

// ------------------------------------------------------------- prune mask (a6)
// After FWD into h and UND into aux: aux = keep flag.  The kept count goes to
// ctr->visited and an order-independent 64-bit hash of the kept set to
// hash_out[list index].

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33;
    return x;
}

__global__ __launch_bounds__(256) void k_mask_combine(const SlotView* __restrict__ slots,
                                                      const int* __restrict__ list,
                                                      unsigned long long* hash_out) {
    const GSlot S = load_slot(slots, list);
    __shared__ uint64_t redu[256 / WAVE];
    const int n = S.g.n;
    uint64_t kept = 0, hsh = 0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int v = (int)i;
        const bool keep = (S.h[v] != n && S.aux[v] != n) || v == S.s || v == S.t;
        S.aux[v] = keep ? 1 : 0;
        if (keep) { ++kept; hsh += mix64((uint64_t)v + 0x9E3779B97F4A7C15ULL); }
    }
    const uint64_t bk = block_sum<uint64_t>(kept, redu);
    const uint64_t bh = block_sum<uint64_t>(hsh, redu);
    if (threadIdx.x == 0) {
        if (bk) add32(&S.ctr->visited, (int32_t)bk);
        if (bh) addu64((uint64_t*)&hash_out[list[blockIdx.y]], bh);
    }
}

void launch_mask_combine(hipStream_t st, const SlotView* slots, const int* list, int nlist,
                         int gridx, unsigned long long* hash_out) {
    hipLaunchKernelGGL(k_mask_combine, dim3(gridx, nlist), dim3(256), 0, st, slots, list, hash_out);
}

// ------------------------------------------------------- capacity updates (a24)
// For entry e (u -> w) with capacity c -> c': the flow on the pair is
// f = cap[e] - res[e] (negative when it runs w -> u).  Setting cap[e] = c'
// keeps f when f <= c': res[e] += c' - c.  When f > c' the excess flow
// f - c' is sent back: res[e] = 0, res[rev] += c - f ... i.e. the pair
// carries exactly c' from u to w, u gains f - c', w loses f - c'.
// Entries are distinct within a batch (the host merges duplicates), so each
// pair has one writer.  The graph's cap array itself is updated by the host.

__global__ __launch_bounds__(256) void k_apply_updates(const SlotView* __restrict__ slots,
                                                       const int* __restrict__ list,
                                                       const int32_t* __restrict__ entries,
                                                       const int64_t* __restrict__ old_cap,
                                                       const int64_t* __restrict__ new_cap, int nupd) {
    const GSlot S = load_slot(slots, list);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nupd; i += gridDim.x * blockDim.x) {
        const int e = entries[i];
        const int r = S.g.rev[e];
        const int64_t c = old_cap[i], c2 = new_cap[i];
        const int64_t res = S.res[e];
        const int64_t f = c - res;  // flow u -> w on this pair (cap[rev] stays)
        if (f <= c2) {
            mark_entry(S, e);
            S.res[e] = res + (c2 - c);
        } else {
            const int64_t back = f - c2;  // flow that no longer fits
            mark_entry(S, e);
            S.res[e] = 0;
            S.res[r] = S.res[r] - back;
            // u = owner of entry e gets the flow back, w loses it
            const int w = S.g.col[e];
            const int u = S.g.col[r];
            add64(&S.excess[u], back);
            add64(&S.excess[w], -back);
        }
    }
}

void launch_apply_updates(hipStream_t st, const SlotView* slots, const int* list, int nlist,
                          const int32_t* entries, const int64_t* old_cap, const int64_t* new_cap,
                          int nupd) {
    const int gx = (nupd + 255) / 256 > 0 ? (nupd + 255) / 256 : 1;
    hipLaunchKernelGGL(k_apply_updates, dim3(gx < 4096 ? gx : 4096, nlist), dim3(256), 0, st, slots,
                       list, entries, old_cap, new_cap, nupd);
}

__global__ void k_cap_scatter(DGraph g, int64_t* __restrict__ cap, const int32_t* __restrict__ entries,
                              const int64_t* __restrict__ new_cap, int64_t* __restrict__ old_cap, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int e = entries[i];
        const int64_t was = cap[e];
        old_cap[i] = was;
        cap[e] = new_cap[i];
        const int64_t d = new_cap[i] - was;
        const int rv = g.rev[e];
        add64(&g.capsum[g.col[rv]], d);  // col[rev e] owns e
        addu64(g.cert_fp, pair_weight((uint64_t)min(e, rv), g.cert_seed) * (uint64_t)d);
    }
}

// After k_cap_scatter (an entry and its pair may both be in one batch):
// the updated pairs' capacity sums.
__global__ void k_pcap_update(DGraph g, const int32_t* __restrict__ entries, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int e = entries[i];
        const int rv = g.rev[e];
        const int64_t p = g.cap[e] + g.cap[rv];
        g.pcap[e] = p;
        g.pcap[rv] = p;
    }
}

void launch_cap_scatter(hipStream_t st, const DGraph& g, int64_t* cap, const int32_t* entries,
                        const int64_t* new_cap, int64_t* old_cap, int n) {
    const int gx = std::max(1, std::min(4096, (n + 255) / 256));
    hipLaunchKernelGGL(k_cap_scatter, dim3(gx), dim3(256), 0, st, g, cap, entries, new_cap, old_cap, n);
    hipLaunchKernelGGL(k_pcap_update, dim3(gx), dim3(256), 0, st, g, entries, n);
}

// Certificate constants of a graph: capsum[v] and cert_fp (see DGraph),
// and the pair capacities pcap.
// One wave per row (rows of any length stay balanced), rows grid-stride.
__global__ __launch_bounds__(256) void k_cert_consts(DGraph g, int64_t* __restrict__ capsum,
                                                     uint64_t* __restrict__ fp_out) {
    __shared__ uint64_t red[256 / WAVE];
    uint64_t fp = 0;
    const int lane = lane_id();
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x / WAVE);
    for (int64_t v = (int64_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE; v < g.n; v += waves) {
        int64_t cs = 0;
        for (int e = g.row_ptr[v] + lane; e < g.row_ptr[v + 1]; e += WAVE) {
            const int64_t c = g.cap[e];
            const int rv = g.rev[e];
            cs += c;
            fp += pair_weight((uint64_t)min(e, rv), g.cert_seed) * (uint64_t)c;
            g.pcap[e] = c + g.cap[rv];
        }
        cs = wave_sum_i64(cs);
        if (lane == 0) capsum[v] = cs;
    }
    const uint64_t b = block_sum<uint64_t>(fp, red);
    if (threadIdx.x == 0 && b) addu64(fp_out, b);
}

// One wave per row, 64 entries per step, stopping at the first step that
// holds a forward entry.
__global__ __launch_bounds__(256) void k_first_forward(DGraph g, int32_t* __restrict__ fstart) {
    const int lane = lane_id();
    const int64_t nw = (int64_t)gridDim.x * (blockDim.x / WAVE);
    for (int64_t v = (int64_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE; v < g.n; v += nw) {
        const int rb = g.row_ptr[v], re = g.row_ptr[v + 1];
        int f = rb;
        for (int base = rb; base < re; base += WAVE) {
            const int e = base + lane;
            const uint64_t m = __ballot(e < re && g.cap[e] > 0);
            if (m) {
                f = base + __ffsll((unsigned long long)m) - 1;
                break;
            }
        }
        if (lane == 0) fstart[v] = f;
    }
}

// One wave per row: the forward entry whose head has the longest row (ties:
// the first such entry).
__global__ __launch_bounds__(256) void k_bu_pick(DGraph g, int32_t* __restrict__ pick) {
    const int lane = lane_id();
    const int64_t nw = (int64_t)gridDim.x * (blockDim.x / WAVE);
    for (int64_t v = (int64_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE; v < g.n; v += nw) {
        const int rb = g.row_ptr[v], re = g.row_ptr[v + 1];
        // (row length << 32) | (~entry): the max takes the longest head row,
        // then the lowest entry
        uint64_t best = 0;
        for (int e = rb + lane; e < re; e += WAVE) {
            if (g.cap[e] <= 0) continue;
            const int w = g.col[e];
            const uint64_t d = (uint64_t)(uint32_t)(g.row_ptr[w + 1] - g.row_ptr[w]) + 1;
            const uint64_t key = (d << 32) | (uint64_t)(~(uint32_t)e);
            best = key > best ? key : best;
        }
        for (int o = WAVE / 2; o > 0; o >>= 1) {
            const uint64_t x = __shfl_xor(best, o, WAVE);
            best = x > best ? x : best;
        }
        if (lane == 0) {
            const int e = best ? (int)~(uint32_t)(best & 0xffffffffu) : -1;
            pick[2 * v] = e;
            pick[2 * v + 1] = e >= 0 ? g.col[e] : -1;
        }
    }
}

// One wave per 64 entries: their forward bits as one 64-bit word (two
// 32-bit words of fwdm; the array is 8-byte aligned).
__global__ __launch_bounds__(256) void k_fwd_mask(DGraph g, uint32_t* __restrict__ fwdm) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(WAVE - 1)); base < g.E; base += stride) {
        const int64_t e = base + lane_id();
        const uint64_t m = __ballot(e < g.E && g.cap[e] > 0);
        if (lane_id() == 0) *(uint64_t*)(fwdm + (base >> 5)) = m;
    }
}

// recv[v] from a given residual: 1 when some entry of v's row differs from
// its capacity (the rule the pushes keep, for a residual loaded from the
// host).
__global__ __launch_bounds__(256) void k_recv_from_res(DGraph g, const int64_t* __restrict__ res,
                                                      uint8_t* __restrict__ recv) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < g.n; v += stride) {
        uint8_t r = 0;
        for (int e = g.row_ptr[v]; e < g.row_ptr[v + 1]; ++e) r |= res[e] != g.cap[e] ? 1 : 0;
        recv[v] = r;
    }
}

void launch_recv_from_res(hipStream_t st, const DGraph& g, const int64_t* res, uint8_t* recv) {
    const int gx = std::max(1, std::min(16384, (g.n + 255) / 256));
    hipLaunchKernelGGL(k_recv_from_res, dim3(gx), dim3(256), 0, st, g, res, recv);
}

void launch_fwd_mask(hipStream_t st, const DGraph& g, uint32_t* fwdm) {
    const int gx = std::max(1, (int)std::min<int64_t>(16384, ((int64_t)g.E + 255) / 256));
    hipLaunchKernelGGL(k_fwd_mask, dim3(gx), dim3(256), 0, st, g, fwdm);
}

void launch_bu_pick(hipStream_t st, const DGraph& g, int32_t* pick) {
    const int gx = std::max(1, std::min(16384, (g.n + 3) / 4));
    hipLaunchKernelGGL(k_bu_pick, dim3(gx), dim3(256), 0, st, g, pick);
}

void launch_first_forward(hipStream_t st, const DGraph& g, int32_t* fstart) {
    const int gx = std::max(1, std::min(16384, (g.n + 3) / 4));
    hipLaunchKernelGGL(k_first_forward, dim3(gx), dim3(256), 0, st, g, fstart);
}

void launch_cert_consts(hipStream_t st, const DGraph& g, int64_t* capsum, uint64_t* fp_out) {
    const int gx = std::max(1, std::min(16384, (g.n + 3) / 4));
    hipLaunchKernelGGL(k_cert_consts, dim3(gx), dim3(256), 0, st, g, capsum, fp_out);
}

}  // namespace mf
