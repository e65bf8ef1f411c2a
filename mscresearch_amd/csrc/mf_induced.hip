// mf_induced.hip -- induced subgraph of a resident graph on a vertex mask,
// built on the GPU (reference: build_induced_subgraph_by_mask, BC:2477-2551).
//
// Kept vertices keep their relative order (so the region-ordered layout and
// the tile structure carry over: a parent tile becomes the run graph's tile
// holding the kept vertices of that tile), kept entries keep their row order,
// and rev is remapped through the parent-entry -> run-entry map.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>

#include "mf_engine.h"

namespace mf {

namespace {

__global__ void k_mask_to_u32(const int32_t* __restrict__ mask, int n, uint32_t* __restrict__ out) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v <= n; v += (int64_t)gridDim.x * blockDim.x)
        out[v] = (v < n && mask[v]) ? 1u : 0u;
}

// kept entries per kept row (0 for dropped rows); kdeg[n] = 0
__global__ void k_kept_deg(const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
                           const int32_t* __restrict__ mask, int n, uint32_t* __restrict__ kdeg) {
    // one wave per vertex: coalesced row scan
    const int lane = threadIdx.x & 63;
    const int64_t wid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t v = wid; v <= n; v += nw) {
        if (v == n) { if (lane == 0) kdeg[n] = 0; continue; }
        uint32_t c = 0;
        if (mask[v]) {
            for (int e = row_ptr[v] + lane; e < row_ptr[v + 1]; e += 64) c += mask[col[e]] ? 1u : 0u;
            for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
        }
        if (lane == 0) kdeg[v] = c;
    }
}

// old2sub / sub2old / tile_of from the exclusive vertex scan
__global__ void k_vertex_maps(const int32_t* __restrict__ mask, const uint32_t* __restrict__ vscan,
                              const int32_t* __restrict__ ptile_of, int n, int32_t* __restrict__ old2sub,
                              int32_t* __restrict__ sub2old, int32_t* __restrict__ tile_of,
                              uint8_t* __restrict__ mask8) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        const bool k = mask[v] != 0;
        mask8[v] = k ? 1 : 0;
        if (k) {
            const int s = (int)vscan[v];
            old2sub[v] = s;
            sub2old[s] = (int)v;
            tile_of[s] = ptile_of[v];
        } else {
            old2sub[v] = -1;
        }
    }
}

// run row_ptr: row_ptr2[s] = escan[sub2old[s]], row_ptr2[k] = E'
__global__ void k_row_ptr(const int32_t* __restrict__ sub2old, const uint32_t* __restrict__ escan,
                          int k, int n, int32_t* __restrict__ row_ptr2) {
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s <= k; s += (int64_t)gridDim.x * blockDim.x)
        row_ptr2[s] = (s < k) ? (int32_t)escan[sub2old[s]] : (int32_t)escan[n];
}

// scatter kept entries; e2s[parent entry] = run entry or -1.  One wave per
// kept row: ballot ranks the kept entries of each 64-entry chunk.
__global__ void k_scatter_edges(const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
                                const int64_t* __restrict__ cap, const int32_t* __restrict__ mask,
                                const int32_t* __restrict__ old2sub, const uint32_t* __restrict__ escan,
                                int n, int32_t* __restrict__ e2s, int32_t* __restrict__ col2,
                                int64_t* __restrict__ cap2) {
    const int lane = threadIdx.x & 63;
    const int64_t wid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t v = wid; v < n; v += nw) {
        const int rb = row_ptr[v], re = row_ptr[v + 1];
        if (!mask[v]) {
            for (int e = rb + lane; e < re; e += 64) e2s[e] = -1;
            continue;
        }
        int pos = (int)escan[v];
        for (int base = rb; base < re; base += 64) {
            const int e = base + lane;
            bool keep = false;
            int w = 0;
            if (e < re) { w = col[e]; keep = mask[w] != 0; }
            const unsigned long long m = __ballot(keep);
            if (e < re) {
                if (keep) {
                    const int p = pos + __popcll(m & ((1ull << lane) - 1ull));
                    e2s[e] = p;
                    col2[p] = old2sub[w];
                    cap2[p] = cap[e];
                } else {
                    e2s[e] = -1;
                }
            }
            pos += __popcll(m);
        }
    }
}

__global__ void k_remap_rev(const int32_t* __restrict__ rev, const int32_t* __restrict__ e2s, int64_t E,
                            int32_t* __restrict__ rev2) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E; e += (int64_t)gridDim.x * blockDim.x) {
        const int p = e2s[e];
        if (p >= 0) rev2[p] = e2s[rev[e]];
    }
}

// run tile starts: kept vertices before each parent tile start
__global__ void k_tiles(const int32_t* __restrict__ ptile_start, const uint32_t* __restrict__ vscan,
                        int ntiles, int32_t* __restrict__ tile_start2) {
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j <= ntiles; j += (int64_t)gridDim.x * blockDim.x)
        tile_start2[j] = (int32_t)vscan[ptile_start[j]];
}

__global__ void k_fwd_from_h(const int32_t* __restrict__ h, int n, uint8_t* __restrict__ mask8) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x)
        mask8[v] = h[v] < n ? 1 : 0;
}

__global__ void k_class_mask(const uint8_t* __restrict__ fwd, const int32_t* __restrict__ wcc, int32_t wt,
                             int n, int32_t* __restrict__ mask) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x)
        mask[v] = (fwd[v] && wcc[v] == wt) ? 1 : 0;
}

__global__ void k_gather_u8(const uint8_t* __restrict__ src, const int32_t* __restrict__ idx, int k,
                            uint8_t* __restrict__ out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < k; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = src[idx[i]];
}

__global__ void k_eq_mask(const int32_t* __restrict__ ids, int32_t want, int n, int32_t* __restrict__ mask) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x)
        mask[v] = ids[v] == want ? 1 : 0;
}

inline int grid_for(int64_t items, int per_block) {
    int64_t g = (items + per_block - 1) / per_block;
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, 16384));
}

}  // namespace

#define HIPCHK(expr)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess) {                                                          \
            set_error(std::string(#expr) + ": " + hipGetErrorString(e_));                \
            return e_ == hipErrorOutOfMemory ? MF_ENOMEM : MF_EHIP;                      \
        }                                                                                \
    } while (0)

void launch_fwd_from_h(hipStream_t st, const int32_t* h, int n, uint8_t* mask8) {
    hipLaunchKernelGGL(k_fwd_from_h, dim3(grid_for(n, 256)), dim3(256), 0, st, h, n, mask8);
}

void launch_class_mask(hipStream_t st, const uint8_t* fwd, const int32_t* wcc, int32_t wt, int n,
                       int32_t* mask) {
    hipLaunchKernelGGL(k_class_mask, dim3(grid_for(n, 256)), dim3(256), 0, st, fwd, wcc, wt, n, mask);
}

void launch_eq_mask(hipStream_t st, const int32_t* ids, int32_t want, int n, int32_t* mask) {
    hipLaunchKernelGGL(k_eq_mask, dim3(grid_for(n, 256)), dim3(256), 0, st, ids, want, n, mask);
}

int build_induced_gpu(const DeviceGraph& P, const int32_t* d_mask, int kept, hipStream_t st,
                      DeviceGraph& out) {
    const int n = P.g.n;
    const int64_t E = P.g.E;
    uint32_t *vflag = nullptr, *vscan = nullptr, *kdeg = nullptr, *escan = nullptr;
    int32_t* e2s = nullptr;
    void* tmp = nullptr;
    size_t tmp_bytes = 0, tb2 = 0;
    HIPCHK(hipMalloc(&vflag, sizeof(uint32_t) * ((size_t)n + 1)));
    HIPCHK(hipMalloc(&vscan, sizeof(uint32_t) * ((size_t)n + 1)));
    HIPCHK(hipMalloc(&kdeg, sizeof(uint32_t) * ((size_t)n + 1)));
    HIPCHK(hipMalloc(&escan, sizeof(uint32_t) * ((size_t)n + 1)));
    HIPCHK(hipMalloc(&e2s, sizeof(int32_t) * std::max<int64_t>(E, 1)));
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, vflag, vscan, n + 1, st));
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb2, kdeg, escan, n + 1, st));
    tmp_bytes = std::max(tmp_bytes, tb2);
    HIPCHK(hipMalloc(&tmp, std::max<size_t>(tmp_bytes, 16)));

    hipLaunchKernelGGL(k_mask_to_u32, dim3(grid_for(n + 1, 256)), dim3(256), 0, st, d_mask, n, vflag);
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, vflag, vscan, n + 1, st));
    hipLaunchKernelGGL(k_kept_deg, dim3(grid_for((int64_t)(n + 1) * 64, 256)), dim3(256), 0, st,
                       P.g.row_ptr, P.g.col, d_mask, n, kdeg);
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, kdeg, escan, n + 1, st));
    uint32_t Ek = 0, nk = 0;
    HIPCHK(hipMemcpyAsync(&Ek, escan + n, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(&nk, vscan + n, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if ((int)nk != kept) {
        set_error("induced subgraph: kept count mismatch");
        (void)hipFree(vflag); (void)hipFree(vscan); (void)hipFree(kdeg); (void)hipFree(escan); (void)hipFree(e2s); (void)hipFree(tmp);
        return MF_EINVAL;
    }
    const int k = (int)nk;
    const int ntiles = P.g.ntiles;
    HIPCHK(hipMalloc(&out.d_row_ptr, sizeof(int32_t) * ((size_t)k + 1)));
    HIPCHK(hipMalloc(&out.d_col, sizeof(int32_t) * std::max<uint32_t>(Ek, 1)));
    HIPCHK(hipMalloc(&out.d_cap, sizeof(int64_t) * std::max<uint32_t>(Ek, 1)));
    HIPCHK(hipMalloc(&out.d_rev, sizeof(int32_t) * std::max<uint32_t>(Ek, 1)));
    HIPCHK(hipMalloc(&out.d_tile_start, sizeof(int32_t) * ((size_t)ntiles + 1)));
    HIPCHK(hipMalloc(&out.d_tile_of, sizeof(int32_t) * std::max(k, 1)));
    HIPCHK(hipMalloc(&out.d_sub2old, sizeof(int32_t) * std::max(k, 1)));
    HIPCHK(hipMalloc(&out.d_old2sub, sizeof(int32_t) * std::max(n, 1)));
    HIPCHK(hipMalloc(&out.d_mask, std::max(n, 1)));
    hipLaunchKernelGGL(k_vertex_maps, dim3(grid_for(n, 256)), dim3(256), 0, st, d_mask, vscan,
                       P.g.tile_of, n, out.d_old2sub, out.d_sub2old, out.d_tile_of, out.d_mask);
    hipLaunchKernelGGL(k_row_ptr, dim3(grid_for(k + 1, 256)), dim3(256), 0, st, out.d_sub2old, escan, k, n,
                       out.d_row_ptr);
    hipLaunchKernelGGL(k_scatter_edges, dim3(grid_for((int64_t)n * 64, 256)), dim3(256), 0, st, P.g.row_ptr,
                       P.g.col, P.g.cap, d_mask, out.d_old2sub, escan, n, e2s, out.d_col, out.d_cap);
    hipLaunchKernelGGL(k_remap_rev, dim3(grid_for(E, 256)), dim3(256), 0, st, P.g.rev, e2s, E, out.d_rev);
    hipLaunchKernelGGL(k_tiles, dim3(grid_for(ntiles + 1, 256)), dim3(256), 0, st, P.g.tile_start, vscan,
                       ntiles, out.d_tile_start);
    if (P.d_live) {
        // the parent's reach of its largest SCC, in run-graph ids
        HIPCHK(hipMalloc(&out.d_live, std::max(k, 1)));
        hipLaunchKernelGGL(k_gather_u8, dim3(grid_for(k, 256)), dim3(256), 0, st, P.d_live, out.d_sub2old, k,
                           out.d_live);
        out.giant_scc = P.giant_scc;
    }
    HIPCHK(hipGetLastError());
    out.old2sub_host.resize(n);
    if (n > 0)
        HIPCHK(hipMemcpyAsync(out.old2sub_host.data(), out.d_old2sub, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    (void)hipFree(vflag); (void)hipFree(vscan); (void)hipFree(kdeg); (void)hipFree(escan); (void)hipFree(e2s); (void)hipFree(tmp);
    out.g.n = k;
    out.g.E = (int)Ek;
    out.g.max_deg = P.g.max_deg;
    out.g.ntiles = ntiles;
    out.g.R = P.g.R;
    out.g.row_ptr = out.d_row_ptr;
    out.g.col = out.d_col;
    out.g.cap = out.d_cap;
    out.g.rev = out.d_rev;
    out.g.tile_start = out.d_tile_start;
    out.g.tile_of = out.d_tile_of;
    out.tile_max = P.tile_max;
    out.bytes = sizeof(int32_t) * ((size_t)k + 1) + (size_t)Ek * 16 + sizeof(int32_t) * (ntiles + 1 + 2 * (size_t)k + n) + n +
                (out.d_live ? (size_t)k : 0);
    return init_cert_consts(out, st);
}

}  // namespace mf
