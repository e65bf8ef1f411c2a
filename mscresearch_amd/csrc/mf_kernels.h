// mf_kernels.h -- launch wrappers for the CDNA4 push-relabel kernels.
//
// Every kernel works on a list of slots (one query each): blockIdx.y selects
// list[blockIdx.y], blockIdx.x splits that slot's work.  A plain solve is a
// list of one slot.
#pragma once
#include <hip/hip_runtime_api.h>

#include "mf_device.h"

namespace mf {

// Counter reset bits for launch_reset.
enum : unsigned {
    RST_BFS = 1u << 0,     // fsize/fedges ring, level, visited
    RST_ACTIVE = 1u << 1,  // active counters, excess sums, band, tile_cnt
    RST_CERT = 1u << 2,    // cert_bad, cut
    RST_RM = 1u << 3,      // fixed_edges
    RST_WORK = 1u << 4,    // work counters
    RST_ERR = 1u << 5,
    RST_TOUCH = 1u << 6,   // new touch epoch, touched list emptied
    RST_SPARSE = 1u << 7,  // like RST_ACTIVE, keeping the touched list for a sparse count
};

enum BfsMode : int { BFS_GR = 0, BFS_FWD = 1, BFS_UND = 2 };

void launch_reset(hipStream_t st, const SlotView* slots, const int* list, int nlist, unsigned what);
void launch_init(hipStream_t st, const SlotView* slots, const int* list, int nlist, int gridx);
// The same for a slab whose residual belongs to the same graph: only pairs
// of entries marked since the last reset are restored.
void launch_init_marked(hipStream_t st, const SlotView* slots, const int* list, int nlist, int gridx);
void launch_preflow(hipStream_t st, const SlotView* slots, const int* list, int nlist);

// Global-relabel seeding: h = 0 on seeds, n elsewhere; seeds go to fa.
void launch_gr_seed(hipStream_t st, const SlotView* slots, const int* list, int nlist, int gridx);
// Local relabel of the tiles holding active vertices (region R), levels below
// lcap <= kLocalLevels.  Only raises heights, so a valid labelling stays valid:
//  mark   : aux = h on R; R' = {v in R : h < lcap}; seeds of R' (h == 0) start
//           the frontier, the rest of R' is marked n + 1;
//  bound  : b(v) = 1 + min h(x) over residual v -> x with x outside R' (the
//           heights around R' enter the BFS as sources at level b(v));
//  bucket : R' vertices with b(v) < lcap grouped by b(v) into hl;
//  bfs    : one workgroup, levels 0.. lcap-1, bucket L joins level L;
//  finish : h = max(aux, level) on reached vertices, max(aux, Lf) on the
//           rest, Lf being the first level the BFS did not complete.
void launch_local_mark(hipStream_t st, const SlotView* slots, const int* list, int nlist, int gridx,
                       int lcap);
void launch_local_bound(hipStream_t st, const SlotView* slots, const int* list, int nlist, int gridx,
                        int lcap);
void launch_local_bucket(hipStream_t st, const SlotView* slots, const int* list, int nlist, int gridx,
                         int lcap);
void launch_local_bfs(hipStream_t st, const SlotView* slots, const int* list, int nlist, int lcap);
void launch_local_finish(hipStream_t st, const SlotView* slots, const int* list, int nlist, int gridx);
// Single-source BFS seeding for pruning: lvl (h for FWD, aux for UND) = n,
// lvl[src] = 0, frontier {src}.
void launch_bfs_seed1(hipStream_t st, const SlotView* slots, const int* list, int nlist, int gridx,
                      int mode);
// One top-down level (level index taken from each slot's ctr->level).
void launch_bfs_td(hipStream_t st, const SlotView* slots, const int* list, int nlist, int gridx,
                   int mode);
// One bottom-up level.
// First bottom-up level of a BFS: the level-L bitmap (SlotView::fbits) from
// the levels.
void launch_bu_bits(hipStream_t st, const SlotView* slots, const int* list, int nlist, int gridx, int mode);
void launch_bfs_bu(hipStream_t st, const SlotView* slots, const int* list, int nlist, int gridx,
                   int mode);
// Multi-level top-down BFS while the frontier stays small: one workgroup
// (wgs == 1), or wgs workgroups per slot meeting at a barrier in global
// memory after each level (spin_ticks: wall-clock ticks a barrier may wait
// before it gives up and flags the slot's error word).  It stops above
// kSmallFrontier vertices or SlotView::bfs_edges frontier row entries.
// cap_total bounds the workgroups of one multi-workgroup launch (all of them
// must be resident at once).
void launch_bfs_small(hipStream_t st, const SlotView* slots, const int* list, int nlist, int mode,
                      int max_levels, int wgs, long long spin_ticks, int cap_total);

// Active list (light entries at fa[0..), heavy at fa[n-1..]) and counters.
void launch_active(hipStream_t st, const SlotView* slots, const int* list, int nlist, int gridx);
// The same from the previous list and the vertices touched since the last
// RST_TOUCH only (after a RST_SPARSE reset); gridx any size.
void launch_active_sparse(hipStream_t st, const SlotView* slots, const int* list, int nlist, int gridx);
// Discharge of the active list: lanes for light rows, waves for heavy rows;
// giants != 0: rows longer than kGiantDeg over the whole grid after it, in
// whole 2048-entry chunks (1) or in rounds that stop once the row is dry (2).
void launch_discharge_list(hipStream_t st, const SlotView* slots, const int* list, int nlist,
                           int gridx, int giants);
// The same discharge for every slot of the list in one persistent launch of
// `workgroups` workgroups taking 256-row blocks off a device work queue
// dq = {0, total, prefix of the slots' blocks [nlist + 1]} that the caller
// uploads with the list; giants as above.
void launch_discharge_queue(hipStream_t st, const SlotView* slots, const int* list, int nlist, int32_t* dq,
                            int workgroups, int giants);
// One workgroup of `threads` (256, 512 or 1024) per tile, tile heights and
// pending excess staged in LDS; split > 1: up to `split` workgroups per tile
// by excess score (k_tile_alloc), each staging its part of the tile.
// Labels of a saturated terminal cut for the certificate: side 0, T = {t};
// side 1, T = V \ {s}.
void launch_cut_labels(hipStream_t st, const SlotView* slots, const int* list, int nlist, int gridx, int side);

// An empty kernel: a phase mark in kernel traces and counter runs.
void launch_trace_marker(hipStream_t st);

void launch_discharge_tile(hipStream_t st, const SlotView* slots, const int* list, int nlist,
                           int gridx, int hot_only, int rounds, int giants, int threads, int split);
// Owner-exclusive path push from the active list.
void launch_path_push(hipStream_t st, const SlotView* slots, const int* list, int nlist, int gridx,
                      int max_hops);
// Hot tiles: the top `k` tiles by active count get tile_hot = 1.
void launch_hot_select(hipStream_t st, const SlotView* slots, const int* list, int nlist, int k);
// RemoveInvalidEdges repair over tiles holding active vertices, band
// [ctr->hmin - margin, ctr->hmax + margin].
void launch_rm(hipStream_t st, const SlotView* slots, const int* list, int nlist, int gridx,
               int margin);
// Min-cut certificate over the final residual and heights.  mode 1: pair
// sums by fingerprint (coalesced); mode 2: pair sums entry by entry; mode 3:
// the fingerprint over the marked pairs only (the slab's marks must cover
// every pair whose residual left its capacities).
void launch_cert(hipStream_t st, const SlotView* slots, const int* list, int nlist, int gridx, int mode);

// ---------------------------------------------------------------- pruning
// aux <- (h != n && aux != n) || v == s || v == t   (after FWD into h, UND into aux)
// ctr->visited <- kept count; hash of the mask into *hash_out (per list entry).
void launch_mask_combine(hipStream_t st, const SlotView* slots, const int* list, int nlist,
                         int gridx, unsigned long long* hash_out);

// Dynamic capacity updates on a residual: for each (entry e, new cap c):
// capacity increase adds to res[e]; a decrease below the current flow moves
// the excess flow back (excess at the tail, deficit at the head).
void launch_apply_updates(hipStream_t st, const SlotView* slots, const int* list, int nlist,
                          const int32_t* entries, const int64_t* old_cap, const int64_t* new_cap,
                          int nupd);

// cap[entries[i]] <- new_cap[i], old value to old_cap[i]; g.capsum and
// *g.cert_fp follow.  Entries must not repeat.
void launch_cap_scatter(hipStream_t st, const DGraph& g, int64_t* cap, const int32_t* entries,
                        const int64_t* new_cap, int64_t* old_cap, int n);
// g.capsum (n) and *fp_out (+=, zero it first) from g's capacities and seed.
void launch_cert_consts(hipStream_t st, const DGraph& g, int64_t* capsum, uint64_t* fp_out);
// fstart[v] = first entry of v's row with cap > 0, row_ptr[v] if none.
void launch_first_forward(hipStream_t st, const DGraph& g, int32_t* fstart);
// DGraph::bu_pick (2n ints)
void launch_bu_pick(hipStream_t st, const DGraph& g, int32_t* pick);
// DGraph::fwdm: one bit per entry, set for forward entries (cap > 0).
void launch_fwd_mask(hipStream_t st, const DGraph& g, uint32_t* fwdm);
// SlotView::recv from a residual loaded from the host (a row that differs
// from its capacities anywhere is marked).
void launch_recv_from_res(hipStream_t st, const DGraph& g, const int64_t* res, uint8_t* recv);

}  // namespace mf
