// Device view and launchers of the batched multi-source kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <type_traits>

namespace bcx {

constexpr int BATCH_W = 64;  // lanes of a wave64; a tile holds W * 64 sources
constexpr int MAX_SPL = 4;   // largest W: sources per lane of the multi-launch tiles

// A tile of the multi-launch kernels holds W * 64 sources (W = spl: 1, 2 or 4):
// lane k owns sources k + 64 h, h < W.  A source mask is W 64-bit words (word
// h bit k = source k + 64 h), stored as W consecutive words per vertex or list
// entry; a per-(vertex, source) row holds W consecutive words per lane, row
// cell (v * 64 + k) * W + h.
struct BatchDev {
    int32_t n;
    int32_t spl;                      // W, sources per lane
    const int32_t *rp, *col, *reach;
    uint64_t *vis, *mb0, *mb1, *mb2;  // [n][W] sources reached; rotating level masks
    void *sig;                        // [n*64*W] u32, u64 or f64 bits (the tile's sigma mode)
    void *tt;                         // [n*64*W] float or double
    int32_t *lvl;                     // optional [n*64*W] levels (probe)
    double *dbg_delta;                // optional [n*64*W] dependencies (probe)
    int32_t *Q;                       // level lists, concatenated
    uint64_t *Qm;                     // [qcap][W] source mask of each list entry
    int32_t *ctr;                     // [0] list end, [1] sigma overflow, [2] list overflow,
                                      // [3] depth (-1 while the forward runs), [4] last
                                      // injection level, [5] sources in the tile
    int32_t *qoff;                    // level offsets into Q (device-maintained)
    int32_t qcap;
    // hub vertices (degree > hub_deg): their arcs are cut into static slices
    // so several waves share one hub; n_hub == 0 disables the split
    int32_t n_hub, n_slice, hub_deg, hub_stride;
    const int32_t *hub_v;             // [n_hub] vertex of each hub
    const int32_t *hub_nsl;           // [n_hub] slices per hub
    const int32_t *sl_hub;            // [n_slice] hub of each slice
    const int32_t *sl_j0, *sl_j1;     // [n_slice] arc range of each slice
    uint64_t *hub_mask;               // [n_hub * hub_stride][W] sources of hub at level l
    double *hub_acc;                  // [n_hub * 64 * W] backward partial sums
    int32_t *hub_cnt;                 // [n_hub] slices finished this level
    const int32_t *srcs;              // [64 W] source of each tile slot (seed kernel only)
    const double *ws;                 // [64 W] source weight r(s)+1 (seed kernel only)
    double *wsl;                      // [64 W] the slot's copy of ws, made by the seed kernel:
                                      // the level kernels read it, so their launches are the
                                      // same for every tile of the slot (replayable graph)
    double *bc;                       // the slot's BC accumulator (folded per component)
    uint64_t limit;                   // integer sigma threshold (0 = the word's range)
    // Contracted chains (n_chain == 0: none).  Vertices [0, n) form the tile
    // graph; the interiors of chain c are the ids n + ch_off[c] + i - 1,
    // i = 1 .. K_c - 1, K_c = ch_off[c+1] - ch_off[c] + 1 arcs from L to R.
    // An endpoint's row keeps one arc per attached chain, at the position of
    // its arc into the chain, encoded as ~c (< 0): the level kernels skip it,
    // and the pull adds a chain source's own path (1) there, in CSR order.
    int32_t n_chain;
    const int32_t *ch_L, *ch_R;       // [n_chain] endpoints
    const int32_t *ch_off;            // [n_chain + 1] first interior of each chain
    const double *ch_pre;             // [interiors] P_c(i) = sum over interiors 1..i of r(x)+1
    const int32_t *ch_of;             // [interiors] chain of each interior
    const int32_t *ep_idx;            // [n] endpoint index of a vertex (-1: none)
    const int32_t *ep_off;            // [n_ep + 1] chains attached to endpoint e:
    const int32_t *ep_list;           //   ep_list[ep_off[e] ..) = 2c + (0: e is L, 1: R)
    int32_t n_ep;                     // endpoints (vertices with an attached chain)
    int32_t *ep_lvl;                  // [n_ep * 64 * W] the tile's level of each endpoint per source
    double *ep_rho;                   // [n_ep * 64 * W] the chain terms of each endpoint's rho per
                                      // source (made after the forward, added by the backward)
    int32_t *lane_ch;                 // [2][64 W] per source slot k + 64 h: chain of its
                                      // source (-1: a tile vertex) and its position in it
    uint64_t *lane_mask;              // [W] sources that are chain interiors
};

// Sigma word of a tile: 32-bit integers, 64-bit integers, or FP64 summed in
// CSR order.  A tile starts in one mode and moves up one mode when a count
// leaves the word's range (or passes `limit`), rerunning its forward.
constexpr int SIG_U32 = 0, SIG_U64 = 1, SIG_F64 = 2;

// Every level kernel reads its list ranges from b.qoff / b.ctr, so a whole
// tile is a fixed launch sequence with no host round trip per level.  The
// launchers pick the kernels of b.spl.
hipError_t launch_b_init(const BatchDev &b, int ns, int sm, int cap, hipStream_t st);
hipError_t launch_b_forward(const BatchDev &b, int sm, int l, int grid, hipStream_t st);
hipError_t launch_b_back(const BatchDev &b, int sm, bool fp32, int l, int grid, hipStream_t st);
hipError_t launch_b_finish(const BatchDev &b, int grid, hipStream_t st);
// Contracted chains: the count check between forward and backward (a tie
// interior's count sigma_L + sigma_R must fit the tile's word), the
// endpoints' chain terms (after the check), and the interiors' dependencies
// and BC after the backward.
hipError_t launch_b_chain_check(const BatchDev &b, int sm, int grid, hipStream_t st);
hipError_t launch_b_chain_terms(const BatchDev &b, int sm, int grid, hipStream_t st);
hipError_t launch_b_chains(const BatchDev &b, int sm, bool fp32, int grid, hipStream_t st);
hipError_t launch_b_clear_all(const BatchDev &b, int end, int grid, hipStream_t st);
hipError_t launch_b_fold(double *dst, double *src, const int32_t *perm, int n, int grid, hipStream_t st);

}  // namespace bcx
