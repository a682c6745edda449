// Runtime types shared by the plain and batched host drivers.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <utility>
#include <vector>

#include "bcx.h"
#include "bcx_device.h"
#include "bcx_kernels.h"

struct bcx_ctx;

namespace bcx {

// Grow-only device buffer accounted against the context's memory cap.
struct DBuf {
    void *p = nullptr;
    size_t bytes = 0;
    int ensure(bcx_ctx *ctx, size_t need, const char *what);
    void release(bcx_ctx *ctx);
};

// HIP-event intervals on the context stream, summed by collect().  Nested
// begin/end pairs record only at the outermost level, so a run whose
// streams overlap is timed as one interval.
struct Timer {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> evs;
    size_t used = 0;
    int depth = 0;
    void begin(bcx_ctx *ctx);
    void end(bcx_ctx *ctx);
    double collect();
    void destroy();
};

// Host copy of a component (planning needs the chains and the degrees).
struct ChainCut;  // multi-launch copy with contracted chains (bcx_batched_host.hip)

struct HostComp {
    int32_t n = 0;
    int64_t arcs = 0;
    std::vector<int32_t> rp, col, reach;
    int32_t n_chains = 0;
    std::vector<int32_t> ch_ptr, ch_nodes, ch_ends;
};

struct DComp {
    int32_t n = 0;
    int64_t arcs = 0;
    int64_t offset = 0;  // into the concatenated BC vector
    int32_t depth = -1;  // BFS depth from the highest-degree vertex (-1: not measured)
    double sigma_est = 0.0;  // largest path count of that BFS (measured with depth)
    std::vector<int32_t> rank;      // batched path: BFS rank of each vertex (built on first use)
    std::vector<uint8_t> on_chain;  // batched path: 1 for chain vertices (built on first use)
    std::vector<int32_t> ball;      // batched path: every vertex in ball-tile order (built on first use)
    int8_t tile_order = -1;         // batched path: 0 BFS-rank, 1 ball, 2 landmark tiles (tuned on first use)
    std::vector<int32_t> lm_pos;    // batched path: position of each vertex in the landmark order
    HostComp host;
    DBuf b_rp, b_col, b_reach, b_chp, b_chn, b_che;
    CompDev dev{};
    ChainDev chains{};
    // multi-launch batched path: the component relabelled in BFS order from
    // its highest-degree vertex (built on first use), perm[new] = old on the
    // device, inv[old] = new
    DComp *relabel = nullptr;
    DBuf b_perm;
    std::vector<int32_t> relabel_inv;
    // multi-launch batched path with contracted chains (built on first use;
    // cut_state -1 not yet tried, 0 the component has too few such chains)
    ChainCut *cut = nullptr;
    int8_t cut_state = -1;
    // batched path: the schedule measured on sample tiles, cached for every
    // later run of the component (sched_kind -1: kernel not measured;
    // sched_multi: the multi-launch fields below were measured)
    int8_t sched_kind = -1;
    bool sched_multi = false;
    int8_t sched_cut = 1;     // multi-launch: 1 contracted chains, 0 the relabelled copy
    int16_t sched_bpc = 0;    // multi-launch: level-kernel blocks per CU (0: default)
    int16_t sched_slots = 0;  // multi-launch: tiles in flight (0: default)
    int8_t sched_spl = 0;     // multi-launch: sources per lane, 1, 2 or 4 (0: default)
};

// One endpoint-row batch: BFS work (endpoints, with slot and mode) then the
// chain sources that combine those rows.
struct PlanBatch {
    std::vector<int32_t> bsrc, bslot;
    std::vector<int8_t> bmode;
    std::vector<int32_t> cci, cpos, cslots;
    int32_t n_rows = 0;
};

struct Plan {
    std::vector<int32_t> free_src;     // BFS sources that need no row
    std::vector<PlanBatch> batches;
    int64_t n_bfs = 0, n_chain = 0, n_row_only = 0;
};

int make_plan(const HostComp &hc, const int32_t *subset, int64_t n_sub, bool use_chains,
              int64_t row_cap, Plan &plan);

}  // namespace bcx

struct BatchedState;  // batched path buffers (bcx_batched_host.hip)
struct SmallState;    // small-component kernel buffers (bcx_batched_host.hip)

struct bcx_ctx {
    int device = 0;
    int n_cu = 0;
    std::string name;
    size_t mem_cap = 0, mem_used = 0;
    std::map<void *, size_t> sizes;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;  // batched: persistent passes beside the multi-launch tiles
    std::string err;
    bool overflow = false;
    int plain_wg_per_cu = 2;
    bcx::Timer timer, timer_chain;
    bcx::DBuf counters, bc_acc, plain_scratch, rows_buf;
    bcx::DBuf w_src, w_aux, w_mode, w_slots;
    bcx::DBuf bc_pack, bc_chunks;  // result download: packed slices and their chunk list
    double *bc_pin = nullptr;      // pinned host staging of the result download
    size_t bc_pin_bytes = 0;
    BatchedState *batched = nullptr;
    SmallState *small = nullptr;
    SmallState *wgk = nullptr;  // workgroup-per-tile kernel buffers
    std::vector<bcx::DBuf *> all_bufs() {
        return {&counters, &bc_acc, &plain_scratch, &rows_buf, &w_src, &w_aux, &w_mode, &w_slots, &bc_pack, &bc_chunks};
    }
};

struct bcx_dgraph {
    std::vector<bcx::DComp> comps;
    std::vector<int64_t> offs;  // comps[i].offset, packed for the source-slot search
    int64_t total_n = 0;
    bcx::DBuf pool;  // every component's CSR, reach and chain arrays, one allocation
};

// shared helpers (bcx_runtime.hip)
int ctx_fail(bcx_ctx *ctx, int code, const std::string &msg);
int ctx_hip(bcx_ctx *ctx, hipError_t e, const char *what);
int dev_alloc(bcx_ctx *ctx, void **p, size_t bytes, const char *what);
void dev_free(bcx_ctx *ctx, void *p);
int upload(bcx_ctx *ctx, const void *host, size_t bytes, bcx::DBuf &buf, const char *what);
int read_counters(bcx_ctx *ctx, bcx_stats *st);
int64_t row_capacity(bcx_ctx *ctx, int32_t n, const bcx_opts *o);
int plain_component(bcx_ctx *ctx, const bcx::DComp &dc, const bcx::Plan &plan, const bcx_opts *o,
                    double *bc, bcx_stats *st);

// Device address of the no-allocation entry point's static scratch array on
// `device` (current), looked up once per device (bcx_noalloc.hip).
int noalloc_init(int device);

// batched path (bcx_batched_host.hip)
struct SmallJob {
    const bcx::DComp *dc;
    std::vector<int32_t> srcs;
};
std::vector<int32_t> batched_sources(bcx::DComp &dc, const int32_t *subset, int64_t n_sub, int64_t *n_chain);
std::vector<int32_t> ball_sources(bcx::DComp &dc, const std::vector<int32_t> &srcs);
std::vector<int32_t> tuned_sources(bcx::DComp &dc, const std::vector<int32_t> &srcs);
// Every vertex of the component in the order a full batched run tiles its
// sources (64 consecutive entries per tile).
std::vector<int32_t> batched_order(bcx::DComp &dc);
// One persistent launch over the tiles of `jobs` (the small-component
// kernel, or with `wg` the workgroup-per-tile kernel), started on stream sp
// without waiting; _finish waits for it and reruns overflowed tiles in FP64.
int batched_persistent_start(bcx_ctx *ctx, const std::vector<SmallJob> &jobs, const bcx_opts *o, double *bc_base,
                             bcx_stats *st, bool wg, hipStream_t sp);
int batched_persistent_finish(bcx_ctx *ctx, bcx_stats *st, bool wg);
// Drops passes left unfinished by an earlier failed run (after waiting for them).
void batched_persistent_reset(bcx_ctx *ctx);
bool batched_narrow(bcx::DComp &dc);
void small_release(bcx_ctx *ctx);
// One-time preparation of a multi-launch component (its relabelled copy, or
// the copy with contracted chains when the options ask for chains).
int batched_prepare(bcx_ctx *ctx, bcx::DComp &dc, const bcx_opts *o);
int batched_component(bcx_ctx *ctx, bcx::DComp &dc, const std::vector<int32_t> &srcs, const bcx_opts *o,
                      double *bc, bcx_stats *st);
// Kernel (BCX_KERNEL_*) the batched path runs a component on, given the
// bcx_opts flags.
int batched_kind(bcx::DComp &dc, int flags);
// The kernel and, for the multi-launch tiles, the schedule of a component of
// the batched path: measured once on sample tiles for components of at least
// TUNE_N vertices (unless the options fix it), else batched_kind's rule.
// Returns a BCX_KERNEL_* or a negative BCX_E_* status.
int batched_schedule(bcx_ctx *ctx, bcx::DComp &dc, const bcx_opts *o, bcx_stats *st);
// Blocks per CU, tiles in flight and sources per lane a multi-launch run of
// dc uses.
void batched_sched_report(const bcx::DComp &dc, const bcx_opts *o, int32_t *bpc, int32_t *slots, int32_t *spl);
int batched_debug_tile(bcx_ctx *ctx, const bcx_component *comp, const int32_t *srcs, int32_t ns, const bcx_opts *o,
                       int32_t *level, uint64_t *sigma_u64, double *sigma_f64, double *delta, int32_t *sigma_is_f64,
                       int32_t *kernel_used);
void batched_release(bcx_ctx *ctx);
void batched_graph_release(bcx_ctx *ctx, bcx_dgraph *g);
