// Host driver of the batched multi-source path: tiles of up to 64 W sources
// (W = 1, 2 or 4 per lane of a wave64) of one component, forward level lists
// built on the device, FP32-compensated (or FP64) rho-form backward.  Path
// counts start in 32-bit words (half the bytes of every sigma gather); a tile
// whose counts leave the word's range is redone with 64-bit words, and past
// those with FP64.  Sources are ordered (BFS rank, landmark distances or graph
// balls, whichever sample tiles show to need less work) so the sources of a
// tile are close together and their frontiers overlap (fewer (vertex, level)
// list entries per source).
//
// Several tiles run at once, one per slot, each slot with its own stream and
// state: the many small levels of one tile (near its sources, along long
// chains) leave most compute units idle, and the other slots fill them.  A
// slot enqueues a tile as one fixed launch sequence and is polled through an
// event; the host never waits on a single tile.
#include <algorithm>
#include <array>
#include <cstring>
#include <thread>
#include <vector>

#include "bcx_batched.h"
#include "bcx_small.h"
#include "bcx_runtime.h"

using namespace bcx;

// Hub vertices of a component: degree above a threshold T, arcs cut into
// slices of T arcs so several waves share one hub in every level kernel.
// Every level kernel scans all slices once per level, so T is the smallest
// power of two >= HUB_MIN whose slices number at most HUB_SLICES_PER_CU per
// compute unit: measured on the Chung-Lu cores of the 1M-vertex mixed graph
// and the many-component graph, T = 512 (7.9k and 5.4k slices) beats 256
// (19.8k, 13.6k) by 17% and 15% and 1024 by 6% and 10%; on the R-MAT core
// T = 1024 (16.9k slices) beats 512 (28.9k) by 12% and 256 (86.6k) by 85%.
constexpr int HUB_MIN = 512, HUB_SLICES_PER_CU = 80, HUB_LEVELS = 1024;
// Tiles in flight on their own streams: 8 measured 4-10% above 4 on the
// Chung-Lu and R-MAT cores (6 equals 4, 12 and 16 within 1.2% of 8).
constexpr int MAX_SLOTS = 8;

// A tile's level launches, captured once and replayed for every tile whose
// launch arguments are the same: the arguments hold no tile index (the seed
// kernel, launched per tile, copies the tile's sources and weights into the
// slot), so a key covers every tile of a component on that slot with the
// same count mode, level count and precision.
struct TileGraph {
    BatchDev key;
    int sm = 0, C = 0, grid = 0;
    bool fp32 = false;
    hipGraphExec_t exec = nullptr;
};
constexpr int GRAPHS_PER_SLOT = 12;

struct Slot {
    DBuf vis, mb, sig, tt, Q, Qm, ctr, qoff, hub_mask, hub_acc, hub_cnt, lvl, dbg, bcacc, wsl;
    DBuf eplv, eprho, lanes;  // contracted chains: endpoint levels and chain terms per lane, lane table
    int32_t n_cap = 0, qcap = 0;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    int32_t *pinned = nullptr;
    // tile in flight
    int tile = -1, ns = 0, L = 0, C = 0, attempts = 0;
    int sm = SIG_U32;  // sigma mode of the tile in flight
    std::vector<TileGraph> graphs;  // most recently used last
};

struct BatchedState {
    Slot slots[MAX_SLOTS];
    int n_slots = 0;
    DBuf srcs, ws, hub_v, hub_nsl, sl_hub, sl_j0, sl_j1;
    int32_t src_cap = 0, n_hub = 0, n_slice = 0, hub_stride = 0, hub_deg = 0;
    bool probe = false;
    // a failed run abandoned tiles part-way: level masks may hold bits of an
    // unfinished forward and accumulators partial BC that was never folded;
    // the next bs_ensure clears every slot before it is used again
    bool dirty = false;
    int last_depth = 8;
    int last_sm = SIG_U32;  // sigma mode the last tile finished in: the next starts there
};

static int slot_ensure(bcx_ctx *ctx, Slot &s, int32_t n, bool probe) {
    int rc;
    const size_t N = (size_t)n;
    if (!s.stream) {
        if (hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess ||
            hipHostMalloc((void **)&s.pinned, 64, hipHostMallocDefault) != hipSuccess)
            return ctx_fail(ctx, BCX_E_HIP, "batched slot stream");
    }
    // capacities are recorded only once every buffer is in place: a failed
    // ensure has released its buffer, so the next call must allocate again
    if (n > s.n_cap) {
        s.n_cap = 0;
        // (every per-vertex and per-entry array holds MAX_SPL sources per lane)
        if ((rc = s.vis.ensure(ctx, 8 * MAX_SPL * N, "batched vis")) ||
            (rc = s.mb.ensure(ctx, 24 * MAX_SPL * N, "batched masks")) ||
            (rc = s.sig.ensure(ctx, 8 * 64 * MAX_SPL * N, "batched sigma")) ||
            (rc = s.tt.ensure(ctx, 8 * 64 * MAX_SPL * N, "batched dependency terms")) ||
            (rc = s.qoff.ensure(ctx, 4 * (N + 8), "batched level offsets")) ||
            (rc = s.bcacc.ensure(ctx, 8 * N, "batched BC accumulator")))
            return rc;
        if ((rc = ctx_hip(ctx, hipMemsetAsync(s.mb.p, 0, s.mb.bytes, s.stream), "memset")) ||
            (rc = ctx_hip(ctx, hipMemsetAsync(s.bcacc.p, 0, s.bcacc.bytes, s.stream), "memset")))
            return rc;
        s.n_cap = n;
    }
    if (s.qcap < 8 * n + 4096) {
        const int64_t q = std::min<int64_t>((int64_t)8 * n + 4096, INT32_MAX / 2);
        s.qcap = 0;
        if ((rc = s.Q.ensure(ctx, 4 * (size_t)q, "batched lists")) ||
            (rc = s.Qm.ensure(ctx, 8 * MAX_SPL * (size_t)q, "batched lists")))
            return rc;
        s.qcap = (int32_t)q;
    }
    if ((rc = s.ctr.ensure(ctx, 64, "batched counters")) ||
        (rc = s.wsl.ensure(ctx, 8 * 64 * MAX_SPL, "tile weights")))
        return rc;
    if (probe && ((rc = s.lvl.ensure(ctx, 4 * 64 * MAX_SPL * N, "probe levels")) ||
                  (rc = s.dbg.ensure(ctx, 8 * 64 * MAX_SPL * N, "probe delta"))))
        return rc;
    return BCX_OK;
}

// Slots sized for n; as many as fit in a quarter of the free capacity.
static int bs_ensure(bcx_ctx *ctx, int32_t n, int32_t n_src, int want_slots, bool probe) {
    if (!ctx->batched) ctx->batched = new BatchedState();
    BatchedState &bs = *ctx->batched;
    const size_t per = (size_t)n * MAX_SPL * (8 + 24 + 512 + 512 + 8 * 12) + (1 << 20);
    const size_t room = ctx->mem_cap > ctx->mem_used ? (ctx->mem_cap - ctx->mem_used) : 0;
    int fit = (int)std::max<size_t>(1, std::min<size_t>(MAX_SLOTS, room / 2 / std::max<size_t>(per, 1)));
    bs.n_slots = std::max(1, std::min(want_slots, fit));
    int rc;
    if (bs.dirty) {
        // after the abandoned work on each slot stream
        for (Slot &s : bs.slots) {
            if (!s.stream || !s.n_cap) continue;
            if ((rc = ctx_hip(ctx, hipMemsetAsync(s.mb.p, 0, s.mb.bytes, s.stream), "memset")) ||
                (rc = ctx_hip(ctx, hipMemsetAsync(s.bcacc.p, 0, s.bcacc.bytes, s.stream), "memset")) ||
                (rc = ctx_hip(ctx, hipStreamSynchronize(s.stream), "slot reset")))
                return rc;
            s.tile = -1;
        }
        bs.dirty = false;
    }
    for (int i = 0; i < bs.n_slots; ++i)
        if ((rc = slot_ensure(ctx, bs.slots[i], n, probe))) return rc;
    if (n_src > bs.src_cap) {
        bs.src_cap = 0;
        if ((rc = bs.srcs.ensure(ctx, 4 * (size_t)n_src, "batched sources")) ||
            (rc = bs.ws.ensure(ctx, 8 * (size_t)n_src, "batched weights")))
            return rc;
        bs.src_cap = n_src;
    }
    bs.probe = probe;
    return BCX_OK;
}

struct TileJob {
    const DComp *dc;
    const int32_t *perm;  // tile vertex id -> component vertex id (nullptr: identity)
    const ChainCut *cut;  // contracted chains of dc (nullptr: none)
    double *bc;
    uint64_t limit;
    bool fp32;
    int grid;
    int spl;      // sources per lane (W): tiles of 64 W sources
    bool graphs;  // replay captured level launches (first chunk of a tile)
};

static void set_chains(BatchDev &b, Slot &s, const ChainCut *cut);

static BatchDev make_dev(bcx_ctx *ctx, Slot &s, const TileJob &j, uint64_t limit) {
    const DComp &dc = *j.dc;
    BatchedState &bs = *ctx->batched;
    BatchDev b{};
    b.n = dc.n;
    b.rp = dc.dev.rp;
    b.col = dc.dev.col;
    b.reach = dc.dev.reach;
    b.vis = (uint64_t *)s.vis.p;
    b.mb0 = (uint64_t *)s.mb.p;
    b.mb1 = b.mb0 + (size_t)s.n_cap * MAX_SPL;
    b.mb2 = b.mb1 + (size_t)s.n_cap * MAX_SPL;
    b.sig = s.sig.p;
    b.tt = s.tt.p;
    b.lvl = bs.probe ? (int32_t *)s.lvl.p : nullptr;
    b.dbg_delta = bs.probe ? (double *)s.dbg.p : nullptr;
    b.Q = (int32_t *)s.Q.p;
    b.Qm = (uint64_t *)s.Qm.p;
    b.ctr = (int32_t *)s.ctr.p;
    b.qoff = (int32_t *)s.qoff.p;
    b.qcap = s.qcap;
    b.n_hub = bs.n_hub;
    b.n_slice = bs.n_slice;
    b.hub_deg = bs.hub_deg;
    b.hub_stride = bs.hub_stride;
    b.hub_v = (const int32_t *)bs.hub_v.p;
    b.hub_nsl = (const int32_t *)bs.hub_nsl.p;
    b.sl_hub = (const int32_t *)bs.sl_hub.p;
    b.sl_j0 = (const int32_t *)bs.sl_j0.p;
    b.sl_j1 = (const int32_t *)bs.sl_j1.p;
    b.hub_mask = (uint64_t *)s.hub_mask.p;
    b.hub_acc = (double *)s.hub_acc.p;
    b.hub_cnt = (int32_t *)s.hub_cnt.p;
    b.spl = j.spl;
    b.srcs = (const int32_t *)bs.srcs.p + (size_t)s.tile * BATCH_W * j.spl;
    b.ws = (const double *)bs.ws.p + (size_t)s.tile * BATCH_W * j.spl;
    b.wsl = (double *)s.wsl.p;
    b.bc = (double *)s.bcacc.p;
    b.limit = limit;
    set_chains(b, s, j.cut);
    return b;
}


// Forward levels [L, L+C), the speculative backward L+C-1 .. 1 (kernels past
// the depth or of an unfinished forward return at once), the cleanup, the
// counter read-back and the completion event.
static int launch_levels(bcx_ctx *ctx, Slot &s, const BatchDev &b, int L0, int C, const TileJob &j) {
    int rc;
    for (int l = L0; l < L0 + C; ++l)
        if ((rc = ctx_hip(ctx, launch_b_forward(b, s.sm, l, j.grid, s.stream), "forward"))) return rc;
    if (b.n_chain > 0 && s.sm != SIG_F64 &&
        (rc = ctx_hip(ctx, launch_b_chain_check(b, s.sm, j.grid, s.stream), "chain check")))
        return rc;
    if (b.n_chain > 0 && (rc = ctx_hip(ctx, launch_b_chain_terms(b, s.sm, j.grid, s.stream), "chain terms")))
        return rc;
    for (int l = L0 + C - 1; l >= 1; --l)
        if ((rc = ctx_hip(ctx, launch_b_back(b, s.sm, j.fp32, l, j.grid, s.stream), "back"))) return rc;
    if (b.n_chain > 0 && (rc = ctx_hip(ctx, launch_b_chains(b, s.sm, j.fp32, j.grid, s.stream), "chains")))
        return rc;
    return ctx_hip(ctx, launch_b_finish(b, j.grid, s.stream), "finish");
}

// The first chunk's launches from the slot's graph cache (captured on a miss).
static int replay_levels(bcx_ctx *ctx, Slot &s, BatchDev b, const TileJob &j) {
    b.srcs = nullptr;  // read by the seed kernel only: not part of the key
    b.ws = nullptr;
    auto hit = std::find_if(s.graphs.begin(), s.graphs.end(), [&](const TileGraph &g) {
        return g.sm == s.sm && g.C == s.C && g.grid == j.grid && g.fp32 == j.fp32 &&
               !std::memcmp(&g.key, &b, sizeof(b));
    });
    if (hit == s.graphs.end()) {
        if ((int)s.graphs.size() >= GRAPHS_PER_SLOT) {
            (void)hipGraphExecDestroy(s.graphs.front().exec);
            s.graphs.erase(s.graphs.begin());
        }
        TileGraph g;
        std::memcpy(&g.key, &b, sizeof(b));
        g.sm = s.sm;
        g.C = s.C;
        g.grid = j.grid;
        g.fp32 = j.fp32;
        hipGraph_t graph = nullptr;
        int rc = ctx_hip(ctx, hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal), "capture");
        if (rc) return rc;
        const int rl = launch_levels(ctx, s, b, 0, s.C, j);
        const hipError_t ee = hipStreamEndCapture(s.stream, &graph);
        if (rl || ee != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return rl ? rl : ctx_hip(ctx, ee, "capture");
        }
        const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) return ctx_hip(ctx, ei, "graph instantiate");
        s.graphs.push_back(g);
        hit = s.graphs.end() - 1;
    } else if (hit != s.graphs.end() - 1) {
        std::rotate(hit, hit + 1, s.graphs.end());  // most recently used last
        hit = s.graphs.end() - 1;
    }
    return ctx_hip(ctx, hipGraphLaunch(hit->exec, s.stream), "graph launch");
}

// Forward levels [L, L+C), the speculative backward L+C-1 .. 1 (kernels past
// the depth or of an unfinished forward return at once), the cleanup, the
// counter read-back and the completion event.  A tile's first chunk replays a
// captured graph; a deeper tile's later chunks are launched one by one.
static int enqueue_chunk(bcx_ctx *ctx, Slot &s, const TileJob &j, bcx_stats *st) {
    BatchDev b = make_dev(ctx, s, j, s.sm == SIG_F64 ? 0 : j.limit);
    int rc;
    if (j.graphs && s.L == 0) {
        if ((rc = replay_levels(ctx, s, b, j))) return rc;
        if (st) st->n_graph_launches++;
    } else if ((rc = launch_levels(ctx, s, b, s.L, s.C, j))) {
        return rc;
    }
    s.L += s.C;
    if (st) st->n_launches += 2 * s.C + (s.L - 1) + 1 + (b.n_chain > 0 ? 3 + (s.sm != SIG_F64) + (b.lvl != nullptr) : 0);
    if ((rc = ctx_hip(ctx, hipMemcpyAsync(s.pinned, b.ctr, 4 * sizeof(int32_t), hipMemcpyDeviceToHost, s.stream),
                      "counters")))
        return rc;
    return ctx_hip(ctx, hipEventRecord(s.done, s.stream), "event");
}

static int start_tile(bcx_ctx *ctx, Slot &s, int tile, int ns, int sm, const TileJob &j, bcx_stats *st) {
    s.tile = tile;
    s.ns = ns;
    s.sm = sm;
    s.L = 0;
    // levels of the first chunk: the last tile's depth + 3, rounded up to a
    // multiple of 8 so that a handful of captured graphs serve every tile
    s.C = (std::max(4, ctx->batched->last_depth + 3) + 7) & ~7;
    BatchDev b = make_dev(ctx, s, j, sm == SIG_F64 ? 0 : j.limit);
    int rc = ctx_hip(ctx, launch_b_init(b, ns, sm, j.grid, s.stream), "init");
    if (rc) return rc;
    if (st) {
        st->n_launches += 2;
        if (sm == SIG_F64) st->n_f64_sigma += ns;
    }
    return enqueue_chunk(ctx, s, j, st);
}

// Returns 1 while the slot's tile is running, 0 when it has completed.
static int poll_tile(bcx_ctx *ctx, Slot &s, const TileJob &j, bcx_stats *st, int *rc_out) {
    *rc_out = BCX_OK;
    hipError_t q = hipEventQuery(s.done);
    if (q == hipErrorNotReady) return 1;
    if (q != hipSuccess) { *rc_out = ctx_hip(ctx, q, "batched tile"); return 0; }
    const int32_t *h = s.pinned;
    int rc;
    if (h[3] < 0) {  // deeper than the levels launched so far
        s.C = std::min(2 * s.C, 4096);
        *rc_out = enqueue_chunk(ctx, s, j, st);
        return *rc_out ? 0 : 1;
    }
    if (++s.attempts > 6) { *rc_out = ctx_fail(ctx, BCX_E_OVERFLOW, "batched tile could not be completed"); return 0; }
    if (h[2]) {
        // lists full: masks may hold entries of vertices never listed
        // (a vertex is an entry at most once per source of the tile)
        const int64_t q2 = std::min<int64_t>((int64_t)s.qcap * 4,
                                             std::min<int64_t>((int64_t)64 * MAX_SPL * j.dc->n + 4096, INT32_MAX / 2));
        if (q2 <= s.qcap) { *rc_out = ctx_fail(ctx, BCX_E_NOMEM, "batched level lists cannot grow"); return 0; }
        if ((rc = ctx_hip(ctx, hipMemsetAsync(s.mb.p, 0, s.mb.bytes, s.stream), "memset")) ||
            (rc = ctx_hip(ctx, hipStreamSynchronize(s.stream), "sync"))) {
            *rc_out = rc;
            return 0;
        }
        s.qcap = 0;  // (a failed ensure has released the list it was growing)
        if ((rc = s.Q.ensure(ctx, 4 * (size_t)q2, "batched lists (growth)")) ||
            (rc = s.Qm.ensure(ctx, 8 * MAX_SPL * (size_t)q2, "batched lists (growth)"))) {
            *rc_out = rc;
            return 0;
        }
        s.qcap = (int32_t)q2;
        if (st && s.sm == SIG_F64) st->n_f64_sigma -= s.ns;  // counted again by the restart
        *rc_out = start_tile(ctx, s, s.tile, s.ns, s.sm, j, st);
        return *rc_out ? 0 : 1;
    }
    if (s.sm != SIG_F64 && h[1]) {
        // a count left the word's range: the backward did not run; redo the
        // tile with the next wider sigma word
        BatchDev b = make_dev(ctx, s, j, 0);
        if ((rc = ctx_hip(ctx, launch_b_clear_all(b, std::min(h[0], s.qcap), j.grid, s.stream), "clear"))) {
            *rc_out = rc;
            return 0;
        }
        s.attempts = 0;
        *rc_out = start_tile(ctx, s, s.tile, s.ns, s.sm + 1, j, st);
        return *rc_out ? 0 : 1;
    }
    ctx->batched->last_depth = h[3];
    ctx->batched->last_sm = s.sm;
    if (st) {
        if (h[3] > st->max_level) st->max_level = h[3];
        st->n_list_entries += h[0];
        st->n_tiles++;
    }
    return 0;
}

// Runs tiles [0, ntiles) over the slots; kernel time is the span between
// events on the context stream that every slot stream is ordered against.
static int run_tiles(bcx_ctx *ctx, const DComp &dc, int64_t n_src, const TileJob &j, bcx_stats *st,
                     int *sm_out) {
    BatchedState &bs = *ctx->batched;
    const int64_t TW = (int64_t)BATCH_W * j.spl;  // sources per tile
    const int64_t ntiles = (n_src + TW - 1) / TW;
    const int ns_slots = (int)std::min<int64_t>(bs.n_slots, ntiles);
    int rc;
    hipEvent_t e0;
    if ((rc = ctx_hip(ctx, hipEventCreateWithFlags(&e0, hipEventDisableTiming), "event"))) return rc;
    ctx->timer.begin(ctx);
    (void)hipEventRecord(e0, ctx->stream);
    for (int i = 0; i < ns_slots; ++i) (void)hipStreamWaitEvent(bs.slots[i].stream, e0, 0);
    int64_t next = 0;
    int active = 0;
    int max_sm = SIG_U32;
    bs.last_sm = SIG_U32;
    for (int i = 0; i < ns_slots; ++i) bs.slots[i].tile = -1;
    rc = BCX_OK;
    while ((next < ntiles || active > 0) && !rc) {
        for (int i = 0; i < ns_slots && !rc; ++i) {
            Slot &s = bs.slots[i];
            if (s.tile < 0) {
                if (next >= ntiles) continue;
                s.attempts = 0;
                const int ns = (int)std::min<int64_t>(TW, n_src - next * TW);
                // start in the mode the last tile needed (neighbouring tiles
                // have similar counts); every 16th tile tries 32-bit words again
                const int sm = (next % 16 == 0) ? SIG_U32 : bs.last_sm;
                rc = start_tile(ctx, s, (int)next++, ns, sm, j, st);
                active++;
                continue;
            }
            int r2;
            if (poll_tile(ctx, s, j, st, &r2) == 0) {
                rc = r2;
                max_sm = std::max(max_sm, s.sm);
                s.tile = -1;
                active--;
            }
        }
    }
    for (int i = 0; i < ns_slots; ++i) {
        (void)hipEventRecord(bs.slots[i].done, bs.slots[i].stream);
        (void)hipStreamWaitEvent(ctx->stream, bs.slots[i].done, 0);
    }
    // each slot accumulated its tiles' BC privately: fold into the component
    for (int i = 0; i < ns_slots && !rc; ++i)
        rc = ctx_hip(ctx, launch_b_fold(j.bc, (double *)bs.slots[i].bcacc.p, j.perm, dc.n, j.grid, ctx->stream),
                     "fold");
    ctx->timer.end(ctx);
    for (int i = 0; i < ns_slots; ++i) (void)hipStreamSynchronize(bs.slots[i].stream);
    (void)hipEventDestroy(e0);
    if (rc) bs.dirty = true;
    if (sm_out) *sm_out = max_sm;
    return rc;
}

// Static hub slice tables of one component.
static int stage_hubs(bcx_ctx *ctx, const DComp &dc) {
    BatchedState &bs = *ctx->batched;
    const HostComp &hc = dc.host;
    const int64_t budget = (int64_t)HUB_SLICES_PER_CU * std::max(1, ctx->n_cu);
    int64_t T = HUB_MIN;
    for (;; T *= 2) {
        int64_t ns = 0;
        for (int32_t v = 0; v < hc.n && ns <= budget; ++v) {
            const int64_t d = hc.rp[v + 1] - hc.rp[v];
            if (d > T) ns += (d + T - 1) / T;
        }
        if (ns <= budget) break;
    }
    bs.hub_deg = (int32_t)T;
    std::vector<int32_t> hv, nsl, slh, j0, j1;
    for (int32_t v = 0; v < hc.n; ++v) {
        const int32_t a = hc.rp[v], e = hc.rp[v + 1];
        if (e - a <= T) continue;
        const int32_t h = (int32_t)hv.size();
        hv.push_back(v);
        int32_t k = 0;
        for (int64_t x = a; x < e; x += T, ++k) {
            slh.push_back(h);
            j0.push_back((int32_t)x);
            j1.push_back((int32_t)std::min<int64_t>(e, x + T));
        }
        nsl.push_back(k);
    }
    const int64_t nh = (int64_t)hv.size();
    bs.n_hub = (int32_t)nh;
    bs.n_slice = (int32_t)slh.size();
    bs.hub_stride = nh ? (int32_t)std::min<int64_t>(HUB_LEVELS, (int64_t)hc.n + 2) : 0;
    if (!nh) return BCX_OK;
    int rc;
    if ((rc = upload(ctx, hv.data(), 4 * hv.size(), bs.hub_v, "hubs")) ||
        (rc = upload(ctx, nsl.data(), 4 * nsl.size(), bs.hub_nsl, "hubs")) ||
        (rc = upload(ctx, slh.data(), 4 * slh.size(), bs.sl_hub, "hub slices")) ||
        (rc = upload(ctx, j0.data(), 4 * j0.size(), bs.sl_j0, "hub slices")) ||
        (rc = upload(ctx, j1.data(), 4 * j1.size(), bs.sl_j1, "hub slices")))
        return rc;
    for (int i = 0; i < bs.n_slots; ++i) {
        Slot &s = bs.slots[i];
        if ((rc = s.hub_mask.ensure(ctx, 8 * MAX_SPL * (size_t)nh * bs.hub_stride, "hub masks")) ||
            (rc = s.hub_acc.ensure(ctx, 8 * 64 * MAX_SPL * (size_t)nh, "hub accumulators")) ||
            (rc = s.hub_cnt.ensure(ctx, 4 * (size_t)nh, "hub counters")))
            return rc;
    }
    return ctx_hip(ctx, hipStreamSynchronize(ctx->stream), "hubs");
}

// Sources and weights of tiles of tw sources (the last one padded).
static int stage_sources(bcx_ctx *ctx, const DComp &dc, const std::vector<int32_t> &srcs, int tw) {
    const size_t ntiles = (srcs.size() + tw - 1) / tw;
    std::vector<int32_t> hs(ntiles * tw, -1);
    std::vector<double> hw(ntiles * tw, 0.0);
    for (size_t i = 0; i < srcs.size(); ++i) {
        hs[i] = srcs[i];
        hw[i] = (double)dc.host.reach[srcs[i]] + 1.0;
    }
    BatchedState &bs = *ctx->batched;
    int rc;
    if ((rc = ctx_hip(ctx, hipMemcpyAsync(bs.srcs.p, hs.data(), 4 * hs.size(), hipMemcpyHostToDevice, ctx->stream),
                      "sources")) ||
        (rc = ctx_hip(ctx, hipMemcpyAsync(bs.ws.p, hw.data(), 8 * hw.size(), hipMemcpyHostToDevice, ctx->stream),
                      "weights")))
        return rc;
    return ctx_hip(ctx, hipStreamSynchronize(ctx->stream), "sources");
}

// BFS order from the highest-degree vertex (host).
static std::vector<int32_t> bfs_rank(const HostComp &hc) {
    const int32_t n = hc.n;
    int32_t root = 0;
    for (int32_t v = 1; v < n; ++v)
        if (hc.rp[v + 1] - hc.rp[v] > hc.rp[root + 1] - hc.rp[root]) root = v;
    std::vector<int32_t> rank((size_t)n, -1), q;
    q.reserve(n);
    q.push_back(root);
    rank[root] = 0;
    for (size_t i = 0; i < q.size(); ++i) {
        const int32_t v = q[i];
        for (int32_t j = hc.rp[v]; j < hc.rp[v + 1]; ++j)
            if (rank[hc.col[j]] < 0) {
                rank[hc.col[j]] = (int32_t)q.size();
                q.push_back(hc.col[j]);
            }
    }
    return rank;
}

// Narrow levels: the average BFS level width from the highest-degree vertex
// is at most WG_WIDTH vertices (grids, meshes, long cycles), so the
// multi-launch tiles would be dominated by per-level launch latency.
constexpr int WG_WIDTH = 4096;

bool batched_narrow(DComp &dc) {
    if (dc.depth < 0) {
        const HostComp &hc = dc.host;
        int32_t root = 0;
        for (int32_t v = 1; v < hc.n; ++v)
            if (hc.rp[v + 1] - hc.rp[v] > hc.rp[root + 1] - hc.rp[root]) root = v;
        std::vector<int32_t> lv((size_t)hc.n, -1), q;
        std::vector<double> sg((size_t)hc.n, 0.0);
        q.reserve(hc.n);
        q.push_back(root);
        lv[root] = 0;
        sg[root] = 1.0;
        int32_t depth = 0;
        double smax = 1.0;
        for (size_t i = 0; i < q.size(); ++i) {
            const int32_t v = q[i];
            depth = std::max(depth, lv[v]);
            smax = std::max(smax, sg[v]);
            for (int32_t j = hc.rp[v]; j < hc.rp[v + 1]; ++j) {
                const int32_t w = hc.col[j];
                if (lv[w] < 0) {
                    lv[w] = lv[v] + 1;
                    q.push_back(w);
                }
                if (lv[w] == lv[v] + 1) sg[w] += sg[v];
            }
        }
        dc.depth = depth;
        dc.sigma_est = smax;
    }
    return (int64_t)dc.n <= (int64_t)WG_WIDTH * (dc.depth + 1);
}

// level kernels' blocks per CU unless tuned or set (2-wave blocks): 8 measured
// best on the Chung-Lu and R-MAT cores with 8 tiles in flight (4, 12 and 16
// within 1-6% below)
constexpr int DEFAULT_BPC = 8;

static TileJob make_job(bcx_ctx *ctx, const DComp &dc, const bcx_opts *o, double *bc, int bpc = DEFAULT_BPC) {
    TileJob j;
    j.dc = &dc;
    j.perm = nullptr;
    j.cut = nullptr;
    j.bc = bc;
    j.limit = (o->sigma_f64_threshold_log2 > 0 && o->sigma_f64_threshold_log2 < 64)
                  ? (1ull << o->sigma_f64_threshold_log2) : 0;
    j.fp32 = o->precision == BCX_PREC_FP32C;
    j.grid = (int)std::max<int64_t>(1, std::min<int64_t>(((int64_t)dc.n + 3) / 4, (int64_t)ctx->n_cu * bpc));
    j.graphs = !(o->flags & BCX_FLAG_NO_GRAPH);
    j.spl = 1;
    return j;
}

// Every requested source (the subset, or every vertex) runs as a BFS lane,
// chain vertices included, ordered by BFS rank from the highest-degree
// vertex.  *n_chain (when given) counts the chain vertices among them.
std::vector<int32_t> batched_sources(DComp &dc, const int32_t *subset, int64_t n_sub, int64_t *n_chain) {
    if (dc.rank.empty()) dc.rank = bfs_rank(dc.host);
    std::vector<int32_t> srcs;
    bool bucketed = false;
    if (subset && n_sub * 8 >= dc.n) {
        // a dense subset is put in rank order by bucketing (O(n), no sort:
        // a step of the 10k-vertex workload spent 240 us sorting); a vertex
        // the BFS did not reach (disconnected input) leaves it to the sort
        std::vector<int32_t> at((size_t)dc.n, -1);
        bucketed = true;
        for (int64_t i = 0; i < n_sub && bucketed; ++i) {
            const int32_t r = dc.rank[subset[i]];
            if (r < 0) bucketed = false;
            else at[(size_t)r] = subset[i];
        }
        if (bucketed) {
            srcs.reserve((size_t)n_sub);
            for (int32_t r = 0; r < dc.n; ++r)
                if (at[(size_t)r] >= 0) srcs.push_back(at[(size_t)r]);
        }
    }
    if (subset && !bucketed) {
        srcs.assign(subset, subset + n_sub);
        std::sort(srcs.begin(), srcs.end(), [&](int32_t a, int32_t b) { return dc.rank[a] < dc.rank[b]; });
        srcs.erase(std::unique(srcs.begin(), srcs.end()), srcs.end());
    } else if (!subset) {
        srcs.resize((size_t)dc.n);
        for (int32_t v = 0; v < dc.n; ++v) srcs[(size_t)dc.rank[v]] = v;
    }
    if (n_chain) {
        if (dc.on_chain.empty()) {
            dc.on_chain.assign((size_t)dc.n, 0);
            for (int32_t v : dc.host.ch_nodes) dc.on_chain[v] = 1;
        }
        int64_t c = 0;
        for (int32_t v : srcs) c += dc.on_chain[v];
        *n_chain = c;
    }
    return srcs;
}

// Tiles as graph balls: each tile is the 64 requested sources nearest (BFS
// over the whole component) to the first unassigned source in `srcs`
// order.  On deep, narrow graphs (grids) the sources of a ball are within a
// few hops of each other, so a vertex is reached at few distinct levels per
// tile; consecutive BFS ranks instead lie along one long BFS layer.
std::vector<int32_t> ball_sources(DComp &dc, const std::vector<int32_t> &srcs) {
    const HostComp &hc = dc.host;
    const bool all = (int64_t)srcs.size() == (int64_t)hc.n;
    if (all && !dc.ball.empty()) return dc.ball;
    // each ball explores at most BALL_CAP vertices: once few requested
    // sources are left, an uncapped ball would sweep the whole component;
    // a ball that collects fewer than 64 leaves its tile to the next one
    constexpr size_t BALL_CAP = 16 * BATCH_W;
    std::vector<int8_t> want((size_t)hc.n, 0);
    for (int32_t s : srcs) want[s] = 1;
    std::vector<int32_t> stamp((size_t)hc.n, -1), q, out;
    q.reserve(BALL_CAP);
    out.reserve(srcs.size());
    int32_t ball = 0;
    for (int32_t u : srcs) {
        if (want[u] != 1) continue;
        const size_t need = BATCH_W - out.size() % BATCH_W;  // rest of the current tile
        size_t got = 0;
        q.clear();
        q.push_back(u);
        stamp[u] = ball;
        for (size_t i = 0; i < q.size() && got < need; ++i) {
            const int32_t v = q[i];
            if (want[v] == 1) {
                want[v] = 2;
                out.push_back(v);
                ++got;
            }
            for (int32_t j = hc.rp[v]; j < hc.rp[v + 1] && q.size() < BALL_CAP; ++j)
                if (stamp[hc.col[j]] != ball) {
                    stamp[hc.col[j]] = ball;
                    q.push_back(hc.col[j]);
                }
        }
        ++ball;
    }
    if (all) dc.ball = out;
    return out;
}

// Work of one tile of ns <= 64 sources: its (vertex, level) entries, each
// weighted by the arcs the pull and the backward walk for it (degree + 1),
// from a host 64-lane level-synchronous BFS (the tile kernels' own scheme,
// on the CPU).  Stops early once the count passes `limit`.
static int64_t tile_entries(const HostComp &hc, const int32_t *src, int ns, int64_t limit,
                            std::vector<uint64_t> &vis, std::vector<uint64_t> &cur, std::vector<uint64_t> &nxt) {
    std::vector<int32_t> fr, nf, seen;
    for (int k = 0; k < ns; ++k) {
        const int32_t v = src[k];
        if (!cur[v]) fr.push_back(v);
        cur[v] |= 1ull << k;
        vis[v] |= 1ull << k;
    }
    seen = fr;
    int64_t entries = 0;
    for (int32_t v : fr) entries += hc.rp[v + 1] - hc.rp[v] + 1;
    while (!fr.empty() && entries <= limit) {
        nf.clear();
        for (int32_t v : fr) {
            const uint64_t F = cur[v];
            for (int32_t j = hc.rp[v]; j < hc.rp[v + 1]; ++j) {
                const int32_t w = hc.col[j];
                const uint64_t m = F & ~vis[w];
                if (!m) continue;
                if (!nxt[w]) nf.push_back(w);
                nxt[w] |= m;
            }
        }
        for (int32_t v : fr) cur[v] = 0;
        for (int32_t w : nf) {
            if (!vis[w]) seen.push_back(w);
            vis[w] |= nxt[w];
            cur[w] = nxt[w];
            nxt[w] = 0;
            entries += hc.rp[w + 1] - hc.rp[w] + 1;
        }
        fr.swap(nf);
    }
    for (int32_t v : seen) vis[v] = 0;
    for (int32_t v : fr) cur[v] = 0;
    return entries;
}

// Landmark order: sources sorted by their distances to the LM highest-degree
// vertices (lexicographically, BFS rank breaking ties).  Sources with equal
// distances to every hub tend to have equal distances to everything else,
// so a tile of them reaches a vertex at fewer distinct levels.  The LM BFS
// passes run on host threads.
constexpr int LM = 16;
static std::vector<int32_t> landmark_order(const HostComp &hc, const std::vector<int32_t> &rank) {
    const int32_t n = hc.n;
    const int K = std::min<int32_t>(LM, n);
    std::vector<int32_t> hubs((size_t)n);
    for (int32_t v = 0; v < n; ++v) hubs[v] = v;
    std::partial_sort(hubs.begin(), hubs.begin() + K, hubs.end(), [&](int32_t a, int32_t b) {
        const int32_t da = hc.rp[a + 1] - hc.rp[a], db = hc.rp[b + 1] - hc.rp[b];
        return da != db ? da > db : a < b;
    });
    // dist[v][k], saturated at 65535
    std::vector<std::array<uint16_t, LM>> dist((size_t)n);
    auto bfs = [&](int k) {
        std::vector<int32_t> d((size_t)n, -1), q;
        q.reserve(n);
        q.push_back(hubs[k]);
        d[hubs[k]] = 0;
        for (size_t i = 0; i < q.size(); ++i) {
            const int32_t v = q[i];
            for (int32_t j = hc.rp[v]; j < hc.rp[v + 1]; ++j)
                if (d[hc.col[j]] < 0) {
                    d[hc.col[j]] = d[v] + 1;
                    q.push_back(hc.col[j]);
                }
        }
        for (int32_t v = 0; v < n; ++v) dist[v][k] = (uint16_t)std::min(d[v] < 0 ? 65535 : d[v], 65535);
    };
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < hw; ++t)
        th.emplace_back([&, t] {
            for (int k = (int)t; k < K; k += (int)hw) bfs(k);
        });
    for (auto &x : th) x.join();
    for (int32_t v = 0; v < n; ++v)
        for (int k = K; k < LM; ++k) dist[v][k] = 0;
    std::vector<int32_t> order((size_t)n);
    for (int32_t v = 0; v < n; ++v) order[(size_t)rank[v]] = v;
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t a, int32_t b) { return dist[a] < dist[b]; });  // ties keep BFS rank
    return order;
}

// Work of the tiles of `order` at nine evenly spaced positions, on host
// threads (the work of single tiles varies by tens of percent).
static int64_t sample_work(const HostComp &hc, const std::vector<int32_t> &order) {
    const int64_t T = ((int64_t)hc.n + BATCH_W - 1) / BATCH_W;
    constexpr int S = 9;
    std::vector<int64_t> w(S, 0);
    const unsigned nt = std::max(1u, std::min<unsigned>(S, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
            std::vector<uint64_t> vis((size_t)hc.n, 0), cur((size_t)hc.n, 0), nxt((size_t)hc.n, 0);
            for (int k = (int)t; k < S; k += (int)nt) {
                const int64_t u = (k + 1) * T / (S + 1);
                const int ns = (int)std::min<int64_t>(BATCH_W, hc.n - u * BATCH_W);
                w[k] = tile_entries(hc, order.data() + u * BATCH_W, ns, INT64_MAX, vis, cur, nxt);
            }
        });
    for (auto &x : th) x.join();
    int64_t sum = 0;
    for (int64_t x : w) sum += x;
    return sum;
}

// Tile-order tuner: on first use, three orders of all the component's
// vertices are compared by the work of sample tiles (see tile_entries).
// BFS rank against a ball: the 64 consecutive ranks in the middle of the
// order against the 64 vertices nearest to that tile's first source; the
// ball order is kept when it needs at least 10% less work (ball tiles are
// built per source set and cost host time; on the 256x256 grid balls need
// 13.8 entries per vertex per tile against 24.4).  Otherwise BFS rank
// against the landmark order, on nine tiles of each: landmarks are kept when
// they need at least 2% less.  Arcs walked per arc per tile on nine tiles
// (BFS rank / landmarks): 1.98 / 1.89 on the 835k-vertex core of the
// 1M-vertex mixed graph, 1.68 / 1.61 on an R-MAT core with chains, 1.82 /
// 1.88 on a 519k-vertex Chung-Lu core without chains.  The choice is cached
// and applied to every later source set of the component.
std::vector<int32_t> tuned_sources(DComp &dc, const std::vector<int32_t> &srcs) {
    if (dc.tile_order < 0) {
        const HostComp &hc = dc.host;
        const int32_t n = hc.n;
        const int64_t T = ((int64_t)n + BATCH_W - 1) / BATCH_W;
        dc.tile_order = 0;
        if (T >= 16) {
            if (dc.rank.empty()) dc.rank = bfs_rank(hc);
            std::vector<int32_t> by_rank((size_t)n);
            for (int32_t v = 0; v < n; ++v) by_rank[(size_t)dc.rank[v]] = v;
            std::vector<uint64_t> vis((size_t)n, 0), cur((size_t)n, 0), nxt((size_t)n, 0);
            std::vector<int32_t> stamp((size_t)n, -1), q;
            const int32_t *tile = by_rank.data() + (T / 2) * BATCH_W;
            const int64_t e_rank = tile_entries(hc, tile, BATCH_W, INT64_MAX, vis, cur, nxt);
            // the 64 vertices nearest to the tile's first source
            q.assign(1, tile[0]);
            stamp[tile[0]] = 0;
            for (size_t i = 0; i < q.size() && q.size() < (size_t)BATCH_W; ++i)
                for (int32_t j = hc.rp[q[i]]; j < hc.rp[q[i] + 1] && q.size() < (size_t)BATCH_W; ++j)
                    if (stamp[hc.col[j]] != 0) {
                        stamp[hc.col[j]] = 0;
                        q.push_back(hc.col[j]);
                    }
            const int64_t e_ball = tile_entries(hc, q.data(), (int)q.size(), e_rank, vis, cur, nxt);
            if (10 * e_ball <= 9 * e_rank) {
                dc.tile_order = 1;
            } else {
                const std::vector<int32_t> by_lm = landmark_order(hc, dc.rank);
                if (50 * sample_work(hc, by_lm) <= 49 * sample_work(hc, by_rank)) {
                    dc.tile_order = 2;
                    dc.lm_pos.assign((size_t)n, 0);
                    for (int32_t i = 0; i < n; ++i) dc.lm_pos[(size_t)by_lm[i]] = i;
                }
            }
        }
    }
    if (dc.tile_order == 1) return ball_sources(dc, srcs);
    if (dc.tile_order == 2) {
        std::vector<int32_t> o(srcs);
        std::sort(o.begin(), o.end(), [&](int32_t a, int32_t b) { return dc.lm_pos[a] < dc.lm_pos[b]; });
        return o;
    }
    return srcs;
}

std::vector<int32_t> batched_order(DComp &dc) {
    std::vector<int32_t> all = batched_sources(dc, nullptr, 0, nullptr);
    return dc.n <= SMALL_N ? all : tuned_sources(dc, all);
}

// The component relabelled in BFS order from its highest-degree vertex: the
// vertices a tile reaches at one level lie in a few consecutive BFS layers
// of that vertex (a tile's sources are close to each other), so each level's
// masks, sigma / term rows and list entries fall in a compact id range, and
// the hub itself and its neighbours lead the arrays.  Each row keeps its arcs
// in the original order, so the FP64 sigma sums keep the oracle's order.
static int relabelled(bcx_ctx *ctx, DComp &dc, DComp *&out) {
    if (dc.relabel) { out = dc.relabel; return BCX_OK; }
    const HostComp &h = dc.host;
    const int32_t n = h.n;
    std::vector<int32_t> perm((size_t)n), inv((size_t)n);
    if (dc.rank.empty()) dc.rank = bfs_rank(h);
    for (int32_t v = 0; v < n; ++v) perm[(size_t)dc.rank[v]] = v;
    for (int32_t i = 0; i < n; ++i) inv[perm[i]] = i;
    DComp *p = new DComp();
    p->n = n;
    p->arcs = dc.arcs;
    p->offset = dc.offset;
    HostComp &q = p->host;
    q.n = n;
    q.arcs = h.arcs;
    q.rp.assign((size_t)n + 1, 0);
    q.col.resize(h.col.size());
    q.reach.resize((size_t)n);
    for (int32_t i = 0; i < n; ++i) {
        const int32_t v = perm[i], a = h.rp[v], d = h.rp[v + 1] - a;
        q.rp[i + 1] = q.rp[i] + d;
        for (int32_t k = 0; k < d; ++k) q.col[q.rp[i] + k] = inv[h.col[a + k]];
        q.reach[i] = h.reach[v];
    }
    q.ch_ptr.assign(1, 0);
    int rc;
    if ((rc = upload(ctx, q.rp.data(), 4 * q.rp.size(), p->b_rp, "relabelled component")) ||
        (rc = upload(ctx, q.col.data(), 4 * q.col.size(), p->b_col, "relabelled component")) ||
        (rc = upload(ctx, q.reach.data(), 4 * q.reach.size(), p->b_reach, "relabelled component")) ||
        (rc = upload(ctx, perm.data(), 4 * perm.size(), dc.b_perm, "relabelling")) ||
        (rc = ctx_hip(ctx, hipStreamSynchronize(ctx->stream), "relabel"))) {
        for (DBuf *b : {&p->b_rp, &p->b_col, &p->b_reach}) b->release(ctx);
        delete p;
        return rc;
    }
    p->dev.n = n;
    p->dev.arcs = p->arcs;
    p->dev.rp = (const int32_t *)p->b_rp.p;
    p->dev.col = (const int32_t *)p->b_col.p;
    p->dev.reach = (const int32_t *)p->b_reach.p;
    dc.relabel_inv.swap(inv);
    dc.relabel = p;
    out = p;
    return BCX_OK;
}

// ------------------------------------------------------------------------
// Contracted chains.  A chain of K arcs between L and R is passive when a
// path of fewer than K arcs joins L and R outside every chain: no shortest
// path between two vertices off the chain then runs through it (the path
// around it is shorter), so its interiors are reached only from its two ends
// and their levels, counts and dependencies follow from the ends' in closed
// form.  The bound is min over landmarks h of d(h, L) + d(h, R), with
// d the BFS distance over the chain-free part of the component from its
// CUT_LM highest-degree vertices; chains with L == R are kept.
// ------------------------------------------------------------------------
constexpr int CUT_LM = 8;

struct bcx::ChainCut {
    DComp g;                   // the copy: [0, n_core) the tile graph in BFS order, then the interiors
    int32_t n_core = 0, n_chain = 0, n_ep = 0;
    std::vector<int32_t> inv;  // component id -> copy id
    DBuf b_perm;               // copy id -> component id
    DBuf b_tab;                // the chain tables, one allocation
    const int32_t *ch_L = nullptr, *ch_R = nullptr, *ch_off = nullptr, *ch_of = nullptr;
    const int32_t *ep_idx = nullptr, *ep_off = nullptr, *ep_list = nullptr;
    const double *ch_pre = nullptr;
};

static void cut_release(bcx_ctx *ctx, DComp &dc) {
    if (!dc.cut) return;
    for (DBuf *b : {&dc.cut->g.b_rp, &dc.cut->g.b_col, &dc.cut->g.b_reach, &dc.cut->b_perm, &dc.cut->b_tab})
        b->release(ctx);
    delete dc.cut;
    dc.cut = nullptr;
}

// Chain ids (into the component's chain tables) of the passive chains.
static std::vector<int32_t> passive_chains(const HostComp &h, const std::vector<uint8_t> &inner) {
    const int32_t n = h.n;
    std::vector<int32_t> lm;
    for (int32_t v = 0; v < n; ++v)
        if (!inner[v]) lm.push_back(v);
    const int K = (int)std::min<size_t>(CUT_LM, lm.size());
    std::partial_sort(lm.begin(), lm.begin() + K, lm.end(), [&](int32_t a, int32_t b) {
        const int32_t da = h.rp[a + 1] - h.rp[a], db = h.rp[b + 1] - h.rp[b];
        return da != db ? da > db : a < b;
    });
    std::vector<std::vector<int32_t>> dist((size_t)K);
    auto bfs = [&](int k) {
        std::vector<int32_t> &d = dist[(size_t)k];
        d.assign((size_t)n, -1);
        std::vector<int32_t> q;
        q.reserve(n);
        q.push_back(lm[(size_t)k]);
        d[lm[(size_t)k]] = 0;
        for (size_t i = 0; i < q.size(); ++i) {
            const int32_t v = q[i];
            for (int32_t j = h.rp[v]; j < h.rp[v + 1]; ++j) {
                const int32_t w = h.col[j];
                if (!inner[w] && d[w] < 0) {
                    d[w] = d[v] + 1;
                    q.push_back(w);
                }
            }
        }
    };
    std::vector<std::thread> th;
    for (int k = 0; k < K; ++k) th.emplace_back(bfs, k);
    for (auto &x : th) x.join();
    std::vector<int32_t> out;
    for (int32_t c = 0; c < h.n_chains; ++c) {
        const int32_t L = h.ch_ends[2 * c], R = h.ch_ends[2 * c + 1];
        const int64_t arcs = (int64_t)h.ch_ptr[c + 1] - h.ch_ptr[c] + 1;
        if (L == R) continue;
        int64_t ub = INT64_MAX;
        for (int k = 0; k < K; ++k)
            if (dist[k][L] >= 0 && dist[k][R] >= 0) ub = std::min<int64_t>(ub, (int64_t)dist[k][L] + dist[k][R]);
        if (arcs > ub) out.push_back(c);
    }
    return out;
}

// The copy of a component with its passive chains contracted (nullptr when
// their interiors are fewer than 1% of the vertices: the plain relabelled
// copy is used then).  Tile vertices keep the BFS order of the relabelled
// copy; each endpoint row keeps its arcs in the original order with the arc
// into a contracted chain c encoded as ~c.
static int chain_cut(bcx_ctx *ctx, DComp &dc, ChainCut *&out) {
    out = nullptr;
    if (dc.cut_state == 1) { out = dc.cut; return BCX_OK; }
    if (dc.cut_state == 0) return BCX_OK;
    const HostComp &h = dc.host;
    const int32_t n = h.n;
    if (h.n_chains == 0) { dc.cut_state = 0; return BCX_OK; }
    std::vector<uint8_t> inner((size_t)n, 0);
    for (int32_t x : h.ch_nodes) inner[x] = 1;
    const std::vector<int32_t> pc = passive_chains(h, inner);
    int64_t n_in = 0;
    for (int32_t c : pc) n_in += h.ch_ptr[c + 1] - h.ch_ptr[c];
    if (n_in == 0 || n_in * 100 < n) { dc.cut_state = 0; return BCX_OK; }
    const int32_t npc = (int32_t)pc.size();
    std::vector<int32_t> cidx((size_t)n, -1);  // contracted chain of each interior
    std::vector<int32_t> off((size_t)npc + 1, 0);
    for (int32_t k = 0; k < npc; ++k) {
        const int32_t c = pc[(size_t)k];
        for (int32_t i = h.ch_ptr[c]; i < h.ch_ptr[c + 1]; ++i) cidx[h.ch_nodes[i]] = k;
        off[(size_t)k + 1] = off[(size_t)k] + (h.ch_ptr[c + 1] - h.ch_ptr[c]);
    }
    if (dc.rank.empty()) dc.rank = bfs_rank(h);
    std::vector<int32_t> byrank((size_t)n), perm((size_t)n), inv((size_t)n);
    for (int32_t v = 0; v < n; ++v) byrank[(size_t)dc.rank[v]] = v;
    int32_t id = 0;
    for (int32_t r = 0; r < n; ++r) {
        const int32_t v = byrank[(size_t)r];
        if (cidx[v] < 0) { inv[v] = id; perm[(size_t)id++] = v; }
    }
    const int32_t n_core = id;
    for (int32_t k = 0; k < npc; ++k) {
        const int32_t c = pc[(size_t)k];
        for (int32_t i = h.ch_ptr[c]; i < h.ch_ptr[c + 1]; ++i) {
            const int32_t x = h.ch_nodes[i];
            inv[x] = n_core + off[(size_t)k] + (i - h.ch_ptr[c]);
            perm[(size_t)inv[x]] = x;
        }
    }
    ChainCut *cut = new ChainCut();
    cut->n_core = n_core;
    cut->n_chain = npc;
    HostComp &q = cut->g.host;
    q.n = n;
    q.rp.assign((size_t)n + 1, 0);
    q.reach.resize((size_t)n);
    for (int32_t i = 0; i < n; ++i) {
        const int32_t v = perm[(size_t)i];
        q.reach[(size_t)i] = h.reach[v];
        q.rp[(size_t)i + 1] = q.rp[(size_t)i];
        if (i >= n_core) continue;  // contracted interiors have no arcs
        for (int32_t j = h.rp[v]; j < h.rp[v + 1]; ++j) {
            const int32_t w = h.col[j];
            q.col.push_back(cidx[w] >= 0 ? ~cidx[w] : inv[w]);
        }
        q.rp[(size_t)i + 1] = (int32_t)q.col.size();
    }
    q.arcs = (int64_t)q.col.size();
    q.ch_ptr.assign(1, 0);
    // chain tables
    std::vector<int32_t> chL((size_t)npc), chR((size_t)npc), chof((size_t)n_in), epidx((size_t)n_core, -1);
    std::vector<double> pre((size_t)n_in);
    for (int32_t k = 0; k < npc; ++k) {
        const int32_t c = pc[(size_t)k];
        chL[(size_t)k] = inv[h.ch_ends[2 * c]];
        chR[(size_t)k] = inv[h.ch_ends[2 * c + 1]];
        double acc = 0.0;
        for (int32_t i = h.ch_ptr[c]; i < h.ch_ptr[c + 1]; ++i) {
            const int32_t x = off[(size_t)k] + (i - h.ch_ptr[c]);
            acc += (double)h.reach[h.ch_nodes[i]] + 1.0;
            pre[(size_t)x] = acc;
            chof[(size_t)x] = k;
        }
        epidx[chL[(size_t)k]] = 0;
        epidx[chR[(size_t)k]] = 0;
    }
    int32_t n_ep = 0;
    for (int32_t v = 0; v < n_core; ++v)
        if (epidx[v] == 0) epidx[v] = ++n_ep;  // (numbered from 1, shifted below)
    for (int32_t v = 0; v < n_core; ++v)
        if (epidx[v] > 0) epidx[v]--;
        else epidx[v] = -1;
    std::vector<int32_t> epoff((size_t)n_ep + 1, 0), eplist((size_t)2 * npc);
    for (int32_t k = 0; k < npc; ++k) {
        epoff[(size_t)epidx[chL[(size_t)k]] + 1]++;
        epoff[(size_t)epidx[chR[(size_t)k]] + 1]++;
    }
    for (int32_t e = 0; e < n_ep; ++e) epoff[(size_t)e + 1] += epoff[(size_t)e];
    {
        std::vector<int32_t> fill(epoff.begin(), epoff.end() - 1);
        for (int32_t k = 0; k < npc; ++k) {
            eplist[(size_t)fill[(size_t)epidx[chL[(size_t)k]]]++] = 2 * k;
            eplist[(size_t)fill[(size_t)epidx[chR[(size_t)k]]]++] = 2 * k + 1;
        }
    }
    cut->n_ep = n_ep;
    // one allocation: the prefix sums (8-byte aligned at the start), then the int tables
    const size_t ints = 2 * (size_t)npc + ((size_t)npc + 1) + (size_t)n_in + (size_t)n_core + ((size_t)n_ep + 1) +
                        2 * (size_t)npc;
    std::vector<unsigned char> tab(8 * (size_t)n_in + 4 * ints);
    std::memcpy(tab.data(), pre.data(), 8 * (size_t)n_in);
    size_t at = 8 * (size_t)n_in;
    std::vector<size_t> pos;
    for (const std::vector<int32_t> *t : {&chL, &chR, &off, &chof, &epidx, &epoff, &eplist}) {
        pos.push_back(at);
        std::memcpy(tab.data() + at, t->data(), 4 * t->size());
        at += 4 * t->size();
    }
    int rc;
    if ((rc = upload(ctx, q.rp.data(), 4 * q.rp.size(), cut->g.b_rp, "contracted component")) ||
        (rc = upload(ctx, q.col.data(), 4 * q.col.size(), cut->g.b_col, "contracted component")) ||
        (rc = upload(ctx, q.reach.data(), 4 * q.reach.size(), cut->g.b_reach, "contracted component")) ||
        (rc = upload(ctx, perm.data(), 4 * perm.size(), cut->b_perm, "contracted relabelling")) ||
        (rc = upload(ctx, tab.data(), tab.size(), cut->b_tab, "chain tables")) ||
        (rc = ctx_hip(ctx, hipStreamSynchronize(ctx->stream), "chain tables"))) {
        dc.cut = cut;
        cut_release(ctx, dc);
        return rc;
    }
    const unsigned char *base = (const unsigned char *)cut->b_tab.p;
    cut->ch_pre = (const double *)base;
    cut->ch_L = (const int32_t *)(base + pos[0]);
    cut->ch_R = (const int32_t *)(base + pos[1]);
    cut->ch_off = (const int32_t *)(base + pos[2]);
    cut->ch_of = (const int32_t *)(base + pos[3]);
    cut->ep_idx = (const int32_t *)(base + pos[4]);
    cut->ep_off = (const int32_t *)(base + pos[5]);
    cut->ep_list = (const int32_t *)(base + pos[6]);
    DComp &g = cut->g;
    g.n = n;
    g.arcs = q.arcs;
    g.offset = dc.offset;
    g.dev.n = n;
    g.dev.arcs = g.arcs;
    g.dev.rp = (const int32_t *)g.b_rp.p;
    g.dev.col = (const int32_t *)g.b_col.p;
    g.dev.reach = (const int32_t *)g.b_reach.p;
    cut->inv.swap(inv);
    dc.cut = cut;
    dc.cut_state = 1;
    out = cut;
    return BCX_OK;
}

static void set_chains(BatchDev &b, Slot &s, const ChainCut *cut) {
    if (!cut) return;  // (make_dev starts from a zeroed view: n_chain 0)
    b.n = cut->n_core;
    b.n_chain = cut->n_chain;
    b.ch_L = cut->ch_L;
    b.ch_R = cut->ch_R;
    b.ch_off = cut->ch_off;
    b.ch_pre = cut->ch_pre;
    b.ch_of = cut->ch_of;
    b.ep_idx = cut->ep_idx;
    b.ep_off = cut->ep_off;
    b.ep_list = cut->ep_list;
    b.n_ep = cut->n_ep;
    b.ep_lvl = (int32_t *)s.eplv.p;
    b.ep_rho = (double *)s.eprho.p;
    b.lane_ch = (int32_t *)s.lanes.p;
    b.lane_mask = (uint64_t *)((char *)s.lanes.p + 2 * 64 * MAX_SPL * sizeof(int32_t));
}

// Per-slot buffers of a run with contracted chains.
static int cut_slots(bcx_ctx *ctx, const ChainCut *cut) {
    BatchedState &bs = *ctx->batched;
    int rc;
    for (int i = 0; i < bs.n_slots; ++i) {
        Slot &s = bs.slots[i];
        if ((rc = s.eplv.ensure(ctx, 4 * 64 * MAX_SPL * (size_t)std::max(cut->n_ep, 1), "endpoint levels")) ||
            (rc = s.eprho.ensure(ctx, 8 * 64 * MAX_SPL * (size_t)std::max(cut->n_ep, 1), "endpoint chain terms")) ||
            (rc = s.lanes.ensure(ctx, (2 * 64 * sizeof(int32_t) + sizeof(uint64_t)) * MAX_SPL, "lane chains")))
            return rc;
    }
    return BCX_OK;
}

static bool want_cut(const bcx_opts *o) { return o->use_chains && !(o->flags & BCX_FLAG_NO_CONTRACT); }

static int tile_graph(bcx_ctx *ctx, DComp &dc, bool want, DComp *&pc, ChainCut *&cut,
                      const std::vector<int32_t> *&inv, const int32_t *&perm);

int batched_prepare(bcx_ctx *ctx, DComp &dc, const bcx_opts *o) {
    DComp *pc = nullptr;
    ChainCut *cut = nullptr;
    const std::vector<int32_t> *inv = nullptr;
    const int32_t *perm = nullptr;
    return tile_graph(ctx, dc, want_cut(o) && dc.sched_cut != 0, pc, cut, inv, perm);
}

// The tile graph of a multi-launch component: the copy with contracted
// chains when asked for and the component has enough of them, else the
// relabelled copy.  inv maps component ids to the copy's.
static int tile_graph(bcx_ctx *ctx, DComp &dc, bool want, DComp *&pc, ChainCut *&cut,
                      const std::vector<int32_t> *&inv, const int32_t *&perm) {
    cut = nullptr;
    int rc = want ? chain_cut(ctx, dc, cut) : BCX_OK;
    if (rc) return rc;
    if (cut) {
        pc = &cut->g;
        inv = &cut->inv;
        perm = (const int32_t *)cut->b_perm.p;
        return BCX_OK;
    }
    if ((rc = relabelled(ctx, dc, pc))) return rc;
    inv = &dc.relabel_inv;
    perm = (const int32_t *)dc.b_perm.p;
    return BCX_OK;
}

// Schedule of a multi-launch component: what the options fix, else what was
// measured for it, else the defaults.
struct MultiSched {
    bool cut;
    int bpc, slots;
    int spl;  // sources per lane
};

// sources per lane unless tuned or fixed: four (256-source tiles) measured
// above two and two above one on the Chung-Lu and R-MAT cores (fewer entries
// and arc walks per source: a tile reaches a vertex at few more distinct
// levels): 165.0k / 147.6k / 126.0k MTEPS on the 1M-vertex mixed graph,
// 276.9k / 244.1k / 214.2k on R-MAT scale 20
constexpr int DEFAULT_SPL = 4;

static bool valid_spl(int w) { return w == 1 || w == 2 || w == 4; }

// Sources per lane of the workgroup kernel's tiles (bcx_opts.sources_per_lane
// when valid).  1: on the 256x256 grid workload, tiles of 128 and 256 sources
// ran 6.6k and 1.5k MTEPS against 10.2k for 64 (a deep tile's work grows with
// its width faster than its levels are shared, and 256-source tiles' lists
// leave room for fewer workgroups).
constexpr int WG_DEFAULT_SPL = 1;
static int wg_spl(const bcx_opts *o) { return valid_spl(o->sources_per_lane) ? o->sources_per_lane : WG_DEFAULT_SPL; }

static MultiSched multi_sched(const DComp &dc, const bcx_opts *o) {
    MultiSched m;
    m.cut = want_cut(o) && dc.sched_cut != 0;
    m.spl = valid_spl(o->sources_per_lane) ? o->sources_per_lane : (dc.sched_spl > 0 ? dc.sched_spl : DEFAULT_SPL);
    m.bpc = o->blocks_per_cu > 0 ? std::min(o->blocks_per_cu, 64) : (dc.sched_bpc > 0 ? dc.sched_bpc : DEFAULT_BPC);
    m.slots = o->batch_width > 0 ? std::max(1, std::min(MAX_SLOTS, o->batch_width))
                                 : (dc.sched_slots > 0 ? dc.sched_slots : MAX_SLOTS);
    return m;
}

static int multi_run(bcx_ctx *ctx, DComp &dc, const std::vector<int32_t> &srcs, const bcx_opts *o,
                     const MultiSched &m, double *bc, bcx_stats *st) {
    DComp *pc = nullptr;
    ChainCut *cut = nullptr;
    const std::vector<int32_t> *inv = nullptr;
    const int32_t *perm = nullptr;
    int rc = tile_graph(ctx, dc, m.cut, pc, cut, inv, perm);
    if (rc) return rc;
    std::vector<int32_t> ps(srcs.size());
    for (size_t i = 0; i < srcs.size(); ++i) ps[i] = (*inv)[srcs[i]];
    const int tw = BATCH_W * m.spl;
    const int64_t ntiles = ((int64_t)ps.size() + tw - 1) / tw;
    rc = bs_ensure(ctx, pc->n, (int32_t)(ntiles * tw), m.slots, false);
    if (rc) return rc;
    if (cut && (rc = cut_slots(ctx, cut))) return rc;
    if ((rc = stage_sources(ctx, *pc, ps, tw)) || (rc = stage_hubs(ctx, *pc))) return rc;
    TileJob j = make_job(ctx, *pc, o, bc, m.bpc);
    j.perm = perm;
    j.cut = cut;
    j.spl = m.spl;
    if (st && cut) st->n_contracted += pc->n - cut->n_core;
    return run_tiles(ctx, *pc, (int64_t)ps.size(), j, st, nullptr);
}

int batched_component(bcx_ctx *ctx, DComp &dc, const std::vector<int32_t> &srcs, const bcx_opts *o, double *bc,
                      bcx_stats *st) {
    return multi_run(ctx, dc, srcs, o, multi_sched(dc, o), bc, st);
}

void batched_sched_report(const DComp &dc, const bcx_opts *o, int32_t *bpc, int32_t *slots, int32_t *spl) {
    const MultiSched m = multi_sched(dc, o);
    *bpc = m.bpc;
    *slots = m.slots;
    *spl = m.spl;
}

// ------------------------------------------------------------------------
// small components: one persistent launch over (component, 64-source) tiles
// ------------------------------------------------------------------------
struct SmallState {
    // tables: a pass's components, tiles, sources and weights (one upload);
    // tiles: the tiles of an FP64 rerun
    DBuf tables, tiles, counter, redo, sig, tt, Q, Qm, masks, qoffs, qr, ro, tag, lvl, dbg, psig;
    unsigned char *pin = nullptr;  // pinned host staging of the tables
    size_t pin_bytes = 0;
    int grid = 0;
    // a pass started and not yet finished (persistent_start / _finish)
    bool pending = false, run_fp32 = false, run_graph = false;
    SmallArgs run{};
    std::vector<SmallTile> run_tiles;
    size_t run_narrow = 0;
    int run_grid = 0;
    hipStream_t run_stream = nullptr;
    // captured launch (counter reset kernel + persistent kernel), replayed
    // while its arguments stay the same
    hipGraphExec_t exec = nullptr;
    SmallArgs key{};
    int key_grid = 0;
    bool key_f64 = false, key_fp32 = false;
};

// Reset + persistent kernel, replayed from a captured graph when possible.
static hipError_t launch_pass(const SmallArgs &a, bool wg, bool f64, bool fp32, int grid, hipStream_t sp) {
    return wg ? launch_wg(a, f64, fp32, grid, sp) : launch_small(a, f64, fp32, grid, sp);
}

static int small_launch(bcx_ctx *ctx, SmallState &ss, const SmallArgs &a, bool wg, bool f64, bool fp32, int grid,
                        bool use_graph, hipStream_t sp) {
    if (!use_graph) {
        int rc = ctx_hip(ctx, launch_small_reset(a.counter, sp), "persistent reset");
        return rc ? rc : ctx_hip(ctx, launch_pass(a, wg, f64, fp32, grid, sp), "persistent kernel");
    }
    const bool same = ss.exec && !std::memcmp(&ss.key, &a, sizeof(a)) && ss.key_grid == grid &&
                      ss.key_f64 == f64 && ss.key_fp32 == fp32;
    if (!same) {
        if (ss.exec) (void)hipGraphExecDestroy(ss.exec);
        ss.exec = nullptr;
        hipGraph_t graph = nullptr;
        int rc = ctx_hip(ctx, hipStreamBeginCapture(sp, hipStreamCaptureModeThreadLocal), "capture");
        if (rc) return rc;
        hipError_t e1 = launch_small_reset(a.counter, sp);
        hipError_t e2 = launch_pass(a, wg, f64, fp32, grid, sp);
        hipError_t e3 = hipStreamEndCapture(sp, &graph);
        if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return ctx_hip(ctx, e1 != hipSuccess ? e1 : (e2 != hipSuccess ? e2 : e3), "capture");
        }
        hipError_t e4 = hipGraphInstantiate(&ss.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e4 != hipSuccess) { ss.exec = nullptr; return ctx_hip(ctx, e4, "graph instantiate"); }
        std::memcpy(&ss.key, &a, sizeof(a));
        ss.key_grid = grid;
        ss.key_f64 = f64;
        ss.key_fp32 = fp32;
    }
    return ctx_hip(ctx, hipGraphLaunch(ss.exec, sp), "graph launch");
}

// Orders stream `sp` after everything enqueued so far on the context stream
// (uploads of the pass's tables).
static int follow_ctx_stream(bcx_ctx *ctx, hipStream_t sp) {
    if (sp == ctx->stream) return BCX_OK;
    hipEvent_t ev;
    int rc = ctx_hip(ctx, hipEventCreateWithFlags(&ev, hipEventDisableTiming), "event");
    if (rc) return rc;
    rc = ctx_hip(ctx, hipEventRecord(ev, ctx->stream), "event");
    if (!rc) rc = ctx_hip(ctx, hipStreamWaitEvent(sp, ev, 0), "event");
    (void)hipEventDestroy(ev);
    return rc;
}

static int persistent_pass(bcx_ctx *ctx, SmallState &ss, bool wg, bool f64, bool graph) {
    return small_launch(ctx, ss, ss.run, wg, f64, ss.run_fp32, ss.run_grid, graph, ss.run_stream);
}

// One persistent launch over the tiles of `jobs`: the small-component kernel
// (LDS state, n <= SMALL_N) or, with `wg`, the workgroup-per-tile kernel
// (HBM slabs, n <= WG_N).  Tiles whose uint64 sigma overflows are rerun in
// FP64 by a second launch.
// Starts the pass on stream sp and returns without waiting for it;
// persistent_finish waits, reruns the overflowed tiles and completes it.
static int persistent_start(bcx_ctx *ctx, const std::vector<SmallJob> &jobs, const bcx_opts *o, double *bc_base,
                            bcx_stats *st, bool wg, hipStream_t sp, bool probe) {
    if (jobs.empty()) return BCX_OK;
    if (!ctx->batched) ctx->batched = new BatchedState();
    SmallState *&state = wg ? ctx->wgk : ctx->small;
    if (!state) state = new SmallState();
    SmallState &ss = *state;
    int64_t nmax = 1, csr = 0;
    for (const SmallJob &jb : jobs) {
        nmax = std::max<int64_t>(nmax, jb.dc->n);
        csr = std::max<int64_t>(csr, jb.dc->n + 1 + jb.dc->arcs);
    }
    // small kernel: a tile's CSR joins the LDS state when it fits beside the
    // masks
    csr = wg ? 0 : small_csr_words(nmax, csr);
    std::vector<SmallComp> comps;
    std::vector<SmallTile> tiles;
    std::vector<int32_t> srcs;
    std::vector<double> ws;
    // workgroup kernel: components whose path counts from their highest-degree
    // vertex already pass 2^62 start with FP64 sigma (their tiles would all
    // overflow the 64-bit pass); their tiles go after the others
    std::vector<const SmallJob *> order;
    for (const SmallJob &jb : jobs) order.push_back(&jb);
    auto wide = [&](const SmallJob *jb) { return wg && jb->dc->sigma_est > 4.6e18; };
    std::stable_partition(order.begin(), order.end(), [&](const SmallJob *jb) { return !wide(jb); });
    // tiles of 64 W sources (source k + 64 h on lane k); the small kernel's are 64
    const int W = wg ? wg_spl(o) : 1;
    const size_t TW = (size_t)BATCH_W * W;
    size_t n_narrow_tiles = 0;
    for (const SmallJob *jbp : order) {
        const SmallJob &jb = *jbp;
        const DComp &dc = *jb.dc;
        if (!wide(jbp)) n_narrow_tiles += (jb.srcs.size() + TW - 1) / TW;
        // (the workgroup kernel stays on the component's own ids: on the
        // BFS-ordered copy the 256x256 grid workload fetched 32% more bytes
        // and ran 8% slower)
        SmallComp c{};
        c.n = dc.n;
        c.rp = dc.dev.rp;
        c.col = dc.dev.col;
        c.reach = dc.dev.reach;
        c.bc = bc_base + dc.offset;
        const int32_t ci = (int32_t)comps.size();
        comps.push_back(c);
        for (size_t i = 0; i < jb.srcs.size(); i += TW) {
            SmallTile t{};
            t.comp = ci;
            t.src0 = (int32_t)srcs.size();
            t.ns = (int32_t)std::min<size_t>(TW, jb.srcs.size() - i);
            tiles.push_back(t);
            for (int k = 0; k < t.ns; ++k) {
                srcs.push_back(jb.srcs[i + k]);
                ws.push_back((double)dc.host.reach[jb.srcs[i + k]] + 1.0);
            }
        }
    }
    // the four tables go up in one copy from pinned staging (the previous
    // pass that used the staging has been waited for)
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t o_tiles = al(sizeof(SmallComp) * comps.size());
    const size_t o_srcs = o_tiles + al(sizeof(SmallTile) * tiles.size());
    const size_t o_ws = o_srcs + al(4 * srcs.size());
    const size_t t_bytes = o_ws + al(8 * ws.size());
    int rc;
    if (ss.pin_bytes < t_bytes) {
        if (ss.pin) (void)hipHostFree(ss.pin);
        ss.pin = nullptr;
        ss.pin_bytes = 0;
        if (hipHostMalloc((void **)&ss.pin, t_bytes, hipHostMallocDefault) == hipSuccess) {
            ss.pin_bytes = t_bytes;
        } else {
            // no pinned memory (a low locked-memory limit): a pageable copy,
            // which HIP stages before returning
            (void)hipGetLastError();
            ss.pin = nullptr;
        }
    }
    std::vector<unsigned char> pageable;
    unsigned char *stage = ss.pin;
    if (!stage) {
        pageable.resize(t_bytes);
        stage = pageable.data();
    }
    std::memcpy(stage, comps.data(), sizeof(SmallComp) * comps.size());
    std::memcpy(stage + o_tiles, tiles.data(), sizeof(SmallTile) * tiles.size());
    std::memcpy(stage + o_srcs, srcs.data(), 4 * srcs.size());
    std::memcpy(stage + o_ws, ws.data(), 8 * ws.size());
    if ((rc = ss.tables.ensure(ctx, t_bytes, "small tables")) ||
        (rc = ctx_hip(ctx, hipMemcpyAsync(ss.tables.p, stage, t_bytes, hipMemcpyHostToDevice, ctx->stream),
                      "small tables")) ||
        (rc = ss.counter.ensure(ctx, 64, "small counters")) ||
        (rc = ss.redo.ensure(ctx, 4 * tiles.size(), "small redo list")))
        return rc;
    // as many small-kernel workgroups per CU as stay resident (LDS,
    // registers) or one workgroup-kernel workgroup per CU, never more than
    // the tiles, nor more slabs than fit in half the free capacity
    const size_t slab = (size_t)nmax * TW;
    // list entries: a vertex joins at most one level per source, so 64 W n
    // bounds a tile
    const size_t qcap = wg ? (size_t)nmax * TW + TW : slab;
    const size_t per_wg = 16 * slab + (4 + 8 * (size_t)W) * qcap +
                          (wg ? 4 * qcap + 24 * (size_t)W * nmax + 12 * (size_t)nmax + 4 * ((size_t)nmax + 2) : 0);
    const int64_t per_cu = wg ? 1 : small_blocks_per_cu(nmax, csr);
    int64_t grid = std::max<int64_t>(1, std::min<int64_t>((int64_t)tiles.size(), per_cu * (int64_t)ctx->n_cu));
    if (wg) {
        const size_t room = ctx->mem_cap > ctx->mem_used ? (ctx->mem_cap - ctx->mem_used) / 2 : 0;
        grid = std::max<int64_t>(1, std::min<int64_t>(grid, (int64_t)(room / per_wg)));
    }
    if ((rc = ss.sig.ensure(ctx, 8 * slab * grid, "persistent sigma")) ||
        (rc = ss.tt.ensure(ctx, 8 * slab * grid, "persistent dependency terms")) ||
        (rc = ss.Q.ensure(ctx, 4 * qcap * grid, "persistent lists")) ||
        (rc = ss.Qm.ensure(ctx, 8 * (size_t)W * qcap * grid, "persistent lists")))
        return rc;
    if (wg && W > 1 && (rc = ss.tag.ensure(ctx, 4 * (size_t)nmax * grid, "workgroup level tags"))) return rc;
    if (wg && ((rc = ss.masks.ensure(ctx, 24 * (size_t)W * nmax * grid, "workgroup masks")) ||
               (rc = ss.qr.ensure(ctx, 4 * qcap * grid, "workgroup record offsets")) ||
               (rc = ss.ro.ensure(ctx, 8 * (size_t)nmax * grid, "workgroup record offsets")) ||
               (rc = ss.qoffs.ensure(ctx, 4 * ((size_t)nmax + 2) * grid, "workgroup level offsets"))))
        return rc;
    if (probe && wg && (rc = ss.psig.ensure(ctx, 8 * slab * grid, "probe path counts"))) return rc;
    if (probe && ((rc = ss.lvl.ensure(ctx, 4 * slab * grid, "probe levels")) ||
                  (rc = ss.dbg.ensure(ctx, 8 * slab * grid, "probe delta"))))
        return rc;
    SmallArgs a{};
    unsigned char *tb = (unsigned char *)ss.tables.p;
    a.comps = (const SmallComp *)tb;
    a.tiles = (const SmallTile *)(tb + o_tiles);
    a.n_tiles = (int32_t)tiles.size();
    a.srcs = (const int32_t *)(tb + o_srcs);
    a.ws = (const double *)(tb + o_ws);
    a.counter = (int32_t *)ss.counter.p;
    a.redo = (int32_t *)ss.redo.p;
    a.sig = (uint64_t *)ss.sig.p;
    a.tt = ss.tt.p;
    a.Q = (int32_t *)ss.Q.p;
    a.Qm = (uint64_t *)ss.Qm.p;
    a.slab = (int64_t)slab;
    a.limit = (o->sigma_f64_threshold_log2 > 0 && o->sigma_f64_threshold_log2 < 64)
                  ? (1ull << o->sigma_f64_threshold_log2) : 0;
    a.masks = wg ? (uint64_t *)ss.masks.p : nullptr;
    a.qoffs = wg ? (int32_t *)ss.qoffs.p : nullptr;
    a.nmax = nmax;
    a.csr_cap = csr;
    a.qcap = (int64_t)qcap;
    a.qr = wg ? (int32_t *)ss.qr.p : nullptr;
    a.ro = wg ? (int32_t *)ss.ro.p : nullptr;
    a.spl = W;
    a.tag = wg && W > 1 ? (int32_t *)ss.tag.p : nullptr;
    a.psig = probe && wg ? (uint64_t *)ss.psig.p : nullptr;
    a.lvl = probe ? (int32_t *)ss.lvl.p : nullptr;
    a.dbg = probe ? (double *)ss.dbg.p : nullptr;
    ss.run = a;
    ss.run_fp32 = o->precision == BCX_PREC_FP32C;
    ss.run_graph = !(o->flags & BCX_FLAG_NO_GRAPH);
    ss.run_tiles = std::move(tiles);
    ss.run_narrow = n_narrow_tiles;
    ss.run_grid = (int)grid;
    ss.run_stream = sp;
    ss.pending = true;
    // uint64 pass over the tiles not predicted to overflow
    ss.run.n_tiles = (int32_t)n_narrow_tiles;
    if (ss.run.n_tiles > 0) {
        if ((rc = follow_ctx_stream(ctx, sp))) return rc;
        ctx->timer.begin(ctx);
        if ((rc = persistent_pass(ctx, ss, wg, false, ss.run_graph))) return rc;
        ctx->timer.end(ctx);
        if (st) {
            st->n_launches += 2;
            if (ss.run_graph) st->n_graph_launches++;
        }
    }
    return BCX_OK;
}

static int persistent_finish(bcx_ctx *ctx, bcx_stats *st, bool wg) {
    SmallState *ps = wg ? ctx->wgk : ctx->small;
    if (!ps || !ps->pending) return BCX_OK;
    SmallState &ss = *ps;
    ss.pending = false;
    const hipStream_t sp = ss.run_stream;
    int rc;
    int32_t h[3] = {0, 0, 0};
    if (ss.run.n_tiles > 0 &&
        ((rc = ctx_hip(ctx, hipMemcpyAsync(h, ss.run.counter, sizeof(h), hipMemcpyDeviceToHost, sp), "counter")) ||
         (rc = ctx_hip(ctx, hipStreamSynchronize(sp), "persistent kernel"))))
        return rc;
    if (h[2]) return ctx_fail(ctx, BCX_E_OVERFLOW, "workgroup kernel: a tile's level lists exceeded their bound");
    if (h[1] > 0 || ss.run_narrow < ss.run_tiles.size()) {
        // tiles whose uint64 sigma overflowed added nothing: rerun them in
        // FP64, with the tiles that start there
        std::vector<int32_t> redo((size_t)h[1]);
        if (h[1] > 0 &&
            ((rc = ctx_hip(ctx, hipMemcpyAsync(redo.data(), ss.run.redo, 4 * redo.size(), hipMemcpyDeviceToHost, sp),
                           "redo")) ||
             (rc = ctx_hip(ctx, hipStreamSynchronize(sp), "redo"))))
            return rc;
        std::vector<SmallTile> again;
        for (int32_t t : redo) again.push_back(ss.run_tiles[t]);
        for (size_t t = ss.run_narrow; t < ss.run_tiles.size(); ++t) again.push_back(ss.run_tiles[t]);
        std::sort(again.begin(), again.end(), [](const SmallTile &x, const SmallTile &y) { return x.src0 < y.src0; });
        if ((rc = upload(ctx, again.data(), sizeof(SmallTile) * again.size(), ss.tiles, "small tiles"))) return rc;
        ss.run.tiles = (const SmallTile *)ss.tiles.p;
        ss.run.n_tiles = (int32_t)again.size();
        ss.run.limit = 0;
        if (st) for (auto &t : again) st->n_f64_sigma += t.ns;
        if ((rc = follow_ctx_stream(ctx, sp))) return rc;
        ctx->timer.begin(ctx);
        if ((rc = persistent_pass(ctx, ss, wg, true, false))) return rc;
        ctx->timer.end(ctx);
        if (st) st->n_launches += 2;
        if ((rc = ctx_hip(ctx, hipMemcpyAsync(h, ss.run.counter, sizeof(h), hipMemcpyDeviceToHost, sp), "counter")) ||
            (rc = ctx_hip(ctx, hipStreamSynchronize(sp), "persistent kernel")))
            return rc;
        if (h[2]) return ctx_fail(ctx, BCX_E_OVERFLOW, "workgroup kernel: a tile's level lists exceeded their bound");
    }
    return BCX_OK;
}

int batched_persistent_start(bcx_ctx *ctx, const std::vector<SmallJob> &jobs, const bcx_opts *o, double *bc_base,
                             bcx_stats *st, bool wg, hipStream_t sp) {
    return persistent_start(ctx, jobs, o, bc_base, st, wg, sp, false);
}

int batched_persistent_finish(bcx_ctx *ctx, bcx_stats *st, bool wg) { return persistent_finish(ctx, st, wg); }

void batched_persistent_reset(bcx_ctx *ctx) {
    if (ctx->side) (void)hipStreamSynchronize(ctx->side);
    if (ctx->small) ctx->small->pending = false;
    if (ctx->wgk) ctx->wgk->pending = false;
}

static void release_state(bcx_ctx *ctx, SmallState *&state) {
    if (!state) return;
    SmallState &ss = *state;
    if (ss.exec) (void)hipGraphExecDestroy(ss.exec);
    if (ss.pin) (void)hipHostFree(ss.pin);
    for (DBuf *b : {&ss.tables, &ss.tiles, &ss.counter, &ss.redo, &ss.sig, &ss.tt, &ss.Q, &ss.Qm, &ss.masks,
                    &ss.qoffs, &ss.qr, &ss.ro, &ss.tag, &ss.lvl, &ss.dbg, &ss.psig})
        b->release(ctx);
    delete state;
    state = nullptr;
}

void small_release(bcx_ctx *ctx) {
    release_state(ctx, ctx->small);
    release_state(ctx, ctx->wgk);
}

// Kernel the batched path runs component dc on, from the bcx_opts flags.
int batched_kind(DComp &dc, int flags) {
    if (dc.n <= SMALL_N && !(flags & BCX_FLAG_NO_SMALL_KERNEL)) return BCX_KERNEL_SMALL;
    if (dc.n <= WG_N && !(flags & BCX_FLAG_NO_WG_KERNEL) && ((flags & BCX_FLAG_WG_KERNEL) || batched_narrow(dc)))
        return BCX_KERNEL_WG;
    return BCX_KERNEL_MULTI;
}

// ------------------------------------------------------------------------
// Schedule tuner.  A component of at least TUNE_N vertices is run once on
// sample tiles (evenly spaced over its tile order, BC into a scratch
// accumulator, timed with HIP events, outside the run's timing) under each
// candidate schedule, and the fastest is cached for every later run:
//  * kernel, for components the workgroup kernel can take (<= WG_N vertices):
//    the workgroup kernel over min(tiles, CUs) sample tiles (its concurrent
//    width) against the multi-launch tiles over TUNE_TILES;
//  * multi-launch: 8 / 4 / 16 level-kernel blocks per CU, 8 / 4 tiles in
//    flight, chain contraction on / off.
// Each candidate runs twice (the first run captures its graphs) and is timed
// on the second; a candidate replaces the default only when its rate
// (sources per second) is at least 3% higher.  Options that fix a choice are
// kept.  (The reference picks its schedule from sampled BFS depths and
// frontier sizes, Codes/kernels_free.cu:447-470, 288, 381.)
// ------------------------------------------------------------------------
constexpr int32_t TUNE_N = 16384;
constexpr int TUNE_TILES = 24;

// Sources of k tiles of `width` evenly spaced over the component's tile order.
static std::vector<int32_t> sample_tiles(DComp &dc, int64_t k, int64_t width) {
    const std::vector<int32_t> order = batched_order(dc);
    const int64_t T = ((int64_t)dc.n + width - 1) / width;
    k = std::min(k, T);
    std::vector<int32_t> out;
    for (int64_t i = 0; i < k; ++i) {
        const int64_t t = (2 * i + 1) * T / (2 * k);
        for (int64_t v = t * width; v < std::min<int64_t>((t + 1) * width, dc.n); ++v)
            out.push_back(order[(size_t)v]);
    }
    return out;
}

int batched_schedule(bcx_ctx *ctx, DComp &dc, const bcx_opts *o, bcx_stats *st) {
    const int kind0 = batched_kind(dc, o->flags);
    if (kind0 == BCX_KERNEL_SMALL) return kind0;
    if ((o->flags & BCX_FLAG_NO_TUNE) || dc.n < TUNE_N) return kind0;
    const bool kind_fixed = (o->flags & (BCX_FLAG_NO_WG_KERNEL | BCX_FLAG_WG_KERNEL)) || dc.n > WG_N;
    int kind = kind_fixed ? kind0 : dc.sched_kind;
    const bool need_kind = kind < 0;
    const bool need_multi = !dc.sched_multi && (need_kind || kind == BCX_KERNEL_MULTI);
    if (!need_kind && !need_multi) return kind;
    // the run's timer is set aside: the samples are timed on their own
    Timer keep;
    std::swap(keep, ctx->timer);
    DBuf scratch;
    int rc = scratch.ensure(ctx, 8 * ((size_t)dc.offset + (size_t)dc.n), "tuner scratch");
    bcx_opts to = *o;
    to.src_subset = nullptr;
    to.n_src = 0;
    double *bc = (double *)scratch.p;
    auto time_multi = [&](const std::vector<int32_t> &srcs, const MultiSched &m, double *rate) {
        double t = 0.0;
        int r = BCX_OK;
        for (int rep = 0; rep < 2 && !r; ++rep) {
            r = multi_run(ctx, dc, srcs, &to, m, bc + dc.offset, nullptr);
            t = ctx->timer.collect();
        }
        *rate = t > 0 ? (double)srcs.size() / t : 0.0;
        return r;
    };
    auto time_wg = [&](const std::vector<int32_t> &srcs, double *rate) {
        double t = 0.0;
        int r = BCX_OK;
        std::vector<SmallJob> jobs{{&dc, srcs}};
        for (int rep = 0; rep < 2 && !r; ++rep) {
            r = persistent_start(ctx, jobs, &to, bc, nullptr, true, ctx->stream, false);
            if (!r) r = persistent_finish(ctx, nullptr, true);
            t = ctx->timer.collect();
        }
        *rate = t > 0 ? (double)srcs.size() / t : 0.0;
        return r;
    };
    const MultiSched base = multi_sched(dc, o);  // the defaults, or what the options fix
    // every multi-launch candidate runs the same sources: TUNE_TILES tiles of
    // the widest lanes, which narrower lanes cut into more tiles (a sample of
    // 64-source tiles left the wide candidates a few tiles for eight slots and
    // picked 128-source tiles where 256-source ones run 10% faster)
    const std::vector<int32_t> samp =
        rc ? std::vector<int32_t>() : sample_tiles(dc, TUNE_TILES, (int64_t)BATCH_W * MAX_SPL);
    double r_multi = 0.0;
    if (!rc) rc = time_multi(samp, base, &r_multi);
    if (!rc && need_kind) {
        double r_wg = 0.0;
        rc = time_wg(sample_tiles(dc, std::max(1, ctx->n_cu), (int64_t)BATCH_W * wg_spl(o)), &r_wg);
        kind = r_wg >= r_multi ? BCX_KERNEL_WG : BCX_KERNEL_MULTI;
    }
    MultiSched best = base;
    const bool tune_multi = need_multi && kind == BCX_KERNEL_MULTI;
    if (!rc && tune_multi) {
        std::vector<MultiSched> cands;
        if (o->blocks_per_cu <= 0)
            for (int bpc : {4, 16}) cands.push_back({base.cut, bpc, base.slots, base.spl});
        if (o->batch_width <= 0) cands.push_back({base.cut, base.bpc, 4, base.spl});
        if (base.cut) cands.push_back({false, base.bpc, base.slots, base.spl});
        if (!valid_spl(o->sources_per_lane))
            for (int w : {1, 2, 4})
                if (w != base.spl) cands.push_back({base.cut, base.bpc, base.slots, w});
        double r_best = r_multi;
        for (const MultiSched &m : cands) {
            double r = 0.0;
            if ((rc = time_multi(samp, m, &r))) break;
            if (r > 1.03 * r_best) {
                r_best = r;
                best = m;
            }
        }
    }
    std::swap(keep, ctx->timer);
    keep.destroy();
    scratch.release(ctx);
    if (rc) return rc;
    if (need_kind) dc.sched_kind = (int8_t)kind;
    if (tune_multi) {
        // (a choice the options fixed is not cached: later runs may leave it free)
        dc.sched_multi = true;
        dc.sched_cut = (base.cut && !best.cut) ? 0 : 1;
        dc.sched_bpc = (int16_t)(o->blocks_per_cu > 0 ? 0 : best.bpc);
        dc.sched_spl = (int8_t)(valid_spl(o->sources_per_lane) ? 0 : best.spl);
        dc.sched_slots = (int16_t)(o->batch_width > 0 ? 0 : best.slots);
    }
    if (st) st->n_tuned++;
    return kind;
}

// One tile of srcs[0..ns) on the kernel bcx_run would pick for the component
// (the multi-launch tiles on the component's BFS-ordered copy, as in a run),
// with the probe arrays of that kernel switched on; every lane is read back.
int batched_debug_tile(bcx_ctx *ctx, const bcx_component *comp, const int32_t *srcs, int32_t ns, const bcx_opts *o,
                       int32_t *level, uint64_t *sigma_u64, double *sigma_f64, double *delta, int32_t *sigma_is_f64,
                       int32_t *kernel_used) {
    bcx_dgraph *g = nullptr;
    int rc = bcx_upload(ctx, comp, 1, &g);
    if (rc) return rc;
    DComp &dc = g->comps[0];
    const int32_t n = dc.n;
    const int kind = batched_kind(dc, o->flags);
    // sources per lane of the tile: the multi-launch tiles' schedule (the
    // other kernels run 64-source tiles)
    const int spl = kind == BCX_KERNEL_MULTI ? multi_sched(dc, o).spl : (kind == BCX_KERNEL_WG ? wg_spl(o) : 1);
    if (ns > BATCH_W * spl) {
        bcx_dgraph_free(ctx, g);
        return ctx_fail(ctx, BCX_E_ARG, "debug tile: more sources than the kernel's tile holds");
    }
    const size_t cells = (size_t)n * BATCH_W * spl;
    std::vector<int32_t> lv(cells, 0), at((size_t)n);  // at[v]: row of v in the device arrays
    std::vector<uint64_t> sg(cells, 0);
    std::vector<double> dl(cells, 0.0);
    bool f64 = false;
    batched_persistent_reset(ctx);
    ctx->timer.depth = 0;
    ctx->timer.used = 0;
    double *bc = nullptr;
    rc = ctx->bc_acc.ensure(ctx, sizeof(double) * n, "bc accumulator");
    if (!rc) {
        bc = (double *)ctx->bc_acc.p;
        rc = ctx_hip(ctx, hipMemsetAsync(bc, 0, sizeof(double) * n, ctx->stream), "memset");
    }
    if (!rc && kind != BCX_KERNEL_MULTI) {
        const bool wg = kind == BCX_KERNEL_WG;
        bcx_stats st{};
        if (!rc) {
            for (int32_t v = 0; v < n; ++v) at[v] = v;
            std::vector<SmallJob> jobs{{&dc, std::vector<int32_t>(srcs, srcs + ns)}};
            rc = persistent_start(ctx, jobs, o, bc, &st, wg, ctx->stream, true);
            if (!rc) rc = persistent_finish(ctx, &st, wg);
        }
        f64 = st.n_f64_sigma > 0;
        // one tile, so one workgroup: its slab is the start of every array
        // (the workgroup kernel keeps lane-compacted records: its probe writes
        // the path counts per (vertex, lane) beside them)
        const SmallState *ss = wg ? ctx->wgk : ctx->small;
        if (!rc && ss) {
            const void *sg_src = wg ? ss->psig.p : ss->sig.p;
            rc = ctx_hip(ctx, hipMemcpyAsync(lv.data(), ss->lvl.p, 4 * cells, hipMemcpyDeviceToHost, ctx->stream),
                         "copy");
            if (!rc) rc = ctx_hip(ctx, hipMemcpyAsync(sg.data(), sg_src, 8 * cells, hipMemcpyDeviceToHost,
                                                      ctx->stream), "copy");
            if (!rc) rc = ctx_hip(ctx, hipMemcpyAsync(dl.data(), ss->dbg.p, 8 * cells, hipMemcpyDeviceToHost,
                                                      ctx->stream), "copy");
            if (!rc) rc = ctx_hip(ctx, hipStreamSynchronize(ctx->stream), "copy");
        }
    } else if (!rc) {
        DComp *pc = nullptr;
        ChainCut *cut = nullptr;
        const std::vector<int32_t> *inv = nullptr;
        const int32_t *perm = nullptr;
        rc = tile_graph(ctx, dc, want_cut(o), pc, cut, inv, perm);
        int sm = SIG_U32;
        if (!rc) {
            std::vector<int32_t> ps((size_t)ns);
            for (int32_t i = 0; i < ns; ++i) ps[(size_t)i] = (*inv)[srcs[i]];
            for (int32_t v = 0; v < n; ++v) at[v] = (*inv)[v];
            rc = bs_ensure(ctx, n, BATCH_W * spl, 1, true);
            if (!rc && cut) rc = cut_slots(ctx, cut);
            if (!rc) rc = stage_sources(ctx, *pc, ps, BATCH_W * spl);
            if (!rc) rc = stage_hubs(ctx, *pc);
            if (!rc) {
                TileJob j = make_job(ctx, *pc, o, bc);
                j.perm = perm;
                j.cut = cut;
                j.spl = spl;
                rc = run_tiles(ctx, *pc, ns, j, nullptr, &sm);
            }
        }
        if (ctx->batched) {
            Slot &s = ctx->batched->slots[0];
            const size_t wsz = sm == SIG_U32 ? 4 : 8;
            // copies are ordered on the slot stream the kernels ran on
            if (!rc) rc = ctx_hip(ctx, hipMemcpyAsync(lv.data(), s.lvl.p, 4 * cells, hipMemcpyDeviceToHost, s.stream),
                                  "copy");
            if (!rc) rc = ctx_hip(ctx, hipMemcpyAsync(sg.data(), s.sig.p, wsz * cells, hipMemcpyDeviceToHost, s.stream),
                                  "copy");
            if (!rc) rc = ctx_hip(ctx, hipMemcpyAsync(dl.data(), s.dbg.p, 8 * cells, hipMemcpyDeviceToHost, s.stream),
                                  "copy");
            if (!rc) rc = ctx_hip(ctx, hipStreamSynchronize(s.stream), "copy");
            ctx->batched->probe = false;
        }
        f64 = sm == SIG_F64;
        if (!rc && sm == SIG_U32) {  // 32-bit words: widen in place, from the top
            const uint32_t *w32 = (const uint32_t *)sg.data();
            for (size_t i = cells; i-- > 0;) sg[i] = w32[i];
        }
    }
    (void)ctx->timer.collect();
    bcx_dgraph_free(ctx, g);
    if (rc) return rc;
    // source i is word h = i / 64 of lane k = i % 64: cell (v * 64 + k) * spl + h
    for (int32_t i = 0; i < ns; ++i)
        for (int32_t v = 0; v < n; ++v) {
            const size_t c = ((size_t)at[v] * BATCH_W + (size_t)(i % BATCH_W)) * spl + (size_t)(i / BATCH_W);
            const size_t r = (size_t)i * (size_t)n + (size_t)v;
            double f;
            std::memcpy(&f, &sg[c], 8);
            if (level) level[r] = lv[c];
            if (sigma_u64) sigma_u64[r] = f64 ? 0 : sg[c];
            if (sigma_f64) sigma_f64[r] = f64 ? f : (double)sg[c];
            // a lane's own source has no dependency (it is never a backward entry)
            if (delta) delta[r] = v == srcs[i] ? 0.0 : dl[c];
        }
    if (sigma_is_f64) *sigma_is_f64 = f64 ? 1 : 0;
    if (kernel_used) *kernel_used = kind;
    return BCX_OK;
}

void batched_release(bcx_ctx *ctx) {
    if (!ctx->batched) return;
    BatchedState &bs = *ctx->batched;
    for (Slot &s : bs.slots) {
        if (s.stream) (void)hipStreamSynchronize(s.stream);
        for (TileGraph &g : s.graphs) (void)hipGraphExecDestroy(g.exec);
        s.graphs.clear();
        for (DBuf *b : {&s.vis, &s.mb, &s.sig, &s.tt, &s.Q, &s.Qm, &s.ctr, &s.qoff, &s.hub_mask, &s.hub_acc,
                        &s.hub_cnt, &s.lvl, &s.dbg, &s.bcacc, &s.wsl, &s.eplv, &s.eprho, &s.lanes})
            b->release(ctx);
        if (s.pinned) (void)hipHostFree(s.pinned);
        if (s.done) (void)hipEventDestroy(s.done);
        if (s.stream) (void)hipStreamDestroy(s.stream);
    }
    for (DBuf *b : {&bs.srcs, &bs.ws, &bs.hub_v, &bs.hub_nsl, &bs.sl_hub, &bs.sl_j0, &bs.sl_j1}) b->release(ctx);
    delete ctx->batched;
    ctx->batched = nullptr;
}

void batched_graph_release(bcx_ctx *ctx, bcx_dgraph *g) {
    for (DComp &dc : g->comps) {
        cut_release(ctx, dc);
        if (dc.relabel) {
            for (DBuf *b : {&dc.relabel->b_rp, &dc.relabel->b_col, &dc.relabel->b_reach}) b->release(ctx);
            delete dc.relabel;
            dc.relabel = nullptr;
        }
        dc.b_perm.release(ctx);
    }
}
