// Host driver of the batched multi-source path: tiles of up to 64 sources
// of one component, forward level lists built on the device, FP32-compensated
// (or FP64) rho-form backward, sigma overflow handled per tile by redoing the
// tile with FP64 sigma.  Sources are ordered by a BFS from the component's
// highest-degree vertex so the 64 sources of a tile are close together and
// their frontiers overlap (fewer (vertex, level) list entries per source).
#include <algorithm>
#include <cstring>
#include <vector>

#include "bcx_batched.h"
#include "bcx_runtime.h"

using namespace bcx;

struct BatchedState {
    DBuf vis, mb, sig, tt, Q, Qm, ctr, qoff, srcs, ws, lvl, dbg;
    int32_t n_cap = 0;
    int32_t qcap = 0;
    int32_t src_cap = 0;
    bool probe = false;
    int32_t *pinned = nullptr;  // ctr read-back
    int last_depth = 8;
};

static int bs_ensure(bcx_ctx *ctx, int32_t n, int32_t n_src, bool probe) {
    if (!ctx->batched) ctx->batched = new BatchedState();
    BatchedState &s = *ctx->batched;
    int rc;
    const size_t N = (size_t)n;
    if (!s.pinned) {
        if (hipHostMalloc((void **)&s.pinned, 64, hipHostMallocDefault) != hipSuccess) {
            s.pinned = nullptr;
            return ctx_fail(ctx, BCX_E_HIP, "pinned counters");
        }
    }
    if (n > s.n_cap) {
        if ((rc = s.vis.ensure(ctx, 8 * N, "batched vis")) ||
            (rc = s.mb.ensure(ctx, 24 * N, "batched masks")) ||
            (rc = s.sig.ensure(ctx, 8 * 64 * N, "batched sigma")) ||
            (rc = s.tt.ensure(ctx, 8 * 64 * N, "batched dependency terms")) ||
            (rc = s.qoff.ensure(ctx, 4 * (N + 8), "batched level offsets")))
            return rc;
        rc = ctx_hip(ctx, hipMemsetAsync(s.mb.p, 0, s.mb.bytes, ctx->stream), "memset");
        if (rc) return rc;
        s.n_cap = n;
    }
    if (s.qcap < 8 * n + 4096) {
        const int64_t q = std::min<int64_t>((int64_t)8 * n + 4096, INT32_MAX / 2);
        if ((rc = s.Q.ensure(ctx, 4 * (size_t)q, "batched lists")) ||
            (rc = s.Qm.ensure(ctx, 8 * (size_t)q, "batched lists")))
            return rc;
        s.qcap = (int32_t)q;
    }
    if (n_src > s.src_cap) {
        if ((rc = s.srcs.ensure(ctx, 4 * (size_t)n_src, "batched sources")) ||
            (rc = s.ws.ensure(ctx, 8 * (size_t)n_src, "batched weights")))
            return rc;
        s.src_cap = n_src;
    }
    if ((rc = s.ctr.ensure(ctx, 64, "batched counters"))) return rc;
    if (probe) {
        if ((rc = s.lvl.ensure(ctx, 2 * 64 * N, "probe levels")) || (rc = s.dbg.ensure(ctx, 8 * 64 * N, "probe delta")))
            return rc;
    }
    s.probe = probe;
    return BCX_OK;
}

static BatchDev make_dev(bcx_ctx *ctx, const DComp &dc, double *bc, uint64_t limit, int tile) {
    BatchedState &s = *ctx->batched;
    BatchDev b{};
    b.n = dc.n;
    b.rp = dc.dev.rp;
    b.col = dc.dev.col;
    b.reach = dc.dev.reach;
    b.vis = (uint64_t *)s.vis.p;
    b.mb0 = (uint64_t *)s.mb.p;
    b.mb1 = b.mb0 + s.n_cap;
    b.mb2 = b.mb1 + s.n_cap;
    b.sig = (uint64_t *)s.sig.p;
    b.tt = s.tt.p;
    b.lvl = s.probe ? (uint16_t *)s.lvl.p : nullptr;
    b.dbg_delta = s.probe ? (double *)s.dbg.p : nullptr;
    b.Q = (int32_t *)s.Q.p;
    b.Qm = (uint64_t *)s.Qm.p;
    b.ctr = (int32_t *)s.ctr.p;
    b.qoff = (int32_t *)s.qoff.p;
    b.qcap = s.qcap;
    b.srcs = (const int32_t *)s.srcs.p + (size_t)tile * BATCH_W;
    b.ws = (const double *)s.ws.p + (size_t)tile * BATCH_W;
    b.bc = bc;
    b.limit = limit;
    return b;
}

// One tile: init, forward levels and a speculative backward launched as one
// sequence, then a single read-back of the counters.  A forward deeper than
// the launched levels continues from where it stopped; a sigma overflow
// (uint64) or full level lists leave the BC untouched and redo the tile.
static int run_tile(bcx_ctx *ctx, const DComp &dc, int tile, int ns, bool fp32, double *bc, uint64_t limit,
                    bcx_stats *st, bool *f64_out) {
    BatchedState &s = *ctx->batched;
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(((int64_t)dc.n + 3) / 4, (int64_t)ctx->n_cu * 8));
    bool f64 = false;
    int rc;
    for (int attempt = 0; attempt < 6; ++attempt) {
        BatchDev b = make_dev(ctx, dc, bc, f64 ? 0 : limit, tile);
        ctx->timer.begin(ctx);
        if ((rc = ctx_hip(ctx, launch_b_init(b, ns, f64, grid, ctx->stream), "init"))) return rc;
        int L = 0, C = std::max(4, s.last_depth + 3);
        int32_t *h = s.pinned;
        int launches = 2;
        while (true) {
            for (int l = L; l < L + C; ++l)
                if ((rc = ctx_hip(ctx, launch_b_forward(b, f64, l, grid, ctx->stream), "forward"))) return rc;
            L += C;
            for (int l = L - 1; l >= 1; --l)
                if ((rc = ctx_hip(ctx, launch_b_back(b, f64, fp32, l, grid, ctx->stream), "back"))) return rc;
            if ((rc = ctx_hip(ctx, launch_b_finish(b, grid, ctx->stream), "finish"))) return rc;
            launches += 2 * C + (L - 1) + 1;
            if ((rc = ctx_hip(ctx, hipMemcpyAsync(h, b.ctr, 4 * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream),
                              "counters")) ||
                (rc = ctx_hip(ctx, hipStreamSynchronize(ctx->stream), "batched kernels")))
                return rc;
            if (h[3] >= 0) break;
            C = std::min(2 * C, 4096);
        }
        ctx->timer.end(ctx);
        if (st) st->n_launches += launches;
        const int D = h[3];
        if (h[2]) {
            // lists full: masks may hold entries of vertices never listed
            if ((rc = ctx_hip(ctx, hipMemsetAsync(s.mb.p, 0, s.mb.bytes, ctx->stream), "memset"))) return rc;
            const int64_t q = std::min<int64_t>((int64_t)s.qcap * 4, std::min<int64_t>((int64_t)64 * dc.n + 4096,
                                                                                        INT32_MAX / 2));
            if (q <= s.qcap) return ctx_fail(ctx, BCX_E_NOMEM, "batched level lists cannot grow");
            if ((rc = ctx_hip(ctx, hipStreamSynchronize(ctx->stream), "sync")) ||
                (rc = s.Q.ensure(ctx, 4 * (size_t)q, "batched lists")) ||
                (rc = s.Qm.ensure(ctx, 8 * (size_t)q, "batched lists")))
                return rc;
            s.qcap = (int32_t)q;
            continue;
        }
        if (!f64 && h[1]) {
            if ((rc = ctx_hip(ctx, launch_b_clear_all(b, std::min(h[0], s.qcap), grid, ctx->stream), "clear")))
                return rc;
            f64 = true;
            if (st) st->n_f64_sigma += ns;
            continue;
        }
        s.last_depth = D;
        if (st && D > st->max_level) st->max_level = D;
        if (f64_out) *f64_out = f64;
        return BCX_OK;
    }
    return ctx_fail(ctx, BCX_E_OVERFLOW, "batched tile could not be completed");
}

static int stage_sources(bcx_ctx *ctx, const DComp &dc, const std::vector<int32_t> &srcs) {
    const size_t ntiles = (srcs.size() + BATCH_W - 1) / BATCH_W;
    std::vector<int32_t> hs(ntiles * BATCH_W, -1);
    std::vector<double> hw(ntiles * BATCH_W, 0.0);
    for (size_t i = 0; i < srcs.size(); ++i) {
        hs[i] = srcs[i];
        hw[i] = (double)dc.host.reach[srcs[i]] + 1.0;
    }
    BatchedState &s = *ctx->batched;
    int rc;
    if ((rc = ctx_hip(ctx, hipMemcpyAsync(s.srcs.p, hs.data(), 4 * hs.size(), hipMemcpyHostToDevice, ctx->stream),
                      "sources")) ||
        (rc = ctx_hip(ctx, hipMemcpyAsync(s.ws.p, hw.data(), 8 * hw.size(), hipMemcpyHostToDevice, ctx->stream),
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

int batched_component(bcx_ctx *ctx, bcx_dgraph *, DComp &dc, const Plan &plan, const bcx_opts *o, double *bc,
                      bcx_stats *st) {
    // every requested source runs as a BFS lane: the plan's BFS sources plus
    // its chain sources
    std::vector<int32_t> srcs(plan.free_src);
    for (auto &b : plan.batches) {
        for (size_t i = 0; i < b.bsrc.size(); ++i)
            if (b.bmode[i] & MODE_BC) srcs.push_back(b.bsrc[i]);
        for (size_t i = 0; i < b.cci.size(); ++i)
            srcs.push_back(dc.host.ch_nodes[dc.host.ch_ptr[b.cci[i]] + b.cpos[i]]);
    }
    const std::vector<int32_t> rank = bfs_rank(dc.host);
    std::sort(srcs.begin(), srcs.end(), [&](int32_t a, int32_t b) { return rank[a] < rank[b]; });
    srcs.erase(std::unique(srcs.begin(), srcs.end()), srcs.end());
    const size_t ntiles = (srcs.size() + BATCH_W - 1) / BATCH_W;
    int rc = bs_ensure(ctx, dc.n, (int32_t)(ntiles * BATCH_W), false);
    if (rc) return rc;
    if ((rc = stage_sources(ctx, dc, srcs))) return rc;
    const uint64_t limit = (o->sigma_f64_threshold_log2 > 0 && o->sigma_f64_threshold_log2 < 64)
                               ? (1ull << o->sigma_f64_threshold_log2) : 0;
    const bool fp32 = o->precision == BCX_PREC_FP32C;
    for (size_t t = 0; t < ntiles; ++t) {
        const int ns = (int)std::min<size_t>(BATCH_W, srcs.size() - t * BATCH_W);
        if ((rc = run_tile(ctx, dc, (int)t, ns, fp32, bc, limit, st, nullptr))) return rc;
    }
    return BCX_OK;
}

int batched_debug_source(bcx_ctx *ctx, const bcx_component *comp, int32_t src, const bcx_opts *o, int32_t *level,
                         uint64_t *sigma_u64, double *sigma_f64, double *delta, int32_t *sigma_is_f64) {
    bcx_dgraph *g = nullptr;
    int rc = bcx_upload(ctx, comp, 1, &g);
    if (rc) return rc;
    DComp &dc = g->comps[0];
    const int32_t n = dc.n;
    rc = bs_ensure(ctx, n, BATCH_W, true);
    if (!rc) rc = ctx->bc_acc.ensure(ctx, sizeof(double) * n, "bc accumulator");
    bool f64 = false;
    std::vector<uint16_t> lv((size_t)n * 64);
    std::vector<uint64_t> sg((size_t)n * 64);
    std::vector<double> dl((size_t)n * 64);
    if (!rc) {
        const uint64_t limit = (o->sigma_f64_threshold_log2 > 0 && o->sigma_f64_threshold_log2 < 64)
                                   ? (1ull << o->sigma_f64_threshold_log2) : 0;
        std::vector<int32_t> tile{src};
        rc = stage_sources(ctx, dc, tile);
        if (!rc)
            rc = run_tile(ctx, dc, 0, 1, o->precision == BCX_PREC_FP32C, (double *)ctx->bc_acc.p, limit, nullptr,
                          &f64);
    }
    // copies are ordered on the context stream (the kernels run there)
    if (!rc) rc = ctx_hip(ctx, hipMemcpyAsync(lv.data(), ctx->batched->lvl.p, 2 * lv.size(), hipMemcpyDeviceToHost,
                                              ctx->stream), "copy");
    if (!rc) rc = ctx_hip(ctx, hipMemcpyAsync(sg.data(), ctx->batched->sig.p, 8 * sg.size(), hipMemcpyDeviceToHost,
                                              ctx->stream), "copy");
    if (!rc) rc = ctx_hip(ctx, hipMemcpyAsync(dl.data(), ctx->batched->dbg.p, 8 * dl.size(), hipMemcpyDeviceToHost,
                                              ctx->stream), "copy");
    if (!rc) rc = ctx_hip(ctx, hipStreamSynchronize(ctx->stream), "copy");
    (void)ctx->timer.collect();
    ctx->batched->probe = false;
    bcx_dgraph_free(ctx, g);
    if (rc) return rc;
    for (int32_t v = 0; v < n; ++v) {
        const size_t i = (size_t)v * 64;
        if (level) level[v] = lv[i];
        double f;
        std::memcpy(&f, &sg[i], 8);
        if (sigma_u64) sigma_u64[v] = f64 ? 0 : sg[i];
        if (sigma_f64) sigma_f64[v] = f64 ? f : (double)sg[i];
        if (delta) delta[v] = v == src ? 0.0 : dl[i];
    }
    if (sigma_is_f64) *sigma_is_f64 = f64 ? 1 : 0;
    return BCX_OK;
}

void batched_release(bcx_ctx *ctx) {
    if (!ctx->batched) return;
    BatchedState &s = *ctx->batched;
    for (DBuf *b : {&s.vis, &s.mb, &s.sig, &s.tt, &s.Q, &s.Qm, &s.ctr, &s.qoff, &s.srcs, &s.ws, &s.lvl, &s.dbg})
        b->release(ctx);
    if (s.pinned) (void)hipHostFree(s.pinned);
    delete ctx->batched;
    ctx->batched = nullptr;
}

void batched_graph_release(bcx_ctx *, bcx_dgraph *) {}
