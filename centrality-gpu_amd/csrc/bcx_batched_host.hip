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
    DBuf vis, mb, sig, tt, Q, Qm, ctr, srcs, ws, lvl, dbg;
    int32_t n_cap = 0;
    int32_t qcap = 0;
    bool probe = false;
};

static int bs_ensure(bcx_ctx *ctx, int32_t n, bool probe) {
    if (!ctx->batched) ctx->batched = new BatchedState();
    BatchedState &s = *ctx->batched;
    int rc;
    const size_t N = (size_t)n;
    if (n > s.n_cap) {
        if ((rc = s.vis.ensure(ctx, 8 * N, "batched vis")) ||
            (rc = s.mb.ensure(ctx, 24 * N, "batched masks")) ||
            (rc = s.sig.ensure(ctx, 8 * 64 * N, "batched sigma")) ||
            (rc = s.tt.ensure(ctx, 8 * 64 * N, "batched dependency terms")))
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
    if ((rc = s.ctr.ensure(ctx, 64, "batched counters")) || (rc = s.srcs.ensure(ctx, 4 * 64, "batched sources")) ||
        (rc = s.ws.ensure(ctx, 8 * 64, "batched weights")))
        return rc;
    if (probe) {
        if ((rc = s.lvl.ensure(ctx, 2 * 64 * N, "probe levels")) || (rc = s.dbg.ensure(ctx, 8 * 64 * N, "probe delta")))
            return rc;
    }
    s.probe = probe;
    return BCX_OK;
}

static BatchDev make_dev(bcx_ctx *ctx, const DComp &dc, double *bc, uint64_t limit) {
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
    b.qcap = s.qcap;
    b.srcs = (const int32_t *)s.srcs.p;
    b.ws = (const double *)s.ws.p;
    b.bc = bc;
    b.limit = limit;
    return b;
}

static int read_ctr(bcx_ctx *ctx, int32_t *h) {
    int rc = ctx_hip(ctx, hipMemcpyAsync(h, ctx->batched->ctr.p, 3 * sizeof(int32_t), hipMemcpyDeviceToHost,
                                         ctx->stream), "batched counters");
    if (rc) return rc;
    return ctx_hip(ctx, hipStreamSynchronize(ctx->stream), "batched kernels");
}

// One tile; on return *qoff_out holds the level offsets of the lists.
static int run_tile(bcx_ctx *ctx, const DComp &dc, const std::vector<int32_t> &src, bool fp32, double *bc,
                    uint64_t limit, bcx_stats *st, std::vector<int32_t> *qoff_out, bool *f64_out) {
    BatchedState &s = *ctx->batched;
    const int ns = (int)src.size();
    int32_t hs[64];
    double hw[64];
    for (int k = 0; k < 64; ++k) {
        hs[k] = k < ns ? src[k] : -1;
        hw[k] = k < ns ? (double)dc.host.reach[src[k]] + 1.0 : 0.0;
    }
    int rc;
    if ((rc = ctx_hip(ctx, hipMemcpyAsync(s.srcs.p, hs, sizeof(hs), hipMemcpyHostToDevice, ctx->stream), "srcs")) ||
        (rc = ctx_hip(ctx, hipMemcpyAsync(s.ws.p, hw, sizeof(hw), hipMemcpyHostToDevice, ctx->stream), "ws")))
        return rc;
    const int cap = ctx->n_cu * 8;
    bool f64 = false;
    for (int attempt = 0; attempt < 4; ++attempt) {
        BatchDev b = make_dev(ctx, dc, bc, f64 ? 0 : limit);
        std::vector<int32_t> qoff{0, ns};
        ctx->timer.begin(ctx);
        if ((rc = ctx_hip(ctx, launch_b_init(b, ns, f64, cap, ctx->stream), "init"))) return rc;
        int l = 0, D = 0;
        int32_t h[3];
        bool qfull = false;
        while (true) {
            const int qm0 = l > 0 ? qoff[l - 1] : 0;
            if ((rc = ctx_hip(ctx, launch_b_expand(b, l, qm0, qoff[l], qoff[l + 1], cap, ctx->stream), "expand")))
                return rc;
            if ((rc = read_ctr(ctx, h))) return rc;
            if (h[2]) { qfull = true; break; }
            qoff.push_back(h[0]);
            if (h[0] == qoff[l + 1]) { D = l; break; }
            if ((rc = ctx_hip(ctx, launch_b_pull(b, f64, l, qoff[l + 1], h[0], cap, ctx->stream), "pull")))
                return rc;
            l++;
        }
        if (!qfull && !f64) {
            if ((rc = read_ctr(ctx, h))) return rc;
        }
        if (qfull || (!f64 && h[1])) {
            // clean every mask entry the aborted forward may have set, then retry
            const int end = std::min<int32_t>(qfull ? s.qcap : h[0], s.qcap);
            if ((rc = ctx_hip(ctx, launch_b_clear2(b, 0, 0, end, 1, 0, end, cap, ctx->stream), "clear")) ||
                (rc = ctx_hip(ctx, launch_b_clear2(b, 2, 0, end, 2, 0, 0, cap, ctx->stream), "clear")))
                return rc;
            ctx->timer.end(ctx);
            if (qfull) {
                s.qcap = 0;
                const int64_t q = std::min<int64_t>((int64_t)64 * dc.n + 4096, INT32_MAX / 2);
                (void)hipStreamSynchronize(ctx->stream);
                if ((rc = s.Q.ensure(ctx, 4 * (size_t)q, "batched lists")) ||
                    (rc = s.Qm.ensure(ctx, 8 * (size_t)q, "batched lists")))
                    return rc;
                s.qcap = (int32_t)q;
            } else {
                f64 = true;
                if (st) st->n_f64_sigma += ns;
            }
            continue;
        }
        if (st && D > st->max_level) st->max_level = D;
        // backward, levels D .. 1
        for (int lb = D; lb >= 1; --lb) {
            const int c0 = lb + 2 <= D ? qoff[lb + 2] : 0, c1 = lb + 2 <= D ? qoff[lb + 3] : 0;
            if ((rc = ctx_hip(ctx, launch_b_back(b, f64, fp32, lb, qoff[lb], qoff[lb + 1], c0, c1, cap, ctx->stream),
                              "back")))
                return rc;
        }
        // masks of levels 1 and 2 remain from the backward
        const int a0 = qoff[1], a1 = D >= 1 ? qoff[2] : qoff[1];
        const int b0 = D >= 2 ? qoff[2] : 0, b1 = D >= 2 ? qoff[3] : 0;
        if ((rc = ctx_hip(ctx, launch_b_clear2(b, 1, a0, a1, 2, b0, b1, cap, ctx->stream), "clear"))) return rc;
        ctx->timer.end(ctx);
        if (st) st->n_launches += 2 + 2 * D + D + 1;
        if (qoff_out) *qoff_out = qoff;
        if (f64_out) *f64_out = f64;
        return BCX_OK;
    }
    return ctx_fail(ctx, BCX_E_OVERFLOW, "batched tile could not be completed");
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
    int rc = bs_ensure(ctx, dc.n, false);
    if (rc) return rc;
    const uint64_t limit = (o->sigma_f64_threshold_log2 > 0 && o->sigma_f64_threshold_log2 < 64)
                               ? (1ull << o->sigma_f64_threshold_log2) : 0;
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
    const bool fp32 = o->precision == BCX_PREC_FP32C;
    std::vector<int32_t> tile;
    for (size_t i = 0; i < srcs.size(); i += BATCH_W) {
        tile.assign(srcs.begin() + i, srcs.begin() + std::min(srcs.size(), i + BATCH_W));
        if ((rc = run_tile(ctx, dc, tile, fp32, bc, limit, st, nullptr, nullptr))) return rc;
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
    rc = bs_ensure(ctx, n, true);
    if (!rc) rc = ctx->bc_acc.ensure(ctx, sizeof(double) * n, "bc accumulator");
    bool f64 = false;
    std::vector<uint16_t> lv((size_t)n * 64);
    std::vector<uint64_t> sg((size_t)n * 64);
    std::vector<double> dl((size_t)n * 64);
    if (!rc) {
        const uint64_t limit = (o->sigma_f64_threshold_log2 > 0 && o->sigma_f64_threshold_log2 < 64)
                                   ? (1ull << o->sigma_f64_threshold_log2) : 0;
        std::vector<int32_t> tile{src};
        rc = run_tile(ctx, dc, tile, o->precision == BCX_PREC_FP32C, (double *)ctx->bc_acc.p, limit, nullptr, nullptr,
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
    for (DBuf *b : {&s.vis, &s.mb, &s.sig, &s.tt, &s.Q, &s.Qm, &s.ctr, &s.srcs, &s.ws, &s.lvl, &s.dbg}) b->release(ctx);
    delete ctx->batched;
    ctx->batched = nullptr;
}

void batched_graph_release(bcx_ctx *, bcx_dgraph *) {}
