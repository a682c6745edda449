// GPU engine half of the C-ABI (include/bcx.h): context, device memory with a
// hard cap, component upload, source planning (free sources, endpoint rows,
// chain batches) and the launch sequence of the plain and batched paths.
//
// Replaces the reference's begin_gpu / bc_gpu_free / bc_gpu_active /
// postprocess host drivers (Codes/begin_gpupgm.cu:42-89,
// Codes/kernels_free.cu:6-114, Codes/kernels_active.cu:18-223,
// Codes/postprocesses.cu:8-203): one upload of all components instead of a
// malloc/copy/free per BCC, status codes instead of exit().
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../host/host_graph.h"
#include "bcx.h"
#include "bcx_batched.h"
#include "bcx_device.h"
#include "bcx_kernels.h"
#include "bcx_runtime.h"
#include "bcx_small.h"

using namespace bcx;

// ------------------------------------------------------------------------
// context and memory
// ------------------------------------------------------------------------

int ctx_fail(bcx_ctx *ctx, int code, const std::string &msg) {
    if (ctx) ctx->err = msg;
    return code;
}

int ctx_hip(bcx_ctx *ctx, hipError_t e, const char *what) {
    if (e == hipSuccess) return BCX_OK;
    return ctx_fail(ctx, e == hipErrorOutOfMemory ? BCX_E_NOMEM : BCX_E_HIP,
                    std::string(what) + ": " + hipGetErrorString(e));
}

int dev_alloc(bcx_ctx *ctx, void **p, size_t bytes, const char *what) {
    *p = nullptr;
    if (bytes == 0) bytes = 64;
    bytes = (bytes + 255) & ~(size_t)255;
    if (ctx->mem_used + bytes > ctx->mem_cap)
        return ctx_fail(ctx, BCX_E_NOMEM,
                        std::string("device memory cap exceeded allocating ") + what + " (" +
                            std::to_string(bytes) + " B, " + std::to_string(ctx->mem_used) +
                            " of " + std::to_string(ctx->mem_cap) + " B in use)");
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return ctx_hip(ctx, e, what);
    }
    ctx->mem_used += bytes;
    ctx->sizes[*p] = bytes;
    return BCX_OK;
}

void dev_free(bcx_ctx *ctx, void *p) {
    if (!p) return;
    auto it = ctx->sizes.find(p);
    if (it != ctx->sizes.end()) {
        ctx->mem_used -= it->second;
        ctx->sizes.erase(it);
    }
    (void)hipFree(p);
}

int DBuf::ensure(bcx_ctx *ctx, size_t need, const char *what) {
    if (need <= bytes && p) return BCX_OK;
    dev_free(ctx, p);
    p = nullptr;
    bytes = 0;
    size_t grow = need + need / 8;
    int rc = dev_alloc(ctx, &p, grow, what);
    if (rc) {
        rc = dev_alloc(ctx, &p, need, what);
        if (rc) return rc;
        bytes = need;
        return BCX_OK;
    }
    bytes = grow;
    return BCX_OK;
}

void DBuf::release(bcx_ctx *ctx) {
    dev_free(ctx, p);
    p = nullptr;
    bytes = 0;
}

int upload(bcx_ctx *ctx, const void *host, size_t bytes, DBuf &buf, const char *what) {
    int rc = buf.ensure(ctx, bytes ? bytes : 4, what);
    if (rc) return rc;
    if (bytes) return ctx_hip(ctx, hipMemcpyAsync(buf.p, host, bytes, hipMemcpyHostToDevice, ctx->stream), what);
    return BCX_OK;
}

// Kernel time: events around every launch, summed after the final sync.
void Timer::begin(bcx_ctx *ctx) {
    if (depth++ > 0) return;
    if (used == evs.size()) {
        hipEvent_t a, b;
        (void)hipEventCreate(&a);
        (void)hipEventCreate(&b);
        evs.push_back({a, b});
    }
    (void)hipEventRecord(evs[used].first, ctx->stream);
}
void Timer::end(bcx_ctx *ctx) {
    if (--depth > 0) return;
    (void)hipEventRecord(evs[used].second, ctx->stream);
    used++;
}
double Timer::collect() {
    double t = 0;
    for (size_t i = 0; i < used; ++i) {
        float ms = 0;
        (void)hipEventSynchronize(evs[i].second);
        (void)hipEventElapsedTime(&ms, evs[i].first, evs[i].second);
        t += ms * 1e-3;
    }
    used = 0;
    return t;
}
void Timer::destroy() {
    for (auto &e : evs) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    evs.clear();
    used = 0;
}

// ------------------------------------------------------------------------
// source planning
// ------------------------------------------------------------------------

// Roles, chain membership and the chain batches for one component.
// Every endpoint of a needed chain gets its own BC exactly once (MODE_BC)
// in the first batch that holds its row; later batches recompute the row
// only (MODE_ROW).  Chains are ordered by the BFS position of their later
// endpoint on the ear graph (endpoints as vertices, chains as edges), the
// role of the ear-graph level batches of Codes/arrangelevels.cpp:56-109, and
// packed greedily while the distinct endpoints fit in row_cap slots.
int bcx::make_plan(const HostComp &hc, const int32_t *subset, int64_t n_sub, bool use_chains,
              int64_t row_cap, Plan &plan) {
    const int32_t n = hc.n;
    const int32_t nch = use_chains ? hc.n_chains : 0;
    plan = Plan();
    std::vector<int32_t> chain_of((size_t)n, -1), pos_of((size_t)n, -1);
    std::vector<uint8_t> is_end((size_t)n, 0);
    for (int32_t c = 0; c < nch; ++c) {
        for (int32_t j = hc.ch_ptr[c]; j < hc.ch_ptr[c + 1]; ++j) {
            chain_of[hc.ch_nodes[j]] = c;
            pos_of[hc.ch_nodes[j]] = j - hc.ch_ptr[c];
        }
        is_end[hc.ch_ends[2 * c]] = 1;
        is_end[hc.ch_ends[2 * c + 1]] = 1;
    }
    std::vector<uint8_t> want_bfs((size_t)n, 0);
    std::vector<std::vector<int32_t>> chain_srcs((size_t)nch);
    auto add = [&](int32_t v) {
        if (v < 0 || v >= n) return;
        if (chain_of[v] >= 0) chain_srcs[chain_of[v]].push_back(pos_of[v]);
        else want_bfs[v] = 1;
    };
    if (subset) for (int64_t i = 0; i < n_sub; ++i) add(subset[i]);
    else for (int32_t v = 0; v < n; ++v) add(v);

    std::vector<int32_t> need;  // needed chains
    for (int32_t c = 0; c < nch; ++c)
        if (!chain_srcs[c].empty()) {
            std::sort(chain_srcs[c].begin(), chain_srcs[c].end());
            chain_srcs[c].erase(std::unique(chain_srcs[c].begin(), chain_srcs[c].end()), chain_srcs[c].end());
            need.push_back(c);
        }
    // endpoints of needed chains
    std::vector<uint8_t> needed_end((size_t)n, 0);
    for (int32_t c : need) {
        needed_end[hc.ch_ends[2 * c]] = 1;
        needed_end[hc.ch_ends[2 * c + 1]] = 1;
    }
    for (int32_t v = 0; v < n; ++v)
        if (want_bfs[v] && !needed_end[v]) plan.free_src.push_back(v);
    plan.n_bfs = (int64_t)plan.free_src.size();
    if (need.empty()) return BCX_OK;

    // ear graph BFS order over needed endpoints
    std::vector<std::vector<int32_t>> eadj;
    std::vector<int32_t> eid((size_t)n, -1), everts;
    for (int32_t c : need)
        for (int k = 0; k < 2; ++k) {
            int32_t v = hc.ch_ends[2 * c + k];
            if (eid[v] < 0) { eid[v] = (int32_t)everts.size(); everts.push_back(v); }
        }
    eadj.assign(everts.size(), {});
    for (int32_t c : need) {
        int32_t a = eid[hc.ch_ends[2 * c]], b = eid[hc.ch_ends[2 * c + 1]];
        eadj[a].push_back(b);
        eadj[b].push_back(a);
    }
    std::vector<int32_t> bpos(everts.size(), -1);
    int32_t cntpos = 0;
    std::vector<int32_t> q;
    for (size_t r = 0; r < everts.size(); ++r) {
        if (bpos[r] >= 0) continue;
        q.clear();
        q.push_back((int32_t)r);
        bpos[r] = cntpos++;
        for (size_t i = 0; i < q.size(); ++i)
            for (int32_t y : eadj[q[i]])
                if (bpos[y] < 0) { bpos[y] = cntpos++; q.push_back(y); }
    }
    std::vector<std::pair<int64_t, int32_t>> order;
    for (int32_t c : need) {
        int32_t a = bpos[eid[hc.ch_ends[2 * c]]], b = bpos[eid[hc.ch_ends[2 * c + 1]]];
        order.push_back({(int64_t)std::max(a, b) * (1ll << 31) + std::min(a, b), c});
    }
    std::sort(order.begin(), order.end());

    if (row_cap < 2) row_cap = 2;
    std::vector<uint8_t> bc_done((size_t)n, 0);
    std::vector<int32_t> slot((size_t)n, -1);
    PlanBatch cur;
    std::vector<int32_t> cur_ends;
    auto flush = [&]() {
        if (cur_ends.empty()) return;
        for (int32_t v : cur_ends) {
            int8_t mode = MODE_ROW;
            if (want_bfs[v] && !bc_done[v]) { mode |= MODE_BC; bc_done[v] = 1; plan.n_bfs++; }
            else plan.n_row_only++;
            cur.bsrc.push_back(v);
            cur.bslot.push_back(slot[v]);
            cur.bmode.push_back(mode);
        }
        for (int32_t v : cur_ends) slot[v] = -1;
        cur.n_rows = (int32_t)cur_ends.size();
        plan.batches.push_back(std::move(cur));
        cur = PlanBatch();
        cur_ends.clear();
    };
    for (auto &pr : order) {
        int32_t c = pr.second;
        int32_t L = hc.ch_ends[2 * c], R = hc.ch_ends[2 * c + 1];
        int extra = (slot[L] < 0) + (slot[R] < 0 && R != L);
        if ((int64_t)cur_ends.size() + extra > row_cap) flush();
        for (int32_t v : {L, R})
            if (slot[v] < 0) { slot[v] = (int32_t)cur_ends.size(); cur_ends.push_back(v); }
        for (int32_t p : chain_srcs[c]) {
            cur.cci.push_back(c);
            cur.cpos.push_back(p);
            cur.cslots.push_back(slot[L]);
            cur.cslots.push_back(slot[R]);
            plan.n_chain++;
        }
    }
    flush();
    // endpoints that wanted BC but were never placed (cannot happen: every
    // needed endpoint is in some batch) -- keep the invariant explicit
    for (int32_t v : everts)
        if (want_bfs[v] && !bc_done[v]) { plan.free_src.push_back(v); plan.n_bfs++; }
    return BCX_OK;
}

// ------------------------------------------------------------------------
// plain path
// ------------------------------------------------------------------------

static int grid_for(bcx_ctx *ctx, int64_t count) {
    int64_t g = (int64_t)ctx->n_cu * ctx->plain_wg_per_cu;
    if (count < g) g = count;
    return (int)std::max<int64_t>(g, 1);
}

static int ensure_plain_scratch(bcx_ctx *ctx, int32_t n, int grid, PlainScratch &sc) {
    const int64_t stride = ((int64_t)n + 2 + 63) & ~(int64_t)63;
    const size_t per = (size_t)stride * (4 + 8 + 8 + 4 + 4);
    int rc = ctx->plain_scratch.ensure(ctx, per * (size_t)grid, "plain scratch");
    if (rc) return rc;
    char *p = (char *)ctx->plain_scratch.p;
    const size_t g = (size_t)grid * stride;
    sc.stride = stride;
    sc.sig = (uint64_t *)p; p += g * 8;
    sc.dl = (double *)p;    p += g * 8;
    sc.d = (int32_t *)p;    p += g * 4;
    sc.S = (int32_t *)p;    p += g * 4;
    sc.off = (int32_t *)p;
    return BCX_OK;
}

static int ensure_rows(bcx_ctx *ctx, int32_t n, int32_t nrows, Rows &rows) {
    const int64_t stride = ((int64_t)n + 63) & ~(int64_t)63;
    const size_t bytes = (size_t)stride * 12 * (size_t)std::max(nrows, 1);
    int rc = ctx->rows_buf.ensure(ctx, bytes, "endpoint rows");
    if (rc) return rc;
    rows.stride = stride;
    rows.sig = (uint64_t *)ctx->rows_buf.p;
    rows.d = (int32_t *)((char *)ctx->rows_buf.p + (size_t)stride * 8 * std::max(nrows, 1));
    return BCX_OK;
}

int64_t row_capacity(bcx_ctx *ctx, int32_t n, const bcx_opts *o) {
    const int64_t per = (((int64_t)n + 63) & ~(int64_t)63) * 12;
    int64_t cap = (o && o->row_cap_bytes > 0) ? o->row_cap_bytes : (int64_t)(ctx->mem_cap / 4);
    int64_t r = cap / std::max<int64_t>(per, 1);
    return std::max<int64_t>(2, std::min<int64_t>(r, 1 << 16));
}

static int run_bfs_list(bcx_ctx *ctx, const DComp &dc, const std::vector<int32_t> &src,
                        const std::vector<int32_t> &slot, const std::vector<int8_t> &mode,
                        bool f64, uint64_t limit, double *bc, const Rows &rows, bcx_stats *st) {
    if (src.empty()) return BCX_OK;
    int rc;
    if ((rc = upload(ctx, src.data(), src.size() * 4, ctx->w_src, "work list"))) return rc;
    if ((rc = upload(ctx, slot.data(), slot.size() * 4, ctx->w_aux, "work list"))) return rc;
    if ((rc = upload(ctx, mode.data(), mode.size(), ctx->w_mode, "work list"))) return rc;
    rc = ctx_hip(ctx, hipMemsetAsync(ctx->counters.p, 0, 4 * sizeof(int32_t), ctx->stream), "memset");
    if (rc) return rc;
    const int grid = grid_for(ctx, (int64_t)src.size());
    PlainScratch sc;
    if ((rc = ensure_plain_scratch(ctx, dc.n, grid, sc))) return rc;
    PlainArgs a{};
    a.src = (const int32_t *)ctx->w_src.p;
    a.aux = (const int32_t *)ctx->w_aux.p;
    a.mode = (const int8_t *)ctx->w_mode.p;
    a.count = (int32_t)src.size();
    a.sigma_f64 = f64;
    a.limit = limit;
    a.counters = (int32_t *)ctx->counters.p;
    a.bc = bc;
    ctx->timer.begin(ctx);
    hipError_t e = launch_plain_bfs(dc.dev, a, sc, rows, grid, ctx->stream);
    ctx->timer.end(ctx);
    if (st) st->n_launches++;
    rc = ctx_hip(ctx, e, "k_plain_bfs launch");
    if (rc) return rc;
    return read_counters(ctx, st);
}

int read_counters(bcx_ctx *ctx, bcx_stats *st) {
    int32_t h[4];
    int rc = ctx_hip(ctx, hipMemcpyAsync(h, ctx->counters.p, sizeof(h), hipMemcpyDeviceToHost, ctx->stream),
                     "counters");
    if (rc) return rc;
    rc = ctx_hip(ctx, hipStreamSynchronize(ctx->stream), "kernel execution");
    if (rc) return rc;
    ctx->overflow |= h[1] != 0;
    if (st && h[2] > st->max_level) st->max_level = h[2];
    return BCX_OK;
}

static int run_chain_list(bcx_ctx *ctx, const DComp &dc, const PlanBatch &b, bool f64, double *bc,
                          const Rows &rows, bcx_stats *st) {
    if (b.cci.empty()) return BCX_OK;
    int rc;
    if ((rc = upload(ctx, b.cci.data(), b.cci.size() * 4, ctx->w_src, "chain list"))) return rc;
    if ((rc = upload(ctx, b.cpos.data(), b.cpos.size() * 4, ctx->w_aux, "chain list"))) return rc;
    if ((rc = upload(ctx, b.cslots.data(), b.cslots.size() * 4, ctx->w_slots, "chain list"))) return rc;
    rc = ctx_hip(ctx, hipMemsetAsync(ctx->counters.p, 0, 4 * sizeof(int32_t), ctx->stream), "memset");
    if (rc) return rc;
    const int grid = grid_for(ctx, (int64_t)b.cci.size());
    PlainScratch sc;
    if ((rc = ensure_plain_scratch(ctx, dc.n, grid, sc))) return rc;
    PlainArgs a{};
    a.src = (const int32_t *)ctx->w_src.p;
    a.aux = (const int32_t *)ctx->w_aux.p;
    a.slots = (const int32_t *)ctx->w_slots.p;
    a.count = (int32_t)b.cci.size();
    a.sigma_f64 = f64;
    a.counters = (int32_t *)ctx->counters.p;
    a.bc = bc;
    ctx->timer_chain.begin(ctx);
    hipError_t e = launch_plain_chain(dc.dev, dc.chains, a, sc, rows, grid, ctx->stream);
    ctx->timer_chain.end(ctx);
    if (st) st->n_launches++;
    rc = ctx_hip(ctx, e, "k_plain_chain launch");
    if (rc) return rc;
    return read_counters(ctx, st);
}

// One component on the plain path in one sigma mode.
static int plain_component_mode(bcx_ctx *ctx, const DComp &dc, const Plan &plan, bool f64,
                                uint64_t limit, double *bc, bcx_stats *st) {
    int rc;
    Rows rows{};
    int32_t max_rows = 1;
    for (auto &b : plan.batches) max_rows = std::max(max_rows, b.n_rows);
    if ((rc = ensure_rows(ctx, dc.n, max_rows, rows))) return rc;
    std::vector<int32_t> slots(plan.free_src.size(), -1);
    std::vector<int8_t> modes(plan.free_src.size(), MODE_BC);
    if ((rc = run_bfs_list(ctx, dc, plan.free_src, slots, modes, f64, limit, bc, rows, st))) return rc;
    if (ctx->overflow && !f64) return BCX_OK;
    for (auto &b : plan.batches) {
        if ((rc = run_bfs_list(ctx, dc, b.bsrc, b.bslot, b.bmode, f64, limit, bc, rows, st))) return rc;
        if (ctx->overflow && !f64) return BCX_OK;
        if ((rc = run_chain_list(ctx, dc, b, f64, bc, rows, st))) return rc;
    }
    return BCX_OK;
}

int plain_component(bcx_ctx *ctx, const DComp &dc, const Plan &plan, const bcx_opts *o,
                    double *bc, bcx_stats *st) {
    const uint64_t limit = (o && o->sigma_f64_threshold_log2 > 0 && o->sigma_f64_threshold_log2 < 64)
                               ? (1ull << o->sigma_f64_threshold_log2) : 0;
    ctx->overflow = false;
    int rc = plain_component_mode(ctx, dc, plan, false, limit, bc, st);
    if (rc) return rc;
    if (!ctx->overflow) return BCX_OK;
    // sigma left uint64: redo the component with FP64 sigma (pull order)
    rc = ctx_hip(ctx, hipMemsetAsync(bc, 0, sizeof(double) * dc.n, ctx->stream), "memset");
    if (rc) return rc;
    if (st) st->n_f64_sigma += plan.n_bfs + plan.n_chain;
    ctx->overflow = false;
    return plain_component_mode(ctx, dc, plan, true, 0, bc, st);
}

// ------------------------------------------------------------------------
// C-ABI
// ------------------------------------------------------------------------
extern "C" {

int bcx_create(bcx_ctx **out, int device, size_t mem_cap) {
    if (!out) return BCX_E_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return BCX_E_NODEV;
    if (device < 0 || device >= count) return BCX_E_ARG;
    if (hipSetDevice(device) != hipSuccess) return BCX_E_HIP;
    bcx_ctx *ctx = new bcx_ctx();
    ctx->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) { delete ctx; return BCX_E_HIP; }
    ctx->n_cu = prop.multiProcessorCount;
    ctx->name = prop.gcnArchName;
    size_t free_b = 0, total_b = 0;
    (void)hipMemGetInfo(&free_b, &total_b);
    ctx->mem_cap = mem_cap ? mem_cap : (size_t)(0.9 * (double)free_b);
    // a failed lookup is recorded there and reported by bcx_bc_small_noalloc;
    // every other entry point works without the array
    (void)noalloc_init(device);
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return BCX_E_HIP;
    }
    if (ctx->counters.ensure(ctx, 64, "counters") != BCX_OK) {
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return BCX_E_NOMEM;
    }
    *out = ctx;
    return BCX_OK;
}

void bcx_destroy(bcx_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (DBuf *b : ctx->all_bufs()) b->release(ctx);
    batched_release(ctx);
    small_release(ctx);
    ctx->timer.destroy();
    ctx->timer_chain.destroy();
    if (ctx->bc_pin) (void)hipHostFree(ctx->bc_pin);
    if (ctx->side) (void)hipStreamDestroy(ctx->side);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *bcx_last_error(const bcx_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int bcx_device_info(bcx_ctx *ctx, int32_t *n_cu, char *name, int32_t name_len) {
    if (!ctx) return BCX_E_ARG;
    if (n_cu) *n_cu = ctx->n_cu;
    if (name && name_len > 0) {
        std::strncpy(name, ctx->name.c_str(), (size_t)name_len - 1);
        name[name_len - 1] = '\0';
    }
    return BCX_OK;
}

int bcx_upload(bcx_ctx *ctx, const bcx_component *comps, int32_t n_comps, bcx_dgraph **out) {
    if (!ctx || !out || n_comps < 0 || (n_comps > 0 && !comps)) return ctx_fail(ctx, BCX_E_ARG, "bad upload args");
    *out = nullptr;
    (void)hipSetDevice(ctx->device);
    bcx_dgraph *g = new bcx_dgraph();
    int64_t off = 0;
    for (int32_t i = 0; i < n_comps; ++i) {
        const bcx_component &c = comps[i];
        if (c.n <= 0 || !c.row_ptr || !c.col || !c.reach || c.row_ptr[0] != 0 || c.arcs != (int64_t)c.row_ptr[c.n] ||
            c.arcs >= INT32_MAX || (c.n_chains > 0 && (!c.chain_ptr || !c.chain_nodes || !c.chain_ends))) {
            bcx_dgraph_free(ctx, g);
            return ctx_fail(ctx, BCX_E_ARG, "component " + std::to_string(i) + " is malformed");
        }
        g->comps.emplace_back();
        DComp &dc = g->comps.back();
        HostComp &hc = dc.host;
        dc.n = c.n;
        dc.arcs = c.arcs;
        dc.offset = off;
        off += c.n;
        hc.n = c.n;
        hc.arcs = c.arcs;
        hc.rp.assign(c.row_ptr, c.row_ptr + c.n + 1);
        hc.col.assign(c.col, c.col + c.arcs);
        hc.reach.assign(c.reach, c.reach + c.n);
        hc.n_chains = std::max(c.n_chains, 0);
        if (hc.n_chains) {
            hc.ch_ptr.assign(c.chain_ptr, c.chain_ptr + c.n_chains + 1);
            hc.ch_nodes.assign(c.chain_nodes, c.chain_nodes + c.chain_ptr[c.n_chains]);
            hc.ch_ends.assign(c.chain_ends, c.chain_ends + 2 * c.n_chains);
        } else {
            hc.ch_ptr.assign(1, 0);
        }
    }
    g->total_n = off;
    for (const DComp &dc : g->comps) g->offs.push_back(dc.offset);
    // every array of every component in one device allocation and one copy
    // (tens of thousands of small components would otherwise cost an
    // allocation and a copy per array); arrays start on 256-byte boundaries
    auto padded = [](size_t k) { return (k + 63) & ~(size_t)63; };
    size_t words = 0;
    for (DComp &dc : g->comps) {
        const HostComp &hc = dc.host;
        words += padded(hc.rp.size()) + padded(hc.col.size()) + padded(hc.reach.size()) + padded(hc.ch_ptr.size()) +
                 padded(hc.ch_nodes.size()) + padded(hc.ch_ends.size());
    }
    int rc = g->pool.ensure(ctx, 4 * std::max<size_t>(words, 1), "components");
    if (rc) { bcx_dgraph_free(ctx, g); return rc; }
    std::vector<int32_t> stage(words, 0);
    int32_t *base = (int32_t *)g->pool.p;
    size_t at = 0;
    auto place = [&](const std::vector<int32_t> &v) -> const int32_t * {
        std::copy(v.begin(), v.end(), stage.begin() + at);
        const int32_t *p = base + at;
        at += padded(v.size());
        return p;
    };
    for (DComp &dc : g->comps) {
        const HostComp &hc = dc.host;
        dc.dev.n = dc.n;
        dc.dev.arcs = dc.arcs;
        dc.dev.rp = place(hc.rp);
        dc.dev.col = place(hc.col);
        dc.dev.reach = place(hc.reach);
        dc.chains.ptr = place(hc.ch_ptr);
        dc.chains.nodes = place(hc.ch_nodes);
        dc.chains.ends = place(hc.ch_ends);
    }
    if (words && (rc = ctx_hip(ctx, hipMemcpyAsync(g->pool.p, stage.data(), 4 * words, hipMemcpyHostToDevice,
                                                   ctx->stream), "components"))) {
        bcx_dgraph_free(ctx, g);
        return rc;
    }
    rc = ctx_hip(ctx, hipStreamSynchronize(ctx->stream), "upload");
    if (rc) { bcx_dgraph_free(ctx, g); return rc; }
    *out = g;
    return BCX_OK;
}

int bcx_batched_order(bcx_ctx *ctx, bcx_dgraph *g, int32_t c, int32_t *order) {
    if (!ctx || !g || !order || c < 0 || c >= (int32_t)g->comps.size()) return BCX_E_ARG;
    const std::vector<int32_t> o = batched_order(g->comps[c]);
    std::memcpy(order, o.data(), sizeof(int32_t) * o.size());
    return BCX_OK;
}

void bcx_dgraph_free(bcx_ctx *ctx, bcx_dgraph *g) {
    if (!g) return;
    if (ctx) {
        (void)hipStreamSynchronize(ctx->stream);
        for (auto &dc : g->comps)
            for (DBuf *b : {&dc.b_rp, &dc.b_col, &dc.b_reach, &dc.b_chp, &dc.b_chn, &dc.b_che}) b->release(ctx);
        batched_graph_release(ctx, g);
        g->pool.release(ctx);
    }
    delete g;
}

// A slice of the BC accumulator that comes back to the host: bc[src, src +
// len) lands at packed[dst, dst + len).
struct BcChunk {
    int64_t src, dst;
    int32_t len, pad;
};
constexpr int32_t BC_CHUNK = 8192;

__global__ void k_bc_gather(const double *bc, double *packed, const BcChunk *ch) {
    const BcChunk c = ch[blockIdx.x];
    for (int i = threadIdx.x; i < c.len; i += blockDim.x) packed[c.dst + i] = bc[c.src + i];
}

// Adds the run's BC to bc_out.  A run over a source subset leaves the
// slices of components without sources at the zeros they were set to, so
// only the other slices come back: gathered into one packed buffer on the
// device, copied through pinned staging and added slice by slice.
static int download_bc(bcx_ctx *ctx, bcx_dgraph *g, const double *bc, const std::vector<char> *touched,
                       double *bc_out) {
    std::vector<BcChunk> chunks;
    int64_t packed = 0;
    for (size_t ci = 0; ci < g->comps.size(); ++ci) {
        if (touched && !(*touched)[ci]) continue;
        const DComp &dc = g->comps[ci];
        for (int64_t i = 0; i < dc.n; i += BC_CHUNK) {
            BcChunk c{};
            c.src = dc.offset + i;
            c.dst = packed;
            c.len = (int32_t)std::min<int64_t>(BC_CHUNK, dc.n - i);
            chunks.push_back(c);
            packed += c.len;
        }
    }
    if (packed == 0) return BCX_OK;
    const size_t bytes = sizeof(double) * (size_t)packed;
    if (ctx->bc_pin_bytes < bytes) {
        if (ctx->bc_pin) (void)hipHostFree(ctx->bc_pin);
        ctx->bc_pin = nullptr;
        ctx->bc_pin_bytes = 0;
        if (hipHostMalloc((void **)&ctx->bc_pin, bytes, hipHostMallocDefault) == hipSuccess) {
            ctx->bc_pin_bytes = bytes;
        } else {
            // no pinned memory (a low locked-memory limit): pageable staging
            (void)hipGetLastError();
            ctx->bc_pin = nullptr;
        }
    }
    std::vector<double> pageable;
    double *stage = ctx->bc_pin;
    if (!stage) {
        pageable.resize((size_t)packed);
        stage = pageable.data();
    }
    const double *from = bc;
    int rc;
    if (packed < g->total_n) {
        if ((rc = upload(ctx, chunks.data(), sizeof(BcChunk) * chunks.size(), ctx->bc_chunks, "result chunks")) ||
            (rc = ctx->bc_pack.ensure(ctx, bytes, "packed result")))
            return rc;
        hipLaunchKernelGGL(k_bc_gather, dim3((unsigned)chunks.size()), dim3(256), 0, ctx->stream, bc,
                           (double *)ctx->bc_pack.p, (const BcChunk *)ctx->bc_chunks.p);
        if ((rc = ctx_hip(ctx, hipGetLastError(), "result gather"))) return rc;
        from = (const double *)ctx->bc_pack.p;
    }
    if ((rc = ctx_hip(ctx, hipMemcpyAsync(stage, from, bytes, hipMemcpyDeviceToHost, ctx->stream), "bc copy")) ||
        (rc = ctx_hip(ctx, hipStreamSynchronize(ctx->stream), "bc copy")))
        return rc;
    for (const BcChunk &c : chunks) {
        double *o = bc_out + c.src;
        const double *h = stage + c.dst;
        for (int32_t i = 0; i < c.len; ++i) o[i] += h[i];
    }
    return BCX_OK;
}

int bcx_run(bcx_ctx *ctx, bcx_dgraph *g, const bcx_opts *opts, double *bc_out, bcx_stats *st) {
    if (!ctx || !g) return ctx_fail(ctx, BCX_E_ARG, "null context or graph");
    (void)hipSetDevice(ctx->device);
    bcx_opts o{};
    if (opts) o = *opts;
    if (o.path != BCX_PATH_PLAIN && o.path != BCX_PATH_BATCHED)
        return ctx_fail(ctx, BCX_E_ARG, "unknown path");
    bcx_stats local{};
    bcx_stats *s = st ? st : &local;
    std::memset(s, 0, sizeof(*s));
    auto t0 = std::chrono::steady_clock::now();
    // a run that failed part-way may have left work on the side stream and
    // timer intervals open or uncollected
    batched_persistent_reset(ctx);
    ctx->timer.depth = 0;
    ctx->timer.used = 0;
    ctx->timer_chain.depth = 0;
    ctx->timer_chain.used = 0;
    int rc = ctx->bc_acc.ensure(ctx, sizeof(double) * (size_t)std::max<int64_t>(g->total_n, 1), "bc accumulator");
    if (rc) return rc;
    double *bc = (double *)ctx->bc_acc.p;
    rc = ctx_hip(ctx, hipMemsetAsync(bc, 0, sizeof(double) * (size_t)std::max<int64_t>(g->total_n, 1), ctx->stream),
                 "memset");
    if (rc) return rc;
    // subset slots -> per-component local lists
    std::vector<std::vector<int32_t>> sub;
    if (o.src_subset) {
        sub.assign(g->comps.size(), {});
        for (int64_t i = 0; i < o.n_src; ++i) {
            const int64_t slot = o.src_subset[i];
            const size_t ci = (size_t)(std::upper_bound(g->offs.begin(), g->offs.end(), slot) - g->offs.begin());
            if (ci == 0) return ctx_fail(ctx, BCX_E_ARG, "source slot out of range");
            const DComp &dc = g->comps[ci - 1];
            if (slot - dc.offset >= dc.n) return ctx_fail(ctx, BCX_E_ARG, "source slot out of range");
            sub[ci - 1].push_back((int32_t)(slot - dc.offset));
        }
    }
    std::vector<SmallJob> small;  // batched path: components for the one-launch small kernel
    std::vector<SmallJob> wg;     // batched path: deep components for the workgroup-per-tile kernel
    std::vector<SmallJob> multi;  // batched path: components for the multi-launch tiles
    int32_t largest_multi = -1;
    for (size_t ci = 0; ci < g->comps.size(); ++ci) {
        DComp &dc = g->comps[ci];
        const int32_t *sp = o.src_subset ? sub[ci].data() : nullptr;
        int64_t sn = o.src_subset ? (int64_t)sub[ci].size() : 0;
        if (o.src_subset && sn == 0) continue;
        // algorithmic work of the reference's method: BFS sources (free and
        // chain endpoints) and chain sources
        auto account = [&](int64_t n_bfs, int64_t n_chain) {
            s->n_bfs_sources += n_bfs;
            s->n_chain_sources += n_chain;
            s->bytes_alg += (double)n_bfs * (16.0 * dc.arcs + 80.0 * dc.n) +
                            (double)n_chain * (8.0 * dc.arcs + 100.0 * dc.n);
            s->teps_work += (double)(n_bfs + n_chain) * (double)dc.arcs * 0.5;
        };
        if (o.path == BCX_PATH_PLAIN) {
            Plan plan;
            rc = make_plan(dc.host, sp, sn, o.use_chains != 0, row_capacity(ctx, dc.n, &o), plan);
            if (rc) return rc;
            account(plan.n_bfs, plan.n_chain);
            s->n_row_only += plan.n_row_only;
            rc = plain_component(ctx, dc, plan, &o, bc + dc.offset, s);
            if (rc) return rc;
            continue;
        }
        int64_t n_chain = 0;
        std::vector<int32_t> srcs = batched_sources(dc, sp, sn, o.use_chains ? &n_chain : nullptr);
        account((int64_t)srcs.size() - n_chain, n_chain);
        const int kind = batched_schedule(ctx, dc, &o, s);  // (one-time sample runs stay outside the timing)
        if (kind < 0) return kind;
        if (kind == BCX_KERNEL_SMALL) {
            s->n_comp_small++;
            small.push_back({&dc, std::move(srcs)});
            continue;
        }
        // larger components: tile order from the tuner (BFS ranks or balls)
        std::vector<int32_t> tiles = tuned_sources(dc, srcs);
        if (dc.tile_order == 1) s->n_ball_tiles++;
        if (kind == BCX_KERNEL_WG) {
            s->n_comp_wg++;
            wg.push_back({&dc, std::move(tiles)});
        } else {
            if ((rc = batched_prepare(ctx, dc, &o))) return rc;  // one-time work stays outside the timing
            s->n_comp_multi++;
            if (dc.n > largest_multi) {
                largest_multi = dc.n;
                batched_sched_report(dc, &o, &s->sched_blocks_per_cu, &s->sched_slots, &s->sched_sources_per_lane);
            }
            multi.push_back({&dc, std::move(tiles)});
        }
    }
    if (o.path == BCX_PATH_BATCHED && (!small.empty() || !wg.empty() || !multi.empty())) {
        // the persistent passes (small components, deep components) run on a
        // side stream while the multi-launch tiles run on theirs; the run is
        // timed as one interval on the context stream around both
        const bool overlap = !multi.empty() && (!small.empty() || !wg.empty());
        if (overlap && !ctx->side &&
            (rc = ctx_hip(ctx, hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking), "side stream")))
            return rc;
        const hipStream_t sp = overlap ? ctx->side : ctx->stream;
        ctx->timer.begin(ctx);
        if (!small.empty() && (rc = batched_persistent_start(ctx, small, &o, bc, s, false, sp))) return rc;
        if (!wg.empty() && (rc = batched_persistent_start(ctx, wg, &o, bc, s, true, sp))) return rc;
        for (SmallJob &jb : multi)
            if ((rc = batched_component(ctx, *const_cast<DComp *>(jb.dc), jb.srcs, &o, bc + jb.dc->offset, s)))
                return rc;
        if ((rc = batched_persistent_finish(ctx, s, false)) || (rc = batched_persistent_finish(ctx, s, true)))
            return rc;
        if (overlap) {
            hipEvent_t ev;
            if ((rc = ctx_hip(ctx, hipEventCreateWithFlags(&ev, hipEventDisableTiming), "event"))) return rc;
            rc = ctx_hip(ctx, hipEventRecord(ev, sp), "event");
            if (!rc) rc = ctx_hip(ctx, hipStreamWaitEvent(ctx->stream, ev, 0), "event");
            (void)hipEventDestroy(ev);
            if (rc) return rc;
        }
        ctx->timer.end(ctx);
    }
    rc = ctx_hip(ctx, hipStreamSynchronize(ctx->stream), "run");
    if (rc) return rc;
    s->t_chain_s = ctx->timer_chain.collect();
    s->t_kernel_s = ctx->timer.collect() + s->t_chain_s;
    if (bc_out && g->total_n > 0) {
        std::vector<char> touched;
        if (o.src_subset) {
            touched.assign(g->comps.size(), 0);
            for (size_t ci = 0; ci < g->comps.size(); ++ci) touched[ci] = !sub[ci].empty();
        }
        if ((rc = download_bc(ctx, g, bc, o.src_subset ? &touched : nullptr, bc_out))) return rc;
    }
    s->t_total_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return BCX_OK;
}

int bcx_bc(bcx_ctx *ctx, const bcx_component *comps, int32_t n_comps, const bcx_opts *opts,
           double *bc_out, bcx_stats *st) {
    bcx_dgraph *g = nullptr;
    int rc = bcx_upload(ctx, comps, n_comps, &g);
    if (rc) return rc;
    rc = bcx_run(ctx, g, opts, bc_out, st);
    bcx_dgraph_free(ctx, g);
    return rc;
}

int bcx_host_bc(bcx_ctx *ctx, const bcx_host *h, const bcx_opts *opts, double *bc_out, bcx_stats *st) {
    if (!ctx || !h || !bc_out || !h->decomposed) return ctx_fail(ctx, BCX_E_ARG, "host graph not decomposed");
    std::vector<bcx_component> comps;
    int64_t total = 0;
    for (auto &c : h->comps) { comps.push_back(c.view()); total += c.n; }
    std::vector<double> local((size_t)std::max<int64_t>(total, 1), 0.0);
    int rc = BCX_OK;
    if (!comps.empty()) rc = bcx_bc(ctx, comps.data(), (int32_t)comps.size(), opts, local.data(), st);
    if (rc) return rc;
    for (int32_t v = 0; v < h->n; ++v) bc_out[v] = (double)h->cut[v];
    int64_t off = 0;
    for (auto &c : h->comps) {
        for (int32_t i = 0; i < c.n; ++i) bc_out[c.gid[i]] += local[off + i];
        off += c.n;
    }
    return BCX_OK;
}

int bcx_debug_source(bcx_ctx *ctx, const bcx_component *comp, int32_t src, const bcx_opts *opts,
                     int32_t *level, uint64_t *sigma_u64, double *sigma_f64, double *delta,
                     int32_t *sigma_is_f64) {
    if (!ctx || !comp || src < 0 || src >= comp->n) return ctx_fail(ctx, BCX_E_ARG, "bad debug args");
    bcx_opts o{};
    if (opts) o = *opts;
    if (o.path == BCX_PATH_BATCHED)
        return batched_debug_tile(ctx, comp, &src, 1, &o, level, sigma_u64, sigma_f64, delta, sigma_is_f64, nullptr);
    bcx_dgraph *g = nullptr;
    int rc = bcx_upload(ctx, comp, 1, &g);
    if (rc) return rc;
    DComp &dc = g->comps[0];
    const int32_t n = dc.n;
    std::vector<int32_t> lv(n);
    std::vector<uint64_t> sg(n);
    std::vector<double> dl(n);
    bool f64 = false;
    const uint64_t limit = (o.sigma_f64_threshold_log2 > 0 && o.sigma_f64_threshold_log2 < 64)
                               ? (1ull << o.sigma_f64_threshold_log2) : 0;
    // is src a chain vertex?
    int32_t ci = -1, pos = -1;
    if (o.use_chains)
        for (int32_t c = 0; c < dc.host.n_chains && ci < 0; ++c)
            for (int32_t j = dc.host.ch_ptr[c]; j < dc.host.ch_ptr[c + 1]; ++j)
                if (dc.host.ch_nodes[j] == src) { ci = c; pos = j - dc.host.ch_ptr[c]; break; }
    rc = ctx->bc_acc.ensure(ctx, sizeof(double) * n, "bc accumulator");
    for (int attempt = 0; attempt < 2 && !rc; ++attempt) {
        ctx->overflow = false;
        Rows rows{};
        if ((rc = ensure_rows(ctx, n, 2, rows))) break;
        PlainScratch sc;
        double *bc = (double *)ctx->bc_acc.p;
        if (ci < 0) {
            std::vector<int32_t> s{src}, sl{0};
            std::vector<int8_t> m{(int8_t)(MODE_BC | MODE_ROW)};
            rc = run_bfs_list(ctx, dc, s, sl, m, f64, limit, bc, rows, nullptr);
            if (rc) break;
            if (ctx->overflow && !f64) { f64 = true; continue; }
            if ((rc = ensure_plain_scratch(ctx, n, 1, sc))) break;
            rc = ctx_hip(ctx, hipMemcpyAsync(lv.data(), rows.d, 4 * n, hipMemcpyDeviceToHost, ctx->stream), "copy");
            if (!rc) rc = ctx_hip(ctx, hipMemcpyAsync(sg.data(), rows.sig, 8 * n, hipMemcpyDeviceToHost, ctx->stream), "copy");
        } else {
            std::vector<int32_t> s{dc.host.ch_ends[2 * ci], dc.host.ch_ends[2 * ci + 1]}, sl{0, 1};
            std::vector<int8_t> m{MODE_ROW, MODE_ROW};
            rc = run_bfs_list(ctx, dc, s, sl, m, f64, limit, bc, rows, nullptr);
            if (rc) break;
            if (ctx->overflow && !f64) { f64 = true; continue; }
            PlanBatch b;
            b.cci = {ci};
            b.cpos = {pos};
            b.cslots = {0, 1};
            rc = run_chain_list(ctx, dc, b, f64, bc, rows, nullptr);
            if (rc) break;
            // the chain launch ran one workgroup: its rows start at offset 0
            if ((rc = ensure_plain_scratch(ctx, n, 1, sc))) break;
            rc = ctx_hip(ctx, hipMemcpyAsync(lv.data(), sc.d, 4 * n, hipMemcpyDeviceToHost, ctx->stream), "copy");
            if (!rc) rc = ctx_hip(ctx, hipMemcpyAsync(sg.data(), sc.sig, 8 * n, hipMemcpyDeviceToHost, ctx->stream), "copy");
        }
        if (!rc) rc = ctx_hip(ctx, hipMemcpyAsync(dl.data(), sc.dl, 8 * n, hipMemcpyDeviceToHost, ctx->stream), "copy");
        if (!rc) rc = ctx_hip(ctx, hipStreamSynchronize(ctx->stream), "copy");
        break;
    }
    (void)ctx->timer.collect();
    (void)ctx->timer_chain.collect();
    bcx_dgraph_free(ctx, g);
    if (rc) return rc;
    dl[src] = 0.0;
    if (level) std::memcpy(level, lv.data(), 4 * n);
    for (int32_t i = 0; i < n; ++i) {
        if (sigma_u64) sigma_u64[i] = f64 ? 0 : sg[i];
        double f;
        std::memcpy(&f, &sg[i], 8);
        if (sigma_f64) sigma_f64[i] = f64 ? f : (double)sg[i];
    }
    if (delta) std::memcpy(delta, dl.data(), 8 * n);
    if (sigma_is_f64) *sigma_is_f64 = f64 ? 1 : 0;
    return BCX_OK;
}

int bcx_debug_tile(bcx_ctx *ctx, const bcx_component *comp, const int32_t *srcs, int32_t ns, const bcx_opts *opts,
                   int32_t *level, uint64_t *sigma_u64, double *sigma_f64, double *delta, int32_t *sigma_is_f64,
                   int32_t *kernel_used) {
    if (!ctx || !comp || !srcs || ns < 1 || ns > BATCH_W * MAX_SPL || comp->n <= 0)
        return ctx_fail(ctx, BCX_E_ARG, "bad debug tile args");
    std::vector<uint8_t> seen((size_t)comp->n, 0);
    for (int32_t i = 0; i < ns; ++i) {
        if (srcs[i] < 0 || srcs[i] >= comp->n || seen[(size_t)srcs[i]])
            return ctx_fail(ctx, BCX_E_ARG, "tile sources must be distinct vertices of the component");
        seen[(size_t)srcs[i]] = 1;
    }
    (void)hipSetDevice(ctx->device);
    bcx_opts o{};
    if (opts) o = *opts;
    return batched_debug_tile(ctx, comp, srcs, ns, &o, level, sigma_u64, sigma_f64, delta, sigma_is_f64,
                              kernel_used);
}

}  // extern "C"
