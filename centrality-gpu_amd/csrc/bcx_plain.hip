// Plain per-source path: one workgroup per source, persistent grid with an
// atomic work counter (the schedule of Codes/kernels_free.cu:154-473 and
// Codes/kernels_active.cu:263-602), written for wave64 gfx950.
//
// Forward: level-synchronous top-down BFS.  A frontier vertex of degree <= 64
// is expanded by one lane, a hub by a whole wave (lanes stride its arcs).
// Discovery by CAS on the level word, sigma pushed with 64-bit atomics (exact
// integer sums, order free) and the wrap of any partial sum flagged; in FP64
// mode sigma is pulled per vertex in CSR order instead (the oracle's order,
// so FP64 sigma is bit-exact).
// Backward: successor method, delta_w = sum_v sigma_w/sigma_v (r(v)+1+delta_v)
// in FP64, pulled per vertex (Codes/kernels_free.cu:378-421); BC[w] gets
// (r(s)+1) delta_w through FP64 atomics (Codes/kernels_free.cu:423-426).
// Chain sources: combine of two endpoint rows (Codes/postprocesses.cu:276-363)
// and level buckets by an LDS counting sort in place of the CUB radix sort
// (Codes/postprocesses.cu:143-155), then the same backward.
#include "bcx_device.h"
#include "bcx_kernels.h"

namespace bcx {

constexpr int PB = 256;            // threads per workgroup
constexpr int PW = PB / 64;        // waves per workgroup
constexpr int HUB = 64;            // degree above which a wave expands a vertex
constexpr int LDS_LEVELS = 4096;   // level histogram held in LDS

struct FwdShared {
    int cnt;
    int ovf;
};

__device__ __forceinline__ uint64_t sig_one(bool f64) { return f64 ? d2u(1.0) : 1ull; }

__device__ __forceinline__ void visit(const CompDev &c, int32_t *d, uint64_t *sg, int32_t *S,
                                      int w, int lvl, int tail, uint64_t sv, bool f64,
                                      uint64_t limit, FwdShared *sh) {
    int dw = ald(&d[w]);
    if (dw == BCX_INF) {
        int old = atomicCAS(&d[w], BCX_INF, lvl + 1);
        if (old == BCX_INF) {
            int pos = atomicAdd(&sh->cnt, 1);
            S[tail + pos] = w;
            dw = lvl + 1;
        } else {
            dw = old;
        }
    }
    if (!f64 && dw == lvl + 1) {
        uint64_t old = atomicAdd((unsigned long long *)&sg[w], (unsigned long long)sv);
        uint64_t nw = old + sv;
        if (nw < old || (limit && nw >= limit)) sh->ovf = 1;
    }
}

// Forward BFS from s.  Returns the deepest level; S holds the vertices in
// level order with off[l]..off[l+1] the range of level l.
__device__ int plain_forward(const CompDev &c, int32_t s, int32_t *d, uint64_t *sg, int32_t *S,
                             int32_t *off, bool f64, uint64_t limit, FwdShared *sh) {
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    for (int k = tid; k < c.n; k += PB) {
        ast(&d[k], BCX_INF);
        ast(&sg[k], (uint64_t)0);
    }
    __syncthreads();
    if (tid == 0) {
        ast(&d[s], 0);
        ast(&sg[s], sig_one(f64));
        S[0] = s;
        off[0] = 0;
        off[1] = 1;
        sh->cnt = 0;
    }
    __syncthreads();
    int head = 0, tail = 1, lvl = 0;
    while (true) {
        for (int i = head + tid; i < tail; i += PB) {
            int v = S[i];
            int j0 = c.rp[v], j1 = c.rp[v + 1];
            if (j1 - j0 > HUB) continue;
            uint64_t sv = f64 ? 0 : ald(&sg[v]);
            for (int j = j0; j < j1; ++j) visit(c, d, sg, S, c.col[j], lvl, tail, sv, f64, limit, sh);
        }
        for (int i = head + wave; i < tail; i += PW) {
            int v = S[i];
            int j0 = c.rp[v], j1 = c.rp[v + 1];
            if (j1 - j0 <= HUB) continue;
            uint64_t sv = f64 ? 0 : ald(&sg[v]);
            for (int j = j0 + lane; j < j1; j += 64) visit(c, d, sg, S, c.col[j], lvl, tail, sv, f64, limit, sh);
        }
        __syncthreads();
        const int cnt = sh->cnt;
        __syncthreads();
        if (cnt == 0) break;
        if (tid == 0) {
            sh->cnt = 0;
            off[lvl + 2] = tail + cnt;
        }
        head = tail;
        tail += cnt;
        lvl++;
        if (f64) {
            // pull: sigma_w = sum over predecessors in CSR order
            for (int i = head + tid; i < tail; i += PB) {
                int w = S[i];
                double acc = 0.0;
                for (int j = c.rp[w]; j < c.rp[w + 1]; ++j) {
                    int v = c.col[j];
                    if (ald(&d[v]) == lvl - 1) acc += u2d(ald(&sg[v]));
                }
                ast(&sg[w], d2u(acc));
            }
        }
        __syncthreads();
    }
    return lvl;
}

__device__ __forceinline__ double sigd(const uint64_t *sg, int v, bool f64) {
    uint64_t x = ald(&sg[v]);
    return f64 ? u2d(x) : (double)x;
}

// Successor-method dependency, FP64, levels maxlvl-1 .. 1; adds ws*delta to bc.
__device__ void plain_backward(const CompDev &c, int maxlvl, const int32_t *d, const uint64_t *sg,
                               double *dl, const int32_t *S, const int32_t *off, bool f64,
                               double ws, double *bc) {
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    for (int i = off[maxlvl] + tid; i < off[maxlvl + 1]; i += PB) dl[S[i]] = 0.0;
    __syncthreads();
    for (int l = maxlvl - 1; l >= 1; --l) {
        const int a = off[l], b = off[l + 1];
        for (int i = a + tid; i < b; i += PB) {
            int w = S[i];
            int j0 = c.rp[w], j1 = c.rp[w + 1];
            if (j1 - j0 > HUB) continue;
            double sw = sigd(sg, w, f64), acc = 0.0;
            for (int j = j0; j < j1; ++j) {
                int v = c.col[j];
                if (ald(&d[v]) == l + 1)
                    acc += sw / sigd(sg, v, f64) * ((double)c.reach[v] + 1.0 + dl[v]);
            }
            dl[w] = acc;
            if (acc != 0.0) unsafeAtomicAdd(&bc[w], ws * acc);
        }
        for (int i = a + wave; i < b; i += PW) {
            int w = S[i];
            int j0 = c.rp[w], j1 = c.rp[w + 1];
            if (j1 - j0 <= HUB) continue;
            double sw = sigd(sg, w, f64), acc = 0.0;
            for (int j = j0 + lane; j < j1; j += 64) {
                int v = c.col[j];
                if (ald(&d[v]) == l + 1)
                    acc += sw / sigd(sg, v, f64) * ((double)c.reach[v] + 1.0 + dl[v]);
            }
            acc = wave_sum(acc);
            if (lane == 0) {
                dl[w] = acc;
                if (acc != 0.0) unsafeAtomicAdd(&bc[w], ws * acc);
            }
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(PB) void k_plain_bfs(CompDev c, PlainArgs a, PlainScratch sc, Rows rows) {
    __shared__ FwdShared sh;
    __shared__ int s_item, s_stop;
    const int tid = threadIdx.x;
    const int64_t base = (int64_t)blockIdx.x * sc.stride;
    const bool f64 = a.sigma_f64 != 0;
    if (tid == 0) sh.ovf = 0;
    while (true) {
        if (tid == 0) {
            s_item = atomicAdd(&a.counters[0], 1);
            s_stop = (!f64 && ald(&a.counters[1]) != 0) ? 1 : 0;
        }
        __syncthreads();
        const int item = s_item;
        if (item >= a.count || s_stop) break;
        const int s = a.src[item];
        const int slot = a.aux[item];
        const int mode = a.mode[item];
        int32_t *d = (mode & MODE_ROW) ? rows.d + (int64_t)slot * rows.stride : sc.d + base;
        uint64_t *sg = (mode & MODE_ROW) ? rows.sig + (int64_t)slot * rows.stride : sc.sig + base;
        int32_t *S = sc.S + base, *off = sc.off + base;
        int maxlvl = plain_forward(c, s, d, sg, S, off, f64, a.limit, &sh);
        if (tid == 0 && sh.ovf) {
            atomicExch(&a.counters[1], 1);
        }
        __syncthreads();
        const int ovf = sh.ovf;
        if (ovf) break;
        if (tid == 0) atomicMax(&a.counters[2], maxlvl);
        if (mode & MODE_BC)
            plain_backward(c, maxlvl, d, sg, sc.dl + base, S, off, f64, (double)c.reach[s] + 1.0, a.bc);
        __syncthreads();
    }
}

__global__ __launch_bounds__(PB) void k_plain_chain(CompDev c, ChainDev ch, PlainArgs a,
                                                    PlainScratch sc, Rows rows) {
    __shared__ int s_item, s_max;
    __shared__ int hist[LDS_LEVELS];
    const int tid = threadIdx.x, lane = tid & 63;
    const int64_t base = (int64_t)blockIdx.x * sc.stride;
    const bool f64 = a.sigma_f64 != 0;
    int32_t *d = sc.d + base, *S = sc.S + base, *off = sc.off + base;
    uint64_t *sg = sc.sig + base;
    while (true) {
        if (tid == 0) {
            s_item = atomicAdd(&a.counters[0], 1);
            s_max = 0;
        }
        __syncthreads();
        const int item = s_item;
        if (item >= a.count) break;
        const int ci = a.src[item], pos = a.aux[item];
        const int p0 = ch.ptr[ci], cnt = ch.ptr[ci + 1] - p0;
        const int x = ch.nodes[p0 + pos];
        const int ld = pos + 1, rd = cnt - pos;
        const int32_t *dL = rows.d + (int64_t)a.slots[2 * item] * rows.stride;
        const int32_t *dR = rows.d + (int64_t)a.slots[2 * item + 1] * rows.stride;
        const uint64_t *sL = rows.sig + (int64_t)a.slots[2 * item] * rows.stride;
        const uint64_t *sR = rows.sig + (int64_t)a.slots[2 * item + 1] * rows.stride;
        // combine (Codes/postprocesses.cu:329-351)
        int mymax = 0;
        for (int k = tid; k < c.n; k += PB) {
            int la = dL[k] + ld, lb = dR[k] + rd;
            int lv;
            uint64_t sv;
            if (la < lb) { lv = la; sv = sL[k]; }
            else if (la > lb) { lv = lb; sv = sR[k]; }
            else { lv = lb; sv = f64 ? d2u(u2d(sL[k]) + u2d(sR[k])) : sL[k] + sR[k]; }
            ast(&d[k], lv);
            ast(&sg[k], sv);
            mymax = max(mymax, lv);
        }
        mymax = wave_max(mymax);
        if (lane == 0) atomicMax(&s_max, mymax);
        __syncthreads();
        // own chain: direct distances (Codes/postprocesses.cu:205-274, 354-360)
        for (int j = tid; j < cnt; j += PB) {
            int y = ch.nodes[p0 + j];
            if (j == pos) {
                ast(&d[y], 0);
                ast(&sg[y], sig_one(f64));
                continue;
            }
            int dist = j > pos ? j - pos : pos - j;
            int dy = ald(&d[y]);
            if (dy > dist) {
                ast(&d[y], dist);
                ast(&sg[y], sig_one(f64));
            } else if (dy == dist) {
                uint64_t sy = ald(&sg[y]);
                ast(&sg[y], f64 ? d2u(u2d(sy) + 1.0) : sy + 1);
            }
        }
        // final levels are BFS distances (< n); the combine maximum can be
        // larger on the source's own chain before the walk above fixes it
        const int maxl = min(s_max, c.n - 1);
        const bool lds = maxl + 2 <= LDS_LEVELS;
        int *h = lds ? hist : off;  // off doubles as the histogram when deep
        for (int l = tid; l < maxl + 2; l += PB) h[l] = 0;
        __syncthreads();
        for (int k = tid; k < c.n; k += PB) atomicAdd(&h[ald(&d[k]) + 1], 1);
        __syncthreads();
        if (tid == 0) {
            int run = 0;
            for (int l = 0; l < maxl + 2; ++l) {
                run += h[l];
                h[l] = run;        // h[l] = start of level l (exclusive prefix over l+1 bins)
                off[l] = run;
            }
        }
        __syncthreads();
        // h[l] now = number of vertices with level < l; use as cursors
        for (int k = tid; k < c.n; k += PB) {
            int l = ald(&d[k]);
            int p = atomicAdd(&h[l], 1);
            S[p] = k;
        }
        __syncthreads();
        if (!lds) {
            // off was consumed as cursors; rebuild it from the sorted levels
            if (tid == 0) off[0] = 0;
            for (int i = tid; i < c.n; i += PB) {
                int li = ald(&d[S[i]]);
                int lp = i ? ald(&d[S[i - 1]]) : -1;
                for (int l = lp + 1; l <= li; ++l) off[l] = i;
                if (i == c.n - 1)
                    for (int l = li + 1; l <= maxl + 1; ++l) off[l] = c.n;
            }
            __syncthreads();
        }
        // the deepest non-empty level
        int top = maxl;
        while (top > 0 && off[top] == off[top + 1]) --top;
        if (tid == 0) atomicMax(&a.counters[2], top);
        plain_backward(c, top, d, sg, sc.dl + base, S, off, f64, (double)c.reach[x] + 1.0, a.bc);
        __syncthreads();
    }
}

hipError_t launch_plain_bfs(const CompDev &c, const PlainArgs &a, const PlainScratch &sc,
                            const Rows &rows, int grid, hipStream_t st) {
    hipLaunchKernelGGL(k_plain_bfs, dim3(grid), dim3(PB), 0, st, c, a, sc, rows);
    return hipGetLastError();
}

hipError_t launch_plain_chain(const CompDev &c, const ChainDev &ch, const PlainArgs &a,
                              const PlainScratch &sc, const Rows &rows, int grid, hipStream_t st) {
    hipLaunchKernelGGL(k_plain_chain, dim3(grid), dim3(PB), 0, st, c, ch, a, sc, rows);
    return hipGetLastError();
}

}  // namespace bcx
