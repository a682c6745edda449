// Small components (n <= SMALL_N) in one persistent launch: a workgroup
// takes a tile = (component, up to 64 sources) from a work counter and runs
// every level of it with workgroup barriers, so thousands of components of
// uneven size share one launch instead of paying launches per level.
//
// Same method as the multi-launch batched kernels (bcx_batched.hip): lane k
// of a wave64 owns source k; lane masks of three levels and the visited
// set live in LDS (32 B per vertex), and so does a copy of the
// tile's CSR when it fits beside them; sigma[v][64] and the dependency
// terms t[v][64] live in a per-workgroup HBM slab; level lists Q / Qm are
// built by a coalesced scan of the next-level masks.  sigma is pulled over
// predecessors in CSR order (exact u64, FP64 in the oracle's order on
// overflow), the backward is the rho-form sum (rho = delta / sigma) in FP32
// with Kahan compensation or FP64, BC(v) += sum_k (r(s_k)+1) sigma_v rho_v
// (Codes/kernels_free.cu:154-473 computes the same, one source per block).
// A tile whose uint64 sigma overflows adds nothing and is listed for an FP64
// rerun by the host.
#include "bcx_device.h"
#include "bcx_small.h"

#include <algorithm>

namespace bcx {

// Threads per workgroup.  A wave takes one list entry at a time through a
// chain of dependent instructions and loads, so a tile's levels run as fast
// as its waves are many: 16 waves per tile ran the 10k-vertex workload of 157 tiles in
// 0.45x the time of 4 waves (8 waves: 0.67x).
constexpr int SMALL_SB = 1024;
// Gathers per round of an entry's predecessor (successor) loop.  The
// rounds are branch-free over their SUNR slots, and an entry of a small
// component has one or two neighbours on the adjacent level for most of its
// sources: 2 ran the 10k-vertex workload's launch in 0.75x the time of 8
// (4: 0.81x, 1: 0.78x).
constexpr int SUNR = 2;

struct KahanS {
    float s = 0.f, c = 0.f;
    __device__ __forceinline__ void add(float x) {
        const float y = x - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __device__ __forceinline__ double value() const { return (double)s - (double)c; }
};
struct SumS {
    double s = 0.0;
    __device__ __forceinline__ void add(double x) { s += x; }
    __device__ __forceinline__ double value() const { return s; }
};

// broadcasts from a lane index that is uniform across the wave
__device__ __forceinline__ int rl(int x, int src) { return __builtin_amdgcn_readlane(x, src); }
__device__ __forceinline__ uint64_t rl64(uint64_t x, int src) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, src);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), src);
    return ((uint64_t)hi << 32) | lo;
}

// LDS bytes of a launch whose largest component has cap vertices, with csr
// words for a copy of a tile's CSR (0: arcs stay in HBM)
size_t small_lds_bytes(int64_t cap, int64_t csr) { return 32 * (size_t)cap + 4 * ((size_t)cap + 2 + (size_t)csr); }

// Workgroup-shared state of a tile: LDS arrays, the tile's CSR (LDS or HBM)
// and the workgroup's HBM slabs.
template <typename TT>
struct TileCtx {
    uint64_t *vis, *mA, *mB, *mC;
    int *qoff, *cnt, *ovf;  // cnt: two list counters, used by alternate levels
    uint64_t *sig;
    TT *tt;
    int32_t *Q;
    uint64_t *Qm;
    int32_t *plv;
    double *pdl;
};

// One tile (component, <= 64 sources) through every level, with the CSR at
// rp / col: a copy in LDS (PI = int32_t; the loops over neighbours then read
// LDS instead of L2) or the component's in HBM (PI = gint).  Returns false
// when the uint64 sigma overflowed.
template <bool F64, typename TT, int SB, typename PI>
__device__ __forceinline__ bool small_tile(const SmallArgs &a, const SmallTile &tl, const SmallComp &cp,
                                           const PI *rp, const PI *col, const TileCtx<TT> &x) {
    constexpr int SW = SB / 64;  // waves per workgroup
    using Acc = typename std::conditional<sizeof(TT) == 4, KahanS, SumS>::type;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const uint64_t me = 1ull << lane;
    const int n = cp.n;
    uint64_t *vis = x.vis, *sig = x.sig, *Qm = x.Qm;
    int *qoff = x.qoff;
    int32_t *Q = x.Q;
    TT *tt = x.tt;
    // weight (reach + 1) of this lane's source
    const double wsrc = lane < tl.ns ? a.ws[tl.src0 + lane] : 0.0;
    for (int v = tid; v < n; v += SB) {
        vis[v] = 0;
        x.mA[v] = 0;
        x.mB[v] = 0;
        x.mC[v] = 0;
    }
    if (tid == 0) {
        x.cnt[0] = 0;
        x.cnt[1] = 0;
        *x.ovf = 0;
        qoff[0] = 0;
        qoff[1] = tl.ns;
    }
    __syncthreads();
    if (tid < tl.ns) {
        const int s = a.srcs[tl.src0 + tid];
        vis[s] = 1ull << tid;
        x.mA[s] = 1ull << tid;
        sig[(int64_t)s * 64 + tid] = F64 ? d2u(1.0) : 1ull;
        Q[tid] = s;
        Qm[tid] = 1ull << tid;
        if (x.plv) x.plv[(int64_t)s * 64 + tid] = 0;
    }
    __syncthreads();
    // three mask arrays in rotation: level l's (cur), level l+1's being
    // built (nxt, all zero at the start of the level) and level l-1's
    // (old), which the level clears while it expands; the clearing then
    // needs no phase and barrier of its own
    uint64_t *cur = x.mA, *nxt = x.mB, *old = x.mC;
    int l = 0;
    // ----------------------------------------------------------- forward
    while (true) {
        const int q0 = qoff[l], q1 = qoff[l + 1];
        for (int v = tid; v < n; v += SB) old[v] = 0;
        // expand: lane per entry, a whole wave for a larger entry
        for (int base = q0 + wv * 64; base < q1; base += SB) {
            const int e = base + lane;
            int j0 = 0, j1 = 0;
            uint64_t F = 0;
            if (e < q1) {
                const int v = Q[e];
                F = Qm[e];
                j0 = rp[v];
                j1 = rp[v + 1];
            }
            if (j1 - j0 <= 16)
                for (int j = j0; j < j1; ++j) {
                    const int w = col[j];
                    const uint64_t m = F & ~vis[w];
                    if (m) atomicOr((unsigned long long *)&nxt[w], (unsigned long long)m);
                }
            uint64_t bal = __ballot(j1 - j0 > 16);
            while (bal) {
                const int src = __ffsll((long long)bal) - 1;
                bal &= bal - 1;
                const uint64_t FF = rl64(F, src);
                const int a0 = rl(j0, src), a1 = rl(j1, src);
                for (int j = a0 + lane; j < a1; j += 64) {
                    const int w = col[j];
                    const uint64_t m = FF & ~vis[w];
                    if (m) atomicOr((unsigned long long *)&nxt[w], (unsigned long long)m);
                }
            }
        }
        __syncthreads();
        // scan: non-zero next-level masks become Q_{l+1}
        for (int base = wv * 64; base < n; base += SB) {
            const int w = base + lane;
            const uint64_t N = w < n ? nxt[w] : 0;
            const uint64_t bal = __ballot(N != 0);
            if (!bal) continue;
            int pos = 0;
            if (lane == 0) pos = atomicAdd(&x.cnt[l & 1], __popcll(bal));
            pos = rl(pos, 0) + __popcll(bal & (me - 1));
            if (N) {
                Q[q1 + pos] = w;
                Qm[q1 + pos] = N;
            }
        }
        __syncthreads();
        // (the other counter, last read two levels ago, starts the next one)
        const int q2 = q1 + x.cnt[l & 1];
        if (tid == 0) {
            qoff[l + 2] = q2;
            x.cnt[(l + 1) & 1] = 0;
        }
        if (q2 == q1) break;
        // pull: sigma of the new level over predecessors (CSR order)
        for (int e = q1 + wv; e < q2; e += SW) {
            const int w = Q[e];
            const uint64_t N = Qm[e];
            const int a0 = rp[w], a1 = rp[w + 1];
            uint64_t acc = 0;
            double facc = 0.0;
            bool ovf = false;
            for (int c = a0; c < a1; c += 64) {
                const int j = c + lane;
                int v = 0;
                uint64_t f = 0;
                if (j < a1) {
                    v = col[j];
                    f = cur[v] & N;
                }
                uint64_t bal = __ballot(f != 0);
                while (bal) {
                    uint64_t xs[SUNR];
                    bool use[SUNR];
#pragma unroll
                    for (int u = 0; u < SUNR; ++u) {
                        use[u] = false;
                        xs[u] = 0;
                        if (bal) {
                            const int src = __ffsll((long long)bal) - 1;
                            bal &= bal - 1;
                            const int vv = rl(v, src);
                            if (rl64(f, src) & me) {
                                use[u] = true;
                                xs[u] = sig[(int64_t)vv * 64 + lane];
                            }
                        }
                    }
#pragma unroll
                    for (int u = 0; u < SUNR; ++u) {
                        if (!use[u]) continue;
                        if (F64) {
                            facc += u2d(xs[u]);
                        } else {
                            const uint64_t s = acc + xs[u];
                            ovf |= (s < acc);
                            acc = s;
                        }
                    }
                }
            }
            if (N & me) {
                sig[(int64_t)w * 64 + lane] = F64 ? d2u(facc) : acc;
                if (!F64 && (ovf || (a.limit && acc >= a.limit))) *x.ovf = 1;
                if (x.plv) x.plv[(int64_t)w * 64 + lane] = l + 1;
            }
            if (lane == 0) vis[w] |= N;
        }
        __syncthreads();
        uint64_t *t = old;
        old = cur;
        cur = nxt;
        nxt = t;
        l++;
    }
    const int D = l;  // deepest non-empty level
    __syncthreads();
    if (*x.ovf) return false;
    // ---------------------------------------------------------- backward
    // nxt (no level D+1) and old (cleared by level D) are zero, cur holds
    // level D's masks.  The same rotation: the successors' masks (succ),
    // this level's, scattered as its entries finish (own, zero at the start
    // of the level), and the masks of two levels down (stale), cleared
    // meanwhile
    uint64_t *succ = nxt, *own = old, *stale = cur;
    for (int lb = D; lb >= 1; --lb) {
        const int q0 = qoff[lb], q1 = qoff[lb + 1];
        for (int v = tid; v < n; v += SB) stale[v] = 0;
        for (int e = q0 + wv; e < q1; e += SW) {
            const int v = Q[e];
            const uint64_t M = Qm[e];
            if (lane == 0) own[v] = M;
            // the vertex's own words load while its successors are summed
            const uint64_t sx = (M & me) ? sig[(int64_t)v * 64 + lane] : 0;
            const int rv = glb(cp.reach)[v];
            const int a0 = rp[v], a1 = rp[v + 1];
            Acc acc;
            if (lb < D) {
                for (int c = a0; c < a1; c += 64) {
                    const int j = c + lane;
                    int w = 0;
                    uint64_t xm = 0;
                    if (j < a1) {
                        w = col[j];
                        xm = succ[w] & M;
                    }
                    uint64_t bal = __ballot(xm != 0);
                    while (bal) {
                        TT t[SUNR];
                        bool use[SUNR];
#pragma unroll
                        for (int u = 0; u < SUNR; ++u) {
                            use[u] = false;
                            t[u] = 0;
                            if (bal) {
                                const int src = __ffsll((long long)bal) - 1;
                                bal &= bal - 1;
                                const int ww = rl(w, src);
                                if (rl64(xm, src) & me) {
                                    use[u] = true;
                                    t[u] = tt[(int64_t)ww * 64 + lane];
                                }
                            }
                        }
#pragma unroll
                        for (int u = 0; u < SUNR; ++u)
                            if (use[u]) acc.add(t[u]);
                    }
                }
            }
            double contrib = 0.0;
            if (M & me) {
                const double sv = F64 ? u2d(sx) : (double)sx;
                const double s = acc.value();  // rho_v
                const double delta = sv * s;
                contrib = wsrc * delta;
                tt[(int64_t)v * 64 + lane] = (TT)(((double)rv + 1.0) / sv + s);
                if (x.pdl) x.pdl[(int64_t)v * 64 + lane] = delta;
            }
            contrib = wave_sum(contrib);
            if (lane == 0 && contrib != 0.0) gadd(glb(cp.bc) + v, contrib);
        }
        __syncthreads();
        uint64_t *t = stale;
        stale = succ;
        succ = own;
        own = t;
    }
    return true;
}

template <bool F64, typename TT, int SB>
__global__ __launch_bounds__(SB) void k_small(SmallArgs a) {
    // visited set, current / next level masks and level offsets of the
    // launch's largest component (a.nmax vertices), then a.csr_cap words for
    // a copy of the tile's CSR (row pointers, then columns) when it fits
    extern __shared__ uint64_t s_dyn[];
    __shared__ int s_item, s_cnt[2], s_ovf;
    const int cap = (int)a.nmax;
    const int tid = threadIdx.x;
    TileCtx<TT> x;
    x.vis = s_dyn;
    x.mA = s_dyn + cap;
    x.mB = s_dyn + 2 * cap;
    x.mC = s_dyn + 3 * cap;
    x.qoff = (int *)(s_dyn + 4 * cap);
    x.cnt = s_cnt;
    x.ovf = &s_ovf;
    int *lrp = x.qoff + cap + 2;
    x.sig = a.sig + (int64_t)blockIdx.x * a.slab;
    x.tt = (TT *)a.tt + (int64_t)blockIdx.x * a.slab;
    x.Q = a.Q + (int64_t)blockIdx.x * a.slab;
    x.Qm = a.Qm + (int64_t)blockIdx.x * a.slab;
    x.plv = a.lvl ? a.lvl + (int64_t)blockIdx.x * a.slab : nullptr;
    x.pdl = a.dbg ? a.dbg + (int64_t)blockIdx.x * a.slab : nullptr;
    const int n_tiles = a.n_tiles_dev ? *a.n_tiles_dev : a.n_tiles;
    while (true) {
        if (tid == 0) s_item = atomicAdd(&a.counter[0], 1);
        __syncthreads();
        const int item = s_item;
        if (item >= n_tiles) break;
        const SmallTile tl = a.tiles[item];
        const SmallComp cp = a.comps[tl.comp];
        const int n = cp.n;
        const gint *grp = glb(cp.rp), *gcol = glb(cp.col);
        const int r0 = grp[0], m = grp[n] - r0;
        bool done;
        if (a.csr_cap > 0 && (int64_t)n + 1 + m <= a.csr_cap) {
            int *lcol = lrp + n + 1;
            for (int v = tid; v <= n; v += SB) lrp[v] = grp[v] - r0;
            for (int j = tid; j < m; j += SB) lcol[j] = gcol[r0 + j];
            // (small_tile's first barrier orders these stores before use)
            done = small_tile<F64, TT, SB, int32_t>(a, tl, cp, lrp, lcol, x);
        } else {
            done = small_tile<F64, TT, SB, gint>(a, tl, cp, grp, gcol, x);
        }
        if (!done && tid == 0) {
            const int k = atomicAdd(&a.counter[1], 1);
            a.redo[k] = item;
        }
        __syncthreads();
    }
}

// FP64-sigma tiles keep FP64 terms in either precision: 1/sigma can leave
// the f32 range.
hipError_t launch_small(const SmallArgs &a, bool f64, bool fp32, int grid, hipStream_t st) {
    if (a.nmax < 1 || a.nmax > SMALL_N || a.csr_cap < 0) return hipErrorInvalidValue;
    constexpr int SB = SMALL_SB;
    const size_t lds = small_lds_bytes(a.nmax, a.csr_cap);
    if (f64) hipLaunchKernelGGL((k_small<true, double, SB>), dim3(grid), dim3(SB), lds, st, a);
    else if (fp32) hipLaunchKernelGGL((k_small<false, float, SB>), dim3(grid), dim3(SB), lds, st, a);
    else hipLaunchKernelGGL((k_small<false, double, SB>), dim3(grid), dim3(SB), lds, st, a);
    return hipGetLastError();
}

// LDS words for tile CSRs in a launch whose largest component has nmax
// vertices and largest CSR csr words: all of csr when the LDS of a
// workgroup holds it beside the masks, else what is left there, else 0
// (every tile then reads its arcs from HBM).
int64_t small_csr_words(int64_t nmax, int64_t csr) {
    const int64_t room = ((int64_t)SMALL_LDS - (int64_t)small_lds_bytes(nmax, 0)) / 4;
    const int64_t words = std::min(csr, room);
    return words < 64 ? 0 : words;
}

// Resident workgroups per CU of a launch whose largest component has nmax
// vertices and csr CSR words in LDS (LDS and registers of the uint64
// instance, which the FP64 and FP64-term instances match).
int small_blocks_per_cu(int64_t nmax, int64_t csr) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void *)&k_small<false, float, SMALL_SB>, SMALL_SB,
                                                     small_lds_bytes(nmax, csr)) != hipSuccess || nb < 1) {
        (void)hipGetLastError();
        nb = 1;
    }
    return nb;
}

// Zeroes the work counters (a kernel, not a memset, so it can sit in a graph).
__global__ void k_small_reset(int32_t *counter) {
    if (threadIdx.x < 4) counter[threadIdx.x] = 0;
}

hipError_t launch_small_reset(int32_t *counter, hipStream_t st) {
    hipLaunchKernelGGL(k_small_reset, dim3(1), dim3(64), 0, st, counter);
    return hipGetLastError();
}

}  // namespace bcx
