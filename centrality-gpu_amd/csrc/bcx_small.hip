// Small components (n <= SMALL_N) in one persistent launch: a workgroup
// takes a tile = (component, up to 64 sources) from a work counter and runs
// every level of it with workgroup barriers, so thousands of components of
// uneven size share one launch instead of paying launches per level.
//
// Same method as the multi-launch batched kernels (bcx_batched.hip): lane k
// of a wave64 owns source k; lane masks of the current / next level and the
// visited set live in LDS (24 B per vertex); sigma[v][64] and the dependency
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

namespace bcx {

constexpr int SB = 256;          // threads per workgroup
constexpr int SW = SB / 64;      // waves per workgroup
constexpr int SUNR = 8;

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

__device__ __forceinline__ uint64_t sh64(uint64_t x, int src) { return (uint64_t)__shfl((long long)x, src, 64); }

template <bool F64, typename TT>
__global__ __launch_bounds__(SB) void k_small(SmallArgs a) {
    __shared__ uint64_t vis[SMALL_N];
    __shared__ uint64_t mA[SMALL_N];
    __shared__ uint64_t mB[SMALL_N];
    __shared__ int qoff[SMALL_N + 2];
    __shared__ int s_item, s_cnt, s_ovf;
    using Acc = typename std::conditional<sizeof(TT) == 4, KahanS, SumS>::type;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const uint64_t me = 1ull << lane;
    uint64_t *sig = a.sig + (int64_t)blockIdx.x * a.slab;
    TT *tt = (TT *)a.tt + (int64_t)blockIdx.x * a.slab;
    int32_t *Q = a.Q + (int64_t)blockIdx.x * a.slab;
    uint64_t *Qm = a.Qm + (int64_t)blockIdx.x * a.slab;
    int32_t *plv = a.lvl ? a.lvl + (int64_t)blockIdx.x * a.slab : nullptr;
    double *pdl = a.dbg ? a.dbg + (int64_t)blockIdx.x * a.slab : nullptr;
    const int n_tiles = a.n_tiles_dev ? *a.n_tiles_dev : a.n_tiles;
    while (true) {
        if (tid == 0) s_item = atomicAdd(&a.counter[0], 1);
        __syncthreads();
        const int item = s_item;
        if (item >= n_tiles) break;
        const SmallTile tl = a.tiles[item];
        const SmallComp cp = a.comps[tl.comp];
        const int n = cp.n;
        const int32_t *rp = cp.rp, *col = cp.col;
        for (int v = tid; v < n; v += SB) {
            vis[v] = 0;
            mA[v] = 0;
            mB[v] = 0;
        }
        if (tid == 0) {
            s_cnt = 0;
            s_ovf = 0;
            qoff[0] = 0;
            qoff[1] = tl.ns;
        }
        __syncthreads();
        if (tid < tl.ns) {
            const int s = a.srcs[tl.src0 + tid];
            vis[s] = 1ull << tid;
            mA[s] = 1ull << tid;
            sig[(int64_t)s * 64 + tid] = F64 ? d2u(1.0) : 1ull;
            Q[tid] = s;
            Qm[tid] = 1ull << tid;
            if (plv) plv[(int64_t)s * 64 + tid] = 0;
        }
        __syncthreads();
        uint64_t *cur = mA, *nxt = mB;
        int l = 0;
        // ------------------------------------------------------- forward
        while (true) {
            const int q0 = qoff[l], q1 = qoff[l + 1];
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
                    const uint64_t FF = sh64(F, src);
                    const int a0 = __shfl(j0, src, 64), a1 = __shfl(j1, src, 64);
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
                if (lane == 0) pos = atomicAdd(&s_cnt, __popcll(bal));
                pos = __shfl(pos, 0, 64) + __popcll(bal & (me - 1));
                if (N) {
                    Q[q1 + pos] = w;
                    Qm[q1 + pos] = N;
                }
            }
            __syncthreads();
            const int q2 = q1 + s_cnt;
            __syncthreads();
            if (tid == 0) {
                qoff[l + 2] = q2;
                s_cnt = 0;
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
                        uint64_t x[SUNR];
                        bool use[SUNR];
#pragma unroll
                        for (int u = 0; u < SUNR; ++u) {
                            use[u] = false;
                            x[u] = 0;
                            if (bal) {
                                const int src = __ffsll((long long)bal) - 1;
                                bal &= bal - 1;
                                const int vv = __shfl(v, src, 64);
                                if (sh64(f, src) & me) {
                                    use[u] = true;
                                    x[u] = sig[(int64_t)vv * 64 + lane];
                                }
                            }
                        }
#pragma unroll
                        for (int u = 0; u < SUNR; ++u) {
                            if (!use[u]) continue;
                            if (F64) {
                                facc += u2d(x[u]);
                            } else {
                                const uint64_t s = acc + x[u];
                                ovf |= (s < acc);
                                acc = s;
                            }
                        }
                    }
                }
                if (N & me) {
                    sig[(int64_t)w * 64 + lane] = F64 ? d2u(facc) : acc;
                    if (!F64 && (ovf || (a.limit && acc >= a.limit))) s_ovf = 1;
                    if (plv) plv[(int64_t)w * 64 + lane] = l + 1;
                }
                if (lane == 0) vis[w] |= N;
            }
            __syncthreads();
            for (int e = q0 + tid; e < q1; e += SB) cur[Q[e]] = 0;
            __syncthreads();
            uint64_t *t = cur;
            cur = nxt;
            nxt = t;
            l++;
        }
        const int D = l;  // deepest non-empty level
        __syncthreads();
        if (s_ovf) {
            if (tid == 0) {
                const int k = atomicAdd(&a.counter[1], 1);
                a.redo[k] = item;
            }
            __syncthreads();
            continue;
        }
        // ------------------------------------------------------ backward
        // cur still holds level D's masks: clear them so both arrays start at
        // zero; each level then scatters its own masks for the level above
        for (int e = qoff[D] + tid; e < qoff[D + 1]; e += SB) cur[Q[e]] = 0;
        __syncthreads();
        uint64_t *succ = cur, *own = nxt;
        for (int lb = D; lb >= 1; --lb) {
            const int q0 = qoff[lb], q1 = qoff[lb + 1];
            for (int e = q0 + wv; e < q1; e += SW) {
                const int v = Q[e];
                const uint64_t M = Qm[e];
                if (lane == 0) own[v] = M;
                const int a0 = rp[v], a1 = rp[v + 1];
                Acc acc;
                if (lb < D) {
                    for (int c = a0; c < a1; c += 64) {
                        const int j = c + lane;
                        int w = 0;
                        uint64_t x = 0;
                        if (j < a1) {
                            w = col[j];
                            x = succ[w] & M;
                        }
                        uint64_t bal = __ballot(x != 0);
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
                                    const int ww = __shfl(w, src, 64);
                                    if (sh64(x, src) & me) {
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
                    const uint64_t sx = sig[(int64_t)v * 64 + lane];
                    const double sv = F64 ? u2d(sx) : (double)sx;
                    const double s = acc.value();  // rho_v
                    const double delta = sv * s;
                    contrib = a.ws[tl.src0 + lane] * delta;
                    tt[(int64_t)v * 64 + lane] = (TT)(((double)cp.reach[v] + 1.0) / sv + s);
                    if (pdl) pdl[(int64_t)v * 64 + lane] = delta;
                }
                contrib = wave_sum(contrib);
                if (lane == 0 && contrib != 0.0) unsafeAtomicAdd(&cp.bc[v], contrib);
            }
            __syncthreads();
            // level lb+1's masks are done with; level lb's become the successors
            const int c0 = qoff[lb + 1], c1 = qoff[lb + 2];
            for (int e = c0 + tid; e < c1; e += SB) succ[Q[e]] = 0;
            __syncthreads();
            uint64_t *t = succ;
            succ = own;
            own = t;
        }
        __syncthreads();
    }
}

// FP64-sigma tiles keep FP64 terms in either precision: 1/sigma can leave
// the f32 range.
hipError_t launch_small(const SmallArgs &a, bool f64, bool fp32, int grid, hipStream_t st) {
    if (f64) {
        hipLaunchKernelGGL((k_small<true, double>), dim3(grid), dim3(SB), 0, st, a);
    } else {
        if (fp32) hipLaunchKernelGGL((k_small<false, float>), dim3(grid), dim3(SB), 0, st, a);
        else hipLaunchKernelGGL((k_small<false, double>), dim3(grid), dim3(SB), 0, st, a);
    }
    return hipGetLastError();
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
