// Batched multi-source Brandes for gfx950: one tile = one component and up
// to 64 sources, one source per lane of a wave64.  State is vertex-major so a
// single gather serves every source of the tile:
//   vis[v]        u64  lanes that have reached v
//   mb[l % 3][v]  u64  lanes for which v is at level l (three rotating dense
//                      mask arrays; only entries of the level lists are ever
//                      non-zero, cleared incrementally)
//   sig[v][64]    u64 (exact) or f64 (overflow mode) path counts per lane
//   tt[v][64]     f32 or f64 backward term t_v = (r(v)+1)/sigma_v + rho_v
//   Q / Qm        per-level lists of (vertex, lane mask), levels concatenated,
//                 level l = Q[qoff[l], qoff[l+1])
//
// Forward level l (two launches):
// expand pushes lane masks from Q_l to neighbours with one
// 64-bit atomic OR per (arc, new lanes) and appends newly touched vertices
// (wave-aggregated) to Q_{l+1}; pull then computes, for each w in Q_{l+1},
// sigma_w = sum of sigma_v over neighbours v in CSR order for the lanes where
// v is at level l -- no sigma atomics, and the FP64 order is the oracle's, so
// FP64 sigma is bit-exact (Codes/kernels_free.cu:286-362 computes the same
// quantities one source per block with atomics).
// Backward level l: rho_v = sum over successors w of t_w, with
// rho = delta / sigma, so the successor-method term
// sigma_v/sigma_w (r(w)+1 + delta_w) (Codes/kernels_free.cu:378-421) becomes a
// plain sum of one gathered word per (arc, lane); FP32 with Kahan
// compensation (or FP64).  BC(v) += sum_lanes (r(s)+1) sigma_v rho_v
// (Codes/kernels_free.cu:423-426), one FP64 atomic per (vertex, level).
//
// Every kernel reads its list ranges from the device (qoff, ctr) and returns
// at once when the forward has finished (ctr[3] >= 0), so the host launches a
// tile as a fixed sequence and synchronises once per tile.
#include "bcx_batched.h"
#include "bcx_device.h"

namespace bcx {

constexpr int BB = 256;  // threads per block of the level kernels
constexpr int UNR = 8;   // gathers in flight per lane in the sum loops

__device__ __forceinline__ int gwave() { return (int)((blockIdx.x * (unsigned)blockDim.x + threadIdx.x) >> 6); }
__device__ __forceinline__ int nwaves() { return (int)((gridDim.x * (unsigned)blockDim.x) >> 6); }

__device__ __forceinline__ uint64_t shfl64(uint64_t x, int src) {
    return (uint64_t)__shfl((long long)x, src, 64);
}

__device__ __forceinline__ uint64_t *mbuf(const BatchDev &b, int l) {
    const int r = l % 3;
    return r == 0 ? b.mb0 : (r == 1 ? b.mb1 : b.mb2);
}

// ---------------------------------------------------------------- init ---
__global__ __launch_bounds__(BB) void k_b_clear_vis(BatchDev b) {
    for (int64_t v = blockIdx.x * (int64_t)BB + threadIdx.x; v < b.n; v += (int64_t)gridDim.x * BB) b.vis[v] = 0;
}

template <bool F64>
__global__ void k_b_seed(BatchDev b, int ns) {
    const int k = threadIdx.x;
    if (k < ns) {
        const int s = b.srcs[k];
        const uint64_t bit = 1ull << k;
        b.vis[s] = bit;
        b.mb0[s] = bit;
        b.sig[(int64_t)s * 64 + k] = F64 ? d2u(1.0) : 1ull;
        b.Q[k] = s;
        b.Qm[k] = bit;
        if (b.lvl) b.lvl[(int64_t)s * 64 + k] = 0;
    }
    if (k == 0) {
        b.ctr[0] = ns;
        b.ctr[1] = 0;
        b.ctr[2] = 0;
        b.ctr[3] = -1;
        b.qoff[0] = 0;
        b.qoff[1] = ns;
    }
}

// ------------------------------------------------------------- expand ---
// Q_l -> candidates of Q_{l+1}; clears Q_{l-1} from mask array (l+2)%3.
__device__ void expand_level(const BatchDev &b, int l) {
    const int q0 = b.qoff[l], q1 = b.qoff[l + 1];
    const int qm0 = l > 0 ? b.qoff[l - 1] : 0;
    uint64_t *dst = mbuf(b, l + 1);
    uint64_t *clr = mbuf(b, l + 2);
    for (int i = qm0 + blockIdx.x * BB + threadIdx.x; i < q0; i += gridDim.x * BB) clr[b.Q[i]] = 0;
    const int lane = threadIdx.x & 63;
    for (int e = q0 + gwave(); e < q1; e += nwaves()) {
        const int v = b.Q[e];
        const uint64_t F = b.Qm[e];
        const int j0 = b.rp[v], j1 = b.rp[v + 1];
        for (int c = j0; c < j1; c += 64) {
            const int j = c + lane;
            bool fresh = false;
            int w = 0;
            if (j < j1) {
                w = b.col[j];
                const uint64_t m = F & ~b.vis[w];
                if (m) {
                    const uint64_t old = atomicOr((unsigned long long *)&dst[w], (unsigned long long)m);
                    fresh = (old == 0);
                }
            }
            const uint64_t bal = __ballot(fresh);
            if (bal) {
                int base = 0;
                if (lane == 0) base = atomicAdd(&b.ctr[0], __popcll(bal));
                base = __shfl(base, 0, 64);
                if (fresh) {
                    const int pos = base + __popcll(bal & ((1ull << lane) - 1));
                    if (pos < b.qcap) b.Q[pos] = w;
                    else b.ctr[2] = 1;
                }
            }
        }
    }
}

// --------------------------------------------------------------- pull ---
// For w in Q_{l+1}: sigma over predecessors (lanes in mb[l][v]), CSR order.
template <bool F64>
__device__ void pull_level(const BatchDev &b, int l, int q1, int q2) {
    const uint64_t *fr = mbuf(b, l);
    const uint64_t *nx = mbuf(b, l + 1);
    const int lane = threadIdx.x & 63;
    const uint64_t me = 1ull << lane;
    for (int e = q1 + gwave(); e < q2; e += nwaves()) {
        const int w = b.Q[e];
        const uint64_t N = nx[w];
        const int j0 = b.rp[w], j1 = b.rp[w + 1];
        uint64_t acc = 0;
        double facc = 0.0;
        bool ovf = false;
        for (int c = j0; c < j1; c += 64) {
            const int j = c + lane;
            int v = 0;
            uint64_t f = 0;
            if (j < j1) {
                v = b.col[j];
                f = fr[v] & N;
            }
            uint64_t bal = __ballot(f != 0);
            while (bal) {
                uint64_t x[UNR];
                bool use[UNR];
#pragma unroll
                for (int u = 0; u < UNR; ++u) {
                    use[u] = false;
                    x[u] = 0;
                    if (bal) {
                        const int src = __ffsll((long long)bal) - 1;
                        bal &= bal - 1;
                        const int vv = __shfl(v, src, 64);
                        const uint64_t ff = shfl64(f, src);
                        if (ff & me) {
                            use[u] = true;
                            x[u] = b.sig[(int64_t)vv * 64 + lane];
                        }
                    }
                }
#pragma unroll
                for (int u = 0; u < UNR; ++u) {
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
            b.sig[(int64_t)w * 64 + lane] = F64 ? d2u(facc) : acc;
            if (b.lvl) b.lvl[(int64_t)w * 64 + lane] = (uint16_t)(l + 1);
            if (!F64 && (ovf || (b.limit && acc >= b.limit))) b.ctr[1] = 1;
        }
        if (lane == 0) {
            b.vis[w] |= N;
            b.Qm[e] = N;
        }
    }
}

// Level l of the forward is two launches: expand (all of Q_l's pushes must
// land before any pull reads the masks of Q_{l+1}) then pull.  Both return at
// once after the forward has ended (depth recorded in ctr[3]).
__global__ __launch_bounds__(BB) void k_b_expand(BatchDev b, int l) {
    if (__hip_atomic_load(&b.ctr[3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= 0) return;
    expand_level(b, l);
}

template <bool F64>
__global__ __launch_bounds__(BB) void k_b_pull(BatchDev b, int l) {
    if (__hip_atomic_load(&b.ctr[3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= 0) return;
    const int q1 = b.qoff[l + 1];
    const int q2 = min(__hip_atomic_load(&b.ctr[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), b.qcap);
    const bool stop = (q2 == q1) || b.ctr[2] != 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        b.qoff[l + 2] = q2;
        if (stop) b.ctr[3] = l;  // depth: Q_{l+1} is empty (or the lists overflowed)
    }
    if (stop) return;
    pull_level<F64>(b, l, q1, q2);
}

// ----------------------------------------------------------- backward ---
struct KahanF {
    float s = 0.f, c = 0.f;
    __device__ __forceinline__ void add(float x) {
        const float y = x - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __device__ __forceinline__ double value() const { return (double)s - (double)c; }
};
struct SumD {
    double s = 0.0;
    __device__ __forceinline__ void add(double x) { s += x; }
    __device__ __forceinline__ double value() const { return s; }
};

// Level l <= depth: entries Q_l; Q_{l+2} is cleared from mb[(l+2)%3] and Q_l
// scattered into mb[l%3] for level l-1.
template <bool F64, typename TT>
__global__ __launch_bounds__(BB) void k_b_back(BatchDev b, int l) {
    // the host launches the backward speculatively behind the forward: a
    // level beyond the depth, an unfinished forward (depth -1), a sigma
    // overflow (the tile is redone in FP64) or full lists all mean no work
    const int D = __hip_atomic_load(&b.ctr[3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (l > D || (!F64 && b.ctr[1]) || b.ctr[2]) return;
    const int q0 = b.qoff[l], q1 = b.qoff[l + 1];
    const int c0 = l + 2 <= D ? b.qoff[l + 2] : 0, c1 = l + 2 <= D ? b.qoff[l + 3] : 0;
    uint64_t *own = mbuf(b, l);
    const uint64_t *succ = mbuf(b, l + 1);
    uint64_t *clr = mbuf(b, l + 2);
    for (int i = c0 + blockIdx.x * BB + threadIdx.x; i < c1; i += gridDim.x * BB) clr[b.Q[i]] = 0;
    TT *tt = (TT *)b.tt;
    const int lane = threadIdx.x & 63;
    const uint64_t me = 1ull << lane;
    using Acc = typename std::conditional<sizeof(TT) == 4, KahanF, SumD>::type;
    for (int e = q0 + gwave(); e < q1; e += nwaves()) {
        const int v = b.Q[e];
        const uint64_t M = b.Qm[e];
        if (lane == 0) own[v] = M;
        const int j0 = b.rp[v], j1 = b.rp[v + 1];
        Acc acc;
        for (int c = j0; c < j1; c += 64) {
            const int j = c + lane;
            int w = 0;
            uint64_t x = 0;
            if (j < j1) {
                w = b.col[j];
                x = succ[w] & M;
            }
            uint64_t bal = __ballot(x != 0);
            while (bal) {
                TT t[UNR];
                bool use[UNR];
#pragma unroll
                for (int u = 0; u < UNR; ++u) {
                    use[u] = false;
                    t[u] = 0;
                    if (bal) {
                        const int src = __ffsll((long long)bal) - 1;
                        bal &= bal - 1;
                        const int ww = __shfl(w, src, 64);
                        const uint64_t xx = shfl64(x, src);
                        if (xx & me) {
                            use[u] = true;
                            t[u] = tt[(int64_t)ww * 64 + lane];
                        }
                    }
                }
#pragma unroll
                for (int u = 0; u < UNR; ++u)
                    if (use[u]) acc.add(t[u]);
            }
        }
        double contrib = 0.0;
        if (M & me) {
            const uint64_t sx = b.sig[(int64_t)v * 64 + lane];
            const double sv = F64 ? u2d(sx) : (double)sx;
            const double rho = acc.value();
            const double delta = sv * rho;
            contrib = b.ws[lane] * delta;
            tt[(int64_t)v * 64 + lane] = (TT)(((double)b.reach[v] + 1.0) / sv + rho);
            if (b.dbg_delta) b.dbg_delta[(int64_t)v * 64 + lane] = delta;
        }
        contrib = wave_sum(contrib);
        if (lane == 0 && contrib != 0.0) unsafeAtomicAdd(&b.bc[v], contrib);
    }
}

// After the backward only Q_1 (mask array 1) and Q_2 (array 2) are set.
__global__ __launch_bounds__(BB) void k_b_finish(BatchDev b) {
    const int D = b.ctr[3];
    if (D < 0 || b.ctr[2]) return;
    const int a0 = b.qoff[1], a1 = D >= 1 ? b.qoff[2] : a0;
    const int c0 = D >= 2 ? b.qoff[2] : 0, c1 = D >= 2 ? b.qoff[3] : 0;
    for (int i = a0 + blockIdx.x * BB + threadIdx.x; i < a1; i += gridDim.x * BB) b.mb1[b.Q[i]] = 0;
    for (int i = c0 + blockIdx.x * BB + threadIdx.x; i < c1; i += gridDim.x * BB) b.mb2[b.Q[i]] = 0;
}

// Clears every mask entry of Q[0, end) (after an abandoned forward).
__global__ __launch_bounds__(BB) void k_b_clear_all(BatchDev b, int end) {
    for (int i = blockIdx.x * BB + threadIdx.x; i < end; i += gridDim.x * BB) {
        const int v = b.Q[i];
        b.mb0[v] = 0;
        b.mb1[v] = 0;
        b.mb2[v] = 0;
    }
}

// ------------------------------------------------------------ launchers ---
hipError_t launch_b_init(const BatchDev &b, int ns, bool f64, int cap, hipStream_t st) {
    int g = (int)std::min<int64_t>(((int64_t)b.n + BB - 1) / BB, cap);
    hipLaunchKernelGGL(k_b_clear_vis, dim3(std::max(g, 1)), dim3(BB), 0, st, b);
    if (f64) hipLaunchKernelGGL(k_b_seed<true>, dim3(1), dim3(64), 0, st, b, ns);
    else hipLaunchKernelGGL(k_b_seed<false>, dim3(1), dim3(64), 0, st, b, ns);
    return hipGetLastError();
}

hipError_t launch_b_forward(const BatchDev &b, bool f64, int l, int grid, hipStream_t st) {
    hipLaunchKernelGGL(k_b_expand, dim3(grid), dim3(BB), 0, st, b, l);
    if (f64) hipLaunchKernelGGL(k_b_pull<true>, dim3(grid), dim3(BB), 0, st, b, l);
    else hipLaunchKernelGGL(k_b_pull<false>, dim3(grid), dim3(BB), 0, st, b, l);
    return hipGetLastError();
}

hipError_t launch_b_back(const BatchDev &b, bool f64, bool fp32, int l, int grid, hipStream_t st) {
    const dim3 g(grid);
    if (f64) {
        if (fp32) hipLaunchKernelGGL((k_b_back<true, float>), g, dim3(BB), 0, st, b, l);
        else hipLaunchKernelGGL((k_b_back<true, double>), g, dim3(BB), 0, st, b, l);
    } else {
        if (fp32) hipLaunchKernelGGL((k_b_back<false, float>), g, dim3(BB), 0, st, b, l);
        else hipLaunchKernelGGL((k_b_back<false, double>), g, dim3(BB), 0, st, b, l);
    }
    return hipGetLastError();
}

hipError_t launch_b_finish(const BatchDev &b, int grid, hipStream_t st) {
    hipLaunchKernelGGL(k_b_finish, dim3(grid), dim3(BB), 0, st, b);
    return hipGetLastError();
}

hipError_t launch_b_clear_all(const BatchDev &b, int end, int grid, hipStream_t st) {
    if (end <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_b_clear_all, dim3(grid), dim3(BB), 0, st, b, end);
    return hipGetLastError();
}

}  // namespace bcx
