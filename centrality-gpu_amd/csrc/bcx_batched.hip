// Batched multi-source Brandes for gfx950: one tile = one component and up
// to 64 W sources (W = 1, 2 or 4 per lane of a wave64: lane k owns sources
// k + 64 h).  State is vertex-major so a single gather serves every source of
// the tile:
//   vis[v]        W u64  sources that have reached v
//   mb[l % 3][v]  W u64  sources for which v is at level l (three rotating
//                        dense mask arrays; only entries of the level lists are
//                        ever non-zero, cleared incrementally)
//   sig[v][64][W] u32 or u64 (exact) or f64 (overflow mode) path counts per
//                 source; the narrowest word that holds the tile's counts
//   tt[v][64][W]  f32 or f64 backward term t_v = (r(v)+1)/sigma_v + rho_v
//   Q / Qm        per-level lists of (vertex, source mask), levels
//                 concatenated, level l = Q[qoff[l], qoff[l+1])
// With W = 2 a (vertex, level) entry, its arc walk and its mask gathers serve
// 128 sources: the tile reaches a vertex at few more distinct levels than a
// 64-source tile does, so the entries and arcs walked per source fall by
// close to half, while a lane's two row words stay one 8- or 16-byte load.
//
// Forward level l, two launches:
//   expand  pushes source masks from Q_l to neighbours, F & ~vis[w], with
//           fire-and-forget 64-bit atomic ORs into mb[(l+1)%3] (a lane walks
//           the arcs of a low-degree entry, a wave those of a larger one);
//   pull    scans mb[(l+1)%3] in coalesced 64-vertex chunks; every non-zero
//           entry w joins Q_{l+1} and gets sigma_w = sum of sigma_v over
//           neighbours v in CSR order for the sources where v is at level l
//           -- no sigma atomics, and the FP64 order is the oracle's, so FP64
//           sigma is bit-exact.  (Codes/kernels_free.cu:286-362 computes the
//           same quantities one source per block with atomics.)
// Backward level l: rho_v = sum over successors w of t_w, with
// rho = delta / sigma, so the successor-method term
// sigma_v/sigma_w (r(w)+1 + delta_w) (Codes/kernels_free.cu:378-421) becomes a
// plain sum of one gathered word per (arc, source); FP32 with Kahan
// compensation (FP64 in FP64 precision and in FP64-sigma tiles).
// BC(v) += sum_sources (r(s)+1) sigma_v rho_v (Codes/kernels_free.cu:423-426),
// into the slot's own FP64 accumulator (a vertex is finished once per level
// launch, so no atomics), folded into the component at the end of a run.
//
// Hubs (degree > hub_deg) are cut into static arc slices shared by several
// waves: integer sigma partials are added with atomics (exact), rho
// partials into FP64 accumulators and the last slice to finish a hub
// finalises it.  FP64-sigma tiles keep hubs whole in the pull (fixed order).
//
// Every kernel reads its list ranges from the device (qoff, ctr) and returns
// at once when the forward has finished (ctr[3] >= 0), so the host launches a
// tile as a fixed sequence and synchronises once per tile.
#include "bcx_batched.h"
#include "bcx_device.h"

namespace bcx {

// threads per block of the level kernels (2 waves; 8 blocks per CU): 128
// measured 2-3.5% above 256 on power-law graphs, 512 5-7% below 256
constexpr int BB = 128;
constexpr int NW = BB / 64;    // waves per block
constexpr int UNR = 8;         // gathers in flight per lane in the sum loops (fewer for W = 4)
constexpr int SMALL = 16;      // degree up to which one lane expands an entry
// The level kernels wait on dependent gathers, so waves per SIMD (occupancy,
// set by VGPRs and LDS per block) decide their speed.  The staging and window
// sizes keep both kernels' LDS at <= 13.3 KB per block (12 blocks per CU);
// uniform values live in scalar registers (readlane / readfirstlane), u32
// counts in 32-bit registers, and the chain terms are made by their own
// kernel, so with u32 counts and FP32 terms both kernels fit 80 VGPRs: 6
// waves per SIMD.  On the 1M-vertex mixed graph: 176.1k MTEPS at 4 waves,
// 198.6k at 5 (188.9k with the pull alone at 5), 221.7k at 6 (215.7k with
// the same registers at 5).
constexpr int STAGE = 176, STAGE_FLUSH = 112;  // pull: list entries staged per wave, flush point
constexpr int FEB = 8;         // list entries per wave block in the backward
constexpr int WIN = 64;        // arcs per compaction window of a backward block

__device__ __forceinline__ int gwave() { return (int)((blockIdx.x * (unsigned)blockDim.x + threadIdx.x) >> 6); }
__device__ __forceinline__ int nwaves() { return (int)((gridDim.x * (unsigned)blockDim.x) >> 6); }

// x on lane src, for a wave-uniform src (every use here: a ballot's lowest
// set bit, or a constant): v_readlane into a scalar register, where __shfl
// would take a round trip through the LDS crossbar (ds_bpermute) into
// vector registers
__device__ __forceinline__ int bcast(int x, int src) { return __builtin_amdgcn_readlane(x, src); }
__device__ __forceinline__ uint64_t shfl64(uint64_t x, int src) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, src);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), src);
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ double bcastd(double x, int src) { return u2d(shfl64(d2u(x), src)); }
// a wave-uniform value (e.g. read from LDS at a uniform address) moved into
// scalar registers
__device__ __forceinline__ uint64_t uni64(uint64_t x) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)x);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(x >> 32));
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t *mbuf(const BatchDev &b, int l) {
    const int r = l % 3;
    return r == 0 ? b.mb0 : (r == 1 ? b.mb1 : b.mb2);
}

__device__ __forceinline__ bool is_hub(const BatchDev &b, int deg) { return b.n_hub > 0 && deg > b.hub_deg; }

__device__ __forceinline__ int ldi(const int32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ------------------------------------------------------- source masks ---
template <int W>
struct Msk {
    uint64_t m[W];
};
template <int W>
__device__ __forceinline__ Msk<W> mzero() {
    Msk<W> r;
#pragma unroll
    for (int h = 0; h < W; ++h) r.m[h] = 0;
    return r;
}
template <int W>
__device__ __forceinline__ bool many(const Msk<W> &a) {
    uint64_t x = 0;
#pragma unroll
    for (int h = 0; h < W; ++h) x |= a.m[h];
    return x != 0;
}
template <int W>
__device__ __forceinline__ Msk<W> mand(const Msk<W> &a, const Msk<W> &c) {
    Msk<W> r;
#pragma unroll
    for (int h = 0; h < W; ++h) r.m[h] = a.m[h] & c.m[h];
    return r;
}
template <int W>
__device__ __forceinline__ Msk<W> mandn(const Msk<W> &a, const Msk<W> &c) {  // a & ~c
    Msk<W> r;
#pragma unroll
    for (int h = 0; h < W; ++h) r.m[h] = a.m[h] & ~c.m[h];
    return r;
}
// the W bits of lane k: bit h = source k + 64 h
template <int W>
__device__ __forceinline__ int mbits(const Msk<W> &a, int k) {
    int r = 0;
#pragma unroll
    for (int h = 0; h < W; ++h) r |= (int)((a.m[h] >> k) & 1) << h;
    return r;
}
template <int W>
__device__ __forceinline__ Msk<W> mshfl(const Msk<W> &a, int src) {
    Msk<W> r;
#pragma unroll
    for (int h = 0; h < W; ++h) r.m[h] = shfl64(a.m[h], src);
    return r;
}
template <int W>
__device__ __forceinline__ Msk<W> mld(const uint64_t *p, int64_t i) {
    Msk<W> r;
    if constexpr (W == 1) {
        r.m[0] = p[i];
    } else {
#pragma unroll
        for (int q = 0; q < W / 2; ++q) {
            const ulonglong2 y = ((const ulonglong2 *)p)[i * (W / 2) + q];
            r.m[2 * q] = y.x;
            r.m[2 * q + 1] = y.y;
        }
    }
    return r;
}
template <int W>
__device__ __forceinline__ void mst(uint64_t *p, int64_t i, const Msk<W> &a) {
    if constexpr (W == 1) {
        p[i] = a.m[0];
    } else {
#pragma unroll
        for (int q = 0; q < W / 2; ++q) {
            ulonglong2 y;
            y.x = a.m[2 * q];
            y.y = a.m[2 * q + 1];
            ((ulonglong2 *)p)[i * (W / 2) + q] = y;
        }
    }
}
// fire-and-forget OR of the non-zero words (results unused: no return)
template <int W>
__device__ __forceinline__ void mor(uint64_t *p, int64_t i, const Msk<W> &a) {
#pragma unroll
    for (int h = 0; h < W; ++h)
        if (a.m[h]) atomicOr((unsigned long long *)&p[i * W + h], (unsigned long long)a.m[h]);
}
template <int W>
__device__ __forceinline__ Msk<W> mbit(int h, int k) {
    Msk<W> r = mzero<W>();
    r.m[h] = 1ull << k;
    return r;
}

// ------------------------------------------------------------ rows ---
// sigma word i of the tile in mode SM, widened to 64 bits (f64 as its bits)
template <int SM>
__device__ __forceinline__ uint64_t sg(const BatchDev &b, int64_t i) {
    if (SM == SIG_U32) return ((const uint32_t *)b.sig)[i];
    return ((const uint64_t *)b.sig)[i];
}
template <int SM>
__device__ __forceinline__ void sg_put(const BatchDev &b, int64_t i, uint64_t x) {
    if (SM == SIG_U32) ((uint32_t *)b.sig)[i] = (uint32_t)x;
    else ((uint64_t *)b.sig)[i] = x;
}
template <int SM>
__device__ __forceinline__ double sg_val(uint64_t x) {
    return SM == SIG_F64 ? u2d(x) : (double)x;
}
// largest count the mode's word holds
template <int SM>
__device__ __forceinline__ uint64_t sg_max() {
    return SM == SIG_U32 ? 0xffffffffull : ~0ull;
}
// register word of a sigma value in mode SM (32 bits for u32 tiles)
template <int SM>
using SgW = typename std::conditional<SM == SIG_U32, uint32_t, uint64_t>::type;
// the W sigma words of lane cell c = v * 64 + k (one 8- or 16-byte load per
// pair or quad of words)
template <int SM, int W>
__device__ __forceinline__ void sg_row(const BatchDev &b, int64_t c, SgW<SM> (&x)[W]) {
    if constexpr (W == 1) {
        x[0] = (SgW<SM>)sg<SM>(b, c);
    } else if constexpr (SM == SIG_U32 && W == 4) {
        const uint4 y = ((const uint4 *)b.sig)[c];
        x[0] = y.x;
        x[1] = y.y;
        x[2] = y.z;
        x[3] = y.w;
    } else if constexpr (SM == SIG_U32) {
        const uint2 y = ((const uint2 *)b.sig)[c];
        x[0] = y.x;
        x[1] = y.y;
    } else {
#pragma unroll
        for (int q = 0; q < W / 2; ++q) {
            const ulonglong2 y = ((const ulonglong2 *)b.sig)[c * (W / 2) + q];
            x[2 * q] = y.x;
            x[2 * q + 1] = y.y;
        }
    }
}
template <typename TT, int W>
__device__ __forceinline__ void tt_row(const BatchDev &b, int64_t c, TT (&x)[W]) {
    if constexpr (W == 1) {
        x[0] = ((const TT *)b.tt)[c];
    } else if constexpr (sizeof(TT) == 4 && W == 4) {
        const float4 y = ((const float4 *)b.tt)[c];
        x[0] = y.x;
        x[1] = y.y;
        x[2] = y.z;
        x[3] = y.w;
    } else {
        using V2 = typename std::conditional<sizeof(TT) == 4, float2, double2>::type;
#pragma unroll
        for (int q = 0; q < W / 2; ++q) {
            const V2 y = ((const V2 *)b.tt)[c * (W / 2) + q];
            x[2 * q] = y.x;
            x[2 * q + 1] = y.y;
        }
    }
}

// ---------------------------------------------------------------- init ---
template <int W>
__global__ __launch_bounds__(BB) void k_b_clear_vis(BatchDev b, int sm) {
    const int64_t stride = (int64_t)gridDim.x * BB;
    const int64_t t0 = blockIdx.x * (int64_t)BB + threadIdx.x;
    for (int64_t v = t0; v < (int64_t)b.n * W; v += stride) b.vis[v] = 0;
    // hub sigma rows and backward accumulators start at zero every tile: each
    // (hub, source) is summed into exactly once
    constexpr int64_t ROW = 64 * W;
    for (int64_t i = t0; i < (int64_t)b.n_hub * ROW; i += stride) {
        const int64_t k = (int64_t)b.hub_v[i / ROW] * ROW + (i % ROW);
        if (sm == SIG_U32) sg_put<SIG_U32>(b, k, 0);
        else sg_put<SIG_U64>(b, k, 0);
        b.hub_acc[i] = 0.0;
    }
    for (int64_t i = t0; i < b.n_hub; i += stride) b.hub_cnt[i] = 0;
}

// Sources that are tile vertices start at level 0.  A source that is the
// interior x_j of a contracted chain (ids >= n) has no vertex at level 0: the
// expand kernel injects it at L and at R, each with a path count of 1 (its
// own path along the chain) on levels j and K - j, and the forward runs on
// through empty levels while an injection into a vertex the source has not
// yet reached is pending.  (Shifting such a source's levels down so that its
// nearer endpoint enters at level 1 cuts a tile's depth but misaligns it with
// the tile's other sources, which the tile order groups by their true
// distances: 115k against 123k MTEPS on the 1M-vertex mixed graph.)
template <int SM, int W>
__global__ void k_b_seed(BatchDev b, int ns) {
    const int k = threadIdx.x;  // (launched with 64 threads)
    int top = 0, base = 0;
    uint64_t chl[W];
#pragma unroll
    for (int h = 0; h < W; ++h) {
        const int i = h * 64 + k;
        b.wsl[i] = i < ns ? b.ws[i] : 0.0;
        int s = i < ns ? b.srcs[i] : -1;
        int c = -1, j = 0;
        if (s >= b.n) {
            const int x = s - b.n;
            c = b.ch_of[x];
            j = x - b.ch_off[c] + 1;
            top = max(top, max(j, b.ch_off[c + 1] - b.ch_off[c] + 1 - j));
            s = -1;
        }
        if (b.n_chain > 0) {
            b.lane_ch[i] = c;
            b.lane_ch[64 * W + i] = j;
        }
        const uint64_t core = __ballot(s >= 0);
        chl[h] = __ballot(c >= 0);
        if (s >= 0) {
            const uint64_t bit = 1ull << k;
            const int pos = base + __popcll(core & (bit - 1));
            const Msk<W> m = mbit<W>(h, k);
            const int64_t cell = ((int64_t)s * 64 + k) * W + h;
            mst<W>(b.vis, s, m);
            mst<W>(b.mb0, s, m);
            sg_put<SM>(b, cell, SM == SIG_F64 ? d2u(1.0) : 1ull);
            b.Q[pos] = s;
            mst<W>(b.Qm, pos, m);
            if (b.lvl) b.lvl[cell] = 0;
            if (b.n_chain > 0 && b.ep_idx[s] >= 0) b.ep_lvl[((int64_t)b.ep_idx[s] * 64 + k) * W + h] = 0;
        }
        base += __popcll(core);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) top = max(top, __shfl_xor(top, o, 64));
    if (k == 0) {
        b.ctr[0] = base;
        b.ctr[1] = 0;
        b.ctr[2] = 0;
        b.ctr[3] = -1;
        b.ctr[4] = top;  // last level with an injected chain source
        b.ctr[5] = ns;
        b.qoff[0] = 0;
        if (b.n_chain > 0)
#pragma unroll
            for (int h = 0; h < W; ++h) b.lane_mask[h] = chl[h];
    }
}

// ------------------------------------------------------------- expand ---
template <int W>
__device__ __forceinline__ void push(const BatchDev &b, uint64_t *dst, int w, const Msk<W> &F) {
    mor<W>(dst, w, mandn<W>(F, mld<W>(b.vis, w)));
}

// Whether some chain source still has an injection due after level l into
// an endpoint it has not reached (vis is not written during an expand launch,
// so every wave of the launch finds the same answer).
template <int W>
__device__ __forceinline__ bool seed_pending(const BatchDev &b, int l) {
    if (b.n_chain == 0) return false;
    const int lane = threadIdx.x & 63;
    bool p = false;
#pragma unroll
    for (int h = 0; h < W; ++h) {
        const int i = h * 64 + lane;
        const int c = b.lane_ch[i];
        if (c < 0) continue;
        const int j = b.lane_ch[64 * W + i], K = b.ch_off[c + 1] - b.ch_off[c] + 1;
        p |= (j > l && !((b.vis[(int64_t)b.ch_L[c] * W + h] >> lane) & 1)) ||
             (K - j > l && !((b.vis[(int64_t)b.ch_R[c] * W + h] >> lane) & 1));
    }
    return __ballot(p) != 0;
}

// Level l: Q_l = Q[qoff[l], ctr[0]).  Records qoff[l+1]; an empty Q_l with
// no pending injection ends the forward with depth l-1.  Clears Q_{l-1} from
// mask array (l+2)%3.
template <int W>
__global__ __launch_bounds__(BB) void k_b_expand(BatchDev b, int l) {
    if (ldi(&b.ctr[3]) >= 0) return;
    const int q0 = b.qoff[l];
    const int q1 = min(ldi(&b.ctr[0]), b.qcap);
    // an empty level ends the forward; so do full lists and a count that left
    // the sigma word (the host reruns the tile either way)
    bool stop = ldi(&b.ctr[2]) != 0 || ldi(&b.ctr[1]) != 0;
    if (!stop && q1 == q0) stop = l >= b.ctr[4] || !seed_pending<W>(b, l);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        b.qoff[l + 1] = q1;
        if (stop) b.ctr[3] = l - 1;
    }
    if (stop) return;
    uint64_t *dst = mbuf(b, l + 1);
    if (l > 0) {
        uint64_t *clr = mbuf(b, l + 2);
        for (int i = b.qoff[l - 1] + blockIdx.x * BB + threadIdx.x; i < q0; i += gridDim.x * BB)
            mst<W>(clr, b.Q[i], mzero<W>());
    }
    const int lane = threadIdx.x & 63;
    if (b.n_chain > 0 && blockIdx.x == 0 && threadIdx.x < 64) {
        // chain sources whose endpoints are due at level l+1
#pragma unroll
        for (int h = 0; h < W; ++h) {
            const int i = h * 64 + lane;
            const int c = b.lane_ch[i];
            if (c < 0) continue;
            const int j = b.lane_ch[64 * W + i], K = b.ch_off[c + 1] - b.ch_off[c] + 1;
            if (j == l + 1) push<W>(b, dst, b.ch_L[c], mbit<W>(h, lane));
            if (K - j == l + 1) push<W>(b, dst, b.ch_R[c], mbit<W>(h, lane));
        }
    }
    for (int base = q0 + gwave() * 64; base < q1; base += nwaves() * 64) {
        const int e = base + lane;
        int j0 = 0, j1 = 0;
        Msk<W> F = mzero<W>();
        if (e < q1) {
            const int v = b.Q[e];
            F = mld<W>(b.Qm, e);
            j0 = b.rp[v];
            j1 = b.rp[v + 1];
        }
        const int d = j1 - j0;
        if (d <= SMALL)
            for (int j = j0; j < j1; ++j) {
                const int w = b.col[j];
                if (w >= 0) push<W>(b, dst, w, F);  // (w < 0: a contracted chain)
            }
        uint64_t bal = __ballot(d > SMALL && !is_hub(b, d));
        while (bal) {
            const int src = __ffsll((long long)bal) - 1;
            bal &= bal - 1;
            const Msk<W> FF = mshfl<W>(F, src);
            const int a0 = bcast(j0, src), a1 = bcast(j1, src);
            for (int j = a0 + lane; j < a1; j += 64) {
                const int w = b.col[j];
                if (w >= 0) push<W>(b, dst, w, FF);
            }
        }
    }
    // hub slices: the hub's level-l sources are its mask in mb[l%3]
    const uint64_t *cur = mbuf(b, l);
    for (int s = gwave(); s < b.n_slice; s += nwaves()) {
        const int v = b.hub_v[b.sl_hub[s]];
        const Msk<W> F = mld<W>(cur, v);
        if (!many<W>(F)) continue;
        for (int j = b.sl_j0[s] + lane; j < b.sl_j1[s]; j += 64) {
            const int w = b.col[j];
            if (w >= 0) push<W>(b, dst, w, F);
        }
    }
}

// --------------------------------------------------------------- pull ---
// sigma_w for the sources N of entry w: the sum over neighbours at level l
// (fr) in CSR order.  A contracted chain's arc (~c) counts 1 for the source
// of word h whose chain is c and that is injected at w on this level
// (sc[h] == c); cm holds the tile's chain sources.
template <int SM, int W>
__device__ __forceinline__ void sigma_sum(const BatchDev &b, const uint64_t *fr, const Msk<W> &N, int a0, int a1,
                                          SgW<SM> (&acc)[W], double (&facc)[W], bool &ovf, const int (&sc)[W],
                                          const Msk<W> &cm) {
    // gathers in flight (W words each): 2 at W = 4, where 4 in flight cost
    // the pull a wave per SIMD (measured 165.0k against 175.2k MTEPS on the
    // 1M-vertex mixed graph, 276.9k against 295.2k on R-MAT)
    constexpr int UW = W == 4 ? UNR / 4 : UNR;
    const int lane = threadIdx.x & 63;
    for (int c = a0; c < a1; c += 64) {
        const int j = c + lane;
        int v = 0;
        Msk<W> f = mzero<W>();
        if (j < a1) {
            v = b.col[j];
            f = v >= 0 ? mand<W>(mld<W>(fr, v), N) : mand<W>(N, cm);
        }
        uint64_t bal = __ballot(many<W>(f));
        while (bal) {
            SgW<SM> x[UW][W];
            int use[UW];
#pragma unroll
            for (int u = 0; u < UW; ++u) {
                use[u] = 0;
#pragma unroll
                for (int h = 0; h < W; ++h) x[u][h] = 0;
                if (bal) {
                    const int src = __ffsll((long long)bal) - 1;
                    bal &= bal - 1;
                    const int vv = bcast(v, src);
                    const int pb = mbits<W>(mshfl<W>(f, src), lane);
                    if (pb) {
                        if (vv >= 0) {
                            use[u] = pb;
                            sg_row<SM, W>(b, (int64_t)vv * 64 + lane, x[u]);
                        } else {
#pragma unroll
                            for (int h = 0; h < W; ++h)
                                if (((pb >> h) & 1) && ~vv == sc[h]) {
                                    use[u] |= 1 << h;
                                    x[u][h] = (SgW<SM>)(SM == SIG_F64 ? d2u(1.0) : 1ull);
                                }
                        }
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < UW; ++u)
#pragma unroll
                for (int h = 0; h < W; ++h) {
                    if (!((use[u] >> h) & 1)) continue;
                    if (SM == SIG_F64) {
                        facc[h] += u2d(x[u][h]);
                    } else {
                        const SgW<SM> s = acc[h] + x[u][h];  // (wrap: the word overflowed)
                        ovf |= (s < acc[h]);
                        acc[h] = s;
                    }
                }
        }
    }
}

template <int W>
__device__ __forceinline__ void stage_flush(const BatchDev &b, int *sw, Msk<W> *sm, int &cnt, int lane) {
    if (cnt == 0) return;
    int base = 0;
    if (lane == 0) base = atomicAdd(&b.ctr[0], cnt);
    base = bcast(base, 0);
    __builtin_amdgcn_wave_barrier();
    for (int i = lane; i < cnt; i += 64) {
        const int pos = base + i;
        if (pos < b.qcap) {
            b.Q[pos] = sw[i];
            mst<W>(b.Qm, pos, sm[i]);
        } else {
            b.ctr[2] = 1;
        }
    }
    __builtin_amdgcn_wave_barrier();
    cnt = 0;
}

// Builds Q_{l+1} from the non-zero entries of mb[(l+1)%3] and their sigma.
template <int SM, int W>
__global__ __launch_bounds__(BB) void k_b_pull(BatchDev b, int l) {
    constexpr bool F64 = SM == SIG_F64;
    __shared__ int s_w[NW][STAGE];
    __shared__ Msk<W> s_m[NW][STAGE];
    if (ldi(&b.ctr[3]) >= 0) return;
    const uint64_t *fr = mbuf(b, l);
    const uint64_t *nx = mbuf(b, l + 1);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint64_t me = 1ull << lane;
    const bool split = !F64 && b.n_hub > 0;
    // chain sources: their chain and their endpoints injected at level l+1
    // (else -1)
    int myc[W], mL[W], mR[W];
    Msk<W> cm = mzero<W>();
#pragma unroll
    for (int h = 0; h < W; ++h) {
        myc[h] = mL[h] = mR[h] = -1;
        if (b.n_chain == 0) continue;
        cm.m[h] = b.lane_mask[h];
        const int i = h * 64 + lane;
        myc[h] = b.lane_ch[i];
        if (myc[h] < 0) continue;
        const int j = b.lane_ch[64 * W + i], K = b.ch_off[myc[h] + 1] - b.ch_off[myc[h]] + 1;
        if (j == l + 1) mL[h] = b.ch_L[myc[h]];
        if (K - j == l + 1) mR[h] = b.ch_R[myc[h]];
    }
    int cnt = 0;  // wave-uniform
    for (int base = gwave() * 64; base < b.n; base += nwaves() * 64) {
        const int w0 = base + lane;
        const Msk<W> N0 = w0 < b.n ? mld<W>(nx, w0) : mzero<W>();
        const uint64_t bal = __ballot(many<W>(N0));
        if (!bal) continue;
        const int k0 = cnt;
        // CSR ranges of the chunk's entries, loaded lane-parallel up front
        int r0 = 0, r1 = 0, e0 = -1;
        if (many<W>(N0)) {
            const int slot = k0 + __popcll(bal & (me - 1));
            s_w[wv][slot] = w0;
            s_m[wv][slot] = N0;
            r0 = b.rp[w0];
            r1 = b.rp[w0 + 1];
            if (b.n_chain > 0) e0 = b.ep_idx[w0];
        }
        cnt += __popcll(bal);
        __builtin_amdgcn_wave_barrier();
        uint64_t rest = bal;
        for (int k = k0; k < cnt; ++k) {
            const int src = __ffsll((long long)rest) - 1;
            rest &= rest - 1;
            // (the entry's vertex and mask from the staging slots, so the
            // chunk's masks need no registers through the entry loop)
            const int w = __builtin_amdgcn_readfirstlane(s_w[wv][k]);
            Msk<W> N;
#pragma unroll
            for (int h = 0; h < W; ++h) N.m[h] = uni64(s_m[wv][k].m[h]);
            const int a0 = bcast(r0, src), a1 = bcast(r1, src);
            const int ew = bcast(e0, src);
            const int pb = mbits<W>(N, lane);
            const int64_t cell = ((int64_t)w * 64 + lane) * W;
            if (!(split && is_hub(b, a1 - a0))) {
                SgW<SM> acc[W];
                double facc[W];
                int sc[W];
#pragma unroll
                for (int h = 0; h < W; ++h) {
                    acc[h] = 0;
                    facc[h] = 0.0;
                    sc[h] = (w == mL[h] || w == mR[h]) ? myc[h] : -1;
                }
                bool ovf = false;
                sigma_sum<SM, W>(b, fr, N, a0, a1, acc, facc, ovf, sc, cm);
#pragma unroll
                for (int h = 0; h < W; ++h) {
                    if (!((pb >> h) & 1)) continue;
                    sg_put<SM>(b, cell + h, F64 ? d2u(facc[h]) : acc[h]);
                    if (!F64 && (ovf || acc[h] > sg_max<SM>() || (b.limit && acc[h] >= b.limit))) b.ctr[1] = 1;
                }
            }
#pragma unroll
            for (int h = 0; h < W; ++h) {
                if (!((pb >> h) & 1)) continue;
                if (b.lvl) b.lvl[cell + h] = l + 1;
                if (ew >= 0) b.ep_lvl[((int64_t)ew * 64 + lane) * W + h] = l + 1;
            }
            if (lane == 0)
#pragma unroll
                for (int h = 0; h < W; ++h) b.vis[(int64_t)w * W + h] |= N.m[h];
        }
        if (cnt >= STAGE_FLUSH) stage_flush<W>(b, s_w[wv], s_m[wv], cnt, lane);
    }
    stage_flush<W>(b, s_w[wv], s_m[wv], cnt, lane);
    // hub slices: the level mask of every hub is recorded for the backward;
    // partial sigma sums are added atomically (exact integers)
    for (int s = gwave(); s < b.n_slice; s += nwaves()) {
        const int hb = b.sl_hub[s];
        const int v = b.hub_v[hb];
        const Msk<W> N = mld<W>(nx, v);
        if (b.sl_j0[s] == b.rp[v] && lane == 0 && l + 1 < b.hub_stride)
            mst<W>(b.hub_mask, (int64_t)hb * b.hub_stride + l + 1, N);
        if (!split || !many<W>(N)) continue;
        constexpr int SI = F64 ? SIG_U64 : SM;  // (not reached in FP64 mode)
        SgW<SI> acc[W];
        double facc[W];
        int sc[W];
#pragma unroll
        for (int h = 0; h < W; ++h) {
            acc[h] = 0;
            facc[h] = 0.0;
            sc[h] = (v == mL[h] || v == mR[h]) ? myc[h] : -1;
        }
        bool ovf = false;
        sigma_sum<SI, W>(b, fr, N, b.sl_j0[s], b.sl_j1[s], acc, facc, ovf, sc, cm);
        const int pb = mbits<W>(N, lane);
#pragma unroll
        for (int h = 0; h < W; ++h) {
            if (!((pb >> h) & 1) || !acc[h]) continue;
            const int64_t k = ((int64_t)v * 64 + lane) * W + h;
            uint64_t old, nw;
            bool o2 = ovf;
            if (SI == SIG_U32) {
                old = atomicAdd((unsigned int *)b.sig + k, (unsigned int)acc[h]);
                nw = old + acc[h];
            } else {
                old = atomicAdd((unsigned long long *)b.sig + k, (unsigned long long)acc[h]);
                nw = old + acc[h];
                o2 |= nw < old;
            }
            if (o2 || acc[h] > sg_max<SI>() || nw > sg_max<SI>() || (b.limit && nw >= b.limit)) b.ctr[1] = 1;
        }
    }
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

// ------------------------------------------------- contracted chains ---
// A chain segment runs from A (level a, path count sA) over K arcs to B
// (level b, count sB); its interiors y_1 .. y_{K-1} are numbered from A, and
// PA(i) = sum of r(y)+1 over y_1 .. y_i.  From the two ends alone: y_i is
// reached from A at level a+i and from B at level b+K-i; with S = b-a+K, the
// interiors up to mA = ceil(S/2)-1 take A's side (count sA), those past it
// B's, and for even S the middle one y_{S/2} takes both (count sA + sB, a
// tie).  Its successors along the chain give dependencies that are prefix
// sums of r(y)+1 (every path count is constant along a side).
//
// Prefix sums of one chain read from a given end: P(x) = sgn * (pre(anchor +
// sgn x) - pre(anchor)), pre(i) = sum over interiors x_1 .. x_i from L.
struct ChainPfx {
    const double *pre;  // the chain's slice of ch_pre (pre(i) at pre[i-1])
    int anchor, sgn;
    __device__ __forceinline__ double at(int i) const { return i > 0 ? pre[i - 1] : 0.0; }
    __device__ __forceinline__ double operator()(int x) const { return sgn * (at(anchor + sgn * x) - at(anchor)); }
};

// The term the segment adds to rho_A: sum over A's successor y_1 and beyond,
// PA(mA) / sA, plus the tie interior's share (r+1) / (sA + sB).
__device__ __forceinline__ double seg_term(int a, int bl, double sA, double sB, int K, const ChainPfx &P) {
    const int S = bl - a + K;
    if (S <= 0) return 0.0;
    int mA;
    bool tie = false;
    int t = 0;
    if ((S & 1) == 0) {
        t = S >> 1;
        mA = t - 1;
        tie = t >= 1 && t <= K - 1;
    } else {
        mA = (S - 1) >> 1;
    }
    mA = min(mA, K - 1);
    double term = P(mA) / sA;
    if (tie) term += (P(t) - P(t - 1)) / (sA + sB);
    return term;
}

// rho_v = sum over successors w (sources in succ[w] & M) of t_w.
template <typename TT, int W>
__device__ __forceinline__ void rho_sum(const BatchDev &b, const uint64_t *succ, const Msk<W> &M, int a0, int a1,
                                        double (&out)[W]) {
    using Acc = typename std::conditional<sizeof(TT) == 4, KahanF, SumD>::type;
    constexpr int UW = W == 4 ? UNR / 2 : UNR;
    const int lane = threadIdx.x & 63;
    Acc acc[W];
    for (int c = a0; c < a1; c += 64) {
        const int j = c + lane;
        int w = 0;
        Msk<W> x = mzero<W>();
        if (j < a1) {
            w = b.col[j];
            if (w >= 0) x = mand<W>(mld<W>(succ, w), M);
        }
        uint64_t bal = __ballot(many<W>(x));
        while (bal) {
            TT t[UW][W];
            int use[UW];
#pragma unroll
            for (int u = 0; u < UW; ++u) {
                use[u] = 0;
#pragma unroll
                for (int h = 0; h < W; ++h) t[u][h] = 0;
                if (bal) {
                    const int src = __ffsll((long long)bal) - 1;
                    bal &= bal - 1;
                    const int ww = bcast(w, src);
                    use[u] = mbits<W>(mshfl<W>(x, src), lane);
                    if (use[u]) tt_row<TT, W>(b, (int64_t)ww * 64 + lane, t[u]);
                }
            }
#pragma unroll
            for (int u = 0; u < UW; ++u)
#pragma unroll
                for (int h = 0; h < W; ++h)
                    if ((use[u] >> h) & 1) acc[h].add(t[u][h]);
        }
    }
#pragma unroll
    for (int h = 0; h < W; ++h) out[h] = acc[h].value();
}

// rho complete for (v, sources M): delta, BC contribution and the term t_v.
template <int SM, typename TT, int W>
__device__ __forceinline__ void finish_vertex(const BatchDev &b, int v, int l, const Msk<W> &M, double (&acc)[W]) {
    const int lane = threadIdx.x & 63;
    const int e = b.n_chain > 0 ? b.ep_idx[v] : -1;
    const int pb = mbits<W>(M, lane);
    double contrib = 0.0;
#pragma unroll
    for (int h = 0; h < W; ++h) {
        if (!((pb >> h) & 1)) continue;
        const int64_t cell = ((int64_t)v * 64 + lane) * W + h;
        const double sv = sg_val<SM>(sg<SM>(b, cell));
        double a = acc[h];
        if (e >= 0) a += b.ep_rho[((int64_t)e * 64 + lane) * W + h];
        const double delta = sv * a;
        contrib += b.wsl[h * 64 + lane] * delta;
        ((TT *)b.tt)[cell] = (TT)(((double)b.reach[v] + 1.0) / sv + a);
        if (b.dbg_delta) b.dbg_delta[cell] = delta;
    }
    contrib = wave_sum(contrib);
    // b.bc is the slot's own accumulator and v is finished once per level
    // kernel, so a plain read-modify-write is race free
    if (lane == 0 && contrib != 0.0) b.bc[v] += contrib;
}

// Level l <= depth: entries Q_l; Q_{l+2} is cleared from mb[(l+2)%3] and Q_l
// scattered into mb[l%3] for level l-1.
template <int SM, typename TT, int W>
__global__ __launch_bounds__(BB) void k_b_back(BatchDev b, int l) {
    constexpr bool F64 = SM == SIG_F64;
    __shared__ int t_w[NW][WIN];
    __shared__ Msk<W> t_x[NW][WIN];
    __shared__ int t_k[NW][WIN];
    // entries per wave block: FEB / W, so a block holds FEB * 64 (entry, source) pairs
    constexpr int EB = FEB / W;
    constexpr int UW = W == 4 ? UNR / 2 : UNR;
    __shared__ double r_rho[NW][EB][64 * W];  // [entry][h * 64 + lane]
    // the host launches the backward speculatively behind the forward: a
    // level beyond the depth, an unfinished forward (depth -1), a sigma
    // overflow (the tile is redone in FP64) or full lists all mean no work
    const int D = ldi(&b.ctr[3]);
    if (l > D || (!F64 && b.ctr[1]) || b.ctr[2]) return;
    const int q0 = b.qoff[l], q1 = b.qoff[l + 1];
    const int c0 = l + 2 <= D ? b.qoff[l + 2] : 0, c1 = l + 2 <= D ? b.qoff[l + 3] : 0;
    uint64_t *own = mbuf(b, l);
    const uint64_t *succ = mbuf(b, l + 1);
    uint64_t *clr = mbuf(b, l + 2);
    for (int i = c0 + blockIdx.x * BB + threadIdx.x; i < c1; i += gridDim.x * BB) mst<W>(clr, b.Q[i], mzero<W>());
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint64_t me = 1ull << lane;
    using Acc = typename std::conditional<sizeof(TT) == 4, KahanF, SumD>::type;
    // Blocks of EB list entries per wave.  The arcs of the block are walked
    // as one flat range (lane per arc) in windows of WIN: successor masks
    // are gathered for 64 arcs at once and the (entry, successor, sources)
    // triples that matter are compacted in LDS; the t rows of those triples
    // are then gathered UW at a time, accumulated per entry in arc order.
    for (int base = q0 + gwave() * EB; base < q1; base += nwaves() * EB) {
        const int e = base + lane;
        int v0 = 0, r0 = 0, r1 = 0, ep0 = -1;
        Msk<W> M0 = mzero<W>();
        if (lane < EB && e < q1) {
            v0 = b.Q[e];
            M0 = mld<W>(b.Qm, e);
            mst<W>(own, v0, M0);
            r0 = b.rp[v0];
            r1 = b.rp[v0 + 1];
            if (b.n_chain > 0) ep0 = b.ep_idx[v0];
        }
        const int cnt = min(EB, q1 - base);
        const bool mine = lane < cnt && !(is_hub(b, r1 - r0) && l < b.hub_stride);  // hubs: slices
        const int d = mine ? r1 - r0 : 0;
        int inc = d;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(inc, o, 64);
            if (lane >= o) inc += y;
        }
        const int pre = inc - d;
        const int T = bcast(inc, 63);
        int P[EB], A[EB], V[EB], EP[EB];
        Msk<W> MM[EB];
        bool OWN[EB];
#pragma unroll
        for (int j = 0; j < EB; ++j) {
            P[j] = j < cnt ? bcast(pre, j) : INT32_MAX;
            A[j] = bcast(r0, j);
            V[j] = bcast(v0, j);
            EP[j] = bcast(ep0, j);
            MM[j] = mshfl<W>(M0, j);
            OWN[j] = bcast((int)mine, j) != 0;
#pragma unroll
            for (int h = 0; h < W; ++h) r_rho[wv][j][h * 64 + lane] = 0.0;
        }
        Acc acc[W];
        int kcur = -1;
        for (int win = 0; win < T; win += WIN) {
            const int wend = min(T, win + WIN);
            int nrel = 0;
            for (int c = win; c < wend; c += 64) {
                const int g = c + lane;
                int k = 0;
#pragma unroll
                for (int j = 1; j < EB; ++j) k += (P[j] <= g);
                int a0k = A[0], pk = P[0];
                Msk<W> mk = MM[0];
#pragma unroll
                for (int j = 1; j < EB; ++j)
                    if (k == j) {
                        a0k = A[j];
                        pk = P[j];
                        mk = MM[j];
                    }
                int w = 0;
                Msk<W> x = mzero<W>();
                if (g < wend) {
                    w = b.col[a0k + g - pk];
                    if (w >= 0) x = mand<W>(mld<W>(succ, w), mk);  // (w < 0: a contracted chain)
                }
                const uint64_t bal = __ballot(many<W>(x));
                if (many<W>(x)) {
                    const int pos = nrel + __popcll(bal & (me - 1));
                    t_w[wv][pos] = w;
                    t_x[wv][pos] = x;
                    t_k[wv][pos] = k;
                }
                nrel += __popcll(bal);
            }
            __builtin_amdgcn_wave_barrier();
            for (int r = 0; r < nrel; r += UW) {
                TT t[UW][W];
                int kk[UW], use[UW];
#pragma unroll
                for (int u = 0; u < UW; ++u) {
                    const int rr = r + u;
                    use[u] = 0;
                    kk[u] = -1;
#pragma unroll
                    for (int h = 0; h < W; ++h) t[u][h] = 0;
                    if (rr < nrel) {
                        kk[u] = __builtin_amdgcn_readfirstlane(t_k[wv][rr]);  // (uniform: a scalar register)
                        use[u] = mbits<W>(t_x[wv][rr], lane);
                        if (use[u]) tt_row<TT, W>(b, (int64_t)t_w[wv][rr] * 64 + lane, t[u]);
                    }
                }
#pragma unroll
                for (int u = 0; u < UW; ++u) {
                    if (kk[u] < 0) continue;
                    if (kk[u] != kcur) {
                        if (kcur >= 0)
#pragma unroll
                            for (int h = 0; h < W; ++h) r_rho[wv][kcur][h * 64 + lane] = acc[h].value();
#pragma unroll
                        for (int h = 0; h < W; ++h) acc[h] = Acc();
                        kcur = kk[u];
                    }
#pragma unroll
                    for (int h = 0; h < W; ++h)
                        if ((use[u] >> h) & 1) acc[h].add(t[u][h]);
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
        if (kcur >= 0)
#pragma unroll
            for (int h = 0; h < W; ++h) r_rho[wv][kcur][h * 64 + lane] = acc[h].value();
        // chain endpoints: their chains' terms join rho
#pragma unroll
        for (int j = 0; j < EB; ++j) {
            if (!OWN[j] || EP[j] < 0) continue;
            const int pb = mbits<W>(MM[j], lane);
#pragma unroll
            for (int h = 0; h < W; ++h)
                if ((pb >> h) & 1) r_rho[wv][j][h * 64 + lane] += b.ep_rho[((int64_t)EP[j] * 64 + lane) * W + h];
        }
        // finish every entry of the block (hubs excepted): sigma rows first
        SgW<SM> SX[EB][W];
#pragma unroll
        for (int j = 0; j < EB; ++j) {
#pragma unroll
            for (int h = 0; h < W; ++h) SX[j][h] = 0;
            if (OWN[j] && mbits<W>(MM[j], lane)) sg_row<SM, W>(b, (int64_t)V[j] * 64 + lane, SX[j]);
        }
#pragma unroll
        for (int j = 0; j < EB; ++j) {
            if (!OWN[j]) continue;
            const int v = V[j];
            const int pb = mbits<W>(MM[j], lane);
            double contrib = 0.0;
#pragma unroll
            for (int h = 0; h < W; ++h) {
                if (!((pb >> h) & 1)) continue;
                const int64_t cell = ((int64_t)v * 64 + lane) * W + h;
                const double sv = sg_val<SM>(SX[j][h]);
                const double a = r_rho[wv][j][h * 64 + lane];  // rho_v
                const double delta = sv * a;
                contrib += b.wsl[h * 64 + lane] * delta;
                ((TT *)b.tt)[cell] = (TT)(((double)b.reach[v] + 1.0) / sv + a);
                if (b.dbg_delta) b.dbg_delta[cell] = delta;
            }
            contrib = wave_sum(contrib);
            if (lane == 0 && contrib != 0.0) b.bc[v] += contrib;
        }
    }
    // hub slices (levels with recorded hub masks): partial rho into FP64
    // accumulators; the last slice of a hub to finish (counter) completes it
    if (l >= b.hub_stride) return;
    for (int s = gwave(); s < b.n_slice; s += nwaves()) {
        const int hb = b.sl_hub[s];
        const Msk<W> M = mld<W>(b.hub_mask, (int64_t)hb * b.hub_stride + l);
        if (!many<W>(M)) continue;
        const int v = b.hub_v[hb];
        double part[W];
        rho_sum<TT, W>(b, succ, M, b.sl_j0[s], b.sl_j1[s], part);
        const int pb = mbits<W>(M, lane);
#pragma unroll
        for (int h = 0; h < W; ++h)
            if (((pb >> h) & 1) && part[h] != 0.0)
                unsafeAtomicAdd(&b.hub_acc[((int64_t)hb * 64 + lane) * W + h], part[h]);
        __threadfence();
        int done = 0;
        if (lane == 0) done = atomicAdd(&b.hub_cnt[hb], 1);
        done = bcast(done, 0);
        if (done != b.hub_nsl[hb] - 1) continue;
        __threadfence();
        double rho[W];
#pragma unroll
        for (int h = 0; h < W; ++h)
            rho[h] = __hip_atomic_load(&b.hub_acc[((int64_t)hb * 64 + lane) * W + h], __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
        finish_vertex<SM, TT, W>(b, v, l, M, rho);
        if (lane == 0) b.hub_cnt[hb] = 0;
    }
}

// Interior y_q (1 <= q <= K-1) of a segment A .. B (see seg_term): its
// level, path-count word and dependency.  tB is B's term t_B = (r(B)+1 +
// delta_B) / sigma_B, used when B itself is reached through the segment
// (b == a + K: only a source's own chain, whose halves end at L and R).
template <int SM>
__device__ __forceinline__ void seg_interior(int a, int bl, double sA, double sB, uint64_t wA, uint64_t wB, int K,
                                             int q, const ChainPfx &P, double tB, int &lev, uint64_t &sw,
                                             double &dl) {
    const int S = bl - a + K;
    int t = -1, mA, b0;
    bool tie = false;
    if ((S & 1) == 0) {
        t = S >> 1;
        mA = t - 1;
        tie = t >= 1 && t <= K - 1;
        b0 = t + 1;
    } else {
        mA = (S - 1) >> 1;
        b0 = mA + 1;
    }
    mA = min(mA, K - 1);  // (B beyond the segment's reach: every interior on A's side)
    const double tw = tie ? P(t) - P(t - 1) : 0.0;
    if (q <= mA) {
        lev = a + q;
        sw = wA;
        dl = P(mA) - P(q) + (tie ? tw * sA / (sA + sB) : 0.0) + (bl == a + K ? sA * tB : 0.0);
    } else if (tie && q == t) {
        lev = a + t;
        sw = SM == SIG_F64 ? d2u(sA + sB) : wA + wB;
        dl = 0.0;
    } else {
        lev = bl + K - q;
        sw = wB;
        dl = P(q - 1) - P(b0 - 1) + (tie ? tw * sB / (sA + sB) : 0.0);
    }
}

// Between the forward and the backward: a tie interior's count sA + sB must
// fit the tile's word (and stay below the test threshold), like every count
// the pull makes; otherwise the tile is redone with the next word.
template <int SM, int W>
__global__ __launch_bounds__(BB) void k_b_chain_check(BatchDev b) {
    if (ldi(&b.ctr[3]) < 0 || b.ctr[1] || b.ctr[2]) return;
    const int lane = threadIdx.x & 63;
    const int ns = b.ctr[5];
    for (int c = gwave(); c < b.n_chain; c += nwaves()) {
        const int L = b.ch_L[c], R = b.ch_R[c], K = b.ch_off[c + 1] - b.ch_off[c] + 1;
        const int64_t eL = (int64_t)b.ep_idx[L] * 64, eR = (int64_t)b.ep_idx[R] * 64;
#pragma unroll
        for (int h = 0; h < W; ++h) {
            const int i = h * 64 + lane;
            if (i >= ns) continue;
            const int myc = b.lane_ch[i], myj = b.lane_ch[64 * W + i];
            const int aL = b.ep_lvl[(eL + lane) * W + h], aR = b.ep_lvl[(eR + lane) * W + h];
            const uint64_t wL = sg<SM>(b, ((int64_t)L * 64 + lane) * W + h);
            const uint64_t wR = sg<SM>(b, ((int64_t)R * 64 + lane) * W + h);
            // the segments of this source: L .. R, or source .. L and source ..
            // R; a tie needs an even S = b - a + K with 2 <= S <= 2K - 2
            auto tie = [](int S, int k) { return (S & 1) == 0 && S >= 2 && S <= 2 * k - 2; };
            auto check = [&](uint64_t x, uint64_t y) {
                const uint64_t z = x + y;
                if (z < x || z > sg_max<SM>() || (b.limit && z >= b.limit)) b.ctr[1] = 1;
            };
            if (c != myc) {
                if (tie(aR - aL + K, K)) check(wL, wR);
            } else {
                if (tie(aL + myj, myj)) check(wL, 1);
                if (tie(aR + K - myj, K - myj)) check(wR, 1);
            }
        }
    }
}

// After the forward and the count check: the chain terms of every endpoint's
// rho for every source, which depend on the endpoints' levels and counts
// alone, into ep_rho; the backward adds them when it finishes an endpoint
// (inlined there, the closed forms raised the backward's register peak by a
// wave per SIMD).  One wave per endpoint.
template <int SM, int W>
__global__ __launch_bounds__(BB) void k_b_chain_terms(BatchDev b) {
    constexpr bool F64 = SM == SIG_F64;
    if (ldi(&b.ctr[3]) < 0 || (!F64 && b.ctr[1]) || b.ctr[2]) return;
    const int lane = threadIdx.x & 63;
    const int ns = b.ctr[5];
    for (int e = gwave(); e < b.n_ep; e += nwaves()) {
        const int q0 = b.ep_off[e], q1 = b.ep_off[e + 1];
        const int code0 = b.ep_list[q0];  // (an endpoint has a chain: its vertex is that chain's end)
        const int v = (code0 & 1) ? b.ch_R[code0 >> 1] : b.ch_L[code0 >> 1];
        // per source slot k + 64 h: the endpoint's level and count, the
        // source's own chain and position (-1: a tile vertex)
        int lv[W], myc[W], myj[W];
        double sv[W], rho[W];
#pragma unroll
        for (int h = 0; h < W; ++h) {
            const int i = h * 64 + lane;
            const bool act = i < ns;
            lv[h] = act ? b.ep_lvl[((int64_t)e * 64 + lane) * W + h] : 0;
            sv[h] = act ? sg_val<SM>(sg<SM>(b, ((int64_t)v * 64 + lane) * W + h)) : 1.0;
            myc[h] = act ? b.lane_ch[i] : -1;
            myj[h] = act ? b.lane_ch[64 * W + i] : 0;
            rho[h] = 0.0;
        }
        // every attached chain is a segment from the endpoint to its other
        // end, except the chain holding the source x_j itself, which is the
        // segment from the endpoint to the source (level 0, count 1, j arcs
        // from L or K - j from R); the other end's level and count are loaded
        // for all W sources of a lane at once
        for (int q = q0; q < q1; ++q) {
            const int code = b.ep_list[q];
            const int c = code >> 1, side = code & 1;
            const int off = b.ch_off[c], K = b.ch_off[c + 1] - off + 1;
            const ChainPfx P{b.ch_pre + off, side ? K - 1 : 0, side ? -1 : 1};
            const int o = side ? b.ch_L[c] : b.ch_R[c];
            const int64_t eo = (int64_t)b.ep_idx[o] * 64 + lane;
            int lo[W];
            double so[W];
#pragma unroll
            for (int h = 0; h < W; ++h) {
                lo[h] = b.ep_lvl[eo * W + h];
                so[h] = sg_val<SM>(sg<SM>(b, ((int64_t)o * 64 + lane) * W + h));
            }
#pragma unroll
            for (int h = 0; h < W; ++h) {
                if (h * 64 + lane >= ns) continue;
                rho[h] += c == myc[h] ? seg_term(lv[h], 0, sv[h], 1.0, side ? K - myj[h] : myj[h], P)
                                      : seg_term(lv[h], lo[h], sv[h], so[h], K, P);
            }
        }
#pragma unroll
        for (int h = 0; h < W; ++h)
            if (h * 64 + lane < ns) b.ep_rho[((int64_t)e * 64 + lane) * W + h] = rho[h];
    }
}

// After the backward: the BC of every contracted interior from the sources
// off its chain, one wave per chain (the launch has a wave for every chain,
// so the chains' dependent loads overlap) and one lane per source slot; BC
// into the slot's accumulator (each interior has one writer).  For such a
// source the segment L .. R has a split mA and possibly a tie interior t, and
// an interior's dependency is cA - P(i) on L's side, 0 at the tie and
// P(i-1) - cB on R's side, with cA, cB set once per chain; an interior then
// costs the chain's running prefix sum (the same word for every lane, loaded
// coalesced once per 64 interiors), a few selects and the wave sum.  The
// sources that lie on a chain (k_b_chains_own) and the probe's rows
// (k_b_chain_rows) have kernels of their own, so this one stays small in
// registers and shares the compute units with the other tiles' level kernels
// (with them inlined it took every register of its SIMD).
template <int SM, int W>
__global__ __launch_bounds__(BB) void k_b_chains(BatchDev b) {
    constexpr bool F64 = SM == SIG_F64;
    if (ldi(&b.ctr[3]) < 0 || (!F64 && b.ctr[1]) || b.ctr[2]) return;
    const int lane = threadIdx.x & 63;
    const int ns = b.ctr[5];
    for (int c = gwave(); c < b.n_chain; c += nwaves()) {
        const int L = b.ch_L[c], R = b.ch_R[c], off = b.ch_off[c], K = b.ch_off[c + 1] - off + 1;
        const double *pre = b.ch_pre + off;
        const ChainPfx P{pre, 0, 1};
        const int64_t eL = (int64_t)b.ep_idx[L] * 64, eR = (int64_t)b.ep_idx[R] * 64;
        int mA[W], tq[W];
        double wgt[W], cA[W], cB[W];
#pragma unroll
        for (int h = 0; h < W; ++h) {
            const int i = h * 64 + lane;
            mA[h] = 0;
            tq[h] = -1;
            wgt[h] = cA[h] = cB[h] = 0.0;  // (weight 0: the slot adds nothing)
            if (i >= ns || b.lane_ch[i] == c) continue;
            const int aL = b.ep_lvl[(eL + lane) * W + h], aR = b.ep_lvl[(eR + lane) * W + h];
            const double sL = sg_val<SM>(sg<SM>(b, ((int64_t)L * 64 + lane) * W + h));
            const double sR = sg_val<SM>(sg<SM>(b, ((int64_t)R * 64 + lane) * W + h));
            // segment L .. R seen from L: split, tie and the prefix sums there
            const int S = aR - aL + K;
            int t = -1, m, b0;
            bool tie = false;
            if ((S & 1) == 0) {
                t = S >> 1;
                m = t - 1;
                tie = t >= 1 && t <= K - 1;
                b0 = t + 1;
            } else {
                m = (S - 1) >> 1;
                b0 = m + 1;
            }
            m = min(m, K - 1);
            const double tw = tie ? P(t) - P(t - 1) : 0.0;
            mA[h] = m;
            tq[h] = tie ? t : -1;
            wgt[h] = b.wsl[i];
            cA[h] = P(m) + (tie ? tw * sL / (sL + sR) : 0.0);
            cB[h] = P(min(b0, K) - 1) - (tie ? tw * sR / (sL + sR) : 0.0);
        }
        double p0 = 0.0;  // P(i - 1)
        // the interiors' BC words and prefix sums, 64 at a time, one per lane
        // (coalesced loads; the loop then reads P(i) from lane i - 1)
        double accbc = 0.0, pv = 0.0;
        for (int i = 1; i < K; ++i) {
            const int64_t x = (int64_t)b.n + off + i - 1;
            if (((i - 1) & 63) == 0) {
                accbc = i - 1 + lane < K - 1 ? b.bc[x + lane] : 0.0;
                pv = i - 1 + lane < K - 1 ? pre[i - 1 + lane] : 0.0;
            }
            const double p1 = bcastd(pv, (i - 1) & 63);  // P(i)
            double contrib = 0.0;
#pragma unroll
            for (int h = 0; h < W; ++h)
                contrib += wgt[h] * (i <= mA[h] ? cA[h] - p1 : (i == tq[h] ? 0.0 : p0 - cB[h]));
            p0 = p1;
            contrib = wave_sum(contrib);
            if (lane == ((i - 1) & 63)) accbc += contrib;
            if ((((i - 1) & 63) == 63 || i == K - 1) && lane <= ((i - 1) & 63))
                b.bc[(int64_t)b.n + off + ((i - 1) & ~63) + lane] = accbc;
        }
    }
}

// The end of a source's own chain segment: its level, count, dependency (and
// the probe's row) at interior x of the chain c holding source x_j: the half
// from the source to L (x < j, numbered from the source) or to R (x > j), each
// a segment from the source (level 0, count 1) to the endpoint.
template <int SM>
__device__ __forceinline__ void own_interior(int x, int j, int K, const double *pre, int aL, int aR, double sL,
                                             double sR, uint64_t wL, uint64_t wR, double tL, double tR, int &lev,
                                             uint64_t &sw, double &dl) {
    const uint64_t one = SM == SIG_F64 ? d2u(1.0) : 1ull;
    lev = 0;
    sw = one;
    dl = 0.0;
    if (x < j)
        seg_interior<SM>(0, aL, 1.0, sL, one, wL, j, j - x, ChainPfx{pre, j - 1, -1}, tL, lev, sw, dl);
    else if (x > j)
        seg_interior<SM>(0, aR, 1.0, sR, one, wR, K - j, x - j, ChainPfx{pre, j, 1}, tR, lev, sw, dl);
}

// After k_b_chains: the interiors' BC from the sources that lie on a chain,
// one wave per source slot (the launch has a wave for each; slots of tile
// vertices return at once) and one lane per interior of the source's chain,
// added with atomics (several sources can lie on one chain).
template <int SM, typename TT, int W>
__global__ __launch_bounds__(BB) void k_b_chains_own(BatchDev b) {
    constexpr bool F64 = SM == SIG_F64;
    if (ldi(&b.ctr[3]) < 0 || (!F64 && b.ctr[1]) || b.ctr[2]) return;
    const int i = gwave();  // source slot k + 64 h
    if (i >= min(b.ctr[5], 64 * W)) return;
    const int c = b.lane_ch[i];
    if (c < 0) return;
    const int lane = threadIdx.x & 63;
    const int j = b.lane_ch[64 * W + i], k = i & 63, h = i >> 6;
    const int L = b.ch_L[c], R = b.ch_R[c], off = b.ch_off[c], K = b.ch_off[c + 1] - off + 1;
    const int64_t cL = ((int64_t)L * 64 + k) * W + h, cR = ((int64_t)R * 64 + k) * W + h;
    const int aL = b.ep_lvl[((int64_t)b.ep_idx[L] * 64 + k) * W + h];
    const int aR = b.ep_lvl[((int64_t)b.ep_idx[R] * 64 + k) * W + h];
    const uint64_t wL = sg<SM>(b, cL), wR = sg<SM>(b, cR);
    const double tL = (double)((const TT *)b.tt)[cL], tR = (double)((const TT *)b.tt)[cR];
    const double wgt = b.wsl[i];
    for (int x = 1 + lane; x < K; x += 64) {
        int lev;
        uint64_t sw;
        double dl;
        own_interior<SM>(x, j, K, b.ch_pre + off, aL, aR, sg_val<SM>(wL), sg_val<SM>(wR), wL, wR, tL, tR, lev, sw,
                         dl);
        if (dl != 0.0) unsafeAtomicAdd(&b.bc[(int64_t)b.n + off + x - 1], wgt * dl);
    }
}

// Probe only: the level, count word and dependency of every (interior,
// source) pair, one wave per chain and one lane per source slot.
template <int SM, typename TT, int W>
__global__ __launch_bounds__(BB) void k_b_chain_rows(BatchDev b) {
    constexpr bool F64 = SM == SIG_F64;
    if (ldi(&b.ctr[3]) < 0 || (!F64 && b.ctr[1]) || b.ctr[2]) return;
    const int lane = threadIdx.x & 63;
    const int ns = b.ctr[5];
    for (int c = gwave(); c < b.n_chain; c += nwaves()) {
        const int L = b.ch_L[c], R = b.ch_R[c], off = b.ch_off[c], K = b.ch_off[c + 1] - off + 1;
        const double *pre = b.ch_pre + off;
        const int64_t eL = (int64_t)b.ep_idx[L] * 64, eR = (int64_t)b.ep_idx[R] * 64;
        for (int h = 0; h < W; ++h) {
            const int i = h * 64 + lane;
            if (i >= ns) continue;
            const int64_t cL = ((int64_t)L * 64 + lane) * W + h, cR = ((int64_t)R * 64 + lane) * W + h;
            const int aL = b.ep_lvl[(eL + lane) * W + h], aR = b.ep_lvl[(eR + lane) * W + h];
            const uint64_t wL = sg<SM>(b, cL), wR = sg<SM>(b, cR);
            const double sL = sg_val<SM>(wL), sR = sg_val<SM>(wR);
            const bool own = b.lane_ch[i] == c;
            const int j = b.lane_ch[64 * W + i];
            const double tL = own ? (double)((const TT *)b.tt)[cL] : 0.0;
            const double tR = own ? (double)((const TT *)b.tt)[cR] : 0.0;
            for (int x = 1; x < K; ++x) {
                int lev;
                uint64_t sw;
                double dl;
                if (own)
                    own_interior<SM>(x, j, K, pre, aL, aR, sL, sR, wL, wR, tL, tR, lev, sw, dl);
                else
                    seg_interior<SM>(aL, aR, sL, sR, wL, wR, K, x, ChainPfx{pre, 0, 1}, 0.0, lev, sw, dl);
                const int64_t cell = (((int64_t)b.n + off + x - 1) * 64 + lane) * W + h;
                b.lvl[cell] = lev;
                sg_put<SM>(b, cell, sw);
                b.dbg_delta[cell] = dl;
            }
        }
    }
}

// After the backward only Q_1 (mask array 1) and Q_2 (array 2) are set.
template <int W>
__global__ __launch_bounds__(BB) void k_b_finish(BatchDev b) {
    const int D = b.ctr[3];
    if (D < 0 || b.ctr[2]) return;
    const int a0 = b.qoff[1], a1 = D >= 1 ? b.qoff[2] : a0;
    const int c0 = D >= 2 ? b.qoff[2] : 0, c1 = D >= 2 ? b.qoff[3] : 0;
    for (int i = a0 + blockIdx.x * BB + threadIdx.x; i < a1; i += gridDim.x * BB) mst<W>(b.mb1, b.Q[i], mzero<W>());
    for (int i = c0 + blockIdx.x * BB + threadIdx.x; i < c1; i += gridDim.x * BB) mst<W>(b.mb2, b.Q[i], mzero<W>());
}

// Clears every mask entry of Q[0, end) (after an abandoned forward).
template <int W>
__global__ __launch_bounds__(BB) void k_b_clear_all(BatchDev b, int end) {
    for (int i = blockIdx.x * BB + threadIdx.x; i < end; i += gridDim.x * BB) {
        const int v = b.Q[i];
        mst<W>(b.mb0, v, mzero<W>());
        mst<W>(b.mb1, v, mzero<W>());
        mst<W>(b.mb2, v, mzero<W>());
    }
}

// dst[v] += src[v]: folds a slot's BC accumulator into the component slice
// and clears it for the next component.
// perm (optional) maps the tile's vertex ids to the component's.
__global__ __launch_bounds__(BB) void k_b_fold(double *dst, double *src, const int32_t *perm, int n) {
    for (int v = blockIdx.x * BB + threadIdx.x; v < n; v += gridDim.x * BB) {
        const double x = src[v];
        if (x != 0.0) dst[perm ? perm[v] : v] += x;
        src[v] = 0.0;
    }
}

// ------------------------------------------------------------ launchers ---
template <int W>
static void init_w(const BatchDev &b, int ns, int sm, int g, hipStream_t st) {
    hipLaunchKernelGGL(k_b_clear_vis<W>, dim3(std::max(g, 1)), dim3(BB), 0, st, b, sm);
    if (sm == SIG_U32) hipLaunchKernelGGL((k_b_seed<SIG_U32, W>), dim3(1), dim3(64), 0, st, b, ns);
    else if (sm == SIG_U64) hipLaunchKernelGGL((k_b_seed<SIG_U64, W>), dim3(1), dim3(64), 0, st, b, ns);
    else hipLaunchKernelGGL((k_b_seed<SIG_F64, W>), dim3(1), dim3(64), 0, st, b, ns);
}

hipError_t launch_b_init(const BatchDev &b, int ns, int sm, int cap, hipStream_t st) {
    const int g = (int)std::min<int64_t>(((int64_t)b.n + BB - 1) / BB, cap);
    if (b.spl == 4) init_w<4>(b, ns, sm, g, st);
    else if (b.spl == 2) init_w<2>(b, ns, sm, g, st);
    else init_w<1>(b, ns, sm, g, st);
    return hipGetLastError();
}

template <int W>
static void forward_w(const BatchDev &b, int sm, int l, int grid, hipStream_t st) {
    hipLaunchKernelGGL(k_b_expand<W>, dim3(grid), dim3(BB), 0, st, b, l);
    if (sm == SIG_U32) hipLaunchKernelGGL((k_b_pull<SIG_U32, W>), dim3(grid), dim3(BB), 0, st, b, l);
    else if (sm == SIG_U64) hipLaunchKernelGGL((k_b_pull<SIG_U64, W>), dim3(grid), dim3(BB), 0, st, b, l);
    else hipLaunchKernelGGL((k_b_pull<SIG_F64, W>), dim3(grid), dim3(BB), 0, st, b, l);
}

hipError_t launch_b_forward(const BatchDev &b, int sm, int l, int grid, hipStream_t st) {
    if (b.spl == 4) forward_w<4>(b, sm, l, grid, st);
    else if (b.spl == 2) forward_w<2>(b, sm, l, grid, st);
    else forward_w<1>(b, sm, l, grid, st);
    return hipGetLastError();
}

// FP64-sigma tiles keep FP64 terms in either precision: 1/sigma can leave
// the f32 range.
template <int W>
static void back_w(const BatchDev &b, int sm, bool fp32, int l, int grid, hipStream_t st) {
    if (sm == SIG_U32 && fp32) hipLaunchKernelGGL((k_b_back<SIG_U32, float, W>), dim3(grid), dim3(BB), 0, st, b, l);
    else if (sm == SIG_U32) hipLaunchKernelGGL((k_b_back<SIG_U32, double, W>), dim3(grid), dim3(BB), 0, st, b, l);
    else if (sm == SIG_U64 && fp32)
        hipLaunchKernelGGL((k_b_back<SIG_U64, float, W>), dim3(grid), dim3(BB), 0, st, b, l);
    else if (sm == SIG_U64) hipLaunchKernelGGL((k_b_back<SIG_U64, double, W>), dim3(grid), dim3(BB), 0, st, b, l);
    else hipLaunchKernelGGL((k_b_back<SIG_F64, double, W>), dim3(grid), dim3(BB), 0, st, b, l);
}

hipError_t launch_b_back(const BatchDev &b, int sm, bool fp32, int l, int grid, hipStream_t st) {
    if (b.spl == 4) back_w<4>(b, sm, fp32, l, grid, st);
    else if (b.spl == 2) back_w<2>(b, sm, fp32, l, grid, st);
    else back_w<1>(b, sm, fp32, l, grid, st);
    return hipGetLastError();
}

hipError_t launch_b_finish(const BatchDev &b, int grid, hipStream_t st) {
    if (b.spl == 4) hipLaunchKernelGGL(k_b_finish<4>, dim3(grid), dim3(BB), 0, st, b);
    else if (b.spl == 2) hipLaunchKernelGGL(k_b_finish<2>, dim3(grid), dim3(BB), 0, st, b);
    else hipLaunchKernelGGL(k_b_finish<1>, dim3(grid), dim3(BB), 0, st, b);
    return hipGetLastError();
}

// one wave per chain (2 per block)
static int chain_grid(const BatchDev &b) {
    return (int)std::max<int64_t>(1, std::min<int64_t>(((int64_t)b.n_chain + NW - 1) / NW, 1 << 16));
}

template <int W>
static void chain_check_w(const BatchDev &b, int sm, int g, hipStream_t st) {
    if (sm == SIG_U32) hipLaunchKernelGGL((k_b_chain_check<SIG_U32, W>), dim3(g), dim3(BB), 0, st, b);
    else if (sm == SIG_U64) hipLaunchKernelGGL((k_b_chain_check<SIG_U64, W>), dim3(g), dim3(BB), 0, st, b);
}

hipError_t launch_b_chain_check(const BatchDev &b, int sm, int grid, hipStream_t st) {
    (void)grid;
    if (b.spl == 4) chain_check_w<4>(b, sm, chain_grid(b), st);
    else if (b.spl == 2) chain_check_w<2>(b, sm, chain_grid(b), st);
    else chain_check_w<1>(b, sm, chain_grid(b), st);
    return hipGetLastError();
}

template <int W>
static void chain_terms_w(const BatchDev &b, int sm, int g, hipStream_t st) {
    if (sm == SIG_U32) hipLaunchKernelGGL((k_b_chain_terms<SIG_U32, W>), dim3(g), dim3(BB), 0, st, b);
    else if (sm == SIG_U64) hipLaunchKernelGGL((k_b_chain_terms<SIG_U64, W>), dim3(g), dim3(BB), 0, st, b);
    else hipLaunchKernelGGL((k_b_chain_terms<SIG_F64, W>), dim3(g), dim3(BB), 0, st, b);
}

// one wave per endpoint
hipError_t launch_b_chain_terms(const BatchDev &b, int sm, int grid, hipStream_t st) {
    (void)grid;
    const int g = (int)std::max<int64_t>(1, std::min<int64_t>(((int64_t)b.n_ep + NW - 1) / NW, 1 << 16));
    if (b.spl == 4) chain_terms_w<4>(b, sm, g, st);
    else if (b.spl == 2) chain_terms_w<2>(b, sm, g, st);
    else chain_terms_w<1>(b, sm, g, st);
    return hipGetLastError();
}

template <int SM, typename TT, int W>
static void chains_k(const BatchDev &b, int g, hipStream_t st) {
    hipLaunchKernelGGL((k_b_chains<SM, W>), dim3(g), dim3(BB), 0, st, b);
    hipLaunchKernelGGL((k_b_chains_own<SM, TT, W>), dim3(64 * W / NW), dim3(BB), 0, st, b);
    if (b.lvl) hipLaunchKernelGGL((k_b_chain_rows<SM, TT, W>), dim3(g), dim3(BB), 0, st, b);
}

template <int W>
static void chains_w(const BatchDev &b, int sm, bool fp32, int g, hipStream_t st) {
    if (sm == SIG_U32 && fp32) chains_k<SIG_U32, float, W>(b, g, st);
    else if (sm == SIG_U32) chains_k<SIG_U32, double, W>(b, g, st);
    else if (sm == SIG_U64 && fp32) chains_k<SIG_U64, float, W>(b, g, st);
    else if (sm == SIG_U64) chains_k<SIG_U64, double, W>(b, g, st);
    else chains_k<SIG_F64, double, W>(b, g, st);
}

// the interiors' BC (and, probing, their rows): one wave per chain, then one
// per source slot for the sources on chains
hipError_t launch_b_chains(const BatchDev &b, int sm, bool fp32, int grid, hipStream_t st) {
    (void)grid;
    if (b.spl == 4) chains_w<4>(b, sm, fp32, chain_grid(b), st);
    else if (b.spl == 2) chains_w<2>(b, sm, fp32, chain_grid(b), st);
    else chains_w<1>(b, sm, fp32, chain_grid(b), st);
    return hipGetLastError();
}

hipError_t launch_b_fold(double *dst, double *src, const int32_t *perm, int n, int grid, hipStream_t st) {
    hipLaunchKernelGGL(k_b_fold, dim3(grid), dim3(BB), 0, st, dst, src, perm, n);
    return hipGetLastError();
}

hipError_t launch_b_clear_all(const BatchDev &b, int end, int grid, hipStream_t st) {
    if (end <= 0) return hipSuccess;
    if (b.spl == 4) hipLaunchKernelGGL(k_b_clear_all<4>, dim3(grid), dim3(BB), 0, st, b, end);
    else if (b.spl == 2) hipLaunchKernelGGL(k_b_clear_all<2>, dim3(grid), dim3(BB), 0, st, b, end);
    else hipLaunchKernelGGL(k_b_clear_all<1>, dim3(grid), dim3(BB), 0, st, b, end);
    return hipGetLastError();
}

}  // namespace bcx
