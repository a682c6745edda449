// Batched multi-source Brandes for gfx950: one tile = one component and up
// to 64 sources, one source per lane of a wave64.  State is vertex-major so a
// single gather serves every source of the tile:
//   vis[v]        u64  lanes that have reached v
//   mb[l % 3][v]  u64  lanes for which v is at level l (three rotating dense
//                      mask arrays; only entries of the level lists are ever
//                      non-zero, cleared incrementally)
//   sig[v][64]    u32 or u64 (exact) or f64 (overflow mode) path counts per
//                 lane; the narrowest word that holds the tile's counts
//   tt[v][64]     f32 or f64 backward term t_v = (r(v)+1)/sigma_v + rho_v
//   Q / Qm        per-level lists of (vertex, lane mask), levels concatenated,
//                 level l = Q[qoff[l], qoff[l+1])
//
// Forward level l, two launches:
//   expand  pushes lane masks from Q_l to neighbours, F & ~vis[w], with
//           fire-and-forget 64-bit atomic ORs into mb[(l+1)%3] (a lane walks
//           the arcs of a low-degree entry, a wave those of a larger one);
//   pull    scans mb[(l+1)%3] in coalesced 64-vertex chunks; every non-zero
//           entry w joins Q_{l+1} and gets sigma_w = sum of sigma_v over
//           neighbours v in CSR order for the lanes where v is at level l --
//           no sigma atomics, and the FP64 order is the oracle's, so FP64
//           sigma is bit-exact.  (Codes/kernels_free.cu:286-362 computes the
//           same quantities one source per block with atomics.)
// Backward level l: rho_v = sum over successors w of t_w, with
// rho = delta / sigma, so the successor-method term
// sigma_v/sigma_w (r(w)+1 + delta_w) (Codes/kernels_free.cu:378-421) becomes a
// plain sum of one gathered word per (arc, lane); FP32 with Kahan
// compensation (FP64 in FP64 precision and in FP64-sigma tiles).
// BC(v) += sum_lanes (r(s)+1) sigma_v rho_v (Codes/kernels_free.cu:423-426),
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
constexpr int UNR = 8;         // gathers in flight per lane in the sum loops
constexpr int SMALL = 16;      // degree up to which one lane expands an entry
constexpr int STAGE = 256, STAGE_FLUSH = 192;
constexpr int FEB = 8;         // list entries per wave block in the backward
constexpr int WIN = 128;       // arcs per window of a block

__device__ __forceinline__ int gwave() { return (int)((blockIdx.x * (unsigned)blockDim.x + threadIdx.x) >> 6); }
__device__ __forceinline__ int nwaves() { return (int)((gridDim.x * (unsigned)blockDim.x) >> 6); }

__device__ __forceinline__ uint64_t shfl64(uint64_t x, int src) {
    return (uint64_t)__shfl((long long)x, src, 64);
}

__device__ __forceinline__ uint64_t *mbuf(const BatchDev &b, int l) {
    const int r = l % 3;
    return r == 0 ? b.mb0 : (r == 1 ? b.mb1 : b.mb2);
}

__device__ __forceinline__ bool is_hub(const BatchDev &b, int deg) { return b.n_hub > 0 && deg > b.hub_deg; }

__device__ __forceinline__ int ldi(const int32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

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

// ---------------------------------------------------------------- init ---
__global__ __launch_bounds__(BB) void k_b_clear_vis(BatchDev b, int sm) {
    const int64_t stride = (int64_t)gridDim.x * BB;
    const int64_t t0 = blockIdx.x * (int64_t)BB + threadIdx.x;
    for (int64_t v = t0; v < b.n; v += stride) b.vis[v] = 0;
    // hub sigma rows and backward accumulators start at zero every tile: each
    // (hub, lane) is summed into exactly once
    for (int64_t i = t0; i < (int64_t)b.n_hub * 64; i += stride) {
        const int64_t k = (int64_t)b.hub_v[i >> 6] * 64 + (i & 63);
        if (sm == SIG_U32) sg_put<SIG_U32>(b, k, 0);
        else sg_put<SIG_U64>(b, k, 0);
        b.hub_acc[i] = 0.0;
    }
    for (int64_t i = t0; i < b.n_hub; i += stride) b.hub_cnt[i] = 0;
}

// Lanes of a tile vertex start at level 0.  A lane whose source is the
// interior x_j of a contracted chain (ids >= n) has no vertex at level 0: the
// expand kernel injects it at L and at R, each with a path count of 1 (its
// own path along the chain) on levels j and K - j, and the forward runs on
// through empty levels while an injection into a vertex the lane has not yet
// reached is pending.  (Shifting such a lane's levels down so that its nearer
// endpoint enters at level 1 cuts a tile's depth but misaligns it with the
// tile's other lanes, which the tile order groups by their true distances:
// 115k against 123k MTEPS on the 1M-vertex mixed graph.)
template <int SM>
__global__ void k_b_seed(BatchDev b, int ns) {
    const int k = threadIdx.x;  // (launched with 64 threads)
    b.wsl[k] = k < ns ? b.ws[k] : 0.0;
    int s = k < ns ? b.srcs[k] : -1;
    int c = -1, j = 0, top = 0;
    if (s >= b.n) {
        const int x = s - b.n;
        c = b.ch_of[x];
        j = x - b.ch_off[c] + 1;
        top = max(j, b.ch_off[c + 1] - b.ch_off[c] + 1 - j);
        s = -1;
    }
    if (b.n_chain > 0) {
        b.lane_ch[k] = c;
        b.lane_ch[64 + k] = j;
    }
    const uint64_t core = __ballot(s >= 0);
    const uint64_t chl = __ballot(c >= 0);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) top = max(top, __shfl_xor(top, o, 64));
    if (s >= 0) {
        const uint64_t bit = 1ull << k;
        const int pos = __popcll(core & (bit - 1));
        b.vis[s] = bit;
        b.mb0[s] = bit;
        sg_put<SM>(b, (int64_t)s * 64 + k, SM == SIG_F64 ? d2u(1.0) : 1ull);
        b.Q[pos] = s;
        b.Qm[pos] = bit;
        if (b.lvl) b.lvl[(int64_t)s * 64 + k] = 0;
        if (b.n_chain > 0 && b.ep_idx[s] >= 0) b.ep_lvl[(int64_t)b.ep_idx[s] * 64 + k] = 0;
    }
    if (k == 0) {
        b.ctr[0] = __popcll(core);
        b.ctr[1] = 0;
        b.ctr[2] = 0;
        b.ctr[3] = -1;
        b.ctr[4] = top;  // last level with an injected chain-source lane
        b.ctr[5] = ns;
        b.qoff[0] = 0;
        if (b.n_chain > 0) *b.lane_mask = chl;
    }
}

// ------------------------------------------------------------- expand ---
__device__ __forceinline__ void push(const BatchDev &b, uint64_t *dst, int w, uint64_t F) {
    const uint64_t m = F & ~b.vis[w];
    if (m) atomicOr((unsigned long long *)&dst[w], (unsigned long long)m);  // result unused: no return
}

// Whether some chain-source lane still has an injection due after level l
// into an endpoint it has not reached (vis is not written during an expand
// launch, so every wave of the launch finds the same answer).
__device__ __forceinline__ bool seed_pending(const BatchDev &b, int l) {
    if (b.n_chain == 0) return false;
    const int lane = threadIdx.x & 63;
    const int c = b.lane_ch[lane];
    bool p = false;
    if (c >= 0) {
        const int j = b.lane_ch[64 + lane], K = b.ch_off[c + 1] - b.ch_off[c] + 1;
        const uint64_t bit = 1ull << lane;
        p = (j > l && !(b.vis[b.ch_L[c]] & bit)) || (K - j > l && !(b.vis[b.ch_R[c]] & bit));
    }
    return __ballot(p) != 0;
}

// Level l: Q_l = Q[qoff[l], ctr[0]).  Records qoff[l+1]; an empty Q_l ends the
// forward with depth l-1.  Clears Q_{l-1} from mask array (l+2)%3.
__global__ __launch_bounds__(BB) void k_b_expand(BatchDev b, int l) {
    if (ldi(&b.ctr[3]) >= 0) return;
    const int q0 = b.qoff[l];
    const int q1 = min(ldi(&b.ctr[0]), b.qcap);
    // an empty level ends the forward; so do full lists and a count that left
    // the sigma word (the host reruns the tile either way)
    bool stop = ldi(&b.ctr[2]) != 0 || ldi(&b.ctr[1]) != 0;
    if (!stop && q1 == q0) stop = l >= b.ctr[4] || !seed_pending(b, l);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        b.qoff[l + 1] = q1;
        if (stop) b.ctr[3] = l - 1;
    }
    if (stop) return;
    uint64_t *dst = mbuf(b, l + 1);
    if (l > 0) {
        uint64_t *clr = mbuf(b, l + 2);
        for (int i = b.qoff[l - 1] + blockIdx.x * BB + threadIdx.x; i < q0; i += gridDim.x * BB) clr[b.Q[i]] = 0;
    }
    const int lane = threadIdx.x & 63;
    if (b.n_chain > 0 && blockIdx.x == 0 && threadIdx.x < 64) {
        // chain-source lanes whose endpoints are due at level l+1
        const int c = b.lane_ch[lane];
        if (c >= 0) {
            const int j = b.lane_ch[64 + lane], K = b.ch_off[c + 1] - b.ch_off[c] + 1;
            if (j == l + 1) push(b, dst, b.ch_L[c], 1ull << lane);
            if (K - j == l + 1) push(b, dst, b.ch_R[c], 1ull << lane);
        }
    }
    for (int base = q0 + gwave() * 64; base < q1; base += nwaves() * 64) {
        const int e = base + lane;
        int j0 = 0, j1 = 0;
        uint64_t F = 0;
        if (e < q1) {
            const int v = b.Q[e];
            F = b.Qm[e];
            j0 = b.rp[v];
            j1 = b.rp[v + 1];
        }
        const int d = j1 - j0;
        if (d <= SMALL)
            for (int j = j0; j < j1; ++j) {
                const int w = b.col[j];
                if (w >= 0) push(b, dst, w, F);  // (w < 0: a contracted chain)
            }
        uint64_t bal = __ballot(d > SMALL && !is_hub(b, d));
        while (bal) {
            const int src = __ffsll((long long)bal) - 1;
            bal &= bal - 1;
            const uint64_t FF = shfl64(F, src);
            const int a0 = __shfl(j0, src, 64), a1 = __shfl(j1, src, 64);
            for (int j = a0 + lane; j < a1; j += 64) {
                const int w = b.col[j];
                if (w >= 0) push(b, dst, w, FF);
            }
        }
    }
    // hub slices: the hub's level-l lanes are its mask in mb[l%3]
    const uint64_t *cur = mbuf(b, l);
    for (int s = gwave(); s < b.n_slice; s += nwaves()) {
        const int v = b.hub_v[b.sl_hub[s]];
        const uint64_t F = cur[v];
        if (!F) continue;
        for (int j = b.sl_j0[s] + lane; j < b.sl_j1[s]; j += 64) {
            const int w = b.col[j];
            if (w >= 0) push(b, dst, w, F);
        }
    }
}

// --------------------------------------------------------------- pull ---
// sigma_w for the lanes N of entry w: the sum over neighbours at level l
// (fr) in CSR order.  A contracted chain's arc (~c) counts 1 for the lane
// whose source lies on chain c and is injected at w on this level (sc == c);
// cm holds the tile's chain-source lanes.
template <int SM>
__device__ __forceinline__ void sigma_sum(const BatchDev &b, const uint64_t *fr, uint64_t N, int a0, int a1,
                                          uint64_t &acc, double &facc, bool &ovf, int sc, uint64_t cm) {
    const int lane = threadIdx.x & 63;
    const uint64_t me = 1ull << lane;
    for (int c = a0; c < a1; c += 64) {
        const int j = c + lane;
        int v = 0;
        uint64_t f = 0;
        if (j < a1) {
            v = b.col[j];
            f = v >= 0 ? fr[v] & N : N & cm;
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
                        if (vv >= 0) {
                            use[u] = true;
                            x[u] = sg<SM>(b, (int64_t)vv * 64 + lane);
                        } else if (~vv == sc) {
                            use[u] = true;
                            x[u] = SM == SIG_F64 ? d2u(1.0) : 1ull;
                        }
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < UNR; ++u) {
                if (!use[u]) continue;
                if (SM == SIG_F64) {
                    facc += u2d(x[u]);
                } else {
                    const uint64_t s = acc + x[u];
                    ovf |= (s < acc);
                    acc = s;
                }
            }
        }
    }
}

__device__ __forceinline__ void stage_flush(const BatchDev &b, int *sw, uint64_t *sm, int &cnt, int lane) {
    if (cnt == 0) return;
    int base = 0;
    if (lane == 0) base = atomicAdd(&b.ctr[0], cnt);
    base = __shfl(base, 0, 64);
    __builtin_amdgcn_wave_barrier();
    for (int i = lane; i < cnt; i += 64) {
        const int pos = base + i;
        if (pos < b.qcap) {
            b.Q[pos] = sw[i];
            b.Qm[pos] = sm[i];
        } else {
            b.ctr[2] = 1;
        }
    }
    __builtin_amdgcn_wave_barrier();
    cnt = 0;
}

// Builds Q_{l+1} from the non-zero entries of mb[(l+1)%3] and their sigma.
template <int SM>
__global__ __launch_bounds__(BB) void k_b_pull(BatchDev b, int l) {
    constexpr bool F64 = SM == SIG_F64;
    __shared__ int s_w[NW][STAGE];
    __shared__ uint64_t s_m[NW][STAGE];
    if (ldi(&b.ctr[3]) >= 0) return;
    const uint64_t *fr = mbuf(b, l);
    const uint64_t *nx = mbuf(b, l + 1);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint64_t me = 1ull << lane;
    const bool split = !F64 && b.n_hub > 0;
    // chain-source lane: its endpoints injected at level l+1 (else -1)
    int myc = -1, mL = -1, mR = -1;
    uint64_t cm = 0;
    if (b.n_chain > 0) {
        cm = *b.lane_mask;
        myc = b.lane_ch[lane];
        if (myc >= 0) {
            const int j = b.lane_ch[64 + lane], K = b.ch_off[myc + 1] - b.ch_off[myc] + 1;
            if (j == l + 1) mL = b.ch_L[myc];
            if (K - j == l + 1) mR = b.ch_R[myc];
        }
    }
    int cnt = 0;  // wave-uniform
    for (int base = gwave() * 64; base < b.n; base += nwaves() * 64) {
        const int w0 = base + lane;
        const uint64_t N0 = w0 < b.n ? nx[w0] : 0;
        const uint64_t bal = __ballot(N0 != 0);
        if (!bal) continue;
        const int k0 = cnt;
        // CSR ranges of the chunk's entries, loaded lane-parallel up front
        int r0 = 0, r1 = 0, e0 = -1;
        if (N0) {
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
            const int w = __shfl(w0, src, 64);
            const uint64_t N = shfl64(N0, src);
            const int a0 = __shfl(r0, src, 64), a1 = __shfl(r1, src, 64);
            const int ew = __shfl(e0, src, 64);
            if (!(split && is_hub(b, a1 - a0))) {
                uint64_t acc = 0;
                double facc = 0.0;
                bool ovf = false;
                sigma_sum<SM>(b, fr, N, a0, a1, acc, facc, ovf, (w == mL || w == mR) ? myc : -1, cm);
                if (N & me) {
                    sg_put<SM>(b, (int64_t)w * 64 + lane, F64 ? d2u(facc) : acc);
                    if (!F64 && (ovf || acc > sg_max<SM>() || (b.limit && acc >= b.limit))) b.ctr[1] = 1;
                }
            }
            if ((N & me) && b.lvl) b.lvl[(int64_t)w * 64 + lane] = l + 1;
            if ((N & me) && ew >= 0) b.ep_lvl[(int64_t)ew * 64 + lane] = l + 1;
            if (lane == 0) b.vis[w] |= N;
        }
        if (cnt >= STAGE_FLUSH) stage_flush(b, s_w[wv], s_m[wv], cnt, lane);
    }
    stage_flush(b, s_w[wv], s_m[wv], cnt, lane);
    // hub slices: the level mask of every hub is recorded for the backward;
    // partial sigma sums are added atomically (exact integers)
    for (int s = gwave(); s < b.n_slice; s += nwaves()) {
        const int h = b.sl_hub[s];
        const int v = b.hub_v[h];
        const uint64_t N = nx[v];
        if (b.sl_j0[s] == b.rp[v] && lane == 0 && l + 1 < b.hub_stride)
            b.hub_mask[(int64_t)h * b.hub_stride + l + 1] = N;
        if (!split || !N) continue;
        uint64_t acc = 0;
        double facc = 0.0;
        bool ovf = false;
        constexpr int SI = F64 ? SIG_U64 : SM;  // (not reached in FP64 mode)
        sigma_sum<SI>(b, fr, N, b.sl_j0[s], b.sl_j1[s], acc, facc, ovf, (v == mL || v == mR) ? myc : -1, cm);
        if ((N & me) && acc) {
            const int64_t k = (int64_t)v * 64 + lane;
            uint64_t old, nw;
            if (SI == SIG_U32) {
                old = atomicAdd((unsigned int *)b.sig + k, (unsigned int)acc);
                nw = old + acc;
            } else {
                old = atomicAdd((unsigned long long *)b.sig + k, (unsigned long long)acc);
                nw = old + acc;
                ovf |= nw < old;
            }
            if (ovf || acc > sg_max<SI>() || nw > sg_max<SI>() || (b.limit && nw >= b.limit)) b.ctr[1] = 1;
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

// The chain terms of endpoint e (vertex at level l, count sv) for this lane:
// every attached chain is a segment from e to its other endpoint, except the
// chain holding the lane's own source x_j, which is the segment from e to the
// source (level 0, count 1, j arcs from L or K - j from R).
template <int SM>
__device__ double chain_rho(const BatchDev &b, int e, int l, double sv, int lane) {
    const int myc = b.lane_ch[lane], myj = b.lane_ch[64 + lane];
    double rho = 0.0;
    for (int q = b.ep_off[e]; q < b.ep_off[e + 1]; ++q) {
        const int code = b.ep_list[q];
        const int c = code >> 1, side = code & 1;
        const int off = b.ch_off[c], K = b.ch_off[c + 1] - off + 1;
        const ChainPfx P{b.ch_pre + off, side ? K - 1 : 0, side ? -1 : 1};
        if (c == myc) {
            rho += seg_term(l, 0, sv, 1.0, side ? K - myj : myj, P);
        } else {
            const int o = side ? b.ch_L[c] : b.ch_R[c];
            const int lv = b.ep_lvl[(int64_t)b.ep_idx[o] * 64 + lane];
            const double so = sg_val<SM>(sg<SM>(b, (int64_t)o * 64 + lane));
            rho += seg_term(l, lv, sv, so, K, P);
        }
    }
    return rho;
}

// rho_v = sum over successors w (lanes in succ[w] & M) of t_w.
template <typename TT>
__device__ __forceinline__ double rho_sum(const BatchDev &b, const uint64_t *succ, uint64_t M, int a0, int a1) {
    using Acc = typename std::conditional<sizeof(TT) == 4, KahanF, SumD>::type;
    const TT *tt = (const TT *)b.tt;
    const int lane = threadIdx.x & 63;
    const uint64_t me = 1ull << lane;
    Acc acc;
    for (int c = a0; c < a1; c += 64) {
        const int j = c + lane;
        int w = 0;
        uint64_t x = 0;
        if (j < a1) {
            w = b.col[j];
            x = w >= 0 ? succ[w] & M : 0;
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
    return acc.value();
}

// rho complete for (v, lanes M): delta, BC contribution and the term t_v.
template <int SM, typename TT>
__device__ __forceinline__ void finish_vertex(const BatchDev &b, int v, int l, uint64_t M, double acc) {
    const int lane = threadIdx.x & 63;
    const int e = b.n_chain > 0 ? b.ep_idx[v] : -1;
    double contrib = 0.0;
    if (M & (1ull << lane)) {
        const double sv = sg_val<SM>(sg<SM>(b, (int64_t)v * 64 + lane));
        if (e >= 0) acc += chain_rho<SM>(b, e, l, sv, lane);
        const double delta = sv * acc;
        contrib = b.wsl[lane] * delta;
        ((TT *)b.tt)[(int64_t)v * 64 + lane] = (TT)(((double)b.reach[v] + 1.0) / sv + acc);
        if (b.dbg_delta) b.dbg_delta[(int64_t)v * 64 + lane] = delta;
    }
    contrib = wave_sum(contrib);
    // b.bc is the slot's own accumulator and v is finished once per level
    // kernel, so a plain read-modify-write is race free
    if (lane == 0 && contrib != 0.0) b.bc[v] += contrib;
}

// Level l <= depth: entries Q_l; Q_{l+2} is cleared from mb[(l+2)%3] and Q_l
// scattered into mb[l%3] for level l-1.
template <int SM, typename TT>
__global__ __launch_bounds__(BB) void k_b_back(BatchDev b, int l) {
    constexpr bool F64 = SM == SIG_F64;
    __shared__ int t_w[NW][WIN];
    __shared__ uint64_t t_x[NW][WIN];
    __shared__ int t_k[NW][WIN];
    __shared__ double r_rho[NW][FEB][64];
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
    for (int i = c0 + blockIdx.x * BB + threadIdx.x; i < c1; i += gridDim.x * BB) clr[b.Q[i]] = 0;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint64_t me = 1ull << lane;
    using Acc = typename std::conditional<sizeof(TT) == 4, KahanF, SumD>::type;
    const TT *tt = (const TT *)b.tt;
    // Blocks of FEB list entries per wave.  The arcs of the block are walked
    // as one flat range (lane per arc) in windows of WIN: successor masks
    // are gathered for 64 arcs at once and the (entry, successor, lanes)
    // triples that matter are compacted in LDS; the t rows of those triples
    // are then gathered UNR at a time, accumulated per entry in arc order.
    for (int base = q0 + gwave() * FEB; base < q1; base += nwaves() * FEB) {
        const int e = base + lane;
        int v0 = 0, r0 = 0, r1 = 0, ep0 = -1;
        uint64_t M0 = 0;
        if (lane < FEB && e < q1) {
            v0 = b.Q[e];
            M0 = b.Qm[e];
            own[v0] = M0;
            r0 = b.rp[v0];
            r1 = b.rp[v0 + 1];
            if (b.n_chain > 0) ep0 = b.ep_idx[v0];
        }
        const int cnt = min(FEB, q1 - base);
        const bool mine = lane < cnt && !(is_hub(b, r1 - r0) && l < b.hub_stride);  // hubs: slices
        const int d = mine ? r1 - r0 : 0;
        int inc = d;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(inc, o, 64);
            if (lane >= o) inc += y;
        }
        const int pre = inc - d;
        const int T = __shfl(inc, 63, 64);
        int P[FEB], A[FEB], V[FEB], EP[FEB];
        uint64_t MM[FEB];
        bool OWN[FEB];
#pragma unroll
        for (int j = 0; j < FEB; ++j) {
            P[j] = j < cnt ? __shfl(pre, j, 64) : INT32_MAX;
            A[j] = __shfl(r0, j, 64);
            V[j] = __shfl(v0, j, 64);
            EP[j] = __shfl(ep0, j, 64);
            MM[j] = shfl64(M0, j);
            OWN[j] = __shfl((int)mine, j, 64) != 0;
            r_rho[wv][j][lane] = 0.0;
        }
        Acc acc;
        int kcur = -1;
        for (int win = 0; win < T; win += WIN) {
            const int wend = min(T, win + WIN);
            int nrel = 0;
            for (int c = win; c < wend; c += 64) {
                const int g = c + lane;
                int k = 0;
#pragma unroll
                for (int j = 1; j < FEB; ++j) k += (P[j] <= g);
                int a0k = A[0], pk = P[0];
                uint64_t mk = MM[0];
#pragma unroll
                for (int j = 1; j < FEB; ++j)
                    if (k == j) { a0k = A[j]; pk = P[j]; mk = MM[j]; }
                int w = 0;
                uint64_t x = 0;
                if (g < wend) {
                    w = b.col[a0k + g - pk];
                    x = w >= 0 ? succ[w] & mk : 0;  // (w < 0: a contracted chain)
                }
                const uint64_t bal = __ballot(x != 0);
                if (x) {
                    const int pos = nrel + __popcll(bal & (me - 1));
                    t_w[wv][pos] = w;
                    t_x[wv][pos] = x;
                    t_k[wv][pos] = k;
                }
                nrel += __popcll(bal);
            }
            __builtin_amdgcn_wave_barrier();
            for (int r = 0; r < nrel; r += UNR) {
                TT t[UNR];
                int kk[UNR];
                bool use[UNR];
#pragma unroll
                for (int u = 0; u < UNR; ++u) {
                    const int rr = r + u;
                    use[u] = false;
                    kk[u] = -1;
                    t[u] = 0;
                    if (rr < nrel) {
                        kk[u] = t_k[wv][rr];
                        if (t_x[wv][rr] & me) {
                            use[u] = true;
                            t[u] = tt[(int64_t)t_w[wv][rr] * 64 + lane];
                        }
                    }
                }
#pragma unroll
                for (int u = 0; u < UNR; ++u) {
                    if (kk[u] < 0) continue;
                    if (kk[u] != kcur) {
                        if (kcur >= 0) r_rho[wv][kcur][lane] = acc.value();
                        acc = Acc();
                        kcur = kk[u];
                    }
                    if (use[u]) acc.add(t[u]);
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
        if (kcur >= 0) r_rho[wv][kcur][lane] = acc.value();
        // finish every entry of the block (hubs excepted): sigma rows first
        uint64_t SX[FEB];
#pragma unroll
        for (int j = 0; j < FEB; ++j) SX[j] = (OWN[j] && (MM[j] & me)) ? sg<SM>(b, (int64_t)V[j] * 64 + lane) : 0;
#pragma unroll
        for (int j = 0; j < FEB; ++j) {
            if (!OWN[j]) continue;
            const int v = V[j];
            double contrib = 0.0;
            if (MM[j] & me) {
                const double sv = sg_val<SM>(SX[j]);
                double a = r_rho[wv][j][lane];  // rho_v
                if (EP[j] >= 0) a += chain_rho<SM>(b, EP[j], l, sv, lane);
                const double delta = sv * a;
                contrib = b.wsl[lane] * delta;
                ((TT *)b.tt)[(int64_t)v * 64 + lane] = (TT)(((double)b.reach[v] + 1.0) / sv + a);
                if (b.dbg_delta) b.dbg_delta[(int64_t)v * 64 + lane] = delta;
            }
            contrib = wave_sum(contrib);
            if (lane == 0 && contrib != 0.0) b.bc[v] += contrib;
        }
    }
    // hub slices (levels with recorded hub masks): partial rho into FP64
    // accumulators; the last slice of a hub to finish (counter) completes it
    if (l >= b.hub_stride) return;
    for (int s = gwave(); s < b.n_slice; s += nwaves()) {
        const int h = b.sl_hub[s];
        const uint64_t M = b.hub_mask[(int64_t)h * b.hub_stride + l];
        if (!M) continue;
        const int v = b.hub_v[h];
        const double part = rho_sum<TT>(b, succ, M, b.sl_j0[s], b.sl_j1[s]);
        if ((M & (1ull << lane)) && part != 0.0) unsafeAtomicAdd(&b.hub_acc[h * 64 + lane], part);
        __threadfence();
        int done = 0;
        if (lane == 0) done = atomicAdd(&b.hub_cnt[h], 1);
        done = __shfl(done, 0, 64);
        if (done != b.hub_nsl[h] - 1) continue;
        __threadfence();
        const double rho = __hip_atomic_load(&b.hub_acc[h * 64 + lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        finish_vertex<SM, TT>(b, v, l, M, rho);
        if (lane == 0) b.hub_cnt[h] = 0;
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
template <int SM>
__global__ __launch_bounds__(BB) void k_b_chain_check(BatchDev b) {
    if (ldi(&b.ctr[3]) < 0 || b.ctr[1] || b.ctr[2]) return;
    const int lane = threadIdx.x & 63;
    if (lane >= b.ctr[5]) return;
    const int myc = b.lane_ch[lane], myj = b.lane_ch[64 + lane];
    for (int c = gwave(); c < b.n_chain; c += nwaves()) {
        const int L = b.ch_L[c], R = b.ch_R[c], K = b.ch_off[c + 1] - b.ch_off[c] + 1;
        const int aL = b.ep_lvl[(int64_t)b.ep_idx[L] * 64 + lane];
        const int aR = b.ep_lvl[(int64_t)b.ep_idx[R] * 64 + lane];
        const uint64_t wL = sg<SM>(b, (int64_t)L * 64 + lane), wR = sg<SM>(b, (int64_t)R * 64 + lane);
        // the segments of this lane: L .. R, or source .. L and source .. R;
        // a tie needs an even S = b - a + K with 2 <= S <= 2K - 2
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

// After the backward: every contracted interior of every chain, one wave per
// chain (the launch has a wave for every chain, so the chains' dependent
// loads overlap) and one lane per source; BC into the slot's accumulator
// (each interior has one writer), and in the probe the interiors' rows.  For a lane
// whose source is off the chain, the segment L .. R's constants (split
// point, tie, the prefix sums at the split) are set once per chain, so an
// interior costs the chain's running prefix sum (the same word for every
// lane), a few selects and the wave sum; the rare lane whose source lies on
// the chain takes the general form for its two halves.
template <int SM, typename TT>
__global__ __launch_bounds__(BB) void k_b_chains(BatchDev b) {
    constexpr bool F64 = SM == SIG_F64;
    if (ldi(&b.ctr[3]) < 0 || (!F64 && b.ctr[1]) || b.ctr[2]) return;
    const int lane = threadIdx.x & 63;
    const bool act = lane < b.ctr[5];
    const double w = b.wsl[lane];
    const int myc = b.lane_ch[lane], myj = b.lane_ch[64 + lane];
    const uint64_t one = F64 ? d2u(1.0) : 1ull;
    const TT *tt = (const TT *)b.tt;
    for (int c = gwave(); c < b.n_chain; c += nwaves()) {
        const int L = b.ch_L[c], R = b.ch_R[c], off = b.ch_off[c], K = b.ch_off[c + 1] - off + 1;
        const double *pre = b.ch_pre + off;
        const ChainPfx P{pre, 0, 1};
        int aL = 0, aR = 0;
        uint64_t wL = 0, wR = 0;
        double sL = 1.0, sR = 1.0, tL = 0.0, tR = 0.0;
        if (act) {
            aL = b.ep_lvl[(int64_t)b.ep_idx[L] * 64 + lane];
            aR = b.ep_lvl[(int64_t)b.ep_idx[R] * 64 + lane];
            wL = sg<SM>(b, (int64_t)L * 64 + lane);
            wR = sg<SM>(b, (int64_t)R * 64 + lane);
            sL = sg_val<SM>(wL);
            sR = sg_val<SM>(wR);
            if (c == myc) {
                tL = (double)tt[(int64_t)L * 64 + lane];
                tR = (double)tt[(int64_t)R * 64 + lane];
            }
        }
        const bool own = act && c == myc;
        const bool any_own = __ballot(own) != 0;
        // segment L .. R seen from L: split, tie and the prefix sums there
        const int S = aR - aL + K;
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
        mA = min(mA, K - 1);
        const double tw = tie ? P(t) - P(t - 1) : 0.0;
        const double cA = P(mA) + (tie ? tw * sL / (sL + sR) : 0.0);
        const double cB = P(min(b0, K) - 1) - (tie ? tw * sR / (sL + sR) : 0.0);
        const uint64_t wT = F64 ? d2u(sL + sR) : wL + wR;
        double p0 = 0.0;  // P(i - 1)
        // the interiors' BC words, 64 at a time, one per lane (coalesced)
        double acc = 0.0;
        for (int i = 1; i < K; ++i) {
            const int64_t x = (int64_t)b.n + off + i - 1;
            if (((i - 1) & 63) == 0 && i - 1 + lane < K - 1) acc = b.bc[x + lane];
            const double p1 = pre[i - 1];  // P(i)
            int lev;
            uint64_t sw;
            double dl;
            if (i <= mA) {
                lev = aL + i;
                sw = wL;
                dl = cA - p1;
            } else if (tie && i == t) {
                lev = aL + t;
                sw = wT;
                dl = 0.0;
            } else {
                lev = aR + K - i;
                sw = wR;
                dl = p0 - cB;
            }
            if (any_own && own) {
                lev = 0;
                sw = one;
                dl = 0.0;
                if (i < myj)  // the half from the source to L, numbered from the source
                    seg_interior<SM>(0, aL, 1.0, sL, one, wL, myj, myj - i, ChainPfx{pre, myj - 1, -1}, tL, lev, sw,
                                     dl);
                else if (i > myj)  // the half from the source to R
                    seg_interior<SM>(0, aR, 1.0, sR, one, wR, K - myj, i - myj, ChainPfx{pre, myj, 1}, tR, lev, sw,
                                     dl);
            }
            p0 = p1;
            const double contrib = wave_sum(act ? w * dl : 0.0);
            if (lane == ((i - 1) & 63)) acc += contrib;
            if ((((i - 1) & 63) == 63 || i == K - 1) && lane <= ((i - 1) & 63))
                b.bc[(int64_t)b.n + off + ((i - 1) & ~63) + lane] = acc;
            if (b.lvl && act) {
                b.lvl[x * 64 + lane] = lev;
                sg_put<SM>(b, x * 64 + lane, sw);
                b.dbg_delta[x * 64 + lane] = dl;
            }
        }
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
hipError_t launch_b_init(const BatchDev &b, int ns, int sm, int cap, hipStream_t st) {
    int g = (int)std::min<int64_t>(((int64_t)b.n + BB - 1) / BB, cap);
    hipLaunchKernelGGL(k_b_clear_vis, dim3(std::max(g, 1)), dim3(BB), 0, st, b, sm);
    if (sm == SIG_U32) hipLaunchKernelGGL(k_b_seed<SIG_U32>, dim3(1), dim3(64), 0, st, b, ns);
    else if (sm == SIG_U64) hipLaunchKernelGGL(k_b_seed<SIG_U64>, dim3(1), dim3(64), 0, st, b, ns);
    else hipLaunchKernelGGL(k_b_seed<SIG_F64>, dim3(1), dim3(64), 0, st, b, ns);
    return hipGetLastError();
}

hipError_t launch_b_forward(const BatchDev &b, int sm, int l, int grid, hipStream_t st) {
    hipLaunchKernelGGL(k_b_expand, dim3(grid), dim3(BB), 0, st, b, l);
    if (sm == SIG_U32) hipLaunchKernelGGL(k_b_pull<SIG_U32>, dim3(grid), dim3(BB), 0, st, b, l);
    else if (sm == SIG_U64) hipLaunchKernelGGL(k_b_pull<SIG_U64>, dim3(grid), dim3(BB), 0, st, b, l);
    else hipLaunchKernelGGL(k_b_pull<SIG_F64>, dim3(grid), dim3(BB), 0, st, b, l);
    return hipGetLastError();
}

template <int SM>
static void back_tt(const BatchDev &b, bool fp32, int l, int grid, hipStream_t st) {
    if (fp32) hipLaunchKernelGGL((k_b_back<SM, float>), dim3(grid), dim3(BB), 0, st, b, l);
    else hipLaunchKernelGGL((k_b_back<SM, double>), dim3(grid), dim3(BB), 0, st, b, l);
}

// FP64-sigma tiles keep FP64 terms in either precision: 1/sigma can leave
// the f32 range.
hipError_t launch_b_back(const BatchDev &b, int sm, bool fp32, int l, int grid, hipStream_t st) {
    if (sm == SIG_U32) back_tt<SIG_U32>(b, fp32, l, grid, st);
    else if (sm == SIG_U64) back_tt<SIG_U64>(b, fp32, l, grid, st);
    else hipLaunchKernelGGL((k_b_back<SIG_F64, double>), dim3(grid), dim3(BB), 0, st, b, l);
    return hipGetLastError();
}

hipError_t launch_b_finish(const BatchDev &b, int grid, hipStream_t st) {
    hipLaunchKernelGGL(k_b_finish, dim3(grid), dim3(BB), 0, st, b);
    return hipGetLastError();
}

// one wave per chain (2 per block)
static int chain_grid(const BatchDev &b) {
    return (int)std::max<int64_t>(1, std::min<int64_t>(((int64_t)b.n_chain + NW - 1) / NW, 1 << 16));
}

hipError_t launch_b_chain_check(const BatchDev &b, int sm, int grid, hipStream_t st) {
    (void)grid;
    grid = chain_grid(b);
    if (sm == SIG_U32) hipLaunchKernelGGL(k_b_chain_check<SIG_U32>, dim3(grid), dim3(BB), 0, st, b);
    else if (sm == SIG_U64) hipLaunchKernelGGL(k_b_chain_check<SIG_U64>, dim3(grid), dim3(BB), 0, st, b);
    return hipGetLastError();
}

hipError_t launch_b_chains(const BatchDev &b, int sm, bool fp32, int grid, hipStream_t st) {
    (void)grid;
    grid = chain_grid(b);
    if (sm == SIG_U32 && fp32) hipLaunchKernelGGL((k_b_chains<SIG_U32, float>), dim3(grid), dim3(BB), 0, st, b);
    else if (sm == SIG_U32) hipLaunchKernelGGL((k_b_chains<SIG_U32, double>), dim3(grid), dim3(BB), 0, st, b);
    else if (sm == SIG_U64 && fp32) hipLaunchKernelGGL((k_b_chains<SIG_U64, float>), dim3(grid), dim3(BB), 0, st, b);
    else if (sm == SIG_U64) hipLaunchKernelGGL((k_b_chains<SIG_U64, double>), dim3(grid), dim3(BB), 0, st, b);
    else hipLaunchKernelGGL((k_b_chains<SIG_F64, double>), dim3(grid), dim3(BB), 0, st, b);
    return hipGetLastError();
}

hipError_t launch_b_fold(double *dst, double *src, const int32_t *perm, int n, int grid, hipStream_t st) {
    hipLaunchKernelGGL(k_b_fold, dim3(grid), dim3(BB), 0, st, dst, src, perm, n);
    return hipGetLastError();
}

hipError_t launch_b_clear_all(const BatchDev &b, int end, int grid, hipStream_t st) {
    if (end <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_b_clear_all, dim3(grid), dim3(BB), 0, st, b, end);
    return hipGetLastError();
}

}  // namespace bcx
