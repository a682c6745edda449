// Workgroup-per-tile kernel for components of up to WG_N vertices with
// narrow levels (deep BFS: grids, meshes, long paths).  The multi-launch
// tiles pay three launches per level, which a BFS hundreds of levels deep
// turns into milliseconds of launch latency per tile; here one workgroup
// takes a tile = (component, up to 64 W sources) from a work counter and runs
// every level of it with workgroup barriers, while hundreds of workgroups run
// other tiles.
//
// Same method as the batched kernels: lane k of a wave64 owns sources
// k + 64 h, h < W (W = 1, 2 or 4 sources per lane); a source mask is W 64-bit
// words.  The visited set and two level masks (W words per vertex), the level
// lists and their offsets live in a per-workgroup HBM slab.  Path counts and
// terms are lane-compacted records: list entry (v, sources M) owns
// popcount(M) consecutive words, word h first (h-major), the source of lane k
// in word h at rank popcount(M[0..h-1]) + popcount(M[h] & (2^k - 1)), from a
// record offset taken when the entry's sigma is pulled; each vertex's offset
// sits beside its level mask.  On deep graphs an entry holds a few sources
// per word, so a record is one cache line where a [v][64 W] row of FP64 words
// is 4 W, and the records of one level are contiguous.  Level l+1's list is
// appended by the first mask OR that reaches a vertex (the atomic OR that
// finds zero; with W > 1 words an exchange of the vertex's level tag keeps
// one of the ORs that found a zero word), so no pass over all n vertices is
// made per level.  Pull and backward walk the arcs of 16 (pull) or 8 / W
// (backward) list entries per wave as one flat range (lane per arc), compact
// the arcs that matter in LDS and then gather one record word per (arc,
// source), summed per entry in CSR order (exact u64, FP64 in the oracle's
// order).  The backward is the rho form in FP32 with Kahan compensation or
// FP64 (always FP64 with FP64 sigma), BC(v) += sum over sources s of
// (r(s)+1) delta_s(v) (Codes/kernels_free.cu:154-473 computes the same, one
// source per block).  A tile whose uint64 sigma overflows stops, adds nothing
// and is listed for an FP64 rerun by the host.
//
// Wider lanes (W > 1) put more sources behind each level of a tile: a tile
// of a deep component costs about the same number of dependent level steps
// whatever its width, and its sources share their entries (a vertex is
// reached at few more distinct levels by 256 nearby sources than by 64).
//
// Mask words (level masks and the visited set) are OR-ed with L2 atomics
// while other waves of the workgroup read them, so every read of a mask
// array goes through agent-scope loads (no stale vector-L1 line).
#include "bcx_device.h"
#include "bcx_small.h"

namespace bcx {

constexpr int GB = 1024;     // threads per workgroup (one workgroup per CU)
constexpr int GW = GB / 64;  // waves per workgroup
constexpr int GEB = 16;      // list entries per wave block (pull)
constexpr int GUNR = 8;      // record words in flight per lane (64-bit words)

// Sizes that depend on the sources per lane W: backward entries per wave
// block (8 W-word entries hold as many (entry, source) pairs as 8 / W ...),
// and arcs per compaction window (two 64-arc chunks at W = 1, one above, so
// the per-wave LDS stays within a workgroup of 16 waves).
template <int W>
struct WgShape {
    static constexpr int BEB = 8 / W;
    static constexpr int WIN = W == 1 ? 128 : 64;
};

struct KahanW {
    float s = 0.f, c = 0.f;
    __device__ __forceinline__ void add(float x) {
        const float y = x - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __device__ __forceinline__ double value() const { return (double)s - (double)c; }
};
struct SumW {
    double s = 0.0;
    __device__ __forceinline__ void add(double x) { s += x; }
    __device__ __forceinline__ double value() const { return s; }
};

// LDS of one wave: the block's entries (first flat arc index P, first CSR
// arc A, vertex V, record R, source mask M), the compacted arcs of a window
// (far end's full level mask fm, its record rr, entry k: the sources that
// matter are fm & M[k]) and the backward's per-entry sums.
template <int W>
struct WaveLds {
    int P[GEB], A[GEB], V[GEB], R[GEB];
    uint64_t M[GEB][W];
    uint64_t fm[WgShape<W>::WIN][W];
    int rr[WgShape<W>::WIN];
    int k[WgShape<W>::WIN];
    double r[WgShape<W>::BEB][64 * W];
};

__device__ __forceinline__ uint64_t wsh64(uint64_t x, int src) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, src);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), src);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t uni64(uint64_t x) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)x);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(x >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// A wave-uniform record mask (read from LDS at a uniform address) in scalar
// registers, with the sources of the words before each word: the source of
// lane k in word h sits at rank pre[h] + popcount(m[h] & (2^k - 1)).
template <int W>
struct RecMask {
    uint64_t m[W];
    int pre[W];
    __device__ __forceinline__ explicit RecMask(const uint64_t *M) {
        int p = 0;
#pragma unroll
        for (int h = 0; h < W; ++h) {
            m[h] = uni64(M[h]);
            pre[h] = p;
            p += __popcll(m[h]);
        }
    }
    __device__ __forceinline__ int rank(int h, uint64_t below) const { return pre[h] + __popcll(m[h] & below); }
};

// Workgroup barrier after every outstanding memory access of the wave has
// completed, so plain stores have reached L2 before other waves read the
// words through L2 (agent-scope loads, atomics).
__device__ __forceinline__ void wg_barrier() {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
}

// Block layout in L: lane j < GEB describes entry j (vertex v, mask m,
// record rec, CSR arcs from r0, d of them walked); P[j] = first flat arc
// index (INT32_MAX past cnt).  Returns the block's flat arc count.
template <int W>
__device__ __forceinline__ int block_layout(int d, int cnt, int r0, int v, const uint64_t (&m)[W], int rec,
                                            WaveLds<W> &L) {
    const int lane = threadIdx.x & 63;
    int inc = d;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(inc, o, 64);
        if (lane >= o) inc += y;
    }
    if (lane < GEB) {
        L.P[lane] = lane < cnt ? inc - d : INT32_MAX;
        L.A[lane] = r0;
        L.V[lane] = v;
#pragma unroll
        for (int h = 0; h < W; ++h) L.M[lane][h] = m[h];
        L.R[lane] = rec;
    }
    __builtin_amdgcn_wave_barrier();
    return __builtin_amdgcn_readlane(inc, 63);
}

// Entry of flat arc index g: the last j with P[j] <= g (binary search).
template <int W>
__device__ __forceinline__ int entry_of(const WaveLds<W> &L, int g) {
    int k = 0;
#pragma unroll
    for (int h = GEB / 2; h > 0; h >>= 1)
        if (L.P[k + h] <= g) k += h;
    return k;
}

// Compacts the arcs [win, wend) of a block whose far-end mask & entry mask is
// non-zero into L, in arc order, with the far end's full mask and record
// offset (rfar: plain stores of this workgroup's waves, so ordinary loads);
// returns the number of compacted arcs.  The window's 64-arc chunks load
// their arcs and masks before any is compacted.
template <int W>
__device__ __forceinline__ int compact(const int32_t *col, const uint64_t *far, const int32_t *rfar, int win,
                                       int wend, WaveLds<W> &L) {
    constexpr int NCH = WgShape<W>::WIN / 64;
    const int lane = threadIdx.x & 63;
    const uint64_t me = 1ull << lane;
    int g[NCH], kk[NCH], w[NCH];
    uint64_t f[NCH][W];
    bool rel[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        g[c] = win + 64 * c + lane;
        kk[c] = entry_of(L, g[c]);
        w[c] = 0;
        if (g[c] < wend) w[c] = col[L.A[kk[c]] + g[c] - L.P[kk[c]]];
    }
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        rel[c] = false;
#pragma unroll
        for (int h = 0; h < W; ++h) {
            f[c][h] = g[c] < wend ? ald(&far[(int64_t)w[c] * W + h]) : 0;
            rel[c] |= (f[c][h] & L.M[kk[c]][h]) != 0;
        }
    }
    int n = 0;
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        const uint64_t bl = __ballot(rel[c]);
        if (rel[c]) {
            const int pos = n + __popcll(bl & (me - 1));
#pragma unroll
            for (int h = 0; h < W; ++h) L.fm[pos][h] = f[c][h];
            L.rr[pos] = rfar[w[c]];
            L.k[pos] = kk[c];
        }
        n += __popcll(bl);
    }
    __builtin_amdgcn_wave_barrier();
    return n;
}

// Metadata of a block's entry on lane j < cnt (loaded a block ahead): vertex,
// CSR range, source mask and (backward) record offset.
template <int W>
struct EntryMeta {
    int v, r0, r1, rec;
    uint64_t m[W];
};

// Pull of list entries [e0, e0+cnt) of level l+1 (lane j: entry j's vertex,
// final source mask and CSR range): records of the block taken from the
// tile's record counter (LDS), sigma over predecessors in CSR order, visited
// update.  Each entry is finished as soon as the walk moves past it (every
// entry has at least one predecessor arc for each of its sources).
template <bool F64, int W>
__device__ __forceinline__ void pull_block(uint64_t limit, const int32_t *col, const uint64_t *cur,
                                           const int32_t *rcur, int32_t *rnxt, uint64_t *Qm, int32_t *Qr,
                                           uint64_t *sig, uint64_t *vis, int e0, int cnt, const EntryMeta<W> &em,
                                           WaveLds<W> &L, int *s_ovf, int *s_rec, int32_t *plv, uint64_t *psig,
                                           int level) {
    constexpr int UW = GUNR / W;  // arcs gathered per round (W words each)
    const int lane = threadIdx.x & 63;
    const uint64_t me = 1ull << lane;
    const int w0 = em.v, r0 = em.r0, r1 = em.r1;
    // record offsets: exclusive prefix of the entries' source counts
    int pc = 0;
    if (lane < cnt)
#pragma unroll
        for (int h = 0; h < W; ++h) pc += __popcll(em.m[h]);
    int inc = pc;
#pragma unroll
    for (int o = 1; o < GEB; o <<= 1) {
        const int y = __shfl_up(inc, o, 64);
        if (lane >= o) inc += y;
    }
    const int total = __builtin_amdgcn_readlane(inc, GEB - 1);  // (lanes >= cnt add 0)
    int base = 0;
    if (lane == 0) base = atomicAdd(s_rec, total);
    const int rec = __builtin_amdgcn_readlane(base, 0) + inc - pc;
    if (lane < cnt) {
#pragma unroll
        for (int h = 0; h < W; ++h) Qm[(int64_t)(e0 + lane) * W + h] = em.m[h];
        Qr[e0 + lane] = rec;
        rnxt[w0] = rec;
    }
    const int T = block_layout<W>(lane < cnt ? r1 - r0 : 0, cnt, r0, w0, em.m, rec, L);
    uint64_t acc[W];
    double facc[W];
#pragma unroll
    for (int h = 0; h < W; ++h) {
        acc[h] = 0;
        facc[h] = 0.0;
    }
    bool ovf = false;
    int kcur = -1;
    auto finish = [&](int j) {
        const int w = __builtin_amdgcn_readfirstlane(L.V[j]);
        const int rec = __builtin_amdgcn_readfirstlane(L.R[j]);
        const RecMask<W> M(L.M[j]);
#pragma unroll
        for (int h = 0; h < W; ++h) {
            const uint64_t val = F64 ? d2u(facc[h]) : acc[h];
            if (M.m[h] & me) {
                sig[rec + M.rank(h, me - 1)] = val;
                if (!F64 && limit && val >= limit) ovf = true;
                const int64_t cell = ((int64_t)w * 64 + lane) * W + h;
                if (plv) plv[cell] = level;
                if (psig) psig[cell] = val;
            }
            acc[h] = 0;
            facc[h] = 0.0;
        }
        if (lane == 0)
#pragma unroll
            for (int h = 0; h < W; ++h)
                if (M.m[h])  // result unused: no return
                    atomicOr((unsigned long long *)&vis[(int64_t)w * W + h], (unsigned long long)M.m[h]);
    };
    for (int win = 0; win < T; win += WgShape<W>::WIN) {
        const int nrel = compact<W>(col, cur, rcur, win, min(T, win + WgShape<W>::WIN), L);
        for (int r = 0; r < nrel; r += UW) {
            uint64_t xv[UW][W];
            int kk[UW], use[UW];
#pragma unroll
            for (int u = 0; u < UW; ++u) {
                const int rr = r + u;
                use[u] = 0;
                kk[u] = -1;
#pragma unroll
                for (int h = 0; h < W; ++h) xv[u][h] = 0;
                if (rr < nrel) {
                    kk[u] = __builtin_amdgcn_readfirstlane(L.k[rr]);
                    const RecMask<W> F(L.fm[rr]);
                    const int rec = __builtin_amdgcn_readfirstlane(L.rr[rr]);
#pragma unroll
                    for (int h = 0; h < W; ++h)
                        if (F.m[h] & uni64(L.M[kk[u]][h]) & me) {
                            use[u] |= 1 << h;
                            xv[u][h] = sig[rec + F.rank(h, me - 1)];
                        }
                }
            }
#pragma unroll
            for (int u = 0; u < UW; ++u) {
                if (kk[u] < 0) continue;
                if (kk[u] != kcur) {  // wave-uniform
                    if (kcur >= 0) finish(kcur);
                    while (++kcur < kk[u]) finish(kcur);  // (entries without arcs: none in a pull)
                }
#pragma unroll
                for (int h = 0; h < W; ++h) {
                    if (!((use[u] >> h) & 1)) continue;
                    if (F64) {
                        facc[h] += u2d(xv[u][h]);
                    } else {
                        const uint64_t s = acc[h] + xv[u][h];
                        ovf |= (s < acc[h]);
                        acc[h] = s;
                    }
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    if (kcur >= 0) finish(kcur);
    while (++kcur < cnt) finish(kcur);
    if (!F64 && __ballot(ovf) && lane == 0) *s_ovf = 1;
}

// Backward of list entries [e0, e0+cnt) (cnt <= BEB) of level lb (lane j:
// entry j's vertex, source mask and CSR range): scatter own masks, sum
// successor terms per entry into LDS (has_succ false at the deepest level),
// then load the block's sigma words together and finish.
template <bool F64, typename TT, int W>
__device__ __forceinline__ void back_block(const int32_t *col, const int32_t *reach, double *bc, const double *ws,
                                           const uint64_t *succ, const int32_t *rsucc, uint64_t *own, int32_t *rown,
                                           const uint64_t *sig, TT *tt, int cnt, bool has_succ,
                                           const EntryMeta<W> &em, WaveLds<W> &L, double *pdl) {
    using Acc = typename std::conditional<sizeof(TT) == 4, KahanW, SumW>::type;
    constexpr int BEB = WgShape<W>::BEB;
    // arcs gathered per round (32-bit words: twice as many in flight)
    constexpr int BUNR = (sizeof(TT) == 8 ? GUNR : 2 * GUNR) / W;
    const int lane = threadIdx.x & 63;
    const uint64_t me = 1ull << lane;
    const int v0 = em.v, r0 = em.r0, r1 = em.r1;
    if (lane < cnt) {
#pragma unroll
        for (int h = 0; h < W; ++h) own[(int64_t)v0 * W + h] = em.m[h];
        rown[v0] = em.rec;
    }
    const int T = block_layout<W>(has_succ && lane < cnt ? r1 - r0 : 0, cnt, r0, v0, em.m, em.rec, L);
#pragma unroll
    for (int j = 0; j < BEB; ++j)
#pragma unroll
        for (int h = 0; h < W; ++h) L.r[j][h * 64 + lane] = 0.0;
    Acc acc[W];
    int kcur = -1;
    for (int win = 0; win < T; win += WgShape<W>::WIN) {
        const int nrel = compact<W>(col, succ, rsucc, win, min(T, win + WgShape<W>::WIN), L);
        for (int r = 0; r < nrel; r += BUNR) {
            TT t[BUNR][W];
            int kk[BUNR], use[BUNR];
#pragma unroll
            for (int u = 0; u < BUNR; ++u) {
                const int rr = r + u;
                use[u] = 0;
                kk[u] = -1;
#pragma unroll
                for (int h = 0; h < W; ++h) t[u][h] = 0;
                if (rr < nrel) {
                    kk[u] = __builtin_amdgcn_readfirstlane(L.k[rr]);
                    const RecMask<W> F(L.fm[rr]);
                    const int rec = __builtin_amdgcn_readfirstlane(L.rr[rr]);
#pragma unroll
                    for (int h = 0; h < W; ++h)
                        if (F.m[h] & uni64(L.M[kk[u]][h]) & me) {
                            use[u] |= 1 << h;
                            t[u][h] = tt[rec + F.rank(h, me - 1)];
                        }
                }
            }
#pragma unroll
            for (int u = 0; u < BUNR; ++u) {
                if (kk[u] < 0) continue;
                if (kk[u] != kcur) {
                    if (kcur >= 0)
#pragma unroll
                        for (int h = 0; h < W; ++h) L.r[kcur][h * 64 + lane] = acc[h].value();
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
        for (int h = 0; h < W; ++h) L.r[kcur][h * 64 + lane] = acc[h].value();
    __builtin_amdgcn_wave_barrier();
    uint64_t sx[BEB][W];
#pragma unroll
    for (int j = 0; j < BEB; ++j) {
        const RecMask<W> M(L.M[j]);
        const int rec = __builtin_amdgcn_readfirstlane(L.R[j]);
#pragma unroll
        for (int h = 0; h < W; ++h) sx[j][h] = (j < cnt && (M.m[h] & me)) ? sig[rec + M.rank(h, me - 1)] : 0;
    }
#pragma unroll
    for (int j = 0; j < BEB; ++j) {
        if (j >= cnt) break;
        const int v = __builtin_amdgcn_readfirstlane(L.V[j]);
        const int rec = __builtin_amdgcn_readfirstlane(L.R[j]);
        const RecMask<W> M(L.M[j]);
        double contrib = 0.0;
#pragma unroll
        for (int h = 0; h < W; ++h) {
            if (!(M.m[h] & me)) continue;
            const double sv = F64 ? u2d(sx[j][h]) : (double)sx[j][h];
            const double rho = L.r[j][h * 64 + lane];
            const double delta = sv * rho;
            contrib += ws[h * 64 + lane] * delta;
            tt[rec + M.rank(h, me - 1)] = (TT)(((double)reach[v] + 1.0) / sv + rho);
            if (pdl) pdl[((int64_t)v * 64 + lane) * W + h] = delta;
        }
        contrib = wave_sum(contrib);
        if (lane == 0 && contrib != 0.0) unsafeAtomicAdd(&bc[v], contrib);
    }
}

// Appends w to level l+1's list after this lane's ORs of m into w's level
// mask returned old: the OR that found the mask zero does it.  With W > 1
// words, every OR that found its word zero is a candidate, and exchanging the
// vertex's level tag keeps one of them.
template <int W>
__device__ __forceinline__ bool first_reach(const uint64_t (&m)[W], const uint64_t (&old)[W], int32_t *tag, int w,
                                            int lvl) {
    bool cand = false;
#pragma unroll
    for (int h = 0; h < W; ++h) cand |= m[h] != 0 && old[h] == 0;
    if (W == 1 || !cand) return cand;
    return atomicExch(&tag[w], lvl) != lvl;
}

template <bool F64, typename TT, int W>
__global__ __launch_bounds__(GB) void k_wg(SmallArgs a) {
    __shared__ WaveLds<W> lds[GW];
    __shared__ int s_item, s_cnt, s_ovf, s_rec;
    constexpr int BEB = WgShape<W>::BEB;
    constexpr int AU = W == 1 ? 4 : (W == 2 ? 2 : 1);  // small-degree arcs pushed per step
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int64_t nm = a.nmax;
    uint64_t *sig = a.sig + (int64_t)blockIdx.x * a.slab;
    TT *tt = (TT *)a.tt + (int64_t)blockIdx.x * a.slab;
    int32_t *Q = a.Q + (int64_t)blockIdx.x * a.qcap;
    uint64_t *Qm = a.Qm + (int64_t)blockIdx.x * a.qcap * W;
    uint64_t *vis = a.masks + (int64_t)blockIdx.x * 3 * nm * W;
    uint64_t *mA = vis + nm * W, *mB = mA + nm * W;
    int32_t *qoff = a.qoffs + (int64_t)blockIdx.x * (nm + 2);
    int32_t *Qr = a.qr + (int64_t)blockIdx.x * a.qcap;               // record offset of each list entry
    int32_t *rA = a.ro + (int64_t)blockIdx.x * 2 * nm, *rB = rA + nm;  // of each vertex, with mA / mB
    int32_t *tag = W > 1 ? a.tag + (int64_t)blockIdx.x * nm : nullptr;  // level a vertex was last listed at
    int32_t *plv = a.lvl ? a.lvl + (int64_t)blockIdx.x * a.slab : nullptr;
    double *pdl = a.dbg ? a.dbg + (int64_t)blockIdx.x * a.slab : nullptr;
    uint64_t *psg = a.psig ? a.psig + (int64_t)blockIdx.x * a.slab : nullptr;
    while (true) {
        if (tid == 0) s_item = atomicAdd(&a.counter[0], 1);
        wg_barrier();
        const int item = s_item;
        if (item >= a.n_tiles) break;
        const SmallTile tl = a.tiles[item];
        const SmallComp cp = a.comps[tl.comp];
        const int n = cp.n;
        const int32_t *rp = cp.rp, *col = cp.col;
        for (int64_t i = tid; i < (int64_t)n * W; i += GB) {
            vis[i] = 0;
            mA[i] = 0;
            mB[i] = 0;
        }
        if (W > 1)
            for (int v = tid; v < n; v += GB) tag[v] = 0;
        if (tid == 0) {
            s_cnt = 0;
            s_ovf = 0;
            s_rec = tl.ns;  // records 0 .. ns-1: the sources, one word each
            qoff[0] = 0;
            qoff[1] = tl.ns;
        }
        wg_barrier();
        if (tid < tl.ns) {
            // slot tid = source of lane tid % 64, word tid / 64; its level-0
            // entry and record (a source vertex holds one source)
            const int s = a.srcs[tl.src0 + tid];
            const int h = tid >> 6;
            const uint64_t bit = 1ull << (tid & 63);
            vis[(int64_t)s * W + h] = bit;
            mA[(int64_t)s * W + h] = bit;
            rA[s] = tid;
            sig[tid] = F64 ? d2u(1.0) : 1ull;
            Q[tid] = s;
#pragma unroll
            for (int q = 0; q < W; ++q) Qm[(int64_t)tid * W + q] = q == h ? bit : 0;
            Qr[tid] = tid;
            const int64_t cell = ((int64_t)s * 64 + (tid & 63)) * W + h;
            if (plv) plv[cell] = 0;
            if (psg) psg[cell] = F64 ? d2u(1.0) : 1ull;
        }
        wg_barrier();
        uint64_t *cur = mA, *nxt = mB;
        int32_t *rcur = rA, *rnxt = rB;
        int l = 0, q0 = 0, q1 = tl.ns;
        // ------------------------------------------------------- forward
        while (true) {
            // expand: lane per entry, a whole wave for an entry of degree > 16;
            // the first OR that reaches a vertex appends it to Q_{l+1}
            for (int base = q0 + wv * 64; base < q1; base += GB) {
                const int e = base + lane;
                int j0 = 0, j1 = 0;
                uint64_t F[W];
#pragma unroll
                for (int h = 0; h < W; ++h) F[h] = 0;
                if (e < q1) {
                    const int v = Q[e];
#pragma unroll
                    for (int h = 0; h < W; ++h) F[h] = Qm[(int64_t)e * W + h];
                    j0 = rp[v];
                    j1 = rp[v + 1];
                }
                if (j1 - j0 <= 16) {
                    // AU arcs at a time: their loads and ORs are independent
                    for (int j = j0; j < j1; j += AU) {
                        int w[AU];
                        uint64_t m[AU][W], old[AU][W];
#pragma unroll
                        for (int u = 0; u < AU; ++u) w[u] = j + u < j1 ? col[j + u] : -1;
#pragma unroll
                        for (int u = 0; u < AU; ++u)
#pragma unroll
                            for (int h = 0; h < W; ++h)
                                m[u][h] = w[u] >= 0 ? F[h] & ~ald(&vis[(int64_t)w[u] * W + h]) : 0;
#pragma unroll
                        for (int u = 0; u < AU; ++u)
#pragma unroll
                            for (int h = 0; h < W; ++h)
                                old[u][h] = m[u][h] ? atomicOr((unsigned long long *)&nxt[(int64_t)w[u] * W + h],
                                                               (unsigned long long)m[u][h])
                                                    : 1;
#pragma unroll
                        for (int u = 0; u < AU; ++u)
                            if (w[u] >= 0 && first_reach<W>(m[u], old[u], tag, w[u], l + 1)) {
                                const int pos = q1 + atomicAdd(&s_cnt, 1);
                                if (pos < a.qcap) Q[pos] = w[u];
                            }
                    }
                }
                uint64_t bal = __ballot(j1 - j0 > 16);
                while (bal) {
                    const int src = __ffsll((long long)bal) - 1;
                    bal &= bal - 1;
                    uint64_t FF[W];
#pragma unroll
                    for (int h = 0; h < W; ++h) FF[h] = wsh64(F[h], src);
                    const int a0 = __builtin_amdgcn_readlane(j0, src), a1 = __builtin_amdgcn_readlane(j1, src);
                    for (int j = a0 + lane; j < a1; j += 64) {
                        const int w = col[j];
                        uint64_t m[W], old[W];
#pragma unroll
                        for (int h = 0; h < W; ++h) m[h] = FF[h] & ~ald(&vis[(int64_t)w * W + h]);
#pragma unroll
                        for (int h = 0; h < W; ++h)
                            old[h] = m[h] ? atomicOr((unsigned long long *)&nxt[(int64_t)w * W + h],
                                                     (unsigned long long)m[h])
                                          : 1;
                        if (first_reach<W>(m, old, tag, w, l + 1)) {
                            const int pos = q1 + atomicAdd(&s_cnt, 1);
                            if (pos < a.qcap) Q[pos] = w;
                        }
                    }
                }
            }
            wg_barrier();
            // (qcap bounds a tile's list entries; past it the host reports an error)
            if (tid == 0 && (int64_t)q1 + s_cnt > a.qcap) atomicOr(&a.counter[2], 1);
            const int q2 = (int)min((int64_t)q1 + s_cnt, a.qcap);
            wg_barrier();
            if (tid == 0) {
                qoff[l + 2] = q2;
                s_cnt = 0;
            }
            if (q2 == q1) break;
            // pull: sigma of Q_{l+1}, GEB entries per wave block; a block's
            // list entries are loaded two blocks ahead, their masks and CSR
            // ranges one block ahead
            {
                constexpr int S = GW * GEB;
                int e0 = q1 + wv * GEB;
                int qn = (e0 + lane < q2 && lane < GEB) ? Q[e0 + lane] : 0;  // next block's vertices
                EntryMeta<W> em{};
                auto meta = [&](int e, int v) {
                    EntryMeta<W> m{};
                    if (e + lane < q2 && lane < GEB) {
                        m.v = v;
#pragma unroll
                        for (int h = 0; h < W; ++h) m.m[h] = ald(&nxt[(int64_t)v * W + h]);
                        m.r0 = rp[v];
                        m.r1 = rp[v + 1];
                    }
                    return m;
                };
                if (e0 < q2) {
                    em = meta(e0, qn);
                    qn = (e0 + S + lane < q2 && lane < GEB) ? Q[e0 + S + lane] : 0;
                }
                for (; e0 < q2; e0 += S) {
                    const EntryMeta<W> cur_em = em;
                    if (e0 + S < q2) {
                        em = meta(e0 + S, qn);
                        qn = (e0 + 2 * S + lane < q2 && lane < GEB) ? Q[e0 + 2 * S + lane] : 0;
                    }
                    pull_block<F64, W>(a.limit, col, cur, rcur, rnxt, Qm, Qr, sig, vis, e0, min(GEB, q2 - e0), cur_em,
                                       lds[wv], &s_ovf, &s_rec, plv, psg, l + 1);
                }
            }
            wg_barrier();
            for (int e = q0 + tid; e < q1; e += GB)
#pragma unroll
                for (int h = 0; h < W; ++h) cur[(int64_t)Q[e] * W + h] = 0;
            wg_barrier();
            uint64_t *t = cur;
            cur = nxt;
            nxt = t;
            int32_t *rt = rcur;
            rcur = rnxt;
            rnxt = rt;
            l++;
            q0 = q1;
            q1 = q2;
            if (s_ovf) break;  // uint64 overflow: the tile is rerun in FP64
        }
        const int D = l;  // deepest non-empty level
        wg_barrier();
        if (s_ovf) {
            if (tid == 0) {
                const int k = atomicAdd(&a.counter[1], 1);
                a.redo[k] = item;
            }
            wg_barrier();
            continue;
        }
        // ------------------------------------------------------ backward
        // cur holds level D's masks (nxt is all zero): clear them so both
        // arrays start at zero; each level scatters its own masks for the
        // level above
        for (int e = q0 + tid; e < q1; e += GB)
#pragma unroll
            for (int h = 0; h < W; ++h) cur[(int64_t)Q[e] * W + h] = 0;
        wg_barrier();
        uint64_t *succ = cur, *own = nxt;
        int32_t *rsucc = rcur, *rown = rnxt;
        for (int lb = D; lb >= 1; --lb) {
            const int b0 = qoff[lb], b1 = qoff[lb + 1];
            {
                // list entries two blocks ahead, CSR ranges one block ahead
                constexpr int S = GW * BEB;
                int e0 = b0 + wv * BEB;
                auto entry = [&](int e, EntryMeta<W> &m) {
                    m = EntryMeta<W>{};
                    if (e + lane < b1 && lane < BEB) {
                        m.v = Q[e + lane];
#pragma unroll
                        for (int h = 0; h < W; ++h) m.m[h] = Qm[(int64_t)(e + lane) * W + h];
                        m.rec = Qr[e + lane];
                    }
                };
                auto ranges = [&](int e, EntryMeta<W> &m) {
                    if (e + lane < b1 && lane < BEB) {
                        m.r0 = rp[m.v];
                        m.r1 = rp[m.v + 1];
                    }
                };
                EntryMeta<W> em{}, en{};
                if (e0 < b1) {
                    entry(e0, em);
                    ranges(e0, em);
                    entry(e0 + S, en);
                }
                for (; e0 < b1; e0 += S) {
                    const EntryMeta<W> cur_em = em;
                    if (e0 + S < b1) {
                        em = en;
                        ranges(e0 + S, em);
                        entry(e0 + 2 * S, en);
                    }
                    back_block<F64, TT, W>(col, cp.reach, cp.bc, a.ws + tl.src0, succ, rsucc, own, rown, sig, tt,
                                           min(BEB, b1 - e0), lb < D, cur_em, lds[wv], pdl);
                }
            }
            wg_barrier();
            // level lb+1's masks are done with; level lb's become the successors
            const int c0 = qoff[lb + 1], c1 = qoff[lb + 2];
            for (int e = c0 + tid; e < c1; e += GB)
#pragma unroll
                for (int h = 0; h < W; ++h) succ[(int64_t)Q[e] * W + h] = 0;
            wg_barrier();
            uint64_t *t = succ;
            succ = own;
            own = t;
            int32_t *rt = rsucc;
            rsucc = rown;
            rown = rt;
        }
        wg_barrier();
    }
}

// FP64-sigma tiles keep FP64 terms in either precision: 1/sigma can leave
// the f32 range.
template <int W>
static void wg_w(const SmallArgs &a, bool f64, bool fp32, int grid, hipStream_t st) {
    if (f64) {
        hipLaunchKernelGGL((k_wg<true, double, W>), dim3(grid), dim3(GB), 0, st, a);
    } else {
        if (fp32) hipLaunchKernelGGL((k_wg<false, float, W>), dim3(grid), dim3(GB), 0, st, a);
        else hipLaunchKernelGGL((k_wg<false, double, W>), dim3(grid), dim3(GB), 0, st, a);
    }
}

hipError_t launch_wg(const SmallArgs &a, bool f64, bool fp32, int grid, hipStream_t st) {
    if (a.spl == 4) wg_w<4>(a, f64, fp32, grid, st);
    else if (a.spl == 2) wg_w<2>(a, f64, fp32, grid, st);
    else wg_w<1>(a, f64, fp32, grid, st);
    return hipGetLastError();
}

}  // namespace bcx
