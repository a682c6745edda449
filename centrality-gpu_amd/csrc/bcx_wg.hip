// Workgroup-per-tile kernel for components of up to WG_N vertices with
// narrow levels (deep BFS: grids, meshes, long paths).  The multi-launch
// tiles pay three launches per level, which a BFS hundreds of levels deep
// turns into milliseconds of launch latency per tile; here one workgroup
// takes a tile = (component, up to 64 sources) from a work counter and runs
// every level of it with workgroup barriers, while hundreds of workgroups run
// other tiles.
//
// Same method as the batched kernels (lane k of a wave64 owns source k): the
// visited set and two level masks (u64 per vertex), the level lists and
// their offsets live in a per-workgroup HBM slab.  Path counts and terms are
// lane-compacted records: list entry (v, lanes M) owns popcount(M)
// consecutive words, lane k at rank popcount(M & (2^k - 1)), from a record
// offset taken when the entry's sigma is pulled; each vertex's offset sits
// beside its level mask.  On deep graphs an entry holds ~5 of the 64 lanes,
// so a record is one cache line where a [v][64] row of FP64 words is four,
// and the records of one level are contiguous.  Level l+1's list is appended
// by the first lane mask that reaches a vertex (the atomic OR that finds
// zero), so no pass over all n vertices is made per level.  Pull and backward
// walk the arcs of 8-16 list entries per wave as one flat range (lane per
// arc), compact the arcs that matter in LDS and then gather one record word
// per (arc, lane), summed per entry in CSR order (exact u64, FP64 in the
// oracle's order).  The backward is the rho
// form in FP32 with Kahan compensation or FP64 (always FP64 with FP64
// sigma), BC(v) += sum_k (r(s_k)+1) delta_k(v)
// (Codes/kernels_free.cu:154-473 computes the same, one source per block).
// A tile whose uint64 sigma overflows stops, adds nothing and is listed for
// an FP64 rerun by the host.
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
constexpr int BEB = 8;       // list entries per wave block (backward)
constexpr int GWIN = 128;    // arcs per compaction window (two 64-arc chunks)
constexpr int GUNR = 8;      // row gathers in flight per lane (64-bit rows)

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
// arc A, vertex V, lane mask M, record R), the compacted arcs of a window
// (far end's lanes that matter x, its full level mask fm and record rr,
// entry k) and the backward's per-entry sums.
struct WaveLds {
    int P[GEB], A[GEB], V[GEB], R[GEB];
    uint64_t M[GEB];
    uint64_t x[GWIN], fm[GWIN];
    int rr[GWIN];
    int k[GWIN];
    double r[BEB][64];
};

// Lane-compacted records: a list entry (vertex v, lanes M) keeps its path
// counts and terms in popcount(M) consecutive words from its record offset,
// lane k of M at rank popcount(M & (2^k - 1)).  The records of one level are
// contiguous, so the gathers of the next level read a small, cache-resident
// range; a dense [v][64] row would cost a 512-byte row (FP64) per gather for
// the ~5 lanes a deep graph's entry holds.
__device__ __forceinline__ int rank_in(uint64_t M, uint64_t me) { return __popcll(M & (me - 1)); }

__device__ __forceinline__ uint64_t wsh64(uint64_t x, int src) { return (uint64_t)__shfl((long long)x, src, 64); }

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
__device__ __forceinline__ int block_layout(int d, int cnt, int r0, int v, uint64_t m, int rec, WaveLds &L) {
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
        L.M[lane] = m;
        L.R[lane] = rec;
    }
    __builtin_amdgcn_wave_barrier();
    return __shfl(inc, 63, 64);
}

// Entry of flat arc index g: the last j with P[j] <= g (binary search).
__device__ __forceinline__ int entry_of(const WaveLds &L, int g) {
    int k = 0;
#pragma unroll
    for (int h = GEB / 2; h > 0; h >>= 1)
        if (L.P[k + h] <= g) k += h;
    return k;
}

// Compacts the arcs [win, wend) of a block whose far-end mask & entry mask is
// non-zero into L, in arc order, with the far end's full mask and record
// offset (rfar: plain stores of this workgroup's waves, so ordinary loads);
// returns the number of compacted arcs.  Both 64-arc chunks of the window
// load their arcs and masks before either is compacted.
__device__ __forceinline__ int compact(const int32_t *col, const uint64_t *far, const int32_t *rfar, int win,
                                       int wend, WaveLds &L) {
    const int lane = threadIdx.x & 63;
    const uint64_t me = 1ull << lane;
    const int g0 = win + lane, g1 = win + 64 + lane;
    const int k0 = entry_of(L, g0), k1 = entry_of(L, g1);
    int w0 = 0, w1 = 0;
    uint64_t f0 = 0, f1 = 0;
    if (g0 < wend) w0 = col[L.A[k0] + g0 - L.P[k0]];
    if (g1 < wend) w1 = col[L.A[k1] + g1 - L.P[k1]];
    if (g0 < wend) f0 = ald(&far[w0]);
    if (g1 < wend) f1 = ald(&far[w1]);
    const uint64_t x0 = f0 & L.M[k0], x1 = f1 & L.M[k1];
    const uint64_t b0 = __ballot(x0 != 0), b1 = __ballot(x1 != 0);
    if (x0) {
        const int pos = __popcll(b0 & (me - 1));
        L.x[pos] = x0;
        L.fm[pos] = f0;
        L.rr[pos] = rfar[w0];
        L.k[pos] = k0;
    }
    if (x1) {
        const int pos = __popcll(b0) + __popcll(b1 & (me - 1));
        L.x[pos] = x1;
        L.fm[pos] = f1;
        L.rr[pos] = rfar[w1];
        L.k[pos] = k1;
    }
    __builtin_amdgcn_wave_barrier();
    return __popcll(b0) + __popcll(b1);
}

// Metadata of a block's entry on lane j < cnt (loaded a block ahead): vertex,
// CSR range, lane mask and (backward) record offset.
struct EntryMeta {
    int v, r0, r1, rec;
    uint64_t m;
};

// Pull of list entries [e0, e0+cnt) of level l+1 (lane j: entry j's vertex,
// final lane mask and CSR range): records of the block taken from the tile's
// record counter (LDS), sigma over predecessors in CSR order, visited
// update.  Each entry is finished as soon as the walk moves past it (every
// entry has at least one predecessor arc for each of its lanes).
template <bool F64>
__device__ __forceinline__ void pull_block(uint64_t limit, const int32_t *col, const uint64_t *cur,
                                           const int32_t *rcur, int32_t *rnxt, uint64_t *Qm, int32_t *Qr,
                                           uint64_t *sig, uint64_t *vis, int e0, int cnt, EntryMeta em, WaveLds &L,
                                           int *s_ovf, int *s_rec, int32_t *plv, uint64_t *psig, int level) {
    const int lane = threadIdx.x & 63;
    const uint64_t me = 1ull << lane;
    const int w0 = em.v, r0 = em.r0, r1 = em.r1;
    const uint64_t N0 = em.m;
    // record offsets: exclusive prefix of the entries' lane counts
    const int pc = lane < cnt ? __popcll(N0) : 0;
    int inc = pc;
#pragma unroll
    for (int o = 1; o < GEB; o <<= 1) {
        const int y = __shfl_up(inc, o, 64);
        if (lane >= o) inc += y;
    }
    const int total = __shfl(inc, GEB - 1, 64);  // (lanes >= cnt add 0)
    int base = 0;
    if (lane == 0) base = atomicAdd(s_rec, total);
    const int rec = __shfl(base, 0, 64) + inc - pc;
    if (lane < cnt) {
        Qm[e0 + lane] = N0;
        Qr[e0 + lane] = rec;
        rnxt[w0] = rec;
    }
    const int T = block_layout(lane < cnt ? r1 - r0 : 0, cnt, r0, w0, N0, rec, L);
    uint64_t acc = 0;
    double facc = 0.0;
    bool ovf = false;
    int kcur = -1;
    auto finish = [&](int j) {
        const int w = L.V[j];
        const uint64_t N = L.M[j];
        const uint64_t val = F64 ? d2u(facc) : acc;
        if (N & me) {
            sig[L.R[j] + rank_in(N, me)] = val;
            if (!F64 && limit && val >= limit) ovf = true;
            if (plv) plv[(int64_t)w * 64 + lane] = level;
            if (psig) psig[(int64_t)w * 64 + lane] = val;
        }
        if (lane == 0) atomicOr((unsigned long long *)&vis[w], (unsigned long long)N);  // result unused: no return
        acc = 0;
        facc = 0.0;
    };
    for (int win = 0; win < T; win += GWIN) {
        const int nrel = compact(col, cur, rcur, win, min(T, win + GWIN), L);
        for (int r = 0; r < nrel; r += GUNR) {
            uint64_t xv[GUNR];
            int kk[GUNR];
            bool use[GUNR];
#pragma unroll
            for (int u = 0; u < GUNR; ++u) {
                const int rr = r + u;
                use[u] = false;
                kk[u] = -1;
                xv[u] = 0;
                if (rr < nrel) {
                    kk[u] = L.k[rr];
                    if (L.x[rr] & me) {
                        use[u] = true;
                        xv[u] = sig[L.rr[rr] + rank_in(L.fm[rr], me)];
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < GUNR; ++u) {
                if (kk[u] < 0) continue;
                if (kk[u] != kcur) {  // wave-uniform
                    if (kcur >= 0) finish(kcur);
                    while (++kcur < kk[u]) finish(kcur);  // (entries without arcs: none in a pull)
                }
                if (!use[u]) continue;
                if (F64) {
                    facc += u2d(xv[u]);
                } else {
                    const uint64_t s = acc + xv[u];
                    ovf |= (s < acc);
                    acc = s;
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
// entry j's vertex, lane mask and CSR range): scatter own masks, sum
// successor terms per entry into LDS (has_succ false at the deepest level),
// then load the block's sigma words together and finish.
template <bool F64, typename TT>
__device__ __forceinline__ void back_block(const int32_t *col, const int32_t *reach, double *bc, const double *ws,
                                           const uint64_t *succ, const int32_t *rsucc, uint64_t *own, int32_t *rown,
                                           const uint64_t *sig, TT *tt, int cnt, bool has_succ, EntryMeta em,
                                           WaveLds &L, double *pdl) {
    using Acc = typename std::conditional<sizeof(TT) == 4, KahanW, SumW>::type;
    constexpr int BUNR = sizeof(TT) == 8 ? GUNR : 2 * GUNR;  // 32-bit words: twice as many in flight
    const int lane = threadIdx.x & 63;
    const uint64_t me = 1ull << lane;
    const int v0 = em.v, r0 = em.r0, r1 = em.r1;
    const uint64_t M0 = em.m;
    if (lane < cnt) {
        own[v0] = M0;
        rown[v0] = em.rec;
    }
    const int T = block_layout(has_succ && lane < cnt ? r1 - r0 : 0, cnt, r0, v0, M0, em.rec, L);
#pragma unroll
    for (int j = 0; j < BEB; ++j) L.r[j][lane] = 0.0;
    Acc acc;
    int kcur = -1;
    for (int win = 0; win < T; win += GWIN) {
        const int nrel = compact(col, succ, rsucc, win, min(T, win + GWIN), L);
        for (int r = 0; r < nrel; r += BUNR) {
            TT t[BUNR];
            int kk[BUNR];
            bool use[BUNR];
#pragma unroll
            for (int u = 0; u < BUNR; ++u) {
                const int rr = r + u;
                use[u] = false;
                kk[u] = -1;
                t[u] = 0;
                if (rr < nrel) {
                    kk[u] = L.k[rr];
                    if (L.x[rr] & me) {
                        use[u] = true;
                        t[u] = tt[L.rr[rr] + rank_in(L.fm[rr], me)];
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < BUNR; ++u) {
                if (kk[u] < 0) continue;
                if (kk[u] != kcur) {
                    if (kcur >= 0) L.r[kcur][lane] = acc.value();
                    acc = Acc();
                    kcur = kk[u];
                }
                if (use[u]) acc.add(t[u]);
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    if (kcur >= 0) L.r[kcur][lane] = acc.value();
    __builtin_amdgcn_wave_barrier();
    uint64_t sx[BEB];
#pragma unroll
    for (int j = 0; j < BEB; ++j) sx[j] = (j < cnt && (L.M[j] & me)) ? sig[L.R[j] + rank_in(L.M[j], me)] : 0;
#pragma unroll
    for (int j = 0; j < BEB; ++j) {
        if (j >= cnt) break;
        const int v = L.V[j];
        double contrib = 0.0;
        if (L.M[j] & me) {
            const double sv = F64 ? u2d(sx[j]) : (double)sx[j];
            const double rho = L.r[j][lane];
            const double delta = sv * rho;
            contrib = ws[lane] * delta;
            tt[L.R[j] + rank_in(L.M[j], me)] = (TT)(((double)reach[v] + 1.0) / sv + rho);
            if (pdl) pdl[(int64_t)v * 64 + lane] = delta;
        }
        contrib = wave_sum(contrib);
        if (lane == 0 && contrib != 0.0) unsafeAtomicAdd(&bc[v], contrib);
    }
}

template <bool F64, typename TT>
__global__ __launch_bounds__(GB) void k_wg(SmallArgs a) {
    __shared__ WaveLds lds[GW];
    __shared__ int s_item, s_cnt, s_ovf, s_rec;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int64_t nm = a.nmax;
    uint64_t *sig = a.sig + (int64_t)blockIdx.x * a.slab;
    TT *tt = (TT *)a.tt + (int64_t)blockIdx.x * a.slab;
    int32_t *Q = a.Q + (int64_t)blockIdx.x * a.qcap;
    uint64_t *Qm = a.Qm + (int64_t)blockIdx.x * a.qcap;
    uint64_t *vis = a.masks + (int64_t)blockIdx.x * 3 * nm;
    uint64_t *mA = vis + nm, *mB = mA + nm;
    int32_t *qoff = a.qoffs + (int64_t)blockIdx.x * (nm + 2);
    int32_t *Qr = a.qr + (int64_t)blockIdx.x * a.qcap;        // record offset of each list entry
    int32_t *rA = a.ro + (int64_t)blockIdx.x * 2 * nm, *rB = rA + nm;  // of each vertex, with mA / mB
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
        for (int v = tid; v < n; v += GB) {
            vis[v] = 0;
            mA[v] = 0;
            mB[v] = 0;
        }
        if (tid == 0) {
            s_cnt = 0;
            s_ovf = 0;
            s_rec = tl.ns;  // records 0 .. ns-1: the sources, one lane each
            qoff[0] = 0;
            qoff[1] = tl.ns;
        }
        wg_barrier();
        if (tid < tl.ns) {
            const int s = a.srcs[tl.src0 + tid];
            vis[s] = 1ull << tid;
            mA[s] = 1ull << tid;
            rA[s] = tid;
            sig[tid] = F64 ? d2u(1.0) : 1ull;
            Q[tid] = s;
            Qm[tid] = 1ull << tid;
            Qr[tid] = tid;
            if (plv) plv[(int64_t)s * 64 + tid] = 0;
            if (psg) psg[(int64_t)s * 64 + tid] = F64 ? d2u(1.0) : 1ull;
        }
        wg_barrier();
        uint64_t *cur = mA, *nxt = mB;
        int32_t *rcur = rA, *rnxt = rB;
        int l = 0, q0 = 0, q1 = tl.ns;
        // ------------------------------------------------------- forward
        while (true) {
            // expand: lane per entry, a whole wave for an entry of degree > 16;
            // the OR that finds a zero word appends the vertex to Q_{l+1}
            for (int base = q0 + wv * 64; base < q1; base += GB) {
                const int e = base + lane;
                int j0 = 0, j1 = 0;
                uint64_t F = 0;
                if (e < q1) {
                    const int v = Q[e];
                    F = Qm[e];
                    j0 = rp[v];
                    j1 = rp[v + 1];
                }
                if (j1 - j0 <= 16) {
                    // four arcs at a time: their loads and ORs are independent
                    for (int j = j0; j < j1; j += 4) {
                        int w[4];
                        uint64_t m[4], old[4];
#pragma unroll
                        for (int u = 0; u < 4; ++u) w[u] = j + u < j1 ? col[j + u] : -1;
#pragma unroll
                        for (int u = 0; u < 4; ++u) m[u] = w[u] >= 0 ? F & ~ald(&vis[w[u]]) : 0;
#pragma unroll
                        for (int u = 0; u < 4; ++u)
                            old[u] = m[u] ? atomicOr((unsigned long long *)&nxt[w[u]], (unsigned long long)m[u]) : 1;
#pragma unroll
                        for (int u = 0; u < 4; ++u)
                            if (old[u] == 0) {
                                const int pos = q1 + atomicAdd(&s_cnt, 1);
                                if (pos < a.qcap) Q[pos] = w[u];
                            }
                    }
                }
                uint64_t bal = __ballot(j1 - j0 > 16);
                while (bal) {
                    const int src = __ffsll((long long)bal) - 1;
                    bal &= bal - 1;
                    const uint64_t FF = wsh64(F, src);
                    const int a0 = __shfl(j0, src, 64), a1 = __shfl(j1, src, 64);
                    for (int j = a0 + lane; j < a1; j += 64) {
                        const int w = col[j];
                        const uint64_t m = FF & ~ald(&vis[w]);
                        if (m && atomicOr((unsigned long long *)&nxt[w], (unsigned long long)m) == 0) {
                            const int pos = q1 + atomicAdd(&s_cnt, 1);
                            if (pos < a.qcap) Q[pos] = w;
                        }
                    }
                }
            }
            wg_barrier();
            const int q2 = min((int64_t)q1 + s_cnt, a.qcap);
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
                EntryMeta em{};
                auto meta = [&](int e, int v) {
                    EntryMeta m{};
                    if (e + lane < q2 && lane < GEB) {
                        m.v = v;
                        m.m = ald(&nxt[v]);
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
                    const EntryMeta cur_em = em;
                    if (e0 + S < q2) {
                        em = meta(e0 + S, qn);
                        qn = (e0 + 2 * S + lane < q2 && lane < GEB) ? Q[e0 + 2 * S + lane] : 0;
                    }
                    pull_block<F64>(a.limit, col, cur, rcur, rnxt, Qm, Qr, sig, vis, e0, min(GEB, q2 - e0), cur_em,
                                    lds[wv], &s_ovf, &s_rec, plv, psg, l + 1);
                }
            }
            wg_barrier();
            for (int e = q0 + tid; e < q1; e += GB) cur[Q[e]] = 0;
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
        for (int e = q0 + tid; e < q1; e += GB) cur[Q[e]] = 0;
        wg_barrier();
        uint64_t *succ = cur, *own = nxt;
        int32_t *rsucc = rcur, *rown = rnxt;
        for (int lb = D; lb >= 1; --lb) {
            const int b0 = qoff[lb], b1 = qoff[lb + 1];
            {
                // list entries two blocks ahead, CSR ranges one block ahead
                constexpr int S = GW * BEB;
                int e0 = b0 + wv * BEB;
                auto entry = [&](int e, EntryMeta &m) {
                    m = EntryMeta{};
                    if (e + lane < b1 && lane < BEB) {
                        m.v = Q[e + lane];
                        m.m = Qm[e + lane];
                        m.rec = Qr[e + lane];
                    }
                };
                auto ranges = [&](int e, EntryMeta &m) {
                    if (e + lane < b1 && lane < BEB) {
                        m.r0 = rp[m.v];
                        m.r1 = rp[m.v + 1];
                    }
                };
                EntryMeta em{}, en{};
                if (e0 < b1) {
                    entry(e0, em);
                    ranges(e0, em);
                    entry(e0 + S, en);
                }
                for (; e0 < b1; e0 += S) {
                    const EntryMeta cur_em = em;
                    if (e0 + S < b1) {
                        em = en;
                        ranges(e0 + S, em);
                        entry(e0 + 2 * S, en);
                    }
                    back_block<F64, TT>(col, cp.reach, cp.bc, a.ws + tl.src0, succ, rsucc, own, rown, sig, tt,
                                        min(BEB, b1 - e0), lb < D, cur_em, lds[wv], pdl);
                }
            }
            wg_barrier();
            // level lb+1's masks are done with; level lb's become the successors
            const int c0 = qoff[lb + 1], c1 = qoff[lb + 2];
            for (int e = c0 + tid; e < c1; e += GB) succ[Q[e]] = 0;
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
hipError_t launch_wg(const SmallArgs &a, bool f64, bool fp32, int grid, hipStream_t st) {
    if (f64) {
        hipLaunchKernelGGL((k_wg<true, double>), dim3(grid), dim3(GB), 0, st, a);
    } else {
        if (fp32) hipLaunchKernelGGL((k_wg<false, float>), dim3(grid), dim3(GB), 0, st, a);
        else hipLaunchKernelGGL((k_wg<false, double>), dim3(grid), dim3(GB), 0, st, a);
    }
    return hipGetLastError();
}

}  // namespace bcx
