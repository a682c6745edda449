/*
 * bcx_oracle.c - CPU restatement of the reference's exact-BC algorithm.
 *
 * TEST INFRASTRUCTURE ONLY.  It computes the results the GPU path is compared
 * against: the tests, the smoke entry and the benchmark's CPU-baseline leg load
 * it; the product path never does.  Plain C99 + OpenMP, FP64 throughout.
 *
 * What each function restates (reference files, path relative to the
 * reference checkout):
 *   orc_source          one Brandes source on a component with reach weights:
 *                       levels and sigma of the forward BFS
 *                       (Codes/kernels_free.cu:245-362), sigma summed by pull
 *                       over the CSR neighbour order (uint64 with an overflow
 *                       flag, and an FP64 shadow summed in the same order),
 *                       dependency of the successor method
 *                       delta_w = sum_v sigma_w/sigma_v ((r(v)+1) + delta_v)
 *                       (Codes/kernels_free.cu:378-421).
 *   orc_sources         orc_source for a list of sources on host threads
 *                       (one call per source, so pinned as orc_source is).
 *   orc_chain_source    the chain-source combine of
 *                       Codes/postprocesses.cu:329-360 (min over the two
 *                       endpoint rows plus ld/rd, sigma summed on ties, then
 *                       the direct walk along the source's own chain,
 *                       Codes/postprocesses.cu:205-274), followed by the same
 *                       dependency pass (Codes/postprocesses.cu:416-466).
 *   orc_component_bc    sum over sources s of (r(s)+1) delta_s
 *                       (Codes/kernels_free.cu:423-426), all sources or a
 *                       subset, OpenMP over sources.
 *   orc_brandes_bc      plain whole-graph Brandes (no reach, no split): the
 *                       end-to-end oracle for the pipeline
 *                       (semantics of Codes/apsp_graph.cpp:54-275, BFS for
 *                       unit weights), ordered pairs.
 *   orc_cut_term_brute  (n-1)^2 - sum |C_i|^2 over the components of G - a,
 *                       by deleting a and counting (what
 *                       Codes/modified_apsp.cpp:157-220 computes with
 *                       2*cntx*cnty terms).
 *   orc_reach_brute     r_B(v): vertices reachable from v without entering
 *                       B's other vertices, minus v itself
 *                       (Codes/modified_apsp.cpp:184-207 semantics).
 *
 * How each is pinned (tests/test_oracle.py): orc_brandes_bc against
 * networkx 2x betweenness_centrality goldens, brute-force all-shortest-path
 * enumeration on tiny graphs and closed forms (path, star, cycle, complete
 * graph, the checksum sum_v BC(v) = sum_{s!=t} (d(s,t)-1)); orc_source sigma
 * against closed forms (grid binomials C(dx+dy,dx), complete graph, cycle)
 * and brute-force path counts; orc_component_bc + cut term against
 * orc_brandes_bc on whole graphs (a dropped reach weight or cut term fails);
 * orc_chain_source against a direct orc_source from the same vertex;
 * orc_cut_term_brute / orc_reach_brute against networkx component sizes.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_INF 0x7fffffff

/* BFS levels; queue out[n] receives vertices in BFS order; returns count. */
static int32_t bfs_levels(int32_t n, const int32_t *rp, const int32_t *col, int32_t s,
                          int32_t *lvl, int32_t *order) {
    for (int32_t i = 0; i < n; ++i) lvl[i] = ORC_INF;
    int32_t head = 0, tail = 0;
    lvl[s] = 0;
    order[tail++] = s;
    while (head < tail) {
        int32_t v = order[head++];
        for (int32_t j = rp[v]; j < rp[v + 1]; ++j) {
            int32_t w = col[j];
            if (lvl[w] == ORC_INF) {
                lvl[w] = lvl[v] + 1;
                order[tail++] = w;
            }
        }
    }
    return tail;
}

/* sigma by pull in CSR order over vertices sorted by level (order[] from BFS
 * is level-sorted).  Writes u64 (overflow flagged) and f64. */
static int sigma_pull(int32_t n, const int32_t *rp, const int32_t *col, int32_t s,
                      const int32_t *lvl, const int32_t *order, int32_t cnt,
                      uint64_t *su, double *sf) {
    int overflow = 0;
    for (int32_t i = 0; i < n; ++i) { su[i] = 0; sf[i] = 0.0; }
    su[s] = 1;
    sf[s] = 1.0;
    for (int32_t i = 1; i < cnt; ++i) {
        int32_t w = order[i];
        uint64_t a = 0;
        double f = 0.0;
        for (int32_t j = rp[w]; j < rp[w + 1]; ++j) {
            int32_t v = col[j];
            if (lvl[v] == lvl[w] - 1) {
                uint64_t b = a + su[v];
                if (b < a) overflow = 1;
                a = b;
                f += sf[v];
            }
        }
        su[w] = a;
        sf[w] = f;
    }
    return overflow;
}

/* Successor-method dependency with reach weights, FP64.  sig: sigma as
 * double (exact u64 converted, or the FP64 pull values). */
static void delta_pass(const int32_t *rp, const int32_t *col, const int32_t *reach,
                       const int32_t *lvl, const int32_t *order, int32_t cnt,
                       const double *sig, double *delta, int32_t n) {
    for (int32_t i = 0; i < n; ++i) delta[i] = 0.0;
    for (int32_t i = cnt; i-- > 1;) {
        int32_t w = order[i];
        double acc = 0.0;
        for (int32_t j = rp[w]; j < rp[w + 1]; ++j) {
            int32_t v = col[j];
            if (lvl[v] == lvl[w] + 1)
                acc += sig[w] / sig[v] * ((double)(reach ? reach[v] : 0) + 1.0 + delta[v]);
        }
        delta[w] = acc;
    }
}

/* One source: outputs may be NULL.  Returns 1 if sigma overflowed uint64. */
int orc_source(int32_t n, const int32_t *rp, const int32_t *col, const int32_t *reach,
               int32_t s, int32_t *level, uint64_t *sigma_u64, double *sigma_f64,
               double *delta) {
    int32_t *lvl = (int32_t *)malloc(sizeof(int32_t) * n);
    int32_t *order = (int32_t *)malloc(sizeof(int32_t) * n);
    uint64_t *su = (uint64_t *)malloc(sizeof(uint64_t) * n);
    double *sf = (double *)malloc(sizeof(double) * n);
    double *sig = (double *)malloc(sizeof(double) * n);
    double *dl = (double *)malloc(sizeof(double) * n);
    int32_t cnt = bfs_levels(n, rp, col, s, lvl, order);
    int ovf = sigma_pull(n, rp, col, s, lvl, order, cnt, su, sf);
    for (int32_t i = 0; i < n; ++i) sig[i] = ovf ? sf[i] : (double)su[i];
    delta_pass(rp, col, reach, lvl, order, cnt, sig, dl, n);
    if (level) memcpy(level, lvl, sizeof(int32_t) * n);
    if (sigma_u64) memcpy(sigma_u64, su, sizeof(uint64_t) * n);
    if (sigma_f64) memcpy(sigma_f64, sf, sizeof(double) * n);
    if (delta) memcpy(delta, dl, sizeof(double) * n);
    free(lvl); free(order); free(su); free(sf); free(sig); free(dl);
    return ovf;
}

/* orc_source for srcs[0..ns), OpenMP over sources: row i of each output
 * ([ns * n], any may be NULL) and ovf[i] belong to srcs[i].               */
void orc_sources(int32_t n, const int32_t *rp, const int32_t *col, const int32_t *reach,
                 const int32_t *srcs, int32_t ns, int32_t *level, uint64_t *sigma_u64,
                 double *sigma_f64, double *delta, int32_t *ovf, int32_t threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel for schedule(dynamic, 1)
    for (int32_t i = 0; i < ns; ++i) {
        const size_t o = (size_t)i * (size_t)n;
        int r = orc_source(n, rp, col, reach, srcs[i], level ? level + o : NULL,
                           sigma_u64 ? sigma_u64 + o : NULL, sigma_f64 ? sigma_f64 + o : NULL,
                           delta ? delta + o : NULL);
        if (ovf) ovf[i] = r;
    }
}

/* Chain source x = chain_nodes[pos] of a chain with endpoints L,R and cnt
 * vertices nodes[0..cnt) listed from L (ld = j+1, rd = cnt-j).  Levels and
 * sigma from the endpoint rows, then the direct walk along the chain; sigma
 * as FP64 (exact for the integer values the tests use) and u64. */
int orc_chain_source(int32_t n, const int32_t *rp, const int32_t *col, const int32_t *reach,
                     int32_t L, int32_t R, int32_t cnt, const int32_t *nodes, int32_t pos,
                     int32_t *level, uint64_t *sigma_u64, double *sigma_f64, double *delta) {
    int32_t *dL = (int32_t *)malloc(sizeof(int32_t) * n);
    int32_t *dR = (int32_t *)malloc(sizeof(int32_t) * n);
    uint64_t *sL = (uint64_t *)malloc(sizeof(uint64_t) * n);
    uint64_t *sR = (uint64_t *)malloc(sizeof(uint64_t) * n);
    double *fL = (double *)malloc(sizeof(double) * n);
    double *fR = (double *)malloc(sizeof(double) * n);
    int ovf = orc_source(n, rp, col, reach, L, dL, sL, fL, NULL);
    ovf |= orc_source(n, rp, col, reach, R, dR, sR, fR, NULL);
    int32_t ld = pos + 1, rd = cnt - pos;
    int32_t *lv = (int32_t *)malloc(sizeof(int32_t) * n);
    uint64_t *su = (uint64_t *)malloc(sizeof(uint64_t) * n);
    double *sf = (double *)malloc(sizeof(double) * n);
    /* Codes/postprocesses.cu:329-351 */
    for (int32_t k = 0; k < n; ++k) {
        int64_t a = (int64_t)dL[k] + ld, b = (int64_t)dR[k] + rd;
        if (a < b) { lv[k] = (int32_t)a; su[k] = sL[k]; sf[k] = fL[k]; }
        else if (a > b) { lv[k] = (int32_t)b; su[k] = sR[k]; sf[k] = fR[k]; }
        else { lv[k] = (int32_t)b; su[k] = sL[k] + sR[k]; if (su[k] < sL[k]) ovf = 1; sf[k] = fL[k] + fR[k]; }
    }
    /* Codes/postprocesses.cu:354-360 + 205-274: own chain, direct distances */
    int32_t x = nodes[pos];
    lv[x] = 0; su[x] = 1; sf[x] = 1.0;
    for (int32_t j = 0; j < cnt; ++j) {
        if (j == pos) continue;
        int32_t y = nodes[j];
        int32_t dist = j > pos ? j - pos : pos - j;
        if (lv[y] > dist) { lv[y] = dist; su[y] = 1; sf[y] = 1.0; }
        else if (lv[y] == dist) { su[y] += 1; sf[y] += 1.0; }
    }
    /* dependency over the level order (counting sort by level) */
    int32_t maxl = 0;
    for (int32_t k = 0; k < n; ++k) if (lv[k] > maxl) maxl = lv[k];
    int32_t *off = (int32_t *)calloc((size_t)maxl + 2, sizeof(int32_t));
    int32_t *order = (int32_t *)malloc(sizeof(int32_t) * n);
    for (int32_t k = 0; k < n; ++k) off[lv[k] + 1]++;
    for (int32_t l = 0; l <= maxl; ++l) off[l + 1] += off[l];
    for (int32_t k = 0; k < n; ++k) order[off[lv[k]]++] = k;
    double *sig = (double *)malloc(sizeof(double) * n);
    double *dl = (double *)malloc(sizeof(double) * n);
    for (int32_t k = 0; k < n; ++k) sig[k] = ovf ? sf[k] : (double)su[k];
    delta_pass(rp, col, reach, lv, order, n, sig, dl, n);
    if (level) memcpy(level, lv, sizeof(int32_t) * n);
    if (sigma_u64) memcpy(sigma_u64, su, sizeof(uint64_t) * n);
    if (sigma_f64) memcpy(sigma_f64, sf, sizeof(double) * n);
    if (delta) memcpy(delta, dl, sizeof(double) * n);
    free(dL); free(dR); free(sL); free(sR); free(fL); free(fR);
    free(lv); free(su); free(sf); free(off); free(order); free(sig); free(dl);
    return ovf;
}

/* Weighted component BC over all sources (srcs == NULL) or a subset.
 * bc[n] is overwritten.  threads <= 0: OpenMP default. */
void orc_component_bc(int32_t n, const int32_t *rp, const int32_t *col, const int32_t *reach,
                      const int32_t *srcs, int64_t n_src, double *bc, int32_t threads) {
    int64_t ns = srcs ? n_src : n;
    for (int32_t i = 0; i < n; ++i) bc[i] = 0.0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel
    {
        int32_t *lvl = (int32_t *)malloc(sizeof(int32_t) * n);
        int32_t *order = (int32_t *)malloc(sizeof(int32_t) * n);
        uint64_t *su = (uint64_t *)malloc(sizeof(uint64_t) * n);
        double *sf = (double *)malloc(sizeof(double) * n);
        double *sig = (double *)malloc(sizeof(double) * n);
        double *dl = (double *)malloc(sizeof(double) * n);
        double *loc = (double *)calloc((size_t)n, sizeof(double));
#pragma omp for schedule(dynamic, 1)
        for (int64_t i = 0; i < ns; ++i) {
            int32_t s = srcs ? srcs[i] : (int32_t)i;
            int32_t cnt = bfs_levels(n, rp, col, s, lvl, order);
            int ovf = sigma_pull(n, rp, col, s, lvl, order, cnt, su, sf);
            for (int32_t k = 0; k < cnt; ++k) {
                int32_t v = order[k];
                sig[v] = ovf ? sf[v] : (double)su[v];
            }
            delta_pass(rp, col, reach, lvl, order, cnt, sig, dl, n);
            double ws = (double)(reach ? reach[s] : 0) + 1.0;
            for (int32_t k = 1; k < cnt; ++k) loc[order[k]] += ws * dl[order[k]];
        }
#pragma omp critical
        for (int32_t k = 0; k < n; ++k) bc[k] += loc[k];
        free(lvl); free(order); free(su); free(sf); free(sig); free(dl); free(loc);
    }
}

/* Plain Brandes on a whole graph (reach 0 everywhere, every vertex a source). */
void orc_brandes_bc(int32_t n, const int32_t *rp, const int32_t *col, double *bc, int32_t threads) {
    orc_component_bc(n, rp, col, NULL, NULL, 0, bc, threads);
}

/* Components of G - a by BFS; rp/col is the whole graph. */
void orc_cut_term_brute(int32_t n, const int32_t *rp, const int32_t *col, int64_t *cut) {
    int32_t *comp = (int32_t *)malloc(sizeof(int32_t) * n);
    int32_t *q = (int32_t *)malloc(sizeof(int32_t) * n);
    for (int32_t a = 0; a < n; ++a) {
        for (int32_t i = 0; i < n; ++i) comp[i] = -1;
        comp[a] = -2;
        int64_t sumsq = 0;
        for (int32_t r = 0; r < n; ++r) {
            if (comp[r] != -1) continue;
            int32_t head = 0, tail = 0;
            comp[r] = r;
            q[tail++] = r;
            while (head < tail) {
                int32_t v = q[head++];
                for (int32_t j = rp[v]; j < rp[v + 1]; ++j)
                    if (comp[col[j]] == -1) { comp[col[j]] = r; q[tail++] = col[j]; }
            }
            sumsq += (int64_t)tail * tail;
        }
        cut[a] = (int64_t)(n - 1) * (n - 1) - sumsq;
    }
    free(comp);
    free(q);
}

/* r_B(v) for every v in block B (in_block[n] flags): BFS from v in G that
 * never enters another vertex of B. */
void orc_reach_brute(int32_t n, const int32_t *rp, const int32_t *col, const uint8_t *in_block,
                     int32_t *reach) {
    int32_t *seen = (int32_t *)malloc(sizeof(int32_t) * n);
    int32_t *q = (int32_t *)malloc(sizeof(int32_t) * n);
    for (int32_t i = 0; i < n; ++i) seen[i] = -1;
    for (int32_t v = 0; v < n; ++v) {
        if (!in_block[v]) { reach[v] = 0; continue; }
        int32_t head = 0, tail = 0;
        seen[v] = v;
        q[tail++] = v;
        while (head < tail) {
            int32_t x = q[head++];
            for (int32_t j = rp[x]; j < rp[x + 1]; ++j) {
                int32_t y = col[j];
                if (seen[y] == v || in_block[y]) continue;
                seen[y] = v;
                q[tail++] = y;
            }
        }
        reach[v] = tail - 1;
    }
    free(seen);
    free(q);
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
