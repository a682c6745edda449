/*
 * bcx.h - C-ABI of the MI355X exact betweenness-centrality (BC) hot path.
 *
 * Two halves, both plain C (no torch or STL types):
 *
 *  1. Host structure (bcx_host_*): edge-list reader, biconnected-component
 *     split in Tarjan order, per-component local relabel, block-cut-tree reach
 *     weights and cut-vertex term, degree-2 chain extraction.  These replace
 *     the host C++ of the reference's L4 layer.
 *
 *  2. GPU engine (bcx_create / bcx_upload / bcx_run / bcx_bc ...): per-source
 *     Brandes forward (levels, sigma) and backward (dependency) kernels on
 *     gfx950, chain-source combine, weighted BC accumulation.  These replace
 *     the reference's host->GPU boundary begin_gpu / bc_gpu_free /
 *     bc_gpu_active / postprocess.
 *
 * Every entry point returns a status code (BCX_OK == 0) and never exits the
 * process; bcx_last_error() / bcx_host_last_error() give the message.
 *
 * BC convention: ordered pairs, not halved, i.e.
 *   BC(v) = sum_{s != v != t} sigma_st(v) / sigma_st
 *         = 2 x networkx.betweenness_centrality(G, normalized=False)
 * (reference: Codes/kernels_free.cu:423-426, Codes/modified_apsp.cpp:209).
 */
#ifndef BCX_H_
#define BCX_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define BCX_OK             0
#define BCX_E_ARG         -1  /* bad argument / shape                        */
#define BCX_E_IO          -2  /* input file missing or malformed             */
#define BCX_E_GRAPH       -3  /* self loop, duplicate edge, disconnected ...  */
#define BCX_E_NOMEM       -4  /* device memory cap exceeded                   */
#define BCX_E_HIP         -5  /* HIP runtime error                            */
#define BCX_E_NODEV       -6  /* no GPU / library built without a device      */
#define BCX_E_OVERFLOW    -7  /* sigma overflow that could not be recovered   */

/* ======================================================================= */
/* 1. Host structure                                                        */
/* ======================================================================= */

typedef struct bcx_host bcx_host;

/* One biconnected component with more than two vertices, in local ids.
 * Local ids are assigned in first-appearance order over the component's edge
 * list in Tarjan pop order (Codes/modified_apsp.cpp:336-357); the adjacency
 * of every local vertex lists its arcs in the order (u->v),(v->u) of that edge
 * list (Codes/to_eargraph_construct.cpp:81-96, Codes/graph_ear_decompos.h:65-83).
 *
 * Chains are the maximal paths of component-degree-2 vertices between two
 * component-degree>=3 endpoints (what Codes/bcc_bfs.cpp:100-236 emits as
 * ear_premble{lnode, rnode, cnt, ld[], rd[], nodes[]}, Codes/ear_graph.h:3-32).
 * Chain i has endpoints chain_ends[2i] (= L) and chain_ends[2i+1] (= R) and
 * vertices chain_nodes[chain_ptr[i] .. chain_ptr[i+1]) listed from the L side;
 * the j-th of them (0-based) has ld = j+1 and rd = cnt-j.                   */
typedef struct {
    int32_t        n;            /* local vertices                           */
    int64_t        arcs;         /* 2 x undirected edges                     */
    const int32_t *row_ptr;      /* [n+1] (arcs < 2^31)                      */
    const int32_t *col;          /* [arcs]                                   */
    const int32_t *reach;        /* [n]  r_B(v) >= 0                         */
    int32_t        n_chains;     /* 0 = no chain reduction                   */
    const int32_t *chain_ptr;    /* [n_chains+1]                             */
    const int32_t *chain_nodes;  /* [chain_ptr[n_chains]]                    */
    const int32_t *chain_ends;   /* [2*n_chains]                             */
} bcx_component;

/* Graph::Graph(char*) (Codes/graph.cpp:38-72): "n m" header then m lines
 * "u v w", 0-based, w == 1.  Adjacency keeps file order.                  */
int  bcx_host_from_file(const char *path, bcx_host **out);
/* Same graph from an in-memory edge list uv[2*m] (file order).            */
int  bcx_host_from_edges(int32_t n, int64_t m, const int32_t *uv, bcx_host **out);
void bcx_host_free(bcx_host *h);
const char *bcx_host_last_error(const bcx_host *h);

int32_t bcx_host_num_vertices(const bcx_host *h);
int64_t bcx_host_num_edges(const bcx_host *h);

/* Run the structural reduction: Tarjan BCC split (Codes/bicc.cpp:39-161),
 * local relabel (Codes/modified_apsp.cpp:266-363), block-cut tree reach and
 * cut term (Codes/modified_apsp.cpp:77-220, restated as a closed form),
 * chains (Codes/bcc_bfs.cpp:100-236, 928-1311) when use_chains != 0.
 * check_input != 0 rejects self loops, duplicate edges and disconnected
 * input with BCX_E_GRAPH.                                                  */
int  bcx_host_decompose(bcx_host *h, int use_chains, int check_input);

/* Number of BCCs of any size, in Tarjan emission order (" n_bcc= ").      */
int32_t bcx_host_num_bcc(const bcx_host *h);
/* Size of BCC i: vertices and undirected edges.                            */
int  bcx_host_bcc_size(const bcx_host *h, int32_t i, int32_t *n, int64_t *edges);
/* Edge list of BCC i in pop order, global ids, uv[2*edges].                */
int  bcx_host_bcc_edges(const bcx_host *h, int32_t i, int32_t *uv);
/* Articulation points: flags[n] (1 = articulation point).                  */
int  bcx_host_articulation(const bcx_host *h, uint8_t *flags);
/* Cut-vertex term (ordered pairs separated by v), int64[n].                */
int  bcx_host_cut_term(const bcx_host *h, int64_t *cut);

/* Components handed to the GPU: the BCCs with more than two vertices, in
 * Tarjan order (Codes/to_eargraph_construct.cpp:54-60).                    */
int32_t bcx_host_num_components(const bcx_host *h);
/* Views into host-owned arrays; valid until bcx_host_free.                 */
int  bcx_host_component(const bcx_host *h, int32_t c, bcx_component *out,
                        const int32_t **global_ids, int32_t *bcc_index);
/* Per-component counts as printed by the reference: free sources (not chain
 * vertex, not chain endpoint) and distinct chain endpoints.                */
int  bcx_host_component_counts(const bcx_host *h, int32_t c,
                               int32_t *n_free, int32_t *n_endpoints,
                               int32_t *n_chain_vertices);

/* Structure cache (no counterpart in the reference, whose text parse and
 * reduction rerun on every call; this is for repeated runs on one input):
 * the reduced host in a binary file.  bcx_host_save needs a decomposed host
 * and records source_path's size and modification time (none when NULL);
 * the file is written beside `path` and renamed into place.
 * bcx_host_load returns BCX_E_IO when the file is damaged or of another
 * format version, when source_path is given and its size or modification
 * time differ from the recorded ones, or when use_chains differs from the
 * saved decomposition: the caller then reads and reduces the text again.
 * A loaded host answers every query except bcx_host_bcc_edges (BCX_E_ARG:
 * edge lists are not stored); bcx_host_decompose on it is a no-op for the
 * same use_chains.                                                         */
int  bcx_host_save(const bcx_host *h, const char *path, const char *source_path);
int  bcx_host_load(const char *path, const char *source_path, int use_chains, bcx_host **out);

/* ======================================================================= */
/* 2. GPU engine                                                            */
/* ======================================================================= */

typedef struct bcx_ctx    bcx_ctx;
typedef struct bcx_dgraph bcx_dgraph;   /* device-resident component set */

/* Precision of the dependency (delta) arithmetic.                          */
#define BCX_PREC_FP64        0   /* FP64 delta, FP64 BC                       */
#define BCX_PREC_FP32C       1   /* FP32 compensated (Kahan) delta, FP64 BC   */

/* Schedules.                                                               */
#define BCX_PATH_PLAIN       0   /* per-component launches, block per source */
#define BCX_PATH_BATCHED     1   /* multi-source batched persistent path     */

typedef struct {
    int32_t        path;          /* BCX_PATH_*                              */
    int32_t        precision;     /* BCX_PREC_*                              */
    int32_t        use_chains;    /* 1: chain vertices via endpoint rows     */
    int32_t        batch_width;   /* batched: concurrent tiles, 1..8 (0 = 8) */
    const int64_t *src_subset;    /* NULL = all sources; else slots          */
    int64_t        n_src;         /* slot = comp_offset[c] + local vertex    */
    int64_t        row_cap_bytes; /* endpoint-row slab cap (0 = auto)        */
    int32_t        sigma_f64_threshold_log2; /* test hook: treat sigma >=
                                     2^k as overflow (0 = 64, i.e. uint64)  */
    int32_t        flags;         /* BCX_FLAG_* bits                          */
    int32_t        blocks_per_cu; /* batched multi-launch tiles: level-kernel
                                     blocks per compute unit (0 = tuned)     */
    int32_t        sources_per_lane; /* batched multi-launch and workgroup
                                     tiles: 1, 2 or 4 sources per lane of a
                                     wave64, tiles of 64, 128 or 256 sources
                                     (0 = tuned / the workgroup default)     */
} bcx_opts;

/* Batched path: run components of up to 2048 vertices through the
 * multi-launch tile path instead of the one-launch small-component kernel
 * (both compute the same; the flag exists for tests and measurements).     */
#define BCX_FLAG_NO_SMALL_KERNEL 1
/* Batched path: launch every kernel eagerly instead of replaying captured
 * hipGraphs (the persistent small-component and workgroup passes, and the
 * level launches of each multi-launch tile).                               */
#define BCX_FLAG_NO_GRAPH        2
/* Batched path: components of 2049..131072 vertices whose BFS levels are
 * narrow (average level width at most 4096 vertices from the highest-degree
 * vertex) run one workgroup per tile through every level.  NO_WG_KERNEL
 * sends them to the multi-launch tiles; WG_KERNEL sends every component of
 * that size range to the workgroup kernel.                                 */
#define BCX_FLAG_NO_WG_KERNEL    4
#define BCX_FLAG_WG_KERNEL       8
/* Batched path, multi-launch tiles: with use_chains, the chains that no
 * shortest path between their two endpoints can use (longer than a landmark
 * bound on the endpoints' distance) are taken out of the tile graph and their
 * interiors' levels, path counts and dependencies follow in closed form from
 * the endpoints'; a source inside such a chain runs as a lane seeded at both
 * endpoints (the reference reduces chain sources through endpoint rows,
 * Codes/postprocesses.cu:329-360).  NO_CONTRACT keeps every chain vertex in
 * the tile graph.                                                          */
#define BCX_FLAG_NO_CONTRACT     16
/* Batched path: components of at least 16384 vertices get a schedule
 * measured once on sample tiles and cached (no counterpart in the
 * reference, which switches schedule from sampled BFS depths,
 * Codes/kernels_free.cu:447-470): the workgroup or multi-launch kernel for
 * components of up to 131072 vertices, and for multi-launch components the
 * sources per lane, the level kernels' blocks per CU, the tiles in flight
 * and chain contraction on or off.  NO_TUNE keeps the defaults (the
 * level-width rule, 4 sources per lane, 8 blocks per CU, 8 tiles,
 * contraction on).  Options that fix a choice (the kernel flags,
 * sources_per_lane, batch_width, blocks_per_cu, NO_CONTRACT) are kept and
 * not tuned.                                                               */
#define BCX_FLAG_NO_TUNE         32

typedef struct {
    double  t_kernel_s;        /* HIP-event time of all GPU work in bcx_run   */
    double  t_total_s;         /* wall time of bcx_run                        */
    int64_t n_bfs_sources;     /* forward+backward sources                    */
    int64_t n_chain_sources;   /* combine+backward sources                    */
    int64_t n_row_only;        /* extra forward-only endpoint rows            */
    int64_t n_f64_sigma;       /* sources redone with FP64 sigma              */
    int64_t n_launches;        /* kernel launches                             */
    int32_t max_level;         /* deepest BFS level seen                      */
    int32_t n_ball_tiles;      /* batched: components whose tiles the tile-order
                                  tuner made graph balls (else BFS-rank runs) */
    double  bytes_alg;         /* algorithmic HBM bytes of the work done      */
    double  teps_work;         /* sum over sources of component arcs/2        */
    double  t_chain_s;         /* part of t_kernel_s spent on chain sources   */
    int64_t n_graph_launches;  /* batched: replays of captured launch sequences
                                  (their kernels are counted in n_launches)   */
    int64_t n_contracted;      /* batched: chain-interior vertices taken out of
                                  the multi-launch tile graphs (per run)      */
    int64_t n_list_entries;    /* batched, multi-launch tiles: (vertex, level)
                                  list entries of the completed tiles         */
    int64_t n_tiles;           /* batched, multi-launch tiles completed       */
    int32_t n_comp_small;      /* batched: components run on the small-component */
    int32_t n_comp_wg;         /*   kernel, the workgroup kernel and the      */
    int32_t n_comp_multi;      /*   multi-launch tiles                        */
    int32_t n_tuned;           /* batched: components whose schedule was measured
                                  (the schedule of each is measured once)     */
    int32_t sched_blocks_per_cu; /* batched: blocks per CU, tiles in flight and */
    int32_t sched_slots;       /*   sources per lane of the largest           */
    int32_t sched_sources_per_lane; /* multi-launch component                 */
} bcx_stats;

/* Context: device, streams, scratch.  mem_cap (bytes, 0 = 90% of free
 * memory) bounds every device allocation the context makes; exceeding it
 * returns BCX_E_NOMEM (the reference exits, Codes/postprocesses.cu:58-62). */
int  bcx_create(bcx_ctx **out, int device, size_t mem_cap);
void bcx_destroy(bcx_ctx *ctx);
const char *bcx_last_error(const bcx_ctx *ctx);
/* Number of compute units and the device name, for reports.               */
int  bcx_device_info(bcx_ctx *ctx, int32_t *n_cu, char *name, int32_t name_len);

/* Upload components once (replaces the per-BCC cudaMemcpy of
 * Codes/kernels_free.cu:43-51 and Codes/kernels_active.cu:69-78).         */
int  bcx_upload(bcx_ctx *ctx, const bcx_component *comps, int32_t n_comps,
                bcx_dgraph **out);
void bcx_dgraph_free(bcx_ctx *ctx, bcx_dgraph *g);

/* Batched path: every local vertex of component c of g in the order a run
 * with default flags tiles its sources (64 consecutive entries per tile; a
 * run over a subset keeps this order).  Runs the per-component tile-order
 * tuner on first use.  No counterpart in the reference (its sources are one
 * per block in id order); for harnesses that time samples of whole tiles.  */
int  bcx_batched_order(bcx_ctx *ctx, bcx_dgraph *g, int32_t c, int32_t *order);

/* Weighted Brandes over every component of g (begin_gpu free + active +
 * postprocess, Codes/begin_gpupgm.cu:42-89).  bc_out is a host array of
 * sum(n_c) doubles, component c at offset sum_{c'<c} n_{c'}; results are
 * ADDED to it (with a source subset, only the slices of components holding
 * sources are touched; the others would only receive zeros).  bc_out may
 * be NULL to leave the result on the device unread.  Per-source weight
 * (r(s)+1), per-target weight (r(t)+1) (Codes/kernels_free.cu:391,409,425). */
int  bcx_run(bcx_ctx *ctx, bcx_dgraph *g, const bcx_opts *opts,
             double *bc_out, bcx_stats *st);

/* Convenience: upload + run + free (one call per component set).          */
int  bcx_bc(bcx_ctx *ctx, const bcx_component *comps, int32_t n_comps,
            const bcx_opts *opts, double *bc_out, bcx_stats *st);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the batched run (bcx_run, BCX_PATH_BATCHED, default
 * flags) of ONE component of at most BCX_NOALLOC_MAX_N vertices -- the
 * small-component kernel -- that allocates nothing.  Every buffer beyond the
 * inputs and the output (tile table, source weights, work counters, and the
 * path-count / term / list slabs of at most BCX_NOALLOC_MAX_WORKGROUPS
 * workgroups) lives in one static device array of BCX_NOALLOC_SCRATCH_BYTES
 * per GPU, held from the first bcx_create on that GPU until the process
 * ends.  Inputs and output are the caller's device memory: d_row_ptr[n+1],
 * d_col[arcs], d_reach[n] (NULL = all zero), d_srcs[n_src] (distinct local
 * vertices; NULL = every vertex, n_src ignored); d_bc[n] is ADDED to, in
 * the FP64 convention of bcx_run.  precision is BCX_PREC_*.  The call runs
 * on the context's stream and returns after the work is done.
 * BCX_E_ARG when n > BCX_NOALLOC_MAX_N, n_src > n or a source is out of
 * range (then nothing is added).  The scratch array keeps the last call's
 * contents and is shared by every context on that GPU: calls that overlap
 * on one GPU race and must be serialised by the caller.                    */
#define BCX_NOALLOC_MAX_N           2048
#define BCX_NOALLOC_MAX_WORKGROUPS  8
#define BCX_NOALLOC_SCRATCH_BYTES   (28u * 1024u * 1024u + 64u * 1024u)
int  bcx_bc_small_noalloc(bcx_ctx *ctx, int32_t n, int64_t arcs, const int32_t *d_row_ptr,
                          const int32_t *d_col, const int32_t *d_reach, const int32_t *d_srcs,
                          int32_t n_src, int32_t precision, double *d_bc);

/* Whole pipeline for a decomposed host graph: GPU BC of every component
 * scattered to global ids plus the cut-vertex term (the reference's
 * bc_result, Codes/to_eargraph_construct.cpp:166-179 without the double
 * addition of the free-source BC).  bc_out[n] is overwritten.             */
int  bcx_host_bc(bcx_ctx *ctx, const bcx_host *h, const bcx_opts *opts,
                 double *bc_out, bcx_stats *st);

/* Parity probe: levels, sigma and delta of one source of one component, as
 * the chosen path computes them (chain vertices go through the chain-source
 * combine when opts->use_chains).  Any output pointer may be NULL.  sigma is
 * reported in both forms; sigma_is_f64 tells which one is authoritative.   */
int  bcx_debug_source(bcx_ctx *ctx, const bcx_component *comp, int32_t src,
                      const bcx_opts *opts, int32_t *level, uint64_t *sigma_u64,
                      double *sigma_f64, double *delta, int32_t *sigma_is_f64);

/* Kernels of the batched path (bcx_debug_tile's *kernel_used).            */
#define BCX_KERNEL_SMALL     1   /* workgroup per tile, state in LDS (n <= 2048)  */
#define BCX_KERNEL_WG        2   /* workgroup per tile, state in HBM (deep BFS)   */
#define BCX_KERNEL_MULTI     3   /* level launches across the device per tile     */

/* Parity probe of one whole tile of the batched path: srcs[0..ns) (distinct
 * local vertices; 1 <= ns <= 64, or up to 64 W on the multi-launch tiles of
 * W sources per lane) run as the sources of one tile on the kernel that bcx_run
 * picks for this component under opts->flags (the flags force a kernel as
 * they do for a run), with that kernel's per-source probe arrays switched
 * on; every source is read back.  Outputs are [ns * n], row i for srcs[i]
 * (any pointer may be NULL): levels, path counts (both forms; the tile's
 * counts are FP64 when *sigma_is_f64 = 1, else exact integers) and
 * dependencies.  *kernel_used receives the BCX_KERNEL_* that ran.  Extends
 * bcx_debug_source (one source; the reference's per-source rows are
 * Codes/kernels_active.cu:544-552) to the sources of a tile, which the
 * reference, one source per block, does not have.                          */
int  bcx_debug_tile(bcx_ctx *ctx, const bcx_component *comp, const int32_t *srcs, int32_t ns,
                    const bcx_opts *opts, int32_t *level, uint64_t *sigma_u64,
                    double *sigma_f64, double *delta, int32_t *sigma_is_f64,
                    int32_t *kernel_used);

/* ======================================================================= */
/* 3. Synthetic workloads                                                    */
/* ======================================================================= */

/* Seeded generators of the five benchmark workloads (1..5) plus small
 * variants; writes an edge list in file order.  scale in (0,1] shrinks the
 * vertex budget (1 = full size).  *uv is allocated with malloc and must be
 * released with bcx_free.                                                  */
int  bcx_generate(int32_t config, uint64_t seed, double scale,
                  int32_t *n, int64_t *m, int32_t **uv);
void bcx_free(void *p);

#ifdef __cplusplus
}
#endif
#endif /* BCX_H_ */
