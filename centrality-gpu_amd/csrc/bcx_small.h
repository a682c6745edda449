// Small-component persistent kernel: device view and launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace bcx {

constexpr int SMALL_N = 2048;  // largest component handled by one workgroup
// LDS of a small-kernel workgroup: 32 B + 4 B per vertex of the largest
// component (72 KiB at SMALL_N), plus room for a tile's CSR as long as the
// total stays within SMALL_LDS bytes (more keeps the multi-launch kernels'
// blocks off the CU while it runs: 2% slower on the 12.8M-vertex workload
// with room up to 160 KiB)
constexpr size_t SMALL_LDS = 65536;
size_t small_lds_bytes(int64_t cap, int64_t csr);
int64_t small_csr_words(int64_t nmax, int64_t csr);

struct SmallComp {
    int32_t n;
    int32_t pad;
    const int32_t *rp, *col, *reach;
    double *bc;  // component slice of the BC accumulator
};

struct SmallTile {
    int32_t comp;  // index into comps
    int32_t src0;  // first source in srcs / ws
    int32_t ns;    // sources in the tile (<= 64; <= 64 W in the workgroup kernel)
    int32_t pad;
};

struct SmallArgs {
    const SmallComp *comps;
    const SmallTile *tiles;
    int32_t n_tiles;
    const int32_t *srcs;
    const double *ws;
    int32_t *counter;   // [0] work counter, [1] tiles listed for an FP64 rerun, [2] a full list
    int32_t *redo;      // [n_tiles] those tiles
    uint64_t *sig;      // per-workgroup slabs of slab words
    void *tt;
    int32_t *Q;
    uint64_t *Qm;
    int64_t slab;       // sigma / term words per workgroup (largest n * 64)
    uint64_t limit;     // u64 sigma threshold (0 = 2^64)
    // workgroup-per-tile kernel for larger components: the visited and level
    // masks and the level offsets live in per-workgroup HBM slabs
    uint64_t *masks;    // [grid][3 * nmax]
    int32_t *qoffs;     // [grid][nmax + 2]
    int64_t nmax;       // largest component
    int64_t qcap;       // list entries per workgroup (64 * nmax + 64)
    // small kernel: LDS words for a copy of a tile's CSR (n + 1 row
    // pointers and the arcs); tiles whose CSR is larger, and every tile when
    // 0, read the CSR from HBM
    int64_t csr_cap;
    // workgroup-per-tile kernel: sig / tt hold lane-compacted records (one
    // word per lane of each list entry) located through these offsets
    int32_t *qr;        // [grid][qcap] record offset of each list entry
    int32_t *ro;        // [grid][2 * nmax] record offset of each vertex's current entry
    // workgroup-per-tile kernel: sources per lane W (0 or 1: 64-source tiles;
    // 2 or 4: tiles of 64 W sources, source k + 64 h on lane k, masks of W
    // words) and, for W > 1, the level each vertex was last listed at
    int32_t spl;
    int32_t *tag;       // [grid][nmax]
    uint64_t *psig;     // probe only: [grid][slab] path counts per (vertex, lane)
    // probe only (null in runs): per-workgroup slabs like sig, holding the
    // level and the dependency of every (vertex, lane) of the last tile
    int32_t *lvl;
    double *dbg;
    // when set, the tile count is read here at kernel start (written by an
    // earlier kernel on the stream) instead of n_tiles
    const int32_t *n_tiles_dev;
};

hipError_t launch_small(const SmallArgs &a, bool f64, bool fp32, int grid, hipStream_t st);
int small_blocks_per_cu(int64_t nmax, int64_t csr);
hipError_t launch_small_reset(int32_t *counter, hipStream_t st);

// Components up to WG_N vertices whose levels are narrow (deep BFS) run one
// workgroup per tile through every level, like the small kernel, with the
// per-vertex state in HBM (a.spl sources per lane).
constexpr int WG_N = 1 << 17;
hipError_t launch_wg(const SmallArgs &a, bool f64, bool fp32, int grid, hipStream_t st);

}  // namespace bcx
