// Small-component persistent kernel: device view and launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace bcx {

constexpr int SMALL_N = 2048;  // largest component handled by one workgroup

struct SmallComp {
    int32_t n;
    int32_t pad;
    const int32_t *rp, *col, *reach;
    double *bc;  // component slice of the BC accumulator
};

struct SmallTile {
    int32_t comp;  // index into comps
    int32_t src0;  // first source in srcs / ws
    int32_t ns;    // sources in the tile (<= 64)
    int32_t pad;
};

struct SmallArgs {
    const SmallComp *comps;
    const SmallTile *tiles;
    int32_t n_tiles;
    const int32_t *srcs;
    const double *ws;
    int32_t *counter;   // [0] work counter, [1] tiles listed for an FP64 rerun
    int32_t *redo;      // [n_tiles] those tiles
    uint64_t *sig;      // per-workgroup slabs of slab words
    void *tt;
    int32_t *Q;
    uint64_t *Qm;
    int64_t slab;       // SMALL_N * 64
    uint64_t limit;     // u64 sigma threshold (0 = 2^64)
};

hipError_t launch_small(const SmallArgs &a, bool f64, bool fp32, int grid, hipStream_t st);
hipError_t launch_small_reset(int32_t *counter, hipStream_t st);

}  // namespace bcx
