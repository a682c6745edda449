// Host-callable launchers of the gfx950 kernels (defined in bcx_plain.hip and
// bcx_batched.hip).  Plain structs only; the runtime owns all memory.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bcx {

struct CompDev;

// Work item modes of the per-source kernel.
enum : int8_t { MODE_BC = 1, MODE_ROW = 2 };

// Per-workgroup scratch rows of the plain path: workgroup g uses
// [g*stride, (g+1)*stride) of each array.
struct PlainScratch {
    int32_t *d;
    uint64_t *sig;
    double *dl;
    int32_t *S;
    int32_t *off;
    int64_t stride;
};

// Endpoint rows (levels + sigma) kept for the chain-source combine.
struct Rows {
    int32_t *d;
    uint64_t *sig;
    int64_t stride;
};

// Chains of one component in device memory.
struct ChainDev {
    const int32_t *ptr;
    const int32_t *nodes;
    const int32_t *ends;
};

// Device counters: [0] work counter, [1] sigma overflow, [2] max level.
struct PlainArgs {
    const int32_t *src;    // [count] source vertex (bfs) | chain index (chain)
    const int32_t *aux;    // bfs: row slot or -1;  chain: position in chain
    const int8_t *mode;    // bfs only: MODE_* bits
    const int32_t *slots;  // chain only: [2*count] row slots of L and R
    int32_t count;
    int32_t sigma_f64;     // 1: FP64 sigma by pull
    uint64_t limit;        // u64 overflow threshold (0 = 2^64)
    int32_t *counters;
    double *bc;            // component slice of the BC accumulator
};

hipError_t launch_plain_bfs(const CompDev &c, const PlainArgs &a, const PlainScratch &sc,
                            const Rows &rows, int grid, hipStream_t st);
hipError_t launch_plain_chain(const CompDev &c, const ChainDev &ch, const PlainArgs &a,
                              const PlainScratch &sc, const Rows &rows, int grid, hipStream_t st);

}  // namespace bcx
