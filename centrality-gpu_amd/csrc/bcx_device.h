// Device-side helpers shared by the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define BCX_INF 0x7fffffff

namespace bcx {

// Component in device memory (CSR + reach weights).
struct CompDev {
    int32_t n;
    int64_t arcs;
    const int32_t *rp;
    const int32_t *col;
    const int32_t *reach;
};

// Loads/stores of words that other lanes of the same workgroup update with
// atomics during a level: agent-scope relaxed atomics, so the read never comes
// from a vector-L1 line the L2 atomic left stale.
__device__ __forceinline__ int32_t ald(const int32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t ald(const uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void ast(int32_t *p, int32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void ast(uint64_t *p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Global-memory views of pointers the compiler cannot place by itself
// (read from device memory rather than from kernel arguments).  Accesses
// through them are global instead of flat instructions: a flat access also
// counts as a pending LDS operation, so every later wait for an LDS result
// would wait for it too.
using gint = __attribute__((address_space(1))) const int32_t;
using gdbl = __attribute__((address_space(1))) double;
__device__ __forceinline__ const gint *glb(const int32_t *p) { return (const gint *)p; }
__device__ __forceinline__ gdbl *glb(double *p) { return (gdbl *)p; }
// FP64 atomic add to global memory (relaxed, device scope)
__device__ __forceinline__ void gadd(gdbl *p, double x) {
    __hip_atomic_fetch_add(p, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ double u2d(uint64_t x) { return __longlong_as_double((long long)x); }
__device__ __forceinline__ uint64_t d2u(double x) { return (uint64_t)__double_as_longlong(x); }

// Wave64 reductions (fixed butterfly order: deterministic).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

}  // namespace bcx
