// bcx_bc_small_noalloc: the batched run of one small component (n <=
// BCX_NOALLOC_MAX_N) on the small-component kernel, with every buffer beyond
// the caller's inputs and output in one static device array, so the call
// allocates nothing.  The launch sequence is kernels only:
//   setup   tile table, sources, weights, zeroed counters, validity check
//   k_small 64-bit path counts over every tile (tile count read on device)
//   remap   the tiles whose counts left 64 bits become the second table
//   k_small FP64 path counts over those tiles (none: returns at once)
// followed by one read-back of the status words and a stream sync.
#include <algorithm>
#include <mutex>
#include <string>

#include "bcx.h"
#include "bcx_runtime.h"
#include "bcx_small.h"

using namespace bcx;

namespace {

constexpr int NA_N = BCX_NOALLOC_MAX_N;
constexpr int NA_WG = BCX_NOALLOC_MAX_WORKGROUPS;
constexpr int NA_TILES = (NA_N + 63) / 64;
constexpr size_t NA_SLAB = (size_t)NA_N * 64;  // words per workgroup slab
static_assert(NA_N <= SMALL_N, "the small-component kernel's LDS holds SMALL_N vertices");

constexpr size_t up(size_t x) { return (x + 255) & ~(size_t)255; }

// byte offsets of the scratch array's parts
constexpr size_t O_COMP = 0;
constexpr size_t O_TILES1 = up(O_COMP + sizeof(SmallComp));
constexpr size_t O_TILES2 = up(O_TILES1 + sizeof(SmallTile) * NA_TILES);
constexpr size_t O_SRCS = up(O_TILES2 + sizeof(SmallTile) * NA_TILES);
constexpr size_t O_WS = up(O_SRCS + 4 * (size_t)NA_N);
constexpr size_t O_ZERO = up(O_WS + 8 * (size_t)NA_N);
constexpr size_t O_CTR = up(O_ZERO + 4 * (size_t)NA_N);
constexpr size_t O_REDO1 = up(O_CTR + 64);
constexpr size_t O_REDO2 = up(O_REDO1 + 4 * (size_t)NA_TILES);
constexpr size_t O_SIG = up(O_REDO2 + 4 * (size_t)NA_TILES);
constexpr size_t O_TT = O_SIG + 8 * NA_SLAB * NA_WG;
constexpr size_t O_Q = O_TT + 8 * NA_SLAB * NA_WG;
constexpr size_t O_QM = O_Q + 4 * NA_SLAB * NA_WG;
constexpr size_t O_END = O_QM + 8 * NA_SLAB * NA_WG;
static_assert(O_END <= BCX_NOALLOC_SCRATCH_BYTES, "scratch layout exceeds the stated size");

// status words at O_CTR: [0..3] first pass counters, [4..7] second pass
// counters, [8] first pass tiles, [9] second pass tiles, [10] invalid input
enum { C1 = 0, C2 = 4, N1 = 8, N2 = 9, BAD = 10 };

struct NaArgs {
    int32_t n, n_src, n_tiles;
    const int32_t *rp, *col, *reach, *srcs;
    double *bc;
    unsigned char *base;  // the scratch array's device address
};

}  // namespace

// The scratch array.  Host code must use its device address from
// hipGetSymbolAddress (kept per device below), never this name.
static __device__ __align__(256) unsigned char scratch[BCX_NOALLOC_SCRATCH_BYTES];

static std::mutex g_mu;
static void *g_ptr[64];      // device address of `scratch` per device
static hipError_t g_err[64];  // result of the lookup per device

// Set-up, from bcx_create with `device` current: looks the array's device
// address up once per device.
int noalloc_init(int device) {
    if (device < 0 || device >= 64) return BCX_E_ARG;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[device]) return BCX_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_err[device] = e;
    if (e != hipSuccess || !p) return BCX_E_HIP;
    g_ptr[device] = p;
    return BCX_OK;
}

__global__ __launch_bounds__(256) void k_na_setup(NaArgs x) {
    unsigned char *b = x.base;
    int32_t *ctr = (int32_t *)(b + O_CTR);
    int32_t *srcs = (int32_t *)(b + O_SRCS);
    double *ws = (double *)(b + O_WS);
    int32_t *zero = (int32_t *)(b + O_ZERO);
    SmallTile *tiles = (SmallTile *)(b + O_TILES1);
    __shared__ int bad;
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    const int32_t *reach = x.reach ? x.reach : zero;
    for (int v = threadIdx.x; v < x.n; v += blockDim.x) zero[v] = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < x.n_src; i += blockDim.x) {
        const int s = x.srcs ? x.srcs[i] : i;
        if (s < 0 || s >= x.n) {
            bad = 1;  // (any thread; all write the same value)
            srcs[i] = 0;
            ws[i] = 0.0;
            continue;
        }
        srcs[i] = s;
        ws[i] = (double)reach[s] + 1.0;
    }
    for (int t = threadIdx.x; t < x.n_tiles; t += blockDim.x) {
        SmallTile tl{};
        tl.comp = 0;
        tl.src0 = 64 * t;
        tl.ns = min(64, x.n_src - 64 * t);
        tiles[t] = tl;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        SmallComp c{};
        c.n = x.n;
        c.rp = x.rp;
        c.col = x.col;
        c.reach = reach;
        c.bc = x.bc;
        *(SmallComp *)(b + O_COMP) = c;
        for (int k = 0; k < 8; ++k) ctr[k] = 0;
        ctr[N1] = bad ? 0 : x.n_tiles;
        ctr[N2] = 0;
        ctr[BAD] = bad;
    }
}

// The first pass listed the tiles whose 64-bit counts overflowed: they make
// the second pass's table.
__global__ __launch_bounds__(64) void k_na_remap(unsigned char *b) {
    int32_t *ctr = (int32_t *)(b + O_CTR);
    const SmallTile *t1 = (const SmallTile *)(b + O_TILES1);
    SmallTile *t2 = (SmallTile *)(b + O_TILES2);
    const int32_t *redo = (const int32_t *)(b + O_REDO1);
    const int k = ctr[C1 + 1];
    for (int i = threadIdx.x; i < k; i += blockDim.x) t2[i] = t1[redo[i]];
    if (threadIdx.x == 0) ctr[N2] = k;
}

static SmallArgs pass_args(unsigned char *b, bool second, int32_t n) {
    SmallArgs a{};
    a.comps = (const SmallComp *)(b + O_COMP);
    a.tiles = (const SmallTile *)(b + (second ? O_TILES2 : O_TILES1));
    a.n_tiles = 0;
    a.n_tiles_dev = (const int32_t *)(b + O_CTR) + (second ? N2 : N1);
    a.srcs = (const int32_t *)(b + O_SRCS);
    a.ws = (const double *)(b + O_WS);
    a.counter = (int32_t *)(b + O_CTR) + (second ? C2 : C1);
    a.redo = (int32_t *)(b + (second ? O_REDO2 : O_REDO1));
    a.sig = (uint64_t *)(b + O_SIG);
    a.tt = b + O_TT;
    a.Q = (int32_t *)(b + O_Q);
    a.Qm = (uint64_t *)(b + O_QM);
    a.slab = (int64_t)NA_SLAB;
    a.limit = 0;
    a.nmax = n;  // sizes the kernel's LDS
    return a;
}

extern "C" int bcx_bc_small_noalloc(bcx_ctx *ctx, int32_t n, int64_t arcs, const int32_t *d_row_ptr,
                                    const int32_t *d_col, const int32_t *d_reach, const int32_t *d_srcs,
                                    int32_t n_src, int32_t precision, double *d_bc) {
    if (!ctx) return BCX_E_ARG;
    if (n < 1 || n > NA_N)
        return ctx_fail(ctx, BCX_E_ARG, "bcx_bc_small_noalloc: n must be 1.." + std::to_string(NA_N) +
                                            " (BCX_NOALLOC_MAX_N), got " + std::to_string(n));
    if (!d_srcs) n_src = n;
    if (n_src < 1 || n_src > n) return ctx_fail(ctx, BCX_E_ARG, "bcx_bc_small_noalloc: n_src must be 1..n");
    if (arcs < 0 || (arcs > 0 && !d_col) || !d_row_ptr || !d_bc ||
        (precision != BCX_PREC_FP64 && precision != BCX_PREC_FP32C))
        return ctx_fail(ctx, BCX_E_ARG, "bcx_bc_small_noalloc: bad arguments");
    const int dev = ctx->device;
    unsigned char *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (dev >= 0 && dev < 64) base = (unsigned char *)g_ptr[dev];
        if (!base)
            return ctx_fail(ctx, BCX_E_HIP, std::string("bcx_bc_small_noalloc: scratch array not available: ") +
                                                hipGetErrorString(dev >= 0 && dev < 64 ? g_err[dev] : hipErrorInvalidDevice));
    }
    int rc = ctx_hip(ctx, hipSetDevice(dev), "set device");
    if (rc) return rc;
    NaArgs x{};
    x.n = n;
    x.n_src = n_src;
    x.n_tiles = (n_src + 63) / 64;
    x.rp = d_row_ptr;
    x.col = d_col;
    x.reach = d_reach;
    x.srcs = d_srcs;
    x.bc = d_bc;
    x.base = base;
    const int grid = std::min(NA_WG, x.n_tiles);
    const bool fp32 = precision == BCX_PREC_FP32C;
    hipStream_t st = ctx->stream;
    hipLaunchKernelGGL(k_na_setup, dim3(1), dim3(256), 0, st, x);
    if ((rc = ctx_hip(ctx, hipGetLastError(), "no-alloc setup"))) return rc;
    if ((rc = ctx_hip(ctx, launch_small(pass_args(base, false, n), false, fp32, grid, st), "no-alloc pass"))) return rc;
    hipLaunchKernelGGL(k_na_remap, dim3(1), dim3(64), 0, st, base);
    if ((rc = ctx_hip(ctx, hipGetLastError(), "no-alloc remap"))) return rc;
    if ((rc = ctx_hip(ctx, launch_small(pass_args(base, true, n), true, fp32, grid, st), "no-alloc FP64 pass")))
        return rc;
    int32_t status[16];
    if ((rc = ctx_hip(ctx, hipMemcpyAsync(status, base + O_CTR, sizeof(status), hipMemcpyDeviceToHost, st),
                      "no-alloc status")) ||
        (rc = ctx_hip(ctx, hipStreamSynchronize(st), "no-alloc run")))
        return rc;
    if (status[BAD]) return ctx_fail(ctx, BCX_E_ARG, "bcx_bc_small_noalloc: a source is out of range");
    return BCX_OK;
}
