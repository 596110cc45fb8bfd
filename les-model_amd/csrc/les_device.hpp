// les_device.hpp -- device-side building blocks shared by the LES kernels.
//
// Grid layout (src/field.h:40): idx = i*ny*nz + j*nz + k, k contiguous.
//
// Work decomposition.  A "row" is one (i, j) line of nz contiguous cells.
// Each workgroup owns a contiguous range of rows; threads run along k, so
// the row's i and j are uniform across the wavefront (scalar ALU) and no
// per-cell division is needed.  When a row has fewer cells than the
// workgroup has threads, one pass covers several rows.
// The workgroup -> row-range map is XCD-aware: MI355X dispatches
// workgroups round-robin over its 8 XCDs, so workgroups b, b+8, b+16, ...
// (one XCD) receive consecutive row ranges and their i+-1 / j+-1 neighbours
// are served by that XCD's L2.
//
// Reductions are deterministic and fence-free: every workgroup reduces its
// rows in a fixed order (lane-sequential, a fixed xor-shuffle tree over the
// wave64, waves in order) and writes one partial; a one-workgroup finalize
// kernel then sums the partials in a fixed order and runs the scalar
// epilogue on the device.  The grid size depends only on the problem size.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace les {

// Threads per workgroup of the row kernels (16 wave64).  With kMaxGrid
// below, a streaming kernel runs one 1024-thread workgroup per CU, each
// walking a long contiguous range of rows.  Measured at 256^3 on MI355X
// against 256 threads x 1024 workgroups: single Chebyshev step 80 -> 72 us
// (5.2 TB/s), fused matvec 115 -> 103 us (FP32); larger grids of smaller
// workgroups were slower still.
#ifndef LES_BLOCK
#define LES_BLOCK 1024
#endif
constexpr int kBlock = LES_BLOCK;
constexpr int kWaves = kBlock / 64;
constexpr int kXcd = 8;          // XCDs on MI355X
// Workgroups of a streaming kernel (multiple of kXcd): one per CU.
#ifndef LES_MAX_GRID
#define LES_MAX_GRID 256
#endif
constexpr int kMaxGrid = LES_MAX_GRID;
constexpr int kMaxRed = 4;       // sums a fused kernel may reduce at once
// Workgroups of a fused Chebyshev pair launch: one round of what its LDS
// ring and registers allow per CU (measured at 256^3 on MI355X; more
// workgroups only add the halo planes each i-run recomputes).
#ifndef LES_PAIR_GRID_CODE32
#define LES_PAIR_GRID_CODE32 512
#endif
#ifndef LES_PAIR_GRID_CODE64
#define LES_PAIR_GRID_CODE64 256
#endif
#ifndef LES_PAIR_GRID_ARRAYS
#define LES_PAIR_GRID_ARRAYS 256
#endif
#ifndef LES_PAIR_MAX_LDS
#define LES_PAIR_MAX_LDS (160 * 1024)
#endif
constexpr int kPairTargetGridCode32 = LES_PAIR_GRID_CODE32;  // 512 threads, 2 per CU
constexpr int kPairTargetGridCode64 = LES_PAIR_GRID_CODE64;  // 1024 threads, 1 per CU
constexpr int kPairTargetGridArrays = LES_PAIR_GRID_ARRAYS;  // 768 threads, 1 per CU
constexpr int kPairMaxGrid = 4096;  // largest target a caller may set (the pair reduces nothing)
constexpr size_t kPairMaxLds = LES_PAIR_MAX_LDS;  // LDS of one workgroup on gfx950

struct Dims {
    int nx, ny, nz;
    int sij;     // ny*nz
    int N;       // nx*ny*nz (< 2^31, src/field.h:40 uses int)
};

// Row ranges of a launch.
struct Chunking {
    int grid;    // workgroups (multiple of kXcd)
    int rpb;     // rows per workgroup
    int rpp;     // rows per pass (kBlock / nz when nz < kBlock, else 1)
};

// V > 1: each thread owns V consecutive cells of a row (vector loads);
// requires nz % V == 0.
inline Chunking make_chunking(const Dims& d, int V = 1)
{
    const int rows = d.nx * d.ny;
    const int tpr = d.nz / V;  // threads per row
    const int rpp = tpr >= kBlock ? 1 : kBlock / tpr;
    // aim for about kBlock cells of work per thread-row pass, capped grid
    int grid = (rows + rpp - 1) / rpp;
    if (grid > kMaxGrid) grid = kMaxGrid;
    grid = ((grid + kXcd - 1) / kXcd) * kXcd;
    int rpb = (rows + grid - 1) / grid;
    rpb = ((rpb + rpp - 1) / rpp) * rpp;
    return Chunking{grid, rpb, rpp};
}

// Logical range of this workgroup: workgroups b, b+8, b+16 ... run on one
// XCD and receive consecutive ranges.
__device__ __forceinline__ int logical_block(int grid)
{
    const int b = blockIdx.x;
    return (b % kXcd) * (grid / kXcd) + b / kXcd;
}

// FOR_CELLS(d, ch, lb) { ... } runs the body for every cell of the
// workgroup's rows with i, j, k, id in scope.
#define FOR_CELLS(d, ch, lb)                                                                      \
    for (int sub_ = ((ch).rpp > 1 ? (int)threadIdx.x / (d).nz : 0),                              \
             k0_ = (int)threadIdx.x - sub_ * (d).nz,                                              \
             row_ = (lb) * (ch).rpb + sub_,                                                       \
             rend_ = min((d).nx * (d).ny, ((lb) + 1) * (ch).rpb);                                 \
         sub_ < (ch).rpp && row_ < rend_; row_ += (ch).rpp)                                       \
        for (int i = row_ / (d).ny, j = row_ - i * (d).ny, k = k0_, id = row_ * (d).nz + k0_;     \
             k < (d).nz; k += ((ch).rpp > 1 ? (d).nz : kBlock),                                   \
                 id += ((ch).rpp > 1 ? (d).nz : kBlock))

// FOR_VCELLS(d, ch, lb, V) { ... }: as FOR_CELLS, with each thread owning
// the V cells k .. k+V-1 (k a multiple of V) of its row; ch from
// make_chunking(d, V).
#define FOR_VCELLS(d, ch, lb, V)                                                                  \
    for (int tpr_ = (d).nz / (V), sub_ = ((ch).rpp > 1 ? (int)threadIdx.x / tpr_ : 0),           \
             k0_ = ((int)threadIdx.x - sub_ * tpr_) * (V),                                        \
             row_ = (lb) * (ch).rpb + sub_,                                                       \
             rend_ = min((d).nx * (d).ny, ((lb) + 1) * (ch).rpb);                                 \
         sub_ < (ch).rpp && row_ < rend_; row_ += (ch).rpp)                                       \
        for (int i = row_ / (d).ny, j = row_ - i * (d).ny, k = k0_, id = row_ * (d).nz + k0_;     \
             k < (d).nz; k += ((ch).rpp > 1 ? (d).nz : kBlock * (V)),                             \
                 id += ((ch).rpp > 1 ? (d).nz : kBlock * (V)))

// ---- wave64 / workgroup reductions in a fixed order ----
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        double o = __shfl_xor(v, off, 64);
        v = v > o ? v : o;
    }
    return v;
}

// Reduces NR values over the workgroup; result valid in thread 0.
template <int NR, bool MAX = false>
__device__ __forceinline__ void block_reduce(double (&v)[NR])
{
    __shared__ double sh[kWaves][kMaxRed];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int r = 0; r < NR; ++r) v[r] = MAX ? wave_max(v[r]) : wave_sum(v[r]);
    if (lane == 0) {
#pragma unroll
        for (int r = 0; r < NR; ++r) sh[wave][r] = v[r];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            double s = sh[0][r];
            for (int w = 1; w < kWaves; ++w) s = MAX ? (s > sh[w][r] ? s : sh[w][r]) : s + sh[w][r];
            v[r] = s;
        }
    }
    __syncthreads();
}

// Workgroup partial of NR sums (or maxima) into partials[lb*kMaxRed + r].
// Every workgroup of the launch must call it.
template <int NR, bool MAX = false>
__device__ __forceinline__ void block_partial(double (&v)[NR], double* partials, int lb)
{
    block_reduce<NR, MAX>(v);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int r = 0; r < NR; ++r) partials[lb * kMaxRed + r] = v[r];
    }
}

// Fixed-order sum (or max) of the grid's partials by one workgroup; result
// valid in every thread.
template <int NR, bool MAX = false>
__device__ __forceinline__ void reduce_partials(const double* partials, int grid, double (&tot)[NR])
{
    double s[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) s[r] = 0.0;
    // thread t takes partials t, t + kBlock, t + 2 kBlock, ... in that
    // order; the loads of a chunk are all issued before the first add, so
    // their latencies overlap instead of adding up
    constexpr int kU = kMaxGrid > kBlock ? kMaxGrid / kBlock : 1;
    for (int b0 = 0; b0 < grid; b0 += kU * kBlock) {
        double x[kU][NR];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int b = b0 + (int)threadIdx.x + u * kBlock;
#pragma unroll
            for (int r = 0; r < NR; ++r) x[u][r] = b < grid ? partials[b * kMaxRed + r] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            if (b0 + (int)threadIdx.x + u * kBlock >= grid) break;
#pragma unroll
            for (int r = 0; r < NR; ++r)
                s[r] = MAX ? (s[r] > x[u][r] ? s[r] : x[u][r]) : s[r] + x[u][r];
        }
    }
    block_reduce<NR, MAX>(s);
    __shared__ double out[kMaxRed];
    if (threadIdx.x == 0) {
#pragma unroll
        for (int r = 0; r < NR; ++r) out[r] = s[r];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < NR; ++r) tot[r] = out[r];
}

}  // namespace les
