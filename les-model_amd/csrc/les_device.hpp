// les_device.hpp -- device-side building blocks shared by the LES kernels.
//
// Grid layout (src/field.h:40): idx = i*ny*nz + j*nz + k, k contiguous.
// Every kernel that touches whole grids walks the flat cell range in
// per-workgroup chunks; the workgroup -> chunk map is XCD-aware (MI355X
// dispatches workgroups round-robin over its 8 XCDs, so consecutive chunks
// are given to workgroups that land on the same XCD and share its L2).
//
// Reductions are deterministic: each workgroup reduces its chunk in a fixed
// order (lane-sequential, then a fixed xor-shuffle tree over the wave64, then
// wave partials in order), writes one partial per workgroup, and the last
// workgroup to finish sums the partials in a fixed order and runs the scalar
// epilogue.  The grid size depends only on the problem size, so results do
// not depend on the device, the run or the launch timing.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace les {

constexpr int kBlock = 256;      // threads per workgroup (4 wave64)
constexpr int kWaves = kBlock / 64;
constexpr int kXcd = 8;          // XCDs on MI355X
constexpr int kMaxGrid = 4096;   // workgroups of a streaming kernel (multiple of kXcd)
constexpr int kMaxRed = 4;       // sums a fused kernel may reduce at once

struct Dims {
    int nx, ny, nz;
    int sij;     // ny*nz
    int N;       // nx*ny*nz (< 2^31, src/field.h:40 uses int)
};

struct Chunking {
    int grid;    // workgroups
    int chunk;   // cells per workgroup (multiple of kBlock)
};

inline Chunking make_chunking(int N)
{
    int blocks = (N + kBlock - 1) / kBlock;
    int grid = blocks < kMaxGrid ? blocks : kMaxGrid;
    grid = ((grid + kXcd - 1) / kXcd) * kXcd;
    int per = (blocks + grid - 1) / grid;
    return Chunking{grid, per * kBlock};
}

// Logical chunk of this workgroup: workgroups b, b+8, b+16 ... run on one XCD
// and receive consecutive chunks.
__device__ __forceinline__ int logical_block(int grid)
{
    const int b = blockIdx.x;
    return (b % kXcd) * (grid / kXcd) + b / kXcd;
}

__device__ __forceinline__ void decode(const Dims& d, int id, int& i, int& j, int& k)
{
    i = id / d.sij;
    const int r = id - i * d.sij;
    j = r / d.nz;
    k = r - j * d.nz;
}

// ---- wave64 / workgroup sums in a fixed order ----
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

// Sums NR values over the workgroup; result valid in thread 0.
template <int NR>
__device__ __forceinline__ void block_sum(double (&v)[NR])
{
    __shared__ double sh[kWaves][kMaxRed];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int r = 0; r < NR; ++r) v[r] = wave_sum(v[r]);
    if (lane == 0) {
#pragma unroll
        for (int r = 0; r < NR; ++r) sh[wave][r] = v[r];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            double s = sh[0][r];
            for (int w = 1; w < kWaves; ++w) s += sh[w][r];
            v[r] = s;
        }
    }
    __syncthreads();
}

template <int NR>
__device__ __forceinline__ void block_max(double (&v)[NR])
{
    __shared__ double sh[kWaves][kMaxRed];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int r = 0; r < NR; ++r) v[r] = wave_max(v[r]);
    if (lane == 0) {
#pragma unroll
        for (int r = 0; r < NR; ++r) sh[wave][r] = v[r];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            double s = sh[0][r];
            for (int w = 1; w < kWaves; ++w) s = s > sh[w][r] ? s : sh[w][r];
            v[r] = s;
        }
    }
    __syncthreads();
}

__device__ __forceinline__ void store_agent(double* p, double v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double load_agent(const double* p)
{
    return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Grid-wide deterministic reduction tail.  Every workgroup calls it once with
// its NR partial sums (valid in thread 0).  Returns true in exactly one
// workgroup -- the last to arrive -- with the NR grid totals in `tot`
// (valid in every thread of that workgroup).  `partials` holds grid*kMaxRed
// doubles; `counter` is zero on entry and is left zero.
template <int NR, bool MAX = false>
__device__ __forceinline__ bool grid_reduce(double (&v)[NR], double* partials,
                                            unsigned int* counter, int grid, int lb,
                                            double (&tot)[NR])
{
    __shared__ int am_last;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int r = 0; r < NR; ++r) store_agent(&partials[lb * kMaxRed + r], v[r]);
        __threadfence();
        unsigned prev = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL,
                                               __HIP_MEMORY_SCOPE_AGENT);
        am_last = (prev == (unsigned)grid - 1u);
    }
    __syncthreads();
    if (!am_last) return false;
    __threadfence();
    double s[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) s[r] = MAX ? 0.0 : 0.0;
    // fixed order: thread t takes partials t, t+kBlock, ...
    for (int b = threadIdx.x; b < grid; b += kBlock) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            double x = load_agent(&partials[b * kMaxRed + r]);
            if (MAX) s[r] = s[r] > x ? s[r] : x;
            else s[r] += x;
        }
    }
    if (MAX) block_max<NR>(s);
    else block_sum<NR>(s);
    __shared__ double out[kMaxRed];
    if (threadIdx.x == 0) {
#pragma unroll
        for (int r = 0; r < NR; ++r) out[r] = s[r];
        *counter = 0u;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < NR; ++r) tot[r] = out[r];
    return true;
}

}  // namespace les
