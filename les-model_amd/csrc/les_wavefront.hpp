// les_wavefront.hpp -- lexicographic sweeps as one launch each: the IC(0)
// factor and its two triangular solves (pressure.cpp:495-556, 442-482) and
// the lexicographic SOR sweep (pressure.cpp:216-234).
//
// Every cell of these sweeps needs its i-1, j-1 and k-1 neighbours' new
// values (i+1, j+1, k+1 for the backward solve), so the cells of one
// i+j+k hyperplane are independent and the hyperplanes are sequential.
// Schedule: one wave64 per 8x8 tile of (i, j) columns.  Lane (a, b) of the
// tile owns column (i0+a, j0+b) and walks it along k, skewed by a+b, so at
// tile level l it updates k = l-a-b: its k-1 neighbour is its own previous
// value, its i-1 and j-1 neighbours were updated by lanes (a-1, b) and
// (a, b-1) at level l-1 and arrive by a cross-lane shuffle.  Lanes on the
// tile's low i (j) edge take them from the tile below instead: tiles hand
// their edge values over in 64-bit packets, 32 bits of the value and the
// launch's 32-bit tag, written and polled with device-scope atomics, so a
// consumer never reads a stale value and needs no flag or fence; the
// buffers are never cleared (a new launch has a new tag).  A wave polls at
// most `spin_limit` times in all; past that it records a stall and goes
// on, so every launch ends.  Tiles are taken from an atomic ticket in
// anti-diagonal order (i+j tile index): a wave only waits for tiles with
// smaller tickets, which are already running, so the launch cannot
// deadlock whatever number of waves the GPU holds at once.  Every cell is
// computed by the reference's formula from the same operands, so the
// results equal the serial loop's bit for bit.
#pragma once

#include "les_kernels.hpp"
#include "les_physics.hpp"

namespace les {

enum WfKind { WF_FACTOR = 0, WF_FWD = 1, WF_BWD = 2, WF_SOR = 3, WF_KINDS = 4 };

// Device state of the sweeps of one context (zeroed once at allocation).
struct WfState {
    unsigned long long ticket[WF_KINDS];  // waves ever started, per kind
    int stalled;                          // 1: a wave ran out of its spin budget
    int pad[3];
    // where the lanes with nothing to write send their results, so that the
    // stores of a chunk need no branch
    alignas(16) unsigned long long sink[2];
};

struct WfBufs {
    WfState* st;
    unsigned long long* fx;  // packets on the i-boundaries between tiles
    unsigned long long* fy;  // packets on the j-boundaries
    int spin_limit;          // polls a wave may make in one launch
};

constexpr int kWfTile = 8;
constexpr int kWfDefaultSpin = 1 << 22;  // about 4 s of polling per wave

// Packets each face buffer must hold for a grid (FP64 values, full grid).
size_t wf_face_packets_x(const Dims& d);
size_t wf_face_packets_y(const Dims& d);

hipError_t launch_wf_ic_factor(hipStream_t s, const Dims& d, Stencil<double> A, double* Lc,
                               double* Lw, double* Ls, double* Lb, const WfBufs& wb);
template <typename T>
hipError_t launch_wf_ic_fwd(hipStream_t s, const Dims& d, const T* r, T* z, const T* Lc,
                            const T* Lw, const T* Ls, const T* Lb, const SolveScalars* sc,
                            const WfBufs& wb);
template <typename T>
hipError_t launch_wf_ic_bwd(hipStream_t s, const Dims& d, T* z, const T* Lc, const T* Lw,
                            const T* Ls, const T* Lb, const SolveScalars* sc, const WfBufs& wb);
hipError_t launch_wf_sor(hipStream_t s, const Dims& d, double* p, const double* rhs,
                         double one_m_omega, double omega_coef, double dx2, double dy2,
                         double dz2, StatScalars* st, const WfBufs& wb);

}  // namespace les
