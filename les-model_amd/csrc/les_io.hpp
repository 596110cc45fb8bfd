// les_io.hpp -- checkpoint, VTK output and time statistics of the device-
// resident state (src/io.cpp, src/statistics.cpp).
//
// The file writers are host code: they format or copy what the device
// fields hold and produce files byte-identical to the reference writers'.
// The statistics sums live on the device and are updated by one streaming
// kernel per step.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "les_device.hpp"

namespace les {

// The ten running sums / results of the statistics module, in the order of
// les_gpu.h's LES_ST_* slots: u, v, w, p, uu, vv, ww, uv, uw, vw.
constexpr int kStatArrays = 10;
struct StatArrays {
    double* s[kStatArrays];
};

// sum += (u, v, w, p, u*u, v*v, w*w, u*v, u*w, v*w) per cell
// (statistics.cpp:98-131).  192 algorithmic bytes per cell.
hipError_t launch_stats_update(hipStream_t s, int n, const double* u, const double* v,
                               const double* w, const double* p, StatArrays sums);
// out = sums * n_inv for the four means; stress = sum * n_inv - mean_a * mean_b
// (statistics.cpp:150-176).  160 algorithmic bytes per cell.
hipError_t launch_stats_finalize(hipStream_t s, int n, StatArrays sums, double n_inv,
                                 StatArrays out);

// Host writers.  Each returns 0, or -1 with a message in err.
struct VtkGrid {
    int nx, ny, nz;
    double dx, dy, dz;
};

// write_field_vtk (io.cpp:225-256) from host arrays.
int vtk_write_fields(const std::string& path, int step, const VtkGrid& g, const double* u,
                     const double* v, const double* w, const double* p, const double* nu_t,
                     const uint8_t* solid, std::string& err);

// write_statistics_vtk (statistics.cpp:11-67): the four means and six
// stresses in StatArrays order, host arrays.
int vtk_write_statistics(const std::string& path, const VtkGrid& g, const double* const* m,
                         std::string& err);

// checkpoint.meta (io.cpp:20-30)
int checkpoint_write_meta(const std::string& path, int next_step, double time_simulated,
                          double dt, std::string& err);

// threads used by the formatters (LES_IO_THREADS, else the hardware's, at most 16)
int io_threads();

}  // namespace les
