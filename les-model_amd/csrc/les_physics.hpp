// les_physics.hpp -- launchers of the physics and stationary-solver kernels.
#pragma once

#include "les_kernels.hpp"

namespace les {

// Host-computed scalars of one step, each formed with the reference's own
// expression (src/main.cpp, src/pressure.cpp) so the device sees the same
// doubles the reference computes.
struct PhysConsts {
    double dt, nu, rho;
    double inv_dx2, inv_dy2, inv_dz2;  // 1/(dx*dx)        main.cpp:569
    double inv_2dx, inv_2dy, inv_2dz;  // 1/(2*dx)         main.cpp:572, 678
    double two_dx, two_dy, two_dz;     // 2*dx (divisor)   main.cpp:436
    double csd2;                       // pow(Cs*Delta,2)  main.cpp:461
    double div_scale;                  // rho/dt           pressure.cpp:260
    double div_ix, div_iy, div_iz;     // 1/(2*dx)         pressure.cpp:263
    double div_ix1, div_iy1, div_iz1;  // 1/dx             pressure.cpp:268
    double dor;                        // dt/rho           main.cpp:677
};

// Optional outputs of the six dead SGS stresses, order xx, xy, xz, yy, yz, zz
// (src/field.h:44-45).  t[0] == nullptr disables them.
struct Tau {
    double* t[6];
};

// Scalars of a stationary sweep loop (Jacobi / SOR).
struct StatScalars {
    double maxr;      // max |dp| of the sweep in progress
    double last_max;  // max |dp| of the last completed sweep
    double thresh;    // 1e-6 (pressure.cpp:165, 240)
    int sweeps;       // sweeps completed
    int done;
    int max_sweeps;
    int pad;
};

// momentum kernel modes
enum { MOM_STEP = 0,      // provisional velocity with BCs + lagged divergence/b
       MOM_RAW = 1,       // solve_momentum alone (interior, no BCs, no divergence)
       MOM_DIV_ONLY = 2 };// calculate_divergence + build_b_vector of the inputs only

// apply_bcs = 0: compute_sgs_stress alone (interior raw values, faces untouched)
hipError_t launch_sgs(hipStream_t s, const Dims& d, const PhysConsts& c, const double* u,
                      const double* v, const double* w, const uint8_t* solid, double* nu_t,
                      Tau tau, int apply_bcs, RedScratch red, double* nut_max);
hipError_t launch_momentum(hipStream_t s, const Dims& d, const PhysConsts& c, const double* u,
                           const double* v, const double* w, const double* nu_t,
                           const uint8_t* solid, double* un, double* vn, double* wn,
                           double* p_rhs, double* b, int mode);
// apply_bcs = 0: correct_velocity alone (all interior cells, faces untouched)
hipError_t launch_correct(hipStream_t s, const Dims& d, const PhysConsts& c, double* u, double* v,
                          double* w, const double* p, const uint8_t* solid, int apply_bcs);
hipError_t launch_velocity_bcs(hipStream_t s, const Dims& d, const uint8_t* solid, double* u,
                               double* v, double* w);
hipError_t launch_nut_bcs(hipStream_t s, const Dims& d, const uint8_t* solid, double* nu_t);
hipError_t launch_build_b(hipStream_t s, const Dims& d, const double* div, const uint8_t* solid,
                          double* b);
hipError_t launch_dt_max(hipStream_t s, const Dims& d, const double* u, const double* v,
                         const double* w, const double* nu_t, RedScratch red, double* out2);
// st may be nullptr (plain apply_pressure_BCs); otherwise the sweep check runs.
hipError_t launch_pressure_bcs(hipStream_t s, const Dims& d, double* p, const uint8_t* solid,
                               StatScalars* st, int sweep);
hipError_t launch_rhs_central(hipStream_t s, const Dims& d, const double* u, const double* v,
                              const double* w, double* rhs, double coef, double two_dx,
                              double two_dy, double two_dz);
hipError_t launch_jacobi_sweep(hipStream_t s, const Dims& d, const double* pc, double* pn,
                               const double* rhs, double dx2, double dy2, double dz2,
                               double denom, StatScalars* st);
hipError_t launch_sor_level(hipStream_t s, const Dims& d, int level, double* p, const double* rhs,
                            double one_m_omega, double omega_coef, double dx2, double dy2,
                            double dz2, StatScalars* st);
hipError_t launch_sor_colour(hipStream_t s, const Dims& d, double* p, const double* rhs,
                             int colour, double one_m_omega, double omega_coef, double dx2,
                             double dy2, double dz2, StatScalars* st);

}  // namespace les
