// les_adjoint.hpp -- launchers of the transposed step pieces used by the
// reverse (adjoint) pass of the LES step map.  The reference has no
// counterpart; each kernel is the exact transpose of one forward piece.
#pragma once

#include "les_physics.hpp"

namespace les {

// The six independent entries of the SGS strain sensitivity at each cell,
// order xx, yy, zz, xy, xz, yz.
struct StrainSens {
    double* g[6];
};

// out = BC^T in for apply_velocity_BCs (main.cpp:323-410) viewed as a map of
// the velocity it is applied to.
hipError_t launch_adj_velocity_bcs(hipStream_t s, const Dims& d, const uint8_t* solid,
                                   const double* iu, const double* iv, const double* iw,
                                   double* ou, double* ov, double* ow);
// rhs = W^-1 (d u/d p)^T y for correct_velocity (main.cpp:671-714);
// W == nullptr means W = 1.
hipError_t launch_adj_correct(hipStream_t s, const Dims& d, const PhysConsts& c, const double* yu,
                              const double* yv, const double* yw, const double* W, double* rhs);
// g = (d momentum / d u)^T m + (d b / d u)^T (W x) + the nu_t path up to the
// strain sensitivity G (gathered by launch_adj_sgs_gather).
hipError_t launch_adj_momentum(hipStream_t s, const Dims& d, const PhysConsts& c,
                               const double* mu, const double* mv, const double* mw,
                               const double* u, const double* v, const double* w,
                               const double* nu_t, const double* x, const double* W,
                               const uint8_t* solid, double* gu, double* gv, double* gw,
                               StrainSens G);
// g += (d nu_t / d u)^T through the gradients of compute_sgs_stress.
hipError_t launch_adj_sgs_gather(hipStream_t s, const Dims& d, const PhysConsts& c, StrainSens G,
                                 double* gu, double* gv, double* gw);

}  // namespace les
