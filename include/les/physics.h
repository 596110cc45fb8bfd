// les/physics.h -- the per-step physics functions the reference declares in
// src/main.cpp:20-29, same signatures, implemented on the GPU by
// liblespressure_gpu.so.  compute_sgs_stress keeps the reference's fixed
// Smagorinsky constant 0.18 (main.cpp:425); les_sgs in include/les_gpu.h
// takes it as a parameter.
#ifndef LES_PHYSICS_H
#define LES_PHYSICS_H

#include "field.h"

double compute_adaptive_dt(const Field& field, const Mesh& mesh, double nu);
void compute_sgs_stress(struct Field& f, struct Stress& sgs, const struct Mesh& m);
void apply_nut_BCs(Field& f, const Mesh& m);
void solve_momentum(struct Field &f, const struct Mesh &m, double dt, double nu);
void apply_velocity_BCs(Field& f, const Mesh &m);
void correct_velocity(struct Field &f, const struct Mesh &m, double dt, double rho);

#endif  // LES_PHYSICS_H
