// les/pressure.h -- the pressure-solver entry points of the reference
// (src/pressure.h:11-63), same names, signatures and default arguments,
// implemented on the GPU by liblespressure_gpu.so over include/les_gpu.h.
//
// Semantics follow the reference function for function: the same
// arithmetic, in-place outputs, iteration counts (max_iter plus a warning on
// std::cerr when PCG does not converge, pressure.cpp:411-413) and the
// "Pressure converged" line of solve_pressure_GS (pressure.cpp:242).
// Device memory is managed per grid size inside the library; every call
// copies its inputs to the device and its outputs back.
#ifndef PRESSURE_H
#define PRESSURE_H

#include "field.h"

#include <cmath>
#include <iostream>
#include <vector>

using Vector = std::vector<double>;

// The 7-point operator: A_c centre, A_n (j+1), A_s (j-1), A_e (i+1),
// A_w (i-1), A_t (k+1), A_b (k-1) per cell.
struct StencilMatrix
{
    Vector A_c, A_n, A_s, A_e, A_w, A_t, A_b;

    StencilMatrix(int n_nodes)
        : A_c(n_nodes, 0.0), A_n(n_nodes, 0.0), A_s(n_nodes, 0.0),
          A_e(n_nodes, 0.0), A_w(n_nodes, 0.0), A_t(n_nodes, 0.0),
          A_b(n_nodes, 0.0)
    {}
};

void apply_pressure_BCs(struct Field &f, const struct Mesh &m);
void solve_pressure_jacobi(struct Field &f, const struct Mesh &m, double dt, double rho);
void solve_pressure_GS(struct Field &f, const struct Mesh &m, double dt, double rho, int maxIter=200);
int solve_pressure_pcg(const StencilMatrix& A, const Vector& b, Vector& x,
              const Field& field, int max_iter, double tol);

void calculate_divergence(Vector& divergence, const Field& f, const Mesh& m, double dt, double rho);
void build_stencil(StencilMatrix& A, const Field& field, const Mesh& m);
void build_b_vector(Vector& b, const Vector& divergence,
                    const Field& field, const Mesh& m);
void fast_apply_laplacian(Vector& q, const Vector& p,
                          const StencilMatrix& A, const Field& field);

void apply_jacobi_preconditioner(Vector& z, const Vector& r, const StencilMatrix& A);
void apply_ic_preconditioner(Vector& z, const Vector& r, const StencilMatrix& L, const Field &field);
void build_ic_preconditioner(StencilMatrix& L, const StencilMatrix& A, const Field& field, const Mesh& m);

int solve_pressure_pcg_chebyshev(const StencilMatrix& A, const Vector& b, Vector& x,
                      const Field& field, double lambda_min, double lambda_max,
                      int poly_degree = 5, int maxIter = 100, double tol = 1e-6);
void apply_chebyshev_preconditioner(Vector& z, const Vector& r,
                                    const StencilMatrix& A, const Field& field,
                                    double lambda_min, double lambda_max,
                                    int poly_degree = 5);
double estimate_lambda_max(const StencilMatrix& A, const Field& field, int n_iter = 20);

void parallel_axpy(Vector& v_out, double a, const Vector& v_in);
double parallel_norm(const Vector& v);
double parallel_dot(const Vector& v1, const Vector& v2);
void parallel_saxpy(Vector& v_out, double a, const Vector& v_in1,
                    double b, const Vector& v_in2);
void parallel_copy(Vector& v_out, const Vector& v_in);

#endif  // PRESSURE_H
