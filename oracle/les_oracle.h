/*
 * les_oracle.h -- CPU oracle for the LES per-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * solver's arithmetic (GabrieLeitao/les-model, src/pressure.cpp and
 * src/main.cpp), written to be the checker for the HIP library.  Only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark may load
 * it; the product path never links or calls it.
 *
 * Layout follows src/field.h:40: idx = i*ny*nz + j*nz + k (k contiguous).
 * Stencil arrays follow the member order of StencilMatrix
 * (src/pressure.h:22): c, n, s, e, w, t, b.
 *
 * Every function reproduces the reference's operand order so that, built
 * without FP contraction and run on one thread, its results are bit-identical
 * to the reference's.  Pinning: see the header comment of les_oracle.c.
 */
#ifndef LES_ORACLE_H
#define LES_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { int nx, ny, nz; double dx, dy, dz; } orc_mesh;

enum { ORC_C = 0, ORC_N = 1, ORC_S = 2, ORC_E = 3, ORC_W = 4, ORC_T = 5, ORC_B = 6 };

enum {
    ORC_PC_NONE = 0,       /* plain CG (z = r) */
    ORC_PC_JACOBI = 1,     /* z = r / A_c (pressure.cpp:420) */
    ORC_PC_IC0 = 2,        /* z = (L L^T)^-1 r with L from build_ic (pressure.cpp:442, 495) */
    ORC_PC_LEGACY_IC_A = 3,/* apply_ic_preconditioner(., ., A) as shipped (pressure.cpp:363, 396) */
    ORC_PC_CHEBYSHEV = 4,  /* apply_chebyshev_preconditioner (pressure.cpp:851) */
    ORC_PC_MG = 5          /* multigrid V-cycle (no reference counterpart; README.md:53 TODO) */
};

void orc_set_threads(int n);

/* ---- BLAS-1 (pressure.cpp:962-1013) ---- */
double orc_dot(const double* a, const double* b, long n);
double orc_norm(const double* a, long n);
void orc_axpy(double* out, double a, const double* in, long n);
void orc_saxpy(double* out, double a, const double* in1, double b, const double* in2, long n);
void orc_copy(double* out, const double* in, long n);

/* ---- operator (pressure.cpp:608-746) ---- */
void orc_build_stencil(double* const A[7], const uint8_t* solid, const orc_mesh* m);
void orc_matvec(double* q, const double* p, double* const A[7], int nx, int ny, int nz);

/* ---- RHS (pressure.cpp:256-329, 566-599) ---- */
void orc_calculate_divergence(double* rhs, const double* un, const double* vn, const double* wn,
                              const uint8_t* solid, const orc_mesh* m, double dt, double rho);
void orc_build_b_vector(double* b, const double* div, const uint8_t* solid, const orc_mesh* m);

/* ---- preconditioners (pressure.cpp:420-556, 851-906) ---- */
void orc_jacobi_precond(double* z, const double* r, const double* Ac, long n);
void orc_ic_apply(double* z, const double* r, double* const L[7], int nx, int ny, int nz);
void orc_ic_build(double* const L[7], double* const A[7], int nx, int ny, int nz);
void orc_cheb_apply(double* z, const double* r, double* const A[7], int nx, int ny, int nz,
                    double lmin, double lmax, int degree);
/* Power iteration (pressure.cpp:917-955) with the mt19937(1337) start vector
 * drawn sequentially (the reference is deterministic only on one thread). */
double orc_estimate_lambda_max(double* const A[7], int nx, int ny, int nz, int n_iter);

/* ---- PCG (pressure.cpp:342-414, 762-835) ----
 * L is used only for ORC_PC_IC0.  res_trace (may be NULL) receives ||r_k||
 * for k = 0..iters (max_iter+1 entries).  Returns the iteration count with
 * the reference's conventions (0 if already converged, max_iter if not). */
int orc_pcg(double* const A[7], const double* b, double* x, int nx, int ny, int nz,
            int precond, double* const L[7], double lmin, double lmax, int degree,
            int max_iter, double tol, double* res_trace);
/* the same with r.z and p.q in the W-weighted inner product (W may be NULL) */
int orc_pcg_w(double* const A[7], const double* b, double* x, int nx, int ny, int nz,
              int precond, double* const L[7], double lmin, double lmax, int degree,
              int max_iter, double tol, double* res_trace, const double* W);

/* ---- multigrid preconditioner (no reference counterpart: the reference's
 * README.md:53 lists it as a TODO) ----
 * Galerkin aggregation multigrid on the 7-point operator.  Level l+1 merges
 * 2x2x2 blocks of level l (ceil(n/2) cells per direction); prolongation P is
 * piecewise constant, restriction is P^T (a sum over the block), and the
 * coarse operator is P^T A P, again a 7-point stencil: a fine coupling inside
 * a block adds to the coarse centre, one across a block face to that face's
 * coefficient.  Levels are added while the largest dimension exceeds
 * coarse_max.  The preconditioner is one V-cycle from x = 0: on every level
 * but the coarsest, `sweeps` damped-Jacobi sweeps x += omega (b - A x) / A_c
 * (the first from x = 0 is x = omega (b / A_c)), the residual restricted,
 * the coarse correction times `scale` added, `sweeps` more sweeps; on the
 * coarsest, coarse_sweeps sweeps from 0.  A fixed linear operator, so PCG
 * can use it.  The summation orders (children in (di, dj, dk) order, a
 * child's couplings in StencilMatrix member order) are part of the
 * definition: the GPU follows them bit for bit.
 * With mirror weights W (W A symmetric: A without rows mirrored around
 * interior obstacles), the cycle is the symmetric one of S = W A applied to
 * W r: the level-0 residual is restricted as P^T (W (b - A x)) and level 1
 * is P^T W A P (each fine coefficient times W of its row); damped Jacobi on
 * A is damped Jacobi on S.  PCG in the W inner product with it is then CG
 * on S with a symmetric preconditioner. */
enum { ORC_MG_MAX_LEVELS = 16 };
typedef struct {
    int sweeps, coarse_sweeps, coarse_max;
    double omega, scale;
} orc_mg_params;
/* the built-in parameters (sweeps 2, coarse_sweeps 32, coarse_max 4,
 * omega 0.85, scale 1.8) */
orc_mg_params orc_mg_default_params(void);
/* level count and the dimensions of every level (dims[3*l + 0..2]) */
int orc_mg_levels(int nx, int ny, int nz, int coarse_max, int* dims);
/* Ac = P^T A P of one level (nx, ny, nz the fine dimensions); with W
 * (may be NULL) P^T W A P */
void orc_mg_coarsen(double* const Ac[7], double* const A[7], const double* W, int nx, int ny,
                    int nz);
/* z = V-cycle(r) on the hierarchy built from A (built, applied, freed); W
 * NULL or the mirror weights */
void orc_mg_apply(double* z, const double* r, double* const A[7], int nx, int ny, int nz,
                  const orc_mg_params* prm, const double* W);
/* PCG with the multigrid preconditioner (orc_pcg's conventions), in the W
 * inner product when W != NULL */
int orc_pcg_mg(double* const A[7], const double* b, double* x, int nx, int ny, int nz,
               const orc_mg_params* prm, int max_iter, double tol, double* res_trace,
               const double* W);

/* ---- stationary solvers (pressure.cpp:14-247) ---- */
void orc_apply_pressure_bcs(double* p, const uint8_t* solid, const orc_mesh* m);
int orc_solve_pressure_jacobi(double* p, const double* u, const double* v, const double* w,
                              const uint8_t* solid, const orc_mesh* m, double dt, double rho);
int orc_solve_pressure_gs(double* p, const double* u, const double* v, const double* w,
                          const uint8_t* solid, const orc_mesh* m, double dt, double rho,
                          int max_iter, double omega);
/* Red-black SOR: a different iteration with no counterpart in the reference
 * (same update formula as pressure.cpp:224-229, colour (i+j+k)%2 order). */
int orc_solve_pressure_rbsor(double* p, const double* u, const double* v, const double* w,
                             const uint8_t* solid, const orc_mesh* m, double dt, double rho,
                             int max_iter, double omega);

/* ---- physics (main.cpp:208-714) ---- */
double orc_adaptive_dt(const double* u, const double* v, const double* w, const double* nu_t,
                       const orc_mesh* m, double nu);
/* (Cs*Delta)^2 exactly as main.cpp:426,461 form it. */
double orc_cs_delta_sq(const orc_mesh* m, double Cs);
void orc_sgs(double* nu_t, double* const tau[6], const double* u, const double* v, const double* w,
             const orc_mesh* m, double Cs);
void orc_nut_bcs(double* nu_t, const uint8_t* solid, const orc_mesh* m);
/* Writes the interior of (un,vn,wn) from (u,v,w); the caller swaps. */
void orc_momentum(double* un, double* vn, double* wn, const double* u, const double* v,
                  const double* w, const double* nu_t, const orc_mesh* m, double dt, double nu);
void orc_velocity_bcs(double* u, double* v, double* w, const uint8_t* solid, const orc_mesh* m);
void orc_correct(double* u, double* v, double* w, const double* p, const orc_mesh* m,
                 double dt, double rho);

/* ---- one LES step in the order of main.cpp:115-144 ---- */
typedef struct {
    double *u, *v, *w, *u_new, *v_new, *w_new, *p, *p_rhs, *nu_t;
    uint8_t* solid;
} orc_fields;

typedef struct {
    double nu, rho, Cs;
    int step;            /* dt = 1e-3 while step <= 20, else adaptive (main.cpp:115-119) */
    double dt_fixed;     /* > 0 overrides the schedule */
    int precond, degree, max_iter;
    double lmin, lmax;
    double tol;          /* absolute, or relative to ||b|| when tol_rel != 0 */
    int tol_rel;
    double* const* L;    /* IC(0) factor for ORC_PC_IC0 */
    double* const* A;    /* the stencil */
    const double* W;     /* NULL: the reference's PCG; else W-weighted inner products */
} orc_step_params;

typedef struct { double dt; int iters; double bnorm; double res; } orc_step_stats;

void orc_step(orc_fields* f, const orc_mesh* m, const orc_step_params* prm, orc_step_stats* st);

/* ---- differentiation (no reference counterpart) ----
 * The step map F: u^n -> u^{n+1} of orc_step with a fixed dt, the pressure
 * solved to a tight relative tolerance (treated as exact), and no solid
 * cells on the inlet plane (their stale ping-pong values are not a function
 * of u^n).  orc_rollout runs nsteps steps from (u,v,w) and records u^n for
 * n = 0..nsteps-1 in traj (3*N doubles per step, may be NULL).
 * Each step's pressure solve starts from p = 0 (same exact map as the warm
 * start of main.cpp:137, but a tight relative tolerance stays reachable).
 * orc_adjoint takes (ubar, vbar, wbar) = dL/du^{nsteps} and returns
 * dL/du^0 in place, re-using traj. */
typedef struct {
    double nu, rho, Cs, dt;
    int nsteps;
    int precond;         /* ORC_PC_JACOBI or ORC_PC_CHEBYSHEV */
    double lmin, lmax;
    int degree, max_iter;
    double tol_rel;      /* relative tolerance of the forward and adjoint solves */
    int weighted;        /* CG in the mirror-weight inner product (obstacle-free grids) */
} orc_adj_params;

/* W = 1/2 per mirrored direction of each row; W A is symmetric on grids
 * without obstacles. */
void orc_mirror_weights(double* W, double* const A[7], const orc_mesh* m);

void orc_transpose_stencil(double* const AT[7], double* const A[7], int nx, int ny, int nz);
void orc_rollout(double* u, double* v, double* w, double* p, const uint8_t* solid,
                 const orc_mesh* m, double* const A[7], const orc_adj_params* prm, double* traj);
void orc_adjoint(double* ubar, double* vbar, double* wbar, const double* traj,
                 const uint8_t* solid, const orc_mesh* m, double* const A[7],
                 const orc_adj_params* prm);

/* the pieces of orc_adjoint, each the transpose of one forward piece */
void orc_adj_velocity_bcs(double* bu, double* bv, double* bw, const uint8_t* solid,
                          const orc_mesh* m);
void orc_adj_correct(double* pbar, const double* yu, const double* yv, const double* yw,
                     const orc_mesh* m, double dt, double rho);
void orc_adj_divergence(double* bu, double* bv, double* bw, const double* bbar,
                        const uint8_t* solid, const orc_mesh* m, double dt, double rho);
void orc_adj_momentum(double* bu, double* bv, double* bw, double* bnut, const double* mu,
                      const double* mv, const double* mw, const double* u, const double* v,
                      const double* w, const double* nu_t, const orc_mesh* m, double dt,
                      double nu);
void orc_adj_nut_bcs(double* braw, const double* bnut, const uint8_t* solid, const orc_mesh* m);
void orc_adj_sgs(double* bu, double* bv, double* bw, const double* braw, const double* u,
                 const double* v, const double* w, const orc_mesh* m, double Cs);

/* ---- time statistics (statistics.cpp:98-131, 133-176) ----
 * sums[10] in the order u, v, w, p, uu, vv, ww, uv, uw, vw. */
void orc_stats_update(double* const sums[10], const double* u, const double* v, const double* w,
                      const double* p, long n);
/* means (slots 0-3) and Reynolds stresses (4-9) from the sums of k samples */
void orc_stats_finalize(double* const out[10], double* const sums[10], int num_samples, long n);

#ifdef __cplusplus
}
#endif
#endif
