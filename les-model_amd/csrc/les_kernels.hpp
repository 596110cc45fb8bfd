// les_kernels.hpp -- host-callable launchers of the LES HIP kernels.
//
// Types here are shared by the kernel translation units and by the C-ABI
// layer (les_api.hip).  Every launcher enqueues on the given stream and
// returns the hipError_t of the launch.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "les_device.hpp"

namespace les {

// Seven stencil coefficient arrays in the member order of StencilMatrix
// (src/pressure.h:22): c, n, s, e, w, t, b.
enum { SC = 0, SN = 1, SS = 2, SE = 3, SW = 4, ST = 5, SB = 6 };

// The operator, either as the 7 coefficient arrays or -- for the operator
// build_stencil produces from a mask -- as one 16-bit code per cell.  Every
// off-diagonal of that operator is 0, -1/d^2 or 2*(-1/d^2) (a mirrored
// face adds the base coefficient to itself, pressure.cpp:662-696) and the
// centre is 2/dx^2+2/dy^2+2/dz^2 or 1 (identity rows, pressure.cpp:634-642),
// so the code reproduces the arrays bit for bit:
//   bits 0-1 w, 2-3 e, 4-5 s, 6-7 n, 8-9 b, 10-11 t: multiplier 0, 1 or 2
//   bit 12: identity row (centre 1)
template <typename T>
struct Stencil {
    const T* a[7];
    const uint16_t* code;  // non-null: coefficients come from the code
    T cfluid;              // centre of a fluid row
    T ox, oy, oz;          // -1/dx^2, -1/dy^2, -1/dz^2; a doubled face is 2*ox ...
};

enum { CODE_IDENTITY = 1 << 12 };

// Scalars of one Krylov solve, resident in device memory so that the
// iteration never needs the host.
struct SolveScalars {
    double rho;      // r.z of the current iterate
    double alpha;
    double beta;
    double rr;       // ||r||^2 of the latest residual
    double bb;       // ||b||^2
    double tol;      // absolute stopping tolerance on ||r||
    double tol_rel;  // > 0: tol = tol_rel * ||b|| (set by the init kernel)
    double err;      // ||r|| of the latest residual
    double err0;     // ||r|| at iteration 0
    double pq;
    int iter;        // iterations completed
    int done;        // 1 once converged or out of iterations
    int converged;
    int max_iter;
    int fault;       // 1 if a non-finite scalar was met
    // mixed precision (FP32 iteration, FP64 reliable residual replacement)
    int reliable_every;  // replace r by the FP64 residual every this many iterations
    double reliable_drop; // > 0: also once ||r|| < reliable_drop * the last true residual
    int reliable_now;    // 1: fold x32 into x64 and replace r in this iteration
    int stop_pending;    // 1: stop after the pending replacement (max_iter reached)
    int replacements;
    int restart_pending; // 1: the next beta is 0 (p restarts from z)
    int restarts;
    double true_err;     // ||b - A x64|| of the last replacement
    double rscale;       // mixed: power of two the FP32 residual carries (set at
                         // iteration 0 from ||b - A x0||, so that ||r32|| ~ 1:
                         // no FP32 overflow or subnormals whatever b's scale)
    double* trace;       // optional: ||r_k|| for k = 0..max_iter
};

// Per-workgroup partials of a reduction (kMaxGrid * kMaxRed doubles); a
// null pointer means the kernel reduces nothing.
struct RedScratch {
    double* partials;
};

// When a gated kernel (and its finalize) runs.
enum Gate {
    GATE_NONE = 0,      // always
    GATE_DONE = 1,      // unless the solve is done
    GATE_RELIABLE = 2,  // only when a reliable replacement is pending
};

// ---- scalar epilogues run by the last workgroup of a reduction ----
enum Epilogue {
    EPI_STORE = 0,       // out[r] = tot[r]
    EPI_INIT = 1,        // tot = {rr, bb}: err0, tol, done if err0 < tol
    EPI_RHO = 2,         // tot = {rz}: rho = rz
    EPI_ALPHA = 3,       // tot = {pq}: alpha = rho / pq
    EPI_CONV = 4,        // tot = {rr}: iter++, err, done if err < tol or iter == max_iter
    EPI_CONV_RZ = 5,     // tot = {rr, rz}: EPI_CONV, then beta = rz/rho, rho = rz
    EPI_BETA = 6,        // tot = {rz}: beta = rz/rho, rho = rz
    EPI_CONV_MIXED = 7,  // tot = {rr}: iter++, err; schedule an FP64 replacement when
                         // err < tol, iter % reliable_every == 0, err has fallen below
                         // reliable_drop * true_err, or iter == max_iter
    EPI_RELIABLE = 8,    // tot = {rr64}: true residual; done if below tol
    EPI_INIT_MIXED = 9,  // tot = {rr64, bb}: as EPI_INIT, also true_err
    EPI_STORE2 = 10,     // out[0..1] = tot[0..1]
    EPI_STORE_SQRT = 11, // out[0] = sqrt(tot[0])
    EPI_PARTIALS = 12,   // no finalize: the partials stay for a consumer that sums them
                         // itself (take_partials_grid() returns their count)
};

// The workgroup count of the last reduction launched with EPI_PARTIALS on
// this host thread (0 if none since the previous call); resets it.
int take_partials_grid();

// The one-workgroup tail of a reduction over `grid` workgroup partials.
hipError_t launch_finalize(hipStream_t s, int grid, int nr, bool max, int epi, int gate,
                           RedScratch red, SolveScalars* sc, double* out);

// ---- operator ----
hipError_t launch_build_code(hipStream_t s, const Dims& d, const uint8_t* solid, uint16_t* code);
// W = 1/2 per mirrored direction of each coded row (W A symmetric without obstacles)
hipError_t launch_mirror_weights(hipStream_t s, const Dims& d, const uint16_t* code, double* W);
hipError_t launch_build_stencil(hipStream_t s, const Dims& d, const uint8_t* solid,
                                double* const A[7], double idx2, double idy2, double idz2);
template <typename T>
hipError_t launch_matvec(hipStream_t s, const Dims& d, Stencil<T> A, const T* x, T* y);
template <typename TO, typename TI>
hipError_t launch_convert(hipStream_t s, int n, const TI* in, TO* out);
hipError_t launch_dot(hipStream_t s, const Dims& d, const double* a, const double* b,
                      RedScratch red, double* out);

// ---- PCG building blocks (pressure.cpp:342-414), fused ----
// r = b - A x; ||r||^2, ||b||^2  (pressure.cpp:355-359)
template <typename T>
hipError_t launch_pcg_init(hipStream_t s, const Dims& d, Stencil<T> A, const T* b, const T* x,
                           T* r, RedScratch red, SolveScalars* sc);
// q = A p; alpha = rho / (p.q)  (pressure.cpp:375-378)
// wgt (may be nullptr): per-cell weights of the inner products r.z and p.q
// (mirror weights W; the stopping test stays the Euclidean ||r||).
template <typename T>
hipError_t launch_pcg_matvec(hipStream_t s, const Dims& d, Stencil<T> A, const T* p, T* q,
                             RedScratch red, SolveScalars* sc, const double* wgt = nullptr);
// The same with the previous iteration's direction update fused in:
// p = z + beta p_old (diag == nullptr) or p = r/diag + beta p_old (z = r,
// Jacobi), formed at every stencil point and written to p (distinct from
// p_old); q = A p.  x non-null: also the previous iteration's deferred
// x += alpha p_old.
// grid_out non-null: no finalize runs; the launch's partials stay in red and
// their count goes to *grid_out (the consumer forms alpha itself).
template <typename T>
hipError_t launch_pcg_matvec_pu(hipStream_t s, const Dims& d, Stencil<T> A, const T* p_old,
                                const T* z, const T* diag, T* p, T* q, RedScratch red,
                                SolveScalars* sc, const double* wgt = nullptr, T* x = nullptr,
                                int* grid_out = nullptr, RedScratch rz_red = RedScratch{nullptr},
                                int rz_grid = 0);
// x += alpha p_last after the last iteration when x's update is deferred to
// the matvec (p_last = p0 for even iterations, p1 for odd; nothing when no
// iteration ran)
template <typename T>
hipError_t launch_pcg_x_pending(hipStream_t s, const Dims& d, T* x, const T* p0, const T* p1,
                                SolveScalars* sc);
// x += alpha p; r -= alpha q; ||r|| test (pressure.cpp:381-392); x null:
// r only (x deferred to the next matvec).  With jacobi_diag != nullptr also
// r.(r/A_c) and beta (pressure.cpp:396-402).
template <typename T, typename TX>
hipError_t launch_pcg_update(hipStream_t s, const Dims& d, TX* x, T* r, const T* p, const T* q,
                             const T* jacobi_diag, RedScratch red, SolveScalars* sc,
                             const double* wgt = nullptr);
// p = M^-1 r + beta p for Jacobi (beta ignored and p overwritten when init);
// init also reduces rho = r.p (pressure.cpp:363-369, 395-405).
// When init, z = r / A_c is written into p and r.z reduced with epilogue
// epi (< 0: no reduction; sc may be nullptr then).
template <typename T>
hipError_t launch_pcg_p_jacobi(hipStream_t s, const Dims& d, const T* r, const T* diag, T* p,
                               int init, int epi, RedScratch red, SolveScalars* sc,
                               const double* wgt = nullptr);
// p = z + beta p (pressure.cpp:405)
template <typename T>
hipError_t launch_pcg_p_update(hipStream_t s, const Dims& d, const T* z, T* p, SolveScalars* sc);
// rho / beta from r.z (pressure.cpp:369, 399-402)
template <typename T>
hipError_t launch_pcg_rz(hipStream_t s, const Dims& d, const T* r, const T* z, int epilogue,
                         RedScratch red, SolveScalars* sc, const double* wgt = nullptr);

// ---- Chebyshev preconditioner (pressure.cpp:851-906) ----
// One inner step k >= 1 of the recurrence:
//   res = r - A z_k ; p = res + beta p ; z_{k+1} = z_k + alpha p
// first: z_1 = 0 + alpha0 r and p_0 = r are formed on the fly from r.
// degree 1: z = 0 + alpha0 r only (launch_cheb_scale).
// When epi >= 0 the kernel also reduces r.z_{k+1} with that epilogue.
template <typename T>
hipError_t launch_cheb_step(hipStream_t s, const Dims& d, Stencil<T> A, const T* r,
                            const T* z_in, T* z_out, T* p, double alpha0, double beta,
                            double alpha, int first, int epi, RedScratch red, SolveScalars* sc,
                            const double* wgt = nullptr);
// Two Chebyshev steps in one pass (k_cheb_pair): steps (beta_a, alpha_a) and
// (beta_b, alpha_b); p_in/p_out and z_in/z_out distinct; no reduction.  Not
// every grid shape is supported (cheb_pair_supported; the geometry depends
// on whether A is coded); the caller then runs two launch_cheb_step.
// target_grid: about how many workgroups to spread the grid over (0: the
// built-in default).
template <typename T>
bool cheb_pair_supported(const Dims& d, int target_grid, bool coded);
template <typename T>
hipError_t launch_cheb_pair(hipStream_t s, const Dims& d, Stencil<T> A, const T* r,
                            const T* z_in, T* z_out, const T* p_in, T* p_out, double alpha0,
                            double beta_a, double alpha_a, double beta_b, double alpha_b,
                            int first, SolveScalars* sc, int target_grid);
template <typename T>
hipError_t launch_cheb_scale(hipStream_t s, const Dims& d, const T* r, T* z, double alpha0,
                             int epi, RedScratch red, SolveScalars* sc,
                             const double* wgt = nullptr);

// the cycle's bottom levels run in one workgroup up to this many cells each
// (one per thread)
constexpr int kMgTailCells = 1024;
constexpr int kMgTailLevels = 8;
// ---- multigrid preconditioner (no reference counterpart: the reference's
// README.md:53 lists it as a TODO) ----
// Galerkin aggregation over 2x2x2 blocks; the level dimensions halve,
// rounding up.  Ac = P^T A P (FP64; f the fine dimensions, cd the coarse);
// with mirror weights W, P^T W A P.
hipError_t launch_mg_coarsen(hipStream_t s, const Dims& f, const Dims& cd, Stencil<double> A,
                             const double* W, double* const Ac[7]);
// One damped-Jacobi sweep on a level (mode MG_ZERO: from x = 0; MG_SWEEP;
// MG_PROLONG: x first corrected by scale * e of its block on the coarser
// level cd; MG_ZS: the first two sweeps from x = 0 in one pass, x unused
// where the level takes vector loads, else the first sweep's output).
// epi >= 0: b.x_out reduced with that epilogue (the PCG's r.z, in the wgt
// inner product when wgt is given).
enum { MG_ZERO = 0, MG_SWEEP = 1, MG_PROLONG = 2, MG_ZS = 3 };
// ring: MG_ZS may run as k_mg_zs_ring where its geometry fits.
template <typename T>
hipError_t launch_mg_smooth(hipStream_t s, const Dims& d, Stencil<T> A, const T* x, const T* b,
                            const T* e, const Dims& cd, T* x_out, int mode, int epi,
                            double omega, double scale, RedScratch red, SolveScalars* sc,
                            const double* wgt = nullptr, bool ring = true);
// The two post-smoothing sweeps (MG_PROLONG, then a plain sweep) in one
// pass where mg_post_fits(d) (FP32 levels whose rows take vector loads);
// x_out must differ from x.  epi as launch_mg_smooth's.
template <typename T>
bool mg_post_fits(const Dims& d);
template <typename T>
hipError_t launch_mg_post(hipStream_t s, const Dims& d, Stencil<T> A, const T* x, const T* b,
                          const T* e, const Dims& cd, T* x_out, int epi, double omega,
                          double scale, RedScratch red, SolveScalars* sc,
                          const double* wgt = nullptr);
// The two pre-smoothing sweeps from x = 0 (into x_out) and the restriction
// bc = P^T (b - A x_out), or P^T W (b - A x_out), in one pass where
// mg_pre_fits(f) (FP32 levels whose rows take vector loads).
template <typename T>
bool mg_pre_fits(const Dims& f);
template <typename T>
hipError_t launch_mg_pre(hipStream_t s, const Dims& f, const Dims& cd, Stencil<T> A, const T* b,
                         T* x_out, T* bc, double omega, SolveScalars* sc,
                         const double* W = nullptr);
// bc = P^T (b - A x), or P^T W (b - A x); ring: as k_mg_restrict_ring where
// its geometry fits
template <typename T>
hipError_t launch_mg_restrict(hipStream_t s, const Dims& f, const Dims& cd, Stencil<T> A,
                              const T* x, const T* b, T* bc, SolveScalars* sc,
                              const double* W = nullptr, bool ring = true);
// Levels dims[0..nlev-1] of the cycle's bottom (each at most kMgTailCells cells,
// arrays format A[l][7]) in one workgroup: from b0 (level 0 of the tail's
// right-hand side) to x0 (its result); the same arithmetic as the
// per-level launches.
template <typename T>
hipError_t launch_mg_tail(hipStream_t s, int nlev, const Dims* dims, T* const* const* A,
                          const T* b0, T* x0, int sweeps, int coarse_sweeps, double omega,
                          double scale, SolveScalars* sc);
// the coarsest level's sweeps from x = 0 in one workgroup (d.N <= kBlock,
// arrays format)
template <typename T>
hipError_t launch_mg_coarsest(hipStream_t s, const Dims& d, Stencil<T> A, const T* b, T* x,
                              int sweeps, double omega, SolveScalars* sc);

// ---- IC(0) (pressure.cpp:442-556), i+j+k hyperplane order ----
hipError_t launch_ic_factor_level(hipStream_t s, const Dims& d, int level, Stencil<double> A,
                                  double* Lc, double* Lw, double* Ls, double* Lb);
template <typename T>
hipError_t launch_ic_fwd_level(hipStream_t s, const Dims& d, int level, const T* r, T* z,
                               const T* Lc, const T* Lw, const T* Ls, const T* Lb,
                               const SolveScalars* sc);
template <typename T>
hipError_t launch_ic_bwd_level(hipStream_t s, const Dims& d, int level, T* z, const T* Lc,
                               const T* Lw, const T* Ls, const T* Lb, const SolveScalars* sc);

// ---- mixed precision (no reference counterpart) ----
// x64 += x32 + alpha p_last; x32 = -alpha p_last (when force or
// sc->reliable_now): the fold with the last update's deferred x term
hipError_t launch_mixed_fold(hipStream_t s, const Dims& d, double* x64, float* x32,
                             const float* p0, const float* p1, int force, SolveScalars* sc);
// r64 = b - A64 x64 -> r32 = (float) r64, ||r64||^2 (and ||b||^2 when init);
// runs when force, init or sc->reliable_now.  Epilogue EPI_INIT_MIXED when
// init, else EPI_RELIABLE (force && !init: EPI_STORE of ||r64|| into out).
hipError_t launch_mixed_residual(hipStream_t s, const Dims& d, Stencil<double> A, const double* b,
                                 const double* x64, float* r32, int init, int force,
                                 RedScratch red, SolveScalars* sc, double* out);
// FP32 update with the mixed convergence epilogue.  pq_grid > 0: alpha =
// rho / (p.q) is formed here from the matvec's pq_grid partials in pq_red
// (the sum and epilogue of its finalize, in every workgroup; workgroup 0
// stores pq and alpha), so the matvec needs no finalize launch.
// With rz_grid > 0 the preconditioner's r.z reduction was left as rz_grid
// partials in rz_red (EPI_PARTIALS): rho is their sum, and workgroup 0 runs
// EPI_BETA's bookkeeping (the matvec formed beta from the same partials).
hipError_t launch_pcg_update_mixed(hipStream_t s, const Dims& d, float* x, float* r, const float* p,
                                   const float* q, RedScratch red, SolveScalars* sc,
                                   RedScratch pq_red = RedScratch{nullptr}, int pq_grid = 0,
                                   RedScratch rz_red = RedScratch{nullptr}, int rz_grid = 0);

// ---- power iteration helpers (pressure.cpp:917-955) ----
// v = w / sqrt(tot) where tot = ||w||^2 is read from device memory
hipError_t launch_scale_by_invnorm(hipStream_t s, int n, const double* w, double* v,
                                   const double* nrm2);
// y = S x, S the geometric-mean symmetrisation of A (same spectrum)
hipError_t launch_matvec_sym(hipStream_t s, const Dims& d, Stencil<double> A, const double* x,
                             double* y);
// w = w - a v - b vp
hipError_t launch_lanczos_axpy(hipStream_t s, int n, double* w, const double* v, const double* vp,
                               double a, double b);

}  // namespace les
