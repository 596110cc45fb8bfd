/*
 * les_gpu.h -- C ABI of the MI355X (gfx950) LES hot-path library
 *              (les-model_amd/liblesgpu.so).
 *
 * The reference (GabrieLeitao/les-model) has no FFI: its boundary is the set
 * of free functions in src/pressure.h:31-63 and the physics functions that
 * src/main.cpp:20-29 declares.  Each entry point below replaces one of them
 * (cited per function) or a composition of them; the C++ drop-in layer
 * (include/les/pressure.h, include/les/physics.h) maps the reference's
 * signatures onto these calls.
 *
 * Conventions
 *   - Grids are flat, idx = i*ny*nz + j*nz + k (src/field.h:40), FP64.
 *   - Masks are uint8 0/1 per cell, the byte encoding the reference's
 *     checkpoint uses for Field::is_solid (src/io.cpp:71-78).
 *   - Stencils are 7 arrays in StencilMatrix member order c,n,s,e,w,t,b
 *     (src/pressure.h:22).
 *   - A context owns all device memory; the caller owns host buffers.
 *     ptr_kind says whether a caller pointer is host (0) or device (1).
 *   - Every function returns 0 on success or a negative LES_ERR_* code;
 *     les_last_error() gives the message (les_last_error(NULL) for a call
 *     made without a context).  Non-convergence of a solver is
 *     not an error (it is reported in the result struct, mirroring
 *     pressure.cpp:411-413).
 *   - Calls are ordered on the context's stream.  Calls that return host
 *     scalars or copy to host memory synchronise that stream.
 *   - A context is not thread-safe: one host thread per context.
 */
#ifndef LES_GPU_H
#define LES_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LES_OK 0
#define LES_ERR_ARG (-1)
#define LES_ERR_HIP (-2)
#define LES_ERR_STATE (-3)
#define LES_ERR_NOMEM (-4)
#define LES_ERR_IO (-5)
/* a wavefront sweep's wave waited longer than LES_OPT_SPIN_LIMIT allows */
#define LES_ERR_STALL (-6)

typedef struct les_ctx les_ctx;

enum { LES_PTR_HOST = 0, LES_PTR_DEVICE = 1 };

/* Field slots of a context, named after the Field members (src/field.h:13-20)
 * plus the pressure right-hand side b (src/main.cpp:135). */
enum {
    LES_F_U = 0, LES_F_V, LES_F_W, LES_F_U_NEW, LES_F_V_NEW, LES_F_W_NEW,
    LES_F_P, LES_F_P_RHS, LES_F_NU_T, LES_F_B, LES_F_COUNT
};

enum { LES_STENCIL_A = 0, LES_STENCIL_L = 1 };

/* Preconditioners of the PCG solve. */
enum {
    LES_PC_NONE = 0,
    LES_PC_JACOBI = 1,      /* apply_jacobi_preconditioner, pressure.cpp:420 */
    LES_PC_IC0 = 2,         /* apply_ic_preconditioner with L from build_ic_preconditioner */
    LES_PC_LEGACY_IC_A = 3, /* apply_ic_preconditioner(., ., A) as shipped, pressure.cpp:363,396 */
    LES_PC_CHEBYSHEV = 4,   /* apply_chebyshev_preconditioner, pressure.cpp:851 */
    LES_PC_MG = 5           /* one multigrid V-cycle: no reference counterpart (the
                               reference's README.md:53 lists multigrid as a TODO);
                               Galerkin aggregation over 2x2x2 blocks, damped-Jacobi
                               sweeps (LES_OPT_MG_SWEEPS, LES_OPT_MG_COARSE_SWEEPS) */
};

enum { LES_PREC_FP64 = 0, LES_PREC_MIXED = 1 };

enum {
    LES_SOLVER_PCG = 0,       /* calculate_divergence + build_b_vector + PCG (main.cpp:134-137) */
    LES_SOLVER_SOR_LEX = 1,   /* solve_pressure_GS, exact lexicographic order (main.cpp:133) */
    LES_SOLVER_SOR_RB = 2,    /* red-black SOR (no reference counterpart) */
    LES_SOLVER_JACOBI = 3     /* solve_pressure_jacobi (main.cpp:132) */
};

/* ---- context ---- */
int les_ctx_create(int nx, int ny, int nz, double dx, double dy, double dz, les_ctx** out);
int les_ctx_destroy(les_ctx* ctx);
/* Message of the last failed call on ctx (or of the last failed
 * les_ctx_create when ctx is NULL). */
const char* les_last_error(const les_ctx* ctx);
/* Use the caller's hipStream_t for all further work (NULL: the context's own). */
int les_set_stream(les_ctx* ctx, void* hip_stream);
int les_synchronize(les_ctx* ctx);

/* Kernel-selection options (no reference counterpart).
 * LES_OPT_CHEB_PAIRS: 1 (default) runs the Chebyshev polynomial as fused
 * pairs of steps where the grid shape allows it, 0 as single steps; the
 * results are bit-identical either way.
 * LES_OPT_PAIR_GRID: about how many workgroups a fused pair launch uses
 * (0: the built-in default; at most 4096).  A tuning knob: the grid shape
 * changes only where the reduction-free pair launches split their work.
 * LES_OPT_WAVEFRONT: 1 (default) runs each lexicographic sweep -- the IC(0)
 * factor, its forward and backward solves, a lexicographic SOR sweep -- as
 * one launch (a wave per 8x8 tile of (i, j) columns, tiles handing their
 * edge values to each other through memory); 0 as one launch per i+j+k
 * hyperplane.  Bit-identical either way.
 * LES_OPT_SPIN_LIMIT: how many times a wave of such a launch may poll for
 * its neighbour tiles' values (default 4194304, about 4 s).  Past it the
 * wave stops waiting, the launch completes with invalid results and the
 * call that ran it returns LES_ERR_STALL.
 * LES_OPT_MG_SWEEPS: damped-Jacobi sweeps (omega 0.85) of the multigrid
 * preconditioner before and after each coarse correction (default 2,
 * 1..16); the correction is added times 1.8.  LES_OPT_MG_COARSE_SWEEPS:
 * sweeps on the coarsest level, at most 4 cells per direction (default 32,
 * 1..4096).
 * LES_OPT_REPLACE_DECADES: mixed precision also replaces the residual by
 * the FP64 one once the recursive residual has fallen this many decades
 * below the last true residual (default 3, 0..15; 0: only every
 * reliable_every iterations, near the tolerance and at max_iter).  The
 * FP32 recursion's drift is proportional to the residual it started from;
 * with a preconditioner that gains a decade in a few iterations (LES_PC_MG)
 * the periodic replacement comes too late and the search direction
 * restarts.
 * LES_OPT_MG_RINGS: 1 (default) runs the multigrid's pre- and
 * post-smoothing sweep pairs as one pass each with the intermediate sweep
 * in LDS rings where the level's shape allows; 0 as separate sweeps.
 * Bit-identical either way (r.z summed in another order). */
enum { LES_OPT_CHEB_PAIRS = 1, LES_OPT_PAIR_GRID = 2, LES_OPT_WAVEFRONT = 3,
       LES_OPT_SPIN_LIMIT = 4, LES_OPT_MG_SWEEPS = 5, LES_OPT_MG_COARSE_SWEEPS = 6,
       LES_OPT_REPLACE_DECADES = 7, LES_OPT_MG_RINGS = 8 };
int les_set_option(les_ctx* ctx, int option, int value);

/* ---- fields (Field, src/field.h:9-41) ---- */
int les_set_field(les_ctx* ctx, int field, const double* src, int ptr_kind);
int les_get_field(les_ctx* ctx, int field, double* dst, int ptr_kind);
/* Device address of a field slot (valid until the next call that swaps
 * buffers: les_momentum, les_step). */
int les_field_ptr(les_ctx* ctx, int field, double** dev_ptr);
int les_set_mask(les_ctx* ctx, const uint8_t* solid, int ptr_kind);
int les_get_mask(les_ctx* ctx, uint8_t* solid, int ptr_kind);

/* ---- operator ---- */
/* build_stencil (pressure.cpp:608).  ignore_mask = 1 builds A as the
 * shipped driver does, before the obstacle mask exists (main.cpp:66, 81). */
int les_build_stencil(les_ctx* ctx, int ignore_mask);
int les_set_stencil(les_ctx* ctx, int which, double* const coef[7], int ptr_kind);
int les_get_stencil(les_ctx* ctx, int which, double* const coef[7], int ptr_kind);
/* y = A x: fast_apply_laplacian (pressure.cpp:709).  which selects A or L. */
int les_matvec(les_ctx* ctx, int which, const double* x, double* y, int ptr_kind);
/* parallel_dot (pressure.cpp:962), deterministic order. */
int les_dot(les_ctx* ctx, const double* a, const double* b, int ptr_kind, double* out);
/* No reference counterpart: les_dot for callers that cannot allocate
 * (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  No context; a, b are device arrays of an nx*ny*nz grid and
 * the sum, bit-identical to les_dot's on the same grid, is written to the
 * device double *out, ordered on `hip_stream` (NULL: the null stream)
 * without synchronising.  Allocates nothing: the reduction's partial sums
 * live in a static 8 KiB device array, one per GPU, shared by every call on
 * that GPU, so calls that overlap on one GPU race.  Grids of more than
 * 2^31-1 cells are rejected (LES_ERR_ARG).  Errors: les_last_error(NULL). */
int les_dot_noalloc(int nx, int ny, int nz, const double* a, const double* b, double* out,
                    void* hip_stream);

/* ---- preconditioners ---- */
/* build_ic_preconditioner (pressure.cpp:495): L from the context's A. */
int les_ic0_factor(les_ctx* ctx);
/* z = M^-1 r for one of LES_PC_*: apply_jacobi_preconditioner,
 * apply_ic_preconditioner (pressure.cpp:442; L or A per precond) or
 * apply_chebyshev_preconditioner (pressure.cpp:851). */
int les_precond_apply(les_ctx* ctx, int precond, const double* r, double* z, int ptr_kind,
                      double lambda_min, double lambda_max, int degree);
/* (no reference counterpart) z = M^-1 r in the arithmetic a mixed-precision
 * PCG applies it: FP32 vectors and FP32 copy of A (stencil_format as in
 * les_pcg_params); LES_PC_MG in inner_product's cycle (LES_IP_MIRROR: the
 * mirror-weighted one).  Sizes: N floats each. */
int les_precond_apply_fp32(les_ctx* ctx, int precond, const float* r, float* z, int ptr_kind,
                           double lambda_min, double lambda_max, int degree, int stencil_format,
                           int inner_product);
/* estimate_lambda_max (pressure.cpp:917): power iteration from the
 * mt19937(1337) U(-1,1) start vector drawn in sequence. */
int les_power_lambda_max(les_ctx* ctx, int n_iter, double* out);
/* Lanczos extreme Ritz values after `steps` steps from a seeded start vector
 * (no reference counterpart; bounds for the Chebyshev polynomial).  A is not
 * symmetric (mirror-Neumann rows), so Lanczos runs on its similarity
 * symmetrisation S_ij = -sqrt(A_ij A_ji), which has A's eigenvalues, and on
 * the fluid subspace when A was built from the mask. */
int les_lanczos_bounds(les_ctx* ctx, int steps, uint64_t seed, double* ritz_min,
                       double* ritz_max);

/* ---- PCG (solve_pressure_pcg, pressure.cpp:342; and
 *      solve_pressure_pcg_chebyshev, pressure.cpp:762) ---- */
typedef struct les_pcg_params {
    int precond;          /* LES_PC_* */
    int max_iter;
    double tol;           /* stop when ||r|| < tol (pressure.cpp:389) ... */
    int tol_is_relative;  /* ... or when ||r|| < tol*||b|| */
    double lambda_min, lambda_max;  /* Chebyshev interval */
    int degree;           /* Chebyshev degree (pressure.h:49: default 5) */
    int precision;        /* LES_PREC_FP64 or LES_PREC_MIXED */
    int reliable_every;   /* mixed: FP64 residual replacement period (0: 50) */
    int check_every;      /* host polls convergence every this many iterations (0: auto);
                             near the iteration its polls predict to meet the
                             tolerance it polls after each one.  No result
                             depends on it except, in the last bits, a mixed
                             solve's replacement near the tolerance */
    int stencil_format;   /* 0 auto: coded when A came from les_build_stencil;
                             1 the 7 coefficient arrays; 2 coded (16 bits per cell) */
    int inner_product;    /* LES_IP_EUCLID: r.z and p.q as the reference forms them;
                             LES_IP_MIRROR: weighted by the mirror weights W
                             (preconditioner NONE, JACOBI, CHEBYSHEV or MG in FP64;
                             MG in mixed precision) */
} les_pcg_params;

/* Inner product of the CG scalars.  W = 1/2 per direction in which a row of
 * build_stencil's operator is mirrored (pressure.cpp:662-696: one
 * neighbour coefficient doubled, the opposite one 0).  W A is symmetric on
 * masks without interior obstacles, so CG in the W inner product is exact
 * CG there; the stopping test stays the Euclidean ||r||. */
enum { LES_IP_EUCLID = 0, LES_IP_MIRROR = 1 };

typedef struct les_pcg_result {
    int iters;            /* reference convention: 0 if converged at start, max_iter if not */
    int converged;
    double res_norm;      /* ||r|| of the recursive residual at exit */
    double res0_norm;     /* ||b - A x0|| */
    double b_norm;
    double true_res_norm; /* ||b - A x|| in FP64 at exit (mixed precision; FP64: = res_norm) */
    int replacements;     /* mixed: FP64 residual replacements performed */
    int fault;            /* a non-finite scalar stopped the iteration */
    int restarts;         /* mixed: replacements after which the search direction
                             restarted (true residual > 2x the recursive one) */
} les_pcg_result;

/* Solves A p = b in place: x0 and the result are the context's LES_F_P,
 * the right-hand side its LES_F_B. */
int les_pcg(les_ctx* ctx, const les_pcg_params* prm, les_pcg_result* res);
/* ||r_k|| for k = 0..iters of the last FP64 les_pcg (n entries at most). */
int les_pcg_trace(les_ctx* ctx, double* trace, int n);

/* ---- stationary pressure solvers (pressure.cpp:14-247) ---- */
int les_pressure_bcs(les_ctx* ctx);                       /* apply_pressure_BCs */
/* solve_pressure_jacobi (pressure.cpp:106): RHS from u,v,w; updates p. */
int les_jacobi_solve(les_ctx* ctx, double dt, double rho, int* sweeps_out);
/* solve_pressure_GS (pressure.cpp:186) with ordering LES_SOLVER_SOR_LEX
 * (bit-exact, i+j+k hyperplane order) or LES_SOLVER_SOR_RB (red-black). */
int les_sor_solve(les_ctx* ctx, int ordering, double dt, double rho, double omega, int max_sweeps,
                  int* sweeps_out);

/* ---- physics (main.cpp:208-714) ---- */
/* compute_sgs_stress (main.cpp:423) with Cs as a parameter (the reference
 * fixes 0.18).  apply_bcs = 1 also applies apply_nut_BCs (main.cpp:483).
 * tau (6 arrays xx,xy,xz,yy,yz,zz; may be NULL) receives the stresses. */
int les_sgs(les_ctx* ctx, double Cs, int apply_bcs, double* const tau[6], int ptr_kind);
int les_nut_bcs(les_ctx* ctx);                            /* apply_nut_BCs */
/* solve_momentum (main.cpp:563) including its u<->u_new buffer swap.
 * apply_bcs = 1 also applies apply_velocity_BCs to the new velocity. */
int les_momentum(les_ctx* ctx, double dt, double nu, int apply_bcs);
int les_velocity_bcs(les_ctx* ctx);                       /* apply_velocity_BCs */
/* calculate_divergence (pressure.cpp:256) of u_new into p_rhs, then
 * build_b_vector (pressure.cpp:566) into b. */
int les_divergence_b(les_ctx* ctx, double dt, double rho);
/* build_b_vector (pressure.cpp:566) alone: b from the context's p_rhs. */
int les_build_b(les_ctx* ctx);
/* correct_velocity (main.cpp:671); apply_bcs = 1 also applies
 * apply_velocity_BCs. */
int les_correct(les_ctx* ctx, double dt, double rho, int apply_bcs);
/* compute_adaptive_dt (main.cpp:208). */
int les_adaptive_dt(les_ctx* ctx, double nu, double* dt_out);

/* ---- one step of the driver loop (main.cpp:115-144) ---- */
typedef struct les_step_params {
    int step;          /* dt = 1e-3 while step <= 20, else adaptive (main.cpp:115-119) */
    double dt_fixed;   /* > 0 overrides that schedule */
    double nu, rho, Cs;
    int solver;        /* LES_SOLVER_* */
    les_pcg_params pcg;
    double omega;      /* SOR relaxation (pressure.cpp:195: 1.7) */
    int sor_sweeps;    /* SOR max sweeps (main.cpp:133) */
} les_step_params;

typedef struct les_step_stats {
    double dt;
    les_pcg_result pcg;
    int sweeps;
    /* max |u| and max nu_t at the start of the step, from the reductions
     * compute_adaptive_dt makes (main.cpp:217-243); -1 on steps whose dt
     * is fixed (no reduction is run) */
    double speed_max;
    double nu_t_max;
} les_step_stats;

int les_step(les_ctx* ctx, const les_step_params* prm, les_step_stats* stats);

/* ---- differentiation (no reference counterpart) ----
 * The step map F: u^n -> u^{n+1} of les_step at a fixed dt with each
 * pressure solve started from p = 0 and run to a relative tolerance (its
 * solution treated as exact).  Inlet-plane solid cells are rejected: the
 * step leaves their ping-pong values stale (main.cpp:340), which makes the
 * map depend on more than u^n.
 *
 * les_rollout advances the context's u, v, w by nsteps such steps and keeps
 * the velocity at steps 0, k, 2k, ... (k = checkpoint_every) on the device.
 * les_adjoint takes dL/du^nsteps (ubar, vbar, wbar; overwritten with
 * dL/du^0) and runs the transposed steps backwards, recomputing each
 * segment of k steps from its checkpoint.  Every pressure solve (forward
 * and adjoint) is CG in the mirror-weighted inner product
 * (pcg.inner_product = LES_IP_MIRROR), and the adjoint one uses
 * A^-T = W A^-1 W^-1.  That CG is exact CG only when A has no rows
 * mirrored around interior obstacles, so A must come from les_build_stencil
 * with ignore_mask = 1 (the shipped driver's order) or an obstacle-free
 * mask; the mask itself may hold obstacles away from the inlet plane.  The
 * context's fields are left as les_rollout left them. */
typedef struct les_adj_params {
    double nu, rho, Cs, dt;
    int nsteps;
    int checkpoint_every;   /* k >= 1 */
    les_pcg_params pcg;     /* FP64; tol_is_relative must be 1 */
} les_adj_params;

typedef struct les_adj_stats {
    int solves;             /* pressure solves run (forward, recompute and adjoint) */
    int iters;              /* their summed iteration counts */
    int unconverged;        /* solves that stopped at max_iter or on a fault */
    double max_rel_res;     /* max over solves of ||r|| / ||b|| at exit */
} les_adj_stats;

int les_rollout(les_ctx* ctx, const les_adj_params* prm, les_adj_stats* stats);
int les_adjoint(les_ctx* ctx, const les_adj_params* prm, double* ubar, double* vbar,
                double* wbar, int ptr_kind, les_adj_stats* stats);

/* ---- checkpoint, VTK output and time statistics (src/io.cpp, src/statistics.cpp) ----
 * The files are byte-identical to the reference writers' for the same field
 * values.  A NULL path or directory selects the reference's own relative
 * one.  Unlike the reference, which prints a message and writes nothing, a
 * file that cannot be opened or written returns LES_ERR_IO. */

/* save_state (io.cpp:14-59): <dir>/checkpoint.meta and <dir>/checkpoint.data
 * (dir NULL: "checkpoint") from the context's u, v, w, p, nu_t and mask.
 * The fields stream through two pinned staging chunks; synchronises. */
int les_save_state(les_ctx* ctx, const char* dir, double time_simulated, double dt,
                   int next_step_to_run);
/* load_state (io.cpp:61-120) into the context's u, v, w, p, nu_t and mask.
 * *next_step receives the meta file's step, or 0 where the reference
 * returns 0: no meta file (time and dt untouched), no data file, or nothing
 * read for the mask.  A short data file changes the fields as the reference
 * changes its arrays: a section read in part is overwritten in part, later
 * sections not at all, and the mask is set to (byte == 1) where it was read
 * and 0 beyond.  The stencil is not rebuilt.  Synchronises. */
int les_load_state(les_ctx* ctx, const char* dir, double* time_simulated, double* dt,
                   int* next_step);

/* write_field_vtk (io.cpp:225-256) from the context's fields (path NULL:
 * "output/output_<step>.vtk").  Synchronises; the formatting runs on several
 * host threads (LES_IO_THREADS, default the hardware's, at most 16). */
int les_write_field_vtk(les_ctx* ctx, const char* path, int step);
/* The same file written in the background: the fields are snapshotted on the
 * device, ordered on the context's stream, and a host thread copies, formats
 * and writes them while further work runs.  A later VTK call waits for it;
 * les_io_wait waits and returns the first background failure. */
int les_write_field_vtk_async(les_ctx* ctx, const char* path, int step);
int les_io_wait(les_ctx* ctx);
/* The same file from host arrays; no context, no device work. */
int les_write_field_vtk_host(const char* path, int step, int nx, int ny, int nz, double dx,
                             double dy, double dz, const double* u, const double* v,
                             const double* w, const double* p, const double* nu_t,
                             const uint8_t* solid);

/* Time statistics on the device (statistics.cpp:70-131): ten running sums
 * per cell, in this slot order. */
enum {
    LES_ST_U = 0, LES_ST_V, LES_ST_W, LES_ST_P,
    LES_ST_UU, LES_ST_VV, LES_ST_WW, LES_ST_UV, LES_ST_UW, LES_ST_VW, LES_ST_COUNT
};
/* init_statistics: zeroes the sums; samples are taken for steps > start_step. */
int les_stats_init(les_ctx* ctx, int start_step);
/* update_statistics: adds the context's u, v, w, p (and their products) when
 * current_step > start_step. */
int les_stats_update(les_ctx* ctx, int current_step);
/* finalize_and_save (statistics.cpp:133-186): the means and Reynolds
 * stresses are formed on the device and written to path (NULL:
 * "output/statistics_final.vtk"; "": formed, not written).  Nothing happens
 * when no sample was taken. */
int les_stats_finalize(les_ctx* ctx, const char* path);
/* One slot: the running sum (final 0) or, after les_stats_finalize, the mean
 * (slots U..P) or stress (UU..VW) (final 1). */
int les_stats_get(les_ctx* ctx, int which, int final, double* out, int ptr_kind);
int les_stats_samples(les_ctx* ctx, int* num_samples);

/* ---- diagnostics (no reference counterpart) ----
 * Brackets every launch of one kernel class with HIP events on the
 * context's stream so its average duration can be read back.  Starting
 * resets the record; LES_KT_NONE stops. */
enum { LES_KT_NONE = 0, LES_KT_CHEB_STEP = 1, LES_KT_PCG_MATVEC = 2, LES_KT_PCG_UPDATE = 3,
       LES_KT_MOMENTUM = 4, LES_KT_IC_SOLVE = 5, LES_KT_CHEB_PAIR = 6, LES_KT_STATS = 7,
       LES_KT_IC_FACTOR = 8, LES_KT_SOR_SWEEP = 9, LES_KT_MG_SWEEP = 10 };
/* LES_KT_MG_SWEEP: the multigrid's level-0 pre-smoothing launch (its first
 * two sweeps in one pass, with the restriction in FP32 with the seven
 * arrays; none with LES_OPT_MG_SWEEPS = 1). */
int les_kernel_timing(les_ctx* ctx, int kernel_class);
/* launches recorded, their summed duration in ms, and the bytes per cell the
 * class is charged (element size of the run) are returned; synchronises. */
int les_kernel_timing_read(les_ctx* ctx, int* launches, double* total_ms, int* elem_bytes);
/* Copies one FP64 grid array to another `reps` times with the plain copy
 * kernel (k_convert<double,double>): the known-size reference for HBM
 * byte counters (8*N read and 8*N written per launch). */
int les_diag_copy(les_ctx* ctx, int reps);

#ifdef __cplusplus
}
#endif
#endif
