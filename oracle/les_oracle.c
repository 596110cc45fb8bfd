/*
 * les_oracle.c -- CPU oracle for the LES per-step hot path (TEST
 * INFRASTRUCTURE ONLY; see les_oracle.h).
 *
 * Each function restates one reference function, cited file:line, keeping
 * the reference's loop order and operand order.  Build with -O2
 * -ffp-contract=off and no -march so no fused multiply-add is formed, as the
 * reference's own build (Makefile:3) forms none.
 *
 * How it is pinned (tests/test_oracle_*.py):
 *   - against the unmodified reference compiled from its sources
 *     (oracle/Makefile -> oracle/_ref/libref.so), bit for bit on one thread,
 *     function by function and for whole steps;
 *   - against the committed fixtures in tests/golden/ written by that build;
 *   - against closed forms (tests/test_oracle_closed_forms.py): the
 *     separable cosine eigenvectors of the obstacle-free stencil, exact
 *     central and one-sided differences of polynomial fields (SGS, momentum,
 *     divergence, projection), L L^T = A on the diagonal and lower pattern
 *     for IC(0), PCG against a dense solve on a 10^3 grid.
 *   orc_solve_pressure_rbsor has no reference counterpart; the same test
 *   file pins it bit for bit against a vectorised numpy restatement of its
 *   two colour half-sweeps (1, 2 and 7 sweeps, and the early-exit count),
 *   and pins its converged result, with orc_solve_pressure_gs's, against
 *   the dense fixed point of the sweep-plus-pressure-BC map.
 */
#include "les_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define IDX(i, j, k) ((i) * sij + (j) * nz + (k))

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------ */
/* BLAS-1: pressure.cpp:962-1013                                       */
/* ------------------------------------------------------------------ */
double orc_dot(const double* a, const double* b, long n)
{
    double sum = 0.0;
#pragma omp parallel for reduction(+ : sum)
    for (long i = 0; i < n; ++i) sum += a[i] * b[i];
    return sum;
}

double orc_norm(const double* a, long n) { return sqrt(orc_dot(a, a, n)); }

void orc_axpy(double* out, double a, const double* in, long n)
{
#pragma omp parallel for
    for (long i = 0; i < n; ++i) out[i] += a * in[i];
}

void orc_saxpy(double* out, double a, const double* in1, double b, const double* in2, long n)
{
#pragma omp parallel for
    for (long i = 0; i < n; ++i) out[i] = a * in1[i] + b * in2[i];
}

void orc_copy(double* out, const double* in, long n)
{
#pragma omp parallel for
    for (long i = 0; i < n; ++i) out[i] = in[i];
}

/* ------------------------------------------------------------------ */
/* build_stencil: pressure.cpp:608-699                                 */
/* ------------------------------------------------------------------ */
void orc_build_stencil(double* const A[7], const uint8_t* solid, const orc_mesh* m)
{
    const int nx = m->nx, ny = m->ny, nz = m->nz, sij = ny * nz;
    const double idx2 = 1.0 / (m->dx * m->dx);
    const double idy2 = 1.0 / (m->dy * m->dy);
    const double idz2 = 1.0 / (m->dz * m->dz);
    const double e_b = -idx2, w_b = -idx2, n_b = -idy2, s_b = -idy2, t_b = -idz2, b_b = -idz2;
    const double c_b = (2.0 * idx2 + 2.0 * idy2 + 2.0 * idz2);
    const long N = (long)nx * ny * nz;
    for (int a = 0; a < 7; ++a) memset(A[a], 0, sizeof(double) * N);
#pragma omp parallel for
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const int id = IDX(i, j, k);
                if (solid[id]) { A[ORC_C][id] = 1.0; continue; }
                if (i == nx - 1) { A[ORC_C][id] = 1.0; continue; }
                double c = c_b, e = e_b, w = w_b, n = n_b, s = s_b, t = t_b, b = b_b;
                if (i == 0) { e += w_b; w = 0.0; }
                else if (solid[IDX(i - 1, j, k)]) { e += w_b; w = 0.0; }
                if (i == nx - 2) { e = 0.0; }
                else if (solid[IDX(i + 1, j, k)]) { w += e_b; e = 0.0; }
                if (j == 0) { n += s_b; s = 0.0; }
                else if (solid[IDX(i, j - 1, k)]) { n += s_b; s = 0.0; }
                if (j == ny - 1) { s += n_b; n = 0.0; }
                else if (solid[IDX(i, j + 1, k)]) { s += n_b; n = 0.0; }
                if (k == 0) { t += b_b; b = 0.0; }
                else if (solid[IDX(i, j, k - 1)]) { t += b_b; b = 0.0; }
                if (k == nz - 1) { b += t_b; t = 0.0; }
                else if (solid[IDX(i, j, k + 1)]) { b += t_b; t = 0.0; }
                A[ORC_C][id] = c; A[ORC_E][id] = e; A[ORC_W][id] = w; A[ORC_N][id] = n;
                A[ORC_S][id] = s; A[ORC_T][id] = t; A[ORC_B][id] = b;
            }
}

/* ------------------------------------------------------------------ */
/* fast_apply_laplacian: pressure.cpp:709-746 (order c,w,e,s,n,b,t)    */
/* ------------------------------------------------------------------ */
void orc_matvec(double* q, const double* p, double* const A[7], int nx, int ny, int nz)
{
    const int sij = ny * nz;
#pragma omp parallel for
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const int id = IDX(i, j, k);
                double val = A[ORC_C][id] * p[id];
                if (i > 0) val += A[ORC_W][id] * p[id - sij];
                if (i < nx - 1) val += A[ORC_E][id] * p[id + sij];
                if (j > 0) val += A[ORC_S][id] * p[id - nz];
                if (j < ny - 1) val += A[ORC_N][id] * p[id + nz];
                if (k > 0) val += A[ORC_B][id] * p[id - 1];
                if (k < nz - 1) val += A[ORC_T][id] * p[id + 1];
                q[id] = val;
            }
}

/* ------------------------------------------------------------------ */
/* calculate_divergence: pressure.cpp:256-329                          */
/* ------------------------------------------------------------------ */
void orc_calculate_divergence(double* rhs, const double* un, const double* vn, const double* wn,
                              const uint8_t* solid, const orc_mesh* m, double dt, double rho)
{
    const int nx = m->nx, ny = m->ny, nz = m->nz, sij = ny * nz;
    const double scale = rho / dt;
    const double ix = 1.0 / (2.0 * m->dx), iy = 1.0 / (2.0 * m->dy), iz = 1.0 / (2.0 * m->dz);
    const double ix1 = 1.0 / m->dx, iy1 = 1.0 / m->dy, iz1 = 1.0 / m->dz;
#pragma omp parallel for
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const int id = IDX(i, j, k);
                if (solid[id] || i == nx - 1) { rhs[id] = 0.0; continue; }
                double dudx, dvdy, dwdz;
                if (i == 0) dudx = (un[IDX(i + 1, j, k)] - un[id]) * ix1;
                else dudx = (un[IDX(i + 1, j, k)] - un[IDX(i - 1, j, k)]) * ix;
                if (j == 0) dvdy = (vn[IDX(i, j + 1, k)] - vn[id]) * iy1;
                else if (j == ny - 1) dvdy = (vn[id] - vn[IDX(i, j - 1, k)]) * iy1;
                else dvdy = (vn[IDX(i, j + 1, k)] - vn[IDX(i, j - 1, k)]) * iy;
                if (k == 0) dwdz = (wn[IDX(i, j, k + 1)] - wn[id]) * iz1;
                else if (k == nz - 1) dwdz = (wn[id] - wn[IDX(i, j, k - 1)]) * iz1;
                else dwdz = (wn[IDX(i, j, k + 1)] - wn[IDX(i, j, k - 1)]) * iz;
                rhs[id] = scale * (dudx + dvdy + dwdz);
            }
}

/* build_b_vector: pressure.cpp:566-599 */
void orc_build_b_vector(double* b, const double* div, const uint8_t* solid, const orc_mesh* m)
{
    const int nx = m->nx, ny = m->ny, nz = m->nz, sij = ny * nz;
#pragma omp parallel for
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const int id = IDX(i, j, k);
                if (solid[id]) b[id] = 0.0;
                else if (i == nx - 1) b[id] = 0.0;
                else b[id] = -div[id];
            }
}

/* ------------------------------------------------------------------ */
/* preconditioners                                                     */
/* ------------------------------------------------------------------ */
/* apply_jacobi_preconditioner: pressure.cpp:420-428 */
void orc_jacobi_precond(double* z, const double* r, const double* Ac, long n)
{
#pragma omp parallel for
    for (long i = 0; i < n; ++i) z[i] = r[i] / Ac[i];
}

/* apply_ic_preconditioner: pressure.cpp:442-482 (serial, lexicographic) */
void orc_ic_apply(double* z, const double* r, double* const L[7], int nx, int ny, int nz)
{
    const int sij = ny * nz;
    const double *Lc = L[ORC_C], *Lw = L[ORC_W], *Ls = L[ORC_S], *Lb = L[ORC_B];
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const int id = IDX(i, j, k);
                double sum = r[id];
                if (i > 0) sum -= Lw[id] * z[id - sij];
                if (j > 0) sum -= Ls[id] * z[id - nz];
                if (k > 0) sum -= Lb[id] * z[id - 1];
                z[id] = sum / Lc[id];
            }
    for (int i = nx - 1; i >= 0; --i)
        for (int j = ny - 1; j >= 0; --j)
            for (int k = nz - 1; k >= 0; --k) {
                const int id = IDX(i, j, k);
                double sum = z[id];
                if (i < nx - 1) sum -= Lw[id + sij] * z[id + sij];
                if (j < ny - 1) sum -= Ls[id + nz] * z[id + nz];
                if (k < nz - 1) sum -= Lb[id + 1] * z[id + 1];
                z[id] = sum / Lc[id];
            }
}

/* build_ic_preconditioner: pressure.cpp:495-556 (serial).  Like the
 * reference, entries of L that the guards skip keep their previous value;
 * the caller passes L zero-initialised as StencilMatrix(n) is. */
void orc_ic_build(double* const L[7], double* const A[7], int nx, int ny, int nz)
{
    const int sij = ny * nz;
    const double small_val = 1e-30;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const int id = IDX(i, j, k);
                double sum_sq = 0.0;
                if (i > 0) {
                    double lc = L[ORC_C][id - sij];
                    if (fabs(lc) > small_val) {
                        L[ORC_W][id] = A[ORC_W][id] / lc;
                        sum_sq += L[ORC_W][id] * L[ORC_W][id];
                    }
                }
                if (j > 0) {
                    double lc = L[ORC_C][id - nz];
                    if (fabs(lc) > small_val) {
                        L[ORC_S][id] = A[ORC_S][id] / lc;
                        sum_sq += L[ORC_S][id] * L[ORC_S][id];
                    }
                }
                if (k > 0) {
                    double lc = L[ORC_C][id - 1];
                    if (fabs(lc) > small_val) {
                        L[ORC_B][id] = A[ORC_B][id] / lc;
                        sum_sq += L[ORC_B][id] * L[ORC_B][id];
                    }
                }
                double diag = A[ORC_C][id] - sum_sq;
                L[ORC_C][id] = (diag <= small_val) ? 1.0 : sqrt(diag);
            }
}

/* apply_chebyshev_preconditioner: pressure.cpp:851-906 */
void orc_cheb_apply(double* z, const double* r, double* const A[7], int nx, int ny, int nz,
                    double lmin, double lmax, int degree)
{
    const long n = (long)nx * ny * nz;
    double* p = (double*)calloc(n, sizeof(double));
    double* res = (double*)calloc(n, sizeof(double));
    const double d = (lmax + lmin) / 2.0;
    const double c = (lmax - lmin) / 2.0;
    double alpha, beta;
#pragma omp parallel for
    for (long i = 0; i < n; ++i) z[i] = 0.0;
    orc_copy(res, r, n);
    alpha = 2.0 / d;
    orc_copy(p, res, n);
    orc_axpy(z, alpha, p, n);
    for (int k = 1; k < degree; ++k) {
        orc_matvec(res, z, A, nx, ny, nz);
        orc_saxpy(res, 1.0, r, -1.0, res, n);
        beta = (c * alpha / 2.0) * (c * alpha / 2.0);
        alpha = 1.0 / (d - beta);
        orc_saxpy(p, 1.0, res, beta, p, n);
        orc_axpy(z, alpha, p, n);
    }
    free(p);
    free(res);
}

/* --- std::mt19937 and libstdc++'s uniform_real_distribution<double> ---
 * MT19937 is the published Matsumoto-Nishimura generator (seed by the
 * 1812433253 recurrence).  libstdc++ (g++ 11) draws a double in [0,1) by
 * generate_canonical<double,53>: two 32-bit outputs x0 + x1*2^32, divided by
 * 2^64, clamped below 1; the distribution returns u*(b-a)+a. */
typedef struct { uint32_t mt[624]; int idx; } orc_mt;

static void mt_seed(orc_mt* g, uint32_t s)
{
    g->mt[0] = s;
    for (int i = 1; i < 624; ++i)
        g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
    g->idx = 624;
}

static uint32_t mt_next(orc_mt* g)
{
    if (g->idx >= 624) {
        for (int i = 0; i < 624; ++i) {
            uint32_t y = (g->mt[i] & 0x80000000u) | (g->mt[(i + 1) % 624] & 0x7fffffffu);
            g->mt[i] = g->mt[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        g->idx = 0;
    }
    uint32_t y = g->mt[g->idx++];
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= (y >> 18);
    return y;
}

static double mt_uniform(orc_mt* g, double a, double b)
{
    double sum = 0.0, tmp = 1.0;
    const double r = 4294967296.0;
    for (int m = 0; m < 2; ++m) {
        sum += (double)mt_next(g) * tmp;
        tmp *= r;
    }
    double ret = sum / tmp;
    if (ret >= 1.0) ret = nextafter(1.0, 0.0);
    return ret * (b - a) + a;
}

/* estimate_lambda_max: pressure.cpp:917-955 */
double orc_estimate_lambda_max(double* const A[7], int nx, int ny, int nz, int n_iter)
{
    const long n = (long)nx * ny * nz;
    double* v = (double*)malloc(n * sizeof(double));
    double* w = (double*)malloc(n * sizeof(double));
    orc_mt g;
    mt_seed(&g, 1337u);
    for (long i = 0; i < n; ++i) v[i] = mt_uniform(&g, -1.0, 1.0);
    orc_copy(w, v, n);
    double lam = 0.0;
    for (int k = 0; k < n_iter; ++k) {
        orc_matvec(w, v, A, nx, ny, nz);
        lam = orc_dot(v, w, n);
        double nrm = orc_norm(w, n);
        orc_copy(v, w, n);
#pragma omp parallel for
        for (long i = 0; i < n; ++i) v[i] /= nrm;
    }
    free(v);
    free(w);
    return lam;
}

/* ------------------------------------------------------------------ */
/* PCG: pressure.cpp:342-414 (and 762-835 for the Chebyshev variant)   */
/* ------------------------------------------------------------------ */
static void orc_precond(int pc, double* z, const double* r, double* const A[7],
                        double* const L[7], int nx, int ny, int nz, double lmin, double lmax,
                        int degree)
{
    const long n = (long)nx * ny * nz;
    switch (pc) {
    case ORC_PC_JACOBI: orc_jacobi_precond(z, r, A[ORC_C], n); break;
    case ORC_PC_IC0: orc_ic_apply(z, r, L, nx, ny, nz); break;
    case ORC_PC_LEGACY_IC_A: orc_ic_apply(z, r, A, nx, ny, nz); break;
    case ORC_PC_CHEBYSHEV: orc_cheb_apply(z, r, A, nx, ny, nz, lmin, lmax, degree); break;
    default: orc_copy(z, r, n); break;
    }
}

static double wdot(const double* a, const double* b, const double* W, long n)
{
    if (!W) return orc_dot(a, b, n);
    double sum = 0.0;
    for (long i = 0; i < n; ++i) sum += W[i] * a[i] * b[i];
    return sum;
}

/* The PCG loop of pressure.cpp:342-414, with the preconditioner as a
 * callback: apply(ctx, z, r). */
typedef void (*orc_precond_fn)(void* ctx, double* z, const double* r);

struct std_pc {
    int pc;
    double* const* A;
    double* const* L;
    int nx, ny, nz, degree;
    double lmin, lmax;
};

static void std_pc_apply(void* ctx, double* z, const double* r)
{
    const struct std_pc* c = (const struct std_pc*)ctx;
    orc_precond(c->pc, z, r, c->A, c->L, c->nx, c->ny, c->nz, c->lmin, c->lmax, c->degree);
}

static int pcg_core(double* const A[7], const double* b, double* x, int nx, int ny, int nz,
                    orc_precond_fn apply, void* ctx, int max_iter, double tol,
                    double* res_trace, const double* W)
{
    const long n = (long)nx * ny * nz;
    double* r = (double*)calloc(n, sizeof(double));
    double* z = (double*)calloc(n, sizeof(double));
    double* p = (double*)calloc(n, sizeof(double));
    double* q = (double*)calloc(n, sizeof(double));
    int ret = max_iter;
    orc_matvec(q, x, A, nx, ny, nz);
    orc_saxpy(r, 1.0, b, -1.0, q, n);
    double err0 = orc_norm(r, n);
    if (res_trace) res_trace[0] = err0;
    if (err0 < tol) { ret = 0; goto out; }
    apply(ctx, p, r);
    double rho = wdot(r, p, W, n);
    for (int k = 1; k <= max_iter; ++k) {
        orc_matvec(q, p, A, nx, ny, nz);
        double alpha = rho / wdot(p, q, W, n);
        orc_axpy(x, alpha, p, n);
        orc_axpy(r, -alpha, q, n);
        double err = orc_norm(r, n);
        if (res_trace) res_trace[k] = err;
        if (err < tol) { ret = k; goto out; }
        apply(ctx, z, r);
        double rho_new = wdot(r, z, W, n);
        double beta = rho_new / rho;
        orc_saxpy(p, 1.0, z, beta, p, n);
        rho = rho_new;
    }
out:
    free(r); free(z); free(p); free(q);
    return ret;
}

int orc_pcg(double* const A[7], const double* b, double* x, int nx, int ny, int nz,
            int precond, double* const L[7], double lmin, double lmax, int degree,
            int max_iter, double tol, double* res_trace)
{
    return orc_pcg_w(A, b, x, nx, ny, nz, precond, L, lmin, lmax, degree, max_iter, tol,
                     res_trace, NULL);
}

/* PCG (pressure.cpp:342-414) with the inner products r.z and p.q taken in
 * the W-weighted inner product when W != NULL (no reference counterpart;
 * the stopping test stays the Euclidean ||r||). */
int orc_pcg_w(double* const A[7], const double* b, double* x, int nx, int ny, int nz,
              int precond, double* const L[7], double lmin, double lmax, int degree,
              int max_iter, double tol, double* res_trace, const double* W)
{
    if (precond == ORC_PC_MG) {
        const orc_mg_params prm = orc_mg_default_params();
        return orc_pcg_mg(A, b, x, nx, ny, nz, &prm, max_iter, tol, res_trace, W);
    }
    struct std_pc c = {precond, A, L, nx, ny, nz, degree, lmin, lmax};
    return pcg_core(A, b, x, nx, ny, nz, std_pc_apply, &c, max_iter, tol, res_trace, W);
}

/* ------------------------------------------------------------------ */
/* multigrid preconditioner (no reference counterpart; README.md:53)   */
/* ------------------------------------------------------------------ */
orc_mg_params orc_mg_default_params(void)
{
    orc_mg_params p;
    p.sweeps = 2;
    p.coarse_sweeps = 32;
    p.coarse_max = 4;
    p.omega = 0.85;
    p.scale = 1.8;
    return p;
}

int orc_mg_levels(int nx, int ny, int nz, int coarse_max, int* dims)
{
    int l = 0;
    dims[0] = nx; dims[1] = ny; dims[2] = nz;
    while (l + 1 < ORC_MG_MAX_LEVELS) {
        const int* d = dims + 3 * l;
        int mx = d[0] > d[1] ? d[0] : d[1];
        if (d[2] > mx) mx = d[2];
        if (mx <= coarse_max) break;
        dims[3 * (l + 1) + 0] = (d[0] + 1) / 2;
        dims[3 * (l + 1) + 1] = (d[1] + 1) / 2;
        dims[3 * (l + 1) + 2] = (d[2] + 1) / 2;
        ++l;
    }
    return l + 1;
}

void orc_mg_coarsen(double* const Ac[7], double* const A[7], const double* W, int nx, int ny,
                    int nz)
{
    const int NX = (nx + 1) / 2, NY = (ny + 1) / 2, NZ = (nz + 1) / 2;
    const int sij = ny * nz;
    for (int I = 0; I < NX; ++I)
        for (int J = 0; J < NY; ++J)
            for (int K = 0; K < NZ; ++K) {
                double c = 0.0, cn = 0.0, cs = 0.0, ce = 0.0, cw = 0.0, ct = 0.0, cb = 0.0;
                for (int di = 0; di < 2; ++di)
                    for (int dj = 0; dj < 2; ++dj)
                        for (int dk = 0; dk < 2; ++dk) {
                            const int i = 2 * I + di, j = 2 * J + dj, k = 2 * K + dk;
                            if (i >= nx || j >= ny || k >= nz) continue;
                            const int id = IDX(i, j, k);
                            /* a fine coefficient, times W of its row when weighted */
#define WA(q) (W ? W[id] * A[q][id] : A[q][id])
                            c += WA(ORC_C);
                            /* member order n, s, e, w, t, b: inside the block
                             * to the centre, across its face to that face */
                            if (j + 1 < ny) { if (dj == 0) c += WA(ORC_N); else cn += WA(ORC_N); }
                            if (j > 0) { if (dj == 1) c += WA(ORC_S); else cs += WA(ORC_S); }
                            if (i + 1 < nx) { if (di == 0) c += WA(ORC_E); else ce += WA(ORC_E); }
                            if (i > 0) { if (di == 1) c += WA(ORC_W); else cw += WA(ORC_W); }
                            if (k + 1 < nz) { if (dk == 0) c += WA(ORC_T); else ct += WA(ORC_T); }
                            if (k > 0) { if (dk == 1) c += WA(ORC_B); else cb += WA(ORC_B); }
#undef WA
                        }
                const int cid = (I * NY + J) * NZ + K;
                Ac[ORC_C][cid] = c;
                Ac[ORC_N][cid] = cn;
                Ac[ORC_S][cid] = cs;
                Ac[ORC_E][cid] = ce;
                Ac[ORC_W][cid] = cw;
                Ac[ORC_T][cid] = ct;
                Ac[ORC_B][cid] = cb;
            }
}

typedef struct {
    int nlev;
    int dims[3 * ORC_MG_MAX_LEVELS];
    double* A[ORC_MG_MAX_LEVELS][7];  /* level 0 borrows the caller's */
    double* x[ORC_MG_MAX_LEVELS];     /* level iterates (level 0: the output z) */
    double* t[ORC_MG_MAX_LEVELS];     /* second buffer of a sweep */
    double* b[ORC_MG_MAX_LEVELS];     /* right-hand sides of levels >= 1 */
    const double* W;                  /* NULL, or the mirror weights of level 0 */
    orc_mg_params prm;
} orc_mg;

static long mg_n(const orc_mg* g, int l)
{
    return (long)g->dims[3 * l] * g->dims[3 * l + 1] * g->dims[3 * l + 2];
}

static void mg_free(orc_mg* g)
{
    for (int l = 0; l < g->nlev; ++l) {
        if (l > 0)
            for (int q = 0; q < 7; ++q) free(g->A[l][q]);
        free(g->t[l]);
        if (l > 0) { free(g->x[l]); free(g->b[l]); }
    }
}

static void mg_build(orc_mg* g, double* const A[7], int nx, int ny, int nz,
                     const orc_mg_params* prm, const double* W)
{
    memset(g, 0, sizeof(*g));
    g->prm = *prm;
    g->W = W;
    g->nlev = orc_mg_levels(nx, ny, nz, prm->coarse_max, g->dims);
    for (int q = 0; q < 7; ++q) g->A[0][q] = A[q];
    for (int l = 0; l < g->nlev; ++l) {
        const long n = mg_n(g, l);
        g->t[l] = (double*)calloc(n, sizeof(double));
        if (l == 0) continue;
        for (int q = 0; q < 7; ++q) g->A[l][q] = (double*)calloc(n, sizeof(double));
        g->x[l] = (double*)calloc(n, sizeof(double));
        g->b[l] = (double*)calloc(n, sizeof(double));
        const int* d = g->dims + 3 * (l - 1);
        orc_mg_coarsen(g->A[l], g->A[l - 1], l == 1 ? W : NULL, d[0], d[1], d[2]);
    }
}

/* (A x)[id] in fast_apply_laplacian's order (pressure.cpp:709-746) */
static double mg_ax(double* const A[7], const double* x, int nx, int ny, int nz, int i, int j,
                    int k)
{
    const int sij = ny * nz, id = IDX(i, j, k);
    double val = A[ORC_C][id] * x[id];
    if (i > 0) val += A[ORC_W][id] * x[id - sij];
    if (i < nx - 1) val += A[ORC_E][id] * x[id + sij];
    if (j > 0) val += A[ORC_S][id] * x[id - nz];
    if (j < ny - 1) val += A[ORC_N][id] * x[id + nz];
    if (k > 0) val += A[ORC_B][id] * x[id - 1];
    if (k < nz - 1) val += A[ORC_T][id] * x[id + 1];
    return val;
}

/* `sweeps` damped-Jacobi sweeps on level l into g->x[l]; from x = 0 when
 * zero (x = omega (b / A_c) first) */
static void mg_smooth(orc_mg* g, int l, double* x, const double* b, int sweeps, int zero)
{
    const int nx = g->dims[3 * l], ny = g->dims[3 * l + 1], nz = g->dims[3 * l + 2];
    const int sij = ny * nz;
    const long n = mg_n(g, l);
    double* const* A = g->A[l];
    const double om = g->prm.omega;
    if (sweeps <= 0 && zero) memset(x, 0, (size_t)n * sizeof(double));
    for (int s = 0; s < sweeps; ++s) {
        if (s == 0 && zero) {
            for (long id = 0; id < n; ++id) x[id] = om * (b[id] / A[ORC_C][id]);
            continue;
        }
        double* t = g->t[l];
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j)
                for (int k = 0; k < nz; ++k) {
                    const int id = IDX(i, j, k);
                    const double q = mg_ax(A, x, nx, ny, nz, i, j, k);
                    t[id] = x[id] + om * ((b[id] - q) / A[ORC_C][id]);
                }
        memcpy(x, t, (size_t)n * sizeof(double));
    }
}

static void mg_vcycle(orc_mg* g, int l, double* x, const double* b)
{
    const orc_mg_params* p = &g->prm;
    if (l == g->nlev - 1) {
        mg_smooth(g, l, x, b, p->coarse_sweeps, 1);
        return;
    }
    const int nx = g->dims[3 * l], ny = g->dims[3 * l + 1], nz = g->dims[3 * l + 2];
    const int NX = g->dims[3 * l + 3], NY = g->dims[3 * l + 4], NZ = g->dims[3 * l + 5];
    const int sij = ny * nz;
    double* const* A = g->A[l];
    mg_smooth(g, l, x, b, p->sweeps, 1);
    /* restriction of the residual: the sum over each block, children in
     * (di, dj, dk) order */
    double* bc = g->b[l + 1];
    for (int I = 0; I < NX; ++I)
        for (int J = 0; J < NY; ++J)
            for (int K = 0; K < NZ; ++K) {
                double acc = 0.0;
                for (int di = 0; di < 2; ++di)
                    for (int dj = 0; dj < 2; ++dj)
                        for (int dk = 0; dk < 2; ++dk) {
                            const int i = 2 * I + di, j = 2 * J + dj, k = 2 * K + dk;
                            if (i >= nx || j >= ny || k >= nz) continue;
                            const int id = IDX(i, j, k);
                            const double res = b[id] - mg_ax(A, x, nx, ny, nz, i, j, k);
                            acc += (l == 0 && g->W) ? g->W[id] * res : res;
                        }
                bc[(I * NY + J) * NZ + K] = acc;
            }
    double* xc = g->x[l + 1];
    mg_vcycle(g, l + 1, xc, bc);
    /* prolongation of the correction, scaled */
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const int id = IDX(i, j, k);
                x[id] = x[id] + p->scale * xc[((i / 2) * NY + j / 2) * NZ + k / 2];
            }
    mg_smooth(g, l, x, b, p->sweeps, 0);
}

void orc_mg_apply(double* z, const double* r, double* const A[7], int nx, int ny, int nz,
                  const orc_mg_params* prm, const double* W)
{
    orc_mg g;
    mg_build(&g, A, nx, ny, nz, prm, W);
    mg_vcycle(&g, 0, z, r);
    mg_free(&g);
}

static void mg_pc_apply(void* ctx, double* z, const double* r)
{
    orc_mg* g = (orc_mg*)ctx;
    mg_vcycle(g, 0, z, r);
}

int orc_pcg_mg(double* const A[7], const double* b, double* x, int nx, int ny, int nz,
               const orc_mg_params* prm, int max_iter, double tol, double* res_trace,
               const double* W)
{
    orc_mg g;
    mg_build(&g, A, nx, ny, nz, prm, W);
    const int it = pcg_core(A, b, x, nx, ny, nz, mg_pc_apply, &g, max_iter, tol, res_trace, W);
    mg_free(&g);
    return it;
}

/* ------------------------------------------------------------------ */
/* stationary solvers                                                  */
/* ------------------------------------------------------------------ */
/* apply_pressure_BCs: pressure.cpp:14-93 */
void orc_apply_pressure_bcs(double* p, const uint8_t* solid, const orc_mesh* m)
{
    const int nx = m->nx, ny = m->ny, nz = m->nz, sij = ny * nz;
    for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k) p[IDX(0, j, k)] = p[IDX(1, j, k)];
    for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k) p[IDX(nx - 1, j, k)] = 0.0;
    for (int i = 0; i < nx; ++i)
        for (int k = 0; k < nz; ++k) {
            p[IDX(i, 0, k)] = p[IDX(i, 1, k)];
            p[IDX(i, ny - 1, k)] = p[IDX(i, ny - 2, k)];
        }
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
            p[IDX(i, j, 0)] = p[IDX(i, j, 1)];
            p[IDX(i, j, nz - 1)] = p[IDX(i, j, nz - 2)];
        }
    for (int i = 1; i < nx - 1; ++i)
        for (int j = 1; j < ny - 1; ++j)
            for (int k = 1; k < nz - 1; ++k) {
                const int id = IDX(i, j, k);
                if (!solid[id]) continue;
                if (i < nx - 1 && !solid[id + sij]) p[id] = p[id + sij];
                else if (i > 0 && !solid[id - sij]) p[id] = p[id - sij];
                else if (j < ny - 1 && !solid[id + nz]) p[id] = p[id + nz];
                else if (j > 0 && !solid[id - nz]) p[id] = p[id - nz];
                else if (k < nz - 1 && !solid[id + 1]) p[id] = p[id + 1];
                else if (k > 0 && !solid[id - 1]) p[id] = p[id - 1];
            }
}

/* RHS of the stationary solvers: pressure.cpp:119-132 and 197-210 */
static void orc_rhs_central(double* rhs, const double* u, const double* v, const double* w,
                            const orc_mesh* m, double coef)
{
    const int nx = m->nx, ny = m->ny, nz = m->nz, sij = ny * nz;
#pragma omp parallel for
    for (int i = 1; i < nx - 1; ++i)
        for (int j = 1; j < ny - 1; ++j)
            for (int k = 1; k < nz - 1; ++k) {
                double div = (u[IDX(i + 1, j, k)] - u[IDX(i - 1, j, k)]) / (2 * m->dx)
                           + (v[IDX(i, j + 1, k)] - v[IDX(i, j - 1, k)]) / (2 * m->dy)
                           + (w[IDX(i, j, k + 1)] - w[IDX(i, j, k - 1)]) / (2 * m->dz);
                rhs[IDX(i, j, k)] = coef * div;
            }
}

/* solve_pressure_jacobi: pressure.cpp:106-172.  p_new is a fresh zeroed
 * buffer per call, swapped with p every sweep, exactly as the reference. */
int orc_solve_pressure_jacobi(double* p, const double* u, const double* v, const double* w,
                              const uint8_t* solid, const orc_mesh* m, double dt, double rho)
{
    const int nx = m->nx, ny = m->ny, nz = m->nz, sij = ny * nz;
    const long n = (long)nx * ny * nz;
    double* pn = (double*)calloc(n, sizeof(double));
    double* rhs = (double*)calloc(n, sizeof(double));
    double* cur = p;  /* the buffer currently holding f.p */
    const double coef = rho / dt;
    const double dx2 = m->dx * m->dx, dy2 = m->dy * m->dy, dz2 = m->dz * m->dz;
    const double denom = 2 * (1.0 / dx2 + 1.0 / dy2 + 1.0 / dz2);
    orc_rhs_central(rhs, u, v, w, m, coef);
    int sweeps = 0;
    for (int it = 0; it < 100; ++it) {
        double maxr = 0.0;
#pragma omp parallel for reduction(max : maxr)
        for (int i = 1; i < nx - 1; ++i)
            for (int j = 1; j < ny - 1; ++j)
                for (int k = 1; k < nz - 1; ++k) {
                    const int id = IDX(i, j, k);
                    double pold = cur[id];
                    pn[id] = ((cur[IDX(i + 1, j, k)] + cur[IDX(i - 1, j, k)]) / dx2 +
                              (cur[IDX(i, j + 1, k)] + cur[IDX(i, j - 1, k)]) / dy2 +
                              (cur[IDX(i, j, k + 1)] + cur[IDX(i, j, k - 1)]) / dz2 - rhs[id]) /
                             denom;
                    double res = fabs(pn[id] - pold);
                    if (res > maxr) maxr = res;
                }
        double* t = cur; cur = pn; pn = t;
        orc_apply_pressure_bcs(cur, solid, m);
        sweeps = it + 1;
        if (maxr < 1e-6) break;
    }
    if (cur != p) {
        memcpy(p, cur, n * sizeof(double));
        free(cur);
    } else {
        free(pn);
    }
    free(rhs);
    return sweeps;
}

/* solve_pressure_GS: pressure.cpp:186-247 (lexicographic, in place) */
int orc_solve_pressure_gs(double* p, const double* u, const double* v, const double* w,
                          const uint8_t* solid, const orc_mesh* m, double dt, double rho,
                          int max_iter, double omega)
{
    const int nx = m->nx, ny = m->ny, nz = m->nz, sij = ny * nz;
    const long n = (long)nx * ny * nz;
    double* rhs = (double*)calloc(n, sizeof(double));
    const double dx2 = m->dx * m->dx, dy2 = m->dy * m->dy, dz2 = m->dz * m->dz;
    const double coef = 1.0 / (2.0 * (1.0 / dx2 + 1.0 / dy2 + 1.0 / dz2));
    orc_rhs_central(rhs, u, v, w, m, rho / dt);
    int sweeps = 0;
    for (int it = 0; it < max_iter; ++it) {
        double maxr = 0.0;
        for (int i = 1; i < nx - 1; ++i)
            for (int j = 1; j < ny - 1; ++j)
                for (int k = 1; k < nz - 1; ++k) {
                    const int id = IDX(i, j, k);
                    double pold = p[id];
                    p[id] = (1 - omega) * p[id] + omega * coef *
                            ((p[IDX(i + 1, j, k)] + p[IDX(i - 1, j, k)]) / dx2 +
                             (p[IDX(i, j + 1, k)] + p[IDX(i, j - 1, k)]) / dy2 +
                             (p[IDX(i, j, k + 1)] + p[IDX(i, j, k - 1)]) / dz2 - rhs[id]);
                    double res = fabs(p[id] - pold);
                    if (res > maxr) maxr = res;
                }
        orc_apply_pressure_bcs(p, solid, m);
        sweeps = it + 1;
        if (maxr < 1e-6) break;
    }
    free(rhs);
    return sweeps;
}

/* Red-black SOR (new): the GS update of pressure.cpp:224-229 applied to the
 * cells with (i+j+k) even, then to those with (i+j+k) odd. */
int orc_solve_pressure_rbsor(double* p, const double* u, const double* v, const double* w,
                             const uint8_t* solid, const orc_mesh* m, double dt, double rho,
                             int max_iter, double omega)
{
    const int nx = m->nx, ny = m->ny, nz = m->nz, sij = ny * nz;
    const long n = (long)nx * ny * nz;
    double* rhs = (double*)calloc(n, sizeof(double));
    const double dx2 = m->dx * m->dx, dy2 = m->dy * m->dy, dz2 = m->dz * m->dz;
    const double coef = 1.0 / (2.0 * (1.0 / dx2 + 1.0 / dy2 + 1.0 / dz2));
    orc_rhs_central(rhs, u, v, w, m, rho / dt);
    int sweeps = 0;
    for (int it = 0; it < max_iter; ++it) {
        double maxr = 0.0;
        for (int colour = 0; colour < 2; ++colour) {
#pragma omp parallel for reduction(max : maxr)
            for (int i = 1; i < nx - 1; ++i)
                for (int j = 1; j < ny - 1; ++j)
                    for (int k = 1; k < nz - 1; ++k) {
                        if (((i + j + k) & 1) != colour) continue;
                        const int id = IDX(i, j, k);
                        double pold = p[id];
                        p[id] = (1 - omega) * p[id] + omega * coef *
                                ((p[IDX(i + 1, j, k)] + p[IDX(i - 1, j, k)]) / dx2 +
                                 (p[IDX(i, j + 1, k)] + p[IDX(i, j - 1, k)]) / dy2 +
                                 (p[IDX(i, j, k + 1)] + p[IDX(i, j, k - 1)]) / dz2 - rhs[id]);
                        double res = fabs(p[id] - pold);
                        if (res > maxr) maxr = res;
                    }
        }
        orc_apply_pressure_bcs(p, solid, m);
        sweeps = it + 1;
        if (maxr < 1e-6) break;
    }
    free(rhs);
    return sweeps;
}

/* ------------------------------------------------------------------ */
/* physics: main.cpp                                                   */
/* ------------------------------------------------------------------ */
/* compute_adaptive_dt: main.cpp:208-250 */
double orc_adaptive_dt(const double* u, const double* v, const double* w, const double* nu_t,
                       const orc_mesh* m, double nu)
{
    const long n = (long)m->nx * m->ny * m->nz;
    double umax2 = 0.0, ntmax = 0.0;
#pragma omp parallel for reduction(max : umax2)
    for (long i = 0; i < n; ++i) {
        double s = u[i] * u[i] + v[i] * v[i] + w[i] * w[i];
        umax2 = umax2 > s ? umax2 : s;
    }
    double umax = sqrt(umax2);
    double mind = m->dx;
    if (m->dy < mind) mind = m->dy;
    if (m->dz < mind) mind = m->dz;
    double dt_conv = 1.7976931348623157e308;
    if (umax > 1e-12) dt_conv = 0.4 * mind / umax;
#pragma omp parallel for reduction(max : ntmax)
    for (long i = 0; i < n; ++i) ntmax = ntmax > nu_t[i] ? ntmax : nu_t[i];
    double dt_visc = 0.2 * (mind * mind) / (nu + ntmax);
    double r = dt_conv;
    if (dt_visc < r) r = dt_visc;
    if (1e-4 < r) r = 1e-4;
    return r;
}

/* (Cs*Delta)^2: main.cpp:426 and 461 */
double orc_cs_delta_sq(const orc_mesh* m, double Cs)
{
    double delta = pow(m->dx * m->dy * m->dz, 1.0 / 3.0);
    return pow(Cs * delta, 2);
}

/* compute_sgs_stress: main.cpp:423-472.  tau may be NULL (the reference
 * writes the six stresses and never reads them). */
void orc_sgs(double* nu_t, double* const tau[6], const double* u, const double* v, const double* w,
             const orc_mesh* m, double Cs)
{
    const int nx = m->nx, ny = m->ny, nz = m->nz, sij = ny * nz;
    const double csd2 = orc_cs_delta_sq(m, Cs);
#pragma omp parallel for
    for (int i = 1; i < nx - 1; i++)
        for (int j = 1; j < ny - 1; j++)
            for (int k = 1; k < nz - 1; k++) {
                const int id = IDX(i, j, k);
                double dudx = (u[IDX(i + 1, j, k)] - u[IDX(i - 1, j, k)]) / (2 * m->dx);
                double dudy = (u[IDX(i, j + 1, k)] - u[IDX(i, j - 1, k)]) / (2 * m->dy);
                double dudz = (u[IDX(i, j, k + 1)] - u[IDX(i, j, k - 1)]) / (2 * m->dz);
                double dvdx = (v[IDX(i + 1, j, k)] - v[IDX(i - 1, j, k)]) / (2 * m->dx);
                double dvdy = (v[IDX(i, j + 1, k)] - v[IDX(i, j - 1, k)]) / (2 * m->dy);
                double dvdz = (v[IDX(i, j, k + 1)] - v[IDX(i, j, k - 1)]) / (2 * m->dz);
                double dwdx = (w[IDX(i + 1, j, k)] - w[IDX(i - 1, j, k)]) / (2 * m->dx);
                double dwdy = (w[IDX(i, j + 1, k)] - w[IDX(i, j - 1, k)]) / (2 * m->dy);
                double dwdz = (w[IDX(i, j, k + 1)] - w[IDX(i, j, k - 1)]) / (2 * m->dz);
                double Sxx = dudx, Syy = dvdy, Szz = dwdz;
                double Sxy = 0.5 * (dudy + dvdx);
                double Sxz = 0.5 * (dudz + dwdx);
                double Syz = 0.5 * (dvdz + dwdy);
                double Smag = sqrt(2.0 * (Sxx * Sxx + Syy * Syy + Szz * Szz) +
                                   4.0 * (Sxy * Sxy + Sxz * Sxz + Syz * Syz));
                double nt = csd2 * Smag;
                nu_t[id] = nt;
                if (tau) {
                    tau[0][id] = -2.0 * nt * Sxx;
                    tau[1][id] = -2.0 * nt * Sxy;
                    tau[2][id] = -2.0 * nt * Sxz;
                    tau[3][id] = -2.0 * nt * Syy;
                    tau[4][id] = -2.0 * nt * Syz;
                    tau[5][id] = -2.0 * nt * Szz;
                }
            }
}

/* apply_nut_BCs: main.cpp:483-547 */
void orc_nut_bcs(double* nu_t, const uint8_t* solid, const orc_mesh* m)
{
    const int nx = m->nx, ny = m->ny, nz = m->nz, sij = ny * nz;
    for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k) nu_t[IDX(0, j, k)] = 0.0;
    for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k) nu_t[IDX(nx - 1, j, k)] = nu_t[IDX(nx - 2, j, k)];
    for (int i = 0; i < nx; ++i)
        for (int k = 0; k < nz; ++k) {
            nu_t[IDX(i, 0, k)] = 0.0;
            nu_t[IDX(i, ny - 1, k)] = 0.0;
        }
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
            nu_t[IDX(i, j, 0)] = 0.0;
            nu_t[IDX(i, j, nz - 1)] = 0.0;
        }
    for (int i = 1; i < nx - 1; ++i)
        for (int j = 1; j < ny - 1; ++j)
            for (int k = 1; k < nz - 1; ++k)
                if (solid[IDX(i, j, k)]) nu_t[IDX(i, j, k)] = 0.0;
}

/* solve_momentum: main.cpp:563-659 (three sweeps, interior only) */
void orc_momentum(double* un, double* vn, double* wn, const double* u, const double* v,
                  const double* w, const double* nu_t, const orc_mesh* m, double dt, double nu)
{
    const int nx = m->nx, ny = m->ny, nz = m->nz, sij = ny * nz;
    const double idx2 = 1.0 / (m->dx * m->dx), idy2 = 1.0 / (m->dy * m->dy),
                 idz2 = 1.0 / (m->dz * m->dz);
    const double i2x = 1.0 / (2.0 * m->dx), i2y = 1.0 / (2.0 * m->dy), i2z = 1.0 / (2.0 * m->dz);
    const double* src[3] = {u, v, w};
    double* dst[3] = {un, vn, wn};
    for (int c = 0; c < 3; ++c) {
        const double* f = src[c];
        double* o = dst[c];
#pragma omp parallel for
        for (int i = 1; i < nx - 1; ++i)
            for (int j = 1; j < ny - 1; ++j)
                for (int k = 1; k < nz - 1; ++k) {
                    const int id = IDX(i, j, k);
                    double nu_eff = nu + nu_t[id];
                    double ddx = (f[IDX(i + 1, j, k)] - f[IDX(i - 1, j, k)]) * i2x;
                    double ddy = (f[IDX(i, j + 1, k)] - f[IDX(i, j - 1, k)]) * i2y;
                    double ddz = (f[IDX(i, j, k + 1)] - f[IDX(i, j, k - 1)]) * i2z;
                    double adv = u[id] * ddx + v[id] * ddy + w[id] * ddz;
                    double lap = (f[IDX(i + 1, j, k)] - 2 * f[id] + f[IDX(i - 1, j, k)]) * idx2
                               + (f[IDX(i, j + 1, k)] - 2 * f[id] + f[IDX(i, j - 1, k)]) * idy2
                               + (f[IDX(i, j, k + 1)] - 2 * f[id] + f[IDX(i, j, k - 1)]) * idz2;
                    o[id] = f[id] + dt * (-adv + nu_eff * lap);
                }
    }
}

/* apply_velocity_BCs: main.cpp:323-410 */
void orc_velocity_bcs(double* u, double* v, double* w, const uint8_t* solid, const orc_mesh* m)
{
    const int nx = m->nx, ny = m->ny, nz = m->nz, sij = ny * nz;
    for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k) {
            const int id = IDX(0, j, k);
            if (solid[id]) continue;
            u[id] = 1.0; v[id] = 0.0; w[id] = 0.0;
        }
    for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k) {
            const int a = IDX(nx - 1, j, k), b = IDX(nx - 2, j, k);
            u[a] = u[b]; v[a] = v[b]; w[a] = w[b];
        }
    for (int i = 0; i < nx; ++i)
        for (int k = 0; k < nz; ++k) {
            const int a = IDX(i, 0, k), b = IDX(i, ny - 1, k);
            u[a] = 0.0; v[a] = 0.0; w[a] = 0.0;
            u[b] = 0.0; v[b] = 0.0; w[b] = 0.0;
        }
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
            const int a = IDX(i, j, 0), b = IDX(i, j, nz - 1);
            u[a] = 0.0; v[a] = 0.0; w[a] = 0.0;
            u[b] = 0.0; v[b] = 0.0; w[b] = 0.0;
        }
    for (int i = 1; i < nx - 1; ++i)
        for (int j = 1; j < ny - 1; ++j)
            for (int k = 1; k < nz - 1; ++k) {
                const int id = IDX(i, j, k);
                if (solid[id]) { u[id] = 0.0; v[id] = 0.0; w[id] = 0.0; }
            }
}

/* correct_velocity: main.cpp:671-714 */
void orc_correct(double* u, double* v, double* w, const double* p, const orc_mesh* m,
                 double dt, double rho)
{
    const int nx = m->nx, ny = m->ny, nz = m->nz, sij = ny * nz;
    const double dor = dt / rho;
    const double i2x = 1.0 / (2.0 * m->dx), i2y = 1.0 / (2.0 * m->dy), i2z = 1.0 / (2.0 * m->dz);
#pragma omp parallel for
    for (int i = 1; i < nx - 1; ++i)
        for (int j = 1; j < ny - 1; ++j)
            for (int k = 1; k < nz - 1; ++k) {
                const int id = IDX(i, j, k);
                u[id] -= dor * ((p[IDX(i + 1, j, k)] - p[IDX(i - 1, j, k)]) * i2x);
            }
#pragma omp parallel for
    for (int i = 1; i < nx - 1; ++i)
        for (int j = 1; j < ny - 1; ++j)
            for (int k = 1; k < nz - 1; ++k) {
                const int id = IDX(i, j, k);
                v[id] -= dor * ((p[IDX(i, j + 1, k)] - p[IDX(i, j - 1, k)]) * i2y);
            }
#pragma omp parallel for
    for (int i = 1; i < nx - 1; ++i)
        for (int j = 1; j < ny - 1; ++j)
            for (int k = 1; k < nz - 1; ++k) {
                const int id = IDX(i, j, k);
                w[id] -= dor * ((p[IDX(i, j, k + 1)] - p[IDX(i, j, k - 1)]) * i2z);
            }
}

/* ------------------------------------------------------------------ */
/* one step: main.cpp:115-144                                          */
/* ------------------------------------------------------------------ */
void orc_step(orc_fields* f, const orc_mesh* m, const orc_step_params* prm, orc_step_stats* st)
{
    const long n = (long)m->nx * m->ny * m->nz;
    double dt;
    if (prm->dt_fixed > 0) dt = prm->dt_fixed;
    else if (prm->step <= 20) dt = 1e-3;
    else dt = orc_adaptive_dt(f->u, f->v, f->w, f->nu_t, m, prm->nu);

    orc_sgs(f->nu_t, NULL, f->u, f->v, f->w, m, prm->Cs);
    orc_nut_bcs(f->nu_t, f->solid, m);
    orc_momentum(f->u_new, f->v_new, f->w_new, f->u, f->v, f->w, f->nu_t, m, dt, prm->nu);
    double* t;
    t = f->u; f->u = f->u_new; f->u_new = t;
    t = f->v; f->v = f->v_new; f->v_new = t;
    t = f->w; f->w = f->w_new; f->w_new = t;
    orc_velocity_bcs(f->u, f->v, f->w, f->solid, m);

    orc_calculate_divergence(f->p_rhs, f->u_new, f->v_new, f->w_new, f->solid, m, dt, prm->rho);
    double* b = (double*)calloc(n, sizeof(double));
    orc_build_b_vector(b, f->p_rhs, f->solid, m);
    double bn = orc_norm(b, n);
    double tol = prm->tol_rel ? prm->tol * bn : prm->tol;
    int its = orc_pcg_w(prm->A, b, f->p, m->nx, m->ny, m->nz, prm->precond, prm->L, prm->lmin,
                        prm->lmax, prm->degree, prm->max_iter, tol, NULL, prm->W);
    /* final true residual, for reporting only */
    double* q = (double*)calloc(n, sizeof(double));
    orc_matvec(q, f->p, prm->A, m->nx, m->ny, m->nz);
    orc_saxpy(q, 1.0, b, -1.0, q, n);
    double res = orc_norm(q, n);
    free(q);
    free(b);

    orc_correct(f->u, f->v, f->w, f->p, m, dt, prm->rho);
    orc_velocity_bcs(f->u, f->v, f->w, f->solid, m);
    if (st) { st->dt = dt; st->iters = its; st->bnorm = bn; st->res = res; }
}

/* ------------------------------------------------------------------ */
/* differentiation: the transpose of each forward piece                */
/* (no reference counterpart; pinned by central finite differences of  */
/* orc_rollout in tests/test_oracle_adjoint.py)                         */
/* ------------------------------------------------------------------ */
static int interior(int i, int j, int k, int nx, int ny, int nz)
{
    return i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2 && k >= 1 && k <= nz - 2;
}

/* (A^T)_w[m] = A_e[m-sx], (A^T)_e[m] = A_w[m+sx], ... (bounds guarded) */
void orc_transpose_stencil(double* const AT[7], double* const A[7], int nx, int ny, int nz)
{
    const int sij = ny * nz;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const int id = IDX(i, j, k);
                AT[ORC_C][id] = A[ORC_C][id];
                AT[ORC_W][id] = i > 0 ? A[ORC_E][id - sij] : 0.0;
                AT[ORC_E][id] = i < nx - 1 ? A[ORC_W][id + sij] : 0.0;
                AT[ORC_S][id] = j > 0 ? A[ORC_N][id - nz] : 0.0;
                AT[ORC_N][id] = j < ny - 1 ? A[ORC_S][id + nz] : 0.0;
                AT[ORC_B][id] = k > 0 ? A[ORC_T][id - 1] : 0.0;
                AT[ORC_T][id] = k < nz - 1 ? A[ORC_B][id + 1] : 0.0;
            }
}

/* W: 1/2 per direction in which a fluid row of build_stencil's operator is
 * mirrored (pressure.cpp:662-696: one coefficient doubled, the opposite one
 * 0), 1 otherwise.  On grids without obstacles W A is symmetric, so CG in
 * the W inner product is exact CG for A. */
void orc_mirror_weights(double* W, double* const A[7], const orc_mesh* m)
{
    const long n = (long)m->nx * m->ny * m->nz;
    const double ib[3] = {m->dx * m->dx, m->dy * m->dy, m->dz * m->dz};
    const int pa[3] = {ORC_E, ORC_N, ORC_T}, pb[3] = {ORC_W, ORC_S, ORC_B};
    for (long id = 0; id < n; ++id) {
        double wt = 1.0;
        for (int d = 0; d < 3; ++d) {
            const double ma = nearbyint(-A[pa[d]][id] * ib[d]);
            const double mb = nearbyint(-A[pb[d]][id] * ib[d]);
            if ((ma == 2.0 && mb == 0.0) || (mb == 2.0 && ma == 0.0)) wt *= 0.5;
        }
        W[id] = wt;
    }
}

void orc_rollout(double* u, double* v, double* w, double* p, const uint8_t* solid,
                 const orc_mesh* m, double* const A[7], const orc_adj_params* prm, double* traj)
{
    const long n = (long)m->nx * m->ny * m->nz;
    double* un = (double*)calloc(n, sizeof(double));
    double* vn = (double*)calloc(n, sizeof(double));
    double* wn = (double*)calloc(n, sizeof(double));
    double* pr = (double*)calloc(n, sizeof(double));
    double* nt = (double*)calloc(n, sizeof(double));
    orc_fields f = {u, v, w, un, vn, wn, p, pr, nt, (uint8_t*)solid};
    orc_step_params sp = {prm->nu, prm->rho, prm->Cs, 0, prm->dt, prm->precond, prm->degree,
                          prm->max_iter, prm->lmin, prm->lmax, prm->tol_rel, 1, NULL, A, NULL};
    double* W = NULL;
    if (prm->weighted) {
        W = (double*)malloc(n * sizeof(double));
        orc_mirror_weights(W, A, m);
        sp.W = W;
    }
    for (int s = 0; s < prm->nsteps; ++s) {
        if (traj) {
            memcpy(traj + (size_t)s * 3 * n, f.u, n * sizeof(double));
            memcpy(traj + (size_t)s * 3 * n + n, f.v, n * sizeof(double));
            memcpy(traj + (size_t)s * 3 * n + 2 * n, f.w, n * sizeof(double));
        }
        sp.step = s;
        /* each Poisson problem from zero: its exact solution does not depend
         * on the start, and a warm start leaves ||r0|| >> ||b||, which puts a
         * tight relative tolerance out of FP64 reach */
        memset(f.p, 0, n * sizeof(double));
        orc_step(&f, m, &sp, NULL);
    }
    /* orc_step swaps pointers: copy the result into the caller's arrays */
    if (f.u != u) { memcpy(u, f.u, n * sizeof(double)); memcpy(v, f.v, n * sizeof(double));
                    memcpy(w, f.w, n * sizeof(double)); }
    /* free the three buffers that are not the caller's */
    if (f.u == u) { free(f.u_new); free(f.v_new); free(f.w_new); }
    else { free(f.u); free(f.v); free(f.w); }
    free(pr);
    free(nt);
    free(W);
}

/* velocity BCs (main.cpp:323-410) as a map y -> BC(y); transpose in place */
void orc_adj_velocity_bcs(double* bu, double* bv, double* bw, const uint8_t* solid,
                          const orc_mesh* m)
{
    const int nx = m->nx, ny = m->ny, nz = m->nz, sij = ny * nz;
    const long n = (long)nx * sij;
    double* in[3] = {bu, bv, bw};
    double* out = (double*)calloc(n, sizeof(double));
    for (int q = 0; q < 3; ++q) {
        double* b = in[q];
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j)
                for (int k = 0; k < nz; ++k) {
                    const int id = IDX(i, j, k);
                    const int wall = (j == 0 || j == ny - 1 || k == 0 || k == nz - 1);
                    double g = 0.0;
                    if (!wall) {
                        if (i == 0) g = solid[id] ? b[id] : 0.0;   /* solid inlet keeps y */
                        else if (i == nx - 1) g = 0.0;              /* overwritten by the copy */
                        else {
                            g = solid[id] ? 0.0 : b[id];
                            if (i == nx - 2) g += b[id + sij];      /* outlet copies raw y */
                        }
                    }
                    out[id] = g;
                }
        memcpy(b, out, n * sizeof(double));
    }
    free(out);
}

/* correct_velocity (main.cpp:671-714): pbar = (d y / d p)^T ybar */
void orc_adj_correct(double* pbar, const double* yu, const double* yv, const double* yw,
                     const orc_mesh* m, double dt, double rho)
{
    const int nx = m->nx, ny = m->ny, nz = m->nz, sij = ny * nz;
    const double dor = dt / rho;
    const double i2x = 1.0 / (2.0 * m->dx), i2y = 1.0 / (2.0 * m->dy), i2z = 1.0 / (2.0 * m->dz);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const int id = IDX(i, j, k);
                double g = 0.0;
                if (interior(i + 1, j, k, nx, ny, nz)) g += dor * i2x * yu[id + sij];
                if (interior(i - 1, j, k, nx, ny, nz)) g -= dor * i2x * yu[id - sij];
                if (interior(i, j + 1, k, nx, ny, nz)) g += dor * i2y * yv[id + nz];
                if (interior(i, j - 1, k, nx, ny, nz)) g -= dor * i2y * yv[id - nz];
                if (interior(i, j, k + 1, nx, ny, nz)) g += dor * i2z * yw[id + 1];
                if (interior(i, j, k - 1, nx, ny, nz)) g -= dor * i2z * yw[id - 1];
                pbar[id] = g;
            }
}

/* calculate_divergence + build_b_vector (pressure.cpp:256-329, 566-599):
 * (bu,bv,bw) += B^T bbar */
void orc_adj_divergence(double* bu, double* bv, double* bw, const double* bbar,
                        const uint8_t* solid, const orc_mesh* m, double dt, double rho)
{
    const int nx = m->nx, ny = m->ny, nz = m->nz, sij = ny * nz;
    const long n = (long)nx * sij;
    const double scale = rho / dt;
    const double ix = 1.0 / (2.0 * m->dx), iy = 1.0 / (2.0 * m->dy), iz = 1.0 / (2.0 * m->dz);
    const double ix1 = 1.0 / m->dx, iy1 = 1.0 / m->dy, iz1 = 1.0 / m->dz;
    double* beta = (double*)calloc(n, sizeof(double));
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const int id = IDX(i, j, k);
                beta[id] = (solid[id] || i == nx - 1) ? 0.0 : -scale * bbar[id];
            }
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const int id = IDX(i, j, k);
                double gu = 0.0, gv = 0.0, gw = 0.0;
                /* x: forward one-sided at i = 0, central elsewhere */
                if (i == 0) gu -= ix1 * beta[id];
                if (i >= 1) gu += (i - 1 == 0 ? ix1 : ix) * beta[id - sij];
                if (i + 1 <= nx - 1) gu -= ix * beta[id + sij];
                /* y: one-sided at j = 0 and j = ny-1 */
                if (j == 0) gv -= iy1 * beta[id];
                if (j == ny - 1) gv += iy1 * beta[id];
                if (j >= 1) gv += (j - 1 == 0 ? iy1 : iy) * beta[id - nz];
                if (j + 1 <= ny - 1) gv -= (j + 1 == ny - 1 ? iy1 : iy) * beta[id + nz];
                /* z: one-sided at k = 0 and k = nz-1 */
                if (k == 0) gw -= iz1 * beta[id];
                if (k == nz - 1) gw += iz1 * beta[id];
                if (k >= 1) gw += (k - 1 == 0 ? iz1 : iz) * beta[id - 1];
                if (k + 1 <= nz - 1) gw -= (k + 1 == nz - 1 ? iz1 : iz) * beta[id + 1];
                bu[id] += gu;
                bv[id] += gv;
                bw[id] += gw;
            }
    free(beta);
}

/* solve_momentum (main.cpp:563-659) linearised at (u,v,w,nu_t):
 * (bu,bv,bw) += (dM/du)^T mbar, bnut = (dM/dnu_t)^T mbar */
void orc_adj_momentum(double* bu, double* bv, double* bw, double* bnut, const double* mu,
                      const double* mv, const double* mw, const double* u, const double* v,
                      const double* w, const double* nu_t, const orc_mesh* m, double dt,
                      double nu)
{
    const int nx = m->nx, ny = m->ny, nz = m->nz, sij = ny * nz;
    const double bx = 1.0 / (m->dx * m->dx), by = 1.0 / (m->dy * m->dy), bz = 1.0 / (m->dz * m->dz);
    const double ax = 1.0 / (2.0 * m->dx), ay = 1.0 / (2.0 * m->dy), az = 1.0 / (2.0 * m->dz);
    const double* mb[3] = {mu, mv, mw};
    const double* fld[3] = {u, v, w};
    double* out[3] = {bu, bv, bw};
    const int str[3] = {sij, nz, 1};
    const double a[3] = {ax, ay, az}, b[3] = {bx, by, bz};
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const int id = IDX(i, j, k);
                const int ijk[3] = {i, j, k};
                const int in_m = interior(i, j, k, nx, ny, nz);
                /* gradient with respect to f through its own stencil */
                for (int q = 0; q < 3; ++q) {
                    double g = 0.0;
                    if (in_m) g += mb[q][id] * (1.0 + dt * (nu + nu_t[id]) * (-2.0 * (bx + by + bz)));
                    for (int dd = 0; dd < 3; ++dd) {
                        int c;
                        int cc[3] = {ijk[0], ijk[1], ijk[2]};
                        /* c = m - s: m is its +s neighbour */
                        cc[dd] = ijk[dd] - 1;
                        if (interior(cc[0], cc[1], cc[2], nx, ny, nz)) {
                            c = id - str[dd];
                            g += mb[q][c] * dt * (-fld[dd][c] * a[dd] + (nu + nu_t[c]) * b[dd]);
                        }
                        cc[dd] = ijk[dd] + 1;
                        if (interior(cc[0], cc[1], cc[2], nx, ny, nz)) {
                            c = id + str[dd];
                            g += mb[q][c] * dt * (fld[dd][c] * a[dd] + (nu + nu_t[c]) * b[dd]);
                        }
                    }
                    out[q][id] += g;
                }
                if (!in_m) { bnut[id] = 0.0; continue; }
                /* gradient with respect to the advecting velocity and nu_t */
                double gnut = 0.0;
                for (int dd = 0; dd < 3; ++dd) {
                    double gv = 0.0;
                    for (int q = 0; q < 3; ++q) {
                        const double* f = fld[q];
                        const double D = (f[id + str[dd]] - f[id - str[dd]]) * a[dd];
                        gv += mb[q][id] * (-dt * D);
                    }
                    out[dd][id] += gv;
                }
                for (int q = 0; q < 3; ++q) {
                    const double* f = fld[q];
                    double lap = 0.0;
                    for (int dd = 0; dd < 3; ++dd)
                        lap += (f[id + str[dd]] - 2 * f[id] + f[id - str[dd]]) * b[dd];
                    gnut += mb[q][id] * dt * lap;
                }
                bnut[id] = gnut;
            }
}

/* apply_nut_BCs (main.cpp:483-547) as a map from the raw interior SGS
 * values: braw = BC^T bnut */
void orc_adj_nut_bcs(double* braw, const double* bnut, const uint8_t* solid, const orc_mesh* m)
{
    const int nx = m->nx, ny = m->ny, nz = m->nz, sij = ny * nz;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const int id = IDX(i, j, k);
                double g = 0.0;
                if (interior(i, j, k, nx, ny, nz)) {
                    g = solid[id] ? 0.0 : bnut[id];
                    if (i == nx - 2) g += bnut[id + sij];
                }
                braw[id] = g;
            }
}

/* compute_sgs_stress (main.cpp:423-461): (bu,bv,bw) += (d nu_t/d u)^T braw */
void orc_adj_sgs(double* bu, double* bv, double* bw, const double* braw, const double* u,
                 const double* v, const double* w, const orc_mesh* m, double Cs)
{
    const int nx = m->nx, ny = m->ny, nz = m->nz, sij = ny * nz;
    const long n = (long)nx * sij;
    const double csd2 = orc_cs_delta_sq(m, Cs);
    const double t2[3] = {2 * m->dx, 2 * m->dy, 2 * m->dz};
    /* G[a][b] = dL/d(d_b f_a) at each interior cell, f = (u,v,w), d = (x,y,z) */
    double* G = (double*)calloc(9 * n, sizeof(double));
    const double* F[3] = {u, v, w};
    const int str[3] = {sij, nz, 1};
    for (int i = 1; i < nx - 1; ++i)
        for (int j = 1; j < ny - 1; ++j)
            for (int k = 1; k < nz - 1; ++k) {
                const int id = IDX(i, j, k);
                double D[3][3];
                for (int a = 0; a < 3; ++a)
                    for (int bb = 0; bb < 3; ++bb)
                        D[a][bb] = (F[a][id + str[bb]] - F[a][id - str[bb]]) / t2[bb];
                const double Sxx = D[0][0], Syy = D[1][1], Szz = D[2][2];
                const double Sxy = 0.5 * (D[0][1] + D[1][0]);
                const double Sxz = 0.5 * (D[0][2] + D[2][0]);
                const double Syz = 0.5 * (D[1][2] + D[2][1]);
                const double Smag = sqrt(2.0 * (Sxx * Sxx + Syy * Syy + Szz * Szz) +
                                         4.0 * (Sxy * Sxy + Sxz * Sxz + Syz * Syz));
                if (!(Smag > 0.0)) continue;
                const double g = braw[id] * csd2 / Smag;
                double* Gc = G + (size_t)9 * id;
                Gc[0] = 2 * g * Sxx; Gc[4] = 2 * g * Syy; Gc[8] = 2 * g * Szz;
                Gc[1] = Gc[3] = 2 * g * Sxy;   /* du/dy, dv/dx */
                Gc[2] = Gc[6] = 2 * g * Sxz;   /* du/dz, dw/dx */
                Gc[5] = Gc[7] = 2 * g * Syz;   /* dv/dz, dw/dy */
            }
    double* out[3] = {bu, bv, bw};
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const int id = IDX(i, j, k);
                const int ijk[3] = {i, j, k};
                for (int a = 0; a < 3; ++a) {
                    double g = 0.0;
                    for (int bb = 0; bb < 3; ++bb) {
                        int cc[3] = {ijk[0], ijk[1], ijk[2]};
                        cc[bb] = ijk[bb] - 1;   /* c = m - s: m is c's + neighbour */
                        if (interior(cc[0], cc[1], cc[2], nx, ny, nz))
                            g += G[(size_t)9 * (id - str[bb]) + 3 * a + bb] / t2[bb];
                        cc[bb] = ijk[bb] + 1;
                        if (interior(cc[0], cc[1], cc[2], nx, ny, nz))
                            g -= G[(size_t)9 * (id + str[bb]) + 3 * a + bb] / t2[bb];
                    }
                    out[a][id] += g;
                }
            }
    free(G);
}

void orc_adjoint(double* ubar, double* vbar, double* wbar, const double* traj,
                 const uint8_t* solid, const orc_mesh* m, double* const A[7],
                 const orc_adj_params* prm)
{
    const int nx = m->nx, ny = m->ny, nz = m->nz;
    const long n = (long)nx * ny * nz;
    double* AT[7];
    for (int q = 0; q < 7; ++q) AT[q] = (double*)calloc(n, sizeof(double));
    orc_transpose_stencil(AT, A, nx, ny, nz);
    double* W = NULL;
    if (prm->weighted) {
        W = (double*)malloc(n * sizeof(double));
        orc_mirror_weights(W, A, m);
    }
    double* nut = (double*)calloc(n, sizeof(double));
    double* mu = (double*)calloc(n, sizeof(double));
    double* mv = (double*)calloc(n, sizeof(double));
    double* mw = (double*)calloc(n, sizeof(double));
    double* pbar = (double*)calloc(n, sizeof(double));
    double* bbar = (double*)calloc(n, sizeof(double));
    double* bnut = (double*)calloc(n, sizeof(double));
    double* braw = (double*)calloc(n, sizeof(double));
    double* su = (double*)calloc(n, sizeof(double));
    double* sv = (double*)calloc(n, sizeof(double));
    double* sw = (double*)calloc(n, sizeof(double));
    for (int s = prm->nsteps - 1; s >= 0; --s) {
        const double* u = traj + (size_t)s * 3 * n;
        const double* v = u + n;
        const double* w = v + n;
        /* recompute nu_t of the step (sgs + BCs) */
        memset(nut, 0, n * sizeof(double));
        orc_sgs(nut, NULL, u, v, w, m, prm->Cs);
        orc_nut_bcs(nut, solid, m);
        /* 7': y-bar from u^{n+1}-bar */
        orc_adj_velocity_bcs(ubar, vbar, wbar, solid, m);
        /* 6': u*-bar = y-bar; p-bar */
        orc_adj_correct(pbar, ubar, vbar, wbar, m, prm->dt, prm->rho);
        /* 5': b-bar = A^-T p-bar; with W A symmetric, A^-T = W A^-1 W^-1 */
        memset(bbar, 0, n * sizeof(double));
        if (W) {
            for (long q = 0; q < n; ++q) pbar[q] /= W[q];
            const double pn = orc_norm(pbar, n);
            if (pn > 0.0)
                orc_pcg_w(A, pbar, bbar, nx, ny, nz, prm->precond, NULL, prm->lmin, prm->lmax,
                          prm->degree, prm->max_iter, prm->tol_rel * pn, NULL, W);
            for (long q = 0; q < n; ++q) bbar[q] *= W[q];
        } else {
            const double pn = orc_norm(pbar, n);
            if (pn > 0.0)
                orc_pcg(AT, pbar, bbar, nx, ny, nz, prm->precond, NULL, prm->lmin, prm->lmax,
                        prm->degree, prm->max_iter, prm->tol_rel * pn, NULL);
        }
        /* 3': m-bar = BC^T u*-bar */
        memcpy(mu, ubar, n * sizeof(double));
        memcpy(mv, vbar, n * sizeof(double));
        memcpy(mw, wbar, n * sizeof(double));
        orc_adj_velocity_bcs(mu, mv, mw, solid, m);
        /* u^n-bar accumulates the momentum, divergence and SGS paths */
        memset(su, 0, n * sizeof(double));
        memset(sv, 0, n * sizeof(double));
        memset(sw, 0, n * sizeof(double));
        orc_adj_momentum(su, sv, sw, bnut, mu, mv, mw, u, v, w, nut, m, prm->dt, prm->nu);
        orc_adj_divergence(su, sv, sw, bbar, solid, m, prm->dt, prm->rho);
        orc_adj_nut_bcs(braw, bnut, solid, m);
        orc_adj_sgs(su, sv, sw, braw, u, v, w, m, prm->Cs);
        memcpy(ubar, su, n * sizeof(double));
        memcpy(vbar, sv, n * sizeof(double));
        memcpy(wbar, sw, n * sizeof(double));
    }
    for (int q = 0; q < 7; ++q) free(AT[q]);
    free(nut); free(mu); free(mv); free(mw); free(pbar); free(bbar); free(bnut); free(braw);
    free(su); free(sv); free(sw);
    free(W);
}

/* ---- time statistics ----
 * update_statistics (statistics.cpp:98-131): per cell, each sum gains the
 * value or product; the reference's loop order over cells is irrelevant
 * (cells are independent). */
void orc_stats_update(double* const sums[10], const double* u, const double* v, const double* w,
                      const double* p, long n)
{
    for (long i = 0; i < n; ++i) {
        const double a = u[i], b = v[i], c = w[i], d = p[i];
        sums[0][i] += a;
        sums[1][i] += b;
        sums[2][i] += c;
        sums[3][i] += d;
        sums[4][i] += a * a;
        sums[5][i] += b * b;
        sums[6][i] += c * c;
        sums[7][i] += a * b;
        sums[8][i] += a * c;
        sums[9][i] += b * c;
    }
}

/* finalize_and_save's arithmetic (statistics.cpp:141-176): n_inv = 1/k, the
 * means sum*n_inv, the stresses (sum*n_inv) - mean_a*mean_b. */
void orc_stats_finalize(double* const out[10], double* const sums[10], int num_samples, long n)
{
    const double n_inv = 1.0 / (double)num_samples;
    static const int pa[6] = {0, 1, 2, 0, 0, 1}, pb[6] = {0, 1, 2, 1, 2, 2};
    for (long i = 0; i < n; ++i) {
        double mean[4];
        for (int q = 0; q < 4; ++q) {
            mean[q] = sums[q][i] * n_inv;
            out[q][i] = mean[q];
        }
        for (int s = 0; s < 6; ++s)
            out[4 + s][i] = (sums[4 + s][i] * n_inv) - (mean[pa[s]] * mean[pb[s]]);
    }
}
