/* Runs every part of the oracle (oracle/les_oracle.c) once on a small
 * anisotropic grid with an obstacle, so that tests/test_oracle_sanitize.py
 * can build the pair with AddressSanitizer and UndefinedBehaviorSanitizer
 * and check that no access leaves its array and no undefined operation
 * happens.  Prints a checksum of the results. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "les_oracle.h"

#define PI 3.14159265358979323846

static double* vec(long n)
{
    double* p = (double*)calloc((size_t)n, sizeof(double));
    if (!p) exit(3);
    return p;
}

static double sum(const double* a, long n)
{
    double s = 0.0;
    for (long q = 0; q < n; ++q) s += fabs(a[q]);
    return s;
}

int main(void)
{
    const orc_mesh m = {11, 9, 7, 1.0 / 10.0, 1.0 / 8.0, 1.0 / 6.0};
    const long n = (long)m.nx * m.ny * m.nz;
    double check = 0.0;
    uint8_t* solid = (uint8_t*)calloc((size_t)n, 1);
    for (int i = 4; i < 7; ++i)
        for (int j = 3; j < 6; ++j)
            for (int k = 2; k < 5; ++k) solid[(i * m.ny + j) * m.nz + k] = 1;
    solid[((m.nx - 2) * m.ny + 4) * m.nz + 3] = 1; /* a solid next to the outlet */
    double *A[7], *L[7], *AT[7], *tau[6], *sums[10], *outs[10];
    for (int q = 0; q < 7; ++q) {
        A[q] = vec(n);
        L[q] = vec(n);
        AT[q] = vec(n);
    }
    for (int q = 0; q < 6; ++q) tau[q] = vec(n);
    for (int q = 0; q < 10; ++q) {
        sums[q] = vec(n);
        outs[q] = vec(n);
    }
    double *u = vec(n), *v = vec(n), *w = vec(n), *un = vec(n), *vn = vec(n), *wn = vec(n);
    double *p = vec(n), *rhs = vec(n), *nut = vec(n), *b = vec(n), *x = vec(n), *z = vec(n);
    double *W = vec(n), *trace = vec(401);
    unsigned s = 12345u;
    for (long q = 0; q < n; ++q) {
        s = s * 1664525u + 1013904223u;
        u[q] = (double)(s >> 8) / 16777216.0 - 0.5;
        s = s * 1664525u + 1013904223u;
        v[q] = (double)(s >> 8) / 16777216.0 - 0.5;
        s = s * 1664525u + 1013904223u;
        w[q] = (double)(s >> 8) / 16777216.0 - 0.5;
    }

    orc_build_stencil(A, solid, &m);
    orc_matvec(z, u, A, m.nx, m.ny, m.nz);
    check += sum(z, n) + orc_dot(u, v, n) + orc_norm(w, n);
    orc_ic_build(L, A, m.nx, m.ny, m.nz);
    orc_ic_apply(z, u, L, m.nx, m.ny, m.nz);
    check += sum(z, n);
    orc_jacobi_precond(z, u, A[ORC_C], n);
    check += sum(z, n);
    const double lmax = orc_estimate_lambda_max(A, m.nx, m.ny, m.nz, 20);
    orc_cheb_apply(z, u, A, m.nx, m.ny, m.nz, 3.0 * PI * PI, lmax, 5);
    check += sum(z, n) + lmax;

    orc_velocity_bcs(u, v, w, solid, &m);
    orc_sgs(nut, tau, u, v, w, &m, 0.18);
    orc_nut_bcs(nut, solid, &m);
    const double dt = orc_adaptive_dt(u, v, w, nut, &m, 1e-3);
    orc_momentum(un, vn, wn, u, v, w, nut, &m, dt, 1e-3);
    orc_calculate_divergence(rhs, un, vn, wn, solid, &m, dt, 1.2);
    orc_build_b_vector(b, rhs, solid, &m);
    check += sum(nut, n) + sum(b, n) + dt + orc_cs_delta_sq(&m, 0.17);
    for (int pc = ORC_PC_NONE; pc <= ORC_PC_CHEBYSHEV; ++pc) {
        memset(x, 0, (size_t)n * sizeof(double));
        const int it = orc_pcg(A, b, x, m.nx, m.ny, m.nz, pc, L, 3.0 * PI * PI, lmax, 4, 400,
                               1e-8 * orc_norm(b, n), trace);
        check += it + sum(x, n);
    }
    orc_mirror_weights(W, A, &m);
    memset(x, 0, (size_t)n * sizeof(double));
    check += orc_pcg_w(A, b, x, m.nx, m.ny, m.nz, ORC_PC_JACOBI, NULL, 0.0, 0.0, 0, 400,
                       1e-8 * orc_norm(b, n), trace, W);
    orc_correct(un, vn, wn, p, &m, dt, 1.2);
    orc_apply_pressure_bcs(p, solid, &m);
    check += orc_solve_pressure_jacobi(p, un, vn, wn, solid, &m, dt, 1.2);
    check += orc_solve_pressure_gs(p, un, vn, wn, solid, &m, dt, 1.2, 30, 1.7);
    check += orc_solve_pressure_rbsor(p, un, vn, wn, solid, &m, dt, 1.2, 30, 1.7);
    check += sum(p, n);

    orc_fields f = {u, v, w, un, vn, wn, p, rhs, nut, solid};
    orc_step_params sp = {1e-3, 1.2, 0.18, 21, 0.0, ORC_PC_CHEBYSHEV, 4, 400,
                          3.0 * PI * PI, lmax, 1e-8, 1, L, A, NULL};
    orc_step_stats st;
    for (int q = 0; q < 2; ++q) {
        sp.step = 21 + q;
        orc_step(&f, &m, &sp, &st);
        check += st.iters + st.dt;
    }

    /* the differentiation path on an obstacle-free operator */
    uint8_t* none = (uint8_t*)calloc((size_t)n, 1);
    orc_build_stencil(A, none, &m);
    orc_transpose_stencil(AT, A, m.nx, m.ny, m.nz);
    orc_adj_params ap = {1e-3, 1.2, 0.17, 1e-4, 3, ORC_PC_JACOBI, 0.0, 0.0, 0, 2000, 1e-12, 1};
    double* traj = vec(3 * n * ap.nsteps);
    memset(p, 0, (size_t)n * sizeof(double));
    orc_rollout(u, v, w, p, solid, &m, A, &ap, traj);
    for (long q = 0; q < n; ++q) un[q] = vn[q] = wn[q] = 1.0;
    orc_adjoint(un, vn, wn, traj, solid, &m, A, &ap);
    check += sum(un, n) + sum(vn, n) + sum(wn, n) + sum(AT[ORC_E], n);

    for (int q = 0; q < 3; ++q) orc_stats_update(sums, u, v, w, p, n);
    orc_stats_finalize(outs, sums, 3, n);
    for (int q = 0; q < 10; ++q) check += sum(outs[q], n);

    for (int q = 0; q < 7; ++q) {
        free(A[q]);
        free(L[q]);
        free(AT[q]);
    }
    for (int q = 0; q < 6; ++q) free(tau[q]);
    for (int q = 0; q < 10; ++q) {
        free(sums[q]);
        free(outs[q]);
    }
    free(u); free(v); free(w); free(un); free(vn); free(wn); free(p); free(rhs); free(nut);
    free(b); free(x); free(z); free(W); free(trace); free(traj); free(solid); free(none);
    printf("oracle checksum %.17g\n", check);
    return isfinite(check) ? 0 : 1;
}
