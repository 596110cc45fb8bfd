// ref_harness.cpp -- C-callable wrappers around the UNMODIFIED reference
// solver (GabrieLeitao/les-model src/*.cpp), compiled from its own sources by
// oracle/Makefile into oracle/_ref/libref.so.  TEST INFRASTRUCTURE ONLY: it
// pins the oracle (les_oracle.c) and writes the fixtures in tests/golden/.
//
// Each wrapper copies raw arrays into the reference's own types (Field,
// StencilMatrix, Vector; src/field.h, src/pressure.h), calls the reference
// function named in its comment, and copies the results back.
#include <chrono>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>
#include <omp.h>
#include <unistd.h>

#include "field.h"
#include "io.h"
#include "pressure.h"
#include "statistics.h"

double compute_adaptive_dt(const Field& field, const Mesh& mesh, double nu);
void compute_sgs_stress(Field& f, Stress& sgs, const Mesh& m);
void apply_nut_BCs(Field& f, const Mesh& m);
void solve_momentum(Field& f, const Mesh& m, double dt, double nu);
void apply_velocity_BCs(Field& f, const Mesh& m);
void correct_velocity(Field& f, const Mesh& m, double dt, double rho);

namespace {

Mesh mk_mesh(int nx, int ny, int nz, double dx, double dy, double dz) { return Mesh{nx, ny, nz, dx, dy, dz}; }

size_t ncell(const Mesh& m) { return (size_t)m.nx * m.ny * m.nz; }

void put(std::vector<double>& v, const double* s, size_t n)
{
    if (s) std::memcpy(v.data(), s, n * sizeof(double));
}
void get(double* d, const std::vector<double>& v, size_t n)
{
    if (d) std::memcpy(d, v.data(), n * sizeof(double));
}
void put_mask(Field& f, const uint8_t* s, size_t n)
{
    if (!s) return;
    for (size_t i = 0; i < n; ++i) f.is_solid[i] = (s[i] != 0);
}
void put_stencil(StencilMatrix& A, double* const* a, size_t n)
{
    put(A.A_c, a[0], n); put(A.A_n, a[1], n); put(A.A_s, a[2], n); put(A.A_e, a[3], n);
    put(A.A_w, a[4], n); put(A.A_t, a[5], n); put(A.A_b, a[6], n);
}
void get_stencil(double* const* a, const StencilMatrix& A, size_t n)
{
    get(a[0], A.A_c, n); get(a[1], A.A_n, n); get(a[2], A.A_s, n); get(a[3], A.A_e, n);
    get(a[4], A.A_w, n); get(a[5], A.A_t, n); get(a[6], A.A_b, n);
}

// The reference's writers use paths relative to the working directory
// (checkpoint/..., output/...): run them with `dir` as that directory.
struct InDir {
    std::string prev;
    bool ok = false;
    explicit InDir(const char* dir)
    {
        char b[4096];
        if (getcwd(b, sizeof b)) prev = b;
        ok = !prev.empty() && chdir(dir) == 0;
    }
    ~InDir()
    {
        if (ok && chdir(prev.c_str()) != 0) std::abort();
    }
};

Field mk_field(const Mesh& m, const double* u, const double* v, const double* w, const double* p,
               const double* nu_t, const uint8_t* solid)
{
    Field f(m.nx, m.ny, m.nz);
    size_t n = ncell(m);
    put(f.u, u, n); put(f.v, v, n); put(f.w, w, n); put(f.p, p, n); put(f.nu_t, nu_t, n);
    put_mask(f, solid, n);
    return f;
}

}  // namespace

extern "C" {

void ref_set_threads(int n) { omp_set_num_threads(n); }

// build_stencil (pressure.cpp:608)
void ref_build_stencil(double* const* A7, const uint8_t* solid, int nx, int ny, int nz, double dx,
                       double dy, double dz)
{
    Mesh m = mk_mesh(nx, ny, nz, dx, dy, dz);
    Field f(nx, ny, nz);
    put_mask(f, solid, ncell(m));
    StencilMatrix A(ncell(m));
    build_stencil(A, f, m);
    get_stencil(A7, A, ncell(m));
}

// fast_apply_laplacian (pressure.cpp:709)
void ref_matvec(double* q, const double* p, double* const* A7, int nx, int ny, int nz)
{
    size_t n = (size_t)nx * ny * nz;
    Field f(nx, ny, nz);
    StencilMatrix A(n);
    put_stencil(A, A7, n);
    Vector vp(n), vq(n);
    put(vp, p, n);
    fast_apply_laplacian(vq, vp, A, f);
    get(q, vq, n);
}

// calculate_divergence (pressure.cpp:256) reading u_new, v_new, w_new
void ref_divergence(double* rhs, const double* un, const double* vn, const double* wn,
                    const uint8_t* solid, int nx, int ny, int nz, double dx, double dy, double dz,
                    double dt, double rho)
{
    Mesh m = mk_mesh(nx, ny, nz, dx, dy, dz);
    size_t n = ncell(m);
    Field f(nx, ny, nz);
    put_mask(f, solid, n);
    put(f.u_new, un, n); put(f.v_new, vn, n); put(f.w_new, wn, n);
    Vector r(n);
    calculate_divergence(r, f, m, dt, rho);
    get(rhs, r, n);
}

// build_b_vector (pressure.cpp:566)
void ref_build_b(double* b, const double* div, const uint8_t* solid, int nx, int ny, int nz)
{
    Mesh m = mk_mesh(nx, ny, nz, 1, 1, 1);
    size_t n = ncell(m);
    Field f(nx, ny, nz);
    put_mask(f, solid, n);
    Vector vd(n), vb(n);
    put(vd, div, n);
    build_b_vector(vb, vd, f, m);
    get(b, vb, n);
}

// apply_jacobi_preconditioner (pressure.cpp:420)
void ref_jacobi_precond(double* z, const double* r, const double* Ac, long n)
{
    StencilMatrix A(n);
    put(A.A_c, Ac, n);
    Vector vr(n), vz(n);
    put(vr, r, n);
    apply_jacobi_preconditioner(vz, vr, A);
    get(z, vz, n);
}

// build_ic_preconditioner (pressure.cpp:495)
void ref_ic_build(double* const* L7, double* const* A7, int nx, int ny, int nz)
{
    Mesh m = mk_mesh(nx, ny, nz, 1, 1, 1);
    size_t n = ncell(m);
    Field f(nx, ny, nz);
    StencilMatrix A(n), L(n);
    put_stencil(A, A7, n);
    build_ic_preconditioner(L, A, f, m);
    get_stencil(L7, L, n);
}

// apply_ic_preconditioner (pressure.cpp:442)
void ref_ic_apply(double* z, const double* r, double* const* L7, int nx, int ny, int nz)
{
    size_t n = (size_t)nx * ny * nz;
    Field f(nx, ny, nz);
    StencilMatrix L(n);
    put_stencil(L, L7, n);
    Vector vr(n), vz(n);
    put(vr, r, n);
    apply_ic_preconditioner(vz, vr, L, f);
    get(z, vz, n);
}

// apply_chebyshev_preconditioner (pressure.cpp:851)
void ref_cheb_apply(double* z, const double* r, double* const* A7, int nx, int ny, int nz,
                    double lmin, double lmax, int degree)
{
    size_t n = (size_t)nx * ny * nz;
    Field f(nx, ny, nz);
    StencilMatrix A(n);
    put_stencil(A, A7, n);
    Vector vr(n), vz(n);
    put(vr, r, n);
    apply_chebyshev_preconditioner(vz, vr, A, f, lmin, lmax, degree);
    get(z, vz, n);
}

// estimate_lambda_max (pressure.cpp:917)
double ref_lambda_max(double* const* A7, int nx, int ny, int nz, int n_iter)
{
    size_t n = (size_t)nx * ny * nz;
    Field f(nx, ny, nz);
    StencilMatrix A(n);
    put_stencil(A, A7, n);
    return estimate_lambda_max(A, f, n_iter);
}

// solve_pressure_pcg (pressure.cpp:342), the shipped IC-with-A variant
int ref_pcg_legacy(double* const* A7, const double* b, double* x, int nx, int ny, int nz,
                   int max_iter, double tol)
{
    size_t n = (size_t)nx * ny * nz;
    Field f(nx, ny, nz);
    StencilMatrix A(n);
    put_stencil(A, A7, n);
    Vector vb(n), vx(n);
    put(vb, b, n); put(vx, x, n);
    int it = solve_pressure_pcg(A, vb, vx, f, max_iter, tol);
    get(x, vx, n);
    return it;
}

// solve_pressure_pcg_chebyshev (pressure.cpp:762)
int ref_pcg_cheb(double* const* A7, const double* b, double* x, int nx, int ny, int nz,
                 double lmin, double lmax, int degree, int max_iter, double tol)
{
    size_t n = (size_t)nx * ny * nz;
    Field f(nx, ny, nz);
    StencilMatrix A(n);
    put_stencil(A, A7, n);
    Vector vb(n), vx(n);
    put(vb, b, n); put(vx, x, n);
    int it = solve_pressure_pcg_chebyshev(A, vb, vx, f, lmin, lmax, degree, max_iter, tol);
    get(x, vx, n);
    return it;
}

// The PCG loop of pressure.cpp:355-409 with a selectable preconditioner,
// composed from the reference's exported building blocks (its Jacobi lines
// are commented out in the shipped code; this loop restores them).
// precond: 1 Jacobi, 2 IC(0) with L, 3 IC applied with A (as shipped),
// 4 Chebyshev.  res_trace receives ||r_k||, k = 0..iters.
int ref_pcg_harness(double* const* A7, const double* b, double* x, int nx, int ny, int nz,
                    int precond, double* const* L7, double lmin, double lmax, int degree,
                    int max_iter, double tol, double* res_trace)
{
    size_t n = (size_t)nx * ny * nz;
    Field f(nx, ny, nz);
    StencilMatrix A(n), L(n);
    put_stencil(A, A7, n);
    if (precond == 2) put_stencil(L, L7, n);
    Vector vb(n), vx(n), r(n), z(n), p(n), q(n);
    put(vb, b, n); put(vx, x, n);
    auto M = [&](Vector& out, const Vector& in) {
        switch (precond) {
        case 1: apply_jacobi_preconditioner(out, in, A); break;
        case 2: apply_ic_preconditioner(out, in, L, f); break;
        case 3: apply_ic_preconditioner(out, in, A, f); break;
        case 4: apply_chebyshev_preconditioner(out, in, A, f, lmin, lmax, degree); break;
        default: parallel_copy(out, in); break;
        }
    };
    int ret = max_iter;
    fast_apply_laplacian(q, vx, A, f);
    parallel_saxpy(r, 1.0, vb, -1.0, q);
    double e0 = parallel_norm(r);
    if (res_trace) res_trace[0] = e0;
    if (e0 < tol) { ret = 0; get(x, vx, n); return ret; }
    M(p, r);
    double rho = parallel_dot(r, p);
    for (int k = 1; k <= max_iter; ++k) {
        fast_apply_laplacian(q, p, A, f);
        double alpha = rho / parallel_dot(p, q);
        parallel_axpy(vx, alpha, p);
        parallel_axpy(r, -alpha, q);
        double e = parallel_norm(r);
        if (res_trace) res_trace[k] = e;
        if (e < tol) { ret = k; break; }
        M(z, r);
        double rho_new = parallel_dot(r, z);
        double beta = rho_new / rho;
        parallel_saxpy(p, 1.0, z, beta, p);
        rho = rho_new;
    }
    get(x, vx, n);
    return ret;
}

// apply_pressure_BCs (pressure.cpp:14)
void ref_pressure_bcs(double* p, const uint8_t* solid, int nx, int ny, int nz)
{
    Mesh m = mk_mesh(nx, ny, nz, 1, 1, 1);
    size_t n = ncell(m);
    Field f(nx, ny, nz);
    put_mask(f, solid, n);
    put(f.p, p, n);
    apply_pressure_BCs(f, m);
    get(p, f.p, n);
}

// solve_pressure_jacobi (pressure.cpp:106) / solve_pressure_GS (pressure.cpp:186)
void ref_jacobi_solve(double* p, const double* u, const double* v, const double* w,
                      const uint8_t* solid, int nx, int ny, int nz, double dx, double dy,
                      double dz, double dt, double rho)
{
    Mesh m = mk_mesh(nx, ny, nz, dx, dy, dz);
    size_t n = ncell(m);
    Field f(nx, ny, nz);
    put_mask(f, solid, n);
    put(f.p, p, n); put(f.u, u, n); put(f.v, v, n); put(f.w, w, n);
    solve_pressure_jacobi(f, m, dt, rho);
    get(p, f.p, n);
}

void ref_gs_solve(double* p, const double* u, const double* v, const double* w,
                  const uint8_t* solid, int nx, int ny, int nz, double dx, double dy, double dz,
                  double dt, double rho, int max_iter)
{
    Mesh m = mk_mesh(nx, ny, nz, dx, dy, dz);
    size_t n = ncell(m);
    Field f(nx, ny, nz);
    put_mask(f, solid, n);
    put(f.p, p, n); put(f.u, u, n); put(f.v, v, n); put(f.w, w, n);
    solve_pressure_GS(f, m, dt, rho, max_iter);
    get(p, f.p, n);
}

// compute_adaptive_dt (main.cpp:208)
double ref_adaptive_dt(const double* u, const double* v, const double* w, const double* nu_t,
                       int nx, int ny, int nz, double dx, double dy, double dz, double nu)
{
    Mesh m = mk_mesh(nx, ny, nz, dx, dy, dz);
    size_t n = ncell(m);
    Field f(nx, ny, nz);
    put(f.u, u, n); put(f.v, v, n); put(f.w, w, n); put(f.nu_t, nu_t, n);
    return compute_adaptive_dt(f, m, nu);
}

// compute_sgs_stress (main.cpp:423; Cs = 0.18 is fixed there)
void ref_sgs(double* nu_t, double* const* tau6, const double* u, const double* v, const double* w,
             int nx, int ny, int nz, double dx, double dy, double dz)
{
    Mesh m = mk_mesh(nx, ny, nz, dx, dy, dz);
    size_t n = ncell(m);
    Field f(nx, ny, nz);
    put(f.u, u, n); put(f.v, v, n); put(f.w, w, n); put(f.nu_t, nu_t, n);
    Stress s(n);
    compute_sgs_stress(f, s, m);
    get(nu_t, f.nu_t, n);
    if (tau6) {
        get(tau6[0], s.tau_xx, n); get(tau6[1], s.tau_xy, n); get(tau6[2], s.tau_xz, n);
        get(tau6[3], s.tau_yy, n); get(tau6[4], s.tau_yz, n); get(tau6[5], s.tau_zz, n);
    }
}

// apply_nut_BCs (main.cpp:483)
void ref_nut_bcs(double* nu_t, const uint8_t* solid, int nx, int ny, int nz)
{
    Mesh m = mk_mesh(nx, ny, nz, 1, 1, 1);
    size_t n = ncell(m);
    Field f(nx, ny, nz);
    put_mask(f, solid, n);
    put(f.nu_t, nu_t, n);
    apply_nut_BCs(f, m);
    get(nu_t, f.nu_t, n);
}

// solve_momentum (main.cpp:563): on return (u,v,w) and (un,vn,wn) hold the
// reference's Field members after its swap.
void ref_momentum(double* u, double* v, double* w, double* un, double* vn, double* wn,
                  const double* nu_t, int nx, int ny, int nz, double dx, double dy, double dz,
                  double dt, double nu)
{
    Mesh m = mk_mesh(nx, ny, nz, dx, dy, dz);
    size_t n = ncell(m);
    Field f(nx, ny, nz);
    put(f.u, u, n); put(f.v, v, n); put(f.w, w, n);
    put(f.u_new, un, n); put(f.v_new, vn, n); put(f.w_new, wn, n);
    put(f.nu_t, nu_t, n);
    solve_momentum(f, m, dt, nu);
    get(u, f.u, n); get(v, f.v, n); get(w, f.w, n);
    get(un, f.u_new, n); get(vn, f.v_new, n); get(wn, f.w_new, n);
}

// apply_velocity_BCs (main.cpp:323)
void ref_velocity_bcs(double* u, double* v, double* w, const uint8_t* solid, int nx, int ny, int nz)
{
    Mesh m = mk_mesh(nx, ny, nz, 1, 1, 1);
    size_t n = ncell(m);
    Field f(nx, ny, nz);
    put_mask(f, solid, n);
    put(f.u, u, n); put(f.v, v, n); put(f.w, w, n);
    apply_velocity_BCs(f, m);
    get(u, f.u, n); get(v, f.v, n); get(w, f.w, n);
}

// correct_velocity (main.cpp:671)
void ref_correct(double* u, double* v, double* w, const double* p, int nx, int ny, int nz,
                 double dx, double dy, double dz, double dt, double rho)
{
    Mesh m = mk_mesh(nx, ny, nz, dx, dy, dz);
    size_t n = ncell(m);
    Field f(nx, ny, nz);
    put(f.u, u, n); put(f.v, v, n); put(f.w, w, n); put(f.p, p, n);
    correct_velocity(f, m, dt, rho);
    get(u, f.u, n); get(v, f.v, n); get(w, f.w, n);
}

// The step loop body of main.cpp:115-144 run for nsteps steps on one
// persistent Field.  solver: 0 shipped solve_pressure_pcg (IC with A),
// 1 harness PCG-Jacobi, 2 harness PCG-IC(0) with L built from A,
// 4 solve_pressure_pcg_chebyshev, 5 solve_pressure_GS (replacing lines
// 134-137, as main.cpp:133 does), 6 solve_pressure_jacobi.
// tol_rel != 0 passes tol*||b|| as the absolute tolerance.  A7 (may be
// NULL) is the stencil to solve with; NULL builds it from the mask.
// iters_out (nsteps entries, may be NULL) gets the per-step PCG count.
// Fields are in/out: u,v,w,u_new,v_new,w_new,p,p_rhs,nu_t.
int ref_run_steps(double* u, double* v, double* w, double* un, double* vn, double* wn, double* p,
                  double* p_rhs, double* nu_t, const uint8_t* solid, double* const* A7, int nx,
                  int ny, int nz, double dx, double dy, double dz, double nu, double rho,
                  int step0, int nsteps, double dt_fixed, int solver, double lmin, double lmax,
                  int degree, int max_iter, double tol, int tol_rel, int gs_iters,
                  int* iters_out, double* dt_out, double* secs_out)
{
    Mesh m = mk_mesh(nx, ny, nz, dx, dy, dz);
    size_t n = ncell(m);
    Field f(nx, ny, nz);
    put_mask(f, solid, n);
    put(f.u, u, n); put(f.v, v, n); put(f.w, w, n);
    put(f.u_new, un, n); put(f.v_new, vn, n); put(f.w_new, wn, n);
    put(f.p, p, n); put(f.p_rhs, p_rhs, n); put(f.nu_t, nu_t, n);
    Stress sgs(n);
    StencilMatrix A(n), L(n);
    if (A7) put_stencil(A, A7, n);
    else build_stencil(A, f, m);
    if (solver == 2) build_ic_preconditioner(L, A, f, m);
    std::vector<double*> a7 = {A.A_c.data(), A.A_n.data(), A.A_s.data(), A.A_e.data(),
                               A.A_w.data(), A.A_t.data(), A.A_b.data()};
    std::vector<double*> l7 = {L.A_c.data(), L.A_n.data(), L.A_s.data(), L.A_e.data(),
                               L.A_w.data(), L.A_t.data(), L.A_b.data()};
    for (int s = 0; s < nsteps; ++s) {
        // wall time of the step as the driver loop runs it (main.cpp:115-144)
        const auto t_start = std::chrono::steady_clock::now();
        int step = step0 + s;
        double dt;
        if (dt_fixed > 0) dt = dt_fixed;
        else if (step <= 20) dt = 1e-3;
        else dt = compute_adaptive_dt(f, m, nu);
        if (dt_out) dt_out[s] = dt;
        compute_sgs_stress(f, sgs, m);
        apply_nut_BCs(f, m);
        solve_momentum(f, m, dt, nu);
        apply_velocity_BCs(f, m);
        int it = 0;
        if (solver == 5) {
            solve_pressure_GS(f, m, dt, rho, gs_iters);
        } else if (solver == 6) {
            solve_pressure_jacobi(f, m, dt, rho);
        } else {
            calculate_divergence(f.p_rhs, f, m, dt, rho);
            Vector b(n);
            build_b_vector(b, f.p_rhs, f, m);
            double t = tol_rel ? tol * parallel_norm(b) : tol;
            if (solver == 0) it = solve_pressure_pcg(A, b, f.p, f, max_iter, t);
            else if (solver == 4)
                it = solve_pressure_pcg_chebyshev(A, b, f.p, f, lmin, lmax, degree, max_iter, t);
            else
                it = ref_pcg_harness(a7.data(), b.data(), f.p.data(), nx, ny, nz, solver,
                                     l7.data(), lmin, lmax, degree, max_iter, t, nullptr);
        }
        if (iters_out) iters_out[s] = it;
        correct_velocity(f, m, dt, rho);
        apply_velocity_BCs(f, m);
        if (secs_out)
            secs_out[s] = std::chrono::duration<double>(std::chrono::steady_clock::now() -
                                                        t_start).count();
    }
    get(u, f.u, n); get(v, f.v, n); get(w, f.w, n);
    get(un, f.u_new, n); get(vn, f.v_new, n); get(wn, f.w_new, n);
    get(p, f.p, n); get(p_rhs, f.p_rhs, n); get(nu_t, f.nu_t, n);
    return 0;
}

// save_state (io.cpp:14) with `dir` as the working directory (it must hold
// checkpoint/).  Returns -1 if dir cannot be entered.
int ref_save_state(const char* dir, const double* u, const double* v, const double* w,
                   const double* p, const double* nu_t, const uint8_t* solid, int nx, int ny,
                   int nz, double time_simulated, double dt, int next_step)
{
    InDir in(dir);
    if (!in.ok) return -1;
    Mesh m = mk_mesh(nx, ny, nz, 1, 1, 1);
    Field f = mk_field(m, u, v, w, p, nu_t, solid);
    save_state(f, m, time_simulated, dt, next_step);
    return 0;
}

// load_state (io.cpp:61) in `dir`: the arrays are the Field's previous
// content on entry and what load_state leaves in it on return.
int ref_load_state(const char* dir, double* u, double* v, double* w, double* p, double* nu_t,
                   uint8_t* solid, int nx, int ny, int nz, double* time_simulated, double* dt)
{
    InDir in(dir);
    if (!in.ok) return -1;
    Mesh m = mk_mesh(nx, ny, nz, 1, 1, 1);
    Field f = mk_field(m, u, v, w, p, nu_t, solid);
    int step = load_state(f, m, *time_simulated, *dt);
    size_t n = ncell(m);
    get(u, f.u, n); get(v, f.v, n); get(w, f.w, n); get(p, f.p, n); get(nu_t, f.nu_t, n);
    for (size_t i = 0; i < n; ++i) solid[i] = f.is_solid[i] ? 1 : 0;
    return step;
}

// write_field_vtk (io.cpp:225) in `dir` (which must hold output/).
int ref_write_field_vtk(const char* dir, const double* u, const double* v, const double* w,
                        const double* p, const double* nu_t, const uint8_t* solid, int nx, int ny,
                        int nz, double dx, double dy, double dz, int step)
{
    InDir in(dir);
    if (!in.ok) return -1;
    Mesh m = mk_mesh(nx, ny, nz, dx, dy, dz);
    Field f = mk_field(m, u, v, w, p, nu_t, solid);
    write_field_vtk(f, m, step);
    return 0;
}

// init_statistics, then update_statistics (statistics.cpp:98) with state s
// (u[s], v[s], w[s], p[s]) at step steps[s] for s < nstates, then -- when
// dir is given -- finalize_and_save (statistics.cpp:133) to
// dir/output/statistics_final.vtk.  sums10 receives the ten running sums in
// the order u, v, w, p, uu, vv, ww, uv, uw, vw; returns num_samples.
int ref_statistics(const char* dir, int nx, int ny, int nz, double dx, double dy, double dz,
                   int start_step, int nstates, const int* steps, const double* const* u,
                   const double* const* v, const double* const* w, const double* const* p,
                   double* const* sums10)
{
    Mesh m = mk_mesh(nx, ny, nz, dx, dy, dz);
    size_t n = ncell(m);
    Statistics st;
    init_statistics(st, m, start_step);
    Field f(nx, ny, nz);
    for (int s = 0; s < nstates; ++s) {
        put(f.u, u[s], n); put(f.v, v[s], n); put(f.w, w[s], n); put(f.p, p[s], n);
        update_statistics(st, f, steps[s]);
    }
    const std::vector<double>* src[10] = {&st.sum_field.u, &st.sum_field.v, &st.sum_field.w,
                                          &st.sum_field.p, &st.uu_sum, &st.vv_sum, &st.ww_sum,
                                          &st.uv_sum, &st.uw_sum, &st.vw_sum};
    for (int a = 0; a < 10; ++a)
        if (sums10 && sums10[a]) get(sums10[a], *src[a], n);
    if (dir) {
        InDir in(dir);
        if (!in.ok) return -1;
        finalize_and_save(st, m, "output/statistics_final.vtk");
    }
    return st.num_samples;
}

}  // extern "C"
