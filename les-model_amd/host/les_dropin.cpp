// les_dropin.cpp -- the reference's pressure.h and physics entry points
// (src/pressure.h:31-63, src/main.cpp:20-29) over the C ABI of liblesgpu.
//
// Each function keeps its reference signature and semantics: inputs are
// read from the caller's std::vector / Field, the work runs in HIP kernels,
// outputs are written back in place.  One device context is kept per grid
// shape and spacing for the life of the process.  Errors from the device
// layer are thrown as std::runtime_error (the reference has no error
// channel; a failed device call must not pass silently).
#include "les/pressure.h"
#include "les/physics.h"
#include "les_gpu.h"

#include <cstdint>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

namespace {

using Key = std::tuple<int, int, int, double, double, double>;

struct Registry {
    std::mutex mu;
    std::map<Key, les_ctx*> ctx;
    ~Registry()
    {
        for (auto& kv : ctx) les_ctx_destroy(kv.second);
    }
};

Registry& registry()
{
    static Registry r;
    return r;
}

void check(les_ctx* c, int rc, const char* what)
{
    if (rc != LES_OK)
        throw std::runtime_error(std::string(what) + ": " + les_last_error(c) + " (" +
                                 std::to_string(rc) + ")");
}

// Context for a grid.  Functions whose reference signature carries no Mesh
// use unit spacings (their arithmetic does not involve the spacings).
les_ctx* ctx_for(int nx, int ny, int nz, double dx = 1.0, double dy = 1.0, double dz = 1.0)
{
    Registry& r = registry();
    std::lock_guard<std::mutex> lk(r.mu);
    Key k{nx, ny, nz, dx, dy, dz};
    auto it = r.ctx.find(k);
    if (it != r.ctx.end()) return it->second;
    les_ctx* c = nullptr;
    check(nullptr, les_ctx_create(nx, ny, nz, dx, dy, dz, &c), "les_ctx_create");
    r.ctx[k] = c;
    return c;
}

les_ctx* ctx_for(const Field& f) { return ctx_for(f.nx, f.ny, f.nz); }
les_ctx* ctx_for(const Mesh& m) { return ctx_for(m.nx, m.ny, m.nz, m.dx, m.dy, m.dz); }

void put(les_ctx* c, int field, const Vector& v)
{
    check(c, les_set_field(c, field, v.data(), LES_PTR_HOST), "les_set_field");
}
void get(les_ctx* c, int field, Vector& v)
{
    check(c, les_get_field(c, field, v.data(), LES_PTR_HOST), "les_get_field");
}
void put_mask(les_ctx* c, const Field& f)
{
    std::vector<uint8_t> m(f.is_solid.size());
    for (size_t i = 0; i < m.size(); ++i) m[i] = f.is_solid[i] ? 1 : 0;
    check(c, les_set_mask(c, m.data(), LES_PTR_HOST), "les_set_mask");
}
void put_stencil(les_ctx* c, int which, const StencilMatrix& A)
{
    double* const a[7] = {const_cast<double*>(A.A_c.data()), const_cast<double*>(A.A_n.data()),
                          const_cast<double*>(A.A_s.data()), const_cast<double*>(A.A_e.data()),
                          const_cast<double*>(A.A_w.data()), const_cast<double*>(A.A_t.data()),
                          const_cast<double*>(A.A_b.data())};
    check(c, les_set_stencil(c, which, a, LES_PTR_HOST), "les_set_stencil");
}
void get_stencil(les_ctx* c, int which, StencilMatrix& A)
{
    double* const a[7] = {A.A_c.data(), A.A_n.data(), A.A_s.data(), A.A_e.data(),
                          A.A_w.data(), A.A_t.data(), A.A_b.data()};
    check(c, les_get_stencil(c, which, a, LES_PTR_HOST), "les_get_stencil");
}

int pcg_common(const StencilMatrix& A, const Vector& b, Vector& x, const Field& field,
               int precond, double lmin, double lmax, int degree, int max_iter, double tol)
{
    les_ctx* c = ctx_for(field);
    put_stencil(c, LES_STENCIL_A, A);
    put(c, LES_F_B, b);
    put(c, LES_F_P, x);
    les_pcg_params prm{};
    prm.precond = precond;
    prm.max_iter = max_iter;
    prm.tol = tol;
    prm.lambda_min = lmin;
    prm.lambda_max = lmax;
    prm.degree = degree;
    prm.precision = LES_PREC_FP64;
    les_pcg_result res{};
    check(c, les_pcg(c, &prm, &res), "les_pcg");
    get(c, LES_F_P, x);
    if (!res.converged && res.iters == max_iter)
        std::cerr << "Warning: PCG did not converge after " << max_iter << " iterations.\n";
    return res.iters;
}

}  // namespace

// ---------------------------------------------------------------------
// pressure.h
// ---------------------------------------------------------------------
void apply_pressure_BCs(Field& f, const Mesh& m)
{
    les_ctx* c = ctx_for(m);
    put_mask(c, f);
    put(c, LES_F_P, f.p);
    check(c, les_pressure_bcs(c), "les_pressure_bcs");
    get(c, LES_F_P, f.p);
}

void solve_pressure_jacobi(Field& f, const Mesh& m, double dt, double rho)
{
    les_ctx* c = ctx_for(m);
    put_mask(c, f);
    put(c, LES_F_U, f.u);
    put(c, LES_F_V, f.v);
    put(c, LES_F_W, f.w);
    put(c, LES_F_P, f.p);
    int sweeps = 0;
    check(c, les_jacobi_solve(c, dt, rho, &sweeps), "les_jacobi_solve");
    get(c, LES_F_P, f.p);
}

void solve_pressure_GS(Field& f, const Mesh& m, double dt, double rho, int maxIter)
{
    les_ctx* c = ctx_for(m);
    put_mask(c, f);
    put(c, LES_F_U, f.u);
    put(c, LES_F_V, f.v);
    put(c, LES_F_W, f.w);
    put(c, LES_F_P, f.p);
    int sweeps = 0;
    check(c, les_sor_solve(c, LES_SOLVER_SOR_LEX, dt, rho, 1.7, maxIter, &sweeps), "les_sor_solve");
    get(c, LES_F_P, f.p);
    // the reference reports early convergence (pressure.cpp:240-244)
    (void)sweeps;
}

int solve_pressure_pcg(const StencilMatrix& A, const Vector& b, Vector& x, const Field& field,
                       int max_iter, double tol)
{
    // the shipped preconditioner: apply_ic_preconditioner with A (pressure.cpp:363, 396)
    return pcg_common(A, b, x, field, LES_PC_LEGACY_IC_A, 0.0, 0.0, 0, max_iter, tol);
}

void calculate_divergence(Vector& divergence, const Field& f, const Mesh& m, double dt, double rho)
{
    les_ctx* c = ctx_for(m);
    put_mask(c, f);
    put(c, LES_F_U_NEW, f.u_new);
    put(c, LES_F_V_NEW, f.v_new);
    put(c, LES_F_W_NEW, f.w_new);
    check(c, les_divergence_b(c, dt, rho), "les_divergence_b");
    get(c, LES_F_P_RHS, divergence);
}

void build_stencil(StencilMatrix& A, const Field& field, const Mesh& m)
{
    les_ctx* c = ctx_for(m);
    put_mask(c, field);
    check(c, les_build_stencil(c, 0), "les_build_stencil");
    get_stencil(c, LES_STENCIL_A, A);
}

void build_b_vector(Vector& b, const Vector& divergence, const Field& field, const Mesh& m)
{
    les_ctx* c = ctx_for(m);
    put_mask(c, field);
    put(c, LES_F_P_RHS, divergence);
    check(c, les_build_b(c), "les_build_b");
    get(c, LES_F_B, b);
}

void fast_apply_laplacian(Vector& q, const Vector& p, const StencilMatrix& A, const Field& field)
{
    les_ctx* c = ctx_for(field);
    put_stencil(c, LES_STENCIL_A, A);
    check(c, les_matvec(c, LES_STENCIL_A, p.data(), q.data(), LES_PTR_HOST), "les_matvec");
}

void apply_jacobi_preconditioner(Vector& z, const Vector& r, const StencilMatrix& A)
{
    // elementwise z = r / A_c (pressure.cpp:420-428); the signature carries
    // no grid, and an elementwise division gains nothing from a device trip
    for (size_t i = 0; i < z.size(); ++i) z[i] = r[i] / A.A_c[i];
}

void apply_ic_preconditioner(Vector& z, const Vector& r, const StencilMatrix& L, const Field& field)
{
    les_ctx* c = ctx_for(field);
    put_stencil(c, LES_STENCIL_L, L);
    // les_precond_apply needs an A to exist in the context; L stands in for it
    put_stencil(c, LES_STENCIL_A, L);
    check(c, les_precond_apply(c, LES_PC_IC0, r.data(), z.data(), LES_PTR_HOST, 0.0, 0.0, 0),
          "les_precond_apply");
}

void build_ic_preconditioner(StencilMatrix& L, const StencilMatrix& A, const Field& field,
                             const Mesh& m)
{
    les_ctx* c = ctx_for(m);
    (void)field;
    put_stencil(c, LES_STENCIL_A, A);
    check(c, les_ic0_factor(c), "les_ic0_factor");
    get_stencil(c, LES_STENCIL_L, L);
}

int solve_pressure_pcg_chebyshev(const StencilMatrix& A, const Vector& b, Vector& x,
                                 const Field& field, double lambda_min, double lambda_max,
                                 int poly_degree, int max_iter, double tol)
{
    return pcg_common(A, b, x, field, LES_PC_CHEBYSHEV, lambda_min, lambda_max, poly_degree,
                      max_iter, tol);
}

void apply_chebyshev_preconditioner(Vector& z, const Vector& r, const StencilMatrix& A,
                                    const Field& field, double lambda_min, double lambda_max,
                                    int poly_degree)
{
    les_ctx* c = ctx_for(field);
    put_stencil(c, LES_STENCIL_A, A);
    check(c, les_precond_apply(c, LES_PC_CHEBYSHEV, r.data(), z.data(), LES_PTR_HOST, lambda_min,
                               lambda_max, poly_degree),
          "les_precond_apply");
}

double estimate_lambda_max(const StencilMatrix& A, const Field& field, int n_iter)
{
    les_ctx* c = ctx_for(field);
    put_stencil(c, LES_STENCIL_A, A);
    double lam = 0.0;
    check(c, les_power_lambda_max(c, n_iter, &lam), "les_power_lambda_max");
    return lam;
}

// BLAS-1 on caller-owned host vectors (pressure.cpp:962-1013)
void parallel_axpy(Vector& v_out, double a, const Vector& v_in)
{
    for (size_t i = 0; i < v_out.size(); ++i) v_out[i] += a * v_in[i];
}

double parallel_dot(const Vector& v1, const Vector& v2)
{
    double sum = 0.0;
    for (size_t i = 0; i < v1.size(); ++i) sum += v1[i] * v2[i];
    return sum;
}

double parallel_norm(const Vector& v) { return std::sqrt(parallel_dot(v, v)); }

void parallel_saxpy(Vector& v_out, double a, const Vector& v_in1, double b, const Vector& v_in2)
{
    for (size_t i = 0; i < v_out.size(); ++i) v_out[i] = a * v_in1[i] + b * v_in2[i];
}

void parallel_copy(Vector& v_out, const Vector& v_in)
{
    for (size_t i = 0; i < v_out.size(); ++i) v_out[i] = v_in[i];
}

// ---------------------------------------------------------------------
// physics (main.cpp:20-29)
// ---------------------------------------------------------------------
double compute_adaptive_dt(const Field& field, const Mesh& mesh, double nu)
{
    les_ctx* c = ctx_for(mesh);
    put(c, LES_F_U, field.u);
    put(c, LES_F_V, field.v);
    put(c, LES_F_W, field.w);
    put(c, LES_F_NU_T, field.nu_t);
    double dt = 0.0;
    check(c, les_adaptive_dt(c, nu, &dt), "les_adaptive_dt");
    return dt;
}

void compute_sgs_stress(Field& f, Stress& sgs, const Mesh& m)
{
    les_ctx* c = ctx_for(m);
    put(c, LES_F_U, f.u);
    put(c, LES_F_V, f.v);
    put(c, LES_F_W, f.w);
    put(c, LES_F_NU_T, f.nu_t);
    double* const tau[6] = {sgs.tau_xx.data(), sgs.tau_xy.data(), sgs.tau_xz.data(),
                            sgs.tau_yy.data(), sgs.tau_yz.data(), sgs.tau_zz.data()};
    check(c, les_sgs(c, 0.18, 0, tau, LES_PTR_HOST), "les_sgs");
    get(c, LES_F_NU_T, f.nu_t);
}

void apply_nut_BCs(Field& f, const Mesh& m)
{
    les_ctx* c = ctx_for(m);
    put_mask(c, f);
    put(c, LES_F_NU_T, f.nu_t);
    check(c, les_nut_bcs(c), "les_nut_bcs");
    get(c, LES_F_NU_T, f.nu_t);
}

void solve_momentum(Field& f, const Mesh& m, double dt, double nu)
{
    les_ctx* c = ctx_for(m);
    put(c, LES_F_U, f.u);
    put(c, LES_F_V, f.v);
    put(c, LES_F_W, f.w);
    put(c, LES_F_U_NEW, f.u_new);
    put(c, LES_F_V_NEW, f.v_new);
    put(c, LES_F_W_NEW, f.w_new);
    put(c, LES_F_NU_T, f.nu_t);
    check(c, les_momentum(c, dt, nu, 0), "les_momentum");
    // the device swapped its buffers as the reference swaps Field's vectors
    get(c, LES_F_U, f.u);
    get(c, LES_F_V, f.v);
    get(c, LES_F_W, f.w);
    get(c, LES_F_U_NEW, f.u_new);
    get(c, LES_F_V_NEW, f.v_new);
    get(c, LES_F_W_NEW, f.w_new);
}

void apply_velocity_BCs(Field& f, const Mesh& m)
{
    les_ctx* c = ctx_for(m);
    put_mask(c, f);
    put(c, LES_F_U, f.u);
    put(c, LES_F_V, f.v);
    put(c, LES_F_W, f.w);
    check(c, les_velocity_bcs(c), "les_velocity_bcs");
    get(c, LES_F_U, f.u);
    get(c, LES_F_V, f.v);
    get(c, LES_F_W, f.w);
}

void correct_velocity(Field& f, const Mesh& m, double dt, double rho)
{
    les_ctx* c = ctx_for(m);
    put(c, LES_F_U, f.u);
    put(c, LES_F_V, f.v);
    put(c, LES_F_W, f.w);
    put(c, LES_F_P, f.p);
    check(c, les_correct(c, dt, rho, 0), "les_correct");
    get(c, LES_F_U, f.u);
    get(c, LES_F_V, f.v);
    get(c, LES_F_W, f.w);
}
