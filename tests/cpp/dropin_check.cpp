// dropin_check.cpp -- exercises the C++ drop-in layer (les/pressure.h,
// les/physics.h over liblespressure_gpu.so) exactly as code written against
// the reference's headers would call it, and dumps every result to a file
// that tests/test_gpu_dropin.py checks against the oracle.
//
// Inputs are generated with integer arithmetic only, so the Python side
// regenerates them bit for bit.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "les/field.h"
#include "les/physics.h"
#include "les/pressure.h"

static double gen(long i, long seed)
{
    return (double)((i * 7919 + seed * 104729) % 2001) / 1000.0 - 1.0;
}

struct Dump {
    FILE* f;
    void put(const char* name, const std::vector<double>& v)
    {
        char nm[32] = {0};
        std::strncpy(nm, name, 31);
        int64_t n = (int64_t)v.size();
        std::fwrite(nm, 1, 32, f);
        std::fwrite(&n, 8, 1, f);
        std::fwrite(v.data(), 8, v.size(), f);
    }
    void put(const char* name, double x) { put(name, std::vector<double>{x}); }
};

int main(int argc, char** argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s out.bin\n", argv[0]);
        return 2;
    }
    Dump d{std::fopen(argv[1], "wb")};
    if (!d.f) return 2;
    Mesh m{12, 10, 9, 1.0 / 12, 0.7 / 10, 1.3 / 9};
    const int N = m.nx * m.ny * m.nz;
    Field f(m.nx, m.ny, m.nz);
    for (int i = 0; i < N; ++i) {
        f.u[i] = gen(i, 1);
        f.v[i] = gen(i, 2);
        f.w[i] = gen(i, 3);
        f.u_new[i] = gen(i, 4);
        f.v_new[i] = gen(i, 5);
        f.w_new[i] = gen(i, 6);
        f.p[i] = gen(i, 7);
        f.nu_t[i] = (gen(i, 8) + 1.0) * 1e-3;
        f.is_solid[i] = ((i * 31) % 17) == 0;
    }
    // operator
    StencilMatrix A(N);
    build_stencil(A, f, m);
    d.put("A_c", A.A_c); d.put("A_n", A.A_n); d.put("A_s", A.A_s); d.put("A_e", A.A_e);
    d.put("A_w", A.A_w); d.put("A_t", A.A_t); d.put("A_b", A.A_b);
    Vector q(N);
    fast_apply_laplacian(q, f.u, A, f);
    d.put("matvec", q);
    Vector div(N), b(N);
    calculate_divergence(div, f, m, 1e-3, 1.2);
    build_b_vector(b, div, f, m);
    d.put("divergence", div);
    d.put("b", b);
    // preconditioners
    Vector z(N);
    apply_jacobi_preconditioner(z, f.v, A);
    d.put("jacobi", z);
    StencilMatrix L(N);
    build_ic_preconditioner(L, A, f, m);
    d.put("L_c", L.A_c); d.put("L_w", L.A_w); d.put("L_s", L.A_s); d.put("L_b", L.A_b);
    apply_ic_preconditioner(z, f.v, L, f);
    d.put("ic_L", z);
    const double lmax = estimate_lambda_max(A, f, 20);
    d.put("lambda_max", lmax);
    apply_chebyshev_preconditioner(z, f.v, A, f, 3 * M_PI * M_PI, lmax, 5);
    d.put("cheb5", z);
    Vector x = f.w;
    int it = solve_pressure_pcg(A, b, x, f, 20, 0.0);
    d.put("pcg_legacy_20", x);
    d.put("pcg_legacy_20_iters", (double)it);
    x = f.w;
    it = solve_pressure_pcg_chebyshev(A, b, x, f, 3 * M_PI * M_PI, lmax, 5, 20, 0.0);
    d.put("pcg_cheb_20", x);
    d.put("pcg_cheb_20_iters", (double)it);
    d.put("dot", parallel_dot(f.u, f.v));
    // stationary
    Field g = f;
    apply_pressure_BCs(g, m);
    d.put("pressure_bcs", g.p);
    g = f;
    for (double& pv : g.p) pv *= 0.01;
    solve_pressure_jacobi(g, m, 1e-3, 1.2);
    d.put("jacobi_solve", g.p);
    g = f;
    for (double& pv : g.p) pv *= 0.01;
    solve_pressure_GS(g, m, 1e-3, 1.2, 10);
    d.put("gs_10", g.p);
    // physics
    d.put("adaptive_dt", compute_adaptive_dt(f, m, 1e-3));
    g = f;
    Stress s(N);
    compute_sgs_stress(g, s, m);
    d.put("sgs_nu_t", g.nu_t);
    d.put("tau_xx", s.tau_xx); d.put("tau_xy", s.tau_xy); d.put("tau_xz", s.tau_xz);
    d.put("tau_yy", s.tau_yy); d.put("tau_yz", s.tau_yz); d.put("tau_zz", s.tau_zz);
    apply_nut_BCs(g, m);
    d.put("nut_bcs", g.nu_t);
    solve_momentum(g, m, 1e-3, 1e-3);
    d.put("mom_u", g.u); d.put("mom_v", g.v); d.put("mom_w", g.w);
    d.put("mom_u_new", g.u_new); d.put("mom_v_new", g.v_new); d.put("mom_w_new", g.w_new);
    apply_velocity_BCs(g, m);
    d.put("vbc_u", g.u); d.put("vbc_v", g.v); d.put("vbc_w", g.w);
    correct_velocity(g, m, 1e-3, 1.2);
    d.put("cor_u", g.u); d.put("cor_v", g.v); d.put("cor_w", g.w);
    std::fclose(d.f);
    std::printf("dropin_check: wrote %s\n", argv[1]);
    return 0;
}
