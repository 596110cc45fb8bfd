// les_adjoint.hip -- reverse-mode kernels of the LES step map for gfx950.
//
// The step (src/main.cpp:115-144) is u^n -> nu_t = SGS(u^n) -> u* =
// BC(M(u^n, nu_t)) -> b = B(u^n) (lagged divergence, src/pressure.cpp:256,
// 566) -> p = A^-1 b -> u^{n+1} = BC(u* - (dt/rho) grad p).  Its vector-
// Jacobian product runs the pieces backwards:
//   y = BC^T g; rhs = W^-1 (grad^T y); x = A^-1 rhs (CG in the W inner
//   product, so W x = A^-T grad^T y);
//   m = BC^T y; g' = M_u^T m + B^T (W x) + SGS^T (BC_nut^T M_nu^T m).
// Every kernel is a gather (each cell sums the terms of the forward cells
// that read it), so no atomics are needed.  All arithmetic is FP64.
#include "les_adjoint.hpp"

#include <math.h>

namespace les {

namespace {

__device__ __forceinline__ bool interior_cell(const Dims& d, int i, int j, int k)
{
    return i >= 1 && i <= d.nx - 2 && j >= 1 && j <= d.ny - 2 && k >= 1 && k <= d.nz - 2;
}

// Forward (main.cpp:323-410): inlet non-solid cells set to constants, the
// outlet plane copies the raw i = nx-2 value, walls zeroed, interior solids
// zeroed; inlet solids keep their value.  Transposed per cell:
//   wall or outlet cell  -> 0
//   inlet cell           -> y if solid, else 0
//   other cell           -> (solid ? 0 : y) (+ y of the outlet cell at i = nx-2)
__global__ __launch_bounds__(kBlock) void k_adj_velocity_bcs(Dims d, Chunking ch,
                                                             const uint8_t* solid, const double* iu,
                                                             const double* iv, const double* iw,
                                                             double* ou, double* ov, double* ow)
{
    const int lb = logical_block(ch.grid);
    FOR_CELLS(d, ch, lb)
    {
        double a = 0.0, b = 0.0, c = 0.0;
        const bool wall = j == 0 || j == d.ny - 1 || k == 0 || k == d.nz - 1;
        if (!wall && i < d.nx - 1) {
            const bool s = solid[id] != 0;
            if (i == 0) {
                if (s) { a = iu[id]; b = iv[id]; c = iw[id]; }
            } else {
                if (!s) { a = iu[id]; b = iv[id]; c = iw[id]; }
                if (i == d.nx - 2) {
                    a += iu[id + d.sij];
                    b += iv[id + d.sij];
                    c += iw[id + d.sij];
                }
            }
        }
        ou[id] = a;
        ov[id] = b;
        ow[id] = c;
    }
}

// correct_velocity (main.cpp:671-714) subtracts (dt/rho) * (p[+] - p[-]) /
// (2 dx) on interior cells; p-bar gathers those terms.
__global__ __launch_bounds__(kBlock) void k_adj_correct(Dims d, Chunking ch, PhysConsts c,
                                                        const double* yu, const double* yv,
                                                        const double* yw, const double* W,
                                                        double* rhs)
{
    const int lb = logical_block(ch.grid);
    const double fx = c.dor * c.inv_2dx, fy = c.dor * c.inv_2dy, fz = c.dor * c.inv_2dz;
    FOR_CELLS(d, ch, lb)
    {
        double g = 0.0;
        if (interior_cell(d, i + 1, j, k)) g += fx * yu[id + d.sij];
        if (interior_cell(d, i - 1, j, k)) g -= fx * yu[id - d.sij];
        if (interior_cell(d, i, j + 1, k)) g += fy * yv[id + d.nz];
        if (interior_cell(d, i, j - 1, k)) g -= fy * yv[id - d.nz];
        if (interior_cell(d, i, j, k + 1)) g += fz * yw[id + 1];
        if (interior_cell(d, i, j, k - 1)) g -= fz * yw[id - 1];
        rhs[id] = W ? g / W[id] : g;
    }
}

// beta = d L / d(divergence) at a cell: build_b_vector (pressure.cpp:566)
// sets b = -rhs on fluid cells, calculate_divergence (pressure.cpp:256)
// zeroes solid cells and the outlet plane; rhs = (rho/dt) * div.
__device__ __forceinline__ double beta_at(const Dims& d, const PhysConsts& c, const double* x,
                                          const double* W, const uint8_t* solid, int id, int i)
{
    if (solid[id] || i == d.nx - 1) return 0.0;
    const double bb = W ? W[id] * x[id] : x[id];
    return -c.div_scale * bb;
}

__global__ __launch_bounds__(kBlock) void k_adj_momentum(Dims d, Chunking ch, PhysConsts c,
                                                         const double* mu, const double* mv,
                                                         const double* mw, const double* u,
                                                         const double* v, const double* w,
                                                         const double* nu_t, const double* x,
                                                         const double* W, const uint8_t* solid,
                                                         double* gu, double* gv, double* gw,
                                                         StrainSens G)
{
    const int lb = logical_block(ch.grid);
    const double bx = c.inv_dx2, by = c.inv_dy2, bz = c.inv_dz2;
    const double ax = c.inv_2dx, ay = c.inv_2dy, az = c.inv_2dz;
    const double nu = c.nu, dt = c.dt;
    const int si = d.sij, sj = d.nz;
    FOR_CELLS(d, ch, lb)
    {
        const bool in_m = interior_cell(d, i, j, k);
        double g[3] = {0.0, 0.0, 0.0};
        const double* mb[3] = {mu, mv, mw};
        // solve_momentum (main.cpp:563-659) at every interior cell c reads
        // f at c and its six neighbours; here f = u, v, w in turn.
        if (in_m) {
            const double self = 1.0 + dt * (nu + nu_t[id]) * (-2.0 * (bx + by + bz));
#pragma unroll
            for (int q = 0; q < 3; ++q) g[q] += mb[q][id] * self;
        }
        // the cell is the +s neighbour of c = id - s, the -s neighbour of c = id + s
#define LES_ADJ_NB(OFF, OK, FLD, A, B, SIGN)                                                     \
        if (OK) {                                                                                \
            const int cc = id + (OFF);                                                           \
            const double coef = dt * ((SIGN) * FLD[cc] * (A) + (nu + nu_t[cc]) * (B));           \
            g[0] += mu[cc] * coef;                                                               \
            g[1] += mv[cc] * coef;                                                               \
            g[2] += mw[cc] * coef;                                                               \
        }
        LES_ADJ_NB(-si, interior_cell(d, i - 1, j, k), u, ax, bx, -1.0)
        LES_ADJ_NB(+si, interior_cell(d, i + 1, j, k), u, ax, bx, 1.0)
        LES_ADJ_NB(-sj, interior_cell(d, i, j - 1, k), v, ay, by, -1.0)
        LES_ADJ_NB(+sj, interior_cell(d, i, j + 1, k), v, ay, by, 1.0)
        LES_ADJ_NB(-1, interior_cell(d, i, j, k - 1), w, az, bz, -1.0)
        LES_ADJ_NB(+1, interior_cell(d, i, j, k + 1), w, az, bz, 1.0)
#undef LES_ADJ_NB
        double Gv[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (in_m) {
            // the advecting velocity: d/d(u_d) of -u_d * (f[+s_d] - f[-s_d]) / (2 d_d)
            const double* fld[3] = {u, v, w};
            const int str[3] = {si, sj, 1};
            const double a3[3] = {ax, ay, az}, b3[3] = {bx, by, bz};
#pragma unroll
            for (int dd = 0; dd < 3; ++dd) {
                double s = 0.0;
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    const double D = (fld[q][id + str[dd]] - fld[q][id - str[dd]]) * a3[dd];
                    s += mb[q][id] * (-dt * D);
                }
                g[dd] += s;
            }
            // nu_t path: d/d(nu_t) of (nu + nu_t) * lap f, then apply_nut_BCs^T
            // (main.cpp:483-547: solid interior cells are zeroed; the outlet
            // copy reads i = nx-2, whose nu_t sensitivity from the outlet
            // cell is zero because momentum reads nu_t on interior cells only)
            double gnut = 0.0;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                double lap = 0.0;
#pragma unroll
                for (int dd = 0; dd < 3; ++dd)
                    lap += (fld[q][id + str[dd]] - 2 * fld[q][id] + fld[q][id - str[dd]]) * b3[dd];
                gnut += mb[q][id] * dt * lap;
            }
            const double braw = solid[id] ? 0.0 : gnut;
            // compute_sgs_stress (main.cpp:436-461): nu_t = (Cs Delta)^2 |S|
            const double dudx = (u[id + si] - u[id - si]) / c.two_dx;
            const double dudy = (u[id + sj] - u[id - sj]) / c.two_dy;
            const double dudz = (u[id + 1] - u[id - 1]) / c.two_dz;
            const double dvdx = (v[id + si] - v[id - si]) / c.two_dx;
            const double dvdy = (v[id + sj] - v[id - sj]) / c.two_dy;
            const double dvdz = (v[id + 1] - v[id - 1]) / c.two_dz;
            const double dwdx = (w[id + si] - w[id - si]) / c.two_dx;
            const double dwdy = (w[id + sj] - w[id - sj]) / c.two_dy;
            const double dwdz = (w[id + 1] - w[id - 1]) / c.two_dz;
            const double Sxx = dudx, Syy = dvdy, Szz = dwdz;
            const double Sxy = 0.5 * (dudy + dvdx);
            const double Sxz = 0.5 * (dudz + dwdx);
            const double Syz = 0.5 * (dvdz + dwdy);
            const double Smag = sqrt(2.0 * (Sxx * Sxx + Syy * Syy + Szz * Szz) +
                                     4.0 * (Sxy * Sxy + Sxz * Sxz + Syz * Syz));
            if (Smag > 0.0) {
                const double gg = braw * c.csd2 / Smag;
                Gv[0] = 2 * gg * Sxx;
                Gv[1] = 2 * gg * Syy;
                Gv[2] = 2 * gg * Szz;
                Gv[3] = 2 * gg * Sxy;
                Gv[4] = 2 * gg * Sxz;
                Gv[5] = 2 * gg * Syz;
            }
        }
#pragma unroll
        for (int q = 0; q < 6; ++q) G.g[q][id] = Gv[q];
        // calculate_divergence^T (pressure.cpp:256-329): central differences
        // on interior faces, one-sided at i = 0, j = 0 / ny-1, k = 0 / nz-1.
        {
            const double ix = c.div_ix, iy = c.div_iy, iz = c.div_iz;
            const double ix1 = c.div_ix1, iy1 = c.div_iy1, iz1 = c.div_iz1;
            double du = 0.0, dv = 0.0, dw = 0.0;
            if (i == 0) du -= ix1 * beta_at(d, c, x, W, solid, id, i);
            if (i >= 1) du += (i - 1 == 0 ? ix1 : ix) * beta_at(d, c, x, W, solid, id - si, i - 1);
            if (i + 1 <= d.nx - 1) du -= ix * beta_at(d, c, x, W, solid, id + si, i + 1);
            if (j == 0) dv -= iy1 * beta_at(d, c, x, W, solid, id, i);
            if (j == d.ny - 1) dv += iy1 * beta_at(d, c, x, W, solid, id, i);
            if (j >= 1) dv += (j - 1 == 0 ? iy1 : iy) * beta_at(d, c, x, W, solid, id - sj, i);
            if (j + 1 <= d.ny - 1)
                dv -= (j + 1 == d.ny - 1 ? iy1 : iy) * beta_at(d, c, x, W, solid, id + sj, i);
            if (k == 0) dw -= iz1 * beta_at(d, c, x, W, solid, id, i);
            if (k == d.nz - 1) dw += iz1 * beta_at(d, c, x, W, solid, id, i);
            if (k >= 1) dw += (k - 1 == 0 ? iz1 : iz) * beta_at(d, c, x, W, solid, id - 1, i);
            if (k + 1 <= d.nz - 1)
                dw -= (k + 1 == d.nz - 1 ? iz1 : iz) * beta_at(d, c, x, W, solid, id + 1, i);
            g[0] += du;
            g[1] += dv;
            g[2] += dw;
        }
        gu[id] = g[0];
        gv[id] = g[1];
        gw[id] = g[2];
    }
}

// G is zero outside interior cells, so the gather only guards the bounds.
__global__ __launch_bounds__(kBlock) void k_adj_sgs_gather(Dims d, Chunking ch, PhysConsts c,
                                                           StrainSens G, double* gu, double* gv,
                                                           double* gw)
{
    const int lb = logical_block(ch.grid);
    const int si = d.sij, sj = d.nz;
    // column b of the symmetric sensitivity for velocity component a
    const double* Gxx = G.g[0];
    const double* Gyy = G.g[1];
    const double* Gzz = G.g[2];
    const double* Gxy = G.g[3];
    const double* Gxz = G.g[4];
    const double* Gyz = G.g[5];
    FOR_CELLS(d, ch, lb)
    {
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        if (i > 0) {
            a0 += Gxx[id - si] / c.two_dx;
            a1 += Gxy[id - si] / c.two_dx;
            a2 += Gxz[id - si] / c.two_dx;
        }
        if (i < d.nx - 1) {
            a0 -= Gxx[id + si] / c.two_dx;
            a1 -= Gxy[id + si] / c.two_dx;
            a2 -= Gxz[id + si] / c.two_dx;
        }
        if (j > 0) {
            a0 += Gxy[id - sj] / c.two_dy;
            a1 += Gyy[id - sj] / c.two_dy;
            a2 += Gyz[id - sj] / c.two_dy;
        }
        if (j < d.ny - 1) {
            a0 -= Gxy[id + sj] / c.two_dy;
            a1 -= Gyy[id + sj] / c.two_dy;
            a2 -= Gyz[id + sj] / c.two_dy;
        }
        if (k > 0) {
            a0 += Gxz[id - 1] / c.two_dz;
            a1 += Gyz[id - 1] / c.two_dz;
            a2 += Gzz[id - 1] / c.two_dz;
        }
        if (k < d.nz - 1) {
            a0 -= Gxz[id + 1] / c.two_dz;
            a1 -= Gyz[id + 1] / c.two_dz;
            a2 -= Gzz[id + 1] / c.two_dz;
        }
        gu[id] += a0;
        gv[id] += a1;
        gw[id] += a2;
    }
}

}  // namespace

hipError_t launch_adj_velocity_bcs(hipStream_t s, const Dims& d, const uint8_t* solid,
                                   const double* iu, const double* iv, const double* iw,
                                   double* ou, double* ov, double* ow)
{
    const Chunking ch = make_chunking(d);
    hipLaunchKernelGGL(k_adj_velocity_bcs, dim3(ch.grid), dim3(kBlock), 0, s, d, ch, solid, iu,
                       iv, iw, ou, ov, ow);
    return hipGetLastError();
}

hipError_t launch_adj_correct(hipStream_t s, const Dims& d, const PhysConsts& c, const double* yu,
                              const double* yv, const double* yw, const double* W, double* rhs)
{
    const Chunking ch = make_chunking(d);
    hipLaunchKernelGGL(k_adj_correct, dim3(ch.grid), dim3(kBlock), 0, s, d, ch, c, yu, yv, yw, W,
                       rhs);
    return hipGetLastError();
}

hipError_t launch_adj_momentum(hipStream_t s, const Dims& d, const PhysConsts& c,
                               const double* mu, const double* mv, const double* mw,
                               const double* u, const double* v, const double* w,
                               const double* nu_t, const double* x, const double* W,
                               const uint8_t* solid, double* gu, double* gv, double* gw,
                               StrainSens G)
{
    const Chunking ch = make_chunking(d);
    hipLaunchKernelGGL(k_adj_momentum, dim3(ch.grid), dim3(kBlock), 0, s, d, ch, c, mu, mv, mw, u,
                       v, w, nu_t, x, W, solid, gu, gv, gw, G);
    return hipGetLastError();
}

hipError_t launch_adj_sgs_gather(hipStream_t s, const Dims& d, const PhysConsts& c, StrainSens G,
                                 double* gu, double* gv, double* gw)
{
    const Chunking ch = make_chunking(d);
    hipLaunchKernelGGL(k_adj_sgs_gather, dim3(ch.grid), dim3(kBlock), 0, s, d, ch, c, G, gu, gv,
                       gw);
    return hipGetLastError();
}

}  // namespace les
