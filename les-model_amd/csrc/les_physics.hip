// les_physics.hip -- per-step physics kernels for gfx950: Smagorinsky nu_t,
// explicit momentum fused with the divergence/RHS of the lagged projection,
// the velocity correction, boundary conditions, the stationary pressure
// solvers and the adaptive time-step reduction.
//
// Per-cell operand order follows src/main.cpp and src/pressure.cpp so FP64
// results are bit-identical to the reference when built with
// -ffp-contract=off.  Boundary conditions that the reference applies as
// separate ordered passes are folded into the producing kernel as the final
// value each boundary cell ends up with after all of those passes.
#include "les_physics.hpp"

#include <math.h>

#include <algorithm>
#include <cmath>

// 1: the step's SGS and momentum kernels march along i (k_phys_march);
// 0: the row kernels run.
#ifndef LES_PHYS_MARCH
#define LES_PHYS_MARCH 1
#endif

namespace les {

namespace {

__device__ __forceinline__ bool on_wall(const Dims& d, int j, int k)
{
    return j == 0 || j == d.ny - 1 || k == 0 || k == d.nz - 1;
}

// nu_t at an interior cell (main.cpp:436-461).  Gradients divide by 2*dx.
__device__ __forceinline__ double sgs_at(const Dims& d, const PhysConsts& c, const double* u,
                                         const double* v, const double* w, int id, double* S)
{
    const int si = d.sij, sj = d.nz;
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
    if (S) {
        S[0] = Sxx; S[1] = Sxy; S[2] = Sxz; S[3] = Syy; S[4] = Syz; S[5] = Szz;
    }
    return c.csd2 * Smag;
}

// compute_sgs_stress + apply_nut_BCs (main.cpp:423-547) in one pass.
// Final nu_t: walls and inlet 0; outlet = raw interior value at i = nx-2;
// interior solids 0; other interior cells the Smagorinsky value.  tau (the
// six stresses the reference writes and never reads) is optional.
__global__ __launch_bounds__(kBlock) void k_sgs(Dims d, Chunking ch, PhysConsts c, const double* u,
                                                const double* v, const double* w,
                                                const uint8_t* solid, double* nu_t, Tau tau,
                                                int apply_bcs, RedScratch red)
{
    const int lb = logical_block(ch.grid);
    double mx[1] = {0.0};
    FOR_CELLS(d, ch, lb)
    {
        double out;
        if (!apply_bcs) {
            // compute_sgs_stress alone: interior cells, raw values
            if (on_wall(d, j, k) || i == 0 || i == d.nx - 1) continue;
            double S[6];
            const double nt = sgs_at(d, c, u, v, w, id, tau.t[0] ? S : nullptr);
            if (tau.t[0]) {
#pragma unroll
                for (int q = 0; q < 6; ++q) tau.t[q][id] = -2.0 * nt * S[q];
            }
            nu_t[id] = nt;
            continue;
        }
        if (on_wall(d, j, k) || i == 0) {
            out = 0.0;
        } else if (i == d.nx - 1) {
            out = sgs_at(d, c, u, v, w, id - d.sij, nullptr);
        } else {
            double S[6];
            const double nt = sgs_at(d, c, u, v, w, id, tau.t[0] ? S : nullptr);
            if (tau.t[0]) {
#pragma unroll
                for (int q = 0; q < 6; ++q) tau.t[q][id] = -2.0 * nt * S[q];
            }
            out = solid[id] ? 0.0 : nt;
        }
        nu_t[id] = out;
        mx[0] = mx[0] > out ? mx[0] : out;
    }
    if (red.partials) block_partial<1, true>(mx, red.partials, lb);
}

// one component of solve_momentum at an interior cell (main.cpp:584-600)
__device__ __forceinline__ double mom_at(const Dims& d, const PhysConsts& c, const double* f,
                                         const double* u, const double* v, const double* w,
                                         double nu_eff, int id)
{
    const int si = d.sij, sj = d.nz;
    const double fc = f[id];
    const double fe = f[id + si], fw = f[id - si], fn = f[id + sj], fs = f[id - sj];
    const double ft = f[id + 1], fb = f[id - 1];
    const double ddx = (fe - fw) * c.inv_2dx;
    const double ddy = (fn - fs) * c.inv_2dy;
    const double ddz = (ft - fb) * c.inv_2dz;
    const double adv = u[id] * ddx + v[id] * ddy + w[id] * ddz;
    const double lap = (fe - 2 * fc + fw) * c.inv_dx2 + (fn - 2 * fc + fs) * c.inv_dy2 +
                       (ft - 2 * fc + fb) * c.inv_dz2;
    return fc + c.dt * (-adv + nu_eff * lap);
}

// solve_momentum + apply_velocity_BCs (main.cpp:563-659, 323-410) writing
// the provisional velocity into (un, vn, wn), plus calculate_divergence and
// build_b_vector (pressure.cpp:256-329, 566-599) of the step's starting
// velocity (u, v, w) -- the reference's lagged RHS reads exactly that field
// (Field.u_new after its buffer swap).
// Final provisional velocity: walls 0; inlet (1,0,0) except solid inlet
// cells, which keep the buffer's previous content; outlet = raw momentum
// value at i = nx-2; interior solids 0.
__global__ __launch_bounds__(kBlock) void k_momentum(Dims d, Chunking ch, PhysConsts c,
                                                     const double* u, const double* v,
                                                     const double* w, const double* nu_t,
                                                     const uint8_t* solid, double* un, double* vn,
                                                     double* wn, double* p_rhs, double* b,
                                                     int mode)
{
    const int lb = logical_block(ch.grid);
    FOR_CELLS(d, ch, lb)
    {
        const bool sol = solid[id];
        // ---- provisional velocity ----
        if (mode == MOM_RAW) {
            // solve_momentum alone: interior cells, no boundary conditions
            if (!(on_wall(d, j, k) || i == 0 || i == d.nx - 1)) {
                const double nu_eff = c.nu + nu_t[id];
                un[id] = mom_at(d, c, u, u, v, w, nu_eff, id);
                vn[id] = mom_at(d, c, v, u, v, w, nu_eff, id);
                wn[id] = mom_at(d, c, w, u, v, w, nu_eff, id);
            }
            continue;
        }
        if (mode == MOM_DIV_ONLY) {
            // fall through to the divergence only
        } else if (on_wall(d, j, k)) {
            un[id] = 0.0; vn[id] = 0.0; wn[id] = 0.0;
        } else if (i == 0) {
            if (!sol) { un[id] = 1.0; vn[id] = 0.0; wn[id] = 0.0; }
        } else {
            const int src = (i == d.nx - 1) ? id - d.sij : id;
            if (i != d.nx - 1 && sol) {
                un[id] = 0.0; vn[id] = 0.0; wn[id] = 0.0;
            } else {
                const double nu_eff = c.nu + nu_t[src];
                un[id] = mom_at(d, c, u, u, v, w, nu_eff, src);
                vn[id] = mom_at(d, c, v, u, v, w, nu_eff, src);
                wn[id] = mom_at(d, c, w, u, v, w, nu_eff, src);
            }
        }
        // ---- lagged divergence and b ----
        double rhs = 0.0, bv = 0.0;
        if (!sol && i != d.nx - 1) {
            double dudx, dvdy, dwdz;
            if (i == 0) dudx = (u[id + d.sij] - u[id]) * c.div_ix1;
            else dudx = (u[id + d.sij] - u[id - d.sij]) * c.div_ix;
            if (j == 0) dvdy = (v[id + d.nz] - v[id]) * c.div_iy1;
            else if (j == d.ny - 1) dvdy = (v[id] - v[id - d.nz]) * c.div_iy1;
            else dvdy = (v[id + d.nz] - v[id - d.nz]) * c.div_iy;
            if (k == 0) dwdz = (w[id + 1] - w[id]) * c.div_iz1;
            else if (k == d.nz - 1) dwdz = (w[id] - w[id - 1]) * c.div_iz1;
            else dwdz = (w[id + 1] - w[id - 1]) * c.div_iz;
            rhs = c.div_scale * (dudx + dvdy + dwdz);
            bv = -rhs;
        }
        if (p_rhs) p_rhs[id] = rhs;
        if (b) b[id] = bv;
    }
}

// correct_velocity (main.cpp:671-714) on interior cells, with the interior
// solid zeroing of apply_velocity_BCs folded in for i <= nx-3 (solids at
// i = nx-2 are zeroed by k_vel_faces after the outlet copy reads them).
__global__ __launch_bounds__(kBlock) void k_correct(Dims d, Chunking ch, PhysConsts c, double* u,
                                                    double* v, double* w, const double* p,
                                                    const uint8_t* solid, int apply_bcs)
{
    const int lb = logical_block(ch.grid);
    FOR_CELLS(d, ch, lb)
    {
        if (i == 0 || i == d.nx - 1 || on_wall(d, j, k)) continue;
        if (apply_bcs && solid[id] && i <= d.nx - 3) {
            u[id] = 0.0; v[id] = 0.0; w[id] = 0.0;
            continue;
        }
        u[id] -= c.dor * ((p[id + d.sij] - p[id - d.sij]) * c.inv_2dx);
        v[id] -= c.dor * ((p[id + d.nz] - p[id - d.nz]) * c.inv_2dy);
        w[id] -= c.dor * ((p[id + 1] - p[id - 1]) * c.inv_2dz);
    }
}

// zero the given arrays at interior solid cells with 1 <= i <= nx-3
__global__ __launch_bounds__(kBlock) void k_zero_solids(Dims d, Chunking ch, const uint8_t* solid,
                                                        double* a0, double* a1, double* a2)
{
    const int lb = logical_block(ch.grid);
    FOR_CELLS(d, ch, lb)
    {
        if (!solid[id]) continue;
        if (i == 0 || i > d.nx - 3 || on_wall(d, j, k)) continue;
        a0[id] = 0.0;
        if (a1) a1[id] = 0.0;
        if (a2) a2[id] = 0.0;
    }
}

// Boundary cells, each owned by exactly one thread: set X (i = 0 or nx-1,
// all j,k), set Y (j = 0 or ny-1, 1 <= i <= nx-2), set Z (k = 0 or nz-1,
// 1 <= i <= nx-2, 1 <= j <= ny-2).
__device__ __forceinline__ bool face_cell(const Dims& d, int t, int& i, int& j, int& k)
{
    const int nX = 2 * d.ny * d.nz;
    const int nY = 2 * (d.nx - 2) * d.nz;
    const int nZ = 2 * (d.nx - 2) * (d.ny - 2);
    if (t < nX) {
        const int f = t / (d.ny * d.nz), r = t - f * (d.ny * d.nz);
        i = f ? d.nx - 1 : 0;
        j = r / d.nz;
        k = r - j * d.nz;
        return true;
    }
    t -= nX;
    if (t < nY) {
        const int f = t / ((d.nx - 2) * d.nz), r = t - f * ((d.nx - 2) * d.nz);
        j = f ? d.ny - 1 : 0;
        i = 1 + r / d.nz;
        k = r - (i - 1) * d.nz;
        return true;
    }
    t -= nY;
    if (t < nZ) {
        const int f = t / ((d.nx - 2) * (d.ny - 2)), r = t - f * ((d.nx - 2) * (d.ny - 2));
        k = f ? d.nz - 1 : 0;
        i = 1 + r / (d.ny - 2);
        j = 1 + r - (i - 1) * (d.ny - 2);
        return true;
    }
    return false;
}

__host__ __device__ inline int face_count(const Dims& d)
{
    return 2 * d.ny * d.nz + 2 * (d.nx - 2) * d.nz + 2 * (d.nx - 2) * (d.ny - 2);
}

// Final boundary values of apply_velocity_BCs (main.cpp:323-410) /
// apply_nut_BCs (main.cpp:483-547).  For the outlet the thread also zeroes
// the solid cell at i = nx-2 after reading it.
template <bool VEL>
__global__ __launch_bounds__(kBlock) void k_faces(Dims d, const uint8_t* solid, double* a0,
                                                  double* a1, double* a2)
{
    int i, j, k;
    if (!face_cell(d, blockIdx.x * kBlock + threadIdx.x, i, j, k)) return;
    const int id = i * d.sij + j * d.nz + k;
    if (on_wall(d, j, k)) {
        a0[id] = 0.0;
        if (VEL) { a1[id] = 0.0; a2[id] = 0.0; }
    } else if (i == 0) {
        if (VEL) {
            if (!solid[id]) { a0[id] = 1.0; a1[id] = 0.0; a2[id] = 0.0; }
        } else {
            a0[id] = 0.0;
        }
    } else {
        const int src = id - d.sij;
        a0[id] = a0[src];
        if (VEL) { a1[id] = a1[src]; a2[id] = a2[src]; }
        if (solid[src]) {
            a0[src] = 0.0;
            if (VEL) { a1[src] = 0.0; a2[src] = 0.0; }
        }
    }
}

// build_b_vector (pressure.cpp:566-599): b = -div on fluid, 0 on solid and
// outlet cells
__global__ __launch_bounds__(kBlock) void k_build_b(Dims d, Chunking ch, const double* div,
                                                    const uint8_t* solid, double* b)
{
    const int lb = logical_block(ch.grid);
    FOR_CELLS(d, ch, lb) b[id] = (solid[id] || i == d.nx - 1) ? 0.0 : -div[id];
}

// max(u^2+v^2+w^2) and max nu_t over all cells (main.cpp:215-245)
__global__ __launch_bounds__(kBlock) void k_dt_max(Dims d, Chunking ch, const double* u,
                                                   const double* v, const double* w,
                                                   const double* nu_t, RedScratch red)
{
    const int lb = logical_block(ch.grid);
    double mx[2] = {0.0, 0.0};
    FOR_CELLS(d, ch, lb)
    {
        const double s = u[id] * u[id] + v[id] * v[id] + w[id] * w[id];
        mx[0] = mx[0] > s ? mx[0] : s;
        mx[1] = mx[1] > nu_t[id] ? mx[1] : nu_t[id];
    }
    block_partial<2, true>(mx, red.partials, lb);
}

// ------------------------------------------------------------------
// apply_pressure_BCs (pressure.cpp:14-93).  The ordered face passes (inlet
// copy, outlet 0, Y walls, Z walls) leave each boundary cell equal to the
// original value at (i', clamp(j,1,ny-2), clamp(k,1,nz-2)) with i' = 1 at
// the inlet, or 0 at the outlet; the solid pass then copies the first fluid
// neighbour in the order +i,-i,+j,-j,+k,-k.
// ------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_pbc_faces(Dims d, double* p, const StatScalars* st)
{
    if (st && st->done) return;
    int i, j, k;
    if (!face_cell(d, blockIdx.x * kBlock + threadIdx.x, i, j, k)) return;
    const int id = i * d.sij + j * d.nz + k;
    if (i == d.nx - 1) {
        p[id] = 0.0;
        return;
    }
    const int i2 = (i == 0) ? 1 : i;
    const int j2 = j < 1 ? 1 : (j > d.ny - 2 ? d.ny - 2 : j);
    const int k2 = k < 1 ? 1 : (k > d.nz - 2 ? d.nz - 2 : k);
    p[id] = p[i2 * d.sij + j2 * d.nz + k2];
}

__global__ __launch_bounds__(kBlock) void k_pbc_solids(Dims d, Chunking ch, double* p,
                                                       const uint8_t* solid, StatScalars* st,
                                                       int sweep)
{
    if (st && st->done) return;
    const int lb = logical_block(ch.grid);
    FOR_CELLS(d, ch, lb)
    {
        if (!solid[id]) continue;
        if (i == 0 || i == d.nx - 1 || on_wall(d, j, k)) continue;
        if (!solid[id + d.sij]) p[id] = p[id + d.sij];
        else if (!solid[id - d.sij]) p[id] = p[id - d.sij];
        else if (!solid[id + d.nz]) p[id] = p[id + d.nz];
        else if (!solid[id - d.nz]) p[id] = p[id - d.nz];
        else if (!solid[id + 1]) p[id] = p[id + 1];
        else if (!solid[id - 1]) p[id] = p[id - 1];
    }
}

// After a sweep's boundary pass: record the sweep and stop when the max
// update is below the threshold (pressure.cpp:165, 240).
__global__ void k_stat_check(StatScalars* st, int sweep)
{
    if (st->done) return;
    st->sweeps = sweep + 1;
    st->last_max = st->maxr;
    if (st->maxr < st->thresh || sweep + 1 >= st->max_sweeps) st->done = 1;
    st->maxr = 0.0;
}

// rhs = coef * central div(u) on interior cells (pressure.cpp:119-132)
__global__ __launch_bounds__(kBlock) void k_rhs_central(Dims d, Chunking ch, const double* u,
                                                        const double* v, const double* w,
                                                        double* rhs, double coef, double two_dx,
                                                        double two_dy, double two_dz)
{
    const int lb = logical_block(ch.grid);
    FOR_CELLS(d, ch, lb)
    {
        if (i == 0 || i == d.nx - 1 || on_wall(d, j, k)) {
            rhs[id] = 0.0;
            continue;
        }
        const double div = (u[id + d.sij] - u[id - d.sij]) / two_dx +
                           (v[id + d.nz] - v[id - d.nz]) / two_dy +
                           (w[id + 1] - w[id - 1]) / two_dz;
        rhs[id] = coef * div;
    }
}

__device__ __forceinline__ void atomic_max_nonneg(double* addr, double v)
{
    // non-negative doubles order like their bit patterns
    __hip_atomic_fetch_max(reinterpret_cast<unsigned long long*>(addr),
                           (unsigned long long)__double_as_longlong(v), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

// One Jacobi sweep (pressure.cpp:138-157) from pc into pn.
__global__ __launch_bounds__(kBlock) void k_jacobi_sweep(Dims d, Chunking ch, const double* pc,
                                                         double* pn, const double* rhs, double dx2,
                                                         double dy2, double dz2, double denom,
                                                         StatScalars* st)
{
    if (st->done) return;
    const int lb = logical_block(ch.grid);
    double mx[1] = {0.0};
    FOR_CELLS(d, ch, lb)
    {
        if (i == 0 || i == d.nx - 1 || on_wall(d, j, k)) continue;
        const double pold = pc[id];
        const double val = ((pc[id + d.sij] + pc[id - d.sij]) / dx2 +
                            (pc[id + d.nz] + pc[id - d.nz]) / dy2 +
                            (pc[id + 1] + pc[id - 1]) / dz2 - rhs[id]) /
                           denom;
        pn[id] = val;
        const double r = fabs(val - pold);
        mx[0] = mx[0] > r ? mx[0] : r;
    }
    block_reduce<1, true>(mx);
    if (threadIdx.x == 0 && mx[0] > 0.0) atomic_max_nonneg(&st->maxr, mx[0]);
}

// One i+j+k hyperplane of the lexicographic SOR sweep (pressure.cpp:216-234):
// west/south/bottom neighbours already hold this sweep's values and
// east/north/top the previous sweep's, as in the serial loop.
__global__ __launch_bounds__(kBlock) void k_sor_level(Dims d, int level, double* p,
                                                      const double* rhs, double one_m_omega,
                                                      double omega_coef, double dx2, double dy2,
                                                      double dz2, StatScalars* st)
{
    if (st->done) return;
    const int t = blockIdx.x * kBlock + threadIdx.x;
    const int mj = d.ny - 2;
    double r = 0.0;
    if (t < (d.nx - 2) * mj) {
        const int i = 1 + t / mj;
        const int j = 1 + t - (i - 1) * mj;
        const int k = level - i - j;
        if (k >= 1 && k <= d.nz - 2) {
            const int id = i * d.sij + j * d.nz + k;
            const double pold = p[id];
            const double val = one_m_omega * p[id] +
                               omega_coef * ((p[id + d.sij] + p[id - d.sij]) / dx2 +
                                             (p[id + d.nz] + p[id - d.nz]) / dy2 +
                                             (p[id + 1] + p[id - 1]) / dz2 - rhs[id]);
            p[id] = val;
            r = fabs(val - pold);
        }
    }
    double mx[1] = {r};
    block_reduce<1, true>(mx);
    if (threadIdx.x == 0 && mx[0] > 0.0) atomic_max_nonneg(&st->maxr, mx[0]);
}

// One colour of red-black SOR: cells with (i+j+k)%2 == colour.
__global__ __launch_bounds__(kBlock) void k_sor_colour(Dims d, Chunking ch, double* p,
                                                       const double* rhs, int colour,
                                                       double one_m_omega, double omega_coef,
                                                       double dx2, double dy2, double dz2,
                                                       StatScalars* st)
{
    if (st->done) return;
    const int lb = logical_block(ch.grid);
    double mx[1] = {0.0};
    FOR_CELLS(d, ch, lb)
    {
        if (i == 0 || i == d.nx - 1 || on_wall(d, j, k) || ((i + j + k) & 1) != colour) continue;
        const double pold = p[id];
        const double val = one_m_omega * p[id] +
                           omega_coef * ((p[id + d.sij] + p[id - d.sij]) / dx2 +
                                         (p[id + d.nz] + p[id - d.nz]) / dy2 +
                                         (p[id + 1] + p[id - 1]) / dz2 - rhs[id]);
        p[id] = val;
        const double r = fabs(val - pold);
        mx[0] = mx[0] > r ? mx[0] : r;
    }
    block_reduce<1, true>(mx);
    if (threadIdx.x == 0 && mx[0] > 0.0) atomic_max_nonneg(&st->maxr, mx[0]);
}

int face_grid(const Dims& d) { return (face_count(d) + kBlock - 1) / kBlock; }

// ------------------------------------------------------------------
// i-marching forms of k_sgs (with the nu_t BCs) and k_momentum (MOM_STEP)
// ------------------------------------------------------------------
// A workgroup owns R consecutive rows j0 .. j0+R-1 (every k) of a run of
// i-planes and marches along i.  Each thread owns V consecutive cells of
// one row.  The u, v, w values of its cells at planes t-1, t, t+1 sit in
// registers (plane t+2 is loaded while plane t is computed); plane t, with
// one halo row on each side, sits in LDS for the j and k neighbours (two
// buffers, one barrier per plane).  Each value is read from HBM once per
// i-run; the halo rows are the own rows of the neighbouring j-blocks of the
// same i-run, which run on the same XCD, so they come mostly from L2.
// Per cell, the arithmetic and its operand order are those of sgs_at and
// mom_at above, so the results are bit-identical to the row kernels'.
struct MarchGeom {
    int tpr;    // threads per row (nz / V)
    int R;      // rows per workgroup
    int nj;     // j-blocks
    int ni;     // i-runs
    int ilen;   // planes per i-run (the last run ends at nx-1)
    int work;   // nj * ni
    int grid;   // work rounded up to a multiple of kXcd
    int block;  // R * tpr
};

#ifndef LES_MARCH_GRID
#define LES_MARCH_GRID 512
#endif
// workgroups aimed at: two 512-thread workgroups per CU (measured at 256^3
// against one of 1024 threads: SGS 141 -> 131 us, momentum 253 -> 248 us,
// host-timed; 384 workgroups or 768 threads were slower)
constexpr int kMarchGrid = LES_MARCH_GRID;
// cells per thread along k (2: 16-byte loads where nz is even)
#ifndef LES_MARCH_V_SGS
#define LES_MARCH_V_SGS 2
#endif
#ifndef LES_MARCH_V_MOM
#define LES_MARCH_V_MOM 1
#endif
constexpr int kMarchBlock = 1024;
// threads of a march workgroup aimed at (rows R = this / threads per row)
#ifndef LES_MARCH_THREADS
#define LES_MARCH_THREADS 512
#endif

template <int V>
size_t march_lds(const Dims& d, const MarchGeom& g)
{
    return 2 * 3 * (size_t)(g.R + 2) * d.nz * sizeof(double);
}

template <int V>
bool march_geom(const Dims& d, MarchGeom& g)
{
    if (d.nz % V) return false;
    g.tpr = d.nz / V;
    if (g.tpr > LES_MARCH_THREADS / 2) return false;  // R >= 2: a thread loads one halo row at most
    g.R = std::min(LES_MARCH_THREADS / g.tpr, d.ny);
    if (g.R < 2) return false;
    g.block = g.R * g.tpr;
    g.nj = (d.ny + g.R - 1) / g.R;
    int ni = std::max(1, std::min(d.nx, kMarchGrid / g.nj));
    g.ilen = (d.nx + ni - 1) / ni;
    ni = (d.nx + g.ilen - 1) / g.ilen;
    // the outlet plane takes the value computed at nx-2, so no i-run may
    // start at nx-1: the last run absorbs that plane
    if (ni > 1 && (ni - 1) * g.ilen == d.nx - 1) --ni;
    g.ni = ni;
    g.work = g.nj * g.ni;
    g.grid = ((g.work + kXcd - 1) / kXcd) * kXcd;
    return march_lds<V>(d, g) <= kPairMaxLds;
}

template <int V>
__device__ __forceinline__ void ld_v(const double* __restrict__ p, double (&o)[V])
{
    if constexpr (V == 2) {
        const double2 x = *reinterpret_cast<const double2*>(p);
        o[0] = x.x;
        o[1] = x.y;
    } else {
        o[0] = *p;
    }
}

template <int V>
__device__ __forceinline__ void st_v(double* __restrict__ p, const double (&a)[V])
{
    if constexpr (V == 2) {
        *reinterpret_cast<double2*>(p) = make_double2(a[0], a[1]);
    } else {
        *p = a[0];
    }
}

// nu_t from the 18 neighbour values of one cell (sgs_at's expressions).
// P2: every divisor 2 d is a power of two, so x / (2 d) and x * (1/(2 d))
// are the same real number -- the reciprocal is exact -- and round to the
// same double; the multiply replaces the long division sequence.
template <bool P2>
__device__ __forceinline__ double grad_div(double x, double two_d, double inv_2d)
{
    return P2 ? x * inv_2d : x / two_d;
}

template <bool P2>
__device__ __forceinline__ double sgs_from(const PhysConsts& c, double ue, double uw, double un,
                                           double us, double ut, double ub, double ve, double vw,
                                           double vn, double vs, double vt, double vb, double we,
                                           double ww, double wn, double ws, double wt, double wb)
{
    const double dudx = grad_div<P2>(ue - uw, c.two_dx, c.inv_2dx);
    const double dudy = grad_div<P2>(un - us, c.two_dy, c.inv_2dy);
    const double dudz = grad_div<P2>(ut - ub, c.two_dz, c.inv_2dz);
    const double dvdx = grad_div<P2>(ve - vw, c.two_dx, c.inv_2dx);
    const double dvdy = grad_div<P2>(vn - vs, c.two_dy, c.inv_2dy);
    const double dvdz = grad_div<P2>(vt - vb, c.two_dz, c.inv_2dz);
    const double dwdx = grad_div<P2>(we - ww, c.two_dx, c.inv_2dx);
    const double dwdy = grad_div<P2>(wn - ws, c.two_dy, c.inv_2dy);
    const double dwdz = grad_div<P2>(wt - wb, c.two_dz, c.inv_2dz);
    const double Sxx = dudx, Syy = dvdy, Szz = dwdz;
    const double Sxy = 0.5 * (dudy + dvdx);
    const double Sxz = 0.5 * (dudz + dwdx);
    const double Syz = 0.5 * (dvdz + dwdy);
    const double Smag = sqrt(2.0 * (Sxx * Sxx + Syy * Syy + Szz * Szz) +
                             4.0 * (Sxy * Sxy + Sxz * Sxz + Syz * Syz));
    return c.csd2 * Smag;
}

// one momentum component from its 7 stencil values (mom_at's expressions)
__device__ __forceinline__ double mom_from(const PhysConsts& c, double fc, double fe, double fw,
                                           double fn, double fs, double ft, double fb, double uc,
                                           double vc, double wc, double nu_eff)
{
    const double ddx = (fe - fw) * c.inv_2dx;
    const double ddy = (fn - fs) * c.inv_2dy;
    const double ddz = (ft - fb) * c.inv_2dz;
    const double adv = uc * ddx + vc * ddy + wc * ddz;
    const double lap = (fe - 2 * fc + fw) * c.inv_dx2 + (fn - 2 * fc + fs) * c.inv_dy2 +
                       (ft - 2 * fc + fb) * c.inv_dz2;
    return fc + c.dt * (-adv + nu_eff * lap);
}

// The inputs of one plane of a thread's cells: u, v, w (index 0..2).
template <int V>
struct PlaneVals {
    double f[3][V];
};

// Per-plane extras: the halo row's u, v, w, and (own cells) nu_t and the
// mask bytes, loaded one plane ahead.
template <int V>
struct PlaneAux {
    double h[3][V];
    double nt[V];
    uint32_t sm;   // mask bytes of the V cells, unpacked where used
};

template <int V>
__device__ __forceinline__ uint32_t ld_mask(const uint8_t* __restrict__ p)
{
    if constexpr (V == 2) return *reinterpret_cast<const uint16_t*>(p);
    else return *p;
}

// MOM = false: k_sgs with apply_bcs (o0 = nu_t).  MOM = true: k_momentum in
// MOM_STEP mode (o0..o2 = the provisional velocity, p_rhs and b optional).
// The loads for plane t+2 (own cells) and t+1 (halo row, nu_t, mask) are
// issued at the top of plane t and not touched before its end (the mask
// stays packed until it is used), so each has the plane's LDS stores,
// barrier and arithmetic to arrive.
template <int V, bool MOM, bool P2>
__global__ __launch_bounds__(kMarchBlock) void k_phys_march(
    Dims d, MarchGeom g, PhysConsts c, const double* __restrict__ u, const double* __restrict__ v,
    const double* __restrict__ w, const double* __restrict__ nu_t,
    const uint8_t* __restrict__ solid, double* __restrict__ o0, double* __restrict__ o1,
    double* __restrict__ o2, double* __restrict__ p_rhs, double* __restrict__ b)
{
    extern __shared__ unsigned char les_march_smem[];
    double* lds = reinterpret_cast<double*>(les_march_smem);
    const int lb = logical_block(gridDim.x);
    if (lb >= g.work) return;  // padding of the grid to a multiple of kXcd
    const int jb = lb % g.nj, ic = lb / g.nj;
    const int j0 = jb * g.R;
    const int i0 = ic * g.ilen;
    const int i1 = (ic == g.ni - 1) ? d.nx : min(d.nx, i0 + g.ilen);
    const int r = threadIdx.x / g.tpr;
    const int k = (threadIdx.x - r * g.tpr) * V;
    const int j = j0 + r;
    const bool live = j < d.ny;
    const int rows = g.R + 2;
    const int plane = 3 * rows * d.nz;  // doubles per LDS buffer
    // the halo row this thread loads: j0-1 (first row's threads) or j0+R
    // (last row's threads, when that row exists)
    const int hj = (r == 0 && j0 > 0) ? j0 - 1
                 : (r == g.R - 1 && j0 + g.R < d.ny) ? j0 + g.R : -1;
    const int hrow = (r == 0) ? 0 : g.R + 1;
    const bool wall_row = (j == 0 || j == d.ny - 1);
    const long sij = d.sij;
    const long col = (long)j * d.nz + k;
    const long hcol = (long)hj * d.nz + k;
    const double* const src[3] = {u, v, w};

    auto load_own = [&](int i, PlaneVals<V>& P) {
#pragma unroll
        for (int a = 0; a < 3; ++a) ld_v<V>(src[a] + i * sij + col, P.f[a]);
    };
    auto load_aux = [&](int i, PlaneAux<V>& X) {
        if (live) {
            X.sm = ld_mask<V>(solid + i * sij + col);
            if (MOM) ld_v<V>(nu_t + i * sij + col, X.nt);
        }
        if (hj >= 0) {
#pragma unroll
            for (int a = 0; a < 3; ++a) ld_v<V>(src[a] + i * sij + hcol, X.h[a]);
        }
    };

    double keep[MOM ? 3 : 1][V];  // raw values of the previous plane (the outlet copies them)
#pragma unroll
    for (int q = 0; q < V; ++q)
#pragma unroll
        for (int a = 0; a < (MOM ? 3 : 1); ++a) keep[a][q] = 0.0;

    // one plane: P, C, X = planes t-1, t, t+1 of the own cells; Y receives
    // plane t+2; A holds plane t's extras, B receives plane t+1's
    auto body = [&](int t, const PlaneVals<V>& P, const PlaneVals<V>& C, const PlaneVals<V>& X,
                    PlaneVals<V>& Y, const PlaneAux<V>& A, PlaneAux<V>& B) {
        double* L = lds + ((t - i0) & 1) * plane;
        if (t + 1 < i1) {
            if (live && t + 2 < d.nx) load_own(t + 2, Y);
            load_aux(t + 1, B);
        }
        // plane t into LDS: own rows, then the halo row
        if (live) {
#pragma unroll
            for (int a = 0; a < 3; ++a) st_v<V>(L + (a * rows + r + 1) * d.nz + k, C.f[a]);
        }
        if (hj >= 0) {
#pragma unroll
            for (int a = 0; a < 3; ++a) st_v<V>(L + (a * rows + hrow) * d.nz + k, A.h[a]);
        }
        __syncthreads();
        if (!live) return;
        const double* Lu = L + (0 * rows + r + 1) * d.nz;
        const double* Lv = L + (1 * rows + r + 1) * d.nz;
        const double* Lw = L + (2 * rows + r + 1) * d.nz;
        const long id0 = t * sij + col;
        double r0[V], r1[V], r2[V], rhs[V], bq[V];
#pragma unroll
        for (int q = 0; q < V; ++q) {
            const int kk = k + q;
            const bool wall = wall_row || kk == 0 || kk == d.nz - 1;
            const bool sol = (A.sm >> (8 * q)) & 0xff;
            const double cu = C.f[0][q], cv = C.f[1][q], cw = C.f[2][q];
            // neighbours (only read where the reference's stencil reads them)
            const double ub = (q > 0) ? C.f[0][q - 1] : (kk > 0 ? Lu[kk - 1] : 0.0);
            const double ut = (q < V - 1) ? C.f[0][q + 1] : (kk < d.nz - 1 ? Lu[kk + 1] : 0.0);
            const double vb = (q > 0) ? C.f[1][q - 1] : (kk > 0 ? Lv[kk - 1] : 0.0);
            const double vt = (q < V - 1) ? C.f[1][q + 1] : (kk < d.nz - 1 ? Lv[kk + 1] : 0.0);
            const double wb = (q > 0) ? C.f[2][q - 1] : (kk > 0 ? Lw[kk - 1] : 0.0);
            const double wt = (q < V - 1) ? C.f[2][q + 1] : (kk < d.nz - 1 ? Lw[kk + 1] : 0.0);
            const double us = j > 0 ? Lu[kk - d.nz] : 0.0;
            const double un = j < d.ny - 1 ? Lu[kk + d.nz] : 0.0;
            const double vs = j > 0 ? Lv[kk - d.nz] : 0.0;
            const double vn = j < d.ny - 1 ? Lv[kk + d.nz] : 0.0;
            const double ws = j > 0 ? Lw[kk - d.nz] : 0.0;
            const double wn = j < d.ny - 1 ? Lw[kk + d.nz] : 0.0;
            const double ue = X.f[0][q], uw = P.f[0][q];
            const double ve = X.f[1][q], vw = P.f[1][q];
            const double we = X.f[2][q], ww = P.f[2][q];
            if constexpr (!MOM) {
                double out;
                if (wall || t == 0) {
                    out = 0.0;
                } else if (t == d.nx - 1) {
                    out = keep[0][q];
                } else {
                    const double raw = sgs_from<P2>(c, ue, uw, un, us, ut, ub, ve, vw, vn, vs, vt,
                                                    vb, we, ww, wn, ws, wt, wb);
                    keep[0][q] = raw;
                    out = sol ? 0.0 : raw;
                }
                r0[q] = out;
            } else {
                if (wall) {
                    r0[q] = r1[q] = r2[q] = 0.0;
                } else if (t == 0) {
                    r0[q] = 1.0;  // stored below only where the cell is fluid
                    r1[q] = r2[q] = 0.0;
                } else if (t == d.nx - 1) {
                    r0[q] = keep[0][q];
                    r1[q] = keep[1][q];
                    r2[q] = keep[2][q];
                } else {
                    const double nu_eff = c.nu + A.nt[q];
                    const double a0 = mom_from(c, cu, ue, uw, un, us, ut, ub, cu, cv, cw, nu_eff);
                    const double a1 = mom_from(c, cv, ve, vw, vn, vs, vt, vb, cu, cv, cw, nu_eff);
                    const double a2 = mom_from(c, cw, we, ww, wn, ws, wt, wb, cu, cv, cw, nu_eff);
                    keep[0][q] = a0;
                    keep[1][q] = a1;
                    keep[2][q] = a2;
                    r0[q] = sol ? 0.0 : a0;
                    r1[q] = sol ? 0.0 : a1;
                    r2[q] = sol ? 0.0 : a2;
                }
                // lagged divergence of (u, v, w) (calculate_divergence) and b
                // (build_b_vector: -div on fluid cells, +0 elsewhere)
                double rv = 0.0, bv = 0.0;
                if (!sol && t != d.nx - 1) {
                    double dudx, dvdy, dwdz;
                    if (t == 0) dudx = (ue - cu) * c.div_ix1;
                    else dudx = (ue - uw) * c.div_ix;
                    if (j == 0) dvdy = (vn - cv) * c.div_iy1;
                    else if (j == d.ny - 1) dvdy = (cv - vs) * c.div_iy1;
                    else dvdy = (vn - vs) * c.div_iy;
                    if (kk == 0) dwdz = (wt - cw) * c.div_iz1;
                    else if (kk == d.nz - 1) dwdz = (cw - wb) * c.div_iz1;
                    else dwdz = (wt - wb) * c.div_iz;
                    rv = c.div_scale * (dudx + dvdy + dwdz);
                    bv = -rv;
                }
                rhs[q] = rv;
                bq[q] = bv;
            }
        }
        if constexpr (!MOM) {
            st_v<V>(o0 + id0, r0);
        } else {
            if (t == 0) {
                // solid inlet cells keep the buffer's previous content
#pragma unroll
                for (int q = 0; q < V; ++q) {
                    const int kk = k + q;
                    const bool wall = wall_row || kk == 0 || kk == d.nz - 1;
                    if (wall || !((A.sm >> (8 * q)) & 0xff)) {
                        o0[id0 + q] = r0[q];
                        o1[id0 + q] = r1[q];
                        o2[id0 + q] = r2[q];
                    }
                }
            } else {
                st_v<V>(o0 + id0, r0);
                st_v<V>(o1 + id0, r1);
                st_v<V>(o2 + id0, r2);
            }
            if (p_rhs) st_v<V>(p_rhs + id0, rhs);
            if (b) st_v<V>(b + id0, bq);
        }
    };

    PlaneVals<V> P, C, X;
    PlaneAux<V> A;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int q = 0; q < V; ++q) P.f[a][q] = C.f[a][q] = X.f[a][q] = A.h[a][q] = 0.0;
#pragma unroll
    for (int q = 0; q < V; ++q) A.nt[q] = 0.0;
    A.sm = 0;
    if (live) {
        if (i0 > 0) load_own(i0 - 1, P);
        load_own(i0, C);
        if (i0 + 1 < d.nx) load_own(i0 + 1, X);
    }
    load_aux(i0, A);
    for (int t = i0; t < i1; ++t) {
        PlaneVals<V> Y;
        PlaneAux<V> B;
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int q = 0; q < V; ++q) Y.f[a][q] = B.h[a][q] = 0.0;
#pragma unroll
        for (int q = 0; q < V; ++q) B.nt[q] = 0.0;
        B.sm = 0;
        body(t, P, C, X, Y, A, B);
        // the register roles move one plane on; the copies wait for the
        // loads issued at the top of this plane, after its work
        P = C;
        C = X;
        X = Y;
        A = B;
    }
}

// x is a positive power of two (a normal number whose significand is 1)
inline bool pow2(double x)
{
    int e = 0;
    return x > 0.0 && std::isfinite(x) && std::frexp(x, &e) == 0.5 &&
           x >= 2.2250738585072014e-308;
}

template <bool MOM>
hipError_t launch_march(hipStream_t s, const Dims& d, const PhysConsts& c, const double* u,
                        const double* v, const double* w, const double* nu_t,
                        const uint8_t* solid, double* o0, double* o1, double* o2, double* p_rhs,
                        double* b, bool& ran)
{
    ran = false;
    MarchGeom g;
    int V = 0;
    if ((MOM ? LES_MARCH_V_MOM : LES_MARCH_V_SGS) >= 2 && march_geom<2>(d, g)) V = 2;
    else if (march_geom<1>(d, g)) V = 1;
    else return hipSuccess;   // the row kernel runs
    const size_t lds = V == 2 ? march_lds<2>(d, g) : march_lds<1>(d, g);
    // SGS divides its differences by 2 dx, 2 dy, 2 dz: by an exact
    // reciprocal multiply where all three are powers of two
    const bool p2 = !MOM && pow2(c.two_dx) && pow2(c.two_dy) && pow2(c.two_dz) &&
                    c.inv_2dx * c.two_dx == 1.0 && c.inv_2dy * c.two_dy == 1.0 &&
                    c.inv_2dz * c.two_dz == 1.0;
    const void* fn =
        V == 2 ? (p2 ? reinterpret_cast<const void*>(&k_phys_march<2, MOM, true>)
                     : reinterpret_cast<const void*>(&k_phys_march<2, MOM, false>))
               : (p2 ? reinterpret_cast<const void*>(&k_phys_march<1, MOM, true>)
                     : reinterpret_cast<const void*>(&k_phys_march<1, MOM, false>));
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
#define LES_MARCH_LAUNCH(VV, PP)                                                                \
    hipLaunchKernelGGL((k_phys_march<VV, MOM, PP>), dim3(g.grid), dim3(g.block), lds, s, d, g, \
                       c, u, v, w, nu_t, solid, o0, o1, o2, p_rhs, b)
    if (V == 2) {
        if (p2) LES_MARCH_LAUNCH(2, true);
        else LES_MARCH_LAUNCH(2, false);
    } else {
        if (p2) LES_MARCH_LAUNCH(1, true);
        else LES_MARCH_LAUNCH(1, false);
    }
#undef LES_MARCH_LAUNCH
    ran = true;
    return hipGetLastError();
}

}  // namespace

// ------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------
hipError_t launch_sgs(hipStream_t s, const Dims& d, const PhysConsts& c, const double* u,
                      const double* v, const double* w, const uint8_t* solid, double* nu_t,
                      Tau tau, int apply_bcs, RedScratch red, double* nut_max)
{
    const Chunking ch = make_chunking(d);
    RedScratch rd = red;
    if (!(apply_bcs && nut_max)) rd.partials = nullptr;
    if (LES_PHYS_MARCH && apply_bcs && !tau.t[0] && !rd.partials) {
        bool ran = false;
        const hipError_t e = launch_march<false>(s, d, c, u, v, w, nullptr, solid, nu_t, nullptr,
                                                 nullptr, nullptr, nullptr, ran);
        if (e != hipSuccess || ran) return e;
    }
    hipLaunchKernelGGL(k_sgs, dim3(ch.grid), dim3(kBlock), 0, s, d, ch, c, u, v, w, solid, nu_t,
                       tau, apply_bcs, rd);
    if (hipGetLastError() != hipSuccess) return hipErrorLaunchFailure;
    if (rd.partials)
        return launch_finalize(s, ch.grid, 1, true, EPI_STORE, GATE_NONE, red, nullptr, nut_max);
    return hipSuccess;
}

hipError_t launch_momentum(hipStream_t s, const Dims& d, const PhysConsts& c, const double* u,
                           const double* v, const double* w, const double* nu_t,
                           const uint8_t* solid, double* un, double* vn, double* wn,
                           double* p_rhs, double* b, int mode)
{
    if (LES_PHYS_MARCH && mode == MOM_STEP) {
        bool ran = false;
        const hipError_t e = launch_march<true>(s, d, c, u, v, w, nu_t, solid, un, vn, wn, p_rhs,
                                                b, ran);
        if (e != hipSuccess || ran) return e;
    }
    const Chunking ch = make_chunking(d);
    hipLaunchKernelGGL(k_momentum, dim3(ch.grid), dim3(kBlock), 0, s, d, ch, c, u, v, w, nu_t,
                       solid, un, vn, wn, p_rhs, b, mode);
    return hipGetLastError();
}

hipError_t launch_correct(hipStream_t s, const Dims& d, const PhysConsts& c, double* u, double* v,
                          double* w, const double* p, const uint8_t* solid, int apply_bcs)
{
    const Chunking ch = make_chunking(d);
    hipLaunchKernelGGL(k_correct, dim3(ch.grid), dim3(kBlock), 0, s, d, ch, c, u, v, w, p, solid,
                       apply_bcs);
    if (hipGetLastError() != hipSuccess) return hipErrorLaunchFailure;
    if (!apply_bcs) return hipSuccess;
    hipLaunchKernelGGL(k_faces<true>, dim3(face_grid(d)), dim3(kBlock), 0, s, d, solid, u, v, w);
    return hipGetLastError();
}

hipError_t launch_velocity_bcs(hipStream_t s, const Dims& d, const uint8_t* solid, double* u,
                               double* v, double* w)
{
    const Chunking ch = make_chunking(d);
    hipLaunchKernelGGL(k_zero_solids, dim3(ch.grid), dim3(kBlock), 0, s, d, ch, solid, u, v, w);
    if (hipGetLastError() != hipSuccess) return hipErrorLaunchFailure;
    hipLaunchKernelGGL(k_faces<true>, dim3(face_grid(d)), dim3(kBlock), 0, s, d, solid, u, v, w);
    return hipGetLastError();
}

hipError_t launch_nut_bcs(hipStream_t s, const Dims& d, const uint8_t* solid, double* nu_t)
{
    const Chunking ch = make_chunking(d);
    hipLaunchKernelGGL(k_zero_solids, dim3(ch.grid), dim3(kBlock), 0, s, d, ch, solid, nu_t,
                       (double*)nullptr, (double*)nullptr);
    if (hipGetLastError() != hipSuccess) return hipErrorLaunchFailure;
    hipLaunchKernelGGL(k_faces<false>, dim3(face_grid(d)), dim3(kBlock), 0, s, d, solid, nu_t,
                       (double*)nullptr, (double*)nullptr);
    return hipGetLastError();
}

hipError_t launch_build_b(hipStream_t s, const Dims& d, const double* div, const uint8_t* solid,
                          double* b)
{
    const Chunking ch = make_chunking(d);
    hipLaunchKernelGGL(k_build_b, dim3(ch.grid), dim3(kBlock), 0, s, d, ch, div, solid, b);
    return hipGetLastError();
}

hipError_t launch_dt_max(hipStream_t s, const Dims& d, const double* u, const double* v,
                         const double* w, const double* nu_t, RedScratch red, double* out2)
{
    const Chunking ch = make_chunking(d);
    hipLaunchKernelGGL(k_dt_max, dim3(ch.grid), dim3(kBlock), 0, s, d, ch, u, v, w, nu_t, red);
    if (hipGetLastError() != hipSuccess) return hipErrorLaunchFailure;
    return launch_finalize(s, ch.grid, 2, true, EPI_STORE2, GATE_NONE, red, nullptr, out2);
}

hipError_t launch_pressure_bcs(hipStream_t s, const Dims& d, double* p, const uint8_t* solid,
                               StatScalars* st, int sweep)
{
    const Chunking ch = make_chunking(d);
    hipLaunchKernelGGL(k_pbc_faces, dim3(face_grid(d)), dim3(kBlock), 0, s, d, p,
                       (const StatScalars*)st);
    if (hipGetLastError() != hipSuccess) return hipErrorLaunchFailure;
    hipLaunchKernelGGL(k_pbc_solids, dim3(ch.grid), dim3(kBlock), 0, s, d, ch, p, solid, st,
                       sweep);
    if (hipGetLastError() != hipSuccess) return hipErrorLaunchFailure;
    if (st) hipLaunchKernelGGL(k_stat_check, dim3(1), dim3(1), 0, s, st, sweep);
    return hipGetLastError();
}

hipError_t launch_rhs_central(hipStream_t s, const Dims& d, const double* u, const double* v,
                              const double* w, double* rhs, double coef, double two_dx,
                              double two_dy, double two_dz)
{
    const Chunking ch = make_chunking(d);
    hipLaunchKernelGGL(k_rhs_central, dim3(ch.grid), dim3(kBlock), 0, s, d, ch, u, v, w, rhs, coef,
                       two_dx, two_dy, two_dz);
    return hipGetLastError();
}

hipError_t launch_jacobi_sweep(hipStream_t s, const Dims& d, const double* pc, double* pn,
                               const double* rhs, double dx2, double dy2, double dz2,
                               double denom, StatScalars* st)
{
    const Chunking ch = make_chunking(d);
    hipLaunchKernelGGL(k_jacobi_sweep, dim3(ch.grid), dim3(kBlock), 0, s, d, ch, pc, pn, rhs, dx2,
                       dy2, dz2, denom, st);
    return hipGetLastError();
}

hipError_t launch_sor_level(hipStream_t s, const Dims& d, int level, double* p, const double* rhs,
                            double one_m_omega, double omega_coef, double dx2, double dy2,
                            double dz2, StatScalars* st)
{
    const int g = ((d.nx - 2) * (d.ny - 2) + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_sor_level, dim3(g), dim3(kBlock), 0, s, d, level, p, rhs, one_m_omega,
                       omega_coef, dx2, dy2, dz2, st);
    return hipGetLastError();
}

hipError_t launch_sor_colour(hipStream_t s, const Dims& d, double* p, const double* rhs,
                             int colour, double one_m_omega, double omega_coef, double dx2,
                             double dy2, double dz2, StatScalars* st)
{
    const Chunking ch = make_chunking(d);
    hipLaunchKernelGGL(k_sor_colour, dim3(ch.grid), dim3(kBlock), 0, s, d, ch, p, rhs, colour,
                       one_m_omega, omega_coef, dx2, dy2, dz2, st);
    return hipGetLastError();
}

}  // namespace les
