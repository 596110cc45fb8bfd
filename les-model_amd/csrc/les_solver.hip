// les_solver.hip -- pressure-operator and Krylov kernels for gfx950.
//
// Operand order inside every per-cell expression follows the reference
// (src/pressure.cpp) so that, built with -ffp-contract=off, FP64 pointwise
// results are bit-identical to it; only the dot-product summation order
// differs (deterministic two-stage reductions, les_device.hpp).
#include "les_kernels.hpp"

#include <math.h>

namespace les {

namespace {

#define FOR_CELLS(d, ch, lb)                                                         \
    for (int id = (lb) * (ch).chunk + (int)threadIdx.x,                              \
             id_end_ = min((d).N, ((lb) + 1) * (ch).chunk);                          \
         id < id_end_; id += kBlock)

// q = A p at one cell, neighbour order c,w,e,s,n,b,t with the reference's
// bounds guards (pressure.cpp:726-744).  Load(n) returns p at flat index n.
template <typename T, typename Load>
__device__ __forceinline__ T apply7(const Stencil<T>& A, const Dims& d, int id, int i, int j, int k,
                                    Load ld)
{
    T val = A.a[SC][id] * ld(id);
    if (i > 0) val += A.a[SW][id] * ld(id - d.sij);
    if (i < d.nx - 1) val += A.a[SE][id] * ld(id + d.sij);
    if (j > 0) val += A.a[SS][id] * ld(id - d.nz);
    if (j < d.ny - 1) val += A.a[SN][id] * ld(id + d.nz);
    if (k > 0) val += A.a[SB][id] * ld(id - 1);
    if (k < d.nz - 1) val += A.a[ST][id] * ld(id + 1);
    return val;
}

__device__ __forceinline__ void epi_init(const double* tot, SolveScalars* sc)
{
    sc->rr = tot[0];
    sc->bb = tot[1];
    const double e = sqrt(tot[0]);
    sc->err0 = e;
    sc->err = e;
    if (sc->tol_rel > 0.0) sc->tol = sc->tol_rel * sqrt(tot[1]);
    if (sc->trace) sc->trace[0] = e;
    sc->iter = 0;
    if (e < sc->tol) {
        sc->done = 1;
        sc->converged = 1;
    } else if (!(e == e) || sc->max_iter <= 0) {
        sc->done = 1;
        sc->fault = !(e == e);
    }
}

__device__ void epilogue(int epi, const double* tot, SolveScalars* sc, double* out)
{
    switch (epi) {
    case EPI_STORE:
        out[0] = tot[0];
        break;
    case EPI_INIT:
        epi_init(tot, sc);
        break;
    case EPI_RHO:
        sc->rho = tot[0];
        break;
    case EPI_ALPHA:
        sc->pq = tot[0];
        sc->alpha = sc->rho / tot[0];
        break;
    case EPI_CONV:
    case EPI_CONV_RZ: {
        const int it = sc->iter + 1;
        sc->iter = it;
        sc->rr = tot[0];
        const double e = sqrt(tot[0]);
        sc->err = e;
        if (sc->trace) sc->trace[it] = e;
        if (e < sc->tol) {
            sc->done = 1;
            sc->converged = 1;
        } else if (!(e == e)) {
            sc->done = 1;
            sc->fault = 1;
        } else if (it >= sc->max_iter) {
            sc->done = 1;
        } else if (epi == EPI_CONV_RZ) {
            sc->beta = tot[1] / sc->rho;
            sc->rho = tot[1];
        }
        break;
    }
    case EPI_BETA:
        sc->beta = tot[0] / sc->rho;
        sc->rho = tot[0];
        break;
    case EPI_CONV_MIXED: {
        const int it = sc->iter + 1;
        sc->iter = it;
        sc->rr = tot[0];
        const double e = sqrt(tot[0]);
        sc->err = e;
        if (sc->trace) sc->trace[it] = e;
        if (!(e == e)) {
            sc->done = 1;
            sc->fault = 1;
            break;
        }
        if (e < sc->tol || (sc->reliable_every > 0 && it % sc->reliable_every == 0))
            sc->reliable_now = 1;
        if (it >= sc->max_iter) {
            sc->reliable_now = 1;
            sc->stop_pending = 1;
        }
        break;
    }
    case EPI_RELIABLE: {
        const double e = sqrt(tot[0]);
        sc->true_err = e;
        sc->err = e;
        sc->rr = tot[0];
        sc->replacements += 1;
        sc->reliable_now = 0;
        if (e < sc->tol) {
            sc->done = 1;
            sc->converged = 1;
        } else if (!(e == e)) {
            sc->done = 1;
            sc->fault = 1;
        } else if (sc->stop_pending) {
            sc->done = 1;
        }
        break;
    }
    case EPI_INIT_MIXED:
        epi_init(tot, sc);
        sc->true_err = sc->err0;
        break;
    default:
        break;
    }
}

__device__ __forceinline__ bool is_done(const SolveScalars* sc) { return sc && sc->done; }

// ------------------------------------------------------------------
// build_stencil (pressure.cpp:608-699)
// ------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_build_stencil(Dims d, Chunking ch, const uint8_t* solid,
                                                          double* Ac, double* An, double* As,
                                                          double* Ae, double* Aw, double* At,
                                                          double* Ab, double idx2, double idy2,
                                                          double idz2)
{
    const int lb = logical_block(ch.grid);
    const double e_b = -idx2, w_b = -idx2, n_b = -idy2, s_b = -idy2, t_b = -idz2, b_b = -idz2;
    const double c_b = (2.0 * idx2 + 2.0 * idy2 + 2.0 * idz2);
    FOR_CELLS(d, ch, lb)
    {
        int i, j, k;
        decode(d, id, i, j, k);
        double c = 1.0, e = 0.0, w = 0.0, n = 0.0, s = 0.0, t = 0.0, b = 0.0;
        if (!solid[id] && i != d.nx - 1) {
            c = c_b; e = e_b; w = w_b; n = n_b; s = s_b; t = t_b; b = b_b;
            if (i == 0) { e += w_b; w = 0.0; }
            else if (solid[id - d.sij]) { e += w_b; w = 0.0; }
            if (i == d.nx - 2) { e = 0.0; }
            else if (solid[id + d.sij]) { w += e_b; e = 0.0; }
            if (j == 0) { n += s_b; s = 0.0; }
            else if (solid[id - d.nz]) { n += s_b; s = 0.0; }
            if (j == d.ny - 1) { s += n_b; n = 0.0; }
            else if (solid[id + d.nz]) { s += n_b; n = 0.0; }
            if (k == 0) { t += b_b; b = 0.0; }
            else if (solid[id - 1]) { t += b_b; b = 0.0; }
            if (k == d.nz - 1) { b += t_b; t = 0.0; }
            else if (solid[id + 1]) { b += t_b; t = 0.0; }
        }
        Ac[id] = c; An[id] = n; As[id] = s; Ae[id] = e; Aw[id] = w; At[id] = t; Ab[id] = b;
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_matvec(Dims d, Chunking ch, Stencil<T> A, const T* x,
                                                   T* y)
{
    const int lb = logical_block(ch.grid);
    FOR_CELLS(d, ch, lb)
    {
        int i, j, k;
        decode(d, id, i, j, k);
        y[id] = apply7(A, d, id, i, j, k, [&](int n) { return x[n]; });
    }
}

template <typename TO, typename TI>
__global__ __launch_bounds__(kBlock) void k_convert(int n, const TI* in, TO* out)
{
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        out[i] = (TO)in[i];
}

__global__ __launch_bounds__(kBlock) void k_dot(Dims d, Chunking ch, const double* a,
                                                const double* b, RedScratch red, double* out)
{
    const int lb = logical_block(ch.grid);
    double acc[1] = {0.0};
    FOR_CELLS(d, ch, lb) acc[0] += a[id] * b[id];
    block_sum<1>(acc);
    double tot[1];
    if (grid_reduce<1>(acc, red.partials, red.counter, ch.grid, lb, tot) && threadIdx.x == 0)
        out[0] = tot[0];
}

// ------------------------------------------------------------------
// PCG (pressure.cpp:342-414)
// ------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_init(Dims d, Chunking ch, Stencil<T> A, const T* b,
                                                     const T* x, T* r, RedScratch red,
                                                     SolveScalars* sc)
{
    const int lb = logical_block(ch.grid);
    double acc[2] = {0.0, 0.0};
    FOR_CELLS(d, ch, lb)
    {
        int i, j, k;
        decode(d, id, i, j, k);
        const T q = apply7(A, d, id, i, j, k, [&](int n) { return x[n]; });
        const T bv = b[id];
        const T rv = bv - q;
        r[id] = rv;
        acc[0] += (double)rv * (double)rv;
        acc[1] += (double)bv * (double)bv;
    }
    block_sum<2>(acc);
    double tot[2];
    if (grid_reduce<2>(acc, red.partials, red.counter, ch.grid, lb, tot) && threadIdx.x == 0)
        epilogue(EPI_INIT, tot, sc, nullptr);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_matvec(Dims d, Chunking ch, Stencil<T> A,
                                                       const T* p, T* q, RedScratch red,
                                                       SolveScalars* sc)
{
    if (sc->done) return;
    const int lb = logical_block(ch.grid);
    double acc[1] = {0.0};
    FOR_CELLS(d, ch, lb)
    {
        int i, j, k;
        decode(d, id, i, j, k);
        const T qv = apply7(A, d, id, i, j, k, [&](int n) { return p[n]; });
        q[id] = qv;
        acc[0] += (double)p[id] * (double)qv;
    }
    block_sum<1>(acc);
    double tot[1];
    if (grid_reduce<1>(acc, red.partials, red.counter, ch.grid, lb, tot) && threadIdx.x == 0)
        epilogue(EPI_ALPHA, tot, sc, nullptr);
}

template <typename T, typename TX>
__global__ __launch_bounds__(kBlock) void k_pcg_update(Dims d, Chunking ch, TX* x, T* r,
                                                       const T* p, const T* q, const T* diag,
                                                       RedScratch red, SolveScalars* sc)
{
    if (sc->done) return;
    const int lb = logical_block(ch.grid);
    const double alpha = sc->alpha;
    const TX ax = (TX)alpha;
    const T ar = (T)(-alpha);
    double acc[2] = {0.0, 0.0};
    FOR_CELLS(d, ch, lb)
    {
        x[id] = x[id] + ax * (TX)p[id];
        const T rv = r[id] + ar * q[id];
        r[id] = rv;
        acc[0] += (double)rv * (double)rv;
        if (diag) acc[1] += (double)rv * (double)(rv / diag[id]);
    }
    double tot[2];
    if (diag) {
        block_sum<2>(acc);
        if (grid_reduce<2>(acc, red.partials, red.counter, ch.grid, lb, tot) && threadIdx.x == 0)
            epilogue(EPI_CONV_RZ, tot, sc, nullptr);
    } else {
        double a1[1] = {acc[0]}, t1[1];
        block_sum<1>(a1);
        if (grid_reduce<1>(a1, red.partials, red.counter, ch.grid, lb, t1) && threadIdx.x == 0)
            epilogue(EPI_CONV, t1, sc, nullptr);
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_p_jacobi(Dims d, Chunking ch, const T* r,
                                                         const T* diag, T* p, int init, int epi,
                                                         RedScratch red, SolveScalars* sc)
{
    if (is_done(sc)) return;
    const int lb = logical_block(ch.grid);
    if (init) {
        // z = r / A_c into p (pressure.cpp:420-428), r.z with the epilogue
        double acc[1] = {0.0};
        FOR_CELLS(d, ch, lb)
        {
            const T rv = r[id];
            const T z = rv / diag[id];
            p[id] = z;
            acc[0] += (double)rv * (double)z;
        }
        if (epi >= 0) {
            block_sum<1>(acc);
            double tot[1];
            if (grid_reduce<1>(acc, red.partials, red.counter, ch.grid, lb, tot) &&
                threadIdx.x == 0)
                epilogue(epi, tot, sc, nullptr);
        }
    } else {
        const T beta = (T)sc->beta;
        FOR_CELLS(d, ch, lb) p[id] = r[id] / diag[id] + beta * p[id];
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_p_update(Dims d, Chunking ch, const T* z, T* p,
                                                         SolveScalars* sc)
{
    if (sc->done) return;
    const int lb = logical_block(ch.grid);
    const T beta = (T)sc->beta;
    FOR_CELLS(d, ch, lb) p[id] = z[id] + beta * p[id];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_rz(Dims d, Chunking ch, const T* r, const T* z,
                                                   int epi, RedScratch red, SolveScalars* sc)
{
    if (sc->done) return;
    const int lb = logical_block(ch.grid);
    double acc[1] = {0.0};
    FOR_CELLS(d, ch, lb) acc[0] += (double)r[id] * (double)z[id];
    block_sum<1>(acc);
    double tot[1];
    if (grid_reduce<1>(acc, red.partials, red.counter, ch.grid, lb, tot) && threadIdx.x == 0)
        epilogue(epi, tot, sc, nullptr);
}

// ------------------------------------------------------------------
// Chebyshev (pressure.cpp:851-906)
// ------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cheb_step(Dims d, Chunking ch, Stencil<T> A,
                                                      const T* r, const T* z_in, T* z_out, T* p,
                                                      double alpha0, double beta, double alpha,
                                                      int first, int epi, RedScratch red,
                                                      SolveScalars* sc)
{
    if (is_done(sc)) return;
    const int lb = logical_block(ch.grid);
    const T a0 = (T)alpha0, bt = (T)beta, at = (T)alpha;
    double acc[1] = {0.0};
    FOR_CELLS(d, ch, lb)
    {
        int i, j, k;
        decode(d, id, i, j, k);
        T az, zc, pp;
        const T rv = r[id];
        if (first) {
            az = apply7(A, d, id, i, j, k, [&](int n) { return T(0) + a0 * r[n]; });
            zc = T(0) + a0 * rv;
            pp = rv;
        } else {
            az = apply7(A, d, id, i, j, k, [&](int n) { return z_in[n]; });
            zc = z_in[id];
            pp = p[id];
        }
        const T res = rv - az;
        const T pn = res + bt * pp;
        p[id] = pn;
        const T zn = zc + at * pn;
        z_out[id] = zn;
        if (epi >= 0) acc[0] += (double)rv * (double)zn;
    }
    if (epi >= 0) {
        block_sum<1>(acc);
        double tot[1];
        if (grid_reduce<1>(acc, red.partials, red.counter, ch.grid, lb, tot) && threadIdx.x == 0)
            epilogue(epi, tot, sc, nullptr);
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_cheb_scale(Dims d, Chunking ch, const T* r, T* z,
                                                       double alpha0, int epi, RedScratch red,
                                                       SolveScalars* sc)
{
    if (is_done(sc)) return;
    const int lb = logical_block(ch.grid);
    const T a0 = (T)alpha0;
    double acc[1] = {0.0};
    FOR_CELLS(d, ch, lb)
    {
        const T rv = r[id];
        const T zn = T(0) + a0 * rv;
        z[id] = zn;
        if (epi >= 0) acc[0] += (double)rv * (double)zn;
    }
    if (epi >= 0) {
        block_sum<1>(acc);
        double tot[1];
        if (grid_reduce<1>(acc, red.partials, red.counter, ch.grid, lb, tot) && threadIdx.x == 0)
            epilogue(epi, tot, sc, nullptr);
    }
}

// ------------------------------------------------------------------
// mixed precision: FP32 iteration with FP64 reliable residual replacement
// ------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_mixed_fold(Dims d, Chunking ch, double* x64,
                                                       float* x32, int force, SolveScalars* sc)
{
    if (!force && (sc->done || !sc->reliable_now)) return;
    const int lb = logical_block(ch.grid);
    FOR_CELLS(d, ch, lb)
    {
        x64[id] = x64[id] + (double)x32[id];
        x32[id] = 0.0f;
    }
}

__global__ __launch_bounds__(kBlock) void k_mixed_residual(Dims d, Chunking ch, Stencil<double> A,
                                                           const double* b, const double* x,
                                                           float* r32, int init, int force,
                                                           RedScratch red, SolveScalars* sc,
                                                           double* out)
{
    if (!force && !init && (sc->done || !sc->reliable_now)) return;
    const int lb = logical_block(ch.grid);
    double acc[2] = {0.0, 0.0};
    FOR_CELLS(d, ch, lb)
    {
        int i, j, k;
        decode(d, id, i, j, k);
        const double q = apply7(A, d, id, i, j, k, [&](int n) { return x[n]; });
        const double bv = b[id];
        const double rv = bv - q;
        r32[id] = (float)rv;
        acc[0] += rv * rv;
        acc[1] += bv * bv;
    }
    block_sum<2>(acc);
    double tot[2];
    if (grid_reduce<2>(acc, red.partials, red.counter, ch.grid, lb, tot) && threadIdx.x == 0) {
        if (init) epilogue(EPI_INIT_MIXED, tot, sc, nullptr);
        else if (force) out[0] = sqrt(tot[0]);
        else epilogue(EPI_RELIABLE, tot, sc, nullptr);
    }
}

__global__ __launch_bounds__(kBlock) void k_pcg_update_mixed(Dims d, Chunking ch, float* x,
                                                             float* r, const float* p,
                                                             const float* q, RedScratch red,
                                                             SolveScalars* sc)
{
    if (sc->done) return;
    const int lb = logical_block(ch.grid);
    const float ax = (float)sc->alpha;
    const float ar = (float)(-sc->alpha);
    double acc[1] = {0.0};
    FOR_CELLS(d, ch, lb)
    {
        x[id] = x[id] + ax * p[id];
        const float rv = r[id] + ar * q[id];
        r[id] = rv;
        acc[0] += (double)rv * (double)rv;
    }
    block_sum<1>(acc);
    double tot[1];
    if (grid_reduce<1>(acc, red.partials, red.counter, ch.grid, lb, tot) && threadIdx.x == 0)
        epilogue(EPI_CONV_MIXED, tot, sc, nullptr);
}

// ------------------------------------------------------------------
// IC(0), one i+j+k hyperplane per launch (pressure.cpp:442-556).  Cells of
// one hyperplane depend only on the previous one, so this order computes
// exactly what the reference's lexicographic loops compute.
// ------------------------------------------------------------------
__device__ __forceinline__ bool level_cell(const Dims& d, int level, int t, int& i, int& j, int& k)
{
    if (t >= d.nx * d.ny) return false;
    i = t / d.ny;
    j = t - i * d.ny;
    k = level - i - j;
    return k >= 0 && k < d.nz;
}

__global__ __launch_bounds__(kBlock) void k_ic_factor_level(Dims d, int level, Stencil<double> A,
                                                            double* Lc, double* Lw, double* Ls,
                                                            double* Lb)
{
    int i, j, k;
    if (!level_cell(d, level, blockIdx.x * kBlock + threadIdx.x, i, j, k)) return;
    const int id = i * d.sij + j * d.nz + k;
    const double small_val = 1e-30;
    double sum_sq = 0.0;
    if (i > 0) {
        const double lc = Lc[id - d.sij];
        if (fabs(lc) > small_val) {
            const double l = A.a[SW][id] / lc;
            Lw[id] = l;
            sum_sq += l * l;
        }
    }
    if (j > 0) {
        const double lc = Lc[id - d.nz];
        if (fabs(lc) > small_val) {
            const double l = A.a[SS][id] / lc;
            Ls[id] = l;
            sum_sq += l * l;
        }
    }
    if (k > 0) {
        const double lc = Lc[id - 1];
        if (fabs(lc) > small_val) {
            const double l = A.a[SB][id] / lc;
            Lb[id] = l;
            sum_sq += l * l;
        }
    }
    const double diag = A.a[SC][id] - sum_sq;
    Lc[id] = (diag <= small_val) ? 1.0 : sqrt(diag);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_ic_fwd_level(Dims d, int level, const T* r, T* z,
                                                         const T* Lc, const T* Lw, const T* Ls,
                                                         const T* Lb, const SolveScalars* sc)
{
    if (is_done(sc)) return;
    int i, j, k;
    if (!level_cell(d, level, blockIdx.x * kBlock + threadIdx.x, i, j, k)) return;
    const int id = i * d.sij + j * d.nz + k;
    T sum = r[id];
    if (i > 0) sum -= Lw[id] * z[id - d.sij];
    if (j > 0) sum -= Ls[id] * z[id - d.nz];
    if (k > 0) sum -= Lb[id] * z[id - 1];
    z[id] = sum / Lc[id];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_ic_bwd_level(Dims d, int level, T* z, const T* Lc,
                                                         const T* Lw, const T* Ls, const T* Lb,
                                                         const SolveScalars* sc)
{
    if (is_done(sc)) return;
    int i, j, k;
    if (!level_cell(d, level, blockIdx.x * kBlock + threadIdx.x, i, j, k)) return;
    const int id = i * d.sij + j * d.nz + k;
    T sum = z[id];
    if (i < d.nx - 1) sum -= Lw[id + d.sij] * z[id + d.sij];
    if (j < d.ny - 1) sum -= Ls[id + d.nz] * z[id + d.nz];
    if (k < d.nz - 1) sum -= Lb[id + 1] * z[id + 1];
    z[id] = sum / Lc[id];
}

// y = S x with S the symmetrisation of A: S_ij = -sqrt(A_ij A_ji) where
// both couplings are non-zero, else 0, and S_ii = A_ii.  When W A is
// symmetric for a positive diagonal W (the mirror-Neumann rows of
// build_stencil, pressure.cpp:662-696, give W = 1/2 per mirrored
// direction), S = W^(1/2) A W^(-1/2) on the fluid block, so S and A share
// their eigenvalues and Lanczos on S bounds the spectrum of A.
__global__ __launch_bounds__(kBlock) void k_matvec_sym(Dims d, Chunking ch, Stencil<double> A,
                                                       const double* x, double* y)
{
    const int lb = logical_block(ch.grid);
    FOR_CELLS(d, ch, lb)
    {
        int i, j, k;
        decode(d, id, i, j, k);
        double val = A.a[SC][id] * x[id];
        auto add = [&](const double* out, const double* back, int nb) {
            const double pr = out[id] * back[nb];
            if (pr > 0.0) val -= sqrt(pr) * x[nb];
        };
        if (i > 0) add(A.a[SW], A.a[SE], id - d.sij);
        if (i < d.nx - 1) add(A.a[SE], A.a[SW], id + d.sij);
        if (j > 0) add(A.a[SS], A.a[SN], id - d.nz);
        if (j < d.ny - 1) add(A.a[SN], A.a[SS], id + d.nz);
        if (k > 0) add(A.a[SB], A.a[ST], id - 1);
        if (k < d.nz - 1) add(A.a[ST], A.a[SB], id + 1);
        y[id] = val;
    }
}

// w = w - a v - b vp  (Lanczos three-term step)
__global__ __launch_bounds__(kBlock) void k_lanczos_axpy(int n, double* w, const double* v,
                                                         const double* vp, double a, double b)
{
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        w[i] = w[i] - a * v[i] - b * vp[i];
}

__global__ __launch_bounds__(kBlock) void k_scale_by_invnorm(int n, const double* w, double* v,
                                                             const double* nrm2)
{
    const double nrm = sqrt(nrm2[0]);
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        v[i] = w[i] / nrm;
}

int level_grid(const Dims& d) { return (d.nx * d.ny + kBlock - 1) / kBlock; }

}  // namespace

// ------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------
hipError_t launch_build_stencil(hipStream_t s, const Dims& d, const uint8_t* solid,
                                double* const A[7], double idx2, double idy2, double idz2)
{
    const Chunking ch = make_chunking(d.N);
    hipLaunchKernelGGL(k_build_stencil, dim3(ch.grid), dim3(kBlock), 0, s, d, ch, solid, A[SC],
                       A[SN], A[SS], A[SE], A[SW], A[ST], A[SB], idx2, idy2, idz2);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_matvec(hipStream_t s, const Dims& d, Stencil<T> A, const T* x, T* y)
{
    const Chunking ch = make_chunking(d.N);
    hipLaunchKernelGGL(k_matvec<T>, dim3(ch.grid), dim3(kBlock), 0, s, d, ch, A, x, y);
    return hipGetLastError();
}

template <typename TO, typename TI>
hipError_t launch_convert(hipStream_t s, int n, const TI* in, TO* out)
{
    int g = (n + kBlock - 1) / kBlock;
    if (g > 8192) g = 8192;
    if (g < 1) g = 1;
    hipLaunchKernelGGL((k_convert<TO, TI>), dim3(g), dim3(kBlock), 0, s, n, in, out);
    return hipGetLastError();
}

hipError_t launch_dot(hipStream_t s, const Dims& d, const double* a, const double* b,
                      RedScratch red, double* out)
{
    const Chunking ch = make_chunking(d.N);
    hipLaunchKernelGGL(k_dot, dim3(ch.grid), dim3(kBlock), 0, s, d, ch, a, b, red, out);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_pcg_init(hipStream_t s, const Dims& d, Stencil<T> A, const T* b, const T* x,
                           T* r, RedScratch red, SolveScalars* sc)
{
    const Chunking ch = make_chunking(d.N);
    hipLaunchKernelGGL(k_pcg_init<T>, dim3(ch.grid), dim3(kBlock), 0, s, d, ch, A, b, x, r, red,
                       sc);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_pcg_matvec(hipStream_t s, const Dims& d, Stencil<T> A, const T* p, T* q,
                             RedScratch red, SolveScalars* sc)
{
    const Chunking ch = make_chunking(d.N);
    hipLaunchKernelGGL(k_pcg_matvec<T>, dim3(ch.grid), dim3(kBlock), 0, s, d, ch, A, p, q, red,
                       sc);
    return hipGetLastError();
}

template <typename T, typename TX>
hipError_t launch_pcg_update(hipStream_t s, const Dims& d, TX* x, T* r, const T* p, const T* q,
                             const T* jacobi_diag, RedScratch red, SolveScalars* sc)
{
    const Chunking ch = make_chunking(d.N);
    hipLaunchKernelGGL((k_pcg_update<T, TX>), dim3(ch.grid), dim3(kBlock), 0, s, d, ch, x, r, p,
                       q, jacobi_diag, red, sc);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_pcg_p_jacobi(hipStream_t s, const Dims& d, const T* r, const T* diag, T* p,
                               int init, int epi, RedScratch red, SolveScalars* sc)
{
    const Chunking ch = make_chunking(d.N);
    hipLaunchKernelGGL(k_pcg_p_jacobi<T>, dim3(ch.grid), dim3(kBlock), 0, s, d, ch, r, diag, p,
                       init, epi, red, sc);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_pcg_p_update(hipStream_t s, const Dims& d, const T* z, T* p, SolveScalars* sc)
{
    const Chunking ch = make_chunking(d.N);
    hipLaunchKernelGGL(k_pcg_p_update<T>, dim3(ch.grid), dim3(kBlock), 0, s, d, ch, z, p, sc);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_pcg_rz(hipStream_t s, const Dims& d, const T* r, const T* z, int epilogue,
                         RedScratch red, SolveScalars* sc)
{
    const Chunking ch = make_chunking(d.N);
    hipLaunchKernelGGL(k_pcg_rz<T>, dim3(ch.grid), dim3(kBlock), 0, s, d, ch, r, z, epilogue, red,
                       sc);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_cheb_step(hipStream_t s, const Dims& d, Stencil<T> A, const T* r,
                            const T* z_in, T* z_out, T* p, double alpha0, double beta,
                            double alpha, int first, int epi, RedScratch red, SolveScalars* sc)
{
    const Chunking ch = make_chunking(d.N);
    hipLaunchKernelGGL(k_cheb_step<T>, dim3(ch.grid), dim3(kBlock), 0, s, d, ch, A, r, z_in,
                       z_out, p, alpha0, beta, alpha, first, epi, red, sc);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_cheb_scale(hipStream_t s, const Dims& d, const T* r, T* z, double alpha0,
                             int epi, RedScratch red, SolveScalars* sc)
{
    const Chunking ch = make_chunking(d.N);
    hipLaunchKernelGGL(k_cheb_scale<T>, dim3(ch.grid), dim3(kBlock), 0, s, d, ch, r, z, alpha0,
                       epi, red, sc);
    return hipGetLastError();
}

hipError_t launch_ic_factor_level(hipStream_t s, const Dims& d, int level, Stencil<double> A,
                                  double* Lc, double* Lw, double* Ls, double* Lb)
{
    hipLaunchKernelGGL(k_ic_factor_level, dim3(level_grid(d)), dim3(kBlock), 0, s, d, level, A,
                       Lc, Lw, Ls, Lb);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_ic_fwd_level(hipStream_t s, const Dims& d, int level, const T* r, T* z,
                               const T* Lc, const T* Lw, const T* Ls, const T* Lb,
                               const SolveScalars* sc)
{
    hipLaunchKernelGGL(k_ic_fwd_level<T>, dim3(level_grid(d)), dim3(kBlock), 0, s, d, level, r, z,
                       Lc, Lw, Ls, Lb, sc);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_ic_bwd_level(hipStream_t s, const Dims& d, int level, T* z, const T* Lc,
                               const T* Lw, const T* Ls, const T* Lb, const SolveScalars* sc)
{
    hipLaunchKernelGGL(k_ic_bwd_level<T>, dim3(level_grid(d)), dim3(kBlock), 0, s, d, level, z,
                       Lc, Lw, Ls, Lb, sc);
    return hipGetLastError();
}

hipError_t launch_mixed_fold(hipStream_t s, const Dims& d, double* x64, float* x32, int force,
                             SolveScalars* sc)
{
    const Chunking ch = make_chunking(d.N);
    hipLaunchKernelGGL(k_mixed_fold, dim3(ch.grid), dim3(kBlock), 0, s, d, ch, x64, x32, force, sc);
    return hipGetLastError();
}

hipError_t launch_mixed_residual(hipStream_t s, const Dims& d, Stencil<double> A, const double* b,
                                 const double* x64, float* r32, int init, int force,
                                 RedScratch red, SolveScalars* sc, double* out)
{
    const Chunking ch = make_chunking(d.N);
    hipLaunchKernelGGL(k_mixed_residual, dim3(ch.grid), dim3(kBlock), 0, s, d, ch, A, b, x64, r32,
                       init, force, red, sc, out);
    return hipGetLastError();
}

hipError_t launch_pcg_update_mixed(hipStream_t s, const Dims& d, float* x, float* r, const float* p,
                                   const float* q, RedScratch red, SolveScalars* sc)
{
    const Chunking ch = make_chunking(d.N);
    hipLaunchKernelGGL(k_pcg_update_mixed, dim3(ch.grid), dim3(kBlock), 0, s, d, ch, x, r, p, q, red,
                       sc);
    return hipGetLastError();
}

hipError_t launch_matvec_sym(hipStream_t s, const Dims& d, Stencil<double> A, const double* x,
                             double* y)
{
    const Chunking ch = make_chunking(d.N);
    hipLaunchKernelGGL(k_matvec_sym, dim3(ch.grid), dim3(kBlock), 0, s, d, ch, A, x, y);
    return hipGetLastError();
}

hipError_t launch_lanczos_axpy(hipStream_t s, int n, double* w, const double* v, const double* vp,
                               double a, double b)
{
    int g = (n + kBlock - 1) / kBlock;
    if (g > 8192) g = 8192;
    if (g < 1) g = 1;
    hipLaunchKernelGGL(k_lanczos_axpy, dim3(g), dim3(kBlock), 0, s, n, w, v, vp, a, b);
    return hipGetLastError();
}

hipError_t launch_scale_by_invnorm(hipStream_t s, int n, const double* w, double* v,
                                   const double* nrm2)
{
    int g = (n + kBlock - 1) / kBlock;
    if (g > 8192) g = 8192;
    if (g < 1) g = 1;
    hipLaunchKernelGGL(k_scale_by_invnorm, dim3(g), dim3(kBlock), 0, s, n, w, v, nrm2);
    return hipGetLastError();
}

// explicit instantiations
#define LES_INST(T)                                                                              \
    template hipError_t launch_matvec<T>(hipStream_t, const Dims&, Stencil<T>, const T*, T*);   \
    template hipError_t launch_pcg_init<T>(hipStream_t, const Dims&, Stencil<T>, const T*,      \
                                           const T*, T*, RedScratch, SolveScalars*);            \
    template hipError_t launch_pcg_matvec<T>(hipStream_t, const Dims&, Stencil<T>, const T*,    \
                                             T*, RedScratch, SolveScalars*);                    \
    template hipError_t launch_pcg_p_jacobi<T>(hipStream_t, const Dims&, const T*, const T*,    \
                                               T*, int, int, RedScratch, SolveScalars*);        \
    template hipError_t launch_pcg_p_update<T>(hipStream_t, const Dims&, const T*, T*,          \
                                               SolveScalars*);                                  \
    template hipError_t launch_pcg_rz<T>(hipStream_t, const Dims&, const T*, const T*, int,     \
                                         RedScratch, SolveScalars*);                            \
    template hipError_t launch_cheb_step<T>(hipStream_t, const Dims&, Stencil<T>, const T*,     \
                                            const T*, T*, T*, double, double, double, int, int, \
                                            RedScratch, SolveScalars*);                         \
    template hipError_t launch_cheb_scale<T>(hipStream_t, const Dims&, const T*, T*, double,    \
                                             int, RedScratch, SolveScalars*);                   \
    template hipError_t launch_ic_fwd_level<T>(hipStream_t, const Dims&, int, const T*, T*,     \
                                               const T*, const T*, const T*, const T*,          \
                                               const SolveScalars*);                            \
    template hipError_t launch_ic_bwd_level<T>(hipStream_t, const Dims&, int, T*, const T*,     \
                                               const T*, const T*, const T*,                    \
                                               const SolveScalars*);

LES_INST(double)
LES_INST(float)
#undef LES_INST

template hipError_t launch_pcg_update<double, double>(hipStream_t, const Dims&, double*, double*,
                                                      const double*, const double*,
                                                      const double*, RedScratch, SolveScalars*);
template hipError_t launch_pcg_update<float, float>(hipStream_t, const Dims&, float*, float*,
                                                    const float*, const float*, const float*,
                                                    RedScratch, SolveScalars*);
template hipError_t launch_convert<float, double>(hipStream_t, int, const double*, float*);
template hipError_t launch_convert<double, double>(hipStream_t, int, const double*, double*);
template hipError_t launch_convert<double, float>(hipStream_t, int, const float*, double*);

}  // namespace les
