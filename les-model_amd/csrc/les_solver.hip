// les_solver.hip -- pressure-operator and Krylov kernels for gfx950.
//
// Operand order inside every per-cell expression follows the reference
// (src/pressure.cpp) so that, built with -ffp-contract=off, FP64 pointwise
// results are bit-identical to it; only the dot-product summation order
// differs (deterministic two-stage reductions, les_device.hpp).
#include "les_kernels.hpp"

#include <type_traits>

#include <math.h>

#include <algorithm>

namespace les {

namespace {

// ------------------------------------------------------------------
// operator application
// ------------------------------------------------------------------
// A coded off-diagonal coefficient: multiplier m (0, 1 or 2) times the base
// coefficient o = (T)(-1/h^2).  The reference's doubled coefficient is
// 2 * (-1/h^2) rounded to T, which equals 2 * o exactly (scaling by two
// commutes with rounding), so m * o is exact; the fused form adds +0 so that
// m = 0 gives +0 (a plain product would give -0 for o < 0).
template <typename T>
__device__ __forceinline__ T mult(unsigned m, T o)
{
    if constexpr (sizeof(T) == 4)
        return __builtin_fmaf((float)m, o, 0.0f);
    else
        return __builtin_fma((double)m, o, 0.0);
}

// The value of a neighbour outside the grid.  With the coded operator the
// coefficient of a missing neighbour is +0, so its product is -0, and adding
// -0 leaves every sum unchanged (-0 included): the coded stencil needs no
// bounds guards.  The arrays format keeps the reference's guards.
template <typename T>
__device__ __forceinline__ constexpr T outside() { return -T(0); }

// The seven coefficients of a cell.  CODED selects the storage at compile
// time (1 coded, 0 arrays; the run-time A.code must agree), -1 at run time.
template <int CODED, typename T>
__device__ __forceinline__ void coefs7(const Stencil<T>& A, int id, T& c, T& w, T& e, T& s,
                                       T& n, T& b, T& t)
{
    if (CODED == 1 || (CODED == -1 && A.code)) {
        const unsigned cd = A.code[id];
        c = (cd & CODE_IDENTITY) ? T(1) : A.cfluid;
        w = mult<T>(cd & 3u, A.ox);
        e = mult<T>((cd >> 2) & 3u, A.ox);
        s = mult<T>((cd >> 4) & 3u, A.oy);
        n = mult<T>((cd >> 6) & 3u, A.oy);
        b = mult<T>((cd >> 8) & 3u, A.oz);
        t = mult<T>((cd >> 10) & 3u, A.oz);
    } else {
        c = A.a[SC][id]; w = A.a[SW][id]; e = A.a[SE][id]; s = A.a[SS][id];
        n = A.a[SN][id]; b = A.a[SB][id]; t = A.a[ST][id];
    }
}

// q = A p at one cell, neighbour order c,w,e,s,n,b,t with the reference's
// bounds guards (pressure.cpp:726-744).  Load(n) returns p at flat index n.
template <int CODED, typename T, typename Load>
__device__ __forceinline__ T apply7(const Stencil<T>& A, const Dims& d, int id, int i, int j, int k,
                                    Load ld)
{
    T c, w, e, s, n, b, t;
    coefs7<CODED>(A, id, c, w, e, s, n, b, t);
    T val = c * ld(id);
    if (i > 0) val += w * ld(id - d.sij);
    if (i < d.nx - 1) val += e * ld(id + d.sij);
    if (j > 0) val += s * ld(id - d.nz);
    if (j < d.ny - 1) val += n * ld(id + d.nz);
    if (k > 0) val += b * ld(id - 1);
    if (k < d.nz - 1) val += t * ld(id + 1);
    return val;
}

// ---- vector forms: V consecutive cells of one row per thread ----
template <typename T, int V>
using vec_t = T __attribute__((ext_vector_type(V)));

template <typename T, int V>
__device__ __forceinline__ void vload(const T* __restrict__ p, T (&out)[V])
{
    if constexpr (V == 1) {
        out[0] = *p;
    } else {
        const vec_t<T, V> x = *reinterpret_cast<const vec_t<T, V>*>(p);
#pragma unroll
        for (int v = 0; v < V; ++v) out[v] = x[v];
    }
}

template <typename T, int V>
__device__ __forceinline__ void vstore(T* __restrict__ p, const T (&in)[V])
{
    if constexpr (V == 1) {
        *p = in[0];
    } else {
        vec_t<T, V> x;
#pragma unroll
        for (int v = 0; v < V; ++v) x[v] = in[v];
        *reinterpret_cast<vec_t<T, V>*>(p) = x;
    }
}

// Streaming (non-temporal) forms for data that is read or written once per
// launch; LES_NT selects them (0: ordinary loads and stores).  Measured at
// 256^3: 3% off the fused Chebyshev pair.
#ifndef LES_NT
#define LES_NT 1
#endif
// Buffer access: a resource over an array from a given element on (the
// range check is left wide open), loads and stores with a byte offset in a
// vector register (a thread's column, shared by every array) and one in a
// scalar register (a plane).  NT: non-temporal, as vload_s / vstore_s.
using Rsrc = __amdgpu_buffer_rsrc_t;
constexpr int kRsrcWord3 = 0x00020000;  // raw 32-bit buffer, CDNA

template <typename U>
__device__ __forceinline__ Rsrc buf_rsrc(const U* base)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<U*>(base), 0, -1, kRsrcWord3);
}

template <typename U, int V, bool NT>
__device__ __forceinline__ void bload(Rsrc rs, unsigned voff, unsigned soff, U (&o)[V])
{
    constexpr int B = V * (int)sizeof(U);
    constexpr int aux = NT ? 2 : 0;
    static_assert(B == 4 || B == 8 || B == 16, "buffer access of 4, 8 or 16 bytes");
    if constexpr (B == 16) {
        const auto x = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)voff, (int)soff, aux);
        __builtin_memcpy(o, &x, B);
    } else if constexpr (B == 8) {
        const auto x = __builtin_amdgcn_raw_buffer_load_b64(rs, (int)voff, (int)soff, aux);
        __builtin_memcpy(o, &x, B);
    } else {
        const auto x = __builtin_amdgcn_raw_buffer_load_b32(rs, (int)voff, (int)soff, aux);
        __builtin_memcpy(o, &x, B);
    }
}

template <typename U, int V, bool NT>
__device__ __forceinline__ void bstore(Rsrc rs, unsigned voff, unsigned soff, const U (&in)[V])
{
    constexpr int B = V * (int)sizeof(U);
    constexpr int aux = NT ? 2 : 0;
    static_assert(B == 4 || B == 8 || B == 16, "buffer access of 4, 8 or 16 bytes");
    if constexpr (B == 16) {
        int __attribute__((ext_vector_type(4))) x;
        __builtin_memcpy(&x, in, B);
        __builtin_amdgcn_raw_buffer_store_b128(x, rs, (int)voff, (int)soff, aux);
    } else if constexpr (B == 8) {
        int __attribute__((ext_vector_type(2))) x;
        __builtin_memcpy(&x, in, B);
        __builtin_amdgcn_raw_buffer_store_b64(x, rs, (int)voff, (int)soff, aux);
    } else {
        int x;
        __builtin_memcpy(&x, in, B);
        __builtin_amdgcn_raw_buffer_store_b32(x, rs, (int)voff, (int)soff, aux);
    }
}

template <typename T, int V>
__device__ __forceinline__ void vload_s(const T* __restrict__ p, T (&out)[V])
{
    if constexpr (LES_NT && V > 1) {
        const vec_t<T, V> x = __builtin_nontemporal_load(reinterpret_cast<const vec_t<T, V>*>(p));
#pragma unroll
        for (int v = 0; v < V; ++v) out[v] = x[v];
    } else {
        vload<T, V>(p, out);
    }
}

template <typename T, int V>
__device__ __forceinline__ void vstore_s(T* __restrict__ p, const T (&in)[V])
{
    if constexpr (LES_NT && V > 1) {
        vec_t<T, V> x;
#pragma unroll
        for (int v = 0; v < V; ++v) x[v] = in[v];
        __builtin_nontemporal_store(x, reinterpret_cast<vec_t<T, V>*>(p));
    } else {
        vstore<T, V>(p, in);
    }
}

// Cells per thread of the streaming kernels: 16-byte vectors when the row
// length allows them, else 1.
template <typename T>
constexpr int vec_width() { return 16 / (int)sizeof(T); }

template <typename T>
inline int pick_width(const Dims& d) { return d.nz % vec_width<T>() == 0 ? vec_width<T>() : 1; }

// The seven-point neighbourhood of cells id .. id+V-1 of one row, each
// loaded value passed through f (the first Chebyshev step scales r).
template <typename T, int V>
struct Nbr {
    T c[V], w[V], e[V], s[V], n[V];
    T km, kp;  // k-1 of the first cell, k+1 of the last
};

// x of lane i-1 (UP) or lane i+1 of the wave, by a DPP move (wave_shr:1 /
// wave_shl:1: one VALU operation, no LDS crossbar); the first / last lane
// gets 0.
template <bool UP, typename T>
__device__ __forceinline__ T wave_shift1(T x)
{
    constexpr int ctrl = UP ? 0x138 : 0x130;
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit values");
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(
            T, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), ctrl, 0xf, 0xf, false));
    } else {
        const unsigned long long u = __builtin_bit_cast(unsigned long long, x);
        const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)u, ctrl, 0xf, 0xf, false);
        const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), ctrl, 0xf, 0xf,
                                                   false);
        return __builtin_bit_cast(T, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
    }
}

// The k-1 / k+V neighbours of a thread's cells are the last / first cell of
// the adjacent lanes in the same row.  LES_KSHFL takes them from those lanes
// (wave_shift1): when a row lies inside one wave (nz / V a power of two up to
// 64) no memory access is left; when rows cross wave edges, SPAN kernels
// load at the edge lanes only and the others keep the two scalar loads
// (a divergent load between the unrolled passes of the fused Chebyshev pair
// costs more than it saves).  Every lane of a row must execute this
// together: callers skip whole rows only.  ld(n) gives the value at flat
// index n where it comes from memory.
#ifndef LES_KSHFL
#define LES_KSHFL 1
#endif
template <typename T, int V, bool SPAN, typename L>
__device__ __forceinline__ void row_edges(const Dims& d, int id, int k, const T (&c)[V], T& km,
                                          T& kp, L ld)
{
    km = outside<T>();
    kp = outside<T>();
    const int tpr = d.nz / V;
    const bool inwave = tpr <= 64 && (tpr & (tpr - 1)) == 0;
    if (LES_KSHFL && (inwave || SPAN)) {
        const int lane = (int)threadIdx.x & 63;
        const T up = wave_shift1<true>(c[V - 1]);
        const T dn = wave_shift1<false>(c[0]);
        if (inwave) {
            if (k > 0) km = up;
            if (k + V < d.nz) kp = dn;
        } else {
            if (k > 0) km = lane > 0 ? up : ld(id - 1);
            if (k + V < d.nz) kp = lane < 63 ? dn : ld(id + V);
        }
    } else {
        if (k > 0) km = ld(id - 1);
        if (k + V < d.nz) kp = ld(id + V);
    }
}

// The centre and the four i/j neighbour vectors (outside<T>() outside the
// grid).
template <typename T, int V>
__device__ __forceinline__ void load_nbr5(const T* __restrict__ x, const Dims& d, int id, int i,
                                          int j, Nbr<T, V>& nb)
{
    vload<T, V>(x + id, nb.c);
#pragma unroll
    for (int v = 0; v < V; ++v) nb.w[v] = nb.e[v] = nb.s[v] = nb.n[v] = outside<T>();
    if (i > 0) vload<T, V>(x + id - d.sij, nb.w);
    if (i < d.nx - 1) vload<T, V>(x + id + d.sij, nb.e);
    if (j > 0) vload<T, V>(x + id - d.nz, nb.s);
    if (j < d.ny - 1) vload<T, V>(x + id + d.nz, nb.n);
}

template <typename T, int V, bool SPAN = true, typename F>
__device__ __forceinline__ void load_nbr(const T* __restrict__ x, const Dims& d, int id, int i,
                                         int j, int k, F f, Nbr<T, V>& nb)
{
    load_nbr5<T, V>(x, d, id, i, j, nb);
    row_edges<T, V, SPAN>(d, id, k, nb.c, nb.km, nb.kp, [&](int n) { return x[n]; });
    // f goes over loaded values only: outside the grid the value stays
    // outside<T>() (an identity f folds the selects away)
#pragma unroll
    for (int v = 0; v < V; ++v) {
        nb.c[v] = f(nb.c[v]);
        if (i > 0) nb.w[v] = f(nb.w[v]);
        if (i < d.nx - 1) nb.e[v] = f(nb.e[v]);
        if (j > 0) nb.s[v] = f(nb.s[v]);
        if (j < d.ny - 1) nb.n[v] = f(nb.n[v]);
    }
    if (k > 0) nb.km = f(nb.km);
    if (k + V < d.nz) nb.kp = f(nb.kp);
}

// The mirror weight of a coded cell: 1/2 for each axis whose two
// coefficients are a doubled face and a missing one (k_mirror_weights' rule;
// products of 1/2, exact)
__device__ __forceinline__ double code_weight(unsigned cd)
{
    double wt = 1.0;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const unsigned lo = (cd >> (4 * ax)) & 3u, hi = (cd >> (4 * ax + 2)) & 3u;
        if ((lo == 2u && hi == 0u) || (hi == 2u && lo == 0u)) wt *= 0.5;
    }
    return wt;
}

// The coded coefficients as LDS lookup tables: (w, e) by code bits 0-3,
// (s, n) by bits 4-7, (b, t, centre) by bits 8-12.  Filled with the same
// operations as the decode in apply7v, so every coefficient is bit-identical;
// a cell then costs three LDS reads in place of about twenty VALU operations.
template <typename T>
struct alignas(4 * sizeof(T)) CodeTab {
    vec_t<T, 2> x[16], y[16];
    vec_t<T, 4> z[32];
};

// Threads 0..31 fill the table; the caller synchronises before use.
template <typename T>
__device__ __forceinline__ void code_tab_fill(CodeTab<T>& tb, const Stencil<T>& A)
{
    const unsigned t = threadIdx.x;
    if (t < 16) {
        tb.x[t] = vec_t<T, 2>{mult<T>(t & 3u, A.ox), mult<T>((t >> 2) & 3u, A.ox)};
        tb.y[t] = vec_t<T, 2>{mult<T>(t & 3u, A.oy), mult<T>((t >> 2) & 3u, A.oy)};
    }
    if (t < 32) {
        const unsigned cd = t << 8;
        tb.z[t] = vec_t<T, 4>{mult<T>(t & 3u, A.oz), mult<T>((t >> 2) & 3u, A.oz),
                              (cd & CODE_IDENTITY) ? T(1) : A.cfluid, T(0)};
    }
}

// q = A x at cells id .. id+V-1, per cell in apply7's operand order.  The
// arrays format keeps apply7's guards; the coded format sums all seven terms,
// its callers padding neighbours outside the grid with outside<T>().  The
// coefficients come from the code (decoded, or looked up in tab when TAB),
// from coef (carried arrays, StencilMatrix member order) or from A's arrays.
template <int CODED, typename T, int V, bool TAB = false>
__device__ __forceinline__ void apply7v(const Stencil<T>& A, const Dims& d, int id, int i, int j,
                                        int k, const Nbr<T, V>& nb, T (&q)[V],
                                        const uint16_t* codes = nullptr,
                                        const T (*coef)[V] = nullptr,
                                        const CodeTab<T>* tab = nullptr)
{
    T cc[V], cw[V], ce[V], cs[V], cn[V], cb[V], ct[V];
    if (CODED == 1) {
        uint16_t cd[V];
        if (codes) {
#pragma unroll
            for (int v = 0; v < V; ++v) cd[v] = codes[v];
        } else {
            vload<uint16_t, V>(A.code + id, cd);
        }
        if constexpr (TAB) {
#pragma unroll
            for (int v = 0; v < V; ++v) {
                const unsigned c = cd[v];
                const vec_t<T, 2> ex = tab->x[c & 15u], ey = tab->y[(c >> 4) & 15u];
                const vec_t<T, 4> ez = tab->z[(c >> 8) & 31u];
                cw[v] = ex[0];
                ce[v] = ex[1];
                cs[v] = ey[0];
                cn[v] = ey[1];
                cb[v] = ez[0];
                ct[v] = ez[1];
                cc[v] = ez[2];
            }
        } else {
#pragma unroll
            for (int v = 0; v < V; ++v) {
                const unsigned c = cd[v];
                cc[v] = (c & CODE_IDENTITY) ? T(1) : A.cfluid;
                cw[v] = mult<T>(c & 3u, A.ox);
                ce[v] = mult<T>((c >> 2) & 3u, A.ox);
                cs[v] = mult<T>((c >> 4) & 3u, A.oy);
                cn[v] = mult<T>((c >> 6) & 3u, A.oy);
                cb[v] = mult<T>((c >> 8) & 3u, A.oz);
                ct[v] = mult<T>((c >> 10) & 3u, A.oz);
            }
        }
    } else if (coef) {  // the seven coefficients, given in StencilMatrix member order
#pragma unroll
        for (int v = 0; v < V; ++v) {
            cc[v] = coef[SC][v];
            cn[v] = coef[SN][v];
            cs[v] = coef[SS][v];
            ce[v] = coef[SE][v];
            cw[v] = coef[SW][v];
            ct[v] = coef[ST][v];
            cb[v] = coef[SB][v];
        }
    } else {
        vload<T, V>(A.a[SC] + id, cc);
        vload<T, V>(A.a[SW] + id, cw);
        vload<T, V>(A.a[SE] + id, ce);
        vload<T, V>(A.a[SS] + id, cs);
        vload<T, V>(A.a[SN] + id, cn);
        vload<T, V>(A.a[SB] + id, cb);
        vload<T, V>(A.a[ST] + id, ct);
    }
    if (CODED == 1) {  // neighbours outside the grid hold outside<T>()
#pragma unroll
        for (int v = 0; v < V; ++v) {
            T val = cc[v] * nb.c[v];
            val += cw[v] * nb.w[v];
            val += ce[v] * nb.e[v];
            val += cs[v] * nb.s[v];
            val += cn[v] * nb.n[v];
            val += cb[v] * (v == 0 ? nb.km : nb.c[v - 1]);
            val += ct[v] * (v == V - 1 ? nb.kp : nb.c[v + 1]);
            q[v] = val;
        }
        return;
    }
#pragma unroll
    for (int v = 0; v < V; ++v) {
        T val = cc[v] * nb.c[v];
        if (i > 0) val += cw[v] * nb.w[v];
        if (i < d.nx - 1) val += ce[v] * nb.e[v];
        if (j > 0) val += cs[v] * nb.s[v];
        if (j < d.ny - 1) val += cn[v] * nb.n[v];
        if (k + v > 0) val += cb[v] * (v == 0 ? nb.km : nb.c[v - 1]);
        if (k + v < d.nz - 1) val += ct[v] * (v == V - 1 ? nb.kp : nb.c[v + 1]);
        q[v] = val;
    }
}

// ------------------------------------------------------------------
// device-resident scalar epilogues
// ------------------------------------------------------------------
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
    // a relative tolerance of a zero right-hand side is 0: an exactly zero
    // residual then counts as converged rather than dividing 0 by 0
    if (e < sc->tol || (sc->tol_rel > 0.0 && e == 0.0)) {
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
    case EPI_STORE2:
        out[0] = tot[0];
        out[1] = tot[1];
        break;
    case EPI_STORE_SQRT:
        out[0] = sqrt(tot[0]);
        break;
    case EPI_INIT:
        epi_init(tot, sc);
        break;
    case EPI_INIT_MIXED: {
        epi_init(tot, sc);
        sc->true_err = sc->err0;
        // r32 = r64 * 2^-e with ||r64|| = m 2^e, 0.5 <= m < 1: the FP32
        // iteration is linear in r, so a power-of-two scale changes no
        // mantissa bit of any in-range value, only keeps the range in FP32
        double s = 1.0;
        const double e = sc->err0;
        if (e > 0.0 && e <= 1.7976931348623157e308) {
            int ex = 0;
            frexp(e, &ex);
            s = ldexp(1.0, -ex);
        }
        sc->rscale = s;
        break;
    }
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
        if (sc->restart_pending) {
            sc->beta = 0.0;
            sc->restart_pending = 0;
            sc->restarts += 1;
        } else {
            sc->beta = tot[0] / sc->rho;
        }
        sc->rho = tot[0];
        break;
    case EPI_CONV_MIXED: {
        const int it = sc->iter + 1;
        sc->iter = it;
        sc->rr = tot[0];
        const double e = sqrt(tot[0]) / sc->rscale;  // r32 carries the scale
        sc->err = e;
        if (sc->trace) sc->trace[it] = e;
        if (!(e == e)) {
            sc->done = 1;
            sc->fault = 1;
            break;
        }
        if (e < sc->tol || (sc->reliable_every > 0 && it % sc->reliable_every == 0))
            sc->reliable_now = 1;
        // the FP32 recursion's drift grows with the residual it started
        // from: once the residual has fallen by reliable_drop since the last
        // true residual, that drift is no longer small against it
        if (e < sc->reliable_drop * sc->true_err) sc->reliable_now = 1;
        if (it >= sc->max_iter) {
            sc->reliable_now = 1;
            sc->stop_pending = 1;
        }
        break;
    }
    case EPI_RELIABLE: {
        const double e = sqrt(tot[0]);
        // A true residual well above the recursive one means the FP32
        // recursion has drifted from the FP64 iterate; the search direction
        // then restarts from the replaced residual (a new refinement
        // solve).  Keeping it, as a small drift allows, lets CG on the
        // non-symmetric operator diverge.
        if (e > 2.0 * sc->err) sc->restart_pending = 1;
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
    default:
        break;
    }
}

__device__ __forceinline__ bool is_done(const SolveScalars* sc) { return sc && sc->done; }

// beta of a p update: EPI_BETA's, formed from the preconditioner's r.z
// partials when they were left for this kernel (rz_grid > 0; every
// workgroup sums them in the finalize's fixed order and divides as it
// does), else the value its finalize stored.  Nothing here is written.
__device__ __forceinline__ double fold_beta(const SolveScalars* sc, RedScratch rz, int rz_grid)
{
    if (rz_grid <= 0) return sc->beta;
    double tot[1];
    reduce_partials<1, false>(rz.partials, rz_grid, tot);
    return sc->restart_pending ? 0.0 : tot[0] / sc->rho;
}

// Whether a kernel gated by `gate` runs.  Producer and finalize apply the
// same test, and neither changes the scalars it reads.
__device__ __forceinline__ bool gate_open(int gate, const SolveScalars* sc)
{
    switch (gate) {
    case GATE_DONE: return !sc->done;
    case GATE_RELIABLE: return !sc->done && sc->reliable_now;
    default: return true;
    }
}

// The one-workgroup tail of every reduction: fixed-order sum (or max) of the
// producer's per-workgroup partials, then the scalar epilogue.
template <int NR, bool MAX>
__global__ __launch_bounds__(kBlock) void k_finalize(int grid, const double* partials, int epi,
                                                     int gate, SolveScalars* sc, double* out)
{
    if (!gate_open(gate, sc)) return;
    double tot[NR];
    reduce_partials<NR, MAX>(partials, grid, tot);
    if (threadIdx.x == 0) epilogue(epi, tot, sc, out);
}

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

// The same case analysis as build_stencil, recording multipliers of the
// base coefficients instead of their values.
__global__ __launch_bounds__(kBlock) void k_build_code(Dims d, Chunking ch, const uint8_t* solid,
                                                       uint16_t* code)
{
    const int lb = logical_block(ch.grid);
    FOR_CELLS(d, ch, lb)
    {
        unsigned cd = CODE_IDENTITY;
        if (!solid[id] && i != d.nx - 1) {
            unsigned e = 1, w = 1, n = 1, s = 1, t = 1, b = 1;
            if (i == 0) { e += 1; w = 0; }
            else if (solid[id - d.sij]) { e += 1; w = 0; }
            if (i == d.nx - 2) { e = 0; }
            else if (solid[id + d.sij]) { w += 1; e = 0; }
            if (j == 0) { n += 1; s = 0; }
            else if (solid[id - d.nz]) { n += 1; s = 0; }
            if (j == d.ny - 1) { s += 1; n = 0; }
            else if (solid[id + d.nz]) { s += 1; n = 0; }
            if (k == 0) { t += 1; b = 0; }
            else if (solid[id - 1]) { t += 1; b = 0; }
            if (k == d.nz - 1) { b += 1; t = 0; }
            else if (solid[id + 1]) { b += 1; t = 0; }
            cd = w | (e << 2) | (s << 4) | (n << 6) | (b << 8) | (t << 10);
        }
        code[id] = (uint16_t)cd;
    }
}

// W = 1/2 per direction in which a row of the coded operator is mirrored
// (one multiplier 2, the opposite 0); W A is symmetric on grids without
// obstacles.
__global__ __launch_bounds__(kBlock) void k_mirror_weights(Dims d, Chunking ch,
                                                           const uint16_t* code, double* W)
{
    const int lb = logical_block(ch.grid);
    FOR_CELLS(d, ch, lb)
    {
        const unsigned cd = code[id];
        double wt = 1.0;
        for (int ax = 0; ax < 3; ++ax) {
            const unsigned lo = (cd >> (4 * ax)) & 3u, hi = (cd >> (4 * ax + 2)) & 3u;
            if ((lo == 2u && hi == 0u) || (hi == 2u && lo == 0u)) wt *= 0.5;
        }
        W[id] = wt;
    }
}

template <typename T, int CODED>
__global__ __launch_bounds__(kBlock) void k_matvec(Dims d, Chunking ch, Stencil<T> A,
                                                   const T* __restrict__ x, T* __restrict__ y)
{
    const int lb = logical_block(ch.grid);
    FOR_CELLS(d, ch, lb) y[id] = apply7<CODED>(A, d, id, i, j, k, [&](int n) { return x[n]; });
}

template <typename TO, typename TI>
__global__ __launch_bounds__(kBlock) void k_convert(int n, const TI* in, TO* out)
{
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        out[i] = (TO)in[i];
}

__global__ __launch_bounds__(kBlock) void k_dot(Dims d, Chunking ch, const double* a,
                                                const double* b, RedScratch red)
{
    const int lb = logical_block(ch.grid);
    double acc[1] = {0.0};
    FOR_CELLS(d, ch, lb) acc[0] += a[id] * b[id];
    block_partial<1>(acc, red.partials, lb);
}

// ------------------------------------------------------------------
// PCG (pressure.cpp:342-414)
// ------------------------------------------------------------------
template <typename T, int CODED>
__global__ __launch_bounds__(kBlock) void k_pcg_init(Dims d, Chunking ch, Stencil<T> A,
                                                     const T* __restrict__ b,
                                                     const T* __restrict__ x, T* __restrict__ r,
                                                     RedScratch red)
{
    const int lb = logical_block(ch.grid);
    double acc[2] = {0.0, 0.0};
    FOR_CELLS(d, ch, lb)
    {
        const T q = apply7<CODED>(A, d, id, i, j, k, [&](int n) { return x[n]; });
        const T bv = b[id];
        const T rv = bv - q;
        r[id] = rv;
        acc[0] += (double)rv * (double)rv;
        acc[1] += (double)bv * (double)bv;
    }
    block_partial<2>(acc, red.partials, lb);
}

// q = A p and p.q.  PU (the previous iteration's direction update fused
// in): 0 p given; 1 p = z + beta p_old; 2 p = r / diag + beta p_old
// (Jacobi), formed at every stencil point from z (or r, diag) and p_old,
// written to p at the cell itself; p_old and p are distinct buffers.
// x (PU != 0, may be null): the previous iteration's x += alpha p_old,
// deferred from its update kernel to here, where p_old is read anyway;
// alpha is still that iteration's (the next one is formed after this
// kernel), and the arithmetic is the update's.
template <typename T, int CODED, int V, int PU>
__global__ __launch_bounds__(kBlock) void k_pcg_matvec(Dims d, Chunking ch, Stencil<T> A,
                                                       const T* __restrict__ p_old,
                                                       const T* __restrict__ z,
                                                       const T* __restrict__ diag,
                                                       T* __restrict__ p,
                                                       T* __restrict__ q, RedScratch red,
                                                       const SolveScalars* sc,
                                                       const double* __restrict__ wgt,
                                                       T* __restrict__ x, RedScratch rz,
                                                       int rz_grid)
{
    if (sc->done) return;
    const int lb = logical_block(ch.grid);
    const T beta = (T)fold_beta(sc, rz, rz_grid);
    const T ax = (T)sc->alpha;
    double acc[1] = {0.0};
    FOR_VCELLS(d, ch, lb, V)
    {
        Nbr<T, V> nb;
        if (PU == 0) {
            load_nbr<T, V>(p, d, id, i, j, k, [](T x) { return x; }, nb);
        } else {
            Nbr<T, V> nz_, np_;
            load_nbr5<T, V>(z, d, id, i, j, nz_);
            load_nbr5<T, V>(p_old, d, id, i, j, np_);
            auto comb = [&](T zz, T pp, int n) {
                if (PU == 1) return zz + beta * pp;
                return zz / diag[n] + beta * pp;
            };
#pragma unroll
            for (int v = 0; v < V; ++v) {
                nb.c[v] = comb(nz_.c[v], np_.c[v], id + v);
                nb.w[v] = i > 0 ? comb(nz_.w[v], np_.w[v], id - d.sij + v) : outside<T>();
                nb.e[v] = i < d.nx - 1 ? comb(nz_.e[v], np_.e[v], id + d.sij + v) : outside<T>();
                nb.s[v] = j > 0 ? comb(nz_.s[v], np_.s[v], id - d.nz + v) : outside<T>();
                nb.n[v] = j < d.ny - 1 ? comb(nz_.n[v], np_.n[v], id + d.nz + v) : outside<T>();
            }
            row_edges<T, V, true>(d, id, k, nb.c, nb.km, nb.kp,
                                  [&](int n) { return comb(z[n], p_old[n], n); });
            vstore_s<T, V>(p + id, nb.c);
            if (x) {
                T xv[V];
                vload<T, V>(x + id, xv);
#pragma unroll
                for (int v = 0; v < V; ++v) xv[v] = xv[v] + ax * np_.c[v];
                vstore_s<T, V>(x + id, xv);
            }
        }
        T qv[V];
        apply7v<CODED, T, V>(A, d, id, i, j, k, nb, qv);
        vstore_s<T, V>(q + id, qv);
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const double pq = (double)nb.c[v] * (double)qv[v];
            acc[0] += wgt ? wgt[id + v] * pq : pq;
        }
    }
    block_partial<1>(acc, red.partials, lb);
}

template <typename T, typename TX, int V>
__global__ __launch_bounds__(kBlock) void k_pcg_update(Dims d, Chunking ch, TX* __restrict__ x,
                                                       T* __restrict__ r,
                                                       const T* __restrict__ p,
                                                       const T* __restrict__ q,
                                                       const T* __restrict__ diag,
                                                       RedScratch red, const SolveScalars* sc,
                                                       const double* __restrict__ wgt)
{
    if (sc->done) return;
    const int lb = logical_block(ch.grid);
    const double alpha = sc->alpha;
    const TX ax = (TX)alpha;
    const T ar = (T)(-alpha);
    double acc[2] = {0.0, 0.0};
    FOR_VCELLS(d, ch, lb, V)
    {
        T rv[V], qv[V];
        if (x) {  // null: x is updated by the next iteration's matvec
            TX xv[V];
            T pv[V];
            vload<TX, V>(x + id, xv);
            vload<T, V>(p + id, pv);
#pragma unroll
            for (int v = 0; v < V; ++v) xv[v] = xv[v] + ax * (TX)pv[v];
            vstore_s<TX, V>(x + id, xv);
        }
        vload<T, V>(r + id, rv);
        vload<T, V>(q + id, qv);
#pragma unroll
        for (int v = 0; v < V; ++v) rv[v] = rv[v] + ar * qv[v];
        vstore_s<T, V>(r + id, rv);
        T dv[V];
        if (diag) vload<T, V>(diag + id, dv);
#pragma unroll
        for (int v = 0; v < V; ++v) {
            acc[0] += (double)rv[v] * (double)rv[v];
            if (diag) {
                const double rz = (double)rv[v] * (double)(rv[v] / dv[v]);
                acc[1] += wgt ? wgt[id + v] * rz : rz;
            }
        }
    }
    block_partial<2>(acc, red.partials, lb);
}

template <typename T, int V>
__global__ __launch_bounds__(kBlock) void k_pcg_p_jacobi(Dims d, Chunking ch,
                                                         const T* __restrict__ r,
                                                         const T* __restrict__ diag,
                                                         T* __restrict__ p, int init,
                                                         RedScratch red, const SolveScalars* sc,
                                                         const double* __restrict__ wgt)
{
    if (is_done(sc)) return;
    const int lb = logical_block(ch.grid);
    if (init) {
        // z = r / A_c into p (pressure.cpp:420-428), r.z partials
        double acc[1] = {0.0};
        FOR_VCELLS(d, ch, lb, V)
        {
            T rv[V], dv[V], z[V];
            vload<T, V>(r + id, rv);
            vload<T, V>(diag + id, dv);
#pragma unroll
            for (int v = 0; v < V; ++v) z[v] = rv[v] / dv[v];
            vstore<T, V>(p + id, z);
#pragma unroll
            for (int v = 0; v < V; ++v) {
                const double rz = (double)rv[v] * (double)z[v];
                acc[0] += wgt ? wgt[id + v] * rz : rz;
            }
        }
        if (red.partials) block_partial<1>(acc, red.partials, lb);
    } else {
        const T beta = (T)sc->beta;
        FOR_VCELLS(d, ch, lb, V)
        {
            T rv[V], dv[V], pv[V];
            vload<T, V>(r + id, rv);
            vload<T, V>(diag + id, dv);
            vload<T, V>(p + id, pv);
#pragma unroll
            for (int v = 0; v < V; ++v) pv[v] = rv[v] / dv[v] + beta * pv[v];
            vstore<T, V>(p + id, pv);
        }
    }
}

template <typename T, int V>
__global__ __launch_bounds__(kBlock) void k_pcg_p_update(Dims d, Chunking ch,
                                                         const T* __restrict__ z,
                                                         T* __restrict__ p, const SolveScalars* sc)
{
    if (sc->done) return;
    const int lb = logical_block(ch.grid);
    const T beta = (T)sc->beta;
    FOR_VCELLS(d, ch, lb, V)
    {
        T zv[V], pv[V];
        vload<T, V>(z + id, zv);
        vload<T, V>(p + id, pv);
#pragma unroll
        for (int v = 0; v < V; ++v) pv[v] = zv[v] + beta * pv[v];
        vstore<T, V>(p + id, pv);
    }
}

template <typename T, int V>
__global__ __launch_bounds__(kBlock) void k_pcg_rz(Dims d, Chunking ch, const T* __restrict__ r,
                                                   const T* __restrict__ z, RedScratch red,
                                                   const SolveScalars* sc,
                                                   const double* __restrict__ wgt)
{
    if (sc->done) return;
    const int lb = logical_block(ch.grid);
    double acc[1] = {0.0};
    FOR_VCELLS(d, ch, lb, V)
    {
        T rv[V], zv[V];
        vload<T, V>(r + id, rv);
        vload<T, V>(z + id, zv);
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const double rz = (double)rv[v] * (double)zv[v];
            acc[0] += wgt ? wgt[id + v] * rz : rz;
        }
    }
    block_partial<1>(acc, red.partials, lb);
}

// ------------------------------------------------------------------
// Chebyshev (pressure.cpp:851-906): one inner step of the recurrence
//   res = r - A z_k ; p = res + beta p ; z_{k+1} = z_k + alpha p
// FIRST forms z_1 = 0 + alpha0 r and p_0 = r on the fly from r.
// ------------------------------------------------------------------
template <typename T, bool FIRST, int CODED, bool REDUCE, int V>
__global__ __launch_bounds__(kBlock) void k_cheb_step(Dims d, Chunking ch, Stencil<T> A,
                                                        const T* __restrict__ r,
                                                        const T* __restrict__ z_in,
                                                        T* __restrict__ z_out, T* __restrict__ p,
                                                        double alpha0, double beta, double alpha,
                                                        RedScratch red, const SolveScalars* sc,
                                                        const double* __restrict__ wgt)
{
    if (is_done(sc)) return;
    const int lb = logical_block(ch.grid);
    const T a0 = (T)alpha0, bt = (T)beta, at = (T)alpha;
    double acc[1] = {0.0};
    FOR_VCELLS(d, ch, lb, V)
    {
        Nbr<T, V> nb;
        T rv[V], az[V], pp[V], zc[V];
        vload_s<T, V>(r + id, rv);
        if (FIRST) {
            load_nbr<T, V>(r, d, id, i, j, k, [&](T x) { return T(0) + a0 * x; }, nb);
#pragma unroll
            for (int v = 0; v < V; ++v) {
                zc[v] = nb.c[v];
                pp[v] = rv[v];
            }
        } else {
            load_nbr<T, V>(z_in, d, id, i, j, k, [](T x) { return x; }, nb);
            vload_s<T, V>(p + id, pp);
#pragma unroll
            for (int v = 0; v < V; ++v) zc[v] = nb.c[v];
        }
        apply7v<CODED, T, V>(A, d, id, i, j, k, nb, az);
        T pn[V], zn[V];
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const T res = rv[v] - az[v];
            pn[v] = res + bt * pp[v];
            zn[v] = zc[v] + at * pn[v];
        }
        // the reducing step is an application's last: its p is never read
        if (!REDUCE) vstore_s<T, V>(p + id, pn);
        vstore_s<T, V>(z_out + id, zn);
        if (REDUCE) {
#pragma unroll
            for (int v = 0; v < V; ++v) {
                const double rz = (double)rv[v] * (double)zn[v];
                acc[0] += wgt ? wgt[id + v] * rz : rz;
            }
        }
    }
    if (REDUCE) block_partial<1>(acc, red.partials, lb);
}

// ------------------------------------------------------------------
// Two Chebyshev steps in one pass over HBM (steps k and k+1):
//   A: res = r - A z_k ; p_k = res + beta_k p_{k-1} ; z_{k+1} = z_k + alpha_k p_k
//   B: res = r - A z_{k+1} ; p_{k+1} = res + beta_{k+1} p_k ;
//      z_{k+2} = z_{k+1} + alpha_{k+1} p_{k+1}
// A workgroup owns J consecutive j-rows (whole rows, all k) of a range of
// i-planes and marches along i.  For plane t it runs A on its rows plus one
// halo row on each side into an R-plane LDS ring of z_{k+1}, then B on plane
// t-1 from that ring; p_k of its own rows stays in registers from one plane
// to the next.  z_{k+1} and p_k never reach HBM.  Each cell's arithmetic is
// exactly k_cheb_step's, so results are bit-identical to two single steps.
// p_in / p_out and z_in / z_out must be distinct buffers: halo rows and
// planes read inputs that neighbouring workgroups own.
// Thread layout: tpr = nz / V threads per row, rpp = TB / tpr rows per
// pass, P passes; row slot q = pass * rpp + sub holds core row q (q < J) or
// one of the two halo rows (q = J: j0 - 1, q = J + 1: j0 + J), J = P*rpp - 2.
// ------------------------------------------------------------------
struct PairGeom {
    int tpr, rpp, J;   // threads per row, rows per pass, core rows
    int njb, irange;   // j-blocks, i-planes per workgroup
    int grid;          // workgroups (multiple of kXcd)
};

// The coded coefficients of the pair (and of the ring matvec, LES_MV_TAB)
// come from LDS lookup tables (measured at 256^3 in FP32: pair 85.2 -> 83.8
// us, ring matvec 90.1 -> 88.7 us).
#ifndef LES_PAIR_CODE_TAB
#define LES_PAIR_CODE_TAB 1
#endif
#ifndef LES_MV_TAB
#define LES_MV_TAB 1
#endif
// The staged forms of the pair and of the reducing last step
// (k_cheb_pair_st) where they apply: FP32, coded operator, rows of one or
// two waves.
// Multigrid sweeps and restriction with vector loads where the rows allow
// (nz % V == 0), else one cell per thread.
#ifndef LES_MG_VEC
#define LES_MG_VEC 1
#endif
// The multigrid's two-sweep pre-smoothing pass marching along i with the
// first sweep's iterate in an LDS ring (k_mg_zs_ring) where the ring
// matvec's geometry fits; 0: k_mg_smooth_v forms it at every neighbour.
#ifndef LES_MG_ZS_RING
#define LES_MG_ZS_RING 1
#endif
#ifndef LES_MG_ZS_GRID
#define LES_MG_ZS_GRID 256
#endif
// The pre-smoothing sweeps and the restriction as one pass with LDS rings
// (k_mg_pre_ring) on FP32 levels where its geometry fits.
#ifndef LES_MG_PRE_RING
#define LES_MG_PRE_RING 1
#endif
// The restriction marching along i with x in an LDS ring
// (k_mg_restrict_ring) where its geometry fits.
#ifndef LES_MG_RESTRICT_RING
#define LES_MG_RESTRICT_RING 1
#endif
// The two post-smoothing sweeps as one pass with two LDS rings
// (k_mg_post_ring) on FP32 levels where its geometry fits.
#ifndef LES_MG_POST_RING
#define LES_MG_POST_RING 1
#endif
constexpr int kMgPostR = 4;
#ifndef LES_PAIR_STAGED
#define LES_PAIR_STAGED 1
#endif
#ifndef LES_LAST_STAGED
#define LES_LAST_STAGED 1
#endif
// Wave priority during one phase of each plane (s_setprio 2, else 0):
// 1 step A (the loads), 2 step B, 0 neither, -1 by element type (measured
// at 256^3: A for FP32, 1.4 % faster; B for FP64, 1.7 % faster).
#ifndef LES_PAIR_PRIO
#define LES_PAIR_PRIO -1
#endif
// Waves per SIMD the pair keeps resident: two workgroups per CU up to 512
// threads, one above (the register budget follows: 128 VGPRs at 4 waves).
template <int TB>
constexpr int pair_waves_per_simd() { return TB <= 512 ? TB / 128 : TB / 256; }

// TB threads per workgroup.  The pair never closes an application (a
// reducing last step runs single), so it reduces nothing.
// R: LDS ring depth in planes.  With 3 a second barrier per plane keeps step
// A of the next plane from overwriting what step B still reads; with 4 the
// plane A(t+1) overwrites was last read by B(t-2), which every wave finished
// before the barrier of plane t, so one barrier per plane suffices.
template <typename T, bool FIRST, int CODED, int V, int P, int TB, int R, bool CC>
__global__ __launch_bounds__(TB, pair_waves_per_simd<TB>()) void k_cheb_pair(Dims d, PairGeom g,
                                                                           Stencil<T> A,
                                                  const T* __restrict__ r,
                                                  const T* __restrict__ z_in,
                                                  T* __restrict__ z_out,
                                                  const T* __restrict__ p_in,
                                                  T* __restrict__ p_out, double alpha0,
                                                  double beta_a, double alpha_a, double beta_b,
                                                  double alpha_b, const SolveScalars* sc)
{
    if (is_done(sc)) return;
    extern __shared__ unsigned char les_smem[];
    T* zs = reinterpret_cast<T*>(les_smem);  // [R][P*rpp][nz]
    // the coded operator's coefficients come from LDS lookup tables
    const CodeTab<T>* tab = nullptr;
    if constexpr (CODED == 1 && LES_PAIR_CODE_TAB) {
        __shared__ CodeTab<T> tab_s;
        code_tab_fill(tab_s, A);
        __syncthreads();
        tab = &tab_s;
    }
    // CC (arrays format): the core rows' seven coefficients of the plane
    // step B works on, [7][J][nz] after the ring.  Step A keeps the ones it
    // loads in registers and stores them here once step B of the previous
    // plane has read its own; every thread reads and writes only its cells.
    T* cfs = zs + R * P * g.rpp * d.nz;
    const int lb = logical_block(g.grid);
    const int jb = lb % g.njb, ic = lb / g.njb;
    const int j0 = jb * g.J;
    const int ia = ic * g.irange, ib = min(ia + g.irange, d.nx);
    const int sub = (int)threadIdx.x / g.tpr;
    const int kk = ((int)threadIdx.x - sub * g.tpr) * V;
    const int slots = P * g.rpp;
    const int plane_sz = slots * d.nz;
    const T a0 = (T)alpha0, bta = (T)beta_a, ata = (T)alpha_a, btb = (T)beta_b, atb = (T)alpha_b;
    constexpr int prio = LES_PAIR_PRIO >= 0 ? LES_PAIR_PRIO : (sizeof(T) == 4 ? 1 : 2);
    // p_k, r and the operator code of the core rows, carried from step A
    // of a plane to step B of the same plane one iteration later
    T pkeep[P][V], pnew[P][V], rkeep[P][V], rnew[P][V];
    uint16_t ckeep[P][V], cnew[P][V];
    T cfnew[CC ? P : 1][7][V];
    bool cfval[CC ? P : 1];
#pragma unroll
    for (int pp = 0; pp < (CC ? P : 1); ++pp) cfval[pp] = false;
#pragma unroll
    for (int pp = 0; pp < P; ++pp)
#pragma unroll
        for (int v = 0; v < V; ++v) {
            pkeep[pp][v] = pnew[pp][v] = rkeep[pp][v] = rnew[pp][v] = T(0);
            ckeep[pp][v] = cnew[pp][v] = 0;
        }
    // the coded step B reads its i/j neighbours from the ring unguarded:
    // rows and planes outside the grid hold outside<T>() there
    T pad[V];
#pragma unroll
    for (int v = 0; v < V; ++v) pad[v] = outside<T>();
    for (int t = ia - 1; t <= ib && ia < d.nx; ++t) {
        if (prio == 1) __builtin_amdgcn_s_setprio(2);
        // ---- A at plane t: z_{k+1} into the ring, p_k of core rows ----
        if (t >= 0 && t < d.nx) {
#pragma unroll
            for (int pp = 0; pp < P; ++pp) {
                const int q = pp * g.rpp + sub;
                const int a = q < g.J ? q + 1 : (q == g.J ? 0 : g.J + 1);
                const int j = j0 - 1 + a;
                if (j < 0 || j >= d.ny) {
                    if (CODED == 1) vstore<T, V>(zs + (t + R) % R * plane_sz + a * d.nz + kk, pad);
                    continue;
                }
                const int i = t, k = kk, id = (i * d.ny + j) * d.nz + kk;
                Nbr<T, V> nb;
                T rv[V], az[V], pp_[V], zc[V];
                vload_s<T, V>(r + id, rv);
                if (FIRST) {
                    load_nbr<T, V, false>(r, d, id, i, j, k,
                                          [&](T x) { return T(0) + a0 * x; }, nb);
#pragma unroll
                    for (int v = 0; v < V; ++v) pp_[v] = rv[v];
                } else {
                    load_nbr<T, V, false>(z_in, d, id, i, j, k, [](T x) { return x; }, nb);
                    vload_s<T, V>(p_in + id, pp_);
                }
#pragma unroll
                for (int v = 0; v < V; ++v) zc[v] = nb.c[v];
                uint16_t cd[V];
                if (CODED == 1) vload<uint16_t, V>(A.code + id, cd);
                if (CC) {
                    T cf[7][V];
#pragma unroll
                    for (int c = 0; c < 7; ++c) vload<T, V>(A.a[c] + id, cf[c]);
                    apply7v<CODED, T, V>(A, d, id, i, j, k, nb, az, nullptr, cf);
                    if (q < g.J) {
#pragma unroll
                        for (int c = 0; c < 7; ++c)
#pragma unroll
                            for (int v = 0; v < V; ++v) cfnew[CC ? pp : 0][c][v] = cf[c][v];
                        cfval[CC ? pp : 0] = true;
                    }
                } else {
                    apply7v<CODED, T, V, CODED == 1 && LES_PAIR_CODE_TAB>(
                        A, d, id, i, j, k, nb, az, CODED == 1 ? cd : nullptr, nullptr, tab);
                }
                T zn[V];
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    const T res = rv[v] - az[v];
                    const T pn = res + bta * pp_[v];
                    zn[v] = zc[v] + ata * pn;
                    pnew[pp][v] = pn;
                    rnew[pp][v] = rv[v];
                    if (CODED == 1) cnew[pp][v] = cd[v];
                }
                vstore<T, V>(zs + (t + R) % R * plane_sz + a * d.nz + kk, zn);
            }
        } else if (CODED == 1) {
#pragma unroll
            for (int pp = 0; pp < P; ++pp) {
                const int q = pp * g.rpp + sub;
                const int a = q < g.J ? q + 1 : (q == g.J ? 0 : g.J + 1);
                vstore<T, V>(zs + (t + R) % R * plane_sz + a * d.nz + kk, pad);
            }
        }
        if (prio == 1) __builtin_amdgcn_s_setprio(0);
        __syncthreads();
        if (prio == 2) __builtin_amdgcn_s_setprio(2);
        // ---- B at plane t-1 from the ring ----
        const int iB = t - 1;
        if (iB >= ia && iB < ib) {
            const T* zc_pl = zs + (iB + R) % R * plane_sz;
            const T* zm_pl = zs + (iB + R - 1) % R * plane_sz;
            const T* zp_pl = zs + (iB + 1) % R * plane_sz;
#pragma unroll
            for (int pp = 0; pp < P; ++pp) {
                const int q = pp * g.rpp + sub;
                if (q >= g.J) continue;
                const int a = q + 1, j = j0 + q;
                if (j >= d.ny) continue;
                const int i = iB, k = kk, id = (i * d.ny + j) * d.nz + kk;
                Nbr<T, V> nb;
                vload<T, V>(zc_pl + a * d.nz + kk, nb.c);
                if (CODED == 1) {
                    vload<T, V>(zm_pl + a * d.nz + kk, nb.w);
                    vload<T, V>(zp_pl + a * d.nz + kk, nb.e);
                    vload<T, V>(zc_pl + (a - 1) * d.nz + kk, nb.s);
                    vload<T, V>(zc_pl + (a + 1) * d.nz + kk, nb.n);
                } else {
#pragma unroll
                    for (int v = 0; v < V; ++v) nb.w[v] = nb.e[v] = nb.s[v] = nb.n[v] = T(0);
                    if (i > 0) vload<T, V>(zm_pl + a * d.nz + kk, nb.w);
                    if (i < d.nx - 1) vload<T, V>(zp_pl + a * d.nz + kk, nb.e);
                    if (j > 0) vload<T, V>(zc_pl + (a - 1) * d.nz + kk, nb.s);
                    if (j < d.ny - 1) vload<T, V>(zc_pl + (a + 1) * d.nz + kk, nb.n);
                }
                row_edges<T, V, false>(d, id, k, nb.c, nb.km, nb.kp,
                                       [&](int n) { return zc_pl[a * d.nz + kk + (n - id)]; });
                T az[V];
                if (CC) {
                    T cf[7][V];
#pragma unroll
                    for (int c = 0; c < 7; ++c) vload<T, V>(cfs + (c * g.J + q) * d.nz + kk, cf[c]);
                    apply7v<CODED, T, V>(A, d, id, i, j, k, nb, az, nullptr, cf);
                } else {
                    apply7v<CODED, T, V, CODED == 1 && LES_PAIR_CODE_TAB>(
                        A, d, id, i, j, k, nb, az, CODED == 1 ? ckeep[pp] : nullptr, nullptr, tab);
                }
                T pn[V], zn[V];
                const T* rv = rkeep[pp];
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    const T res = rv[v] - az[v];
                    pn[v] = res + btb * pkeep[pp][v];
                    zn[v] = nb.c[v] + atb * pn[v];
                }
                vstore_s<T, V>(p_out + id, pn);
                vstore_s<T, V>(z_out + id, zn);
            }
        }
        if (CC) {
#pragma unroll
            for (int pp = 0; pp < (CC ? P : 1); ++pp) {
                if (!cfval[pp]) continue;
                const int q = pp * g.rpp + sub;
#pragma unroll
                for (int c = 0; c < 7; ++c) vstore<T, V>(cfs + (c * g.J + q) * d.nz + kk, cfnew[pp][c]);
                cfval[pp] = false;
            }
        }
        if (R == 3) __syncthreads();
        if (prio == 2) __builtin_amdgcn_s_setprio(0);
#pragma unroll
        for (int pp = 0; pp < P; ++pp)
#pragma unroll
            for (int v = 0; v < V; ++v) {
                pkeep[pp][v] = pnew[pp][v];
                rkeep[pp][v] = rnew[pp][v];
                ckeep[pp][v] = cnew[pp][v];
            }
    }
}

// ------------------------------------------------------------------
// The fused Chebyshev pair with its input staged (coded operator)
// ------------------------------------------------------------------
// k_cheb_pair's step A fetches z_k at the cell and its four i/j neighbours
// from global memory (mostly L2 hits) when it needs them.  Here each thread
// owns one A-row (the J core rows and the halo row on each side) and V cells,
// keeps z_k of its cells at planes t-1, t, t+1 in registers, and gets the
// j neighbours from an LDS copy of plane t (its A-rows plus the rows j0-2 and
// j0+J+1).  z_k of plane t+2, the extra rows of plane t+2, and r, p and the
// code of plane t+1 are loaded while plane t is computed, and the LDS copy of
// plane t+1 is written from registers before the plane's barrier (two
// buffers), so every global load has a plane of work to arrive and one
// barrier per plane still orders all LDS use.  One 1024-thread workgroup per
// CU.  Step B and every cell's arithmetic are k_cheb_pair's, so the results
// are bit-identical.  A row inside one wave takes its k-edges by lane
// shuffle, a row of two waves (nz = 512) from the LDS copies of z_k and
// z_{k+1}.
// SINGLE: one reducing step (an application's last; its p is never read):
// the core rows' z_{k+1} goes to z_out and r.z_{k+1} to red's partials, with
// no ring and no step B.
// NZ > 0: the kernel for grids with nz = NZ (row strides and LDS offsets
// then fold into the instructions); 0: any nz.
template <typename T, bool FIRST, int V, int R, bool SINGLE, int NZ>
__global__ __launch_bounds__(1024, 1) void k_cheb_pair_st(
    Dims dn, PairGeom g, Stencil<T> A, const T* __restrict__ r, const T* __restrict__ z_in,
    T* __restrict__ z_out, const T* __restrict__ p_in, T* __restrict__ p_out, double alpha0,
    double beta_a, double alpha_a, double beta_b, double alpha_b, const SolveScalars* sc,
    RedScratch red)
{
    static_assert(SINGLE || R == 4, "the ring slot of a plane is its place in a four-plane pass");
    Dims d = dn;
    if constexpr (NZ > 0) d.nz = NZ;
    if (is_done(sc)) return;
    extern __shared__ unsigned char les_smem[];
    const int arows = g.J + 2;  // A-rows j0-1 .. j0+J
    const int zrows = arows + 2;
    T* zs = reinterpret_cast<T*>(les_smem);        // ring of z_{k+1}: [R][arows][nz]
    T* zi = zs + (SINGLE ? 0 : (size_t)R * arows * d.nz);  // plane t of z_k: [2][zrows][nz]
    __shared__ CodeTab<T> tab_s;
    code_tab_fill(tab_s, A);
    __syncthreads();
    const CodeTab<T>* tab = &tab_s;
    const int lb = logical_block(g.grid);
    const int jb = lb % g.njb, ic = lb / g.njb;
    const int j0 = jb * g.J;
    const int ia = ic * g.irange, ib = min(ia + g.irange, d.nx);
    double acc[1] = {0.0};
    if (ia >= d.nx) {  // padding of the grid (the whole workgroup)
        if (SINGLE) block_partial<1>(acc, red.partials, lb);
        return;
    }
    const int a = (int)threadIdx.x / g.tpr;  // A-row
    const int kk = ((int)threadIdx.x - a * g.tpr) * V;
    const int j = j0 - 1 + a;
    const bool jin = j >= 0 && j < d.ny;
    const bool core = a >= 1 && a <= g.J && j < d.ny;
    // the extra input row this thread also loads: j0-2 (A-row 0) or j0+J+1
    // (the last A-row); its LDS row index and whether it lies in the grid
    const int xa = SINGLE ? -1 : (a == 0) ? 0 : (a == arows - 1 ? zrows - 1 : -1);
    const int xj = (a == 0) ? j0 - 2 : j0 + g.J + 1;
    const bool xin = xa >= 0 && xj >= 0 && xj < d.ny;
    // 32-bit cell offsets (the grid holds fewer than 2^32 cells)
    const unsigned usij = (unsigned)d.sij;
    const unsigned col = (unsigned)(j * d.nz + kk), xcol = (unsigned)(xj * d.nz + kk);
    // plane i's first cell, wave-uniform and opaque to the compiler's
    // induction-variable rewriting, so it stays in a scalar register
    auto pl = [&](int i) {
        return (unsigned)__builtin_amdgcn_readfirstlane((int)((unsigned)i * usij));
    };
    auto off = [&](int i, unsigned c0) { return pl(i) + c0; };
    // the arrays as buffer resources from the first plane the workgroup
    // touches (ia - 2): plane i's byte offset from there is scalar, a
    // thread's column is its fixed vector offset
    const int pbase = max(ia - 2, 0);
    const size_t cbase = (size_t)pbase * d.sij;
    auto rsrc = [&](const auto* p) { return buf_rsrc(p ? p + cbase : p); };
    const Rsrc rz = rsrc(FIRST ? r : z_in), rr = rsrc(r), rp = rsrc(p_in), rcd = rsrc(A.code),
               rzo = rsrc(z_out), rpo = rsrc(p_out);
    auto so = [&](int i, unsigned es) { return (unsigned)(i - pbase) * usij * es; };
    constexpr unsigned ET = sizeof(T), EC = sizeof(uint16_t);
    const unsigned colt = col * ET, colc = col * EC;
    const int zplane = zrows * d.nz, rplane = arows * d.nz;
    const T a0 = (T)alpha0, bta = (T)beta_a, ata = (T)alpha_a, btb = (T)beta_b, atb = (T)alpha_b;
    constexpr int prio = LES_PAIR_PRIO >= 0 ? LES_PAIR_PRIO : (sizeof(T) == 4 ? 1 : 2);
    T pad[V];
#pragma unroll
    for (int v = 0; v < V; ++v) pad[v] = outside<T>();
    // z_k as stored (r for the first pair, scaled where used), outside<T>()
    // outside the grid
    auto load_z = [&](int i, bool in, unsigned c0, T (&o)[V]) {
        if (in && i >= 0 && i < d.nx) bload<T, V, false>(rz, c0 * ET, so(i, ET), o);
        else {
#pragma unroll
            for (int v = 0; v < V; ++v) o[v] = pad[v];
        }
    };
    // the value step A uses: alpha0 r in the first pair for loaded values
    auto zval = [&](bool loaded, const T (&x)[V], T (&o)[V]) {
#pragma unroll
        for (int v = 0; v < V; ++v) o[v] = (FIRST && loaded) ? T(0) + a0 * x[v] : x[v];
    };
    auto write_zi = [&](int i, int buf, const T (&own)[V], const T (&ext)[V]) {
        T* zt = zi + buf * zplane;
        const bool pin = i >= 0 && i < d.nx;
        T o[V];
        zval(pin && jin, own, o);
        vstore<T, V>(zt + (a + 1) * d.nz + kk, o);
        if (xa >= 0) {
            zval(pin && xin, ext, o);
            vstore<T, V>(zt + xa * d.nz + kk, o);
        }
    };

    // Registers of one plane step; the roles rotate from plane to plane
    // (z_k over four buffers, the rest over two), and the march runs four
    // planes per loop pass with the roles renamed, so no copies are made.
    T zb[4][V], xb[2][V];                    // z_k at t-1, t, t+1, t+2; extra row at t+1, t+2
    T rb[2][V], pb[2][V];                    // r and p of plane t, t+1 (non-first)
    uint16_t cb[2][V];                       // code of plane t, t+1
    T pk[2][V], rk[2][V];                    // p_k, r of core rows for step B: plane t-1, t
    uint16_t ck[2][V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
        rb[0][v] = rb[1][v] = pb[0][v] = pb[1][v] = pk[0][v] = pk[1][v] = rk[0][v] = rk[1][v] =
            T(0);
        cb[0][v] = cb[1][v] = ck[0][v] = ck[1][v] = 0;
    }
    const int t0 = SINGLE ? ia : ia - 1;
    const int t1 = SINGLE ? ib - 1 : ib;
    load_z(t0 - 1, jin, col, zb[0]);
    load_z(t0, jin, col, zb[1]);
    load_z(t0 + 1, jin, col, zb[2]);
    {
        T x0[V];
        load_z(t0, xin, xcol, x0);
        load_z(t0 + 1, xin, xcol, xb[0]);
        write_zi(t0, 0, zb[1], x0);
    }
    if (jin && t0 >= 0) {
        if (!FIRST) {
            bload<T, V, LES_NT>(rr, colt, so(t0, ET), rb[0]);
            bload<T, V, LES_NT>(rp, colt, so(t0, ET), pb[0]);
        }
        bload<uint16_t, V, false>(rcd, colc, so(t0, EC), cb[0]);
    }
    __syncthreads();
    // one plane: zw, zc, ze, zn are z_k at t-1 .. t+2 (zn loaded here), xz/xn
    // the extra row at t+1/t+2, rc/rn, pc/pn, cc/cn plane t/t+1, pkeep etc.
    // step B's inputs from plane t-1 and pnew etc. step A's outputs
    // s: the plane's place in the four-plane pass, which is also its ring
    // slot and (s & 1) its LDS copy of z_k
    auto plane = [&](int t, int s, const T (&zw)[V], const T (&zc)[V], const T (&ze)[V], T (&zn)[V],
                     const T (&xz)[V], T (&xn)[V], const T (&rc)[V], T (&rn)[V],
                     const T (&pc)[V], T (&pn_)[V], const uint16_t (&cc)[V], uint16_t (&cn)[V],
                     const T (&pkeep)[V], T (&pnew)[V], const T (&rkeep)[V], T (&rnew)[V],
                     const uint16_t (&ckeep)[V], uint16_t (&cnew)[V])
                     __attribute__((always_inline)) {
        if (prio == 1) __builtin_amdgcn_s_setprio(2);
        // ---- loads for the next planes ----
#pragma unroll
        for (int v = 0; v < V; ++v) {
            rn[v] = pn_[v] = T(0);
            xn[v] = pad[v];
            cn[v] = 0;
        }
        load_z(t + 2, jin, col, zn);
        if (xa >= 0) load_z(t + 2, xin, xcol, xn);
        if (jin && t + 1 < d.nx) {
            if (!FIRST) {
                bload<T, V, LES_NT>(rr, colt, so(t + 1, ET), rn);
                bload<T, V, LES_NT>(rp, colt, so(t + 1, ET), pn_);
            }
            bload<uint16_t, V, false>(rcd, colc, so(t + 1, EC), cn);
        }
        // ---- A at plane t: z_{k+1} into the ring, p_k of core rows ----
        T* ring_t = zs + s * rplane + a * d.nz + kk;
        if (t >= 0 && t < d.nx && (SINGLE ? core : jin)) {
            Nbr<T, V> nb;
            zval(true, zc, nb.c);
            if (t > 0) zval(true, zw, nb.w);
            else {
#pragma unroll
                for (int v = 0; v < V; ++v) nb.w[v] = pad[v];
            }
            if (t < d.nx - 1) zval(true, ze, nb.e);
            else {
#pragma unroll
                for (int v = 0; v < V; ++v) nb.e[v] = pad[v];
            }
            const T* zt = zi + (s & 1) * zplane;
            vload<T, V>(zt + a * d.nz + kk, nb.s);
            vload<T, V>(zt + (a + 2) * d.nz + kk, nb.n);
            const int id = (int)off(t, col);
            row_edges<T, V, false>(d, (int)id, kk, nb.c, nb.km, nb.kp,
                                   [&](int n) { return zt[(a + 1) * d.nz + kk + (n - (int)id)]; });
            T az[V];
            apply7v<1, T, V, true>(A, d, (int)id, t, j, kk, nb, az, cc, nullptr, tab);
            const T* rv = FIRST ? zc : rc;
            const T* pp = FIRST ? zc : pc;
            T zo[V];
#pragma unroll
            for (int v = 0; v < V; ++v) {
                const T res = rv[v] - az[v];
                const T pkv = res + bta * pp[v];
                zo[v] = nb.c[v] + ata * pkv;
                pnew[v] = pkv;
                rnew[v] = rv[v];
                cnew[v] = cc[v];
            }
            if (SINGLE) {
                bstore<T, V, LES_NT>(rzo, colt, so(t, ET), zo);
#pragma unroll
                for (int v = 0; v < V; ++v) acc[0] += (double)rv[v] * (double)zo[v];
            } else {
                vstore<T, V>(ring_t, zo);
            }
        } else if (!SINGLE) {
            vstore<T, V>(ring_t, pad);
        }
        // ---- plane t+1 of z_k into the other LDS buffer ----
        write_zi(t + 1, (s + 1) & 1, ze, xz);
        if (prio == 1) __builtin_amdgcn_s_setprio(0);
        __syncthreads();
        if (prio == 2) __builtin_amdgcn_s_setprio(2);
        // ---- B at plane t-1 from the ring ----
        const int iB = t - 1;
        if (!SINGLE && core && iB >= ia && iB < ib) {
            const T* zc_pl = zs + ((s + 3) & 3) * rplane;
            const T* zm_pl = zs + ((s + 2) & 3) * rplane;
            const T* zp_pl = zs + s * rplane;
            Nbr<T, V> nb;
            vload<T, V>(zc_pl + a * d.nz + kk, nb.c);
            vload<T, V>(zm_pl + a * d.nz + kk, nb.w);
            vload<T, V>(zp_pl + a * d.nz + kk, nb.e);
            vload<T, V>(zc_pl + (a - 1) * d.nz + kk, nb.s);
            vload<T, V>(zc_pl + (a + 1) * d.nz + kk, nb.n);
            const int id = (int)off(iB, col);
            row_edges<T, V, false>(d, (int)id, kk, nb.c, nb.km, nb.kp,
                                   [&](int n) { return zc_pl[a * d.nz + kk + (n - (int)id)]; });
            T az[V];
            apply7v<1, T, V, true>(A, d, (int)id, iB, j, kk, nb, az, ckeep, nullptr, tab);
            T pn[V], zo[V];
#pragma unroll
            for (int v = 0; v < V; ++v) {
                const T res = rkeep[v] - az[v];
                pn[v] = res + btb * pkeep[v];
                zo[v] = nb.c[v] + atb * pn[v];
            }
            bstore<T, V, LES_NT>(rpo, colt, so(iB, ET), pn);
            bstore<T, V, LES_NT>(rzo, colt, so(iB, ET), zo);
        }
        if (prio == 2) __builtin_amdgcn_s_setprio(0);
    };
    for (int t = t0; t <= t1; t += 4) {
        plane(t, 0, zb[0], zb[1], zb[2], zb[3], xb[0], xb[1], rb[0], rb[1], pb[0], pb[1], cb[0],
              cb[1], pk[0], pk[1], rk[0], rk[1], ck[0], ck[1]);
        if (t + 1 > t1) break;
        plane(t + 1, 1, zb[1], zb[2], zb[3], zb[0], xb[1], xb[0], rb[1], rb[0], pb[1], pb[0], cb[1],
              cb[0], pk[1], pk[0], rk[1], rk[0], ck[1], ck[0]);
        if (t + 2 > t1) break;
        plane(t + 2, 2, zb[2], zb[3], zb[0], zb[1], xb[0], xb[1], rb[0], rb[1], pb[0], pb[1], cb[0],
              cb[1], pk[0], pk[1], rk[0], rk[1], ck[0], ck[1]);
        if (t + 3 > t1) break;
        plane(t + 3, 3, zb[3], zb[0], zb[1], zb[2], xb[1], xb[0], rb[1], rb[0], pb[1], pb[0], cb[1],
              cb[0], pk[1], pk[0], rk[1], rk[0], ck[1], ck[0]);
    }
    if (SINGLE) block_partial<1>(acc, red.partials, lb);
}

// ------------------------------------------------------------------
// q = A p with the direction update fused (PU 0: p given; PU 1:
// p = z + beta p_old, stored) and the deferred x update, marching along i
// like the Chebyshev pair: for plane t a workgroup forms p on its J rows and
// one halo row on each side into an R-plane LDS ring (once per cell, where
// k_pcg_matvec forms it again at every neighbour), then applies A at plane
// t-1 from the ring.  Every cell's arithmetic is k_pcg_matvec's, so p, q and
// x are bit-identical to it; only the order in which p.q is summed differs.
// One kBlock workgroup per CU and at most kMaxGrid of them, so the partials
// fit the finalize.  p_old, z and p must be distinct buffers.
// ------------------------------------------------------------------
template <typename T, int CODED, int V, int PU, int P, int R>
__global__ __launch_bounds__(kBlock) void k_pcg_matvec_ring(Dims d, PairGeom g, Stencil<T> A,
                                                            const T* __restrict__ p_old,
                                                            const T* __restrict__ z,
                                                            T* __restrict__ p,
                                                            T* __restrict__ q, RedScratch red,
                                                            const SolveScalars* sc,
                                                            const double* __restrict__ wgt,
                                                            T* __restrict__ x, RedScratch rz,
                                                            int rz_grid)
{
    if (sc->done) return;
    extern __shared__ unsigned char les_smem[];
    T* ps = reinterpret_cast<T*>(les_smem);  // [R][P*rpp][nz]
    const CodeTab<T>* tab = nullptr;
    if constexpr (CODED == 1 && LES_MV_TAB) {
        __shared__ CodeTab<T> tab_s;
        code_tab_fill(tab_s, A);
        __syncthreads();
        tab = &tab_s;
    }
    const int lb = logical_block(g.grid);
    const int jb = lb % g.njb, ic = lb / g.njb;
    const int j0 = jb * g.J;
    const int ia = ic * g.irange, ib = min(ia + g.irange, d.nx);
    const int sub = (int)threadIdx.x / g.tpr;
    const int kk = ((int)threadIdx.x - sub * g.tpr) * V;
    const int plane_sz = P * g.rpp * d.nz;
    const T beta = (T)fold_beta(sc, rz, rz_grid);
    const T ax = (T)sc->alpha;
    double acc[1] = {0.0};
    // p_old of the core rows, kept for the x update of the same cells one
    // plane later
    T okeep[P][V], onew[P][V];
    T pad[V];
#pragma unroll
    for (int v = 0; v < V; ++v) pad[v] = outside<T>();
#pragma unroll
    for (int pp = 0; pp < P; ++pp)
#pragma unroll
        for (int v = 0; v < V; ++v) okeep[pp][v] = onew[pp][v] = T(0);
    for (int t = ia - 1; t <= ib && ia < d.nx; ++t) {
        // ---- p at plane t into the ring (the coded operator reads rows and
        // planes outside the grid as outside<T>() from it) ----
#pragma unroll
        for (int pp = 0; pp < P; ++pp) {
            const int qr = pp * g.rpp + sub;
            const int a = qr < g.J ? qr + 1 : (qr == g.J ? 0 : g.J + 1);
            const int j = j0 - 1 + a;
            T* slot = ps + (t + R) % R * plane_sz + a * d.nz + kk;
            if (t < 0 || t >= d.nx || j < 0 || j >= d.ny) {
                if (CODED == 1) vstore<T, V>(slot, pad);
                continue;
            }
            const int id = (t * d.ny + j) * d.nz + kk;
            T pv[V];
            if (PU == 0) {
                vload<T, V>(p + id, pv);
            } else {
                T zv[V], ov[V];
                vload_s<T, V>(z + id, zv);
                vload_s<T, V>(p_old + id, ov);
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    pv[v] = zv[v] + beta * ov[v];
                    onew[pp][v] = ov[v];
                }
                if (qr < g.J && t >= ia && t < ib) vstore_s<T, V>(p + id, pv);
            }
            vstore<T, V>(slot, pv);
        }
        __syncthreads();
        // ---- q = A p at plane t-1 from the ring ----
        const int iB = t - 1;
        if (iB >= ia && iB < ib) {
            const T* pc_pl = ps + (iB + R) % R * plane_sz;
            const T* pm_pl = ps + (iB + R - 1) % R * plane_sz;
            const T* pp_pl = ps + (iB + 1) % R * plane_sz;
#pragma unroll
            for (int pp = 0; pp < P; ++pp) {
                const int qr = pp * g.rpp + sub;
                if (qr >= g.J) continue;
                const int a = qr + 1, j = j0 + qr;
                if (j >= d.ny) continue;
                const int i = iB, k = kk, id = (i * d.ny + j) * d.nz + kk;
                Nbr<T, V> nb;
                vload<T, V>(pc_pl + a * d.nz + kk, nb.c);
                if (CODED == 1) {
                    vload<T, V>(pm_pl + a * d.nz + kk, nb.w);
                    vload<T, V>(pp_pl + a * d.nz + kk, nb.e);
                    vload<T, V>(pc_pl + (a - 1) * d.nz + kk, nb.s);
                    vload<T, V>(pc_pl + (a + 1) * d.nz + kk, nb.n);
                } else {
#pragma unroll
                    for (int v = 0; v < V; ++v) nb.w[v] = nb.e[v] = nb.s[v] = nb.n[v] = T(0);
                    if (i > 0) vload<T, V>(pm_pl + a * d.nz + kk, nb.w);
                    if (i < d.nx - 1) vload<T, V>(pp_pl + a * d.nz + kk, nb.e);
                    if (j > 0) vload<T, V>(pc_pl + (a - 1) * d.nz + kk, nb.s);
                    if (j < d.ny - 1) vload<T, V>(pc_pl + (a + 1) * d.nz + kk, nb.n);
                }
                row_edges<T, V, false>(d, id, k, nb.c, nb.km, nb.kp,
                                       [&](int n) { return pc_pl[a * d.nz + kk + (n - id)]; });
                T qv[V];
                apply7v<CODED, T, V, CODED == 1 && LES_MV_TAB>(A, d, id, i, j, k, nb, qv, nullptr,
                                                                nullptr, tab);
                vstore_s<T, V>(q + id, qv);
                if (PU == 1 && x) {
                    T xv[V];
                    vload<T, V>(x + id, xv);
#pragma unroll
                    for (int v = 0; v < V; ++v) xv[v] = xv[v] + ax * okeep[pp][v];
                    vstore_s<T, V>(x + id, xv);
                }
                // the coded operator's mirror weights come from its code
                // (the array was made from the same code)
                uint16_t cw[V];
                if (CODED == 1 && wgt) vload<uint16_t, V>(A.code + id, cw);
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    const double pq = (double)nb.c[v] * (double)qv[v];
                    const double wv = !wgt ? 1.0 : (CODED == 1 ? code_weight(cw[v]) : wgt[id + v]);
                    acc[0] += wgt ? wv * pq : pq;
                }
            }
        }
#pragma unroll
        for (int pp = 0; pp < P; ++pp)
#pragma unroll
            for (int v = 0; v < V; ++v) okeep[pp][v] = onew[pp][v];
    }
    block_partial<1>(acc, red.partials, lb);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_cheb_scale(Dims d, Chunking ch,
                                                       const T* __restrict__ r,
                                                       T* __restrict__ z, double alpha0,
                                                       RedScratch red, const SolveScalars* sc,
                                                       const double* __restrict__ wgt)
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
        const double rz = (double)rv * (double)zn;
        acc[0] += wgt ? wgt[id] * rz : rz;
    }
    if (red.partials) block_partial<1>(acc, red.partials, lb);
}

// ------------------------------------------------------------------
// mixed precision: FP32 iteration with FP64 reliable residual replacement
// ------------------------------------------------------------------
// x64 += x32 + alpha p_last, the last update's x term, which the next
// matvec would otherwise add (p_last: the direction of iteration iter-1,
// p0 for even iterations, p1 for odd ones).  x32 is left at
// -(alpha p_last), so that next matvec brings it to exactly 0 -- the value
// an eager update followed by this fold would leave.
template <int V>
__global__ __launch_bounds__(kBlock) void k_mixed_fold(Dims d, Chunking ch,
                                                       double* __restrict__ x64,
                                                       float* __restrict__ x32,
                                                       const float* __restrict__ p0,
                                                       const float* __restrict__ p1, int gate,
                                                       const SolveScalars* sc)
{
    if (!gate_open(gate, sc)) return;
    const int lb = logical_block(ch.grid);
    const int it = sc->iter;
    const float ax = (float)sc->alpha;
    const double unscale = 1.0 / sc->rscale;  // exact: a power of two
    const float* pl = ((it - 1) & 1) ? p1 : p0;
    FOR_VCELLS(d, ch, lb, V)
    {
        float xs[V], pv[V];
        double xd[V];
        vload<float, V>(x32 + id, xs);
        vload<double, V>(x64 + id, xd);
        if (it > 0) vload<float, V>(pl + id, pv);
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const float ap = it > 0 ? ax * pv[v] : 0.0f;
            const float xv = it > 0 ? xs[v] + ap : xs[v];
            xd[v] = xd[v] + (double)xv * unscale;
            xs[v] = -ap;
        }
        vstore<double, V>(x64 + id, xd);
        vstore<float, V>(x32 + id, xs);
    }
}

// x += alpha p_last (FP64 solve): the last update's x term, deferred to the
// next matvec, which does not run once the solve has stopped.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_x_pending(Dims d, Chunking ch, T* __restrict__ x,
                                                          const T* __restrict__ p0,
                                                          const T* __restrict__ p1,
                                                          const SolveScalars* sc)
{
    const int it = sc->iter;
    if (it <= 0) return;
    const int lb = logical_block(ch.grid);
    const T ax = (T)sc->alpha;
    const T* pl = ((it - 1) & 1) ? p1 : p0;
    FOR_CELLS(d, ch, lb) x[id] = x[id] + ax * pl[id];
}

// V cells of a row per thread (V = 1 where nz is not a multiple of the
// vector width), the neighbours as the matvec loads them.
template <int CODED, int V>
__global__ __launch_bounds__(kBlock) void k_mixed_residual(Dims d, Chunking ch, Stencil<double> A,
                                                           const double* __restrict__ b,
                                                           const double* __restrict__ x,
                                                           float* __restrict__ r32, int gate,
                                                           RedScratch red, const SolveScalars* sc)
{
    if (!gate_open(gate, sc)) return;
    const int lb = logical_block(ch.grid);
    const double rs = r32 ? sc->rscale : 1.0;
    double acc[2] = {0.0, 0.0};
    FOR_VCELLS(d, ch, lb, V)
    {
        Nbr<double, V> nb;
        load_nbr<double, V>(x, d, id, i, j, k, [](double v) { return v; }, nb);
        double q[V], bv[V];
        apply7v<CODED, double, V>(A, d, id, i, j, k, nb, q);
        vload<double, V>(b + id, bv);
        float rv32[V];
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const double rv = bv[v] - q[v];
            rv32[v] = (float)(rv * rs);
            acc[0] += rv * rv;
            acc[1] += bv[v] * bv[v];
        }
        if (r32) vstore_s<float, V>(r32 + id, rv32);
    }
    if (red.partials) block_partial<2>(acc, red.partials, lb);
}

template <int V, bool FOLD>
__global__ __launch_bounds__(kBlock) void k_pcg_update_mixed(Dims d, Chunking ch,
                                                             float* __restrict__ x,
                                                             float* __restrict__ r,
                                                             const float* __restrict__ p,
                                                             const float* __restrict__ q,
                                                             RedScratch red,
                                                             const SolveScalars* sc,
                                                             RedScratch pq_red, int pq_grid,
                                                             RedScratch rz_red, int rz_grid)
{
    if (sc->done) return;
    const int lb = logical_block(ch.grid);
    double alpha;
    if (FOLD) {
        // the matvec's finalize (EPI_ALPHA) in every workgroup: the same
        // fixed-order sum, the same division; workgroup 0 stores the result
        // for the kernels that follow.  No thread of this launch reads the
        // stored pq or alpha.  With rz_grid > 0, rho (this iteration's r.z)
        // is likewise the sum of the preconditioner's partials, and
        // workgroup 0 also does what EPI_BETA's finalize did (the matvec
        // already formed beta the same way); then no other thread reads
        // rho, beta or the restart flag.
        double tot[1];
        reduce_partials<1, false>(pq_red.partials, pq_grid, tot);
        double rho;
        if (rz_grid > 0) {
            double rz[1];
            reduce_partials<1, false>(rz_red.partials, rz_grid, rz);
            rho = rz[0];
        } else {
            rho = sc->rho;
        }
        alpha = rho / tot[0];
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            SolveScalars* w = const_cast<SolveScalars*>(sc);
            if (rz_grid > 0) {
                if (w->restart_pending) {
                    w->beta = 0.0;
                    w->restart_pending = 0;
                    w->restarts += 1;
                } else {
                    w->beta = rho / w->rho;
                }
                w->rho = rho;
            }
            w->pq = tot[0];
            w->alpha = alpha;
        }
    } else {
        alpha = sc->alpha;
    }
    const float ax = (float)alpha;
    const float ar = (float)(-alpha);
    double acc[1] = {0.0};
    FOR_VCELLS(d, ch, lb, V)
    {
        float rv[V], qv[V];
        if (x) {  // null: x is updated by the next iteration's matvec
            float xv[V], pv[V];
            vload<float, V>(x + id, xv);
            vload<float, V>(p + id, pv);
#pragma unroll
            for (int v = 0; v < V; ++v) xv[v] = xv[v] + ax * pv[v];
            vstore_s<float, V>(x + id, xv);
        }
        vload<float, V>(r + id, rv);
        vload<float, V>(q + id, qv);
#pragma unroll
        for (int v = 0; v < V; ++v) {
            rv[v] = rv[v] + ar * qv[v];
            acc[0] += (double)rv[v] * (double)rv[v];
        }
        vstore_s<float, V>(r + id, rv);
    }
    block_partial<1>(acc, red.partials, lb);
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

int flat_grid(int n)
{
    int g = (n + kBlock - 1) / kBlock;
    if (g > 8192) g = 8192;
    return g < 1 ? 1 : g;
}


// ------------------------------------------------------------------
// Multigrid preconditioner (no reference counterpart: the reference's
// README.md:53 lists multigrid as a TODO).  Galerkin aggregation over 2x2x2
// blocks: piecewise-constant prolongation P, restriction P^T (the block
// sum), coarse operator P^T A P (again 7-point: a coupling inside a block
// adds to the coarse centre, one across a face to that face's coefficient).
// One V-cycle of damped-Jacobi sweeps from x = 0 is the preconditioner.
// Every per-cell operation order below (children in (di, dj, dk) order, a
// child's couplings in StencilMatrix member order, the sweep's apply7 order)
// is the oracle's, so the FP64 V-cycle is bit-identical to it.
// ------------------------------------------------------------------
template <typename T>
struct Arr7 {
    T* a[7];
};

// Ac = P^T A P, one thread per coarse cell (f: the fine dimensions); with
// W, P^T W A P (each fine coefficient times W of its row)
__global__ __launch_bounds__(kBlock) void k_mg_coarsen(Dims f, Dims cd, Chunking ch,
                                                       Stencil<double> A, const double* W,
                                                       Arr7<double> Ac)
{
    const int lb = logical_block(ch.grid);
    FOR_CELLS(cd, ch, lb)
    {
        double c = 0.0, cn = 0.0, cs = 0.0, ce = 0.0, cw = 0.0, ct = 0.0, cb = 0.0;
        for (int di = 0; di < 2; ++di)
            for (int dj = 0; dj < 2; ++dj)
                for (int dk = 0; dk < 2; ++dk) {
                    const int fi = 2 * i + di, fj = 2 * j + dj, fk = 2 * k + dk;
                    if (fi >= f.nx || fj >= f.ny || fk >= f.nz) continue;
                    const int fid = (fi * f.ny + fj) * f.nz + fk;
                    auto wa = [&](int q) { return W ? W[fid] * A.a[q][fid] : A.a[q][fid]; };
                    c += wa(SC);
                    if (fj + 1 < f.ny) { if (dj == 0) c += wa(SN); else cn += wa(SN); }
                    if (fj > 0) { if (dj == 1) c += wa(SS); else cs += wa(SS); }
                    if (fi + 1 < f.nx) { if (di == 0) c += wa(SE); else ce += wa(SE); }
                    if (fi > 0) { if (di == 1) c += wa(SW); else cw += wa(SW); }
                    if (fk + 1 < f.nz) { if (dk == 0) c += wa(ST); else ct += wa(ST); }
                    if (fk > 0) { if (dk == 1) c += wa(SB); else cb += wa(SB); }
                }
        Ac.a[SC][id] = c;
        Ac.a[SN][id] = cn;
        Ac.a[SS][id] = cs;
        Ac.a[SE][id] = ce;
        Ac.a[SW][id] = cw;
        Ac.a[ST][id] = ct;
        Ac.a[SB][id] = cb;
    }
}

// One damped-Jacobi sweep x_out = x + omega ((b - A x) / A_c) on a level.
// MG_ZERO: from x = 0, x_out = omega (b / A_c).  MG_PROLONG: x is first
// corrected by scale * e of the cell's block on the next coarser level (cd),
// at the cell and at each neighbour the stencil reads.  RZ: b.x_out partials
// into red (the PCG's r.z, b being the level-0 residual; W-weighted when
// wgt is given).
template <typename T, int CODED, int MODE, bool RZ>
__global__ __launch_bounds__(kBlock) void k_mg_smooth(Dims d, Chunking ch, Stencil<T> A,
                                                      const T* __restrict__ x,
                                                      const T* __restrict__ b,
                                                      const T* __restrict__ e, Dims cd,
                                                      T* __restrict__ x_out, double omega,
                                                      double scale, RedScratch red,
                                                      const SolveScalars* sc,
                                                      const double* __restrict__ wgt)
{
    if (is_done(sc)) return;
    const int lb = logical_block(ch.grid);
    const T om = (T)omega, sf = (T)scale;
    double acc[1] = {0.0};
    FOR_CELLS(d, ch, lb)
    {
        T c, w, ea, s, n, bb, t;
        coefs7<CODED>(A, id, c, w, ea, s, n, bb, t);
        T out;
        if constexpr (MODE == MG_ZERO) {
            out = om * (b[id] / c);
        } else {
            // x at (i+di, j+dj, k+dk), corrected in MG_PROLONG
            auto xv = [&](int di, int dj, int dk) -> T {
                const int nid = id + di * d.sij + dj * d.nz + dk;
                if constexpr (MODE == MG_PROLONG) {
                    const int pi = (i + di) >> 1, pj = (j + dj) >> 1, pk = (k + dk) >> 1;
                    return x[nid] + sf * e[(pi * cd.ny + pj) * cd.nz + pk];
                } else {
                    return x[nid];
                }
            };
            const T xc = xv(0, 0, 0);
            T q = c * xc;
            if (i > 0) q += w * xv(-1, 0, 0);
            if (i < d.nx - 1) q += ea * xv(1, 0, 0);
            if (j > 0) q += s * xv(0, -1, 0);
            if (j < d.ny - 1) q += n * xv(0, 1, 0);
            if (k > 0) q += bb * xv(0, 0, -1);
            if (k < d.nz - 1) q += t * xv(0, 0, 1);
            out = xc + om * ((b[id] - q) / c);
        }
        x_out[id] = out;
        if (RZ) {
            const double bz = (double)b[id] * (double)out;
            acc[0] += wgt ? wgt[id] * bz : bz;
        }
    }
    if (RZ) block_partial<1>(acc, red.partials, lb);
}

// b_c = P^T (b - A x), or P^T W (b - A x): one thread per coarse cell,
// children in (di, dj, dk) order
template <typename T, int CODED>
__global__ __launch_bounds__(kBlock) void k_mg_restrict(Dims f, Dims cd, Chunking ch,
                                                        Stencil<T> A, const T* __restrict__ x,
                                                        const T* __restrict__ b,
                                                        T* __restrict__ bc,
                                                        const SolveScalars* sc,
                                                        const double* __restrict__ W)
{
    if (is_done(sc)) return;
    const int lb = logical_block(ch.grid);
    FOR_CELLS(cd, ch, lb)
    {
        T acc = T(0);
        for (int di = 0; di < 2; ++di)
            for (int dj = 0; dj < 2; ++dj)
                for (int dk = 0; dk < 2; ++dk) {
                    const int fi = 2 * i + di, fj = 2 * j + dj, fk = 2 * k + dk;
                    if (fi >= f.nx || fj >= f.ny || fk >= f.nz) continue;
                    const int fid = (fi * f.ny + fj) * f.nz + fk;
                    const T q = apply7<CODED>(A, f, fid, fi, fj, fk, [&](int m) { return x[m]; });
                    const T res = b[fid] - q;
                    acc += W ? (T)W[fid] * res : res;
                }
        bc[id] = acc;
    }
}

// The diagonal of V cells (the coded centre, or A_c)
template <int CODED, typename T, int V>
__device__ __forceinline__ void diag7v(const Stencil<T>& A, int id, T (&dg)[V])
{
    if constexpr (CODED == 1) {
        uint16_t cd[V];
        vload<uint16_t, V>(A.code + id, cd);
#pragma unroll
        for (int v = 0; v < V; ++v) dg[v] = (cd[v] & CODE_IDENTITY) ? T(1) : A.cfluid;
    } else {
        vload<T, V>(A.a[SC] + id, dg);
    }
}

// k_mg_smooth with V consecutive cells of a row per thread (16-byte loads;
// k-neighbours by lane shifts as in the Chebyshev step).  In MG_PROLONG each
// neighbour vector is corrected by scale * e of its cells' blocks (cells
// k..k+V-1 lie in blocks k/2 ..), the k-edge values through the shifted
// corrected centres.  Each cell's arithmetic is k_mg_smooth's.  nz % V == 0.
template <typename T, int CODED, int MODE, bool RZ, int V>
__global__ __launch_bounds__(kBlock) void k_mg_smooth_v(Dims d, Chunking ch, Stencil<T> A,
                                                        const T* __restrict__ x,
                                                        const T* __restrict__ b,
                                                        const T* __restrict__ e, Dims cd,
                                                        T* __restrict__ x_out, double omega,
                                                        double scale, RedScratch red,
                                                        const SolveScalars* sc,
                                                        const double* __restrict__ wgt)
{
    if (is_done(sc)) return;
    const int lb = logical_block(ch.grid);
    const T om = (T)omega, sf = (T)scale;
    double acc[1] = {0.0};
    FOR_VCELLS(d, ch, lb, V)
    {
        T bv[V], dg[V], out[V];
        vload_s<T, V>(b + id, bv);
        diag7v<CODED, T, V>(A, id, dg);
        if constexpr (MODE == MG_ZERO) {
#pragma unroll
            for (int v = 0; v < V; ++v) out[v] = om * (bv[v] / dg[v]);
        } else {
            Nbr<T, V> nb;
            if constexpr (MODE == MG_PROLONG) {
                // x + scale * e at the cells (ii, jj, kk..), e of their blocks
                auto eat = [&](int ii, int jj, int kk) {
                    return e[((ii >> 1) * cd.ny + (jj >> 1)) * cd.nz + (kk >> 1)];
                };
                // cells k..k+V-1 lie in blocks k/2 .. (k+V-1)/2: one load
                // for V = 2, an aligned pair for V = 4 (k % 4 == 0)
                auto corr = [&](int ii, int jj, T (&vals)[V]) {
                    const T* er = e + ((ii >> 1) * cd.ny + (jj >> 1)) * cd.nz + (k >> 1);
                    if constexpr (V == 4) {
                        T ev[2];
                        vload<T, 2>(er, ev);
#pragma unroll
                        for (int v = 0; v < V; ++v) vals[v] = vals[v] + sf * ev[v >> 1];
                    } else if constexpr (V == 2) {
                        const T ev = er[0];
#pragma unroll
                        for (int v = 0; v < V; ++v) vals[v] = vals[v] + sf * ev;
                    } else {
#pragma unroll
                        for (int v = 0; v < V; ++v) vals[v] = vals[v] + sf * eat(ii, jj, k + v);
                    }
                };
                load_nbr5<T, V>(x, d, id, i, j, nb);
                corr(i, j, nb.c);
                if (i > 0) corr(i - 1, j, nb.w);
                if (i < d.nx - 1) corr(i + 1, j, nb.e);
                if (j > 0) corr(i, j - 1, nb.s);
                if (j < d.ny - 1) corr(i, j + 1, nb.n);
                row_edges<T, V, true>(d, id, k, nb.c, nb.km, nb.kp, [&](int n) {
                    return x[n] + sf * eat(i, j, k + (n - id));
                });
            } else if constexpr (MODE == MG_ZS) {
                // the first sweep's iterate om * (b / A_c) at the cells and
                // their neighbours, formed where it is read (bit for bit the
                // values an MG_ZERO launch would have stored)
                auto diag1 = [&](int n) {
                    if constexpr (CODED == 1)
                        return (A.code[n] & CODE_IDENTITY) ? T(1) : A.cfluid;
                    else
                        return A.a[SC][n];
                };
                auto zero_at = [&](int n, T (&vals)[V]) {
                    T g[V];
                    diag7v<CODED, T, V>(A, n, g);
#pragma unroll
                    for (int v = 0; v < V; ++v) vals[v] = om * (vals[v] / g[v]);
                };
                load_nbr5<T, V>(b, d, id, i, j, nb);
#pragma unroll
                for (int v = 0; v < V; ++v) nb.c[v] = om * (nb.c[v] / dg[v]);
                if (i > 0) zero_at(id - d.sij, nb.w);
                if (i < d.nx - 1) zero_at(id + d.sij, nb.e);
                if (j > 0) zero_at(id - d.nz, nb.s);
                if (j < d.ny - 1) zero_at(id + d.nz, nb.n);
                row_edges<T, V, true>(d, id, k, nb.c, nb.km, nb.kp,
                                      [&](int n) { return om * (b[n] / diag1(n)); });
            } else {
                load_nbr<T, V>(x, d, id, i, j, k, [](T v) { return v; }, nb);
            }
            T q[V];
            apply7v<CODED, T, V>(A, d, id, i, j, k, nb, q);
#pragma unroll
            for (int v = 0; v < V; ++v) out[v] = nb.c[v] + om * ((bv[v] - q[v]) / dg[v]);
        }
        vstore_s<T, V>(x_out + id, out);
        if (RZ) {
            // the coded operator's weights come from its code (the array was
            // made from the same code)
            uint16_t cw[V];
            if (CODED == 1 && wgt) vload<uint16_t, V>(A.code + id, cw);
#pragma unroll
            for (int v = 0; v < V; ++v) {
                const double bz = (double)bv[v] * (double)out[v];
                const double wv = !wgt ? 1.0 : (CODED == 1 ? code_weight(cw[v]) : wgt[id + v]);
                acc[0] += wgt ? wv * bz : bz;
            }
        }
    }
    if (RZ) block_partial<1>(acc, red.partials, lb);
}

// The first two sweeps from x = 0 (MG_ZS) marching along i like the ring
// matvec: for plane t a workgroup forms the first sweep's iterate
// om * (b / A_c) on its J rows and one halo row on each side into an R-plane
// LDS ring (once per cell, where k_mg_smooth_v forms it again at each of a
// cell's four i/j neighbours), then applies the second sweep at plane t-1
// from the ring.  b of the core rows is carried from plane to plane.  Every
// cell's arithmetic is k_mg_smooth_v's, so x_out is bit-identical to it.
template <typename T, int CODED, int V, int P, int R>
__global__ __launch_bounds__(kBlock) void k_mg_zs_ring(Dims d, PairGeom g, Stencil<T> A,
                                                       const T* __restrict__ b,
                                                       T* __restrict__ x_out, double omega,
                                                       const SolveScalars* sc)
{
    if (is_done(sc)) return;
    extern __shared__ unsigned char les_smem[];
    T* xs = reinterpret_cast<T*>(les_smem);  // [R][P*rpp][nz]
    const CodeTab<T>* tab = nullptr;
    if constexpr (CODED == 1 && LES_MV_TAB) {
        __shared__ CodeTab<T> tab_s;
        code_tab_fill(tab_s, A);
        __syncthreads();
        tab = &tab_s;
    }
    const int lb = logical_block(g.grid);
    const int jb = lb % g.njb, ic = lb / g.njb;
    const int j0 = jb * g.J;
    const int ia = ic * g.irange, ib = min(ia + g.irange, d.nx);
    const int sub = (int)threadIdx.x / g.tpr;
    const int kk = ((int)threadIdx.x - sub * g.tpr) * V;
    const int plane_sz = P * g.rpp * d.nz;
    const T om = (T)omega;
    // b and the diagonal's source (the code, or A_c) of this thread's rows:
    // plane t-1 (b only, for the second sweep), t, t+1 and t+2 (loaded
    // while plane t is computed, so the loads have two planes of work to
    // arrive)
    using DS = std::conditional_t<CODED == 1, uint16_t, T>;
    T bkeep[P][V], bcur[P][V], bnx[P][V], bnx2[P][V];
    DS dkeep[P][V], dcur[P][V], dnx[P][V], dnx2[P][V];
    T pad[V];
#pragma unroll
    for (int v = 0; v < V; ++v) pad[v] = outside<T>();
    auto row_of = [&](int pp, int& a, int& j) {
        const int qr = pp * g.rpp + sub;
        a = qr < g.J ? qr + 1 : (qr == g.J ? 0 : g.J + 1);
        j = j0 - 1 + a;
    };
    auto load_plane = [&](int t, T (&bo)[P][V], DS (&dso)[P][V]) {
#pragma unroll
        for (int pp = 0; pp < P; ++pp) {
            int a, j;
            row_of(pp, a, j);
#pragma unroll
            for (int v = 0; v < V; ++v) {
                bo[pp][v] = T(0);
                dso[pp][v] = DS(0);
            }
            if (t < 0 || t >= d.nx || j < 0 || j >= d.ny) continue;
            const int id = (t * d.ny + j) * d.nz + kk;
            vload_s<T, V>(b + id, bo[pp]);
            if constexpr (CODED == 1) vload<uint16_t, V>(A.code + id, dso[pp]);
            else vload<T, V>(A.a[SC] + id, dso[pp]);
        }
    };
#pragma unroll
    for (int pp = 0; pp < P; ++pp)
#pragma unroll
        for (int v = 0; v < V; ++v) {
            bkeep[pp][v] = T(0);
            dkeep[pp][v] = DS(0);
        }
    if (ia < d.nx) {
        load_plane(ia - 1, bcur, dcur);
        load_plane(ia, bnx, dnx);
    }
    for (int t = ia - 1; t <= ib && ia < d.nx; ++t) {
        load_plane(t + 2, bnx2, dnx2);
        // ---- the first sweep's iterate at plane t into the ring (the
        // coded operator reads rows and planes outside the grid as
        // outside<T>() from it) ----
#pragma unroll
        for (int pp = 0; pp < P; ++pp) {
            int a, j;
            row_of(pp, a, j);
            T* slot = xs + (t + R) % R * plane_sz + a * d.nz + kk;
            if (t < 0 || t >= d.nx || j < 0 || j >= d.ny) {
                if (CODED == 1) vstore<T, V>(slot, pad);
                continue;
            }
            T xv[V];
#pragma unroll
            for (int v = 0; v < V; ++v) {
                T dg;
                if constexpr (CODED == 1) dg = (dcur[pp][v] & CODE_IDENTITY) ? T(1) : A.cfluid;
                else dg = dcur[pp][v];
                xv[v] = om * (bcur[pp][v] / dg);
            }
            vstore<T, V>(slot, xv);
        }
        __syncthreads();
        // ---- the second sweep at plane t-1 from the ring ----
        const int iB = t - 1;
        if (iB >= ia && iB < ib) {
            const T* pc_pl = xs + (iB + R) % R * plane_sz;
            const T* pm_pl = xs + (iB + R - 1) % R * plane_sz;
            const T* pp_pl = xs + (iB + 1) % R * plane_sz;
#pragma unroll
            for (int pp = 0; pp < P; ++pp) {
                const int qr = pp * g.rpp + sub;
                if (qr >= g.J) continue;
                const int a = qr + 1, j = j0 + qr;
                if (j >= d.ny) continue;
                const int i = iB, k = kk, id = (i * d.ny + j) * d.nz + kk;
                Nbr<T, V> nb;
                vload<T, V>(pc_pl + a * d.nz + kk, nb.c);
                if (CODED == 1) {
                    vload<T, V>(pm_pl + a * d.nz + kk, nb.w);
                    vload<T, V>(pp_pl + a * d.nz + kk, nb.e);
                    vload<T, V>(pc_pl + (a - 1) * d.nz + kk, nb.s);
                    vload<T, V>(pc_pl + (a + 1) * d.nz + kk, nb.n);
                } else {
#pragma unroll
                    for (int v = 0; v < V; ++v) nb.w[v] = nb.e[v] = nb.s[v] = nb.n[v] = T(0);
                    if (i > 0) vload<T, V>(pm_pl + a * d.nz + kk, nb.w);
                    if (i < d.nx - 1) vload<T, V>(pp_pl + a * d.nz + kk, nb.e);
                    if (j > 0) vload<T, V>(pc_pl + (a - 1) * d.nz + kk, nb.s);
                    if (j < d.ny - 1) vload<T, V>(pc_pl + (a + 1) * d.nz + kk, nb.n);
                }
                row_edges<T, V, false>(d, id, k, nb.c, nb.km, nb.kp,
                                       [&](int n) { return pc_pl[a * d.nz + kk + (n - id)]; });
                T qv[V], dg[V], out[V];
                if constexpr (CODED == 1) {
                    // the code of plane t-1, carried from its step A
                    apply7v<1, T, V, LES_MV_TAB>(A, d, id, i, j, k, nb, qv, dkeep[pp], nullptr,
                                                 tab);
#pragma unroll
                    for (int v = 0; v < V; ++v)
                        dg[v] = (dkeep[pp][v] & CODE_IDENTITY) ? T(1) : A.cfluid;
                } else {
                    apply7v<0, T, V>(A, d, id, i, j, k, nb, qv);
                    diag7v<0, T, V>(A, id, dg);
                }
#pragma unroll
                for (int v = 0; v < V; ++v) out[v] = nb.c[v] + om * ((bkeep[pp][v] - qv[v]) / dg[v]);
                vstore_s<T, V>(x_out + id, out);
            }
        }
#pragma unroll
        for (int pp = 0; pp < P; ++pp)
#pragma unroll
            for (int v = 0; v < V; ++v) {
                bkeep[pp][v] = bcur[pp][v];
                dkeep[pp][v] = dcur[pp][v];
                bcur[pp][v] = bnx[pp][v];
                dcur[pp][v] = dnx[pp][v];
                bnx[pp][v] = bnx2[pp][v];
                dnx[pp][v] = dnx2[pp][v];
            }
    }
}

// The two post-smoothing sweeps of a level in one pass (MG_PROLONG, then a
// plain sweep), marching along i with two LDS rings: for plane t+1 the
// corrected iterate x + scale * e of its block goes into ring 0 (the
// workgroup's J rows and two halo rows on each side, once per cell, where
// the separate prolongation sweep forms it again at every neighbour); the
// first sweep runs at plane t from ring 0 into ring 1 (J rows and one halo
// row on each side); the second at plane t-1 from ring 1 for the J rows,
// with b.x_out reduced when RZ.  b and the diagonal's source of the rows
// are loaded with ring 0's plane and carried.  Every cell's arithmetic is
// k_mg_smooth_v's, so x_out is bit-identical to the two sweeps apart; the
// r.z partials are summed in another order.
template <typename T, int CODED, int V, int P, int R, bool RZ>
__global__ __launch_bounds__(kBlock) void k_mg_post_ring(Dims d, PairGeom g, Stencil<T> A,
                                                         const T* __restrict__ x,
                                                         const T* __restrict__ b,
                                                         const T* __restrict__ e, Dims cd,
                                                         T* __restrict__ x_out, double omega,
                                                         double scale, RedScratch red,
                                                         const SolveScalars* sc,
                                                         const double* __restrict__ wgt)
{
    static_assert(V == 2 || V == 4, "the block corrections of V cells come as one or two loads");
    if (is_done(sc)) return;
    extern __shared__ unsigned char les_smem[];
    const int Q = P * g.rpp;  // rows per plane: J + 4
    const int plane_sz = Q * d.nz;
    T* r0 = reinterpret_cast<T*>(les_smem);  // x + scale e: [R][Q][nz]
    T* r1 = r0 + (size_t)R * plane_sz;       // first sweep:  [R][Q][nz]
    const CodeTab<T>* tab = nullptr;
    if constexpr (CODED == 1 && LES_MV_TAB) {
        __shared__ CodeTab<T> tab_s;
        code_tab_fill(tab_s, A);
        __syncthreads();
        tab = &tab_s;
    }
    const int lb = logical_block(g.grid);
    const int jb = lb % g.njb, ic = lb / g.njb;
    const int j0 = jb * g.J;
    const int ia = ic * g.irange, ib = min(ia + g.irange, d.nx);
    const int sub = (int)threadIdx.x / g.tpr;
    const int kk = ((int)threadIdx.x - sub * g.tpr) * V;
    const T om = (T)omega, sf = (T)scale;
    using DS = std::conditional_t<CODED == 1, uint16_t, T>;
    // b and the diagonal's source of this thread's rows at plane t+1 (just
    // loaded), t (the first sweep) and t-1 (the second)
    T bn[P][V], bA[P][V], bB[P][V];
    DS dn[P][V], dA[P][V], dB[P][V];
    double acc[1] = {0.0};
    T pad[V];
#pragma unroll
    for (int v = 0; v < V; ++v) pad[v] = outside<T>();
    // this thread's own row of x + scale e at planes t-1, t, t+1 and of the
    // first sweep at t-2, t-1, t: the i-neighbours come from registers, only
    // the j-neighbours (and k-edges across waves) from the rings
    T xo[P][3][V], yo[P][3][V];
#pragma unroll
    for (int pp = 0; pp < P; ++pp)
#pragma unroll
        for (int v = 0; v < V; ++v) {
            bn[pp][v] = bA[pp][v] = bB[pp][v] = T(0);
            dn[pp][v] = dA[pp][v] = dB[pp][v] = DS(0);
#pragma unroll
            for (int w = 0; w < 3; ++w) xo[pp][w][v] = yo[pp][w][v] = pad[v];
        }
    auto dgv = [&](const DS (&ds)[V], T (&dg)[V]) {
#pragma unroll
        for (int v = 0; v < V; ++v) {
            if constexpr (CODED == 1) dg[v] = (ds[v] & CODE_IDENTITY) ? T(1) : A.cfluid;
            else dg[v] = ds[v];
        }
    };
    // ---- plane u of x + scale * e into ring 0; b, diagonal of plane u ----
    auto stage0 = [&](int u) {
#pragma unroll
        for (int pp = 0; pp < P; ++pp) {
            const int a = pp * g.rpp + sub;
            const int j = j0 - 2 + a;
            T* slot = r0 + (u + R) % R * plane_sz + a * d.nz + kk;
#pragma unroll
            for (int v = 0; v < V; ++v) {
                bn[pp][v] = T(0);
                dn[pp][v] = DS(0);
            }
#pragma unroll
            for (int v = 0; v < V; ++v) xo[pp][2][v] = pad[v];
            if (u < 0 || u >= d.nx || j < 0 || j >= d.ny) {
                if (CODED == 1) vstore<T, V>(slot, pad);
                continue;
            }
            const int id = (u * d.ny + j) * d.nz + kk;
            T xv[V];
            vload<T, V>(x + id, xv);
            const T* er = e + ((u >> 1) * cd.ny + (j >> 1)) * cd.nz + (kk >> 1);
            if constexpr (V == 4) {
                T ev[2];
                vload<T, 2>(er, ev);
#pragma unroll
                for (int v = 0; v < V; ++v) xv[v] = xv[v] + sf * ev[v >> 1];
            } else {
                const T ev = er[0];
#pragma unroll
                for (int v = 0; v < V; ++v) xv[v] = xv[v] + sf * ev;
            }
            vstore<T, V>(slot, xv);
#pragma unroll
            for (int v = 0; v < V; ++v) xo[pp][2][v] = xv[v];
            if (a >= 1 && a <= Q - 2) {
                vload_s<T, V>(b + id, bn[pp]);
                if constexpr (CODED == 1) vload<uint16_t, V>(A.code + id, dn[pp]);
                else vload<T, V>(A.a[SC] + id, dn[pp]);
            }
        }
    };
    auto shift = [&]() {
#pragma unroll
        for (int pp = 0; pp < P; ++pp)
#pragma unroll
            for (int v = 0; v < V; ++v) {
                bB[pp][v] = bA[pp][v];
                dB[pp][v] = dA[pp][v];
                bA[pp][v] = bn[pp][v];
                dA[pp][v] = dn[pp][v];
                xo[pp][0][v] = xo[pp][1][v];
                xo[pp][1][v] = xo[pp][2][v];
                yo[pp][0][v] = yo[pp][1][v];
                yo[pp][1][v] = yo[pp][2][v];
            }
    };
    // one sweep at plane i for row a from ring rg (planes i-1, i, i+1)
    auto sweep = [&](const T* rg, int i, int a, int j, const T (&own)[3][V], const T (&bv)[V],
                     const DS (&ds)[V], T (&out)[V]) {
        const T* pc_pl = rg + (i + R) % R * plane_sz;
        const int k = kk, id = (i * d.ny + j) * d.nz + kk;
        Nbr<T, V> nb;
#pragma unroll
        for (int v = 0; v < V; ++v) {
            nb.w[v] = own[0][v];
            nb.c[v] = own[1][v];
            nb.e[v] = own[2][v];
        }
        if (CODED == 1) {
            vload<T, V>(pc_pl + (a - 1) * d.nz + kk, nb.s);
            vload<T, V>(pc_pl + (a + 1) * d.nz + kk, nb.n);
        } else {
#pragma unroll
            for (int v = 0; v < V; ++v) nb.s[v] = nb.n[v] = T(0);
            if (j > 0) vload<T, V>(pc_pl + (a - 1) * d.nz + kk, nb.s);
            if (j < d.ny - 1) vload<T, V>(pc_pl + (a + 1) * d.nz + kk, nb.n);
        }
        row_edges<T, V, false>(d, id, k, nb.c, nb.km, nb.kp,
                               [&](int n) { return pc_pl[a * d.nz + kk + (n - id)]; });
        T q[V], dg[V];
        if constexpr (CODED == 1)
            apply7v<1, T, V, LES_MV_TAB>(A, d, id, i, j, k, nb, q, ds, nullptr, tab);
        else
            apply7v<0, T, V>(A, d, id, i, j, k, nb, q);
        dgv(ds, dg);
#pragma unroll
        for (int v = 0; v < V; ++v) out[v] = nb.c[v] + om * ((bv[v] - q[v]) / dg[v]);
    };
    if (ia < d.nx) {
        stage0(ia - 2);
        shift();
        stage0(ia - 1);
        shift();
    }
    for (int t = ia - 1; t <= ib && ia < d.nx; ++t) {
        stage0(t + 1);
        __syncthreads();
        // ---- the first sweep at plane t, rows 1 .. Q-2, into ring 1 ----
#pragma unroll
        for (int pp = 0; pp < P; ++pp) {
            const int a = pp * g.rpp + sub;
            const int j = j0 - 2 + a;
#pragma unroll
            for (int v = 0; v < V; ++v) yo[pp][2][v] = pad[v];
            if (a < 1 || a > Q - 2) continue;
            T* slot = r1 + (t + R) % R * plane_sz + a * d.nz + kk;
            if (t < 0 || t >= d.nx || j < 0 || j >= d.ny) {
                if (CODED == 1) vstore<T, V>(slot, pad);
                continue;
            }
            T o[V];
            sweep(r0, t, a, j, xo[pp], bA[pp], dA[pp], o);
            vstore<T, V>(slot, o);
#pragma unroll
            for (int v = 0; v < V; ++v) yo[pp][2][v] = o[v];
        }
        __syncthreads();
        // ---- the second sweep at plane t-1, rows 2 .. Q-3 ----
        const int iB = t - 1;
        if (iB >= ia && iB < ib) {
#pragma unroll
            for (int pp = 0; pp < P; ++pp) {
                const int a = pp * g.rpp + sub;
                const int j = j0 - 2 + a;
                if (a < 2 || a > Q - 3 || j >= d.ny) continue;
                T o[V];
                sweep(r1, iB, a, j, yo[pp], bB[pp], dB[pp], o);
                const int id = (iB * d.ny + j) * d.nz + kk;
                vstore_s<T, V>(x_out + id, o);
                if (RZ) {
#pragma unroll
                    for (int v = 0; v < V; ++v) {
                        const double bz = (double)bB[pp][v] * (double)o[v];
                        double wv = 1.0;
                        if (wgt) {
                            if constexpr (CODED == 1) wv = code_weight(dB[pp][v]);
                            else wv = wgt[id + v];
                        }
                        acc[0] += wgt ? wv * bz : bz;
                    }
                }
            }
        }
        shift();
    }
    if (RZ) block_partial<1>(acc, red.partials, lb);
}

// The restriction bc = P^T (b - A x) (or P^T W (b - A x)) marching along i
// like the ring matvec: x of plane t goes into an LDS ring (J rows and one
// halo row on each side), the residual of plane t-1 is formed for the J
// rows from it, and each coarse cell sums its eight children in (di, dj, dk)
// order: the thread of the even fine row holds the sum over the plane pair,
// adds its own two children of a plane and then its partner row's, which
// come through LDS.  J and the i-runs are even, so a block of 2 x 2 x 2
// children lies in one workgroup.  Bit-identical to k_mg_restrict_v.
template <typename T, int CODED, int V, int P, int R>
__global__ __launch_bounds__(kBlock) void k_mg_restrict_ring(Dims f, Dims cd, PairGeom g,
                                                             Stencil<T> A,
                                                             const T* __restrict__ x,
                                                             const T* __restrict__ b,
                                                             T* __restrict__ bc,
                                                             const SolveScalars* sc,
                                                             const double* __restrict__ W)
{
    static_assert(V == 2 || V == 4, "a thread's cells are whole coarse k-cells");
    constexpr int VC = V / 2;  // coarse cells of a thread's V fine cells
    if (is_done(sc)) return;
    extern __shared__ unsigned char les_smem[];
    const int Q = P * g.rpp;  // J + 2 rows per plane
    const int plane_sz = Q * f.nz;
    T* xs = reinterpret_cast<T*>(les_smem);  // [R][Q][nz]
    T* rs = xs + (size_t)R * plane_sz;       // the odd rows' residuals: [Q][nz]
    const CodeTab<T>* tab = nullptr;
    if constexpr (CODED == 1 && LES_MV_TAB) {
        __shared__ CodeTab<T> tab_s;
        code_tab_fill(tab_s, A);
        __syncthreads();
        tab = &tab_s;
    }
    const int lb = logical_block(g.grid);
    const int jb = lb % g.njb, ic = lb / g.njb;
    const int j0 = jb * g.J;
    const int ia = ic * g.irange, ib = min(ia + g.irange, f.nx);
    const int sub = (int)threadIdx.x / g.tpr;
    const int kk = ((int)threadIdx.x - sub * g.tpr) * V;
    // b and the code of the core rows: plane t (loaded with x) and t-1 (its
    // residual)
    T bcur[P][V], bkeep[P][V];
    uint16_t ccur[P][V], ckeep[P][V];
    T acc[P][VC];
    T pad[V];
#pragma unroll
    for (int v = 0; v < V; ++v) pad[v] = outside<T>();
#pragma unroll
    for (int pp = 0; pp < P; ++pp) {
#pragma unroll
        for (int v = 0; v < V; ++v) {
            bcur[pp][v] = bkeep[pp][v] = T(0);
            ccur[pp][v] = ckeep[pp][v] = 0;
        }
#pragma unroll
        for (int u = 0; u < VC; ++u) acc[pp][u] = T(0);
    }
    for (int t = ia - 1; t <= ib && ia < f.nx; ++t) {
        // ---- x of plane t into the ring; b and the code of its core rows ----
#pragma unroll
        for (int pp = 0; pp < P; ++pp) {
            const int qr = pp * g.rpp + sub;
            const int a = qr < g.J ? qr + 1 : (qr == g.J ? 0 : g.J + 1);
            const int j = j0 - 1 + a;
            T* slot = xs + (t + R) % R * plane_sz + a * f.nz + kk;
            if (t < 0 || t >= f.nx || j < 0 || j >= f.ny) {
                if (CODED == 1) vstore<T, V>(slot, pad);
                continue;
            }
            const int id = (t * f.ny + j) * f.nz + kk;
            T xv[V];
            vload<T, V>(x + id, xv);
            vstore<T, V>(slot, xv);
            if (qr < g.J && t >= ia && t < ib) {
                vload_s<T, V>(b + id, bcur[pp]);
                if (CODED == 1) vload<uint16_t, V>(A.code + id, ccur[pp]);
            }
        }
        __syncthreads();
        // ---- residuals of plane t-1; the odd rows' into LDS ----
        const int iB = t - 1;
        const bool act = iB >= ia && iB < ib;
        T res[P][V];
        if (act) {
            const T* pc_pl = xs + (iB + R) % R * plane_sz;
            const T* pm_pl = xs + (iB + R - 1) % R * plane_sz;
            const T* pp_pl = xs + (iB + 1) % R * plane_sz;
#pragma unroll
            for (int pp = 0; pp < P; ++pp) {
                const int qr = pp * g.rpp + sub;
                const int a = qr + 1, j = j0 + qr;
#pragma unroll
                for (int v = 0; v < V; ++v) res[pp][v] = T(0);
                if (qr >= g.J || j >= f.ny) continue;
                const int i = iB, k = kk, id = (i * f.ny + j) * f.nz + kk;
                Nbr<T, V> nb;
                vload<T, V>(pc_pl + a * f.nz + kk, nb.c);
                if (CODED == 1) {
                    vload<T, V>(pm_pl + a * f.nz + kk, nb.w);
                    vload<T, V>(pp_pl + a * f.nz + kk, nb.e);
                    vload<T, V>(pc_pl + (a - 1) * f.nz + kk, nb.s);
                    vload<T, V>(pc_pl + (a + 1) * f.nz + kk, nb.n);
                } else {
#pragma unroll
                    for (int v = 0; v < V; ++v) nb.w[v] = nb.e[v] = nb.s[v] = nb.n[v] = T(0);
                    if (i > 0) vload<T, V>(pm_pl + a * f.nz + kk, nb.w);
                    if (i < f.nx - 1) vload<T, V>(pp_pl + a * f.nz + kk, nb.e);
                    if (j > 0) vload<T, V>(pc_pl + (a - 1) * f.nz + kk, nb.s);
                    if (j < f.ny - 1) vload<T, V>(pc_pl + (a + 1) * f.nz + kk, nb.n);
                }
                row_edges<T, V, false>(f, id, k, nb.c, nb.km, nb.kp,
                                       [&](int n) { return pc_pl[a * f.nz + kk + (n - id)]; });
                T q[V];
                if constexpr (CODED == 1)
                    apply7v<1, T, V, LES_MV_TAB>(A, f, id, i, j, k, nb, q, ckeep[pp], nullptr, tab);
                else
                    apply7v<0, T, V>(A, f, id, i, j, k, nb, q);
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    const T r = bkeep[pp][v] - q[v];
                    double wv = 1.0;
                    if (W) {
                        if constexpr (CODED == 1) wv = code_weight(ckeep[pp][v]);
                        else wv = W[id + v];
                    }
                    res[pp][v] = W ? (T)wv * r : r;
                }
                if (qr & 1) vstore<T, V>(rs + qr * f.nz + kk, res[pp]);
            }
        }
        __syncthreads();
        // ---- the even rows sum their children and their partner's ----
        if (act) {
            const bool last = (iB & 1) || iB + 1 >= f.nx;  // the block's last fine plane
#pragma unroll
            for (int pp = 0; pp < P; ++pp) {
                const int qr = pp * g.rpp + sub;
                const int j = j0 + qr;
                if (qr >= g.J || (qr & 1) || j >= f.ny) continue;
                if ((iB & 1) == 0) {
#pragma unroll
                    for (int u = 0; u < VC; ++u) acc[pp][u] = T(0);
                }
#pragma unroll
                for (int v = 0; v < V; ++v) acc[pp][v >> 1] += res[pp][v];
                if (j + 1 < f.ny) {
                    T rp[V];
                    vload<T, V>(rs + (qr + 1) * f.nz + kk, rp);
#pragma unroll
                    for (int v = 0; v < V; ++v) acc[pp][v >> 1] += rp[v];
                }
                if (last) {
                    const int cid = ((iB >> 1) * cd.ny + (j >> 1)) * cd.nz + (kk >> 1);
                    vstore<T, VC>(bc + cid, acc[pp]);
                }
            }
        }
#pragma unroll
        for (int pp = 0; pp < P; ++pp)
#pragma unroll
            for (int v = 0; v < V; ++v) {
                bkeep[pp][v] = bcur[pp][v];
                ckeep[pp][v] = ccur[pp][v];
            }
    }
}

// The pre-smoothing and the restriction of an FP32 level in one pass,
// marching along i with three LDS buffers: the first sweep's iterate
// om * (b / A_c) of plane t into ring 1 (the J rows and two halo rows on
// each side), the second sweep at plane t-1 from it into ring 2 (and x_out
// for the J rows), the residual b - A x of plane t-2 from ring 2 for the J
// rows, whose 2 x 2 x 2 blocks the even rows sum in (di, dj, dk) order with
// their partner rows' residuals from the third buffer.  b and the code (or
// A_c) of the rows are loaded a plane ahead and carried.  J and the i-runs
// are even.  Bit-identical to k_mg_zs_ring followed by k_mg_restrict_v.
template <typename T, int CODED, int V, int R>
__global__ __launch_bounds__(kBlock) void k_mg_pre_ring(Dims f, Dims cd, PairGeom g,
                                                        Stencil<T> A, const T* __restrict__ b,
                                                        T* __restrict__ x_out,
                                                        T* __restrict__ bc, double omega,
                                                        const SolveScalars* sc,
                                                        const double* __restrict__ W)
{
    static_assert(V == 4, "a thread's cells are two whole coarse k-cells");
    constexpr int VC = V / 2;
    if (is_done(sc)) return;
    extern __shared__ unsigned char les_smem[];
    const int Q = g.rpp;  // J + 4 rows per plane
    const int plane_sz = Q * f.nz;
    T* r1 = reinterpret_cast<T*>(les_smem);  // first sweep:  [R][Q][nz]
    T* r2 = r1 + (size_t)R * plane_sz;       // second sweep: [R][Q][nz]
    T* rs = r2 + (size_t)R * plane_sz;       // the odd rows' residuals: [Q][nz]
    const CodeTab<T>* tab = nullptr;
    if constexpr (CODED == 1 && LES_MV_TAB) {
        __shared__ CodeTab<T> tab_s;
        code_tab_fill(tab_s, A);
        __syncthreads();
        tab = &tab_s;
    }
    const int lb = logical_block(g.grid);
    const int jb = lb % g.njb, ic = lb / g.njb;
    const int j0 = jb * g.J;
    const int ia = ic * g.irange, ib = min(ia + g.irange, f.nx);
    const int a = (int)threadIdx.x / g.tpr;  // this thread's row: j0 - 2 + a
    const int kk = ((int)threadIdx.x - a * g.tpr) * V;
    const int j = j0 - 2 + a;
    const bool jin = j >= 0 && j < f.ny;
    const int c = a - 2;  // core row index (0 .. J-1 for the J rows)
    const T om = (T)omega;
    using DS = std::conditional_t<CODED == 1, uint16_t, T>;
    // b and the diagonal's source of this row: plane t+1 (loaded), t (first
    // sweep), t-1 (second), t-2 (residual)
    T bn[V], bA[V], bB[V], bC[V];
    DS dn[V], dA[V], dB[V], dC[V];
    T acc[VC];
    T pad[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
        pad[v] = outside<T>();
        bn[v] = bA[v] = bB[v] = bC[v] = T(0);
        dn[v] = dA[v] = dB[v] = dC[v] = DS(0);
    }
#pragma unroll
    for (int u = 0; u < VC; ++u) acc[u] = T(0);
    auto in_plane = [&](int i) { return i >= 0 && i < f.nx; };
    auto load = [&](int i) {
#pragma unroll
        for (int v = 0; v < V; ++v) {
            bn[v] = T(0);
            dn[v] = DS(0);
        }
        if (!jin || !in_plane(i)) return;
        const int id = (i * f.ny + j) * f.nz + kk;
        vload_s<T, V>(b + id, bn);
        if constexpr (CODED == 1) vload<uint16_t, V>(A.code + id, dn);
        else vload<T, V>(A.a[SC] + id, dn);
    };
    auto shift = [&]() {
#pragma unroll
        for (int v = 0; v < V; ++v) {
            bC[v] = bB[v];
            dC[v] = dB[v];
            bB[v] = bA[v];
            dB[v] = dA[v];
            bA[v] = bn[v];
            dA[v] = dn[v];
        }
    };
    auto dgv = [&](const DS (&ds)[V], T (&dg)[V]) {
#pragma unroll
        for (int v = 0; v < V; ++v) {
            if constexpr (CODED == 1) dg[v] = (ds[v] & CODE_IDENTITY) ? T(1) : A.cfluid;
            else dg[v] = ds[v];
        }
    };
    // A y at plane i of this row from ring rg (planes i-1, i, i+1)
    auto apply_ring = [&](const T* rg, int i, const DS (&ds)[V], Nbr<T, V>& nb, T (&q)[V]) {
        const T* pc_pl = rg + (i + R) % R * plane_sz;
        const T* pm_pl = rg + (i + R - 1) % R * plane_sz;
        const T* pp_pl = rg + (i + 1) % R * plane_sz;
        const int id = (i * f.ny + j) * f.nz + kk;
        vload<T, V>(pc_pl + a * f.nz + kk, nb.c);
        if (CODED == 1) {
            vload<T, V>(pm_pl + a * f.nz + kk, nb.w);
            vload<T, V>(pp_pl + a * f.nz + kk, nb.e);
            vload<T, V>(pc_pl + (a - 1) * f.nz + kk, nb.s);
            vload<T, V>(pc_pl + (a + 1) * f.nz + kk, nb.n);
        } else {
#pragma unroll
            for (int v = 0; v < V; ++v) nb.w[v] = nb.e[v] = nb.s[v] = nb.n[v] = T(0);
            if (i > 0) vload<T, V>(pm_pl + a * f.nz + kk, nb.w);
            if (i < f.nx - 1) vload<T, V>(pp_pl + a * f.nz + kk, nb.e);
            if (j > 0) vload<T, V>(pc_pl + (a - 1) * f.nz + kk, nb.s);
            if (j < f.ny - 1) vload<T, V>(pc_pl + (a + 1) * f.nz + kk, nb.n);
        }
        row_edges<T, V, false>(f, id, kk, nb.c, nb.km, nb.kp,
                               [&](int n) { return pc_pl[a * f.nz + kk + (n - id)]; });
        if constexpr (CODED == 1)
            apply7v<1, T, V, LES_MV_TAB>(A, f, id, i, j, kk, nb, q, ds, nullptr, tab);
        else
            apply7v<0, T, V>(A, f, id, i, j, kk, nb, q);
    };
    if (ia < f.nx) {
        load(ia - 2);
        shift();
    }
    for (int t = ia - 2; t <= ib + 1 && ia < f.nx; ++t) {
        load(t + 1);
        // ---- the first sweep at plane t, all rows, into ring 1 ----
        {
            T* slot = r1 + (t + R) % R * plane_sz + a * f.nz + kk;
            if (jin && in_plane(t)) {
                T dg[V], xv[V];
                dgv(dA, dg);
#pragma unroll
                for (int v = 0; v < V; ++v) xv[v] = om * (bA[v] / dg[v]);
                vstore<T, V>(slot, xv);
            } else if (CODED == 1) {
                vstore<T, V>(slot, pad);
            }
        }
        __syncthreads();
        // ---- the second sweep at plane t-1, rows 1 .. Q-2, into ring 2 ----
        const int iB = t - 1;
        if (a >= 1 && a <= Q - 2 && iB >= ia - 1 && iB <= ib) {
            T* slot = r2 + (iB + R) % R * plane_sz + a * f.nz + kk;
            if (jin && in_plane(iB)) {
                Nbr<T, V> nb;
                T q[V], dg[V], o[V];
                apply_ring(r1, iB, dB, nb, q);
                dgv(dB, dg);
#pragma unroll
                for (int v = 0; v < V; ++v) o[v] = nb.c[v] + om * ((bB[v] - q[v]) / dg[v]);
                vstore<T, V>(slot, o);
                if (c >= 0 && c < g.J && iB >= ia && iB < ib)
                    vstore_s<T, V>(x_out + (iB * f.ny + j) * f.nz + kk, o);
            } else if (CODED == 1) {
                vstore<T, V>(slot, pad);
            }
        }
        __syncthreads();
        // ---- the residual at plane t-2 for the J rows; the odd rows' into LDS ----
        const int iC = t - 2;
        const bool act = iC >= ia && iC < ib;
        const bool live = act && c >= 0 && c < g.J && jin;
        T res[V];
#pragma unroll
        for (int v = 0; v < V; ++v) res[v] = T(0);
        if (live) {
            Nbr<T, V> nb;
            T q[V];
            apply_ring(r2, iC, dC, nb, q);
            const int id = (iC * f.ny + j) * f.nz + kk;
#pragma unroll
            for (int v = 0; v < V; ++v) {
                const T r = bC[v] - q[v];
                double wv = 1.0;
                if (W) {
                    if constexpr (CODED == 1) wv = code_weight(dC[v]);
                    else wv = W[id + v];
                }
                res[v] = W ? (T)wv * r : r;
            }
            if (c & 1) vstore<T, V>(rs + a * f.nz + kk, res);
        }
        __syncthreads();
        // ---- the even rows sum their children and their partner's ----
        if (live && !(c & 1)) {
            if ((iC & 1) == 0) {
#pragma unroll
                for (int u = 0; u < VC; ++u) acc[u] = T(0);
            }
#pragma unroll
            for (int v = 0; v < V; ++v) acc[v >> 1] += res[v];
            if (j + 1 < f.ny) {
                T rp[V];
                vload<T, V>(rs + (a + 1) * f.nz + kk, rp);
#pragma unroll
                for (int v = 0; v < V; ++v) acc[v >> 1] += rp[v];
            }
            if ((iC & 1) || iC + 1 >= f.nx)
                vstore<T, VC>(bc + ((iC >> 1) * cd.ny + (j >> 1)) * cd.nz + (kk >> 1), acc);
        }
        shift();
    }
}

// k_mg_restrict with two coarse cells of a row per thread: the four fine
// rows of their blocks are read four cells at a time (k-neighbours by lane
// shifts), the residual formed per cell as apply7 forms it, and each coarse
// cell sums its children in (di, dj, dk) order.  Fine nz % 4 == 0 (so the
// coarse nz is even).
template <typename T, int CODED>
__global__ __launch_bounds__(kBlock) void k_mg_restrict_v(Dims f, Dims cd, Chunking ch,
                                                          Stencil<T> A, const T* __restrict__ x,
                                                          const T* __restrict__ b,
                                                          T* __restrict__ bc,
                                                          const SolveScalars* sc,
                                                          const double* __restrict__ W)
{
    if (is_done(sc)) return;
    const int lb = logical_block(ch.grid);
    FOR_VCELLS(cd, ch, lb, 2)
    {
        T acc[2] = {T(0), T(0)};
        const int fk = 2 * k;
        for (int di = 0; di < 2; ++di)
            for (int dj = 0; dj < 2; ++dj) {
                const int fi = 2 * i + di, fj = 2 * j + dj;
                // a missing fine row contributes nothing, but its lanes still
                // take part in the k-edge shifts (row_edges reads neighbours'
                // lanes); only the sums skip it
                const bool live = fi < f.nx && fj < f.ny;
                const int ci = live ? fi : 2 * i, cj = live ? fj : 2 * j;
                const int fid = (ci * f.ny + cj) * f.nz + fk;
                Nbr<T, 4> nb;
                load_nbr<T, 4>(x, f, fid, ci, cj, fk, [](T v) { return v; }, nb);
                T q[4], bv[4];
                apply7v<CODED, T, 4>(A, f, fid, ci, cj, fk, nb, q);
                vload<T, 4>(b + fid, bv);
                if (!live) continue;
                uint16_t cw[4];
                if (CODED == 1 && W) vload<uint16_t, 4>(A.code + fid, cw);
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const T res = bv[v] - q[v];
                    // the coded operator's weights from its code
                    const double wv = !W ? 1.0 : (CODED == 1 ? code_weight(cw[v]) : W[fid + v]);
                    acc[v >> 1] += W ? (T)wv * res : res;
                }
            }
        T o[2] = {acc[0], acc[1]};
        vstore<T, 2>(bc + id, o);
    }
}

// The bottom of the V-cycle in one workgroup: levels l0 .. L-1 (each with
// at most kMgTailCells = kBlock cells, so one cell per thread; l0 >= 1,
// arrays format), their iterates, right-hand sides and coefficients in LDS.
// b of level l0 comes from global memory (the restriction from level
// l0 - 1), x of level l0 goes back (for its prolongation).  Every phase is
// one pass over the level's cells with a barrier after it; each cell's
// arithmetic is k_mg_smooth's and k_mg_restrict's, in their order.  Saves
// the 5 launches per level that the small levels spend mostly waiting.
template <typename T>
struct MgTail {
    int nlev;                        // levels in the tail (the last is the coarsest)
    Dims d[kMgTailLevels];
    Arr7<T> A[kMgTailLevels];
    int off[kMgTailLevels];          // LDS offset (elements) of level l: x, t, b, 7 coefficients
    const T* b0;
    T* x0;
    int sweeps, coarse_sweeps;
    double omega, scale;
};

template <typename T>
__global__ __launch_bounds__(kBlock) void k_mg_tail(MgTail<T> g, const SolveScalars* sc)
{
    if (is_done(sc)) return;
    extern __shared__ unsigned char les_mg_smem[];
    T* lds = reinterpret_cast<T*>(les_mg_smem);
    const T om = (T)g.omega, sf = (T)g.scale;
    const int c = threadIdx.x;  // this thread's cell on every level
    auto X = [&](int l) { return lds + g.off[l]; };
    auto Tb = [&](int l) { return lds + g.off[l] + g.d[l].N; };
    auto B = [&](int l) { return lds + g.off[l] + 2 * g.d[l].N; };
    auto Ac = [&](int l, int q) { return lds + g.off[l] + (3 + q) * g.d[l].N; };
    auto ijk = [&](const Dims& d, int n, int& i, int& j, int& k) {
        i = n / d.sij;
        j = (n - i * d.sij) / d.nz;
        k = n - i * d.sij - j * d.nz;
    };
    // A x at cell n of level l (apply7's order and guards)
    auto ax = [&](int l, const T* x, int n, int i, int j, int k) {
        const Dims& d = g.d[l];
        T q = Ac(l, SC)[n] * x[n];
        if (i > 0) q += Ac(l, SW)[n] * x[n - d.sij];
        if (i < d.nx - 1) q += Ac(l, SE)[n] * x[n + d.sij];
        if (j > 0) q += Ac(l, SS)[n] * x[n - d.nz];
        if (j < d.ny - 1) q += Ac(l, SN)[n] * x[n + d.nz];
        if (k > 0) q += Ac(l, SB)[n] * x[n - 1];
        if (k < d.nz - 1) q += Ac(l, ST)[n] * x[n + 1];
        return q;
    };
    const int L = g.nlev;
    for (int l = 0; l < L; ++l)
        if (c < g.d[l].N)
            for (int q = 0; q < 7; ++q) Ac(l, q)[c] = g.A[l].a[q][c];
    if (c < g.d[0].N) B(0)[c] = g.b0[c];
    __syncthreads();
    // one sweep of level l: mode 0 from zero into x; 1 plain; 2 with the
    // correction from level l+1 (x corrected in place first, then a plain
    // sweep, as the separate prolongation leaves it); the result is in X(l)
    auto sweep = [&](int l, int mode) {
        const Dims& d = g.d[l];
        const bool live = c < d.N;
        int i = 0, j = 0, k = 0;
        if (live) ijk(d, c, i, j, k);
        T* x = X(l);
        if (mode == 0) {
            if (live) x[c] = om * (B(l)[c] / Ac(l, SC)[c]);
            __syncthreads();
            return;
        }
        if (mode == 2) {
            const Dims& cd = g.d[l + 1];
            if (live) x[c] = x[c] + sf * X(l + 1)[((i >> 1) * cd.ny + (j >> 1)) * cd.nz + (k >> 1)];
            __syncthreads();
        }
        if (live) Tb(l)[c] = x[c] + om * ((B(l)[c] - ax(l, x, c, i, j, k)) / Ac(l, SC)[c]);
        __syncthreads();
        if (live) x[c] = Tb(l)[c];
        __syncthreads();
    };
    for (int l = 0; l < L - 1; ++l) {
        sweep(l, 0);
        for (int s2 = 1; s2 < g.sweeps; ++s2) sweep(l, 1);
        // restriction into level l+1's b, children in (di, dj, dk) order
        const Dims& f = g.d[l];
        const Dims& cd = g.d[l + 1];
        if (c < cd.N) {
            int I, J, K;
            ijk(cd, c, I, J, K);
            T acc = T(0);
            for (int di = 0; di < 2; ++di)
                for (int dj = 0; dj < 2; ++dj)
                    for (int dk = 0; dk < 2; ++dk) {
                        const int fi = 2 * I + di, fj = 2 * J + dj, fk = 2 * K + dk;
                        if (fi >= f.nx || fj >= f.ny || fk >= f.nz) continue;
                        const int fid = (fi * f.ny + fj) * f.nz + fk;
                        acc += B(l)[fid] - ax(l, X(l), fid, fi, fj, fk);
                    }
            B(l + 1)[c] = acc;
        }
        __syncthreads();
    }
    sweep(L - 1, 0);
    for (int s2 = 1; s2 < g.coarse_sweeps; ++s2) sweep(L - 1, 1);
    for (int l = L - 2; l >= 0; --l) {
        sweep(l, 2);
        for (int s2 = 1; s2 < g.sweeps; ++s2) sweep(l, 1);
    }
    if (c < g.d[0].N) g.x0[c] = X(0)[c];
}

// The coarsest level: `sweeps` damped-Jacobi sweeps from x = 0 in one
// workgroup, the iterate in LDS (at most kBlock cells)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mg_coarsest(Dims d, Stencil<T> A,
                                                        const T* __restrict__ b,
                                                        T* __restrict__ x, int sweeps,
                                                        double omega, const SolveScalars* sc)
{
    if (is_done(sc)) return;
    __shared__ T xs[2][kBlock];
    const int id = threadIdx.x;
    const bool live = id < d.N;
    const T om = (T)omega;
    int i = 0, j = 0, k = 0;
    T c = T(1), w = T(0), ea = T(0), s = T(0), n = T(0), bb = T(0), t = T(0), bv = T(0);
    if (live) {
        i = id / d.sij;
        j = (id - i * d.sij) / d.nz;
        k = id - i * d.sij - j * d.nz;
        coefs7<0>(A, id, c, w, ea, s, n, bb, t);
        bv = b[id];
    }
    int cur = 0;
    for (int sw = 0; sw < sweeps; ++sw) {
        T out = T(0);
        if (live) {
            if (sw == 0) {
                out = om * (bv / c);
            } else {
                const T* xp = xs[cur];
                const T xc = xp[id];
                T q = c * xc;
                if (i > 0) q += w * xp[id - d.sij];
                if (i < d.nx - 1) q += ea * xp[id + d.sij];
                if (j > 0) q += s * xp[id - d.nz];
                if (j < d.ny - 1) q += n * xp[id + d.nz];
                if (k > 0) q += bb * xp[id - 1];
                if (k < d.nz - 1) q += t * xp[id + 1];
                out = xc + om * ((bv - q) / c);
            }
        }
        const int nxt = sw == 0 ? 0 : cur ^ 1;
        xs[nxt][id] = out;
        cur = nxt;
        __syncthreads();
    }
    if (live) x[id] = sweeps > 0 ? xs[cur][id] : T(0);
}

}  // namespace

// ------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------
namespace {
thread_local int g_partials_grid = 0;
}

int take_partials_grid()
{
    const int g = g_partials_grid;
    g_partials_grid = 0;
    return g;
}

hipError_t launch_finalize(hipStream_t s, int grid, int nr, bool max, int epi, int gate,
                           RedScratch red, SolveScalars* sc, double* out)
{
    if (epi == EPI_PARTIALS) {
        g_partials_grid = grid;
        return hipSuccess;
    }
#define LES_FIN(NR, MX)                                                                         \
    hipLaunchKernelGGL((k_finalize<NR, MX>), dim3(1), dim3(kBlock), 0, s, grid, red.partials,  \
                       epi, gate, sc, out)
    if (max) {
        if (nr == 1) LES_FIN(1, true);
        else LES_FIN(2, true);
    } else {
        if (nr == 1) LES_FIN(1, false);
        else LES_FIN(2, false);
    }
#undef LES_FIN
    return hipGetLastError();
}

hipError_t launch_build_stencil(hipStream_t s, const Dims& d, const uint8_t* solid,
                                double* const A[7], double idx2, double idy2, double idz2)
{
    const Chunking ch = make_chunking(d);
    hipLaunchKernelGGL(k_build_stencil, dim3(ch.grid), dim3(kBlock), 0, s, d, ch, solid, A[SC],
                       A[SN], A[SS], A[SE], A[SW], A[ST], A[SB], idx2, idy2, idz2);
    return hipGetLastError();
}

hipError_t launch_build_code(hipStream_t s, const Dims& d, const uint8_t* solid, uint16_t* code)
{
    const Chunking ch = make_chunking(d);
    hipLaunchKernelGGL(k_build_code, dim3(ch.grid), dim3(kBlock), 0, s, d, ch, solid, code);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_matvec(hipStream_t s, const Dims& d, Stencil<T> A, const T* x, T* y)
{
    const Chunking ch = make_chunking(d);
    if (A.code)
        hipLaunchKernelGGL((k_matvec<T, 1>), dim3(ch.grid), dim3(kBlock), 0, s, d, ch, A, x, y);
    else
        hipLaunchKernelGGL((k_matvec<T, 0>), dim3(ch.grid), dim3(kBlock), 0, s, d, ch, A, x, y);
    return hipGetLastError();
}

template <typename TO, typename TI>
hipError_t launch_convert(hipStream_t s, int n, const TI* in, TO* out)
{
    hipLaunchKernelGGL((k_convert<TO, TI>), dim3(flat_grid(n)), dim3(kBlock), 0, s, n, in, out);
    return hipGetLastError();
}

hipError_t launch_dot(hipStream_t s, const Dims& d, const double* a, const double* b,
                      RedScratch red, double* out)
{
    const Chunking ch = make_chunking(d);
    hipLaunchKernelGGL(k_dot, dim3(ch.grid), dim3(kBlock), 0, s, d, ch, a, b, red);
    if (hipGetLastError() != hipSuccess) return hipErrorLaunchFailure;
    return launch_finalize(s, ch.grid, 1, false, EPI_STORE, GATE_NONE, red, nullptr, out);
}

template <typename T>
hipError_t launch_pcg_init(hipStream_t s, const Dims& d, Stencil<T> A, const T* b, const T* x,
                           T* r, RedScratch red, SolveScalars* sc)
{
    const Chunking ch = make_chunking(d);
    if (A.code)
        hipLaunchKernelGGL((k_pcg_init<T, 1>), dim3(ch.grid), dim3(kBlock), 0, s, d, ch, A, b, x,
                           r, red);
    else
        hipLaunchKernelGGL((k_pcg_init<T, 0>), dim3(ch.grid), dim3(kBlock), 0, s, d, ch, A, b, x,
                           r, red);
    if (hipGetLastError() != hipSuccess) return hipErrorLaunchFailure;
    return launch_finalize(s, ch.grid, 2, false, EPI_INIT, GATE_NONE, red, sc, nullptr);
}

hipError_t launch_mirror_weights(hipStream_t s, const Dims& d, const uint16_t* code, double* W)
{
    const Chunking ch = make_chunking(d);
    hipLaunchKernelGGL(k_mirror_weights, dim3(ch.grid), dim3(kBlock), 0, s, d, ch, code, W);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_pcg_matvec(hipStream_t s, const Dims& d, Stencil<T> A, const T* p, T* q,
                             RedScratch red, SolveScalars* sc, const double* wgt)
{
    return launch_pcg_matvec_pu<T>(s, d, A, nullptr, nullptr, nullptr, const_cast<T*>(p), q, red,
                                   sc, wgt, nullptr);
}

// The ring form of the matvec (k_pcg_matvec_ring): passes of rows per
// workgroup (FP32 rows fill a wave at nz = 256, FP64 rows two), ring depth.
#ifndef LES_MATVEC_RING
#define LES_MATVEC_RING 1
#endif
template <typename T>
struct MvRingCfg {
    static constexpr int P = sizeof(T) == 4 ? 1 : 2;
    static constexpr int R = 4;
};

template <typename T>
size_t mv_ring_lds(const Dims& d, const PairGeom& g)
{
    return (size_t)MvRingCfg<T>::R * MvRingCfg<T>::P * g.rpp * d.nz * sizeof(T);
}

// Workgroups: njb j-blocks times i-runs, at most target_grid in all (a
// multiple of kXcd; kMaxGrid for the kernels that leave partials).
template <typename T>
bool mv_ring_geom(const Dims& d, PairGeom& g, int target_grid = kMaxGrid)
{
    constexpr int V = vec_width<T>(), P = MvRingCfg<T>::P;
    if (d.nz % V) return false;
    g.tpr = d.nz / V;
    if (g.tpr > kBlock || kBlock % g.tpr) return false;
    g.rpp = kBlock / g.tpr;
    g.J = P * g.rpp - 2;
    if (g.J < 2) return false;
    if (mv_ring_lds<T>(d, g) + (LES_MV_TAB ? sizeof(CodeTab<T>) : 0) > kPairMaxLds) return false;
    g.njb = (d.ny + g.J - 1) / g.J;
    if (g.njb > kMaxGrid) return false;
    int nxc = std::max(1, std::min(d.nx, target_grid / g.njb));
    g.irange = (d.nx + nxc - 1) / nxc;
    nxc = (d.nx + g.irange - 1) / g.irange;
    g.grid = ((g.njb * nxc + kXcd - 1) / kXcd) * kXcd;  // <= target_grid, a multiple of kXcd
    return true;
}

template <typename T>
hipError_t launch_pcg_matvec_pu(hipStream_t s, const Dims& d, Stencil<T> A, const T* p_old,
                                const T* z, const T* diag, T* p, T* q, RedScratch red,
                                SolveScalars* sc, const double* wgt, T* x, int* grid_out,
                                RedScratch rz_red, int rz_grid)
{
    if (rz_grid > 0 && (rz_red.partials == red.partials || p_old == nullptr))
        return hipErrorInvalidValue;
    constexpr int W = vec_width<T>();
    const int V = pick_width<T>(d);
    const int pu = p_old == nullptr ? 0 : (diag ? 2 : 1);
    PairGeom g;
    if (LES_MATVEC_RING && pu != 2 && V == W && mv_ring_geom<T>(d, g)) {
        constexpr int P = MvRingCfg<T>::P, R = MvRingCfg<T>::R;
        const size_t lds = mv_ring_lds<T>(d, g);
        const bool cd = A.code != nullptr;
        const void* fn =
            cd ? (pu ? reinterpret_cast<const void*>(&k_pcg_matvec_ring<T, 1, W, 1, P, R>)
                     : reinterpret_cast<const void*>(&k_pcg_matvec_ring<T, 1, W, 0, P, R>))
               : (pu ? reinterpret_cast<const void*>(&k_pcg_matvec_ring<T, 0, W, 1, P, R>)
                     : reinterpret_cast<const void*>(&k_pcg_matvec_ring<T, 0, W, 0, P, R>));
        if (lds > 64 * 1024) {
            hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                               (int)lds);
            if (e != hipSuccess) return e;
        }
#define LES_MVR(CD, PU)                                                                         \
    hipLaunchKernelGGL((k_pcg_matvec_ring<T, CD, W, PU, P, R>), dim3(g.grid), dim3(kBlock), lds, \
                       s, d, g, A, p_old, z, p, q, red, sc, wgt, x, rz_red, rz_grid)
        if (cd) {
            if (pu) LES_MVR(1, 1);
            else LES_MVR(1, 0);
        } else {
            if (pu) LES_MVR(0, 1);
            else LES_MVR(0, 0);
        }
#undef LES_MVR
        if (hipGetLastError() != hipSuccess) return hipErrorLaunchFailure;
        if (grid_out) {
            *grid_out = g.grid;
            return hipSuccess;
        }
        return launch_finalize(s, g.grid, 1, false, EPI_ALPHA, GATE_DONE, red, sc, nullptr);
    }
    const Chunking ch = make_chunking(d, V);
#define LES_MV(CD, VV, PU)                                                                      \
    hipLaunchKernelGGL((k_pcg_matvec<T, CD, VV, PU>), dim3(ch.grid), dim3(kBlock), 0, s, d, ch, \
                       A, p_old, z, diag, p, q, red, sc, wgt, x, rz_red, rz_grid)
#define LES_MV_V(CD, PU)                 \
    do {                                 \
        if (V == W) LES_MV(CD, W, PU);   \
        else LES_MV(CD, 1, PU);          \
    } while (0)
#define LES_MV_P(CD)                     \
    do {                                 \
        if (pu == 0) LES_MV_V(CD, 0);    \
        else if (pu == 1) LES_MV_V(CD, 1); \
        else LES_MV_V(CD, 2);            \
    } while (0)
    if (A.code) LES_MV_P(1);
    else LES_MV_P(0);
#undef LES_MV_P
#undef LES_MV_V
#undef LES_MV
    if (hipGetLastError() != hipSuccess) return hipErrorLaunchFailure;
    if (grid_out) {
        *grid_out = ch.grid;
        return hipSuccess;
    }
    return launch_finalize(s, ch.grid, 1, false, EPI_ALPHA, GATE_DONE, red, sc, nullptr);
}

template <typename T, typename TX>
hipError_t launch_pcg_update(hipStream_t s, const Dims& d, TX* x, T* r, const T* p, const T* q,
                             const T* jacobi_diag, RedScratch red, SolveScalars* sc,
                             const double* wgt)
{
    static_assert(sizeof(T) == sizeof(TX), "x and r share the vector width");
    constexpr int W = vec_width<T>();
    const int V = pick_width<T>(d);
    const Chunking ch = make_chunking(d, V);
    if (V == W)
        hipLaunchKernelGGL((k_pcg_update<T, TX, W>), dim3(ch.grid), dim3(kBlock), 0, s, d, ch, x,
                           r, p, q, jacobi_diag, red, sc, wgt);
    else
        hipLaunchKernelGGL((k_pcg_update<T, TX, 1>), dim3(ch.grid), dim3(kBlock), 0, s, d, ch, x,
                           r, p, q, jacobi_diag, red, sc, wgt);
    if (hipGetLastError() != hipSuccess) return hipErrorLaunchFailure;
    return launch_finalize(s, ch.grid, 2, false, jacobi_diag ? EPI_CONV_RZ : EPI_CONV, GATE_DONE,
                           red, sc, nullptr);
}

template <typename T>
hipError_t launch_pcg_p_jacobi(hipStream_t s, const Dims& d, const T* r, const T* diag, T* p,
                               int init, int epi, RedScratch red, SolveScalars* sc,
                               const double* wgt)
{
    constexpr int W = vec_width<T>();
    const int V = pick_width<T>(d);
    const Chunking ch = make_chunking(d, V);
    RedScratch rd = red;
    if (!(init && epi >= 0)) rd.partials = nullptr;
    if (V == W)
        hipLaunchKernelGGL((k_pcg_p_jacobi<T, W>), dim3(ch.grid), dim3(kBlock), 0, s, d, ch, r,
                           diag, p, init, rd, sc, wgt);
    else
        hipLaunchKernelGGL((k_pcg_p_jacobi<T, 1>), dim3(ch.grid), dim3(kBlock), 0, s, d, ch, r,
                           diag, p, init, rd, sc, wgt);
    if (hipGetLastError() != hipSuccess) return hipErrorLaunchFailure;
    if (init && epi >= 0)
        return launch_finalize(s, ch.grid, 1, false, epi, sc ? GATE_DONE : GATE_NONE, red, sc,
                               nullptr);
    return hipSuccess;
}

template <typename T>
hipError_t launch_pcg_p_update(hipStream_t s, const Dims& d, const T* z, T* p, SolveScalars* sc)
{
    constexpr int W = vec_width<T>();
    const int V = pick_width<T>(d);
    const Chunking ch = make_chunking(d, V);
    if (V == W)
        hipLaunchKernelGGL((k_pcg_p_update<T, W>), dim3(ch.grid), dim3(kBlock), 0, s, d, ch, z, p,
                           sc);
    else
        hipLaunchKernelGGL((k_pcg_p_update<T, 1>), dim3(ch.grid), dim3(kBlock), 0, s, d, ch, z, p,
                           sc);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_pcg_rz(hipStream_t s, const Dims& d, const T* r, const T* z, int epilogue,
                         RedScratch red, SolveScalars* sc, const double* wgt)
{
    constexpr int W = vec_width<T>();
    const int V = pick_width<T>(d);
    const Chunking ch = make_chunking(d, V);
    if (V == W)
        hipLaunchKernelGGL((k_pcg_rz<T, W>), dim3(ch.grid), dim3(kBlock), 0, s, d, ch, r, z, red,
                           sc, wgt);
    else
        hipLaunchKernelGGL((k_pcg_rz<T, 1>), dim3(ch.grid), dim3(kBlock), 0, s, d, ch, r, z, red,
                           sc, wgt);
    if (hipGetLastError() != hipSuccess) return hipErrorLaunchFailure;
    return launch_finalize(s, ch.grid, 1, false, epilogue, GATE_DONE, red, sc, nullptr);
}

template <typename T>
hipError_t launch_cheb_last_st(hipStream_t s, const Dims& d, Stencil<T> A, const T* r,
                               const T* z_in, T* z_out, const T* p, double beta, double alpha,
                               int epi, RedScratch red, SolveScalars* sc);

template <typename T>
hipError_t launch_cheb_step(hipStream_t s, const Dims& d, Stencil<T> A, const T* r,
                            const T* z_in, T* z_out, T* p, double alpha0, double beta,
                            double alpha, int first, int epi, RedScratch red, SolveScalars* sc,
                            const double* wgt)
{
    if (LES_PAIR_STAGED && LES_LAST_STAGED && A.code && sizeof(T) == 4 && !first && epi >= 0 &&
        !wgt) {
        const hipError_t e = launch_cheb_last_st<T>(s, d, A, r, z_in, z_out, p, beta, alpha, epi,
                                                    red, sc);
        if (e != hipErrorNotSupported) return e;
    }
    constexpr int VW = vec_width<T>();
    const bool vec = pick_width<T>(d) == VW;
    const Chunking ch = make_chunking(d, vec ? VW : 1);
#define LES_CHEB(F, CD, RD)                                                                     \
    do {                                                                                        \
        if (vec)                                                                                \
            hipLaunchKernelGGL((k_cheb_step<T, F, CD, RD, VW>), dim3(ch.grid), dim3(kBlock), 0,   \
                               s, d, ch, A, r, z_in, z_out, p, alpha0, beta, alpha, red, sc,    \
                               wgt);                                                            \
        else                                                                                    \
            hipLaunchKernelGGL((k_cheb_step<T, F, CD, RD, 1>), dim3(ch.grid), dim3(kBlock), 0, s, \
                               d, ch, A, r, z_in, z_out, p, alpha0, beta, alpha, red, sc, wgt); \
    } while (0)
#define LES_CHEB_R(F, CD)                   \
    do {                                    \
        if (epi >= 0) LES_CHEB(F, CD, true);  \
        else LES_CHEB(F, CD, false);          \
    } while (0)
    if (first) {
        if (A.code) LES_CHEB_R(true, 1);
        else LES_CHEB_R(true, 0);
    } else {
        if (A.code) LES_CHEB_R(false, 1);
        else LES_CHEB_R(false, 0);
    }
#undef LES_CHEB_R
#undef LES_CHEB
    if (hipGetLastError() != hipSuccess) return hipErrorLaunchFailure;
    if (epi >= 0)
        return launch_finalize(s, ch.grid, 1, false, epi, sc ? GATE_DONE : GATE_NONE, red, sc,
                               nullptr);
    return hipSuccess;
}

// Geometry of the pair kernel per element type and operator storage (CD 1:
// the code, 0: the seven arrays): workgroup size TB, row passes per thread P
// (J = P * rows-per-pass - 2 core rows per workgroup), LDS ring depth R and
// the number of workgroups aimed for -- one round of what the ring and the
// registers allow per CU.  Measured at 256^3 on MI355X:
//   FP32, code:   512 x 2 (J = 14, 64 KiB ring, 2 per CU); 256 x 2, 768 x 2
//                 and 1024 x 2 are 2-8% slower;
//   FP64, code:   1024 x 2 (J = 14, 128 KiB ring, 1 per CU), 12% faster than
//                 512 x 2 and 20% faster than 256 x 3;
//   arrays, both: 768 x 1 (J = 10 FP32, 4 FP64) with the core rows'
//                 coefficients carried through LDS (118 / 104 KiB, 1 per
//                 CU): 15% (FP32) and 20% (FP64) faster than 1024 x 1
//                 re-reading them; 512, 896 and 768 x 2 are slower in one
//                 precision or both.
// Aiming for more workgroups than one round is slower.  Issuing all passes'
// loads before their arithmetic saves a memory round trip per plane but
// costs occupancy, so no configuration does it.
#ifndef LES_PAIR_BLOCK_CODE32
#define LES_PAIR_BLOCK_CODE32 512
#endif
#ifndef LES_PAIR_BLOCK_CODE64
#define LES_PAIR_BLOCK_CODE64 1024
#endif
#ifndef LES_PAIR_PASSES_CODE
#define LES_PAIR_PASSES_CODE 2
#endif
#ifndef LES_PAIR_BLOCK_ARRAYS
#define LES_PAIR_BLOCK_ARRAYS 768
#endif
#ifndef LES_PAIR_PASSES_ARRAYS
#define LES_PAIR_PASSES_ARRAYS 1
#endif
#ifndef LES_PAIR_RING
#define LES_PAIR_RING 4
#endif
#ifndef LES_PAIR_CARRY_COEF64
#define LES_PAIR_CARRY_COEF64 1
#endif
#ifndef LES_PAIR_CARRY_COEF32
#define LES_PAIR_CARRY_COEF32 1
#endif
template <typename T, int CD>
struct PairCfg {
    static constexpr bool F32 = sizeof(T) == 4;
    // arrays format: carry the core rows' coefficients from step A to step
    // B through LDS (one workgroup per CU leaves the room) instead of
    // reading the seven arrays again
    static constexpr bool CC = !CD && (F32 ? LES_PAIR_CARRY_COEF32 : LES_PAIR_CARRY_COEF64);
    static constexpr int TB = CD ? (F32 ? LES_PAIR_BLOCK_CODE32 : LES_PAIR_BLOCK_CODE64)
                                 : LES_PAIR_BLOCK_ARRAYS;
    static constexpr int P = CD ? LES_PAIR_PASSES_CODE : LES_PAIR_PASSES_ARRAYS;
    static constexpr int R = LES_PAIR_RING;
    static constexpr int GRID = CD ? (F32 ? kPairTargetGridCode32 : kPairTargetGridCode64)
                                   : kPairTargetGridArrays;
};

// LDS of one pair workgroup: the ring, plus the carried coefficients (cc)
template <typename T, int CD>
size_t pair_lds(const Dims& d, const PairGeom& g, bool cc)
{
    using C = PairCfg<T, CD>;
    return ((size_t)C::R * C::P * g.rpp + (cc ? 7 * (size_t)g.J : 0)) * d.nz * sizeof(T);
}

template <typename T, int CD>
bool cheb_pair_geom(const Dims& d, PairGeom& g, int target_grid)
{
    using C = PairCfg<T, CD>;
    constexpr int V = vec_width<T>(), P = C::P, TB = C::TB;
    if (d.nz % V) return false;
    g.tpr = d.nz / V;
    if (g.tpr > TB || TB % g.tpr) return false;
    g.rpp = TB / g.tpr;
    g.J = P * g.rpp - 2;
    if (g.J < 2) return false;
    // the coded pair also holds its decode tables in static LDS
    const size_t tab = CD && LES_PAIR_CODE_TAB ? sizeof(CodeTab<T>) : 0;
    if (pair_lds<T, CD>(d, g, false) + tab > kPairMaxLds) return false;
    g.njb = (d.ny + g.J - 1) / g.J;
    if (target_grid <= 0) target_grid = C::GRID;
    int nxc = std::max(1, std::min(d.nx, target_grid / g.njb));
    g.irange = (d.nx + nxc - 1) / nxc;
    nxc = (d.nx + g.irange - 1) / g.irange;
    g.grid = ((g.njb * nxc + kXcd - 1) / kXcd) * kXcd;
    return true;
}

template <typename T>
bool cheb_pair_supported(const Dims& d, int target_grid, bool coded)
{
    PairGeom g;
    return coded ? cheb_pair_geom<T, 1>(d, g, target_grid)
                 : cheb_pair_geom<T, 0>(d, g, target_grid);
}

template <typename T, int CD>
hipError_t launch_cheb_pair_cd(hipStream_t s, const Dims& d, Stencil<T> A, const T* r,
                               const T* z_in, T* z_out, const T* p_in, T* p_out, double alpha0,
                               double beta_a, double alpha_a, double beta_b, double alpha_b,
                               int first, SolveScalars* sc, int target_grid)
{
    using C = PairCfg<T, CD>;
    constexpr int V = vec_width<T>(), P = C::P, TB = C::TB, R = C::R;
    PairGeom g;
    if (!cheb_pair_geom<T, CD>(d, g, target_grid)) return hipErrorNotSupported;
    // carry the coefficients where configured and where they fit
    const bool cc = C::CC && pair_lds<T, CD>(d, g, true) <= kPairMaxLds;
    const size_t lds = pair_lds<T, CD>(d, g, cc);
    auto kernel = [&](auto F, auto CCv) {
        return reinterpret_cast<const void*>(
            &k_cheb_pair<T, decltype(F)::value, CD, V, P, TB, R, decltype(CCv)::value>);
    };
    using t_ = std::true_type;
    using f_ = std::false_type;
    using c_ = std::integral_constant<bool, C::CC>;  // no carrying variant where CC is off
    const void* fn = first ? (cc ? kernel(t_{}, c_{}) : kernel(t_{}, f_{}))
                           : (cc ? kernel(f_{}, c_{}) : kernel(f_{}, f_{}));
    if (lds > 64 * 1024) {
        // rings above the default dynamic LDS limit need the kernel's opt-in
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)lds);
        if (e != hipSuccess) return e;
    }
#define LES_PAIR(F, K)                                                                          \
    hipLaunchKernelGGL((k_cheb_pair<T, F, CD, V, P, TB, R, K>), dim3(g.grid), dim3(TB), lds, s, \
                       d, g, A, r, z_in, z_out, p_in, p_out, alpha0, beta_a, alpha_a, beta_b,   \
                       alpha_b, sc)
    if (first) {
        if (cc) LES_PAIR(true, C::CC);
        else LES_PAIR(true, false);
    } else {
        if (cc) LES_PAIR(false, C::CC);
        else LES_PAIR(false, false);
    }
#undef LES_PAIR
    return hipGetLastError();
}

// The staged pair (k_cheb_pair_st): FP32 with the coded operator, rows of
// one or two waves, one 1024-thread workgroup per CU (LES_PAIR_STAGED and
// LES_LAST_STAGED are defined with the pair's other switches).
#ifndef LES_PAIR_ST_GRID
#define LES_PAIR_ST_GRID 256
#endif
constexpr int kPairStBlock = 1024;

template <typename T>
size_t pair_st_lds(const Dims& d, const PairGeom& g, bool single = false)
{
    return ((single ? 0 : (size_t)LES_PAIR_RING * (g.J + 2)) + 2 * (size_t)(g.J + 4)) * d.nz *
           sizeof(T);
}

template <typename T>
bool cheb_pair_st_geom(const Dims& d, PairGeom& g, int target_grid)
{
    constexpr int V = vec_width<T>();
    if (sizeof(T) != 4 || d.nz % V) return false;
    g.tpr = d.nz / V;
    // rows of one wave take their k-edges by lane shuffle, rows of two
    // waves from the LDS copies; a power of two fills the workgroup
    if (g.tpr > 128 || (g.tpr & (g.tpr - 1))) return false;
    g.rpp = kPairStBlock / g.tpr;
    g.J = g.rpp - 2;
    if (g.J < 2) return false;
    if (pair_st_lds<T>(d, g) + sizeof(CodeTab<T>) > kPairMaxLds) return false;
    g.njb = (d.ny + g.J - 1) / g.J;
    if (target_grid <= 0) target_grid = LES_PAIR_ST_GRID;
    int nxc = std::max(1, std::min(d.nx, target_grid / g.njb));
    g.irange = (d.nx + nxc - 1) / nxc;
    nxc = (d.nx + g.irange - 1) / g.irange;
    g.grid = ((g.njb * nxc + kXcd - 1) / kXcd) * kXcd;
    return true;
}

// The staged kernel for the grid's nz: specialised for nz = 32 .. 512 (row
// strides and LDS offsets fold into the instructions), generic otherwise.
template <typename T, bool FIRST, bool SINGLE, int NZ>
hipError_t run_pair_st_nz(hipStream_t s, const Dims& d, const PairGeom& g, size_t lds,
                          Stencil<T> A, const T* r, const T* z_in, T* z_out, const T* p_in,
                          T* p_out, double alpha0, double beta_a, double alpha_a, double beta_b,
                          double alpha_b, SolveScalars* sc, RedScratch red)
{
    constexpr int V = vec_width<T>(), R = LES_PAIR_RING;
    const void* fn = reinterpret_cast<const void*>(&k_cheb_pair_st<T, FIRST, V, R, SINGLE, NZ>);
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_cheb_pair_st<T, FIRST, V, R, SINGLE, NZ>), dim3(g.grid),
                       dim3(kPairStBlock), lds, s, d, g, A, r, z_in, z_out, p_in, p_out, alpha0,
                       beta_a, alpha_a, beta_b, alpha_b, sc, red);
    return hipGetLastError();
}

template <typename T, bool FIRST, bool SINGLE>
hipError_t run_pair_st(hipStream_t s, const Dims& d, const PairGeom& g, size_t lds,
                       Stencil<T> A, const T* r, const T* z_in, T* z_out, const T* p_in,
                       T* p_out, double alpha0, double beta_a, double alpha_a, double beta_b,
                       double alpha_b, SolveScalars* sc, RedScratch red)
{
    if constexpr (sizeof(T) != 4) {
        return hipErrorNotSupported;  // the staged kernels run FP32 only
    } else {
#define LES_RUN_NZ(NZ)                                                                         \
    return run_pair_st_nz<T, FIRST, SINGLE, NZ>(s, d, g, lds, A, r, z_in, z_out, p_in, p_out, \
                                                alpha0, beta_a, alpha_a, beta_b, alpha_b, sc,  \
                                                red)
        switch (d.nz) {
        case 512: LES_RUN_NZ(512);
        case 256: LES_RUN_NZ(256);
        case 128: LES_RUN_NZ(128);
        case 64: LES_RUN_NZ(64);
        case 32: LES_RUN_NZ(32);
        default: LES_RUN_NZ(0);
        }
#undef LES_RUN_NZ
    }
}

template <typename T>
hipError_t launch_cheb_pair_st(hipStream_t s, const Dims& d, Stencil<T> A, const T* r,
                               const T* z_in, T* z_out, const T* p_in, T* p_out, double alpha0,
                               double beta_a, double alpha_a, double beta_b, double alpha_b,
                               int first, SolveScalars* sc, const PairGeom& g)
{
    const size_t lds = pair_st_lds<T>(d, g);
    if (first)
        return run_pair_st<T, true, false>(s, d, g, lds, A, r, z_in, z_out, p_in, p_out, alpha0,
                                           beta_a, alpha_a, beta_b, alpha_b, sc,
                                           RedScratch{nullptr});
    return run_pair_st<T, false, false>(s, d, g, lds, A, r, z_in, z_out, p_in, p_out, alpha0,
                                        beta_a, alpha_a, beta_b, alpha_b, sc, RedScratch{nullptr});
}

// The reducing last Chebyshev step as the staged kernel's single mode (FP32,
// coded, not the first step, no weights), then its finalize.  Returns
// hipErrorNotSupported where it does not apply.
template <typename T>
hipError_t launch_cheb_last_st(hipStream_t s, const Dims& d, Stencil<T> A, const T* r,
                               const T* z_in, T* z_out, const T* p, double beta, double alpha,
                               int epi, RedScratch red, SolveScalars* sc)
{
    PairGeom g;
    if (!cheb_pair_st_geom<T>(d, g, kMaxGrid) || g.grid > kMaxGrid) return hipErrorNotSupported;
    const size_t lds = pair_st_lds<T>(d, g, true);
    const hipError_t e = run_pair_st<T, false, true>(s, d, g, lds, A, r, z_in, z_out, p, nullptr,
                                                     0.0, beta, alpha, 0.0, 0.0, sc, red);
    if (e != hipSuccess) return e == hipErrorNotSupported ? hipErrorLaunchFailure : e;
    return launch_finalize(s, g.grid, 1, false, epi, sc ? GATE_DONE : GATE_NONE, red, sc,
                           nullptr);
}

template <typename T>
hipError_t launch_cheb_pair(hipStream_t s, const Dims& d, Stencil<T> A, const T* r,
                            const T* z_in, T* z_out, const T* p_in, T* p_out, double alpha0,
                            double beta_a, double alpha_a, double beta_b, double alpha_b,
                            int first, SolveScalars* sc, int target_grid)
{
    if (LES_PAIR_STAGED && A.code && sizeof(T) == 4) {
        PairGeom g;
        if (cheb_pair_st_geom<T>(d, g, target_grid))
            return launch_cheb_pair_st<T>(s, d, A, r, z_in, z_out, p_in, p_out, alpha0, beta_a,
                                          alpha_a, beta_b, alpha_b, first, sc, g);
    }
    if (A.code)
        return launch_cheb_pair_cd<T, 1>(s, d, A, r, z_in, z_out, p_in, p_out, alpha0, beta_a,
                                         alpha_a, beta_b, alpha_b, first, sc, target_grid);
    return launch_cheb_pair_cd<T, 0>(s, d, A, r, z_in, z_out, p_in, p_out, alpha0, beta_a,
                                     alpha_a, beta_b, alpha_b, first, sc, target_grid);
}

template <typename T>
hipError_t launch_cheb_scale(hipStream_t s, const Dims& d, const T* r, T* z, double alpha0,
                             int epi, RedScratch red, SolveScalars* sc, const double* wgt)
{
    const Chunking ch = make_chunking(d);
    RedScratch rd = red;
    if (epi < 0) rd.partials = nullptr;
    hipLaunchKernelGGL(k_cheb_scale<T>, dim3(ch.grid), dim3(kBlock), 0, s, d, ch, r, z, alpha0,
                       rd, sc, wgt);
    if (hipGetLastError() != hipSuccess) return hipErrorLaunchFailure;
    if (epi >= 0)
        return launch_finalize(s, ch.grid, 1, false, epi, sc ? GATE_DONE : GATE_NONE, red, sc,
                               nullptr);
    return hipSuccess;
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

hipError_t launch_mixed_fold(hipStream_t s, const Dims& d, double* x64, float* x32,
                             const float* p0, const float* p1, int force, SolveScalars* sc)
{
    constexpr int W = vec_width<float>();
    const int V = pick_width<float>(d);
    const Chunking ch = make_chunking(d, V);
    const int gate = force ? (int)GATE_NONE : (int)GATE_RELIABLE;
    if (V == W)
        hipLaunchKernelGGL((k_mixed_fold<W>), dim3(ch.grid), dim3(kBlock), 0, s, d, ch, x64, x32, p0,
                           p1, gate, sc);
    else
        hipLaunchKernelGGL((k_mixed_fold<1>), dim3(ch.grid), dim3(kBlock), 0, s, d, ch, x64, x32, p0,
                           p1, gate, sc);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_pcg_x_pending(hipStream_t s, const Dims& d, T* x, const T* p0, const T* p1,
                                SolveScalars* sc)
{
    const Chunking ch = make_chunking(d);
    hipLaunchKernelGGL(k_pcg_x_pending<T>, dim3(ch.grid), dim3(kBlock), 0, s, d, ch, x, p0, p1,
                       sc);
    return hipGetLastError();
}
template hipError_t launch_pcg_x_pending<double>(hipStream_t, const Dims&, double*, const double*,
                                                 const double*, SolveScalars*);

hipError_t launch_mixed_residual(hipStream_t s, const Dims& d, Stencil<double> A, const double* b,
                                 const double* x64, float* r32, int init, int force,
                                 RedScratch red, SolveScalars* sc, double* out)
{
    constexpr int W = vec_width<double>();
    const int V = pick_width<double>(d);
    const Chunking ch = make_chunking(d, V);
    const int gate = (init || force) ? (int)GATE_NONE : (int)GATE_RELIABLE;
    auto pass = [&](float* dst, RedScratch rd) {
#define LES_MRES(CC, VV)                                                                          \
    hipLaunchKernelGGL((k_mixed_residual<CC, VV>), dim3(ch.grid), dim3(kBlock), 0, s, d, ch, A, b, \
                       x64, dst, gate, rd, sc)
        if (A.code) {
            if (V == W) LES_MRES(1, W);
            else LES_MRES(1, 1);
        } else {
            if (V == W) LES_MRES(0, W);
            else LES_MRES(0, 1);
        }
#undef LES_MRES
        return hipGetLastError();
    };
    if (init) {
        // ||b - A x0|| first: it sets the scale r32 carries; then r32
        if (pass(nullptr, red) != hipSuccess) return hipErrorLaunchFailure;
        hipError_t e = launch_finalize(s, ch.grid, 2, false, EPI_INIT_MIXED, gate, red, sc, out);
        if (e != hipSuccess) return e;
        return pass(r32, RedScratch{nullptr});
    }
    if (pass(r32, red) != hipSuccess) return hipErrorLaunchFailure;
    const int epi = force ? EPI_STORE_SQRT : EPI_RELIABLE;
    return launch_finalize(s, ch.grid, 2, false, epi, gate, red, sc, out);
}

hipError_t launch_pcg_update_mixed(hipStream_t s, const Dims& d, float* x, float* r, const float* p,
                                   const float* q, RedScratch red, SolveScalars* sc,
                                   RedScratch pq_red, int pq_grid, RedScratch rz_red, int rz_grid)
{
    constexpr int W = vec_width<float>();
    const int V = pick_width<float>(d);
    const Chunking ch = make_chunking(d, V);
    const bool fold = pq_grid > 0;
    if (fold && pq_red.partials == red.partials) return hipErrorInvalidValue;
    // the partials read here must not be the ones this launch writes
    if (rz_grid > 0 && (!fold || rz_red.partials == red.partials)) return hipErrorInvalidValue;
#define LES_UPM(VV, FF)                                                                           \
    hipLaunchKernelGGL((k_pcg_update_mixed<VV, FF>), dim3(ch.grid), dim3(kBlock), 0, s, d, ch, x, \
                       r, p, q, red, sc, pq_red, pq_grid, rz_red, rz_grid)
    if (V == W) {
        if (fold) LES_UPM(W, true);
        else LES_UPM(W, false);
    } else {
        if (fold) LES_UPM(1, true);
        else LES_UPM(1, false);
    }
#undef LES_UPM
    if (hipGetLastError() != hipSuccess) return hipErrorLaunchFailure;
    return launch_finalize(s, ch.grid, 1, false, EPI_CONV_MIXED, GATE_DONE, red, sc, nullptr);
}

hipError_t launch_matvec_sym(hipStream_t s, const Dims& d, Stencil<double> A, const double* x,
                             double* y)
{
    const Chunking ch = make_chunking(d);
    hipLaunchKernelGGL(k_matvec_sym, dim3(ch.grid), dim3(kBlock), 0, s, d, ch, A, x, y);
    return hipGetLastError();
}

hipError_t launch_lanczos_axpy(hipStream_t s, int n, double* w, const double* v, const double* vp,
                               double a, double b)
{
    hipLaunchKernelGGL(k_lanczos_axpy, dim3(flat_grid(n)), dim3(kBlock), 0, s, n, w, v, vp, a, b);
    return hipGetLastError();
}

hipError_t launch_scale_by_invnorm(hipStream_t s, int n, const double* w, double* v,
                                   const double* nrm2)
{
    hipLaunchKernelGGL(k_scale_by_invnorm, dim3(flat_grid(n)), dim3(kBlock), 0, s, n, w, v, nrm2);
    return hipGetLastError();
}

// explicit instantiations

// ---- multigrid preconditioner ----
hipError_t launch_mg_coarsen(hipStream_t s, const Dims& f, const Dims& cd, Stencil<double> A,
                             const double* W, double* const Ac[7])
{
    const Chunking ch = make_chunking(cd);
    Arr7<double> out;
    for (int q = 0; q < 7; ++q) out.a[q] = Ac[q];
    hipLaunchKernelGGL(k_mg_coarsen, dim3(ch.grid), dim3(kBlock), 0, s, f, cd, ch, A, W, out);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_mg_smooth(hipStream_t s, const Dims& d, Stencil<T> A, const T* x, const T* b,
                            const T* e, const Dims& cd, T* x_out, int mode, int epi,
                            double omega, double scale, RedScratch red, SolveScalars* sc,
                            const double* wgt, bool ring)
{
    constexpr int VW = vec_width<T>();
    const bool vec = LES_MG_VEC && pick_width<T>(d) == VW;
    const Chunking ch = make_chunking(d, vec ? VW : 1);
    const bool rz = epi >= 0;
    PairGeom g;
    if (LES_MG_ZS_RING && ring && mode == MG_ZS && vec && !rz &&
        mv_ring_geom<T>(d, g, LES_MG_ZS_GRID)) {
        constexpr int P = MvRingCfg<T>::P, R = MvRingCfg<T>::R;
        const size_t lds = mv_ring_lds<T>(d, g);
        const void* fn = A.code ? reinterpret_cast<const void*>(&k_mg_zs_ring<T, 1, VW, P, R>)
                                : reinterpret_cast<const void*>(&k_mg_zs_ring<T, 0, VW, P, R>);
        if (lds > 64 * 1024) {
            const hipError_t e = hipFuncSetAttribute(
                fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) return e;
        }
        if (A.code)
            hipLaunchKernelGGL((k_mg_zs_ring<T, 1, VW, P, R>), dim3(g.grid), dim3(kBlock), lds, s,
                               d, g, A, b, x_out, omega, sc);
        else
            hipLaunchKernelGGL((k_mg_zs_ring<T, 0, VW, P, R>), dim3(g.grid), dim3(kBlock), lds, s,
                               d, g, A, b, x_out, omega, sc);
        return hipGetLastError();
    }
    if (mode == MG_ZS && !vec) {
        // the scalar kernels run the two sweeps apart, the first into x
        if (!x || x == x_out) return hipErrorInvalidValue;
        const hipError_t e0 = launch_mg_smooth<T>(s, d, A, nullptr, b, e, cd, const_cast<T*>(x),
                                                  MG_ZERO, -1, omega, scale, red, sc, nullptr);
        if (e0 != hipSuccess) return e0;
        return launch_mg_smooth<T>(s, d, A, x, b, e, cd, x_out, MG_SWEEP, epi, omega, scale, red,
                                   sc, wgt);
    }
#define LES_MGS(CD, MD, RZ)                                                                       \
    do {                                                                                          \
        if (vec)                                                                                  \
            hipLaunchKernelGGL((k_mg_smooth_v<T, CD, MD, RZ, VW>), dim3(ch.grid), dim3(kBlock), 0, \
                               s, d, ch, A, x, b, e, cd, x_out, omega, scale, red, sc, wgt);      \
        else                                                                                      \
            hipLaunchKernelGGL((k_mg_smooth<T, CD, MD, RZ>), dim3(ch.grid), dim3(kBlock), 0, s, d, \
                               ch, A, x, b, e, cd, x_out, omega, scale, red, sc, wgt);            \
    } while (0)
#define LES_MGS_ZS(CD, RZ)                                                                        \
    hipLaunchKernelGGL((k_mg_smooth_v<T, CD, MG_ZS, RZ, VW>), dim3(ch.grid), dim3(kBlock), 0, s, d, \
                       ch, A, x, b, e, cd, x_out, omega, scale, red, sc, wgt)
#define LES_MGS_M(CD, RZ)                             \
    do {                                              \
        if (mode == MG_ZERO) LES_MGS(CD, MG_ZERO, RZ);     \
        else if (mode == MG_SWEEP) LES_MGS(CD, MG_SWEEP, RZ); \
        else if (mode == MG_ZS) LES_MGS_ZS(CD, RZ);       \
        else LES_MGS(CD, MG_PROLONG, RZ);                 \
    } while (0)
    if (A.code) {
        if (rz) LES_MGS_M(1, true);
        else LES_MGS_M(1, false);
    } else {
        if (rz) LES_MGS_M(0, true);
        else LES_MGS_M(0, false);
    }
#undef LES_MGS_M
#undef LES_MGS_ZS
#undef LES_MGS
    if (hipGetLastError() != hipSuccess) return hipErrorLaunchFailure;
    if (rz)
        return launch_finalize(s, ch.grid, 1, false, epi, sc ? GATE_DONE : GATE_NONE, red, sc,
                               nullptr);
    return hipSuccess;
}

// The post-smoothing ring: FP32, rows of V cells, J = rows per plane - 4
// core rows (at least 8), two four-plane rings in LDS; at most kMaxGrid
// workgroups (the r.z partials).
template <typename T>
bool post_ring_geom(const Dims& d, PairGeom& g)
{
    constexpr int V = vec_width<T>();
    if (sizeof(T) != 4 || !LES_MG_POST_RING || !LES_MG_VEC || d.nz % V) return false;
    g.tpr = d.nz / V;
    if (g.tpr > kBlock || kBlock % g.tpr) return false;
    g.rpp = kBlock / g.tpr;
    g.J = g.rpp - 4;
    if (g.J < 8) return false;
    if (2 * (size_t)kMgPostR * g.rpp * d.nz * sizeof(T) + sizeof(CodeTab<T>) > kPairMaxLds)
        return false;
    g.njb = (d.ny + g.J - 1) / g.J;
    if (g.njb > kMaxGrid) return false;
    int nxc = std::max(1, std::min(d.nx, kMaxGrid / g.njb));
    g.irange = (d.nx + nxc - 1) / nxc;
    nxc = (d.nx + g.irange - 1) / g.irange;
    g.grid = ((g.njb * nxc + kXcd - 1) / kXcd) * kXcd;
    return true;
}

template <typename T>
bool mg_post_fits(const Dims& d)
{
    PairGeom g;
    return post_ring_geom<T>(d, g);
}

template <typename T>
hipError_t launch_mg_post(hipStream_t s, const Dims& d, Stencil<T> A, const T* x, const T* b,
                          const T* e, const Dims& cd, T* x_out, int epi, double omega,
                          double scale, RedScratch red, SolveScalars* sc, const double* wgt)
{
    PairGeom g;
    if (!post_ring_geom<T>(d, g)) return hipErrorNotSupported;
    if (x == x_out) return hipErrorInvalidValue;
    if constexpr (sizeof(T) != 4) {
        return hipErrorNotSupported;
    } else {
        constexpr int V = vec_width<T>(), R = kMgPostR;
        const size_t lds = 2 * (size_t)R * g.rpp * d.nz * sizeof(T);
        const bool rz = epi >= 0;
        const bool cdd = A.code != nullptr;
        const void* fn =
            cdd ? (rz ? reinterpret_cast<const void*>(&k_mg_post_ring<T, 1, V, 1, R, true>)
                      : reinterpret_cast<const void*>(&k_mg_post_ring<T, 1, V, 1, R, false>))
                : (rz ? reinterpret_cast<const void*>(&k_mg_post_ring<T, 0, V, 1, R, true>)
                      : reinterpret_cast<const void*>(&k_mg_post_ring<T, 0, V, 1, R, false>));
        if (lds > 64 * 1024) {
            const hipError_t er = hipFuncSetAttribute(
                fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (er != hipSuccess) return er;
        }
#define LES_MGP(CD, RZ)                                                                          \
    hipLaunchKernelGGL((k_mg_post_ring<T, CD, V, 1, R, RZ>), dim3(g.grid), dim3(kBlock), lds, s, \
                       d, g, A, x, b, e, cd, x_out, omega, scale, red, sc, wgt)
        if (cdd) {
            if (rz) LES_MGP(1, true);
            else LES_MGP(1, false);
        } else {
            if (rz) LES_MGP(0, true);
            else LES_MGP(0, false);
        }
#undef LES_MGP
        if (hipGetLastError() != hipSuccess) return hipErrorLaunchFailure;
        if (rz)
            return launch_finalize(s, g.grid, 1, false, epi, sc ? GATE_DONE : GATE_NONE, red, sc,
                                   nullptr);
        return hipSuccess;
    }
}

// The pre-smoothing + restriction ring: FP32, four cells per thread, J =
// rows per plane - 4 core rows (at least 8, even), even i-runs, two rings
// and one plane of residuals in LDS.
template <typename T>
bool pre_ring_geom(const Dims& f, PairGeom& g)
{
    if (sizeof(T) != 4 || !LES_MG_PRE_RING || !LES_MG_VEC || f.nz % 4) return false;
    g.tpr = f.nz / 4;
    if (g.tpr > kBlock || kBlock % g.tpr) return false;
    g.rpp = kBlock / g.tpr;
    g.J = g.rpp - 4;
    if (g.J < 8 || (g.J & 1)) return false;
    if ((size_t)(2 * kMgPostR + 1) * g.rpp * f.nz * sizeof(T) + sizeof(CodeTab<T>) > kPairMaxLds)
        return false;
    g.njb = (f.ny + g.J - 1) / g.J;
    if (g.njb > kMaxGrid) return false;
    int nxc = std::max(1, std::min(f.nx, kMaxGrid / g.njb));
    g.irange = (f.nx + nxc - 1) / nxc;
    g.irange += g.irange & 1;
    nxc = (f.nx + g.irange - 1) / g.irange;
    g.grid = ((g.njb * nxc + kXcd - 1) / kXcd) * kXcd;
    return true;
}

template <typename T>
bool mg_pre_fits(const Dims& f)
{
    PairGeom g;
    return pre_ring_geom<T>(f, g);
}

template <typename T>
hipError_t launch_mg_pre(hipStream_t s, const Dims& f, const Dims& cd, Stencil<T> A, const T* b,
                         T* x_out, T* bc, double omega, SolveScalars* sc, const double* W)
{
    PairGeom g;
    if (!pre_ring_geom<T>(f, g)) return hipErrorNotSupported;
    if constexpr (sizeof(T) != 4) {
        return hipErrorNotSupported;
    } else {
        constexpr int R = kMgPostR;
        const size_t lds = (size_t)(2 * R + 1) * g.rpp * f.nz * sizeof(T);
        const void* fn = A.code ? reinterpret_cast<const void*>(&k_mg_pre_ring<T, 1, 4, R>)
                                : reinterpret_cast<const void*>(&k_mg_pre_ring<T, 0, 4, R>);
        if (lds > 64 * 1024) {
            const hipError_t e = hipFuncSetAttribute(
                fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) return e;
        }
        if (A.code)
            hipLaunchKernelGGL((k_mg_pre_ring<T, 1, 4, R>), dim3(g.grid), dim3(kBlock), lds, s, f,
                               cd, g, A, b, x_out, bc, omega, sc, W);
        else
            hipLaunchKernelGGL((k_mg_pre_ring<T, 0, 4, R>), dim3(g.grid), dim3(kBlock), lds, s, f,
                               cd, g, A, b, x_out, bc, omega, sc, W);
        return hipGetLastError();
    }
}

// The restriction ring: the ring matvec's rows (J = rows per plane - 2 core
// rows, even), even i-runs, the ring and one plane of residuals in LDS.
template <typename T>
bool restrict_ring_geom(const Dims& f, PairGeom& g)
{
    constexpr int V = vec_width<T>(), P = MvRingCfg<T>::P, R = MvRingCfg<T>::R;
    if (!LES_MG_RESTRICT_RING || !LES_MG_VEC || f.nz % V) return false;
    g.tpr = f.nz / V;
    if (g.tpr > kBlock || kBlock % g.tpr) return false;
    g.rpp = kBlock / g.tpr;
    g.J = P * g.rpp - 2;
    if (g.J < 2 || (g.J & 1)) return false;
    if ((size_t)(R + 1) * P * g.rpp * f.nz * sizeof(T) + sizeof(CodeTab<T>) > kPairMaxLds)
        return false;
    g.njb = (f.ny + g.J - 1) / g.J;
    if (g.njb > kMaxGrid) return false;
    int nxc = std::max(1, std::min(f.nx, kMaxGrid / g.njb));
    g.irange = (f.nx + nxc - 1) / nxc;
    g.irange += g.irange & 1;
    nxc = (f.nx + g.irange - 1) / g.irange;
    g.grid = ((g.njb * nxc + kXcd - 1) / kXcd) * kXcd;
    return true;
}

template <typename T>
hipError_t launch_mg_restrict(hipStream_t s, const Dims& f, const Dims& cd, Stencil<T> A,
                              const T* x, const T* b, T* bc, SolveScalars* sc, const double* W,
                              bool ring)
{
    PairGeom g;
    if (ring && restrict_ring_geom<T>(f, g)) {
        constexpr int V = vec_width<T>(), P = MvRingCfg<T>::P, R = MvRingCfg<T>::R;
        const size_t lds = (size_t)(R + 1) * P * g.rpp * f.nz * sizeof(T);
        const void* fn = A.code
                             ? reinterpret_cast<const void*>(&k_mg_restrict_ring<T, 1, V, P, R>)
                             : reinterpret_cast<const void*>(&k_mg_restrict_ring<T, 0, V, P, R>);
        if (lds > 64 * 1024) {
            const hipError_t e = hipFuncSetAttribute(
                fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) return e;
        }
        if (A.code)
            hipLaunchKernelGGL((k_mg_restrict_ring<T, 1, V, P, R>), dim3(g.grid), dim3(kBlock),
                               lds, s, f, cd, g, A, x, b, bc, sc, W);
        else
            hipLaunchKernelGGL((k_mg_restrict_ring<T, 0, V, P, R>), dim3(g.grid), dim3(kBlock),
                               lds, s, f, cd, g, A, x, b, bc, sc, W);
        return hipGetLastError();
    }
    if (LES_MG_VEC && f.nz % 4 == 0) {
        const Chunking ch = make_chunking(cd, 2);
        if (A.code)
            hipLaunchKernelGGL((k_mg_restrict_v<T, 1>), dim3(ch.grid), dim3(kBlock), 0, s, f, cd,
                               ch, A, x, b, bc, sc, W);
        else
            hipLaunchKernelGGL((k_mg_restrict_v<T, 0>), dim3(ch.grid), dim3(kBlock), 0, s, f, cd,
                               ch, A, x, b, bc, sc, W);
        return hipGetLastError();
    }
    const Chunking ch = make_chunking(cd);
    if (A.code)
        hipLaunchKernelGGL((k_mg_restrict<T, 1>), dim3(ch.grid), dim3(kBlock), 0, s, f, cd, ch, A,
                           x, b, bc, sc, W);
    else
        hipLaunchKernelGGL((k_mg_restrict<T, 0>), dim3(ch.grid), dim3(kBlock), 0, s, f, cd, ch, A,
                           x, b, bc, sc, W);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_mg_tail(hipStream_t s, int nlev, const Dims* dims, T* const* const* A,
                          const T* b0, T* x0, int sweeps, int coarse_sweeps, double omega,
                          double scale, SolveScalars* sc)
{
    if (nlev < 1 || nlev > kMgTailLevels) return hipErrorInvalidValue;
    MgTail<T> g{};
    g.nlev = nlev;
    size_t off = 0;
    for (int l = 0; l < nlev; ++l) {
        if (dims[l].N > kMgTailCells) return hipErrorInvalidValue;
        g.d[l] = dims[l];
        for (int q = 0; q < 7; ++q) g.A[l].a[q] = A[l][q];
        g.off[l] = (int)off;
        off += 10 * (size_t)dims[l].N;
    }
    const size_t lds = off * sizeof(T);
    if (lds > kPairMaxLds) return hipErrorInvalidValue;
    g.b0 = b0;
    g.x0 = x0;
    g.sweeps = sweeps;
    g.coarse_sweeps = coarse_sweeps;
    g.omega = omega;
    g.scale = scale;
    const void* fn = reinterpret_cast<const void*>(&k_mg_tail<T>);
    if (lds > 64 * 1024) {
        const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                 (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_mg_tail<T>, dim3(1), dim3(kBlock), lds, s, g, sc);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_mg_coarsest(hipStream_t s, const Dims& d, Stencil<T> A, const T* b, T* x,
                              int sweeps, double omega, SolveScalars* sc)
{
    if (d.N > kBlock || A.code) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_mg_coarsest<T>, dim3(1), dim3(kBlock), 0, s, d, A, b, x, sweeps, omega,
                       sc);
    return hipGetLastError();
}

#define LES_INST(T)                                                                              \
    template hipError_t launch_matvec<T>(hipStream_t, const Dims&, Stencil<T>, const T*, T*);   \
    template hipError_t launch_pcg_init<T>(hipStream_t, const Dims&, Stencil<T>, const T*,      \
                                           const T*, T*, RedScratch, SolveScalars*);            \
    template hipError_t launch_pcg_matvec<T>(hipStream_t, const Dims&, Stencil<T>, const T*,    \
                                             T*, RedScratch, SolveScalars*, const double*);     \
    template hipError_t launch_pcg_matvec_pu<T>(hipStream_t, const Dims&, Stencil<T>, const T*, \
                                                const T*, const T*, T*, T*, RedScratch,         \
                                                SolveScalars*, const double*, T*, int*,         \
                                                RedScratch, int);                               \
    template hipError_t launch_pcg_p_jacobi<T>(hipStream_t, const Dims&, const T*, const T*,    \
                                               T*, int, int, RedScratch, SolveScalars*,         \
                                               const double*);                                  \
    template hipError_t launch_pcg_p_update<T>(hipStream_t, const Dims&, const T*, T*,          \
                                               SolveScalars*);                                  \
    template hipError_t launch_pcg_rz<T>(hipStream_t, const Dims&, const T*, const T*, int,     \
                                         RedScratch, SolveScalars*, const double*);             \
    template hipError_t launch_cheb_step<T>(hipStream_t, const Dims&, Stencil<T>, const T*,     \
                                            const T*, T*, T*, double, double, double, int, int, \
                                            RedScratch, SolveScalars*, const double*);          \
    template hipError_t launch_cheb_scale<T>(hipStream_t, const Dims&, const T*, T*, double,    \
                                             int, RedScratch, SolveScalars*, const double*);    \
    template bool cheb_pair_supported<T>(const Dims&, int, bool);                               \
    template hipError_t launch_cheb_pair<T>(hipStream_t, const Dims&, Stencil<T>, const T*,     \
                                            const T*, T*, const T*, T*, double, double, double, \
                                            double, double, int, SolveScalars*, int);           \
    template hipError_t launch_ic_fwd_level<T>(hipStream_t, const Dims&, int, const T*, T*,     \
                                               const T*, const T*, const T*, const T*,          \
                                               const SolveScalars*);                            \
    template hipError_t launch_ic_bwd_level<T>(hipStream_t, const Dims&, int, T*, const T*,     \
                                               const T*, const T*, const T*,                    \
                                               const SolveScalars*);                            \
    template hipError_t launch_mg_smooth<T>(hipStream_t, const Dims&, Stencil<T>, const T*,     \
                                            const T*, const T*, const Dims&, T*, int, int,      \
                                            double, double, RedScratch, SolveScalars*,          \
                                            const double*, bool);                               \
    template bool mg_post_fits<T>(const Dims&);                                                \
    template bool mg_pre_fits<T>(const Dims&);                                                 \
    template hipError_t launch_mg_pre<T>(hipStream_t, const Dims&, const Dims&, Stencil<T>,     \
                                         const T*, T*, T*, double, SolveScalars*,              \
                                         const double*);                                       \
    template hipError_t launch_mg_post<T>(hipStream_t, const Dims&, Stencil<T>, const T*,       \
                                          const T*, const T*, const Dims&, T*, int, double,     \
                                          double, RedScratch, SolveScalars*, const double*);    \
    template hipError_t launch_mg_restrict<T>(hipStream_t, const Dims&, const Dims&, Stencil<T>, \
                                              const T*, const T*, T*, SolveScalars*,            \
                                              const double*, bool);                             \
    template hipError_t launch_mg_coarsest<T>(hipStream_t, const Dims&, Stencil<T>, const T*,   \
                                              T*, int, double, SolveScalars*);                  \
    template hipError_t launch_mg_tail<T>(hipStream_t, int, const Dims*, T* const* const*,      \
                                          const T*, T*, int, int, double, double,              \
                                          SolveScalars*);

LES_INST(double)
LES_INST(float)
#undef LES_INST

template hipError_t launch_pcg_update<double, double>(hipStream_t, const Dims&, double*, double*,
                                                      const double*, const double*,
                                                      const double*, RedScratch, SolveScalars*,
                                                      const double*);
template hipError_t launch_pcg_update<float, float>(hipStream_t, const Dims&, float*, float*,
                                                    const float*, const float*, const float*,
                                                    RedScratch, SolveScalars*, const double*);
template hipError_t launch_convert<float, double>(hipStream_t, int, const double*, float*);
template hipError_t launch_convert<double, float>(hipStream_t, int, const float*, double*);
template hipError_t launch_convert<double, double>(hipStream_t, int, const double*, double*);

}  // namespace les
