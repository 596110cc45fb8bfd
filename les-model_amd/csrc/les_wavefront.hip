// les_wavefront.hip -- the IC(0) factor, its triangular solves and the
// lexicographic SOR sweep, one launch each (schedule in les_wavefront.hpp).
#include "les_wavefront.hpp"

namespace les {
namespace {

constexpr int kT = kWfTile;
constexpr int kSkew = 2 * (kT - 1);  // largest a+b of a tile
// Levels a wave's loads run ahead of its updates.
#ifndef LES_WF_DEPTH
#define LES_WF_DEPTH 8
#endif
constexpr int kDepth = LES_WF_DEPTH;

__device__ __forceinline__ unsigned long long ld_pk(const unsigned long long* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void st_pk(unsigned long long* p, unsigned long long v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ unsigned long long tagged(unsigned payload, unsigned tag)
{
    return (unsigned long long)payload | ((unsigned long long)tag << 32);
}

// One value from beyond the tile's input edge: a neighbour tile's packets.
template <typename T>
struct Edge;

template <>
struct Edge<float> {
    static constexpr int n = 1;
    unsigned long long w[1];
    __device__ void get(const unsigned long long* p) { w[0] = ld_pk(p); }
    __device__ bool fresh(unsigned tag) const { return (unsigned)(w[0] >> 32) == tag; }
    __device__ float value() const { return __uint_as_float((unsigned)w[0]); }
    __device__ static void put(unsigned long long* p, float v, unsigned tag)
    {
        st_pk(p, tagged(__float_as_uint(v), tag));
    }
};

template <>
struct Edge<double> {
    static constexpr int n = 2;
    unsigned long long w[2];
    __device__ void get(const unsigned long long* p)
    {
        w[0] = ld_pk(p);
        w[1] = ld_pk(p + 1);
    }
    __device__ bool fresh(unsigned tag) const
    {
        return (unsigned)(w[0] >> 32) == tag && (unsigned)(w[1] >> 32) == tag;
    }
    __device__ double value() const
    {
        return __longlong_as_double((long long)((w[1] << 32) | (w[0] & 0xffffffffull)));
    }
    __device__ static void put(unsigned long long* p, double v, unsigned tag)
    {
        const unsigned long long bits = (unsigned long long)__double_as_longlong(v);
        st_pk(p, tagged((unsigned)bits, tag));
        st_pk(p + 1, tagged((unsigned)(bits >> 32), tag));
    }
};

template <typename T>
struct WfArgs {
    Dims d;
    int lo, klo;      // first swept i (and j) and k
    int ni, nj, nk;   // swept extents
    int NI, NJ;       // tiles along i and j
    const T* in[5];
    T* out[4];
    const SolveScalars* sc;
    StatScalars* st;
    double omc, oc, dx2, dy2, dz2;
    WfBufs wb;
};

template <int KIND>
constexpr int n_streamed() { return KIND == WF_SOR ? 2 : (KIND == WF_FACTOR ? 4 : 5); }

// The loads of one level, issued kDepth levels ahead.
template <typename T, int KIND>
struct Slot {
    T v[n_streamed<KIND>()];
    Edge<T> ei, ej;  // packets from the tiles beyond the input i and j edges
    T west, south;   // SOR: the boundary cells i-1 = 0 and j-1 = 0
    T east, north;   // SOR: old values beyond the tile's far i and j edges
};

// Tile of ticket rank r: anti-diagonals I+J in increasing order.
__device__ __forceinline__ void tile_of_rank(int r, int NI, int NJ, int& I, int& J)
{
    I = 0;
    J = 0;
    for (int dg = 0; dg < NI + NJ - 1; ++dg) {
        const int lo = dg - NJ + 1 > 0 ? dg - NJ + 1 : 0;
        const int hi = dg < NI - 1 ? dg : NI - 1;
        const int cnt = hi - lo + 1;
        if (r < cnt) {
            I = lo + r;
            J = dg - I;
            return;
        }
        r -= cnt;
    }
}

__device__ __forceinline__ void max_nonneg(double* addr, double v)
{
    // non-negative doubles order like their bit patterns
    __hip_atomic_fetch_max(reinterpret_cast<unsigned long long*>(addr),
                           (unsigned long long)__double_as_longlong(v), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T, int KIND>
__global__ __launch_bounds__(64) void k_wf(WfArgs<T> g)
{
    constexpr bool REV = KIND == WF_BWD;  // the backward solve walks from the high corner
    constexpr int P = Edge<T>::n;
    using S = Slot<T, KIND>;
    const int lane = threadIdx.x;

    // The ticket comes first, so that every wave of every launch takes one.
    unsigned long long t = 0;
    if (lane == 0) t = atomicAdd(&g.wb.st->ticket[KIND], 1ull);
    t = __shfl(t, 0);
    const unsigned long long ntiles = (unsigned long long)g.NI * (unsigned long long)g.NJ;
    const unsigned long long launch = t / ntiles;
    int rank = (int)(t - launch * ntiles);
    // distinct for every launch of every kind (never 0, the buffers' initial tag)
    const unsigned tag = (unsigned)(launch * WF_KINDS + KIND + 1);
    if constexpr (KIND == WF_FWD || KIND == WF_BWD) {
        if (g.sc && g.sc->done) return;
    }
    if constexpr (KIND == WF_SOR) {
        if (g.st->done) return;
    }
    if (REV) rank = (int)ntiles - 1 - rank;
    int I, J;
    tile_of_rank(rank, g.NI, g.NJ, I, J);

    const int a = lane >> 3, b = lane & 7;
    const int i = g.lo + I * kT + a, j = g.lo + J * kT + b;
    const bool col = i < g.lo + g.ni && j < g.lo + g.nj;  // a swept column
    const bool mem = i < g.d.nx && j < g.d.ny;            // a column of the grid
    const int s = REV ? kSkew - a - b : a + b;
    const int base = mem ? i * g.d.sij + j * g.d.nz : 0;
    const int nx = g.d.nx, ny = g.d.ny, nz = g.d.nz, sij = g.d.sij;

    const int e_in = REV ? kT - 1 : 0;  // the tile edge whose neighbours come first
    const bool in_i = a == e_in, in_j = b == e_in;
    const bool has_in_i = REV ? I < g.NI - 1 : I > 0;
    const bool has_in_j = REV ? J < g.NJ - 1 : J > 0;
    const bool has_out_i = REV ? I > 0 : I < g.NI - 1;
    const bool has_out_j = REV ? J > 0 : J < g.NJ - 1;
    const bool pk_in_i = col && in_i && has_in_i;
    const bool pk_in_j = col && in_j && has_in_j;
    const bool bd_in_i = KIND == WF_SOR && col && in_i && !has_in_i;
    const bool bd_in_j = KIND == WF_SOR && col && in_j && !has_in_j;
    const bool pk_out_i = col && a == kT - 1 - e_in && has_out_i;
    const bool pk_out_j = col && b == kT - 1 - e_in && has_out_j;
    // packet rows: boundary q lies between tiles q and q+1
    const size_t row_in_i = pk_in_i ? ((size_t)(REV ? I : I - 1) * ny + j) * nz : 0;
    const size_t row_in_j = pk_in_j ? ((size_t)(REV ? J : J - 1) * nx + i) * nz : 0;
    const size_t row_out_i = pk_out_i ? ((size_t)(REV ? I - 1 : I) * ny + j) * nz : 0;
    const size_t row_out_j = pk_out_j ? ((size_t)(REV ? J - 1 : J) * nx + i) * nz : 0;

    auto kof = [&](int l) { return REV ? g.klo + g.nk - 1 - (l - s) : g.klo + l - s; };

    // Loads are issued by every lane at every level, from a cell inside the
    // grid (inactive lanes read cell 0 or the first packet), so that no load
    // sits under a branch and the compiler can keep kDepth levels in flight.
    auto load = [&](S& q, int l) {
        const int m = l - s;
        const bool act = col && m >= 0 && m < g.nk;
        const int k = kof(l);
        const int id = act ? base + k : 0;
        const unsigned long long* pxi = (act && pk_in_i) ? g.wb.fx + (row_in_i + k) * P : g.wb.fx;
        const unsigned long long* pxj = (act && pk_in_j) ? g.wb.fy + (row_in_j + k) * P : g.wb.fy;
        if constexpr (KIND == WF_SOR) {
            const T* p = g.in[0];
            // the old value, also on the boundary planes k = 0 and nz-1
            const int kc = k < 0 ? 0 : (k > nz - 1 ? nz - 1 : k);
            q.v[0] = p[base + kc];
            q.v[1] = g.in[1][id];
            q.ei.get(pxi);
            q.ej.get(pxj);
            q.west = p[act && bd_in_i ? id - sij : id];
            q.south = p[act && bd_in_j ? id - nz : id];
            q.east = p[act && a == kT - 1 ? id + sij : id];
            q.north = p[act && b == kT - 1 ? id + nz : id];
        } else {
            q.ei.get(pxi);
            q.ej.get(pxj);
            if constexpr (KIND == WF_FACTOR || KIND == WF_FWD) {
                for (int v = 0; v < n_streamed<KIND>(); ++v) q.v[v] = g.in[v][id];
            } else {
                q.v[0] = g.in[0][id];
                q.v[1] = g.in[1][id];
                q.v[2] = g.in[2][act && i < nx - 1 ? id + sij : id];
                q.v[3] = g.in[3][act && j < ny - 1 ? id + nz : id];
                q.v[4] = g.in[4][act && k < nz - 1 ? id + 1 : id];
            }
        }
    };

    // this lane's value of the previous level; SOR's first level (a+b = 0)
    // reads the boundary cell k = 0 through it
    T prev = 0;
    if constexpr (KIND == WF_SOR) prev = g.in[0][base];
    int budget = g.wb.spin_limit;
    bool stall = false;
    double rmax = 0.0;   // SOR: max |p_new - p_old| of this lane

    auto step = [&](S& q, const S& qn, int l) {
        const int m = l - s;
        const bool act = col && m >= 0 && m < g.nk;
        const int k = kof(l);
        const int id = base + k;
        // level l-1 of the lanes before this one along i and j
        T fi = REV ? __shfl_down(prev, 8) : __shfl_up(prev, 8);
        T fj = REV ? __shfl_down(prev, 1) : __shfl_up(prev, 1);
        T east = 0, north = 0;
        if constexpr (KIND == WF_SOR) {
            // old values of (i+1, j, k) and (i, j+1, k): the next lanes' level l+1
            east = __shfl_down(qn.v[0], 8);
            north = __shfl_down(qn.v[0], 1);
        }
        // the neighbour tiles' packets must carry this launch's tag (the
        // checks are evaluated for every lane, with no branch, so that the
        // compiler knows every reloaded packet has arrived when the loop ends)
        const bool need_i = act && pk_in_i, need_j = act && pk_in_j;
        int stale = ((int)need_i & (int)!q.ei.fresh(tag)) | ((int)need_j & (int)!q.ej.fresh(tag));
        if (__any(stale)) {
            const unsigned long long* pxi = need_i ? g.wb.fx + (row_in_i + k) * P : g.wb.fx;
            const unsigned long long* pxj = need_j ? g.wb.fy + (row_in_j + k) * P : g.wb.fy;
            for (;;) {
                if (budget <= 0) {
                    stall = true;
                    break;
                }
                --budget;
                __builtin_amdgcn_s_sleep(2);
                q.ei.get(pxi);
                q.ej.get(pxj);
                stale = ((int)need_i & (int)!q.ei.fresh(tag)) | ((int)need_j & (int)!q.ej.fresh(tag));
                if (!__any(stale)) break;
            }
        }
        if (pk_in_i) fi = q.ei.value();
        if (pk_in_j) fj = q.ej.value();
        if constexpr (KIND == WF_SOR) {
            if (bd_in_i) fi = q.west;
            if (bd_in_j) fj = q.south;
            if (a == kT - 1) east = q.east;
            if (b == kT - 1) north = q.north;
        }
        T val = prev;
        if (act) {
            if constexpr (KIND == WF_FACTOR) {
                // pressure.cpp:506-555
                const double small_val = 1e-30;
                double sum_sq = 0.0;
                if (i > 0 && fabs(fi) > small_val) {
                    const double lv = q.v[1] / fi;
                    g.out[1][id] = lv;
                    sum_sq += lv * lv;
                }
                if (j > 0 && fabs(fj) > small_val) {
                    const double lv = q.v[2] / fj;
                    g.out[2][id] = lv;
                    sum_sq += lv * lv;
                }
                if (k > 0 && fabs(prev) > small_val) {
                    const double lv = q.v[3] / prev;
                    g.out[3][id] = lv;
                    sum_sq += lv * lv;
                }
                const double diag = q.v[0] - sum_sq;
                val = (diag <= small_val) ? 1.0 : sqrt(diag);
                g.out[0][id] = val;
            } else if constexpr (KIND == WF_FWD) {
                // pressure.cpp:451-463
                T sum = q.v[0];
                if (i > 0) sum -= q.v[2] * fi;
                if (j > 0) sum -= q.v[3] * fj;
                if (k > 0) sum -= q.v[4] * prev;
                val = sum / q.v[1];
                g.out[0][id] = val;
            } else if constexpr (KIND == WF_BWD) {
                // pressure.cpp:469-481
                T sum = q.v[0];
                if (i < nx - 1) sum -= q.v[2] * fi;
                if (j < ny - 1) sum -= q.v[3] * fj;
                if (k < nz - 1) sum -= q.v[4] * prev;
                val = sum / q.v[1];
                g.out[0][id] = val;
            } else {
                // pressure.cpp:216-234
                const double pold = q.v[0];
                val = g.omc * pold + g.oc * ((east + fi) / g.dx2 + (north + fj) / g.dy2 +
                                             (qn.v[0] + prev) / g.dz2 - q.v[1]);
                g.out[0][id] = val;
                const double r = fabs(val - pold);
                rmax = rmax > r ? rmax : r;
            }
            if (pk_out_i) Edge<T>::put(g.wb.fx + (row_out_i + k) * P, val, tag);
            if (pk_out_j) Edge<T>::put(g.wb.fy + (row_out_j + k) * P, val, tag);
        } else if constexpr (KIND == WF_SOR) {
            val = q.v[0];  // the old value: the k = 0 boundary is the first level's k-1
        }
        prev = val;
    };

    S q[kDepth];
#pragma unroll
    for (int u = 0; u < kDepth; ++u) load(q[u], u);
    const int nlev = g.nk + kSkew;
    for (int l0 = 0; l0 < nlev; l0 += kDepth) {
#pragma unroll
        for (int u = 0; u < kDepth; ++u) {
            step(q[u], q[(u + 1) % kDepth], l0 + u);
            load(q[u], l0 + u + kDepth);
        }
    }
    if constexpr (KIND == WF_SOR) {
        double mx = rmax;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const double x = __shfl_xor(mx, o);
            mx = mx > x ? mx : x;
        }
        if (lane == 0 && mx > 0.0) max_nonneg(&g.st->maxr, mx);
    }
    if (stall && lane == 0) atomicOr(&g.wb.st->stalled, 1);
}

template <typename T, int KIND>
hipError_t launch_wf(hipStream_t s, WfArgs<T>& g)
{
    if (g.ni <= 0 || g.nj <= 0 || g.nk <= 0) return hipSuccess;
    g.NI = (g.ni + kT - 1) / kT;
    g.NJ = (g.nj + kT - 1) / kT;
    hipLaunchKernelGGL((k_wf<T, KIND>), dim3(g.NI * g.NJ), dim3(64), 0, s, g);
    return hipGetLastError();
}

template <typename T>
WfArgs<T> full_grid(const Dims& d, const WfBufs& wb)
{
    WfArgs<T> g{};
    g.d = d;
    g.lo = 0;
    g.klo = 0;
    g.ni = d.nx;
    g.nj = d.ny;
    g.nk = d.nz;
    g.wb = wb;
    return g;
}

}  // namespace

size_t wf_face_packets_x(const Dims& d)
{
    const size_t NI = (size_t)(d.nx + kT - 1) / kT;
    return NI > 1 ? (NI - 1) * (size_t)d.ny * d.nz * 2 : 0;
}

size_t wf_face_packets_y(const Dims& d)
{
    const size_t NJ = (size_t)(d.ny + kT - 1) / kT;
    return NJ > 1 ? (NJ - 1) * (size_t)d.nx * d.nz * 2 : 0;
}

hipError_t launch_wf_ic_factor(hipStream_t s, const Dims& d, Stencil<double> A, double* Lc,
                               double* Lw, double* Ls, double* Lb, const WfBufs& wb)
{
    WfArgs<double> g = full_grid<double>(d, wb);
    g.in[0] = A.a[SC];
    g.in[1] = A.a[SW];
    g.in[2] = A.a[SS];
    g.in[3] = A.a[SB];
    g.out[0] = Lc;
    g.out[1] = Lw;
    g.out[2] = Ls;
    g.out[3] = Lb;
    return launch_wf<double, WF_FACTOR>(s, g);
}

template <typename T>
hipError_t launch_wf_ic_fwd(hipStream_t s, const Dims& d, const T* r, T* z, const T* Lc,
                            const T* Lw, const T* Ls, const T* Lb, const SolveScalars* sc,
                            const WfBufs& wb)
{
    WfArgs<T> g = full_grid<T>(d, wb);
    g.in[0] = r;
    g.in[1] = Lc;
    g.in[2] = Lw;
    g.in[3] = Ls;
    g.in[4] = Lb;
    g.out[0] = z;
    g.sc = sc;
    return launch_wf<T, WF_FWD>(s, g);
}

template <typename T>
hipError_t launch_wf_ic_bwd(hipStream_t s, const Dims& d, T* z, const T* Lc, const T* Lw,
                            const T* Ls, const T* Lb, const SolveScalars* sc, const WfBufs& wb)
{
    WfArgs<T> g = full_grid<T>(d, wb);
    g.in[0] = z;
    g.in[1] = Lc;
    g.in[2] = Lw;
    g.in[3] = Ls;
    g.in[4] = Lb;
    g.out[0] = z;
    g.sc = sc;
    return launch_wf<T, WF_BWD>(s, g);
}

hipError_t launch_wf_sor(hipStream_t s, const Dims& d, double* p, const double* rhs,
                         double one_m_omega, double omega_coef, double dx2, double dy2,
                         double dz2, StatScalars* st, const WfBufs& wb)
{
    WfArgs<double> g{};
    g.d = d;
    g.lo = 1;
    g.klo = 1;
    g.ni = d.nx - 2;
    g.nj = d.ny - 2;
    g.nk = d.nz - 2;
    g.in[0] = p;
    g.in[1] = rhs;
    g.out[0] = p;
    g.st = st;
    g.omc = one_m_omega;
    g.oc = omega_coef;
    g.dx2 = dx2;
    g.dy2 = dy2;
    g.dz2 = dz2;
    g.wb = wb;
    return launch_wf<double, WF_SOR>(s, g);
}

#define LES_WF_INSTANTIATE(T)                                                                    \
    template hipError_t launch_wf_ic_fwd<T>(hipStream_t, const Dims&, const T*, T*, const T*,   \
                                            const T*, const T*, const T*, const SolveScalars*,  \
                                            const WfBufs&);                                     \
    template hipError_t launch_wf_ic_bwd<T>(hipStream_t, const Dims&, T*, const T*, const T*,   \
                                            const T*, const T*, const SolveScalars*,            \
                                            const WfBufs&);
LES_WF_INSTANTIATE(float)
LES_WF_INSTANTIATE(double)
#undef LES_WF_INSTANTIATE

}  // namespace les
