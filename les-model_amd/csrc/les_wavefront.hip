// les_wavefront.hip -- the IC(0) factor, its triangular solves and the
// lexicographic SOR sweep, one launch each (schedule in les_wavefront.hpp).
//
// Memory traffic.  A lane's cells are one k-column, and the skew puts the
// 64 lanes of a wave at 15 different k, so reading the operands cell by cell
// would touch 64 cache lines per load.  Instead the wave stages its 64
// columns through an LDS ring: every C levels it loads the next C k-values
// of each column with coalesced loads (C*sizeof(T) = 64 contiguous bytes per
// column) and each lane reads its own, skewed, k from LDS.  The tile-edge
// packets are laid out per (boundary, tile, level, lane) so that the eight
// lanes of an edge read and write one contiguous 64/128-byte run.
#include "les_wavefront.hpp"

namespace les {
namespace {

constexpr int kT = kWfTile;          // 8 x 8 columns per wave
constexpr int kSkew = 2 * (kT - 1);  // largest lane skew a+b
// Levels the packet loads run ahead of their use.
#ifndef LES_WF_DEPTH
#define LES_WF_DEPTH 4
#endif
constexpr int kDepth = LES_WF_DEPTH;

// The LDS ring: C k-values of a column per chunk, M chunks, column pitch
// PITCH elements.  A chunk is read during the C+14 levels its k-values are
// in use by some lane and is written two levels before the first of them,
// so a ring slot is free again when (M-1)*C >= 14 + 2.  The pitches make
// the skewed reads at most 2-way and the staging writes free of bank
// conflicts (checked for every level phase).
template <typename T>
struct Ring;
template <>
struct Ring<double> {
    static constexpr int C = 8, M = 3, PITCH = 24;
};
template <>
struct Ring<float> {
    static constexpr int C = 16, M = 2, PITCH = 48;
};

// Arrays a kind stages for its own 64 columns, and whether it also stages
// the 16 columns beyond its far edges, (i0+8, j0+b) and (i0+a, j0+8).
template <int KIND>
struct Staged {
    static constexpr int n = KIND == WF_SOR ? 2 : (KIND == WF_FACTOR ? 4 : 5);
    static constexpr bool edge = KIND == WF_SOR || KIND == WF_BWD;
};

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

// A value crosses a tile edge as Pk<T>::n 64-bit packets: 32 payload bits
// and the launch's tag each.
template <typename T>
struct Pk;

template <>
struct Pk<float> {
    static constexpr int n = 1;
    __device__ static float value(const unsigned long long* w)
    {
        return __uint_as_float((unsigned)w[0]);
    }
    __device__ static void put(unsigned long long* p, float v, unsigned tag)
    {
        st_pk(p, tagged(__float_as_uint(v), tag));
    }
};

template <>
struct Pk<double> {
    static constexpr int n = 2;
    __device__ static double value(const unsigned long long* w)
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

template <typename T, int V>
struct VecOf;
template <typename T>
struct VecOf<T, 1> {
    using type = T;
};
template <>
struct VecOf<double, 2> {
    using type = double2;
};
template <>
struct VecOf<float, 4> {
    using type = float4;
};

template <typename T, int V>
__device__ __forceinline__ void ld_vec(const T* p, T* out)
{
    if constexpr (V == 1) {
        out[0] = *p;
    } else {
        using Vec = typename VecOf<T, V>::type;
        const Vec x = *reinterpret_cast<const Vec*>(p);
        out[0] = x.x;
        out[1] = x.y;
        if constexpr (V == 4) {
            out[2] = x.z;
            out[3] = x.w;
        }
    }
}

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

// LDS written by some lanes of the wave and then read by others: the LDS
// instructions of one wave complete in order, so a compiler-level fence is
// all that is needed.
__device__ __forceinline__ void wave_lds_fence()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T, int KIND, int V>
__global__ __launch_bounds__(64) void k_wf(WfArgs<T> g)
{
    constexpr bool REV = KIND == WF_BWD;  // the backward solve walks from the high corner
    constexpr bool SOR = KIND == WF_SOR;
    constexpr int P = Pk<T>::n;
    constexpr int C = Ring<T>::C, M = Ring<T>::M, PITCH = Ring<T>::PITCH, R = C * M;
    constexpr int NA = Staged<KIND>::n;
    constexpr bool EDGE = Staged<KIND>::edge;
    constexpr int CPL = C / V;     // lanes sharing one column in a staging load
    constexpr int CPI = 64 / CPL;  // columns one staging load covers
    constexpr int G = C / V;       // staging loads per array
    constexpr int GE = EDGE ? 16 / CPI : 0;
    static_assert((M - 1) * C >= kSkew + 2, "LDS ring too short for the skew");
    static_assert(16 % CPI == 0 && C % kDepth == 0, "staging shape");
    __shared__ T sm[(NA * 64 + (EDGE ? 16 : 0)) * PITCH];
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
    if constexpr (SOR) {
        if (g.st->done) return;
    }
    if (REV) rank = (int)ntiles - 1 - rank;
    int I, J;
    tile_of_rank(rank, g.NI, g.NJ, I, J);

    const int nx = g.d.nx, ny = g.d.ny, nz = g.d.nz, sij = g.d.sij;
    const int a = lane >> 3, b = lane & 7;
    const int i0 = g.lo + I * kT, j0 = g.lo + J * kT;
    const int i = i0 + a, j = j0 + b;
    auto swept = [&](int ci, int cj) { return ci < g.lo + g.ni && cj < g.lo + g.nj; };
    const bool col = swept(i, j);
    const int s = REV ? kSkew - a - b : a + b;
    const int base = (i < nx && j < ny) ? i * sij + j * nz : 0;
    const int K1 = g.klo + g.nk - 1;
    const int nlev = g.nk + kSkew;
    auto kof = [&](int l, int sk) { return REV ? K1 - (l - sk) : g.klo + l - sk; };
    auto active = [&](bool c, int l, int sk) { return c && l - sk >= 0 && l - sk < g.nk; };
    auto slot = [&](int k) { return (k + 2 * R) % R; };  // every level reads k >= -R

    // ---- edge roles.  A lane on the input edge waits for one neighbour
    // tile's value, a lane on the output edge hands one over; the corner
    // lane's second (j) edge is served by a helper lane, so that every lane
    // issues one packet load and one packet store per level.
    const int ein = REV ? kT - 1 : 0, eout = kT - 1 - ein;
    const bool has_in_i = REV ? I < g.NI - 1 : I > 0;
    const bool has_in_j = REV ? J < g.NJ - 1 : J > 0;
    const bool has_out_i = REV ? I > 0 : I < g.NI - 1;
    const bool has_out_j = REV ? J > 0 : J < g.NJ - 1;
    // SOR reads the fixed boundary cells i-1 = 0 and j-1 = 0 where there is
    // no neighbour tile (the IC kinds have no term there)
    const bool need_i = has_in_i || SOR, need_j = has_in_j || SOR;
    const int corner_in = REV ? 63 : 0, helper_in = REV ? 54 : 9;
    const int corner_out = REV ? 0 : 63, helper_out = REV ? 9 : 54;
    // input role: 0 none, 1 packet, 2 boundary cell (raw bits)
    int in_mode = 0, in_s = s;
    bool in_is_i = false, in_col = col;
    const bool in_helper = lane == helper_in && need_j && need_i;  // loads for the corner
    const unsigned long long* in_a0 = g.wb.fx;  // level-0 packet / boundary column
    if (a == ein && need_i) {
        in_is_i = true;
        if (has_in_i) {
            in_mode = 1;
            in_a0 = g.wb.fx + (((size_t)(REV ? I : I - 1) * g.NJ + J) * nlev * kT + b) * P;
        } else {
            in_mode = 2;
            in_a0 = reinterpret_cast<const unsigned long long*>(g.in[0] + (i - 1) * sij + j * nz);
        }
    } else if ((b == ein && need_j) || in_helper) {
        // the helper works for the corner lane: its cell, skew and column
        const int ta = in_helper ? (corner_in >> 3) : a;
        const int tb = in_helper ? (corner_in & 7) : b;
        in_s = REV ? kSkew - ta - tb : ta + tb;
        in_col = swept(i0 + ta, j0 + tb);
        if (has_in_j) {
            in_mode = 1;
            in_a0 = g.wb.fy + (((size_t)(REV ? J : J - 1) * g.NI + I) * nlev * kT + ta) * P;
        } else {
            in_mode = 2;
            in_a0 = reinterpret_cast<const unsigned long long*>(g.in[0] + (i0 + ta) * sij +
                                                                (j0 + tb - 1) * nz);
        }
    }
    const bool corner_split = lane == corner_in && need_i && need_j;  // fj from helper_in
    const size_t in_step = in_mode == 1 ? (size_t)kT * P : 1;        // per level / per k
    // output role: 1 own value, 2 the corner lane's (helper)
    int out_mode = 0, out_s = s;
    bool out_col = col;
    unsigned long long* out_a0 = g.wb.fx;
    if (a == eout && has_out_i) {
        out_mode = 1;
        out_a0 = g.wb.fx + (((size_t)(REV ? I - 1 : I) * g.NJ + J) * nlev * kT + b) * P;
    } else if ((b == eout && has_out_j) || (lane == helper_out && has_out_j && has_out_i)) {
        const int ta = lane == helper_out ? (corner_out >> 3) : a;
        const int tb = lane == helper_out ? (corner_out & 7) : b;
        out_mode = lane == helper_out ? 2 : 1;
        out_s = REV ? kSkew - ta - tb : ta + tb;
        out_col = swept(i0 + ta, j0 + tb);
        out_a0 = g.wb.fy + (((size_t)(REV ? J - 1 : J) * g.NI + I) * nlev * kT + ta) * P;
    }

    // this lane's input for level l (a dummy in-bounds address when there is
    // nothing to read, so that every lane loads, with no branch)
    auto in_addr = [&](int l) -> const unsigned long long* {
        if (in_mode == 0 || !active(in_col, l, in_s)) return g.wb.fx;
        if (in_mode == 1) return in_a0 + (size_t)l * in_step;
        return in_a0 + kof(l, in_s);
    };
    unsigned long long pq[kDepth][P];
    auto pq_load = [&](int d, int l) {
        const unsigned long long* p = in_addr(l);
        pq[d][0] = ld_pk(p);
        if constexpr (P == 2) pq[d][1] = ld_pk(in_mode == 1 ? p + 1 : p);
    };

    // ---- staging of the 64 columns (and the 16 edge columns) through LDS
    const int cmax = (nz - 1) / C;
    const int ctop = K1 / C;
    const int F0 = REV ? (K1 % C) - C + 1 : -g.klo;  // first level of the walk's chunk 0
    constexpr int NSTG = NA * G * V + (EDGE ? GE * V : 0);
    T stg0[NSTG], stg1[NSTG];  // two chunks in flight: loaded two phases before use
    auto chunk_of = [&](int q) { return REV ? ctop - q : q; };
    auto load_chunk = [&](int q, T* stg) {
        const int ck = chunk_of(q);
        const bool okc = ck >= 0 && ck <= cmax;
        const int kk = ck * C + (lane % CPL) * V;
#pragma unroll
        for (int x = 0; x < NA; ++x) {
#pragma unroll
            for (int gi = 0; gi < G; ++gi) {
                const int c = gi * CPI + lane / CPL;
                const int ic = i0 + (c >> 3), jc = j0 + (c & 7);
                const bool ok = okc && ic < nx && jc < ny && kk < nz;
                ld_vec<T, V>(g.in[x] + (ok ? ic * sij + jc * nz + kk : 0), &stg[(x * G + gi) * V]);
            }
        }
        if constexpr (EDGE) {
            // BWD: Lw at (i0+8, j0+b) and Ls at (i0+a, j0+8); SOR: p at both
            const T* ex = KIND == WF_BWD ? g.in[2] : g.in[0];
            const T* ey = KIND == WF_BWD ? g.in[3] : g.in[0];
#pragma unroll
            for (int ge = 0; ge < GE; ++ge) {
                const int c = ge * CPI + lane / CPL;  // 0..15
                const int ic = c < 8 ? i0 + 8 : i0 + c - 8;
                const int jc = c < 8 ? j0 + c : j0 + 8;
                const bool ok = okc && ic < nx && jc < ny && kk < nz;
                ld_vec<T, V>((c < 8 ? ex : ey) + (ok ? ic * sij + jc * nz + kk : 0),
                             &stg[NA * G * V + ge * V]);
            }
        }
    };
    auto store_chunk = [&](int q, const T* stg) {
        const int ck = chunk_of(q);
        const int sb = ((ck % M) + M) % M * C + (lane % CPL) * V;
#pragma unroll
        for (int x = 0; x < NA; ++x) {
#pragma unroll
            for (int gi = 0; gi < G; ++gi) {
                const int c = gi * CPI + lane / CPL;
#pragma unroll
                for (int v = 0; v < V; ++v)
                    sm[(x * 64 + c) * PITCH + sb + v] = stg[(x * G + gi) * V + v];
            }
        }
        if constexpr (EDGE) {
#pragma unroll
            for (int ge = 0; ge < GE; ++ge) {
                const int c = ge * CPI + lane / CPL;
#pragma unroll
                for (int v = 0; v < V; ++v)
                    sm[(NA * 64 + c) * PITCH + sb + v] = stg[NA * G * V + ge * V + v];
            }
        }
    };

    // Results go back into ring 0 (the FACTOR: rings 0-3) in place of the
    // operand they replace, and a chunk is written out with coalesced stores
    // M phases after it was staged, when every lane is done with it: chunk q
    // is complete after level F(q)+C+13 < F(q)-2+M*C.  Cells outside the
    // swept box keep their staged (old) values, so writing them back is a
    // no-op.
    constexpr int NO = KIND == WF_FACTOR ? 4 : 1;
    auto flush_chunk = [&](int q) {
        const int ck = chunk_of(q);
        const bool okc = q >= 0 && ck >= 0 && ck <= cmax;
        const int sb = ((ck % M) + M) % M * C + (lane % CPL) * V;
        const int kk = ck * C + (lane % CPL) * V;
#pragma unroll
        for (int o = 0; o < NO; ++o) {
#pragma unroll
            for (int gi = 0; gi < G; ++gi) {
                const int c = gi * CPI + lane / CPL;
                const int ic = i0 + (c >> 3), jc = j0 + (c & 7);
                {
                    T* dst = (okc && ic < nx && jc < ny && kk < nz)
                                 ? g.out[o] + ic * sij + jc * nz + kk
                                 : reinterpret_cast<T*>(g.wb.st->sink);
                    if constexpr (V == 1) {
                        *dst = sm[(o * 64 + c) * PITCH + sb];
                    } else {
                        using Vec = typename VecOf<T, V>::type;
                        Vec x;
                        x.x = sm[(o * 64 + c) * PITCH + sb];
                        x.y = sm[(o * 64 + c) * PITCH + sb + 1];
                        if constexpr (V == 4) {
                            x.z = sm[(o * 64 + c) * PITCH + sb + 2];
                            x.w = sm[(o * 64 + c) * PITCH + sb + 3];
                        }
                        *reinterpret_cast<Vec*>(dst) = x;
                    }
                }
            }
        }
    };

    // this lane's value of the previous level; SOR's first level (a+b = 0)
    // reads the boundary cell k = 0 through it
    T prev = 0;
    if constexpr (SOR) prev = g.in[0][base];
    T prev_lb = 0;  // BWD: Lb of the cell above (k+1), read on the previous level
    int budget = g.wb.spin_limit;
    bool stall = false;
    double rmax = 0.0;  // SOR: max |p_new - p_old| of this lane

    auto step = [&](int d, int l) {
        const bool act = active(col, l, s);
        const int k = kof(l, s);
        const int sl = slot(k);
        // level l-1 of the lanes before this one along i and j
        T fi = REV ? __shfl_down(prev, 8) : __shfl_up(prev, 8);
        T fj = REV ? __shfl_down(prev, 1) : __shfl_up(prev, 1);
        // the neighbour tile's packets must carry this launch's tag (checked
        // for every lane, with no branch, so that the compiler knows every
        // reloaded packet has arrived when the loop ends)
        const bool need = in_mode == 1 && active(in_col, l, in_s);
        auto fresh = [&]() {
            bool f = (unsigned)(pq[d][0] >> 32) == tag;
            if constexpr (P == 2) f = f && (unsigned)(pq[d][1] >> 32) == tag;
            return f;
        };
        int stale = (int)need & (int)!fresh();
        if (__any(stale)) {
            for (;;) {
                if (budget <= 0) {
                    stall = true;
                    break;
                }
                --budget;
                __builtin_amdgcn_s_sleep(2);
                pq_load(d, l);
                stale = (int)need & (int)!fresh();
                if (!__any(stale)) break;
            }
        }
        T pv;
        if constexpr (P == 2) {
            // a boundary cell's raw bits arrive whole in the first word
            const unsigned long long raw[2] = {pq[d][0], pq[d][0] >> 32};
            pv = in_mode == 2 ? Pk<T>::value(raw) : Pk<T>::value(pq[d]);
        } else {
            pv = Pk<T>::value(pq[d]);
        }
        const T hv = __shfl(pv, helper_in);  // the corner lane's j value
        if (in_mode != 0 && !in_helper) {
            if (in_is_i) fi = pv;
            else fj = pv;
        }
        if (corner_split) fj = hv;

        T v[NA];
#pragma unroll
        for (int x = 0; x < NA; ++x) v[x] = sm[(x * 64 + lane) * PITCH + sl];
        T val = prev;
        if constexpr (KIND == WF_FACTOR) {
            // pressure.cpp:506-555
            if (act) {
                const double small_val = 1e-30;
                double sum_sq = 0.0;
                // an entry the reference leaves alone keeps StencilMatrix's 0
                double lw = 0.0, ls = 0.0, lb = 0.0;
                if (i > 0 && fabs(fi) > small_val) {
                    lw = v[1] / fi;
                    sum_sq += lw * lw;
                }
                if (j > 0 && fabs(fj) > small_val) {
                    ls = v[2] / fj;
                    sum_sq += ls * ls;
                }
                if (k > 0 && fabs(prev) > small_val) {
                    lb = v[3] / prev;
                    sum_sq += lb * lb;
                }
                const double diag = v[0] - sum_sq;
                val = (diag <= small_val) ? 1.0 : sqrt(diag);
                sm[(0 * 64 + lane) * PITCH + sl] = val;
                sm[(1 * 64 + lane) * PITCH + sl] = lw;
                sm[(2 * 64 + lane) * PITCH + sl] = ls;
                sm[(3 * 64 + lane) * PITCH + sl] = lb;
            }
        } else if constexpr (KIND == WF_FWD) {
            // pressure.cpp:451-463
            T sum = v[0];
            if (i > 0) sum -= v[2] * fi;
            if (j > 0) sum -= v[3] * fj;
            if (k > 0) sum -= v[4] * prev;
            if (act) {
                val = sum / v[1];
                sm[lane * PITCH + sl] = val;
            }
        } else if constexpr (KIND == WF_BWD) {
            // pressure.cpp:469-481: the i+1 and j+1 coefficients sit in the
            // next lanes' columns (or the edge columns); k+1's was read last level
            const T lw1 = sm[(a < kT - 1 ? 2 * 64 + lane + 8 : NA * 64 + b) * PITCH + sl];
            const T ls1 = sm[(b < kT - 1 ? 3 * 64 + lane + 1 : NA * 64 + 8 + a) * PITCH + sl];
            T sum = v[0];
            if (i < nx - 1) sum -= lw1 * fi;
            if (j < ny - 1) sum -= ls1 * fj;
            if (k < nz - 1) sum -= prev_lb * prev;
            if (act) {
                val = sum / v[1];
                sm[lane * PITCH + sl] = val;
            }
            prev_lb = v[4];
        } else {
            // pressure.cpp:216-234: old values of (i+1, j, k) and (i, j+1, k)
            // from the next lanes' columns (or the edge columns), (i, j, k+1)
            // from this column's next slot
            const T east = sm[(a < kT - 1 ? lane + 8 : NA * 64 + b) * PITCH + sl];
            const T north = sm[(b < kT - 1 ? lane + 1 : NA * 64 + 8 + a) * PITCH + sl];
            const T top = sm[lane * PITCH + slot(k + 1)];
            const double pold = v[0];
            if (act) {
                val = g.omc * pold + g.oc * ((east + fi) / g.dx2 + (north + fj) / g.dy2 +
                                             (top + prev) / g.dz2 - v[1]);
                // the next lanes read this slot as their old west/south value
                // one level earlier, this lane as its old top value last level
                sm[lane * PITCH + sl] = val;
                const double r = fabs(val - pold);
                rmax = rmax > r ? rmax : r;
            } else {
                val = pold;  // the old value: the k = 0 boundary is the first level's k-1
            }
        }
        // hand the output edge over (the helper stores the corner lane's value);
        // a value of level l is read by the next tile at its level l-7
        const T cv = __shfl(val, corner_out);
        if (out_mode != 0 && active(out_col, l, out_s))
            Pk<T>::put(out_a0 + (size_t)(l - (kT - 1)) * kT * P, out_mode == 2 ? cv : val, tag);
        prev = val;
    };

    // ---- the walk: phase q writes chunk q-M out, chunk q into the ring,
    // starts loading chunk q+2, then runs its C levels; phases go in pairs so
    // that each staging buffer stays in registers
    const int w0 = F0 - 2;
    const int qlast = REV ? ctop - g.klo / C : K1 / C;  // the last chunk with swept cells
    auto phase = [&](int q, T* stg) {
        const int w = w0 + q * C;
        flush_chunk(q - M);
        store_chunk(q, stg);
        wave_lds_fence();
        load_chunk(q + 2, stg);
#pragma unroll
        for (int u = 0; u < C; ++u) {
            const int l = w + u;
            if (l >= 0 && l < nlev) step(u % kDepth, l);
            pq_load(u % kDepth, l + kDepth);
        }
        wave_lds_fence();
    };
    load_chunk(0, stg0);
    load_chunk(1, stg1);
#pragma unroll
    for (int d = 0; d < kDepth; ++d) pq_load(d, w0 + d);
    for (int q = 0; q <= qlast + M; q += 2) {
        phase(q, stg0);
        phase(q + 1, stg1);
    }
    if constexpr (SOR) {
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

// Width of the staging loads: 16-byte loads when every staged column chunk
// starts 16-byte aligned.
template <typename T>
int stage_width(const WfArgs<T>& g, int nptr)
{
    constexpr int VMAX = 16 / sizeof(T);
    if (g.d.nz % VMAX != 0) return 1;
    for (int x = 0; x < nptr; ++x)
        if ((reinterpret_cast<uintptr_t>(g.in[x]) % 16) != 0) return 1;
    return VMAX;
}

template <typename T, int KIND>
hipError_t launch_wf(hipStream_t s, WfArgs<T>& g, int nptr)
{
    if (g.ni <= 0 || g.nj <= 0 || g.nk <= 0) return hipSuccess;
    g.NI = (g.ni + kT - 1) / kT;
    g.NJ = (g.nj + kT - 1) / kT;
    const dim3 grid(g.NI * g.NJ), block(64);
    if (stage_width(g, nptr) > 1)
        hipLaunchKernelGGL((k_wf<T, KIND, 16 / sizeof(T)>), grid, block, 0, s, g);
    else
        hipLaunchKernelGGL((k_wf<T, KIND, 1>), grid, block, 0, s, g);
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

// packets of one face buffer: boundaries x tiles along them x levels x 8
// lanes x 2 (the full grid's tiles and levels bound the SOR sweep's)
size_t face_packets(int n_across, int n_along, int nz)
{
    const size_t NA = (size_t)(n_across + kT - 1) / kT, NB = (size_t)(n_along + kT - 1) / kT;
    return NA > 1 ? (NA - 1) * NB * (size_t)(nz + kSkew) * kT * 2 : 0;
}

}  // namespace

size_t wf_face_packets_x(const Dims& d) { return face_packets(d.nx, d.ny, d.nz); }
size_t wf_face_packets_y(const Dims& d) { return face_packets(d.ny, d.nx, d.nz); }

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
    return launch_wf<double, WF_FACTOR>(s, g, 4);
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
    return launch_wf<T, WF_FWD>(s, g, 5);
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
    return launch_wf<T, WF_BWD>(s, g, 5);
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
    return launch_wf<double, WF_SOR>(s, g, 2);
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
