// les_api.hip -- the C ABI (include/les_gpu.h) over a device-resident context.
//
// The context owns every device buffer: the Field members, the mask, the
// stencil A (FP64, plus an FP32 copy for mixed precision), the IC(0) factor
// L, Krylov workspaces, and the device-resident solver scalars.  Host
// scalars the reference forms with its own expressions (1/(dx*dx), rho/dt,
// the Chebyshev alpha/beta sequence, the mt19937 start vector ...) are
// formed here on the host the same way, so device results match the
// reference's bit for bit wherever no reduction is involved.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <mutex>
#include <new>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "../../include/les_gpu.h"
#include "les_adjoint.hpp"
#include "les_io.hpp"
#include "les_mixed.hpp"
#include "les_physics.hpp"
#include "les_wavefront.hpp"

using namespace les;

// Multigrid preconditioner: levels while the largest dimension exceeds
// kMgCoarseMax (so the coarsest has at most 64 cells, one workgroup's), the
// damping and the coarse-correction factor (chosen on the CPU for PCG at
// 64^3 and 128^3 with and without solid masks: 19-24 iterations).
constexpr int kMgMaxLevels = 16;
constexpr int kMgCoarseMax = 4;
constexpr double kMgOmega = 0.85;
constexpr double kMgScale = 1.8;
// 1: the levels of at most kMgTailCells cells below level 0 run as one
// launch (k_mg_tail); 0: the per-level launches
#ifndef LES_MG_TAIL
#define LES_MG_TAIL 1
#endif

struct les_ctx {
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    Dims d{};
    double dx = 0, dy = 0, dz = 0;
    double* f[LES_F_COUNT] = {};
    uint8_t* solid = nullptr;
    double* A[7] = {};
    bool have_A = false;
    bool A_from_mask = false;  // A was built by les_build_stencil from the mask
    bool A_has_obstacles = false;  // ... and that mask had solid cells
    uint16_t* code = nullptr;  // the same operator as one 16-bit code per cell
    bool have_code = false;
    double* L[7] = {};
    bool have_L = false;
    float* A32[7] = {};
    bool have_A32 = false;
    float* L32[7] = {};
    bool have_L32 = false;
    // FP64 Krylov workspaces
    double *r = nullptr, *z = nullptr, *p = nullptr, *p2 = nullptr, *q = nullptr, *za = nullptr,
           *zb = nullptr, *pc = nullptr, *pc2 = nullptr, *t1 = nullptr, *t2 = nullptr;
    MixedWork mw{};
    RedScratch red{};
    RedScratch red2{};                // a second partials buffer (a producer whose
                                      // consumer reduces while writing its own partials)
    RedScratch red3{};                // a third: the r update's while red holds r.z
    SolveScalars* sc = nullptr;       // device
    SolveScalars* sc_host = nullptr;  // pinned, two slots
    StatScalars* st = nullptr;        // device
    StatScalars* st_host = nullptr;   // pinned
    double* dsc = nullptr;            // device scalar slots
    double* hsc = nullptr;            // pinned scalar slots
    double* trace = nullptr;          // device, trace_cap entries
    int trace_cap = 0;
    int last_trace_n = 0;
    hipEvent_t ev[2] = {};
    // differentiation
    double* W = nullptr;              // mirror weights (LES_IP_MIRROR)
    bool have_W = false;
    double* ckpt = nullptr;           // checkpoints, 3N doubles each
    int ckpt_cap = 0;
    double* seg = nullptr;            // states of the segment being recomputed
    int seg_cap = 0;
    double* adjb[16] = {};            // adjoint work arrays
    double* shadow[LES_F_COUNT] = {}; // field set used while recomputing
    bool have_roll = false;
    int roll_nsteps = 0, roll_every = 0;
    double roll_dt = 0, roll_nu = 0, roll_rho = 0, roll_Cs = 0;
    // multigrid preconditioner: the levels below the grid and their P^T A P
    // operators, rebuilt when A changes; x, t, b per level and precision
    struct Mg {
        int nlev = 0;
        Dims dims[kMgMaxLevels] = {};
        double* A64[kMgMaxLevels][7] = {};
        float* A32[kMgMaxLevels][7] = {};
        double* v64[kMgMaxLevels][3] = {};
        float* v32[kMgMaxLevels][3] = {};
        bool built = false, have32 = false;
        bool weighted = false;  // level 1 is P^T W A P (mirror weights)
    } mg;
    int mg_sweeps = 2;                // damped-Jacobi sweeps before and after the coarse correction
    int mg_coarse_sweeps = 32;        // sweeps on the coarsest level
    int replace_decades = 3;          // mixed: replacement after this many decades of drop (0: off)
    bool mg_rings = true;             // multigrid sweep pairs as one pass (LDS rings) where they fit
    bool cheb_pairs = true;           // fuse Chebyshev steps in pairs where the grid allows
    int pair_grid = 0;                // workgroups of a pair launch (0: built-in default)
    // lexicographic sweeps (IC(0), SOR): one wavefront launch each, or one
    // launch per i+j+k hyperplane
    bool wavefront = true;
    int spin_limit = kWfDefaultSpin;
    WfState* wf = nullptr;
    unsigned long long* wf_fx = nullptr;
    unsigned long long* wf_fy = nullptr;
    bool wf_ran = false;              // a wavefront launch since the last stall check
    // checkpoint, VTK output, statistics
    void* io_stage[2] = {};           // pinned staging chunks of the checkpoint streams
    hipEvent_t io_ev[2] = {};         // completion of each chunk's copy
    double* stat_sum[kStatArrays] = {};   // running sums (device)
    double* stat_out[kStatArrays] = {};   // means and stresses (device)
    bool have_stats = false, stat_final = false;
    int stat_start = 0, stat_samples = 0;
    unsigned char* vtk_snap = nullptr;    // device snapshot: u, v, w, p, nu_t, mask
    std::vector<unsigned char> vtk_host;  // host copy of the same
    hipStream_t io_stream = nullptr;      // the writer thread's copies
    hipEvent_t snap_ev = nullptr;
    std::thread io_thread;
    int io_rc = 0;                        // first failure of a background write
    std::string io_err;
    // kernel timing diagnostics
    int kt_class = 0;
    int kt_elem = 0;
    std::vector<hipEvent_t> kt_events;  // pairs (start, stop)
    size_t kt_used = 0;
    std::string err;
};

static std::string g_create_err;

namespace {

constexpr size_t kScalarSlots = 64;

int fail(les_ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    else g_create_err = msg;
    return code;
}

#define HIPCHK(ctx, expr)                                                                     \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(ctx, LES_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

size_t nbytes(const les_ctx* c, size_t elem) { return (size_t)c->d.N * elem; }

template <typename T>
int alloc(les_ctx* c, T** p, size_t count, bool zero = true)
{
    if (*p) return LES_OK;
    hipError_t e = hipMalloc((void**)p, count * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(c, LES_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    if (zero) {
        e = hipMemsetAsync(*p, 0, count * sizeof(T), c->stream);
        if (e != hipSuccess) return fail(c, LES_ERR_HIP, hipGetErrorString(e));
    }
    return LES_OK;
}

int ensure_krylov(les_ctx* c)
{
    int rc;
    double** bufs[] = {&c->r,  &c->z,  &c->p,  &c->p2,  &c->q, &c->za,
                       &c->zb, &c->pc, &c->pc2, &c->t1, &c->t2};
    for (double** b : bufs)
        if ((rc = alloc(c, b, c->d.N)) != LES_OK) return rc;
    return LES_OK;
}

int ensure_trace(les_ctx* c, int n)
{
    if (c->trace_cap >= n) return LES_OK;
    if (c->trace) hipFree(c->trace);
    c->trace = nullptr;
    c->trace_cap = 0;
    int rc = alloc(c, &c->trace, (size_t)n);
    if (rc == LES_OK) c->trace_cap = n;
    return rc;
}

hipMemcpyKind kind_in(int ptr_kind) { return ptr_kind ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice; }
hipMemcpyKind kind_out(int ptr_kind) { return ptr_kind ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost; }

Stencil<double> stencil_of(double* const a[7])
{
    Stencil<double> s{};
    for (int q = 0; q < 7; ++q) s.a[q] = a[q];
    return s;
}
Stencil<float> stencil_of(float* const a[7])
{
    Stencil<float> s{};
    for (int q = 0; q < 7; ++q) s.a[q] = a[q];
    return s;
}

// The coded form of the operator build_stencil makes (pressure.cpp:613-623):
// the base coefficients formed with the reference's expressions, then
// rounded to T.
template <typename T>
Stencil<T> coded_stencil(const les_ctx* c)
{
    Stencil<T> s{};
    const double idx2 = 1.0 / (c->dx * c->dx), idy2 = 1.0 / (c->dy * c->dy),
                 idz2 = 1.0 / (c->dz * c->dz);
    s.code = c->code;
    s.cfluid = (T)(2.0 * idx2 + 2.0 * idy2 + 2.0 * idz2);
    s.ox = (T)(-idx2);
    s.oy = (T)(-idy2);
    s.oz = (T)(-idz2);
    // a mirrored face's (-idx2 + -idx2) rounded to T is exactly 2 * s.ox:
    // the kernels form it as the multiplier times s.ox
    return s;
}

// stencil_format: 0 auto (coded when A came from build_stencil), 1 the 7
// arrays, 2 coded (error when A did not come from build_stencil)
int pick_coded(les_ctx* c, int fmt, bool& coded);

PhysConsts make_consts(const les_ctx* c, double dt, double nu, double rho, double Cs)
{
    PhysConsts k{};
    k.dt = dt;
    k.nu = nu;
    k.rho = rho;
    k.inv_dx2 = 1.0 / (c->dx * c->dx);
    k.inv_dy2 = 1.0 / (c->dy * c->dy);
    k.inv_dz2 = 1.0 / (c->dz * c->dz);
    k.inv_2dx = 1.0 / (2.0 * c->dx);
    k.inv_2dy = 1.0 / (2.0 * c->dy);
    k.inv_2dz = 1.0 / (2.0 * c->dz);
    k.two_dx = 2 * c->dx;
    k.two_dy = 2 * c->dy;
    k.two_dz = 2 * c->dz;
    const double delta = std::pow(c->dx * c->dy * c->dz, 1.0 / 3.0);
    k.csd2 = std::pow(Cs * delta, 2);
    k.div_scale = rho / dt;
    k.div_ix = 1.0 / (2.0 * c->dx);
    k.div_iy = 1.0 / (2.0 * c->dy);
    k.div_iz = 1.0 / (2.0 * c->dz);
    k.div_ix1 = 1.0 / c->dx;
    k.div_iy1 = 1.0 / c->dy;
    k.div_iz1 = 1.0 / c->dz;
    k.dor = dt / rho;
    return k;
}

// Chebyshev scalars (pressure.cpp:863-898): alpha_0, then (beta_k, alpha_k)
// for k = 1..degree-1, formed with the reference's expressions.
void cheb_coeffs(double lmin, double lmax, int degree, double& a0, std::vector<double>& beta,
                 std::vector<double>& alpha)
{
    const double d = (lmax + lmin) / 2.0;
    const double c = (lmax - lmin) / 2.0;
    double a = 2.0 / d;
    a0 = a;
    beta.clear();
    alpha.clear();
    for (int k = 1; k < degree; ++k) {
        const double b = (c * a / 2.0) * (c * a / 2.0);
        a = 1.0 / (d - b);
        beta.push_back(b);
        alpha.push_back(a);
    }
}

int last_level(const Dims& d) { return d.nx + d.ny + d.nz - 3; }


// Sizes the batches of PCG iterations the host issues between polls of the
// device's scalars.  Away from the end the host checks batch b-1 while batch
// b runs, so the GPU never waits for it; but a solve that has converged then
// still has up to two batches of iterations queued, whose launches exit at
// once and yet cost several microseconds of dispatch each (about 1.6 ms a
// solve for a multigrid iteration's 30 launches).  From the residuals seen
// at the polls the planner predicts the iteration that meets the tolerance:
// the batch that would pass it stops there and is waited for directly, and
// from then on single iterations are issued and waited for until the solve
// is done (one host round trip each, a few tens of microseconds).
struct BatchPlanner {
    int check;                      // iterations per batch away from the end
    long predict = 1L << 40;        // predicted iteration that meets the tolerance
    int seen_iter = -1;
    double seen_err = 0.0;
    bool tight = false;             // the batch just issued is waited for directly
    explicit BatchPlanner(int chk) : check(chk) {}
    int next(int issued, int left)
    {
        int n = std::min(check, left);
        tight = predict <= (long)issued + n;
        if (tight) n = (int)std::max<long>(1, std::min<long>(n, predict - issued));
        return n;
    }
    // a poll's state: iterations done, ||r||, the tolerance; issued so far
    void observe(int iter, double err, double tol, int issued)
    {
        if (seen_iter >= 0 && iter > seen_iter && err > 0.0 && err < seen_err && tol > 0.0) {
            const double rate = std::log(err / seen_err) / (iter - seen_iter);  // < 0
            const double m = std::max(0.0, std::log(tol / err) / rate);
            predict = iter + (long)std::ceil(std::min(m, 1e9));
        }
        if (tight) predict = std::max<long>(predict, (long)issued + 1);
        seen_iter = iter;
        seen_err = err;
    }
};

// Event bracket around the launches of the timed kernel class.
struct KtScope {
    les_ctx* c;
    bool on;
    hipEvent_t stop = nullptr;
    KtScope(les_ctx* ctx, int cls, int elem) : c(ctx), on(ctx->kt_class == cls)
    {
        if (!on) return;
        if (c->kt_used + 2 > c->kt_events.size()) {
            for (int q = 0; q < 64; ++q) {
                hipEvent_t e;
                if (hipEventCreate(&e) != hipSuccess) { on = false; return; }
                c->kt_events.push_back(e);
            }
        }
        c->kt_elem = elem;
        (void)hipEventRecord(c->kt_events[c->kt_used], c->stream);
        stop = c->kt_events[c->kt_used + 1];
        c->kt_used += 2;
    }
    ~KtScope()
    {
        if (on) (void)hipEventRecord(stop, c->stream);
    }
};

// Buffers of the wavefront sweeps: the ticket counters and the tiles' face
// packets (zeroed once; every launch writes packets with a new tag).
int ensure_wf(les_ctx* c, WfBufs* wb)
{
    int rc;
    if ((rc = alloc(c, &c->wf, 1)) != LES_OK) return rc;
    // at least one packet each: lanes with nothing to read load the first
    if ((rc = alloc(c, &c->wf_fx, std::max<size_t>(wf_face_packets_x(c->d), 2))) != LES_OK)
        return rc;
    if ((rc = alloc(c, &c->wf_fy, std::max<size_t>(wf_face_packets_y(c->d), 2))) != LES_OK)
        return rc;
    *wb = WfBufs{c->wf, c->wf_fx, c->wf_fy, c->spin_limit};
    c->wf_ran = true;
    return LES_OK;
}

// After the stream has drained: a wave that ran out of its spin budget left
// a flag (and results that are not to be trusted).  Clears it.
int wf_collect(les_ctx* c)
{
    if (!c->wf_ran) return LES_OK;
    c->wf_ran = false;
    int stalled = 0;
    HIPCHK(c, hipMemcpyAsync(&stalled, &c->wf->stalled, sizeof(int), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (!stalled) return LES_OK;
    HIPCHK(c, hipMemsetAsync(&c->wf->stalled, 0, sizeof(int), c->stream));
    return fail(c, LES_ERR_STALL,
                "wavefront sweep: a wave polled its neighbour tiles more than the spin limit "
                "(LES_OPT_SPIN_LIMIT) allows; the results of that launch are invalid");
}

// z = (L L^T)^-1 r by forward and backward substitution
// (apply_ic_preconditioner, pressure.cpp:442-482).
template <typename T>
int ic_solve(les_ctx* c, const T* r, T* z, const T* const* L, const SolveScalars* sc)
{
    const Dims& d = c->d;
    hipStream_t s = c->stream;
    KtScope kt(c, LES_KT_IC_SOLVE, (int)sizeof(T));
    if (c->wavefront) {
        WfBufs wb;
        int rc;
        if ((rc = ensure_wf(c, &wb)) != LES_OK) return rc;
        HIPCHK(c, launch_wf_ic_fwd<T>(s, d, r, z, L[SC], L[SW], L[SS], L[SB], sc, wb));
        HIPCHK(c, launch_wf_ic_bwd<T>(s, d, z, L[SC], L[SW], L[SS], L[SB], sc, wb));
        return LES_OK;
    }
    for (int l = 0; l <= last_level(d); ++l)
        HIPCHK(c, launch_ic_fwd_level<T>(s, d, l, r, z, L[SC], L[SW], L[SS], L[SB], sc));
    for (int l = last_level(d); l >= 0; --l)
        HIPCHK(c, launch_ic_bwd_level<T>(s, d, l, z, L[SC], L[SW], L[SS], L[SB], sc));
    return LES_OK;
}

double tridiag_eig(const std::vector<double>& a, const std::vector<double>& b, int which);

int pick_coded(les_ctx* c, int fmt, bool& coded)
{
    if (fmt == 2 && !c->have_code)
        return fail(c, LES_ERR_STATE, "coded stencil requested but A was not built from the mask");
    if (fmt < 0 || fmt > 2) return fail(c, LES_ERR_ARG, "unknown stencil format");
    coded = (fmt == 2) || (fmt == 0 && c->have_code);
    return LES_OK;
}

// The Chebyshev polynomial (pressure.cpp:851-906) as a chain of launches:
// fused pairs of steps while two or more remain (when the grid shape allows
// them), then single steps; the last launch reduces r.z with `epi`.  A
// reducing last step always runs as a single step, so the reduction sums in
// the same order with or without pairs and results are bit-identical.  z
// ping-pongs between za and zb, the recurrence's p between pa and pb.
template <typename T>
int cheb_chain(les_ctx* c, const Stencil<T>& A, const T* r, T* zout, T* za, T* zb, T* pa, T* pb,
               double lmin, double lmax, int degree, int epi, SolveScalars* sc, const double* wgt)
{
    const Dims& d = c->d;
    hipStream_t s = c->stream;
    double a0;
    std::vector<double> be, al;
    cheb_coeffs(lmin, lmax, degree, a0, be, al);
    const int elem = (int)sizeof(T);
    if (degree <= 1) {
        HIPCHK(c, launch_cheb_scale<T>(s, d, r, zout, a0, epi, c->red, sc, wgt));
        return LES_OK;
    }
    const bool pairs = c->cheb_pairs && cheb_pair_supported<T>(d, c->pair_grid, A.code != nullptr);
    const T* zin = nullptr;
    T* zbuf[2] = {za, zb};
    int nz = 0;
    T* pin = pa;
    T* pout = pb;
    for (int k = 1; k < degree;) {
        // a pair never runs the reducing last step
        const bool pair = pairs && degree - k >= 2 && !(k + 1 == degree - 1 && epi >= 0);
        const int klast = pair ? k + 1 : k;
        const bool last = klast == degree - 1;
        T* zo = last ? zout : zbuf[nz];
        nz ^= 1;
        // the timing brackets leave out the reducing launch (it is followed
        // by its finalize kernel)
        const bool timed = !(last && epi >= 0);
        if (pair) {
            KtScope kt(c, timed ? LES_KT_CHEB_PAIR : -1, elem);
            HIPCHK(c, launch_cheb_pair<T>(s, d, A, r, zin, zo, pin, pout, a0, be[k - 1],
                                          al[k - 1], be[k], al[k], k == 1, sc, c->pair_grid));
            std::swap(pin, pout);
        } else {
            KtScope kt(c, timed ? LES_KT_CHEB_STEP : -1, elem);
            HIPCHK(c, launch_cheb_step<T>(s, d, A, r, zin, zo, pin, a0, be[k - 1], al[k - 1],
                                          k == 1, last ? epi : -1, c->red, sc, wgt));
        }
        zin = zo;
        k = klast + 1;
    }
    return LES_OK;
}

// z = M^-1 r for FP64 preconditioners.  When epi >= 0 the final kernel also
// reduces r.z into the solve scalars with that epilogue.
// The multigrid hierarchy of the current A: level dimensions, the P^T A P
// operators in FP64 (and FP32 for mixed solves), and each level's x, t, b.
// W (the mirror weights, or null): level 1 is P^T W A P, for solves in the
// W inner product.
int mg_setup(les_ctx* c, bool fp32, const double* W)
{
    les_ctx::Mg& g = c->mg;
    hipStream_t s = c->stream;
    int rc;
    if (g.built && g.weighted != (W != nullptr)) g.built = false;
    if (!g.built) {
        g.nlev = 1;
        g.dims[0] = c->d;
        while (g.nlev < kMgMaxLevels) {
            const Dims& p = g.dims[g.nlev - 1];
            if (std::max(p.nx, std::max(p.ny, p.nz)) <= kMgCoarseMax) break;
            Dims q{};
            q.nx = (p.nx + 1) / 2;
            q.ny = (p.ny + 1) / 2;
            q.nz = (p.nz + 1) / 2;
            q.sij = q.ny * q.nz;
            q.N = q.nx * q.sij;
            g.dims[g.nlev++] = q;
        }
        if (g.dims[g.nlev - 1].N > kBlock)
            return fail(c, LES_ERR_ARG, "multigrid: the coarsest level exceeds one workgroup");
        for (int l = 1; l < g.nlev; ++l) {
            for (int q = 0; q < 7; ++q)
                if ((rc = alloc(c, &g.A64[l][q], g.dims[l].N, false)) != LES_OK) return rc;
            HIPCHK(c, launch_mg_coarsen(s, g.dims[l - 1], g.dims[l],
                                        stencil_of(l == 1 ? c->A : g.A64[l - 1]),
                                        l == 1 ? W : nullptr, g.A64[l]));
        }
        g.built = true;
        g.weighted = W != nullptr;
        g.have32 = false;
    }
    for (int l = 1; l < g.nlev; ++l)
        for (int q = 0; q < 3; ++q) {
            if (fp32) {
                if ((rc = alloc(c, &g.v32[l][q], g.dims[l].N)) != LES_OK) return rc;
            } else {
                if ((rc = alloc(c, &g.v64[l][q], g.dims[l].N)) != LES_OK) return rc;
            }
        }
    if (fp32 && !g.have32) {
        for (int l = 1; l < g.nlev; ++l)
            for (int q = 0; q < 7; ++q) {
                if ((rc = alloc(c, &g.A32[l][q], g.dims[l].N, false)) != LES_OK) return rc;
                HIPCHK(c, (launch_convert<float, double>(s, g.dims[l].N, g.A64[l][q],
                                                         g.A32[l][q])));
            }
        g.have32 = true;
    }
    return LES_OK;
}

template <typename T>
T* mg_vec(les_ctx::Mg& g, int l, int q);
template <>
double* mg_vec<double>(les_ctx::Mg& g, int l, int q) { return g.v64[l][q]; }
template <>
float* mg_vec<float>(les_ctx::Mg& g, int l, int q) { return g.v32[l][q]; }
template <typename T>
T* const* mg_arr(les_ctx::Mg& g, int l);
template <>
double* const* mg_arr<double>(les_ctx::Mg& g, int l) { return g.A64[l]; }
template <>
float* const* mg_arr<float>(les_ctx::Mg& g, int l) { return g.A32[l]; }
template <typename T>
Stencil<T> mg_op(les_ctx::Mg& g, int l);
template <>
Stencil<double> mg_op<double>(les_ctx::Mg& g, int l) { return stencil_of(g.A64[l]); }
template <>
Stencil<float> mg_op<float>(les_ctx::Mg& g, int l) { return stencil_of(g.A32[l]); }

// z = one V-cycle on r (A0 the level-0 operator, arrays or coded; t0 a
// level-0 work vector distinct from r and z).  Each level's 2 nu sweeps
// alternate between two buffers, the first writing the one that is not the
// level's result, so the last writes z on level 0.  epi >= 0: the last
// sweep also reduces r.z with that epilogue.  W: the hierarchy's mirror
// weights (the level-0 restriction and r.z are weighted).
template <typename T>
int mg_vcycle(les_ctx* c, const Stencil<T>& A0, const T* r, T* z, T* t0, int epi,
              SolveScalars* sc, const double* W)
{
    les_ctx::Mg& g = c->mg;
    hipStream_t s = c->stream;
    const int L = g.nlev, nu = c->mg_sweeps;
    auto op = [&](int l) { return l == 0 ? A0 : mg_op<T>(g, l); };
    // the two buffers of level l (the result lands in [1]) and its b
    auto buf = [&](int l, int k) -> T* { return l == 0 ? (k ? z : t0) : mg_vec<T>(g, l, k); };
    auto bvec = [&](int l) -> const T* { return l == 0 ? r : mg_vec<T>(g, l, 2); };
    if (L == 1) {  // no coarser level: the coarsest sweeps on the grid itself
        const int sw = c->mg_coarse_sweeps;
        T* cur = nullptr;
        for (int k = 1; k <= sw; ++k) {
            T* out = buf(0, (k + sw + 1) & 1);  // the last (k = sw) writes z
            HIPCHK(c, launch_mg_smooth<T>(s, g.dims[0], A0, cur, r, nullptr, g.dims[0], out,
                                          k == 1 ? MG_ZERO : MG_SWEEP, k == sw ? epi : -1,
                                          kMgOmega, kMgScale, c->red, sc, W));
            cur = out;
        }
        return LES_OK;
    }
    // the bottom levels (from lt on, each small enough) run in one launch
    int lt = L - 1;
    while (lt > 1 && g.dims[lt - 1].N <= kMgTailCells && L - (lt - 1) <= kMgTailLevels) --lt;
    const bool tail = LES_MG_TAIL && g.dims[lt].N <= kMgTailCells && L - lt <= kMgTailLevels;
    if (!tail) lt = L - 1;
    T* cur[kMgMaxLevels] = {};
    // levels whose two post-smoothing sweeps run as one pass: their
    // pre-smoothing lands in buf(l, 0), so that pass writes buf(l, 1)
    bool post[kMgMaxLevels] = {};
    for (int l = 0; l < lt; ++l) post[l] = c->mg_rings && nu == 2 && mg_post_fits<T>(g.dims[l]);
    for (int l = 0; l < lt; ++l) {
        // the two sweeps and the restriction in one pass pay with the seven
        // arrays (the coarse levels, and level 0 of the arrays format); with
        // the code the separate passes are faster at 256^3 (49 + 47 against
        // 110 us); the level-0 launch is the timed class either way
        if (c->mg_rings && nu == 2 && !op(l).code && mg_pre_fits<T>(g.dims[l])) {
            T* out = buf(l, post[l] ? 0 : 1);
            KtScope kt(c, l == 0 ? LES_KT_MG_SWEEP : -1, (int)sizeof(T));
            HIPCHK(c, launch_mg_pre<T>(s, g.dims[l], g.dims[l + 1], op(l), bvec(l), out,
                                       mg_vec<T>(g, l + 1, 2), kMgOmega, sc,
                                       l == 0 ? W : nullptr));
            cur[l] = out;
            continue;
        }
        int k = 1;
        if (nu >= 2) {
            // the first two sweeps in one pass (the other buffer holds the
            // first where the level runs the scalar kernels); the level-0
            // one is the timed class
            T* out = buf(l, post[l] ? 0 : 1);
            KtScope kt(c, l == 0 ? LES_KT_MG_SWEEP : -1, (int)sizeof(T));
            HIPCHK(c, launch_mg_smooth<T>(s, g.dims[l], op(l), buf(l, post[l] ? 1 : 0), bvec(l),
                                          nullptr, g.dims[l], out, MG_ZS, -1, kMgOmega, kMgScale,
                                          c->red, sc, nullptr, c->mg_rings));
            cur[l] = out;
            k = 3;
        }
        for (; k <= nu; ++k) {
            T* out = buf(l, (k + 1) & 1);
            HIPCHK(c, launch_mg_smooth<T>(s, g.dims[l], op(l), cur[l], bvec(l), nullptr, g.dims[l],
                                          out, k == 1 ? MG_ZERO : MG_SWEEP, -1, kMgOmega,
                                          kMgScale, c->red, sc));
            cur[l] = out;
        }
        // the restriction ring pays on the coarse levels (9 against 13 us at
        // 128^3 .. 16^3) and not on level 0 (60 against 47 us at 256^3)
        HIPCHK(c, launch_mg_restrict<T>(s, g.dims[l], g.dims[l + 1], op(l), cur[l], bvec(l),
                                        mg_vec<T>(g, l + 1, 2), sc, l == 0 ? W : nullptr,
                                        c->mg_rings && l > 0));
    }
    cur[lt] = buf(lt, 1);
    if (tail) {
        T* const* Al[kMgTailLevels];
        for (int l = lt; l < L; ++l) Al[l - lt] = mg_arr<T>(g, l);
        HIPCHK(c, launch_mg_tail<T>(s, L - lt, g.dims + lt, Al, bvec(lt), cur[lt], nu,
                                    c->mg_coarse_sweeps, kMgOmega, kMgScale, sc));
    } else {
        HIPCHK(c, launch_mg_coarsest<T>(s, g.dims[L - 1], op(L - 1), bvec(L - 1), cur[L - 1],
                                        c->mg_coarse_sweeps, kMgOmega, sc));
    }
    for (int l = lt - 1; l >= 0; --l) {
        if (post[l]) {
            HIPCHK(c, launch_mg_post<T>(s, g.dims[l], op(l), cur[l], bvec(l), cur[l + 1],
                                        g.dims[l + 1], buf(l, 1), l == 0 ? epi : -1, kMgOmega,
                                        kMgScale, c->red, sc, l == 0 ? W : nullptr));
            cur[l] = buf(l, 1);
            continue;
        }
        for (int k = nu + 1; k <= 2 * nu; ++k) {
            T* out = buf(l, (k + 1) & 1);
            const bool last = l == 0 && k == 2 * nu;
            HIPCHK(c, launch_mg_smooth<T>(s, g.dims[l], op(l), cur[l], bvec(l), cur[l + 1],
                                          g.dims[l + 1], out, k == nu + 1 ? MG_PROLONG : MG_SWEEP,
                                          last ? epi : -1, kMgOmega, kMgScale, c->red, sc,
                                          last ? W : nullptr));
            cur[l] = out;
        }
    }
    return LES_OK;
}

int precond_fp64(les_ctx* c, int pc, const double* r, double* zout, double lmin, double lmax,
                 int degree, int epi, SolveScalars* sc, const Stencil<double>& Aop,
                 const double* wgt = nullptr)
{
    const Dims& d = c->d;
    hipStream_t s = c->stream;
    switch (pc) {
    case LES_PC_NONE:
        HIPCHK(c, hipMemcpyAsync(zout, r, nbytes(c, 8), hipMemcpyDeviceToDevice, s));
        break;
    case LES_PC_JACOBI:
        HIPCHK(c, launch_pcg_p_jacobi<double>(s, d, r, c->A[SC], zout, 1, epi, c->red, sc, wgt));
        return LES_OK;
    case LES_PC_IC0:
    case LES_PC_LEGACY_IC_A: {
        double* const* L = (pc == LES_PC_IC0) ? c->L : c->A;
        if (pc == LES_PC_IC0 && !c->have_L) return fail(c, LES_ERR_STATE, "IC(0) factor not built");
        int rc;
        if ((rc = ic_solve<double>(c, r, zout, L, sc)) != LES_OK) return rc;
        break;
    }
    case LES_PC_CHEBYSHEV:
        return cheb_chain<double>(c, Aop, r, zout, c->za, c->zb, c->pc, c->pc2, lmin, lmax,
                                  degree, epi, sc, wgt);
    case LES_PC_MG: {
        int rc;
        if ((rc = mg_setup(c, false, wgt)) != LES_OK) return rc;
        return mg_vcycle<double>(c, Aop, r, zout, c->za, epi, sc, wgt);
    }
    default:
        return fail(c, LES_ERR_ARG, "unknown preconditioner");
    }
    if (epi >= 0) HIPCHK(c, launch_pcg_rz<double>(s, d, r, zout, epi, c->red, sc, wgt));
    return LES_OK;
}

}  // namespace

// ====================================================================
// C ABI
// ====================================================================
extern "C" {

int les_ctx_create(int nx, int ny, int nz, double dx, double dy, double dz, les_ctx** out)
{
    if (!out) return fail(nullptr, LES_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (nx < 3 || ny < 3 || nz < 3)
        return fail(nullptr, LES_ERR_ARG, "grid must be at least 3 cells in each direction");
    const long long N = (long long)nx * ny * nz;
    if (N >= (1LL << 31)) return fail(nullptr, LES_ERR_ARG, "grid exceeds 2^31 cells (int indices)");
    if (!(dx > 0 && dy > 0 && dz > 0)) return fail(nullptr, LES_ERR_ARG, "spacings must be > 0");
    les_ctx* c = new (std::nothrow) les_ctx();
    if (!c) return fail(nullptr, LES_ERR_NOMEM, "host allocation failed");
    c->d = Dims{nx, ny, nz, ny * nz, (int)N};
    c->dx = dx;
    c->dy = dy;
    c->dz = dz;
    auto bail = [&](hipError_t e, const char* what) {
        g_create_err = std::string(what) + ": " + hipGetErrorString(e);
        les_ctx_destroy(c);
        return LES_ERR_HIP;
    };
    hipError_t e;
    if ((e = hipGetDevice(&c->device)) != hipSuccess) return bail(e, "hipGetDevice");
    if ((e = hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking)) != hipSuccess)
        return bail(e, "hipStreamCreate");
    c->stream = c->own;
    for (int q = 0; q < LES_F_COUNT; ++q)
        if (alloc(c, &c->f[q], N) != LES_OK) { g_create_err = c->err; les_ctx_destroy(c); return LES_ERR_NOMEM; }
    if (alloc(c, &c->solid, N) != LES_OK) { g_create_err = c->err; les_ctx_destroy(c); return LES_ERR_NOMEM; }
    if (alloc(c, &c->red.partials, (size_t)kMaxGrid * kMaxRed) != LES_OK ||
        alloc(c, &c->red2.partials, (size_t)kMaxGrid * kMaxRed) != LES_OK ||
        alloc(c, &c->red3.partials, (size_t)kMaxGrid * kMaxRed) != LES_OK ||
        alloc(c, &c->sc, 1) != LES_OK ||
        alloc(c, &c->st, 1) != LES_OK || alloc(c, &c->dsc, kScalarSlots) != LES_OK) {
        g_create_err = c->err;
        les_ctx_destroy(c);
        return LES_ERR_NOMEM;
    }
    if ((e = hipHostMalloc((void**)&c->sc_host, 2 * sizeof(SolveScalars), 0)) != hipSuccess)
        return bail(e, "hipHostMalloc");
    if ((e = hipHostMalloc((void**)&c->st_host, 2 * sizeof(StatScalars), 0)) != hipSuccess)
        return bail(e, "hipHostMalloc");
    if ((e = hipHostMalloc((void**)&c->hsc, kScalarSlots * sizeof(double), 0)) != hipSuccess)
        return bail(e, "hipHostMalloc");
    for (int q = 0; q < 2; ++q)
        if ((e = hipEventCreateWithFlags(&c->ev[q], hipEventDisableTiming)) != hipSuccess)
            return bail(e, "hipEventCreate");
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return bail(e, "hipStreamSynchronize");
    *out = c;
    return LES_OK;
}

int les_ctx_destroy(les_ctx* c)
{
    if (!c) return LES_OK;
    if (c->io_thread.joinable()) c->io_thread.join();
    if (c->stream) hipStreamSynchronize(c->stream);
    auto fr = [](void* p) {
        if (p) hipFree(p);
    };
    for (auto* p : c->f) fr(p);
    fr(c->solid);
    fr(c->code);
    for (int q = 0; q < 7; ++q) {
        fr(c->A[q]);
        fr(c->L[q]);
        fr(c->A32[q]);
        fr(c->L32[q]);
    }
    for (double* p : {c->r, c->z, c->p, c->p2, c->q, c->za, c->zb, c->pc, c->pc2, c->t1, c->t2})
        fr(p);
    for (int l = 0; l < kMgMaxLevels; ++l) {
        for (int q = 0; q < 7; ++q) {
            fr(c->mg.A64[l][q]);
            fr(c->mg.A32[l][q]);
        }
        for (int q = 0; q < 3; ++q) {
            fr(c->mg.v64[l][q]);
            fr(c->mg.v32[l][q]);
        }
    }
    fr(c->W);
    fr(c->ckpt);
    fr(c->seg);
    for (auto* p : c->adjb) fr(p);
    for (auto* p : c->shadow) fr(p);
    free_mixed(c->mw);
    fr(c->red.partials);
    fr(c->red2.partials);
    fr(c->red3.partials);
    fr(c->sc);
    fr(c->st);
    fr(c->dsc);
    fr(c->trace);
    fr(c->wf);
    fr(c->wf_fx);
    fr(c->wf_fy);
    if (c->sc_host) hipHostFree(c->sc_host);
    if (c->st_host) hipHostFree(c->st_host);
    if (c->hsc) hipHostFree(c->hsc);
    for (auto ev : c->ev)
        if (ev) hipEventDestroy(ev);
    for (auto ev : c->kt_events) hipEventDestroy(ev);
    for (int q = 0; q < 2; ++q) {
        if (c->io_stage[q]) hipHostFree(c->io_stage[q]);
        if (c->io_ev[q]) hipEventDestroy(c->io_ev[q]);
    }
    for (int a = 0; a < kStatArrays; ++a) {
        fr(c->stat_sum[a]);
        fr(c->stat_out[a]);
    }
    fr(c->vtk_snap);
    if (c->snap_ev) hipEventDestroy(c->snap_ev);
    if (c->io_stream) hipStreamDestroy(c->io_stream);
    if (c->own) hipStreamDestroy(c->own);
    delete c;
    return LES_OK;
}

const char* les_last_error(const les_ctx* c) { return c ? c->err.c_str() : g_create_err.c_str(); }

int les_set_stream(les_ctx* c, void* stream)
{
    if (!c) return LES_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->stream = stream ? (hipStream_t)stream : c->own;
    return LES_OK;
}

int les_set_option(les_ctx* c, int option, int value)
{
    if (!c) return LES_ERR_ARG;
    switch (option) {
    case LES_OPT_CHEB_PAIRS:
        c->cheb_pairs = value != 0;
        return LES_OK;
    case LES_OPT_PAIR_GRID:
        if (value < 0 || value > kPairMaxGrid) return fail(c, LES_ERR_ARG, "pair grid out of range");
        c->pair_grid = value;
        return LES_OK;
    case LES_OPT_WAVEFRONT:
        c->wavefront = value != 0;
        return LES_OK;
    case LES_OPT_SPIN_LIMIT:
        if (value < 0) return fail(c, LES_ERR_ARG, "spin limit < 0");
        c->spin_limit = value;
        return LES_OK;
    case LES_OPT_MG_SWEEPS:
        if (value < 1 || value > 16) return fail(c, LES_ERR_ARG, "multigrid sweeps out of 1..16");
        c->mg_sweeps = value;
        return LES_OK;
    case LES_OPT_MG_COARSE_SWEEPS:
        if (value < 1 || value > 4096)
            return fail(c, LES_ERR_ARG, "multigrid coarsest sweeps out of 1..4096");
        c->mg_coarse_sweeps = value;
        return LES_OK;
    case LES_OPT_MG_RINGS:
        c->mg_rings = value != 0;
        return LES_OK;
    case LES_OPT_REPLACE_DECADES:
        if (value < 0 || value > 15) return fail(c, LES_ERR_ARG, "replacement decades out of 0..15");
        c->replace_decades = value;
        return LES_OK;
    default:
        return fail(c, LES_ERR_ARG, "unknown option");
    }
}

int les_synchronize(les_ctx* c)
{
    if (!c) return LES_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LES_OK;
}

// ---------------- fields ----------------
int les_set_field(les_ctx* c, int field, const double* src, int ptr_kind)
{
    if (!c || !src || field < 0 || field >= LES_F_COUNT) return fail(c, LES_ERR_ARG, "bad field/src");
    HIPCHK(c, hipMemcpyAsync(c->f[field], src, nbytes(c, 8), kind_in(ptr_kind), c->stream));
    if (!ptr_kind) HIPCHK(c, hipStreamSynchronize(c->stream));
    return LES_OK;
}

int les_get_field(les_ctx* c, int field, double* dst, int ptr_kind)
{
    if (!c || !dst || field < 0 || field >= LES_F_COUNT) return fail(c, LES_ERR_ARG, "bad field/dst");
    HIPCHK(c, hipMemcpyAsync(dst, c->f[field], nbytes(c, 8), kind_out(ptr_kind), c->stream));
    if (!ptr_kind) HIPCHK(c, hipStreamSynchronize(c->stream));
    return LES_OK;
}

int les_field_ptr(les_ctx* c, int field, double** dev_ptr)
{
    if (!c || !dev_ptr || field < 0 || field >= LES_F_COUNT) return fail(c, LES_ERR_ARG, "bad field");
    *dev_ptr = c->f[field];
    return LES_OK;
}

int les_set_mask(les_ctx* c, const uint8_t* solid, int ptr_kind)
{
    if (!c || !solid) return fail(c, LES_ERR_ARG, "mask is NULL");
    HIPCHK(c, hipMemcpyAsync(c->solid, solid, nbytes(c, 1), kind_in(ptr_kind), c->stream));
    if (!ptr_kind) HIPCHK(c, hipStreamSynchronize(c->stream));
    c->have_roll = false;
    return LES_OK;
}

int les_get_mask(les_ctx* c, uint8_t* solid, int ptr_kind)
{
    if (!c || !solid) return fail(c, LES_ERR_ARG, "mask is NULL");
    HIPCHK(c, hipMemcpyAsync(solid, c->solid, nbytes(c, 1), kind_out(ptr_kind), c->stream));
    if (!ptr_kind) HIPCHK(c, hipStreamSynchronize(c->stream));
    return LES_OK;
}

// ---------------- operator ----------------
int les_build_stencil(les_ctx* c, int ignore_mask)
{
    if (!c) return LES_ERR_ARG;
    int rc;
    for (int q = 0; q < 7; ++q)
        if ((rc = alloc(c, &c->A[q], c->d.N, false)) != LES_OK) return rc;
    const uint8_t* m = c->solid;
    uint8_t* zeros = nullptr;
    if (ignore_mask) {
        if ((rc = alloc(c, &zeros, c->d.N)) != LES_OK) return rc;
        m = zeros;
    }
    HIPCHK(c, launch_build_stencil(c->stream, c->d, m, c->A, 1.0 / (c->dx * c->dx),
                                   1.0 / (c->dy * c->dy), 1.0 / (c->dz * c->dz)));
    if ((rc = alloc(c, &c->code, c->d.N, false)) != LES_OK) return rc;
    HIPCHK(c, launch_build_code(c->stream, c->d, m, c->code));
    if (zeros) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        hipFree(zeros);
    }
    c->have_A = true;
    c->have_code = true;
    c->A_from_mask = !ignore_mask;
    c->have_A32 = false;
    c->mg.built = false;
    c->have_L = false;
    c->have_L32 = false;
    c->have_W = false;
    c->have_roll = false;
    c->A_has_obstacles = false;
    if (!ignore_mask) {
        // whether A has mirror rows around interior obstacles (then no
        // diagonal W symmetrises it)
        std::vector<uint8_t> h((size_t)c->d.N);
        HIPCHK(c, hipMemcpyAsync(h.data(), m, h.size(), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        c->A_has_obstacles = std::any_of(h.begin(), h.end(), [](uint8_t v) { return v != 0; });
    }
    return LES_OK;
}

int les_set_stencil(les_ctx* c, int which, double* const coef[7], int ptr_kind)
{
    if (!c || !coef || (which != LES_STENCIL_A && which != LES_STENCIL_L))
        return fail(c, LES_ERR_ARG, "bad stencil arguments");
    double** dst = which == LES_STENCIL_A ? c->A : c->L;
    int rc;
    for (int q = 0; q < 7; ++q) {
        if ((rc = alloc(c, &dst[q], c->d.N, false)) != LES_OK) return rc;
        if (!coef[q]) return fail(c, LES_ERR_ARG, "stencil array is NULL");
        HIPCHK(c, hipMemcpyAsync(dst[q], coef[q], nbytes(c, 8), kind_in(ptr_kind), c->stream));
    }
    if (!ptr_kind) HIPCHK(c, hipStreamSynchronize(c->stream));
    if (which == LES_STENCIL_A) {
        c->have_A = true;
        c->have_code = false;
        c->A_from_mask = false;
        c->have_A32 = false;
        c->mg.built = false;
        c->have_W = false;
        c->have_roll = false;
    } else {
        c->have_L = true;
        c->have_L32 = false;
    }
    return LES_OK;
}

int les_get_stencil(les_ctx* c, int which, double* const coef[7], int ptr_kind)
{
    if (!c || !coef) return fail(c, LES_ERR_ARG, "bad stencil arguments");
    const bool isA = which == LES_STENCIL_A;
    if ((isA && !c->have_A) || (!isA && !c->have_L)) return fail(c, LES_ERR_STATE, "stencil not set");
    double* const* src = isA ? c->A : c->L;
    for (int q = 0; q < 7; ++q)
        HIPCHK(c, hipMemcpyAsync(coef[q], src[q], nbytes(c, 8), kind_out(ptr_kind), c->stream));
    if (!ptr_kind) HIPCHK(c, hipStreamSynchronize(c->stream));
    return LES_OK;
}

int les_matvec(les_ctx* c, int which, const double* x, double* y, int ptr_kind)
{
    if (!c || !x || !y) return fail(c, LES_ERR_ARG, "NULL vector");
    const bool isA = which == LES_STENCIL_A;
    if ((isA && !c->have_A) || (!isA && !c->have_L)) return fail(c, LES_ERR_STATE, "stencil not set");
    int rc;
    if ((rc = ensure_krylov(c)) != LES_OK) return rc;
    const double* xd = x;
    double* yd = y;
    if (!ptr_kind) {
        HIPCHK(c, hipMemcpyAsync(c->t1, x, nbytes(c, 8), hipMemcpyHostToDevice, c->stream));
        xd = c->t1;
        yd = c->t2;
    }
    HIPCHK(c, launch_matvec<double>(c->stream, c->d, stencil_of(isA ? c->A : c->L), xd, yd));
    if (!ptr_kind) {
        HIPCHK(c, hipMemcpyAsync(y, yd, nbytes(c, 8), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return LES_OK;
}

int les_dot(les_ctx* c, const double* a, const double* b, int ptr_kind, double* out)
{
    if (!c || !a || !b || !out) return fail(c, LES_ERR_ARG, "NULL argument");
    int rc;
    if ((rc = ensure_krylov(c)) != LES_OK) return rc;
    const double *ad = a, *bd = b;
    if (!ptr_kind) {
        HIPCHK(c, hipMemcpyAsync(c->t1, a, nbytes(c, 8), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->t2, b, nbytes(c, 8), hipMemcpyHostToDevice, c->stream));
        ad = c->t1;
        bd = c->t2;
    }
    HIPCHK(c, launch_dot(c->stream, c->d, ad, bd, c->red, c->dsc));
    HIPCHK(c, hipMemcpyAsync(c->hsc, c->dsc, 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *out = c->hsc[0];
    return LES_OK;
}

}  // extern "C"

// ---------------- parallel_dot without allocation ----------------
// Partial sums of les_dot_noalloc's reduction: one row of kMaxRed per
// workgroup, kMaxGrid rows (8 KiB).  A file-scope device array, so the
// entry point needs no allocation; each GPU holds its own copy for the life
// of the process.  Every call on a GPU uses that GPU's copy: calls that
// overlap on one GPU (different streams or threads) race on it.  Each call
// writes every partial row its finalize reads, so nothing is cleared.
static __device__ __align__(256) double scratch[kMaxGrid * kMaxRed];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// The device address of the current device's copy of `scratch`, looked up
// once per device (the array's name is a host-side stand-in, not a device
// address).
int scratch_of_current_device(double** out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(nullptr, LES_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return fail(nullptr, LES_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(nullptr, LES_ERR_HIP,
                        std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return LES_OK;
}

}  // namespace

extern "C" {

int les_dot_noalloc(int nx, int ny, int nz, const double* a, const double* b, double* out,
                    void* stream)
{
    if (!a || !b || !out) return fail(nullptr, LES_ERR_ARG, "NULL argument");
    if (nx < 1 || ny < 1 || nz < 1) return fail(nullptr, LES_ERR_ARG, "grid dimensions must be >= 1");
    const long long n = (long long)nx * ny * nz;
    if (n > 2147483647LL)
        return fail(nullptr, LES_ERR_ARG,
                    "grid of more than 2^31-1 cells: cell indices are 32-bit (field.h:40)");
    double* scr = nullptr;
    int rc;
    if ((rc = scratch_of_current_device(&scr)) != LES_OK) return rc;
    const Dims d{nx, ny, nz, ny * nz, (int)n};
    // the reduction's grid never exceeds kMaxGrid, the rows scratch holds
    if (make_chunking(d).grid > kMaxGrid) return fail(nullptr, LES_ERR_STATE, "reduction grid too large");
    const hipError_t e = launch_dot(static_cast<hipStream_t>(stream), d, a, b, RedScratch{scr}, out);
    if (e != hipSuccess) return fail(nullptr, LES_ERR_HIP, std::string("les_dot_noalloc: ") + hipGetErrorString(e));
    return LES_OK;
}

// ---------------- preconditioners ----------------
int les_ic0_factor(les_ctx* c)
{
    if (!c) return LES_ERR_ARG;
    if (!c->have_A) return fail(c, LES_ERR_STATE, "stencil A not built");
    int rc;
    for (int q = 0; q < 7; ++q) {
        if ((rc = alloc(c, &c->L[q], c->d.N, false)) != LES_OK) return rc;
        // StencilMatrix(n) starts at zero; guarded entries keep that value
        HIPCHK(c, hipMemsetAsync(c->L[q], 0, nbytes(c, 8), c->stream));
    }
    const Stencil<double> A = stencil_of(c->A);
    {
        KtScope kt(c, LES_KT_IC_FACTOR, 8);
        if (c->wavefront) {
            WfBufs wb;
            if ((rc = ensure_wf(c, &wb)) != LES_OK) return rc;
            HIPCHK(c, launch_wf_ic_factor(c->stream, c->d, A, c->L[SC], c->L[SW], c->L[SS],
                                          c->L[SB], wb));
        } else {
            for (int l = 0; l <= last_level(c->d); ++l)
                HIPCHK(c, launch_ic_factor_level(c->stream, c->d, l, A, c->L[SC], c->L[SW],
                                                 c->L[SS], c->L[SB]));
        }
    }
    c->have_L = false;
    if ((rc = wf_collect(c)) != LES_OK) return rc;
    c->have_L = true;
    c->have_L32 = false;
    return LES_OK;
}

int les_precond_apply(les_ctx* c, int pc, const double* r, double* z, int ptr_kind, double lmin,
                      double lmax, int degree)
{
    if (!c || !r || !z) return fail(c, LES_ERR_ARG, "NULL vector");
    if (!c->have_A) return fail(c, LES_ERR_STATE, "stencil A not built");
    int rc;
    if ((rc = ensure_krylov(c)) != LES_OK) return rc;
    const double* rd = r;
    double* zd = z;
    // host vectors, and device vectors the 16-byte vector loads cannot
    // address, go through the context's own buffers
    const bool stage = !ptr_kind || (((uintptr_t)r | (uintptr_t)z) % 16) != 0;
    if (stage) {
        HIPCHK(c, hipMemcpyAsync(c->t1, r, nbytes(c, 8), kind_in(ptr_kind), c->stream));
        rd = c->t1;
        zd = c->t2;
    }
    if ((rc = precond_fp64(c, pc, rd, zd, lmin, lmax, degree, -1, nullptr, stencil_of(c->A))) !=
        LES_OK)
        return rc;
    if (stage) {
        HIPCHK(c, hipMemcpyAsync(z, zd, nbytes(c, 8), kind_out(ptr_kind), c->stream));
        if (!ptr_kind) HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return wf_collect(c);
}

int les_power_lambda_max(les_ctx* c, int n_iter, double* out)
{
    if (!c || !out) return fail(c, LES_ERR_ARG, "NULL argument");
    if (!c->have_A) return fail(c, LES_ERR_STATE, "stencil A not built");
    int rc;
    if ((rc = ensure_krylov(c)) != LES_OK) return rc;
    // start vector: pressure.cpp:926-931, drawn in sequence
    std::vector<double> v0(c->d.N);
    std::mt19937 gen(1337);
    std::uniform_real_distribution<> dis(-1.0, 1.0);
    for (int i = 0; i < c->d.N; ++i) v0[i] = dis(gen);
    double *v = c->t1, *w = c->t2;
    HIPCHK(c, hipMemcpyAsync(v, v0.data(), nbytes(c, 8), hipMemcpyHostToDevice, c->stream));
    const Stencil<double> A = stencil_of(c->A);
    for (int k = 0; k < n_iter; ++k) {
        HIPCHK(c, launch_matvec<double>(c->stream, c->d, A, v, w));
        HIPCHK(c, launch_dot(c->stream, c->d, v, w, c->red, c->dsc + 0));
        HIPCHK(c, launch_dot(c->stream, c->d, w, w, c->red, c->dsc + 1));
        HIPCHK(c, launch_scale_by_invnorm(c->stream, c->d.N, w, v, c->dsc + 1));
    }
    HIPCHK(c, hipMemcpyAsync(c->hsc, c->dsc, 16, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *out = n_iter > 0 ? c->hsc[0] : 0.0;
    return LES_OK;
}


int les_lanczos_bounds(les_ctx* c, int steps, uint64_t seed, double* ritz_min, double* ritz_max)
{
    if (!c || !ritz_min || !ritz_max || steps < 1) return fail(c, LES_ERR_ARG, "bad arguments");
    if (!c->have_A) return fail(c, LES_ERR_STATE, "stencil A not built");
    int rc;
    if ((rc = ensure_krylov(c)) != LES_OK) return rc;
    const int N = c->d.N;
    // seeded start vector (SplitMix64 per cell), normalised on the device
    std::vector<double> v0(N);
    for (int i = 0; i < N; ++i) {
        uint64_t z = seed * 0x9E3779B97F4A7C15ull + (uint64_t)i + 1;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        v0[i] = (double)(z >> 11) * (1.0 / 9007199254740992.0) - 0.5;
    }
    if (c->A_from_mask) {
        // Solid and outlet rows of A are decoupled identity rows (eigenvalue
        // 1, pressure.cpp:634-642); start in the fluid subspace so the Ritz
        // values bound the fluid block the Chebyshev polynomial acts on.
        std::vector<uint8_t> m(N);
        HIPCHK(c, hipMemcpyAsync(m.data(), c->solid, (size_t)N, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (int i = 0; i < N; ++i)
            if (m[i] || i / c->d.sij == c->d.nx - 1) v0[i] = 0.0;
    }
    double *v = c->t1, *w = c->t2, *vp = c->za;
    hipStream_t s = c->stream;
    HIPCHK(c, hipMemcpyAsync(w, v0.data(), (size_t)N * 8, hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemsetAsync(vp, 0, (size_t)N * 8, s));
    HIPCHK(c, launch_dot(s, c->d, w, w, c->red, c->dsc + 1));
    HIPCHK(c, launch_scale_by_invnorm(s, N, w, v, c->dsc + 1));
    const Stencil<double> A = stencil_of(c->A);
    std::vector<double> al, be;
    double beta_prev = 0.0;
    for (int j = 0; j < steps; ++j) {
        HIPCHK(c, launch_matvec_sym(s, c->d, A, v, w));
        HIPCHK(c, launch_dot(s, c->d, w, v, c->red, c->dsc + 0));
        HIPCHK(c, hipMemcpyAsync(c->hsc, c->dsc, 8, hipMemcpyDeviceToHost, s));
        HIPCHK(c, hipStreamSynchronize(s));
        const double alpha = c->hsc[0];
        al.push_back(alpha);
        HIPCHK(c, launch_lanczos_axpy(s, N, w, v, vp, alpha, beta_prev));
        if (j == steps - 1) break;
        HIPCHK(c, launch_dot(s, c->d, w, w, c->red, c->dsc + 1));
        HIPCHK(c, hipMemcpyAsync(c->hsc + 1, c->dsc + 1, 8, hipMemcpyDeviceToHost, s));
        HIPCHK(c, hipStreamSynchronize(s));
        const double beta = std::sqrt(c->hsc[1]);
        if (!(beta > 1e-300)) break;
        be.push_back(beta);
        // vp <- v ; v <- w / beta
        HIPCHK(c, hipMemcpyAsync(vp, v, (size_t)N * 8, hipMemcpyDeviceToDevice, s));
        HIPCHK(c, launch_scale_by_invnorm(s, N, w, v, c->dsc + 1));
        beta_prev = beta;
    }
    be.resize(al.size() > 0 ? al.size() - 1 : 0);
    *ritz_min = tridiag_eig(al, be, 0);
    *ritz_max = tridiag_eig(al, be, (int)al.size() - 1);
    return LES_OK;
}

// ---------------- PCG ----------------
}  // extern "C"

namespace {

// op (may be nullptr: the context's A) is the operator solved with; wgt
// (may be nullptr) weights the r.z and p.q inner products.
int pcg_fp64(les_ctx* c, const les_pcg_params* prm, les_pcg_result* res,
             const Stencil<double>* op = nullptr, const double* wgt = nullptr)
{
    const Dims& d = c->d;
    hipStream_t s = c->stream;
    const int pc = prm->precond;
    int rc;
    if ((rc = ensure_krylov(c)) != LES_OK) return rc;
    if ((rc = ensure_trace(c, prm->max_iter + 1)) != LES_OK) return rc;
    if (pc == LES_PC_IC0 && !c->have_L) return fail(c, LES_ERR_STATE, "IC(0) factor not built");

    SolveScalars init{};
    init.tol = prm->tol_is_relative ? 0.0 : prm->tol;
    init.tol_rel = prm->tol_is_relative ? prm->tol : 0.0;
    init.max_iter = prm->max_iter;
    init.trace = c->trace;
    c->sc_host[0] = init;
    HIPCHK(c, hipMemcpyAsync(c->sc, &c->sc_host[0], sizeof(SolveScalars), hipMemcpyHostToDevice, s));

    bool coded = false;
    if ((rc = pick_coded(c, prm->stencil_format, coded)) != LES_OK) return rc;
    const Stencil<double> A = op ? *op : (coded ? coded_stencil<double>(c) : stencil_of(c->A));
    double* x = c->f[LES_F_P];
    const double* b = c->f[LES_F_B];
    double *r = c->r, *q = c->q, *z = c->z;
    // the search direction ping-pongs: each iteration's matvec forms the new
    // p from the previous one (the direction update fused into it)
    double* pb[2] = {c->p, c->p2};
    int cur = 0;
    double* p = pb[0];
    SolveScalars* sc = c->sc;

    // r = b - A x, ||r||, stop if converged (pressure.cpp:355-359)
    HIPCHK(c, launch_pcg_init<double>(s, d, A, b, x, r, c->red, sc));
    // p = M^-1 r, rho = r.p (pressure.cpp:363-369)
    if (pc == LES_PC_JACOBI)
        HIPCHK(c, launch_pcg_p_jacobi<double>(s, d, r, c->A[SC], p, 1, EPI_RHO, c->red, sc, wgt));
    else if ((rc = precond_fp64(c, pc, r, p, prm->lambda_min, prm->lambda_max, prm->degree, EPI_RHO,
                                sc, A, wgt)) != LES_OK)
        return rc;

    BatchPlanner plan(prm->check_every > 0 ? prm->check_every : (d.N >= (1 << 22) ? 4 : 16));
    int issued = 0, batch = 0;
    bool finished = false;
    while (!finished) {
        const int n = plan.next(issued, prm->max_iter - issued);
        for (int it = 0; it < n; ++it) {
            {
                KtScope kt(c, LES_KT_PCG_MATVEC, 8);
                if (issued + it == 0) {
                    HIPCHK(c, launch_pcg_matvec<double>(s, d, A, p, q, c->red, sc, wgt));
                } else {
                    // p = M^-1 r + beta p_old, fused into q = A p
                    const bool jac = pc == LES_PC_JACOBI;
                    // (with the previous iteration's x += alpha p_old)
                    HIPCHK(c, launch_pcg_matvec_pu<double>(s, d, A, pb[cur], jac ? r : z,
                                                           jac ? c->A[SC] : nullptr, pb[cur ^ 1],
                                                           q, c->red, sc, wgt, x));
                    cur ^= 1;
                    p = pb[cur];
                }
            }
            // r -= alpha q; x's term waits for the next matvec
            if (pc == LES_PC_JACOBI) {
                HIPCHK(c, (launch_pcg_update<double, double>(s, d, nullptr, r, nullptr, q, c->A[SC],
                                                             c->red, sc, wgt)));
            } else {
                HIPCHK(c, (launch_pcg_update<double, double>(s, d, nullptr, r, nullptr, q, nullptr,
                                                             c->red, sc)));
                if ((rc = precond_fp64(c, pc, r, z, prm->lambda_min, prm->lambda_max, prm->degree,
                                       EPI_BETA, sc, A, wgt)) != LES_OK)
                    return rc;
            }
        }
        issued += n;
        const int slot = batch & 1;
        HIPCHK(c, hipMemcpyAsync(&c->sc_host[slot], sc, sizeof(SolveScalars), hipMemcpyDeviceToHost, s));
        HIPCHK(c, hipEventRecord(c->ev[slot], s));
        if (issued >= prm->max_iter) {
            HIPCHK(c, hipEventSynchronize(c->ev[slot]));
            finished = true;
        } else if (plan.tight || batch > 0) {
            // the previous batch while this one runs, or this one when the
            // tolerance is predicted within it
            const int seen = plan.tight ? slot : slot ^ 1;
            HIPCHK(c, hipEventSynchronize(c->ev[seen]));
            const SolveScalars& h = c->sc_host[seen];
            if (h.done) finished = true;
            plan.observe(h.iter, h.err, h.tol, issued);
        }
        if (prm->max_iter <= 0) finished = true;
        ++batch;
    }
    // the last update's x term (its matvec never ran)
    HIPCHK(c, launch_pcg_x_pending<double>(s, d, x, pb[0], pb[1], sc));
    HIPCHK(c, hipMemcpyAsync(&c->sc_host[0], sc, sizeof(SolveScalars), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    const SolveScalars& h = c->sc_host[0];
    res->iters = h.converged ? h.iter : (h.err0 < h.tol ? 0 : prm->max_iter);
    if (h.converged && h.iter == 0) res->iters = 0;
    res->converged = h.converged;
    res->res_norm = h.err;
    res->res0_norm = h.err0;
    res->b_norm = std::sqrt(h.bb);
    res->true_res_norm = h.err;
    res->replacements = 0;
    res->fault = h.fault;
    c->last_trace_n = h.iter + 1;
    return LES_OK;
}


int ensure_mirror_weights(les_ctx* c)
{
    if (c->have_W) return LES_OK;
    if (!c->have_code)
        return fail(c, LES_ERR_STATE, "mirror weights need A built from the mask (les_build_stencil)");
    int rc;
    if ((rc = alloc(c, &c->W, c->d.N, false)) != LES_OK) return rc;
    HIPCHK(c, launch_mirror_weights(c->stream, c->d, c->code, c->W));
    c->have_W = true;
    return LES_OK;
}

int ensure_fp32_ops(les_ctx* c, int pc, bool coded)
{
    int rc;
    const bool need_arrays = !coded || pc == LES_PC_LEGACY_IC_A || pc == LES_PC_JACOBI;
    if (need_arrays && !c->have_A32) {
        for (int q = 0; q < 7; ++q) {
            if ((rc = alloc(c, &c->A32[q], c->d.N, false)) != LES_OK) return rc;
            HIPCHK(c, (launch_convert<float, double>(c->stream, c->d.N, c->A[q], c->A32[q])));
        }
        c->have_A32 = true;
    }
    if (pc == LES_PC_IC0 && !c->have_L32) {
        if (!c->have_L) return fail(c, LES_ERR_STATE, "IC(0) factor not built");
        for (int q = 0; q < 7; ++q) {
            if ((rc = alloc(c, &c->L32[q], c->d.N, false)) != LES_OK) return rc;
            HIPCHK(c, (launch_convert<float, double>(c->stream, c->d.N, c->L[q], c->L32[q])));
        }
        c->have_L32 = true;
    }
    MixedWork& w = c->mw;
    for (float** b : {&w.r, &w.p, &w.p2, &w.q, &w.z, &w.za, &w.zb, &w.pc, &w.pc2, &w.x})
        if ((rc = alloc(c, b, c->d.N)) != LES_OK) return rc;
    return LES_OK;
}

// z = M^-1 r in FP32 (operator and vectors FP32, r.z accumulated in FP64)
int precond_fp32(les_ctx* c, int pc, const float* r, float* zout, double lmin, double lmax,
                 int degree, int epi, SolveScalars* sc, const Stencil<float>& Aop,
                 const double* wgt = nullptr)
{
    if (wgt && pc != LES_PC_MG)
        return fail(c, LES_ERR_ARG, "mixed precision in the mirror-weighted inner product needs "
                                    "the multigrid preconditioner");
    const Dims& d = c->d;
    hipStream_t s = c->stream;
    switch (pc) {
    case LES_PC_NONE:
        HIPCHK(c, hipMemcpyAsync(zout, r, nbytes(c, 4), hipMemcpyDeviceToDevice, s));
        break;
    case LES_PC_JACOBI:
        HIPCHK(c, launch_pcg_p_jacobi<float>(s, d, r, c->A32[SC], zout, 1, epi, c->red, sc));
        return LES_OK;
    case LES_PC_IC0:
    case LES_PC_LEGACY_IC_A: {
        float* const* L = (pc == LES_PC_IC0) ? c->L32 : c->A32;
        int rc;
        if ((rc = ic_solve<float>(c, r, zout, L, sc)) != LES_OK) return rc;
        break;
    }
    case LES_PC_CHEBYSHEV:
        return cheb_chain<float>(c, Aop, r, zout, c->mw.za, c->mw.zb, c->mw.pc, c->mw.pc2, lmin,
                                 lmax, degree, epi, sc, nullptr);
    case LES_PC_MG: {
        int rc;
        if ((rc = mg_setup(c, true, wgt)) != LES_OK) return rc;
        return mg_vcycle<float>(c, Aop, r, zout, c->mw.za, epi, sc, wgt);
    }
    default:
        return fail(c, LES_ERR_ARG, "unknown preconditioner");
    }
    if (epi >= 0) HIPCHK(c, launch_pcg_rz<float>(s, d, r, zout, epi, c->red, sc));
    return LES_OK;
}

// Mixed-precision PCG: the Krylov iteration runs on the FP32 copy of A with
// FP32 vectors and FP64 dot accumulation; the iterate is held as
// x64 + x32.  Every `reliable_every` iterations, once the recursive
// residual has fallen `replace_decades` decades below the last true one,
// whenever it meets the tolerance, and at max_iter, x32 is folded into x64
// and r is replaced by the FP64 residual b - A x64 (FP64 stencil).
// Convergence is declared only on that FP64 residual, so the result meets
// the tolerance in FP64 (outer iterative refinement to the requested tol).
int pcg_mixed(les_ctx* c, const les_pcg_params* prm, les_pcg_result* res)
{
    const Dims& d = c->d;
    hipStream_t s = c->stream;
    const int pc = prm->precond;
    int rc;
    if ((rc = ensure_krylov(c)) != LES_OK) return rc;
    bool coded = false;
    if ((rc = pick_coded(c, prm->stencil_format, coded)) != LES_OK) return rc;
    if ((rc = ensure_fp32_ops(c, pc, coded)) != LES_OK) return rc;
    if ((rc = ensure_trace(c, prm->max_iter + 1)) != LES_OK) return rc;
    MixedWork& w = c->mw;

    SolveScalars init{};
    init.tol = prm->tol_is_relative ? 0.0 : prm->tol;
    init.tol_rel = prm->tol_is_relative ? prm->tol : 0.0;
    init.max_iter = prm->max_iter;
    init.reliable_every = prm->reliable_every > 0 ? prm->reliable_every : 50;
    init.reliable_drop = c->replace_decades > 0 ? std::pow(10.0, -c->replace_decades) : 0.0;
    init.trace = c->trace;
    c->sc_host[0] = init;
    // A multigrid iteration takes the residual down about 2.5x, so a drop
    // request can stand on any iteration: its gated replacement kernels are
    // enqueued on every one (a few microseconds against the V-cycle).  The
    // slower preconditioners serve it at the next periodic replacement.
    const bool gate_every = pc == LES_PC_MG && init.reliable_drop > 0.0;
    HIPCHK(c, hipMemcpyAsync(c->sc, &c->sc_host[0], sizeof(SolveScalars), hipMemcpyHostToDevice, s));
    HIPCHK(c, hipMemsetAsync(w.x, 0, nbytes(c, 4), s));

    const Stencil<double> A64 = coded ? coded_stencil<double>(c) : stencil_of(c->A);
    const Stencil<float> A32 = coded ? coded_stencil<float>(c) : stencil_of(c->A32);
    double* x64 = c->f[LES_F_P];
    const double* b = c->f[LES_F_B];
    SolveScalars* sc = c->sc;
    float* pb[2] = {w.p, w.p2};
    int cur = 0;
    // r.z and p.q in the mirror-weighted inner product (multigrid only)
    const double* wgt = prm->inner_product == LES_IP_MIRROR ? c->W : nullptr;

    // r = b - A x0 in FP64, rounded to FP32; ||r||, ||b||
    HIPCHK(c, launch_mixed_residual(s, d, A64, b, x64, w.r, 1, 0, c->red, sc, nullptr));
    if ((rc = precond_fp32(c, pc, w.r, w.p, prm->lambda_min, prm->lambda_max, prm->degree, EPI_RHO,
                           sc, A32, wgt)) != LES_OK)
        return rc;

    BatchPlanner plan(prm->check_every > 0 ? prm->check_every : (d.N >= (1 << 22) ? 4 : 16));
    int issued = 0, batch = 0;
    bool finished = prm->max_iter <= 0;
    bool near = false;  // replacement kernels on every iteration from then on
    int pq_grid = 0;    // > 0: the matvec left its p.q partials in red2 for the update
    int rz_grid = 0;    // > 0: the preconditioner left its r.z partials in red for the
                        // next matvec (beta) and r update (rho, EPI_BETA's bookkeeping)
    double last_err = 0.0;
    while (!finished) {
        const int n = plan.next(issued, prm->max_iter - issued);
        // a batch predicted to meet the tolerance carries the replacement
        // that confirms it
        if (plan.tight) near = true;
        for (int it = 0; it < n; ++it) {
            {
                KtScope kt(c, LES_KT_PCG_MATVEC, 4);
                if (issued + it == 0) {
                    HIPCHK(c, launch_pcg_matvec<float>(s, d, A32, pb[cur], w.q, c->red, sc, wgt));
                } else {
                    // p = z + beta p_old fused into q = A p, with the
                    // previous iteration's x32 += alpha p_old; its p.q
                    // partials go to red2, and the r update forms alpha
                    HIPCHK(c, launch_pcg_matvec_pu<float>(s, d, A32, pb[cur], w.z, nullptr,
                                                          pb[cur ^ 1], w.q, c->red2, sc, wgt,
                                                          w.x, &pq_grid, c->red, rz_grid));
                    cur ^= 1;
                }
            }
            {
                KtScope kt(c, LES_KT_PCG_UPDATE, 4);
                HIPCHK(c, launch_pcg_update_mixed(s, d, nullptr, w.r, nullptr, w.q, c->red3, sc,
                                                  c->red2, pq_grid, c->red, rz_grid));
                pq_grid = 0;
                rz_grid = 0;
            }
            // The replacement kernels run gated on the device's request; they
            // are enqueued on the iterations where a request can stand: the
            // periodic ones, max_iter, every iteration once the recursive
            // residual is near the tolerance or a request was seen pending,
            // and every iteration of a multigrid solve.
            const int itn = issued + it + 1;
            if (near || gate_every || itn % init.reliable_every == 0 || itn >= prm->max_iter) {
                HIPCHK(c, launch_mixed_fold(s, d, x64, w.x, pb[0], pb[1], 0, sc));
                HIPCHK(c, launch_mixed_residual(s, d, A64, b, x64, w.r, 0, 0, c->red, sc,
                                                nullptr));
            }
            // r.z stays as partials: the next matvec forms beta and the
            // next r update rho from them (one finalize launch fewer)
            take_partials_grid();
            if ((rc = precond_fp32(c, pc, w.r, w.z, prm->lambda_min, prm->lambda_max, prm->degree,
                                   EPI_PARTIALS, sc, A32, wgt)) != LES_OK)
                return rc;
            rz_grid = take_partials_grid();
            if (rz_grid <= 0) return fail(c, LES_ERR_STATE, "preconditioner left no r.z partials");
        }
        issued += n;
        const int slot = batch & 1;
        HIPCHK(c, hipMemcpyAsync(&c->sc_host[slot], sc, sizeof(SolveScalars), hipMemcpyDeviceToHost, s));
        HIPCHK(c, hipEventRecord(c->ev[slot], s));
        if (issued >= prm->max_iter) {
            HIPCHK(c, hipEventSynchronize(c->ev[slot]));
            finished = true;
        } else if (plan.tight || batch > 0) {
            const int seen = plan.tight ? slot : slot ^ 1;
            HIPCHK(c, hipEventSynchronize(c->ev[seen]));
            const SolveScalars& h = c->sc_host[seen];
            if (h.done) finished = true;
            plan.observe(h.iter, h.err, h.tol, issued);
            // near: the residual would reach the tolerance within about two
            // more polls at the rate of the last one (the host sees a request
            // up to two batches late)
            const double drop = last_err > 0.0 && h.err > 0.0 ? last_err / h.err : 1.0;
            const double reach = std::max(4.0, drop * drop * drop);
            if (h.reliable_now || h.err < reach * h.tol) near = true;
            last_err = h.err;
        }
        ++batch;
    }
    // fold whatever correction is left and measure the FP64 residual
    HIPCHK(c, launch_mixed_fold(s, d, x64, w.x, pb[0], pb[1], 1, sc));
    HIPCHK(c, launch_mixed_residual(s, d, A64, b, x64, nullptr, 0, 1, c->red, sc, c->dsc));
    HIPCHK(c, hipMemcpyAsync(&c->sc_host[0], sc, sizeof(SolveScalars), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipMemcpyAsync(c->hsc, c->dsc, 8, hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    const SolveScalars& h = c->sc_host[0];
    // reference convention (pressure.cpp:411-413) except on a fault, where
    // the iteration that met the non-finite scalar is reported
    res->iters = (h.converged || h.fault) ? h.iter : prm->max_iter;
    res->converged = h.converged;
    res->res_norm = h.err;
    res->res0_norm = h.err0;
    res->b_norm = std::sqrt(h.bb);
    res->true_res_norm = c->hsc[0];
    res->replacements = h.replacements;
    res->fault = h.fault;
    res->restarts = h.restarts;
    c->last_trace_n = h.iter + 1;
    return LES_OK;
}

// eigenvalue of a symmetric tridiagonal matrix (diag a, off-diag b) with
// index `which` (0 = smallest) by Sturm-sequence bisection
double tridiag_eig(const std::vector<double>& a, const std::vector<double>& b, int which)
{
    const int m = (int)a.size();
    double lo = a[0], hi = a[0];
    for (int i = 0; i < m; ++i) {
        const double r = (i > 0 ? std::fabs(b[i - 1]) : 0.0) + (i < m - 1 ? std::fabs(b[i]) : 0.0);
        lo = std::min(lo, a[i] - r);
        hi = std::max(hi, a[i] + r);
    }
    auto count_below = [&](double x) {
        int cnt = 0;
        double q = 1.0;
        for (int i = 0; i < m; ++i) {
            const double bb = i > 0 ? b[i - 1] * b[i - 1] : 0.0;
            q = a[i] - x - (i > 0 ? bb / q : 0.0);
            if (q == 0.0) q = -1e-300;
            if (q < 0) ++cnt;
        }
        return cnt;
    };
    for (int it = 0; it < 200; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (mid == lo || mid == hi) break;
        if (count_below(mid) > which) hi = mid;
        else lo = mid;
    }
    return 0.5 * (lo + hi);
}

}  // namespace

extern "C" {

int les_precond_apply_fp32(les_ctx* c, int pc, const float* r, float* z, int ptr_kind, double lmin,
                           double lmax, int degree, int stencil_format, int inner_product)
{
    if (!c || !r || !z) return fail(c, LES_ERR_ARG, "NULL vector");
    if (!c->have_A) return fail(c, LES_ERR_STATE, "stencil A not built");
    if (pc < LES_PC_NONE || pc > LES_PC_MG) return fail(c, LES_ERR_ARG, "unknown preconditioner");
    if (inner_product != LES_IP_EUCLID && inner_product != LES_IP_MIRROR)
        return fail(c, LES_ERR_ARG, "unknown inner product");
    int rc;
    if ((rc = ensure_krylov(c)) != LES_OK) return rc;
    bool coded = false;
    if ((rc = pick_coded(c, stencil_format, coded)) != LES_OK) return rc;
    if ((rc = ensure_fp32_ops(c, pc, coded)) != LES_OK) return rc;
    const double* wgt = nullptr;
    if (inner_product == LES_IP_MIRROR) {
        if ((rc = ensure_mirror_weights(c)) != LES_OK) return rc;
        wgt = c->W;
    }
    const Stencil<float> A32 = coded ? coded_stencil<float>(c) : stencil_of(c->A32);
    MixedWork& w = c->mw;
    HIPCHK(c, hipMemcpyAsync(w.r, r, nbytes(c, 4), kind_in(ptr_kind), c->stream));
    if ((rc = precond_fp32(c, pc, w.r, w.z, lmin, lmax, degree, -1, nullptr, A32, wgt)) != LES_OK)
        return rc;
    HIPCHK(c, hipMemcpyAsync(z, w.z, nbytes(c, 4), kind_out(ptr_kind), c->stream));
    if (!ptr_kind) HIPCHK(c, hipStreamSynchronize(c->stream));
    return wf_collect(c);
}

int les_pcg(les_ctx* c, const les_pcg_params* prm, les_pcg_result* res)
{
    if (!c || !prm || !res) return fail(c, LES_ERR_ARG, "NULL argument");
    if (!c->have_A) return fail(c, LES_ERR_STATE, "stencil A not built");
    if (prm->precond < LES_PC_NONE || prm->precond > LES_PC_MG)
        return fail(c, LES_ERR_ARG, "unknown preconditioner");
    if (prm->max_iter < 0) return fail(c, LES_ERR_ARG, "max_iter < 0");
    std::memset(res, 0, sizeof(*res));
    if (prm->inner_product == LES_IP_MIRROR) {
        int rc;
        if (prm->precision != LES_PREC_FP64 &&
            !(prm->precision == LES_PREC_MIXED && prm->precond == LES_PC_MG))
            return fail(c, LES_ERR_ARG, "the mirror-weighted inner product is FP64 only (mixed "
                                        "precision: with the multigrid preconditioner)");
        if (prm->precond == LES_PC_IC0 || prm->precond == LES_PC_LEGACY_IC_A)
            return fail(c, LES_ERR_ARG, "the mirror-weighted inner product needs a preconditioner "
                                        "self-adjoint in it (NONE, JACOBI, CHEBYSHEV or MG)");
        if ((rc = ensure_mirror_weights(c)) != LES_OK) return rc;
        if (prm->precision == LES_PREC_MIXED) {
            if ((rc = pcg_mixed(c, prm, res)) != LES_OK) return rc;
            return wf_collect(c);
        }
        return pcg_fp64(c, prm, res, nullptr, c->W);
    }
    if (prm->inner_product != LES_IP_EUCLID) return fail(c, LES_ERR_ARG, "unknown inner product");
    if (prm->precision != LES_PREC_MIXED && prm->precision != LES_PREC_FP64)
        return fail(c, LES_ERR_ARG, "unknown precision");
    const int rc = prm->precision == LES_PREC_MIXED ? pcg_mixed(c, prm, res) : pcg_fp64(c, prm, res);
    if (rc != LES_OK) return rc;
    return wf_collect(c);
}

int les_pcg_trace(les_ctx* c, double* trace, int n)
{
    if (!c || !trace) return fail(c, LES_ERR_ARG, "NULL argument");
    const int m = std::min(n, c->last_trace_n);
    if (m <= 0) return LES_OK;
    HIPCHK(c, hipMemcpyAsync(trace, c->trace, (size_t)m * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LES_OK;
}

// ---------------- stationary solvers ----------------
int les_pressure_bcs(les_ctx* c)
{
    if (!c) return LES_ERR_ARG;
    HIPCHK(c, launch_pressure_bcs(c->stream, c->d, c->f[LES_F_P], c->solid, nullptr, 0));
    return LES_OK;
}

}  // extern "C"

namespace {

int stat_begin(les_ctx* c, int max_sweeps)
{
    StatScalars init{};
    init.thresh = 1e-6;
    init.max_sweeps = max_sweeps;
    c->st_host[0] = init;
    HIPCHK(c, hipMemcpyAsync(c->st, &c->st_host[0], sizeof(StatScalars), hipMemcpyHostToDevice,
                             c->stream));
    return LES_OK;
}

// Runs `sweep(k)` for k = 0..max_sweeps-1 with device-side early exit,
// polling the host every few sweeps.
template <typename F>
int stat_loop(les_ctx* c, int max_sweeps, F sweep)
{
    const int check = 8;
    int issued = 0, batch = 0;
    bool finished = max_sweeps <= 0;
    while (!finished) {
        const int n = std::min(check, max_sweeps - issued);
        for (int k = 0; k < n; ++k) {
            int rc = sweep(issued + k);
            if (rc != LES_OK) return rc;
        }
        issued += n;
        const int slot = batch & 1;
        HIPCHK(c, hipMemcpyAsync(&c->st_host[slot], c->st, sizeof(StatScalars),
                                 hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipEventRecord(c->ev[slot], c->stream));
        if (issued >= max_sweeps) {
            HIPCHK(c, hipEventSynchronize(c->ev[slot]));
            finished = true;
        } else if (batch > 0) {
            HIPCHK(c, hipEventSynchronize(c->ev[slot ^ 1]));
            if (c->st_host[slot ^ 1].done) finished = true;
        }
        ++batch;
    }
    HIPCHK(c, hipMemcpyAsync(&c->st_host[0], c->st, sizeof(StatScalars), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LES_OK;
}

}  // namespace

extern "C" {

int les_jacobi_solve(les_ctx* c, double dt, double rho, int* sweeps_out)
{
    if (!c) return LES_ERR_ARG;
    int rc;
    if ((rc = ensure_krylov(c)) != LES_OK) return rc;
    const Dims& d = c->d;
    // pressure.cpp:111-132
    const double coef = rho / dt;
    const double dx2 = c->dx * c->dx, dy2 = c->dy * c->dy, dz2 = c->dz * c->dz;
    const double denom = 2 * (1.0 / dx2 + 1.0 / dy2 + 1.0 / dz2);
    double* rhs = c->t1;
    HIPCHK(c, launch_rhs_central(c->stream, d, c->f[LES_F_U], c->f[LES_F_V], c->f[LES_F_W], rhs,
                                 coef, 2 * c->dx, 2 * c->dy, 2 * c->dz));
    double* buf[2] = {c->f[LES_F_P], c->t2};
    HIPCHK(c, hipMemsetAsync(c->t2, 0, nbytes(c, 8), c->stream));
    if ((rc = stat_begin(c, 100)) != LES_OK) return rc;
    rc = stat_loop(c, 100, [&](int k) -> int {
        HIPCHK(c, launch_jacobi_sweep(c->stream, d, buf[k & 1], buf[(k + 1) & 1], rhs, dx2, dy2,
                                      dz2, denom, c->st));
        HIPCHK(c, launch_pressure_bcs(c->stream, d, buf[(k + 1) & 1], c->solid, c->st, k));
        return LES_OK;
    });
    if (rc != LES_OK) return rc;
    const int sw = c->st_host[0].sweeps;
    if (sw & 1)
        HIPCHK(c, hipMemcpyAsync(c->f[LES_F_P], c->t2, nbytes(c, 8), hipMemcpyDeviceToDevice,
                                 c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (sweeps_out) *sweeps_out = sw;
    return LES_OK;
}

int les_sor_solve(les_ctx* c, int ordering, double dt, double rho, double omega, int max_sweeps,
                  int* sweeps_out)
{
    if (!c) return LES_ERR_ARG;
    if (ordering != LES_SOLVER_SOR_LEX && ordering != LES_SOLVER_SOR_RB)
        return fail(c, LES_ERR_ARG, "ordering must be LES_SOLVER_SOR_LEX or LES_SOLVER_SOR_RB");
    int rc;
    if ((rc = ensure_krylov(c)) != LES_OK) return rc;
    const Dims& d = c->d;
    // pressure.cpp:193-210
    const double dx2 = c->dx * c->dx, dy2 = c->dy * c->dy, dz2 = c->dz * c->dz;
    const double coef = 1.0 / (2.0 * (1.0 / dx2 + 1.0 / dy2 + 1.0 / dz2));
    double* rhs = c->t1;
    HIPCHK(c, launch_rhs_central(c->stream, d, c->f[LES_F_U], c->f[LES_F_V], c->f[LES_F_W], rhs,
                                 rho / dt, 2 * c->dx, 2 * c->dy, 2 * c->dz));
    const double omc = 1 - omega, oc = omega * coef;
    double* p = c->f[LES_F_P];
    if ((rc = stat_begin(c, max_sweeps)) != LES_OK) return rc;
    const int lmax = (d.nx - 2) + (d.ny - 2) + (d.nz - 2);
    rc = stat_loop(c, max_sweeps, [&](int k) -> int {
        if (ordering == LES_SOLVER_SOR_LEX) {
            KtScope kt(c, LES_KT_SOR_SWEEP, 8);
            if (c->wavefront) {
                WfBufs wb;
                int rc2;
                if ((rc2 = ensure_wf(c, &wb)) != LES_OK) return rc2;
                HIPCHK(c, launch_wf_sor(c->stream, d, p, rhs, omc, oc, dx2, dy2, dz2, c->st, wb));
            } else {
                for (int l = 3; l <= lmax; ++l)
                    HIPCHK(c, launch_sor_level(c->stream, d, l, p, rhs, omc, oc, dx2, dy2, dz2,
                                               c->st));
            }
        } else {
            for (int colour = 0; colour < 2; ++colour)
                HIPCHK(c, launch_sor_colour(c->stream, d, p, rhs, colour, omc, oc, dx2, dy2, dz2,
                                            c->st));
        }
        HIPCHK(c, launch_pressure_bcs(c->stream, d, p, c->solid, c->st, k));
        return LES_OK;
    });
    if (rc != LES_OK) return rc;
    if ((rc = wf_collect(c)) != LES_OK) return rc;
    if (sweeps_out) *sweeps_out = c->st_host[0].sweeps;
    return LES_OK;
}

// ---------------- physics ----------------
int les_sgs(les_ctx* c, double Cs, int apply_bcs, double* const tau[6], int ptr_kind)
{
    if (!c) return LES_ERR_ARG;
    Tau t{};
    std::vector<double*> tmp;
    if (tau) {
        for (int q = 0; q < 6; ++q) {
            if (!tau[q]) return fail(c, LES_ERR_ARG, "tau array is NULL");
            if (ptr_kind) {
                t.t[q] = tau[q];
            } else {
                double* dptr = nullptr;
                int rc = alloc(c, &dptr, c->d.N);
                if (rc != LES_OK) {
                    for (double* p : tmp) hipFree(p);
                    return rc;
                }
                HIPCHK(c, hipMemcpyAsync(dptr, tau[q], nbytes(c, 8), hipMemcpyHostToDevice, c->stream));
                tmp.push_back(dptr);
                t.t[q] = dptr;
            }
        }
    }
    const PhysConsts k = make_consts(c, 1.0, 0.0, 1.0, Cs);
    HIPCHK(c, launch_sgs(c->stream, c->d, k, c->f[LES_F_U], c->f[LES_F_V], c->f[LES_F_W], c->solid,
                         c->f[LES_F_NU_T], t, apply_bcs, c->red, nullptr));
    if (tau && !ptr_kind) {
        for (int q = 0; q < 6; ++q)
            HIPCHK(c, hipMemcpyAsync(tau[q], tmp[q], nbytes(c, 8), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (double* p : tmp) hipFree(p);
    }
    return LES_OK;
}

int les_nut_bcs(les_ctx* c)
{
    if (!c) return LES_ERR_ARG;
    HIPCHK(c, launch_nut_bcs(c->stream, c->d, c->solid, c->f[LES_F_NU_T]));
    return LES_OK;
}

static void swap_velocity(les_ctx* c)
{
    std::swap(c->f[LES_F_U], c->f[LES_F_U_NEW]);
    std::swap(c->f[LES_F_V], c->f[LES_F_V_NEW]);
    std::swap(c->f[LES_F_W], c->f[LES_F_W_NEW]);
}

int les_momentum(les_ctx* c, double dt, double nu, int apply_bcs)
{
    if (!c) return LES_ERR_ARG;
    const PhysConsts k = make_consts(c, dt, nu, 1.0, 0.18);
    if (apply_bcs) {
        // provisional velocity with BCs; the lagged divergence is not wanted here
        HIPCHK(c, launch_momentum(c->stream, c->d, k, c->f[LES_F_U], c->f[LES_F_V], c->f[LES_F_W],
                                  c->f[LES_F_NU_T], c->solid, c->f[LES_F_U_NEW], c->f[LES_F_V_NEW],
                                  c->f[LES_F_W_NEW], nullptr, nullptr, MOM_STEP));
    } else {
        HIPCHK(c, launch_momentum(c->stream, c->d, k, c->f[LES_F_U], c->f[LES_F_V], c->f[LES_F_W],
                                  c->f[LES_F_NU_T], c->solid, c->f[LES_F_U_NEW], c->f[LES_F_V_NEW],
                                  c->f[LES_F_W_NEW], nullptr, nullptr, MOM_RAW));
    }
    swap_velocity(c);
    return LES_OK;
}

int les_velocity_bcs(les_ctx* c)
{
    if (!c) return LES_ERR_ARG;
    HIPCHK(c, launch_velocity_bcs(c->stream, c->d, c->solid, c->f[LES_F_U], c->f[LES_F_V],
                                  c->f[LES_F_W]));
    return LES_OK;
}

int les_divergence_b(les_ctx* c, double dt, double rho)
{
    if (!c) return LES_ERR_ARG;
    const PhysConsts k = make_consts(c, dt, 0.0, rho, 0.18);
    HIPCHK(c, launch_momentum(c->stream, c->d, k, c->f[LES_F_U_NEW], c->f[LES_F_V_NEW],
                              c->f[LES_F_W_NEW], c->f[LES_F_NU_T], c->solid, nullptr, nullptr,
                              nullptr, c->f[LES_F_P_RHS], c->f[LES_F_B], MOM_DIV_ONLY));
    return LES_OK;
}

int les_build_b(les_ctx* c)
{
    if (!c) return LES_ERR_ARG;
    HIPCHK(c, launch_build_b(c->stream, c->d, c->f[LES_F_P_RHS], c->solid, c->f[LES_F_B]));
    return LES_OK;
}

int les_correct(les_ctx* c, double dt, double rho, int apply_bcs)
{
    if (!c) return LES_ERR_ARG;
    const PhysConsts k = make_consts(c, dt, 0.0, rho, 0.18);
    HIPCHK(c, launch_correct(c->stream, c->d, k, c->f[LES_F_U], c->f[LES_F_V], c->f[LES_F_W],
                             c->f[LES_F_P], c->solid, apply_bcs));
    return LES_OK;
}

static double dt_from_max(const les_ctx* c, double nu, double umax2, double ntmax)
{
    // compute_adaptive_dt (main.cpp:227-249)
    const double CFL_CONV = 0.4, CFL_VISC = 0.2, DT_MAX = 1e-4;
    const double u_max = std::sqrt(umax2);
    const double min_dx = std::min({c->dx, c->dy, c->dz});
    double dt_conv = std::numeric_limits<double>::max();
    if (u_max > 1e-12) dt_conv = CFL_CONV * min_dx / u_max;
    const double min_dx_sq = min_dx * min_dx;
    const double dt_visc = CFL_VISC * min_dx_sq / (nu + ntmax);
    return std::min({dt_conv, dt_visc, DT_MAX});
}

int les_adaptive_dt(les_ctx* c, double nu, double* dt_out)
{
    if (!c || !dt_out) return fail(c, LES_ERR_ARG, "NULL argument");
    HIPCHK(c, launch_dt_max(c->stream, c->d, c->f[LES_F_U], c->f[LES_F_V], c->f[LES_F_W],
                            c->f[LES_F_NU_T], c->red, c->dsc));
    HIPCHK(c, hipMemcpyAsync(c->hsc, c->dsc, 16, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *dt_out = dt_from_max(c, nu, c->hsc[0], c->hsc[1]);
    return LES_OK;
}

// ---------------- one step (main.cpp:115-144) ----------------
int les_step(les_ctx* c, const les_step_params* prm, les_step_stats* stats)
{
    if (!c || !prm) return fail(c, LES_ERR_ARG, "NULL argument");
    les_step_stats st{};
    st.speed_max = -1.0;
    st.nu_t_max = -1.0;
    int rc;
    double dt;
    if (prm->dt_fixed > 0) dt = prm->dt_fixed;
    else if (prm->step <= 20) dt = 1e-3;
    else {
        if ((rc = les_adaptive_dt(c, prm->nu, &dt)) != LES_OK) return rc;
        st.speed_max = std::sqrt(c->hsc[0]);  // the reductions les_adaptive_dt read back
        st.nu_t_max = c->hsc[1];
    }
    st.dt = dt;
    const PhysConsts k = make_consts(c, dt, prm->nu, prm->rho, prm->Cs);
    hipStream_t s = c->stream;
    // compute_sgs_stress + apply_nut_BCs
    HIPCHK(c, launch_sgs(s, c->d, k, c->f[LES_F_U], c->f[LES_F_V], c->f[LES_F_W], c->solid,
                         c->f[LES_F_NU_T], Tau{}, 1, c->red, nullptr));
    const bool pcg = prm->solver == LES_SOLVER_PCG;
    // solve_momentum + apply_velocity_BCs, and (PCG path) calculate_divergence
    // + build_b_vector of the step's starting velocity
    {
        KtScope ktm(c, LES_KT_MOMENTUM, 8);
        HIPCHK(c, launch_momentum(s, c->d, k, c->f[LES_F_U], c->f[LES_F_V], c->f[LES_F_W],
                                  c->f[LES_F_NU_T], c->solid, c->f[LES_F_U_NEW],
                                  c->f[LES_F_V_NEW], c->f[LES_F_W_NEW],
                                  pcg ? c->f[LES_F_P_RHS] : nullptr,
                                  pcg ? c->f[LES_F_B] : nullptr, MOM_STEP));
    }
    swap_velocity(c);
    switch (prm->solver) {
    case LES_SOLVER_PCG:
        if ((rc = les_pcg(c, &prm->pcg, &st.pcg)) != LES_OK) return rc;
        break;
    case LES_SOLVER_SOR_LEX:
    case LES_SOLVER_SOR_RB:
        if ((rc = les_sor_solve(c, prm->solver, dt, prm->rho, prm->omega, prm->sor_sweeps,
                                &st.sweeps)) != LES_OK)
            return rc;
        break;
    case LES_SOLVER_JACOBI:
        if ((rc = les_jacobi_solve(c, dt, prm->rho, &st.sweeps)) != LES_OK) return rc;
        break;
    default:
        return fail(c, LES_ERR_ARG, "unknown solver");
    }
    // correct_velocity + apply_velocity_BCs
    HIPCHK(c, launch_correct(s, c->d, k, c->f[LES_F_U], c->f[LES_F_V], c->f[LES_F_W],
                             c->f[LES_F_P], c->solid, 1));
    if (stats) *stats = st;
    return LES_OK;
}

}  // extern "C"

// ---------------- differentiation ----------------
namespace {

constexpr int kAdjBufs = 16;  // g(3) y(3) m(3) G(6) nu_t(1)

int adj_check(les_ctx* c, const les_adj_params* prm)
{
    if (!prm) return fail(c, LES_ERR_ARG, "NULL params");
    if (!c->have_A) return fail(c, LES_ERR_STATE, "stencil A not built");
    if (!(prm->dt > 0)) return fail(c, LES_ERR_ARG, "dt must be > 0");
    if (prm->nsteps < 1 || prm->checkpoint_every < 1)
        return fail(c, LES_ERR_ARG, "nsteps and checkpoint_every must be >= 1");
    const les_pcg_params& q = prm->pcg;
    if (q.precision != LES_PREC_FP64) return fail(c, LES_ERR_ARG, "the adjoint path is FP64");
    if (!q.tol_is_relative || !(q.tol > 0))
        return fail(c, LES_ERR_ARG, "the adjoint path needs a relative pressure tolerance > 0");
    if (q.precond != LES_PC_NONE && q.precond != LES_PC_JACOBI && q.precond != LES_PC_CHEBYSHEV)
        return fail(c, LES_ERR_ARG, "the adjoint path takes preconditioner NONE, JACOBI or CHEBYSHEV");
    if (q.inner_product != LES_IP_MIRROR)
        return fail(c, LES_ERR_ARG, "the adjoint path solves in the mirror-weighted inner product "
                                    "(pcg.inner_product = LES_IP_MIRROR)");
    if (!c->have_code)
        return fail(c, LES_ERR_STATE, "the adjoint path needs A built by les_build_stencil");
    if (c->A_has_obstacles)
        return fail(c, LES_ERR_STATE, "A has mirror rows around interior obstacles, which no "
                                      "diagonal weight symmetrises, so CG would not give the exact "
                                      "map; build A with ignore_mask = 1 (as the shipped driver "
                                      "does) or from an obstacle-free mask");
    return LES_OK;
}

int ensure_states(les_ctx* c, double** buf, int* cap, int count)
{
    if (*cap >= count) return LES_OK;
    if (*buf) hipFree(*buf);
    *buf = nullptr;
    *cap = 0;
    int rc = alloc(c, buf, (size_t)count * 3 * c->d.N, false);
    if (rc == LES_OK) *cap = count;
    return rc;
}

void copy_state(les_ctx* c, double* dst, const double* u, const double* v, const double* w,
                hipError_t& e)
{
    const size_t b = nbytes(c, 8), n = (size_t)c->d.N;
    if (e == hipSuccess) e = hipMemcpyAsync(dst, u, b, hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dst + n, v, b, hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dst + 2 * n, w, b, hipMemcpyDeviceToDevice, c->stream);
}

void adj_account(les_adj_stats* st, const les_pcg_result& r)
{
    if (!st) return;
    st->solves += 1;
    st->iters += r.iters;
    if (!r.converged) st->unconverged += 1;
    const double rel = r.b_norm > 0 ? r.res_norm / r.b_norm : 0.0;
    st->max_rel_res = std::max(st->max_rel_res, rel);
}

// One step of the differentiated map: p from zero, fixed dt.
int roll_step(les_ctx* c, const les_adj_params* prm, les_adj_stats* st)
{
    HIPCHK(c, hipMemsetAsync(c->f[LES_F_P], 0, nbytes(c, 8), c->stream));
    les_step_params sp{};
    sp.dt_fixed = prm->dt;
    sp.nu = prm->nu;
    sp.rho = prm->rho;
    sp.Cs = prm->Cs;
    sp.solver = LES_SOLVER_PCG;
    sp.pcg = prm->pcg;
    les_step_stats ss{};
    const int rc = les_step(c, &sp, &ss);
    if (rc != LES_OK) return rc;
    adj_account(st, ss.pcg);
    return LES_OK;
}

// The transposed step at state (u, v, w): g <- (dF/du)^T g.
int adj_step(les_ctx* c, const les_adj_params* prm, const double* u, const double* v,
             const double* w, les_adj_stats* st)
{
    const Dims& d = c->d;
    hipStream_t s = c->stream;
    const PhysConsts k = make_consts(c, prm->dt, prm->nu, prm->rho, prm->Cs);
    double** b = c->adjb;
    double *gu = b[0], *gv = b[1], *gw = b[2], *yu = b[3], *yv = b[4], *yw = b[5];
    double *mu = b[6], *mv = b[7], *mw = b[8], *nut = b[15];
    const StrainSens G{{b[9], b[10], b[11], b[12], b[13], b[14]}};
    const double* W = c->W;
    // nu_t of the step (compute_sgs_stress + apply_nut_BCs)
    HIPCHK(c, launch_sgs(s, d, k, u, v, w, c->solid, nut, Tau{}, 1, c->red, nullptr));
    // u^{n+1} = BC(u* - (dt/rho) grad p): y = BC^T g, rhs = W^-1 grad^T y
    HIPCHK(c, launch_adj_velocity_bcs(s, d, c->solid, gu, gv, gw, yu, yv, yw));
    HIPCHK(c, launch_adj_correct(s, d, k, yu, yv, yw, W, c->f[LES_F_B]));
    // x = A^-1 rhs by CG in the W inner product (A^-T = W A^-1 W^-1)
    HIPCHK(c, hipMemsetAsync(c->f[LES_F_P], 0, nbytes(c, 8), s));
    les_pcg_result r{};
    const int rc = pcg_fp64(c, &prm->pcg, &r, nullptr, c->W);
    if (rc != LES_OK) return rc;
    adj_account(st, r);
    // u* = BC(M(u^n, nu_t)): m = BC^T y, then the momentum, lagged
    // divergence and nu_t paths back to u^n
    HIPCHK(c, launch_adj_velocity_bcs(s, d, c->solid, yu, yv, yw, mu, mv, mw));
    HIPCHK(c, launch_adj_momentum(s, d, k, mu, mv, mw, u, v, w, nut, c->f[LES_F_P], W, c->solid,
                                  gu, gv, gw, G));
    HIPCHK(c, launch_adj_sgs_gather(s, d, k, G, gu, gv, gw));
    return LES_OK;
}

// Swaps the context's field set with the shadow set for its lifetime.
struct ShadowFields {
    les_ctx* c;
    explicit ShadowFields(les_ctx* ctx) : c(ctx) { std::swap(c->f, c->shadow); }
    ~ShadowFields() { std::swap(c->f, c->shadow); }
};

}  // namespace

extern "C" {

int les_rollout(les_ctx* c, const les_adj_params* prm, les_adj_stats* stats)
{
    if (!c) return LES_ERR_ARG;
    int rc;
    if ((rc = adj_check(c, prm)) != LES_OK) return rc;
    {
        std::vector<uint8_t> inlet((size_t)c->d.sij);
        HIPCHK(c, hipMemcpyAsync(inlet.data(), c->solid, inlet.size(), hipMemcpyDeviceToHost,
                                 c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (std::any_of(inlet.begin(), inlet.end(), [](uint8_t v) { return v != 0; }))
            return fail(c, LES_ERR_STATE, "solid cells on the inlet plane keep stale ping-pong "
                                          "values (main.cpp:340); the step map is not a function "
                                          "of the velocity alone");
    }
    if ((rc = ensure_mirror_weights(c)) != LES_OK) return rc;
    const int every = prm->checkpoint_every;
    const int nck = (prm->nsteps - 1) / every + 1;
    if ((rc = ensure_states(c, &c->ckpt, &c->ckpt_cap, nck)) != LES_OK) return rc;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    c->have_roll = false;
    for (int t = 0; t < prm->nsteps; ++t) {
        if (t % every == 0) {
            hipError_t e = hipSuccess;
            copy_state(c, c->ckpt + (size_t)(t / every) * 3 * c->d.N, c->f[LES_F_U],
                       c->f[LES_F_V], c->f[LES_F_W], e);
            HIPCHK(c, e);
        }
        if ((rc = roll_step(c, prm, stats)) != LES_OK) return rc;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->have_roll = true;
    c->roll_nsteps = prm->nsteps;
    c->roll_every = every;
    c->roll_dt = prm->dt;
    c->roll_nu = prm->nu;
    c->roll_rho = prm->rho;
    c->roll_Cs = prm->Cs;
    return LES_OK;
}

int les_adjoint(les_ctx* c, const les_adj_params* prm, double* ubar, double* vbar, double* wbar,
                int ptr_kind, les_adj_stats* stats)
{
    if (!c) return LES_ERR_ARG;
    int rc;
    if ((rc = adj_check(c, prm)) != LES_OK) return rc;
    if (!ubar || !vbar || !wbar) return fail(c, LES_ERR_ARG, "NULL gradient array");
    if (!c->have_roll || c->roll_nsteps != prm->nsteps || c->roll_every != prm->checkpoint_every ||
        c->roll_dt != prm->dt || c->roll_nu != prm->nu || c->roll_rho != prm->rho ||
        c->roll_Cs != prm->Cs)
        return fail(c, LES_ERR_STATE, "no les_rollout with these parameters since the operator, "
                                      "mask or parameters last changed");
    if ((rc = ensure_mirror_weights(c)) != LES_OK) return rc;
    for (int q = 0; q < kAdjBufs; ++q)
        if ((rc = alloc(c, &c->adjb[q], c->d.N, false)) != LES_OK) return rc;
    for (int q = 0; q < LES_F_COUNT; ++q)
        if ((rc = alloc(c, &c->shadow[q], c->d.N)) != LES_OK) return rc;
    const int every = prm->checkpoint_every;
    if ((rc = ensure_states(c, &c->seg, &c->seg_cap, every)) != LES_OK) return rc;
    if ((rc = ensure_krylov(c)) != LES_OK) return rc;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    const size_t n = (size_t)c->d.N;
    double* gin[3] = {ubar, vbar, wbar};
    for (int q = 0; q < 3; ++q)
        HIPCHK(c, hipMemcpyAsync(c->adjb[q], gin[q], nbytes(c, 8), kind_in(ptr_kind), c->stream));
    {
        ShadowFields sh(c);
        const int nck = (prm->nsteps - 1) / every + 1;
        for (int sg = nck - 1; sg >= 0; --sg) {
            const int len = std::min(every, prm->nsteps - sg * every);
            const double* ck = c->ckpt + (size_t)sg * 3 * n;
            HIPCHK(c, hipMemcpyAsync(c->seg, ck, 3 * nbytes(c, 8), hipMemcpyDeviceToDevice,
                                     c->stream));
            // recompute the states before steps 1..len-1 of the segment
            if (len > 1) {
                HIPCHK(c, hipMemcpyAsync(c->f[LES_F_U], ck, nbytes(c, 8), hipMemcpyDeviceToDevice, c->stream));
                HIPCHK(c, hipMemcpyAsync(c->f[LES_F_V], ck + n, nbytes(c, 8), hipMemcpyDeviceToDevice, c->stream));
                HIPCHK(c, hipMemcpyAsync(c->f[LES_F_W], ck + 2 * n, nbytes(c, 8), hipMemcpyDeviceToDevice, c->stream));
            }
            for (int t = 1; t < len; ++t) {
                if ((rc = roll_step(c, prm, stats)) != LES_OK) return rc;
                hipError_t e = hipSuccess;
                copy_state(c, c->seg + (size_t)t * 3 * n, c->f[LES_F_U], c->f[LES_F_V],
                           c->f[LES_F_W], e);
                HIPCHK(c, e);
            }
            for (int t = len - 1; t >= 0; --t) {
                const double* st = c->seg + (size_t)t * 3 * n;
                if ((rc = adj_step(c, prm, st, st + n, st + 2 * n, stats)) != LES_OK) return rc;
            }
        }
    }
    for (int q = 0; q < 3; ++q)
        HIPCHK(c, hipMemcpyAsync(gin[q], c->adjb[q], nbytes(c, 8), kind_out(ptr_kind), c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LES_OK;
}

}  // extern "C"

extern "C" {

int les_kernel_timing(les_ctx* c, int kernel_class)
{
    if (!c) return LES_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->kt_class = kernel_class;
    c->kt_used = 0;
    c->kt_elem = 0;
    return LES_OK;
}

int les_kernel_timing_read(les_ctx* c, int* launches, double* total_ms, int* elem_bytes)
{
    if (!c || !launches || !total_ms) return fail(c, LES_ERR_ARG, "NULL argument");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double tot = 0.0;
    for (size_t q = 0; q + 1 < c->kt_used; q += 2) {
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, c->kt_events[q], c->kt_events[q + 1]));
        tot += ms;
    }
    *launches = (int)(c->kt_used / 2);
    *total_ms = tot;
    if (elem_bytes) *elem_bytes = c->kt_elem;
    return LES_OK;
}


int les_diag_copy(les_ctx* c, int reps)
{
    if (!c || reps < 0) return LES_ERR_ARG;
    int rc;
    if ((rc = ensure_krylov(c)) != LES_OK) return rc;
    for (int r = 0; r < reps; ++r)
        HIPCHK(c, (launch_convert<double, double>(c->stream, c->d.N, c->t1, c->t2)));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LES_OK;
}

}  // extern "C"

// ------------------------------------------------------------------
// checkpoint (io.cpp:14-120), VTK output (io.cpp:225-256) and time
// statistics (statistics.cpp:70-186) of the device-resident state
// ------------------------------------------------------------------
namespace {

constexpr size_t kIoChunk = size_t(32) << 20;  // bytes per staging chunk

int ensure_stage(les_ctx* c)
{
    for (int q = 0; q < 2; ++q) {
        if (!c->io_stage[q]) HIPCHK(c, hipHostMalloc(&c->io_stage[q], kIoChunk, 0));
        if (!c->io_ev[q]) HIPCHK(c, hipEventCreateWithFlags(&c->io_ev[q], hipEventDisableTiming));
    }
    return LES_OK;
}

std::string join_path(const char* dir, const char* def_dir, const char* name)
{
    return std::string(dir ? dir : def_dir) + "/" + name;
}

// Appends `bytes` device bytes at src to f through the two staging chunks:
// the copy of chunk k+1 runs while chunk k is written.  mask01 writes each
// byte as 1 if nonzero, else 0 (Field::is_solid as io.cpp:71-74 stores it).
int device_to_file(les_ctx* c, FILE* f, const void* src, size_t bytes, bool mask01,
                   const std::string& path)
{
    const size_t nch = (bytes + kIoChunk - 1) / kIoChunk;
    auto enqueue = [&](size_t k) -> hipError_t {
        const size_t off = k * kIoChunk, len = std::min(kIoChunk, bytes - off);
        hipError_t e = hipMemcpyAsync(c->io_stage[k & 1], (const char*)src + off, len,
                                      hipMemcpyDeviceToHost, c->stream);
        return e == hipSuccess ? hipEventRecord(c->io_ev[k & 1], c->stream) : e;
    };
    if (nch) HIPCHK(c, enqueue(0));
    for (size_t k = 0; k < nch; ++k) {
        if (k + 1 < nch) HIPCHK(c, enqueue(k + 1));
        HIPCHK(c, hipEventSynchronize(c->io_ev[k & 1]));
        const size_t len = std::min(kIoChunk, bytes - k * kIoChunk);
        unsigned char* b = (unsigned char*)c->io_stage[k & 1];
        if (mask01)
            for (size_t i = 0; i < len; ++i) b[i] = b[i] ? 1 : 0;
        if (std::fwrite(b, 1, len, f) != len)
            return fail(c, LES_ERR_IO, "write to " + path + " failed");
    }
    return LES_OK;
}

// Reads up to `bytes` from `in` into device memory at dst the way one
// std::istream::read of that size fills an array: a short read overwrites
// the prefix it got and leaves the stream failed, so later reads get
// nothing.  mask: the section is the reference's zero-initialised byte
// buffer turned into is_solid = (byte == 1) (io.cpp:103-107), so every byte
// is written, 0 past the data.  *got receives the bytes read.
int file_to_device(les_ctx* c, std::ifstream& in, void* dst, size_t bytes, bool mask,
                   size_t* got)
{
    size_t total = 0;
    for (size_t off = 0, k = 0; off < bytes; off += kIoChunk, ++k) {
        const size_t len = std::min(kIoChunk, bytes - off);
        HIPCHK(c, hipEventSynchronize(c->io_ev[k & 1]));  // the chunk's last copy is done
        char* b = (char*)c->io_stage[k & 1];
        size_t n = 0;
        if (in) {
            in.read(b, (std::streamsize)len);
            n = (size_t)in.gcount();
        }
        total += n;
        size_t up = n;
        if (mask) {
            for (size_t i = 0; i < n; ++i) b[i] = b[i] == 1 ? 1 : 0;
            std::memset(b + n, 0, len - n);
            up = len;
        }
        if (up) {
            HIPCHK(c, hipMemcpyAsync((char*)dst + off, b, up, hipMemcpyHostToDevice, c->stream));
            HIPCHK(c, hipEventRecord(c->io_ev[k & 1], c->stream));
        }
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *got = total;
    return LES_OK;
}

VtkGrid vtk_grid(const les_ctx* c) { return VtkGrid{c->d.nx, c->d.ny, c->d.nz, c->dx, c->dy, c->dz}; }

constexpr int kVtkFields[5] = {LES_F_U, LES_F_V, LES_F_W, LES_F_P, LES_F_NU_T};

// Waits for a background VTK write; its failure, if any, is kept for les_io_wait.
void join_writer(les_ctx* c)
{
    if (c->io_thread.joinable()) c->io_thread.join();
}

int write_vtk_from_host(les_ctx* c, const std::string& path, int step)
{
    const size_t N = (size_t)c->d.N;
    const double* h = (const double*)c->vtk_host.data();
    std::string err;
    if (vtk_write_fields(path, step, vtk_grid(c), h, h + N, h + 2 * N, h + 3 * N, h + 4 * N,
                         c->vtk_host.data() + 5 * N * sizeof(double), err))
        return fail(c, LES_ERR_IO, err);
    return LES_OK;
}

int ensure_vtk_host(les_ctx* c)
{
    const size_t bytes = (size_t)c->d.N * (5 * sizeof(double) + 1);
    if (c->vtk_host.size() != bytes) {
        try {
            c->vtk_host.resize(bytes);
        } catch (const std::bad_alloc&) {
            return fail(c, LES_ERR_NOMEM, "host buffer for the VTK fields");
        }
    }
    return LES_OK;
}

std::string vtk_path(const char* path, int step)
{
    return path ? std::string(path) : "output/output_" + std::to_string(step) + ".vtk";
}

int ensure_stats(les_ctx* c, double** arr)
{
    int rc;
    for (int a = 0; a < kStatArrays; ++a)
        if ((rc = alloc(c, &arr[a], c->d.N)) != LES_OK) return rc;
    return LES_OK;
}

}  // namespace

extern "C" {

int les_save_state(les_ctx* c, const char* dir, double time_simulated, double dt,
                   int next_step_to_run)
{
    if (!c) return LES_ERR_ARG;
    int rc;
    if ((rc = ensure_stage(c)) != LES_OK) return rc;
    std::string err;
    if (checkpoint_write_meta(join_path(dir, "checkpoint", "checkpoint.meta"), next_step_to_run,
                              time_simulated, dt, err))
        return fail(c, LES_ERR_IO, err);
    const std::string path = join_path(dir, "checkpoint", "checkpoint.data");
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return fail(c, LES_ERR_IO, "cannot open " + path + " for writing");
    for (int q : kVtkFields)
        if ((rc = device_to_file(c, f, c->f[q], nbytes(c, 8), false, path)) != LES_OK) {
            std::fclose(f);
            return rc;
        }
    rc = device_to_file(c, f, c->solid, nbytes(c, 1), true, path);
    if (std::fclose(f) != 0 && rc == LES_OK) rc = fail(c, LES_ERR_IO, "closing " + path + " failed");
    return rc;
}

int les_load_state(les_ctx* c, const char* dir, double* time_simulated, double* dt,
                   int* next_step)
{
    if (!c || !time_simulated || !dt || !next_step) return fail(c, LES_ERR_ARG, "NULL argument");
    *next_step = 0;
    std::ifstream meta(join_path(dir, "checkpoint", "checkpoint.meta"));
    if (!meta) return LES_OK;  // no checkpoint: start from scratch
    int start = 0;
    meta >> start;
    meta >> *time_simulated;
    meta >> *dt;
    meta.close();
    std::ifstream data(join_path(dir, "checkpoint", "checkpoint.data"), std::ios::binary);
    if (!data) {
        c->err = "found checkpoint.meta but not checkpoint.data";
        return LES_OK;
    }
    int rc;
    if ((rc = ensure_stage(c)) != LES_OK) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    size_t got = 0;
    for (int q : kVtkFields)
        if ((rc = file_to_device(c, data, c->f[q], nbytes(c, 8), false, &got)) != LES_OK) return rc;
    if ((rc = file_to_device(c, data, c->solid, nbytes(c, 1), true, &got)) != LES_OK) return rc;
    c->have_roll = false;
    if (got == 0) {
        c->err = "error reading checkpoint data: the file might be corrupt";
        return LES_OK;
    }
    *next_step = start;
    return LES_OK;
}

int les_write_field_vtk(les_ctx* c, const char* path, int step)
{
    if (!c) return LES_ERR_ARG;
    join_writer(c);
    int rc;
    if ((rc = ensure_vtk_host(c)) != LES_OK) return rc;
    const size_t N = (size_t)c->d.N;
    unsigned char* h = c->vtk_host.data();
    for (int q = 0; q < 5; ++q)
        HIPCHK(c, hipMemcpyAsync(h + q * N * sizeof(double), c->f[kVtkFields[q]], N * sizeof(double),
                                 hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(h + 5 * N * sizeof(double), c->solid, N, hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return write_vtk_from_host(c, vtk_path(path, step), step);
}

int les_write_field_vtk_async(les_ctx* c, const char* path, int step)
{
    if (!c) return LES_ERR_ARG;
    join_writer(c);
    int rc;
    if ((rc = ensure_vtk_host(c)) != LES_OK) return rc;
    const size_t N = (size_t)c->d.N, fb = N * sizeof(double);
    if (!c->vtk_snap && (rc = alloc(c, &c->vtk_snap, 5 * fb + N, false)) != LES_OK) return rc;
    if (!c->io_stream) HIPCHK(c, hipStreamCreateWithFlags(&c->io_stream, hipStreamNonBlocking));
    if (!c->snap_ev) HIPCHK(c, hipEventCreateWithFlags(&c->snap_ev, hipEventDisableTiming));
    // the snapshot is ordered after the work already queued on the context's
    // stream and before whatever is queued after this call
    for (int q = 0; q < 5; ++q)
        HIPCHK(c, hipMemcpyAsync(c->vtk_snap + q * fb, c->f[kVtkFields[q]], fb,
                                 hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->vtk_snap + 5 * fb, c->solid, N, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipEventRecord(c->snap_ev, c->stream));
    const std::string p = vtk_path(path, step);
    try {
        c->io_thread = std::thread([c, p, step, N, fb]() {
            hipError_t e = hipSetDevice(c->device);
            if (e == hipSuccess) e = hipStreamWaitEvent(c->io_stream, c->snap_ev, 0);
            if (e == hipSuccess)
                e = hipMemcpyAsync(c->vtk_host.data(), c->vtk_snap, 5 * fb + N,
                                   hipMemcpyDeviceToHost, c->io_stream);
            if (e == hipSuccess) e = hipStreamSynchronize(c->io_stream);
            std::string err;
            if (e != hipSuccess) {
                err = std::string("background VTK copy: ") + hipGetErrorString(e);
            } else {
                const double* h = (const double*)c->vtk_host.data();
                vtk_write_fields(p, step, vtk_grid(c), h, h + N, h + 2 * N, h + 3 * N, h + 4 * N,
                                 c->vtk_host.data() + 5 * fb, err);
            }
            if (!err.empty() && c->io_rc == 0) {
                c->io_rc = e != hipSuccess ? LES_ERR_HIP : LES_ERR_IO;
                c->io_err = err;
            }
        });
    } catch (const std::system_error& ex) {
        return fail(c, LES_ERR_STATE, std::string("cannot start the writer thread: ") + ex.what());
    }
    return LES_OK;
}

int les_io_wait(les_ctx* c)
{
    if (!c) return LES_ERR_ARG;
    join_writer(c);
    const int rc = c->io_rc;
    if (rc != LES_OK) fail(c, rc, c->io_err);
    c->io_rc = 0;
    c->io_err.clear();
    return rc;
}

int les_write_field_vtk_host(const char* path, int step, int nx, int ny, int nz, double dx,
                             double dy, double dz, const double* u, const double* v,
                             const double* w, const double* p, const double* nu_t,
                             const uint8_t* solid)
{
    if (!u || !v || !w || !p || !nu_t || !solid || nx < 1 || ny < 1 || nz < 1)
        return fail(nullptr, LES_ERR_ARG, "bad grid or NULL array");
    if ((long long)nx * ny * nz >= (1LL << 31))
        return fail(nullptr, LES_ERR_ARG, "grid exceeds 2^31 cells (int indices)");
    std::string err;
    if (vtk_write_fields(vtk_path(path, step), step, VtkGrid{nx, ny, nz, dx, dy, dz}, u, v, w, p,
                         nu_t, solid, err))
        return fail(nullptr, LES_ERR_IO, err);
    return LES_OK;
}

int les_stats_init(les_ctx* c, int start_step)
{
    if (!c) return LES_ERR_ARG;
    int rc;
    if ((rc = ensure_stats(c, c->stat_sum)) != LES_OK) return rc;
    for (int a = 0; a < kStatArrays; ++a)
        HIPCHK(c, hipMemsetAsync(c->stat_sum[a], 0, nbytes(c, 8), c->stream));
    c->stat_start = start_step;
    c->stat_samples = 0;
    c->have_stats = true;
    c->stat_final = false;
    return LES_OK;
}

int les_stats_update(les_ctx* c, int current_step)
{
    if (!c) return LES_ERR_ARG;
    if (!c->have_stats) return fail(c, LES_ERR_STATE, "les_stats_init has not been called");
    if (current_step <= c->stat_start) return LES_OK;
    ++c->stat_samples;
    StatArrays S;
    for (int a = 0; a < kStatArrays; ++a) S.s[a] = c->stat_sum[a];
    KtScope kt(c, LES_KT_STATS, 8);
    HIPCHK(c, launch_stats_update(c->stream, c->d.N, c->f[LES_F_U], c->f[LES_F_V], c->f[LES_F_W],
                                  c->f[LES_F_P], S));
    return LES_OK;
}

int les_stats_finalize(les_ctx* c, const char* path)
{
    if (!c) return LES_ERR_ARG;
    if (!c->have_stats) return fail(c, LES_ERR_STATE, "les_stats_init has not been called");
    if (c->stat_samples == 0) return LES_OK;  // statistics.cpp:135-138: nothing to save
    int rc;
    if ((rc = ensure_stats(c, c->stat_out)) != LES_OK) return rc;
    StatArrays S, O;
    for (int a = 0; a < kStatArrays; ++a) {
        S.s[a] = c->stat_sum[a];
        O.s[a] = c->stat_out[a];
    }
    const double n_inv = 1.0 / static_cast<double>(c->stat_samples);
    HIPCHK(c, launch_stats_finalize(c->stream, c->d.N, S, n_inv, O));
    c->stat_final = true;
    if (path && !*path) return LES_OK;  // "": compute only
    const size_t N = (size_t)c->d.N;
    std::vector<double> h;
    try {
        h.resize(kStatArrays * N);
    } catch (const std::bad_alloc&) {
        return fail(c, LES_ERR_NOMEM, "host buffer for the statistics");
    }
    const double* m[kStatArrays];
    for (int a = 0; a < kStatArrays; ++a) {
        HIPCHK(c, hipMemcpyAsync(h.data() + a * N, c->stat_out[a], N * sizeof(double),
                                 hipMemcpyDeviceToHost, c->stream));
        m[a] = h.data() + a * N;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::string err;
    if (vtk_write_statistics(path ? path : "output/statistics_final.vtk", vtk_grid(c), m, err))
        return fail(c, LES_ERR_IO, err);
    return LES_OK;
}

int les_stats_get(les_ctx* c, int which, int final, double* out, int ptr_kind)
{
    if (!c || !out || which < 0 || which >= kStatArrays)
        return fail(c, LES_ERR_ARG, "bad statistics slot or NULL out");
    if (!c->have_stats || (final && !c->stat_final))
        return fail(c, LES_ERR_STATE, final ? "statistics not finalised" : "no statistics");
    const double* src = final ? c->stat_out[which] : c->stat_sum[which];
    HIPCHK(c, hipMemcpyAsync(out, src, nbytes(c, 8), kind_out(ptr_kind), c->stream));
    if (!ptr_kind) HIPCHK(c, hipStreamSynchronize(c->stream));
    return LES_OK;
}

int les_stats_samples(les_ctx* c, int* num_samples)
{
    if (!c || !num_samples) return fail(c, LES_ERR_ARG, "NULL argument");
    *num_samples = c->have_stats ? c->stat_samples : 0;
    return LES_OK;
}

}  // extern "C"
