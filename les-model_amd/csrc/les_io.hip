// les_io.hip -- time statistics on the device and the host-side writers of
// the checkpoint, VTK and statistics files (src/io.cpp, src/statistics.cpp).
//
// Number formatting.  The reference writes doubles with std::ostream's
// default floatfield, i.e. printf's %.<precision>g in the "C" locale
// (precision 6 for io.cpp's VTK, 10 for the statistics file after its
// header, 17 = max_digits10 for checkpoint.meta).  std::to_chars with
// chars_format::general and the same precision is specified to give exactly
// that text, without the locale and stream machinery, so the files here are
// byte-identical and are formatted on several threads at once.
#include "les_io.hpp"

#include <algorithm>
#include <charconv>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <system_error>
#include <thread>
#include <vector>

namespace les {

namespace {

// ---- statistics kernels: two cells per thread (16-byte loads), one
// workgroup per CU on a grid-stride loop (measured at 256^3 on MI355X: 256
// workgroups 638 us, 2048 workgroups 679 us) ----
#ifndef LES_STAT_GRID
#define LES_STAT_GRID 256
#endif
constexpr int kStatGrid = LES_STAT_GRID;

__global__ __launch_bounds__(kBlock) void k_stats_update(int n, const double* __restrict__ u,
                                                         const double* __restrict__ v,
                                                         const double* __restrict__ w,
                                                         const double* __restrict__ p,
                                                         StatArrays S)
{
    auto cell = [&](double uu, double vv, double ww, double pp, int i, auto&& ld, auto&& st) {
        const double add[kStatArrays] = {uu,      vv,      ww,      pp,      uu * uu,
                                         vv * vv, ww * ww, uu * vv, uu * ww, vv * ww};
#pragma unroll
        for (int a = 0; a < kStatArrays; ++a) st(a, i, ld(a, i) + add[a]);
    };
    const int n2 = n / 2;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n2; i += gridDim.x * kBlock) {
        const double2 u2 = reinterpret_cast<const double2*>(u)[i];
        const double2 v2 = reinterpret_cast<const double2*>(v)[i];
        const double2 w2 = reinterpret_cast<const double2*>(w)[i];
        const double2 p2 = reinterpret_cast<const double2*>(p)[i];
        double2 s2[kStatArrays];
#pragma unroll
        for (int a = 0; a < kStatArrays; ++a) s2[a] = reinterpret_cast<const double2*>(S.s[a])[i];
        const double ax[kStatArrays] = {u2.x,        v2.x,        w2.x,        p2.x,
                                        u2.x * u2.x, v2.x * v2.x, w2.x * w2.x, u2.x * v2.x,
                                        u2.x * w2.x, v2.x * w2.x};
        const double ay[kStatArrays] = {u2.y,        v2.y,        w2.y,        p2.y,
                                        u2.y * u2.y, v2.y * v2.y, w2.y * w2.y, u2.y * v2.y,
                                        u2.y * w2.y, v2.y * w2.y};
#pragma unroll
        for (int a = 0; a < kStatArrays; ++a) {
            s2[a].x = s2[a].x + ax[a];
            s2[a].y = s2[a].y + ay[a];
            reinterpret_cast<double2*>(S.s[a])[i] = s2[a];
        }
    }
    if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
        const int i = n - 1;
        cell(u[i], v[i], w[i], p[i], i, [&](int a, int k) { return S.s[a][k]; },
             [&](int a, int k, double x) { S.s[a][k] = x; });
    }
}

__device__ __forceinline__ void stats_final_cell(const double (&s)[kStatArrays], double n_inv,
                                                 double (&o)[kStatArrays])
{
    const double um = s[0] * n_inv, vm = s[1] * n_inv, wm = s[2] * n_inv, pm = s[3] * n_inv;
    o[0] = um;
    o[1] = vm;
    o[2] = wm;
    o[3] = pm;
    o[4] = (s[4] * n_inv) - (um * um);
    o[5] = (s[5] * n_inv) - (vm * vm);
    o[6] = (s[6] * n_inv) - (wm * wm);
    o[7] = (s[7] * n_inv) - (um * vm);
    o[8] = (s[8] * n_inv) - (um * wm);
    o[9] = (s[9] * n_inv) - (vm * wm);
}

__global__ __launch_bounds__(kBlock) void k_stats_finalize(int n, StatArrays S, double n_inv,
                                                           StatArrays O)
{
    const int n2 = n / 2;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n2; i += gridDim.x * kBlock) {
        double sx[kStatArrays], sy[kStatArrays], ox[kStatArrays], oy[kStatArrays];
#pragma unroll
        for (int a = 0; a < kStatArrays; ++a) {
            const double2 t = reinterpret_cast<const double2*>(S.s[a])[i];
            sx[a] = t.x;
            sy[a] = t.y;
        }
        stats_final_cell(sx, n_inv, ox);
        stats_final_cell(sy, n_inv, oy);
#pragma unroll
        for (int a = 0; a < kStatArrays; ++a)
            reinterpret_cast<double2*>(O.s[a])[i] = make_double2(ox[a], oy[a]);
    }
    if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
        const int i = n - 1;
        double s[kStatArrays], o[kStatArrays];
        for (int a = 0; a < kStatArrays; ++a) s[a] = S.s[a][i];
        stats_final_cell(s, n_inv, o);
        for (int a = 0; a < kStatArrays; ++a) O.s[a][i] = o[a];
    }
}

int stat_grid(int n)
{
    const int need = (n / 2 + kBlock - 1) / kBlock;
    return std::max(1, std::min(kStatGrid, need));
}

// ---- formatting ----
constexpr int kNumMax = 40;  // bytes of one formatted double, with margin

inline char* put_g(char* p, double v, int prec)
{
    return std::to_chars(p, p + kNumMax, v, std::chars_format::general, prec).ptr;
}

std::string str_g(double v, int prec)
{
    char b[kNumMax];
    return std::string(b, put_g(b, v, prec));
}

struct FileCloser {
    void operator()(FILE* f) const
    {
        if (f) std::fclose(f);
    }
};
using File = std::unique_ptr<FILE, FileCloser>;

File open_w(const std::string& path, std::string& err)
{
    File f(std::fopen(path.c_str(), "wb"));
    if (!f) err = "cannot open " + path + " for writing: " + std::strerror(errno);
    else std::setvbuf(f.get(), nullptr, _IOFBF, 1 << 22);
    return f;
}

int put_str(FILE* f, const std::string& s, const std::string& path, std::string& err)
{
    if (std::fwrite(s.data(), 1, s.size(), f) != s.size()) {
        err = "write to " + path + " failed";
        return -1;
    }
    return 0;
}

// n lines, line i formatted by fmt(out, i) -> end (at most max_line bytes),
// written in order.  Blocks of lines are formatted by io_threads() threads,
// each into its own buffer, then written in thread order.
template <typename F>
int write_lines(FILE* f, size_t n, size_t max_line, F fmt, const std::string& path,
                std::string& err)
{
    const int T = io_threads();
    const size_t block = size_t(1) << 20;
    const size_t per = (block + T - 1) / T;
    std::vector<std::vector<char>> buf(T, std::vector<char>(per * max_line));
    std::vector<size_t> len(T, 0);
    for (size_t b0 = 0; b0 < n; b0 += block) {
        const size_t b1 = std::min(n, b0 + block);
        auto work = [&, b0, b1](int t) {
            const size_t s = b0 + (b1 - b0) * t / T, e = b0 + (b1 - b0) * (t + 1) / T;
            char* p = buf[t].data();
            for (size_t i = s; i < e; ++i) p = fmt(p, i);
            len[t] = (size_t)(p - buf[t].data());
        };
        std::vector<std::thread> th;
        for (int t = 1; t < T; ++t) {
            try {
                th.emplace_back(work, t);
            } catch (const std::system_error&) {
                work(t);  // no thread available: format this part here
            }
        }
        work(0);
        for (auto& x : th) x.join();
        for (int t = 0; t < T; ++t)
            if (std::fwrite(buf[t].data(), 1, len[t], f) != len[t]) {
                err = "write to " + path + " failed";
                return -1;
            }
    }
    return 0;
}

// The header both VTK writers share (io.cpp:228-235, statistics.cpp:26-33);
// SPACING uses the stream's default precision 6.
std::string vtk_header(const std::string& title, const VtkGrid& g)
{
    return "# vtk DataFile Version 3.0\n" + title + "\nASCII\nDATASET STRUCTURED_POINTS\n" +
           "DIMENSIONS " + std::to_string(g.nx) + " " + std::to_string(g.ny) + " " +
           std::to_string(g.nz) + "\nORIGIN 0 0 0\nSPACING " + str_g(g.dx, 6) + " " +
           str_g(g.dy, 6) + " " + str_g(g.dz, 6) + "\nPOINT_DATA " +
           std::to_string(g.nx * g.ny * g.nz) + "\n";
}

int write_scalars(FILE* f, const std::string& head, const double* a, size_t n, int prec,
                  const std::string& path, std::string& err)
{
    if (put_str(f, head, path, err)) return -1;
    return write_lines(
        f, n, kNumMax + 1,
        [&](char* p, size_t i) {
            p = put_g(p, a[i], prec);
            *p++ = '\n';
            return p;
        },
        path, err);
}

int write_vectors(FILE* f, const std::string& head, const double* u, const double* v,
                  const double* w, size_t n, int prec, const std::string& path,
                  std::string& err)
{
    if (put_str(f, head, path, err)) return -1;
    return write_lines(
        f, n, 3 * (kNumMax + 1),
        [&](char* p, size_t i) {
            p = put_g(p, u[i], prec);
            *p++ = ' ';
            p = put_g(p, v[i], prec);
            *p++ = ' ';
            p = put_g(p, w[i], prec);
            *p++ = '\n';
            return p;
        },
        path, err);
}

int close_file(File& f, const std::string& path, std::string& err)
{
    FILE* raw = f.release();
    if (std::fclose(raw) != 0) {
        err = "closing " + path + " failed";
        return -1;
    }
    return 0;
}

}  // namespace

int io_threads()
{
    if (const char* e = std::getenv("LES_IO_THREADS")) {
        const int t = std::atoi(e);
        if (t > 0) return std::min(t, 256);
    }
    const int hw = (int)std::thread::hardware_concurrency();
    return std::max(1, std::min(16, hw));
}

hipError_t launch_stats_update(hipStream_t s, int n, const double* u, const double* v,
                               const double* w, const double* p, StatArrays sums)
{
    hipLaunchKernelGGL(k_stats_update, dim3(stat_grid(n)), dim3(kBlock), 0, s, n, u, v, w, p,
                       sums);
    return hipGetLastError();
}

hipError_t launch_stats_finalize(hipStream_t s, int n, StatArrays sums, double n_inv,
                                 StatArrays out)
{
    hipLaunchKernelGGL(k_stats_finalize, dim3(stat_grid(n)), dim3(kBlock), 0, s, n, sums, n_inv,
                       out);
    return hipGetLastError();
}

int vtk_write_fields(const std::string& path, int step, const VtkGrid& g, const double* u,
                     const double* v, const double* w, const double* p, const double* nu_t,
                     const uint8_t* solid, std::string& err)
{
    File f = open_w(path, err);
    if (!f) return -1;
    const size_t n = (size_t)g.nx * g.ny * g.nz;
    if (put_str(f.get(), vtk_header("LES output step " + std::to_string(step), g), path, err))
        return -1;
    if (write_vectors(f.get(), "VECTORS velocity double\n", u, v, w, n, 6, path, err)) return -1;
    if (write_scalars(f.get(), "SCALARS pressure double 1\nLOOKUP_TABLE default\n", p, n, 6,
                      path, err))
        return -1;
    if (write_scalars(f.get(), "SCALARS nu_t double 1\nLOOKUP_TABLE default\n", nu_t, n, 6, path,
                      err))
        return -1;
    if (put_str(f.get(), "SCALARS solid int 1\nLOOKUP_TABLE default\n", path, err)) return -1;
    if (write_lines(
            f.get(), n, 2,
            [&](char* o, size_t i) {
                *o++ = solid[i] ? '1' : '0';
                *o++ = '\n';
                return o;
            },
            path, err))
        return -1;
    return close_file(f, path, err);
}

int vtk_write_statistics(const std::string& path, const VtkGrid& g, const double* const* m,
                         std::string& err)
{
    static const char* const kStress[6] = {"uu", "vv", "ww", "uv", "uw", "vw"};
    File f = open_w(path, err);
    if (!f) return -1;
    const size_t n = (size_t)g.nx * g.ny * g.nz;
    if (put_str(f.get(), vtk_header("Time-averaged statistics", g), path, err)) return -1;
    if (write_vectors(f.get(), "VECTORS mean_velocity double\n", m[0], m[1], m[2], n, 10, path,
                      err))
        return -1;
    if (write_scalars(f.get(), "SCALARS mean_pressure double 1\nLOOKUP_TABLE default\n", m[3], n,
                      10, path, err))
        return -1;
    for (int s = 0; s < 6; ++s)
        if (write_scalars(f.get(),
                          std::string("SCALARS ") + kStress[s] +
                              "_stress double 1\nLOOKUP_TABLE default\n",
                          m[4 + s], n, 10, path, err))
            return -1;
    return close_file(f, path, err);
}

int checkpoint_write_meta(const std::string& path, int next_step, double time_simulated,
                          double dt, std::string& err)
{
    File f = open_w(path, err);
    if (!f) return -1;
    const std::string s = std::to_string(next_step) + "\n" + str_g(time_simulated, 17) + "\n" +
                          str_g(dt, 17) + "\n";
    if (put_str(f.get(), s, path, err)) return -1;
    return close_file(f, path, err);
}

}  // namespace les
