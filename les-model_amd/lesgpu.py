"""Python binding of the MI355X LES hot-path library (liblesgpu.so).

A thin ctypes layer over the C ABI in include/les_gpu.h.  Host arrays are
numpy float64/uint8; device arrays may be passed as torch tensors (their
data_ptr() is handed over with ptr_kind = device).  The library is required:
importing this module on a machine without the built library raises, there
is no CPU fallback.

The method names follow the reference's functions (src/pressure.h,
src/main.cpp) they replace.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblesgpu.so")

LES_PTR_HOST, LES_PTR_DEVICE = 0, 1
FIELDS = ("u", "v", "w", "u_new", "v_new", "w_new", "p", "p_rhs", "nu_t", "b")
FIELD_ID = {n: i for i, n in enumerate(FIELDS)}
STENCIL_A, STENCIL_L = 0, 1
STENCIL_AUTO, STENCIL_ARRAYS, STENCIL_CODED = 0, 1, 2
PC_NONE, PC_JACOBI, PC_IC0, PC_LEGACY_IC_A, PC_CHEBYSHEV, PC_MG = 0, 1, 2, 3, 4, 5
PREC_FP64, PREC_MIXED = 0, 1
SOLVER_PCG, SOLVER_SOR_LEX, SOLVER_SOR_RB, SOLVER_JACOBI = 0, 1, 2, 3
KT_NONE, KT_CHEB_STEP, KT_PCG_MATVEC, KT_PCG_UPDATE, KT_MOMENTUM, KT_IC_SOLVE = 0, 1, 2, 3, 4, 5
KT_CHEB_PAIR, KT_STATS, KT_IC_FACTOR, KT_SOR_SWEEP, KT_MG_SWEEP = 6, 7, 8, 9, 10
IP_EUCLID, IP_MIRROR = 0, 1
OPT_CHEB_PAIRS, OPT_PAIR_GRID, OPT_WAVEFRONT, OPT_SPIN_LIMIT = 1, 2, 3, 4
OPT_MG_SWEEPS, OPT_MG_COARSE_SWEEPS, OPT_REPLACE_DECADES, OPT_MG_RINGS = 5, 6, 7, 8
ERR_ARG, ERR_HIP, ERR_STATE, ERR_NOMEM, ERR_IO, ERR_STALL = -1, -2, -3, -4, -5, -6
# statistics slots (statistics.h:15-31): means of u, v, w, p, then the sums
# of the six products
STATS = ("u", "v", "w", "p", "uu", "vv", "ww", "uv", "uw", "vw")
STAT_ID = {n: i for i, n in enumerate(STATS)}

DP = C.POINTER(C.c_double)


class PcgParams(C.Structure):
    _fields_ = [("precond", C.c_int), ("max_iter", C.c_int), ("tol", C.c_double),
                ("tol_is_relative", C.c_int), ("lambda_min", C.c_double),
                ("lambda_max", C.c_double), ("degree", C.c_int), ("precision", C.c_int),
                ("reliable_every", C.c_int), ("check_every", C.c_int),
                ("stencil_format", C.c_int), ("inner_product", C.c_int)]


class PcgResult(C.Structure):
    _fields_ = [("iters", C.c_int), ("converged", C.c_int), ("res_norm", C.c_double),
                ("res0_norm", C.c_double), ("b_norm", C.c_double),
                ("true_res_norm", C.c_double), ("replacements", C.c_int), ("fault", C.c_int),
                ("restarts", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class StepParams(C.Structure):
    _fields_ = [("step", C.c_int), ("dt_fixed", C.c_double), ("nu", C.c_double),
                ("rho", C.c_double), ("Cs", C.c_double), ("solver", C.c_int),
                ("pcg", PcgParams), ("omega", C.c_double), ("sor_sweeps", C.c_int)]


class StepStats(C.Structure):
    _fields_ = [("dt", C.c_double), ("pcg", PcgResult), ("sweeps", C.c_int),
                ("speed_max", C.c_double), ("nu_t_max", C.c_double)]


class AdjParams(C.Structure):
    _fields_ = [("nu", C.c_double), ("rho", C.c_double), ("Cs", C.c_double), ("dt", C.c_double),
                ("nsteps", C.c_int), ("checkpoint_every", C.c_int), ("pcg", PcgParams)]


class AdjStats(C.Structure):
    _fields_ = [("solves", C.c_int), ("iters", C.c_int), ("unconverged", C.c_int),
                ("max_rel_res", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class LesError(RuntimeError):
    pass


_lib = None


def load_library(path: str = LIB_PATH):
    """Loads liblesgpu.so (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise LesError(f"{path} is not built: run __graft_entry__.build() "
                       "(make -C les-model_amd); there is no CPU fallback")
    # PyTorch-ROCm carries its own HIP runtime.  When this library (linked
    # against the system one) loads first and torch initialises its GPU
    # afterwards, this library's runtime no longer finds the device; with
    # torch's runtime loaded first, both share it.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(path)
    vp = C.c_void_p
    L.les_ctx_create.argtypes = [C.c_int] * 3 + [C.c_double] * 3 + [C.POINTER(vp)]
    L.les_ctx_destroy.argtypes = [vp]
    L.les_last_error.argtypes = [vp]
    L.les_last_error.restype = C.c_char_p
    L.les_set_stream.argtypes = [vp, vp]
    L.les_synchronize.argtypes = [vp]
    L.les_set_option.argtypes = [vp, C.c_int, C.c_int]
    L.les_set_field.argtypes = [vp, C.c_int, vp, C.c_int]
    L.les_get_field.argtypes = [vp, C.c_int, vp, C.c_int]
    L.les_field_ptr.argtypes = [vp, C.c_int, C.POINTER(vp)]
    L.les_set_mask.argtypes = [vp, vp, C.c_int]
    L.les_get_mask.argtypes = [vp, vp, C.c_int]
    L.les_build_stencil.argtypes = [vp, C.c_int]
    L.les_set_stencil.argtypes = [vp, C.c_int, C.POINTER(vp), C.c_int]
    L.les_get_stencil.argtypes = [vp, C.c_int, C.POINTER(vp), C.c_int]
    L.les_matvec.argtypes = [vp, C.c_int, vp, vp, C.c_int]
    L.les_dot.argtypes = [vp, vp, vp, C.c_int, DP]
    L.les_dot_noalloc.argtypes = [C.c_int, C.c_int, C.c_int, vp, vp, vp, vp]
    L.les_ic0_factor.argtypes = [vp]
    L.les_precond_apply.argtypes = [vp, C.c_int, vp, vp, C.c_int, C.c_double, C.c_double,
                                    C.c_int]
    L.les_precond_apply_fp32.argtypes = [vp, C.c_int, vp, vp, C.c_int, C.c_double, C.c_double,
                                         C.c_int, C.c_int, C.c_int]
    L.les_power_lambda_max.argtypes = [vp, C.c_int, DP]
    L.les_lanczos_bounds.argtypes = [vp, C.c_int, C.c_uint64, DP, DP]
    L.les_pcg.argtypes = [vp, C.POINTER(PcgParams), C.POINTER(PcgResult)]
    L.les_pcg_trace.argtypes = [vp, DP, C.c_int]
    L.les_pressure_bcs.argtypes = [vp]
    L.les_jacobi_solve.argtypes = [vp, C.c_double, C.c_double, C.POINTER(C.c_int)]
    L.les_sor_solve.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                C.POINTER(C.c_int)]
    L.les_sgs.argtypes = [vp, C.c_double, C.c_int, C.POINTER(vp), C.c_int]
    L.les_nut_bcs.argtypes = [vp]
    L.les_momentum.argtypes = [vp, C.c_double, C.c_double, C.c_int]
    L.les_velocity_bcs.argtypes = [vp]
    L.les_divergence_b.argtypes = [vp, C.c_double, C.c_double]
    L.les_correct.argtypes = [vp, C.c_double, C.c_double, C.c_int]
    L.les_adaptive_dt.argtypes = [vp, C.c_double, DP]
    L.les_step.argtypes = [vp, C.POINTER(StepParams), C.POINTER(StepStats)]
    L.les_rollout.argtypes = [vp, C.POINTER(AdjParams), C.POINTER(AdjStats)]
    L.les_adjoint.argtypes = [vp, C.POINTER(AdjParams), vp, vp, vp, C.c_int, C.POINTER(AdjStats)]
    L.les_kernel_timing.argtypes = [vp, C.c_int]
    L.les_kernel_timing_read.argtypes = [vp, C.POINTER(C.c_int), DP, C.POINTER(C.c_int)]
    L.les_diag_copy.argtypes = [vp, C.c_int]
    cp = C.c_char_p
    L.les_save_state.argtypes = [vp, cp, C.c_double, C.c_double, C.c_int]
    L.les_load_state.argtypes = [vp, cp, DP, DP, C.POINTER(C.c_int)]
    L.les_write_field_vtk.argtypes = [vp, cp, C.c_int]
    L.les_write_field_vtk_async.argtypes = [vp, cp, C.c_int]
    L.les_io_wait.argtypes = [vp]
    L.les_write_field_vtk_host.argtypes = [cp, C.c_int] + [C.c_int] * 3 + [C.c_double] * 3 + [vp] * 6
    L.les_stats_init.argtypes = [vp, C.c_int]
    L.les_stats_update.argtypes = [vp, C.c_int]
    L.les_stats_finalize.argtypes = [vp, cp]
    L.les_stats_get.argtypes = [vp, C.c_int, C.c_int, vp, C.c_int]
    L.les_stats_samples.argtypes = [vp, C.POINTER(C.c_int)]
    _lib = L
    return L


def _host_ptr(a, dtype):
    assert isinstance(a, np.ndarray) and a.dtype == dtype and a.flags["C_CONTIGUOUS"], (
        a.dtype, dtype)
    return a.ctypes.data_as(C.c_void_p)


def _ptr(a, dtype=np.float64):
    """(pointer, ptr_kind) for a numpy array or a torch CUDA tensor."""
    if isinstance(a, np.ndarray):
        return _host_ptr(a, dtype), LES_PTR_HOST
    # torch tensor
    if not a.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return C.c_void_p(a.data_ptr()), (LES_PTR_DEVICE if a.is_cuda else LES_PTR_HOST)


def _path(p):
    return None if p is None else os.fsencode(p)


def dot_noalloc(nx, ny, nz, a, b, out, stream=None):
    """les_dot_noalloc: <a, b> of two device arrays (torch CUDA tensors)
    into the device tensor `out` (first element), ordered on `stream` (a
    hipStream_t handle; None: the null stream), with no allocation."""
    L = load_library()
    for t in (a, b, out):
        if not getattr(t, "is_cuda", False) or not t.is_contiguous():
            raise ValueError("les_dot_noalloc takes contiguous device tensors")
    rc = L.les_dot_noalloc(int(nx), int(ny), int(nz), C.c_void_p(a.data_ptr()),
                           C.c_void_p(b.data_ptr()), C.c_void_p(out.data_ptr()),
                           C.c_void_p(stream))
    if rc != 0:
        raise LesError(f"les_dot_noalloc: {L.les_last_error(None).decode()} ({rc})")


def write_field_vtk_host(path, step, m, u, v, w, p, nu_t, solid):
    """write_field_vtk (io.cpp:225-256) from host arrays with the library's
    multi-threaded writer; no GPU is used."""
    L = load_library()
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (u, v, w, p, nu_t)]
    sm = np.ascontiguousarray(solid, dtype=np.uint8)
    rc = L.les_write_field_vtk_host(_path(path), int(step), m["nx"], m["ny"], m["nz"], m["dx"],
                                    m["dy"], m["dz"], *[_host_ptr(a, np.float64) for a in arrs],
                                    _host_ptr(sm, np.uint8))
    if rc != 0:
        raise LesError(f"les_write_field_vtk_host: {L.les_last_error(None).decode()} ({rc})")


class LesGpu:
    """One device-resident LES context (Field + stencil + solver state)."""

    def __init__(self, nx, ny, nz, dx, dy, dz):
        self.lib = load_library()
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.N = self.nx * self.ny * self.nz
        h = C.c_void_p()
        rc = self.lib.les_ctx_create(self.nx, self.ny, self.nz, dx, dy, dz, C.byref(h))
        if rc != 0:
            raise LesError(f"les_ctx_create: {self.lib.les_last_error(None).decode()} ({rc})")
        self.h = h

    @classmethod
    def from_mesh(cls, m):
        return cls(m["nx"], m["ny"], m["nz"], m["dx"], m["dy"], m["dz"])

    def close(self):
        if getattr(self, "h", None):
            self.lib.les_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc, what):
        if rc != 0:
            err = LesError(f"{what}: {self.lib.les_last_error(self.h).decode()} ({rc})")
            err.code = rc
            raise err

    # ---- fields ----
    def set_field(self, name, a):
        p, k = _ptr(a)
        self._chk(self.lib.les_set_field(self.h, FIELD_ID[name], p, k), "les_set_field")

    def get_field(self, name, out=None):
        out = np.empty(self.N) if out is None else out
        p, k = _ptr(out)
        self._chk(self.lib.les_get_field(self.h, FIELD_ID[name], p, k), "les_get_field")
        return out

    def field_ptr(self, name):
        q = C.c_void_p()
        self._chk(self.lib.les_field_ptr(self.h, FIELD_ID[name], C.byref(q)), "les_field_ptr")
        return q.value

    def set_mask(self, m):
        p, k = _ptr(m, np.uint8)
        self._chk(self.lib.les_set_mask(self.h, p, k), "les_set_mask")

    def get_mask(self):
        out = np.empty(self.N, dtype=np.uint8)
        self._chk(self.lib.les_get_mask(self.h, _host_ptr(out, np.uint8), 0), "les_get_mask")
        return out

    def set_state(self, st):
        for k, v in st.items():
            self.set_field(k, np.ascontiguousarray(v, dtype=np.float64))

    def get_state(self, names=FIELDS[:9]):
        return {n: self.get_field(n) for n in names}

    def set_option(self, option, value):
        self._chk(self.lib.les_set_option(self.h, option, int(value)), "les_set_option")

    def synchronize(self):
        self._chk(self.lib.les_synchronize(self.h), "les_synchronize")

    def set_stream(self, stream_handle):
        self._chk(self.lib.les_set_stream(self.h, C.c_void_p(stream_handle)), "les_set_stream")

    # ---- checkpoint, VTK output, statistics (io.cpp, statistics.cpp) ----
    def save_state(self, time_simulated, dt, next_step_to_run, directory=None):
        """save_state (io.cpp:14-59) from the device fields."""
        self._chk(self.lib.les_save_state(self.h, _path(directory), float(time_simulated),
                                          float(dt), int(next_step_to_run)), "les_save_state")

    def load_state(self, directory=None, time_simulated=0.0, dt=0.0):
        """load_state (io.cpp:61-120): returns (step, time_simulated, dt);
        step 0 where the reference returns 0 (time and dt then as it leaves
        them)."""
        t, d, s = C.c_double(time_simulated), C.c_double(dt), C.c_int(0)
        self._chk(self.lib.les_load_state(self.h, _path(directory), C.byref(t), C.byref(d),
                                          C.byref(s)), "les_load_state")
        return s.value, t.value, d.value

    def write_field_vtk(self, step, path=None, background=False):
        """write_field_vtk (io.cpp:225-256) from the device fields; with
        background=True the call returns once the fields are snapshotted on
        the device and io_wait() joins the writer."""
        f = self.lib.les_write_field_vtk_async if background else self.lib.les_write_field_vtk
        self._chk(f(self.h, _path(path), int(step)), "les_write_field_vtk")

    def io_wait(self):
        self._chk(self.lib.les_io_wait(self.h), "les_io_wait")

    def stats_init(self, start_step):
        self._chk(self.lib.les_stats_init(self.h, int(start_step)), "les_stats_init")

    def stats_update(self, current_step):
        self._chk(self.lib.les_stats_update(self.h, int(current_step)), "les_stats_update")

    def stats_finalize(self, path=None):
        self._chk(self.lib.les_stats_finalize(self.h, _path(path)), "les_stats_finalize")

    def stats_get(self, name, final=False):
        out = np.empty(self.N)
        self._chk(self.lib.les_stats_get(self.h, STAT_ID[name], 1 if final else 0,
                                         _host_ptr(out, np.float64), 0), "les_stats_get")
        return out

    def stats_samples(self):
        n = C.c_int(0)
        self._chk(self.lib.les_stats_samples(self.h, C.byref(n)), "les_stats_samples")
        return n.value

    # ---- operator ----
    def build_stencil(self, ignore_mask=False):
        self._chk(self.lib.les_build_stencil(self.h, 1 if ignore_mask else 0), "les_build_stencil")

    def set_stencil(self, arrs, which=STENCIL_A):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in arrs]
        ptrs = (C.c_void_p * 7)(*[_host_ptr(a, np.float64) for a in arrs])
        self._chk(self.lib.les_set_stencil(self.h, which, ptrs, 0), "les_set_stencil")

    def get_stencil(self, which=STENCIL_A):
        arrs = [np.empty(self.N) for _ in range(7)]
        ptrs = (C.c_void_p * 7)(*[_host_ptr(a, np.float64) for a in arrs])
        self._chk(self.lib.les_get_stencil(self.h, which, ptrs, 0), "les_get_stencil")
        return arrs

    def matvec(self, x, which=STENCIL_A, out=None):
        if out is None:
            out = np.empty(self.N) if isinstance(x, np.ndarray) else x.new_empty(x.shape)
        px, k = _ptr(x)
        py, _ = _ptr(out)
        self._chk(self.lib.les_matvec(self.h, which, px, py, k), "les_matvec")
        return out

    def dot(self, a, b):
        pa, k = _ptr(a)
        pb, _ = _ptr(b)
        r = C.c_double()
        self._chk(self.lib.les_dot(self.h, pa, pb, k, C.byref(r)), "les_dot")
        return r.value

    # ---- preconditioners ----
    def ic0_factor(self):
        self._chk(self.lib.les_ic0_factor(self.h), "les_ic0_factor")

    def precond_apply(self, pc, r, lmin=0.0, lmax=0.0, degree=5):
        z = np.empty(self.N)
        self._chk(self.lib.les_precond_apply(self.h, pc, _host_ptr(r, np.float64),
                                             _host_ptr(z, np.float64), 0, lmin, lmax, degree),
                  "les_precond_apply")
        return z

    def precond_apply_fp32(self, pc, r, lmin=0.0, lmax=0.0, degree=5,
                           stencil_format=STENCIL_AUTO, inner_product=IP_EUCLID):
        """z = M^-1 r in FP32, as a mixed-precision PCG applies it."""
        r = np.ascontiguousarray(r, dtype=np.float32)
        z = np.empty(self.N, dtype=np.float32)
        self._chk(self.lib.les_precond_apply_fp32(self.h, pc, _host_ptr(r, np.float32),
                                                  _host_ptr(z, np.float32), 0, lmin, lmax, degree,
                                                  stencil_format, inner_product),
                  "les_precond_apply_fp32")
        return z

    def power_lambda_max(self, n_iter=20):
        r = C.c_double()
        self._chk(self.lib.les_power_lambda_max(self.h, n_iter, C.byref(r)), "les_power_lambda_max")
        return r.value

    def lanczos_bounds(self, steps=20, seed=0):
        a, b = C.c_double(), C.c_double()
        self._chk(self.lib.les_lanczos_bounds(self.h, steps, seed, C.byref(a), C.byref(b)),
                  "les_lanczos_bounds")
        return a.value, b.value

    # ---- PCG ----
    @staticmethod
    def pcg_params(precond=PC_LEGACY_IC_A, max_iter=1000, tol=1e-6, tol_rel=False, lmin=0.0,
                   lmax=0.0, degree=5, precision=PREC_FP64, reliable_every=50, check_every=0,
                   stencil_format=STENCIL_AUTO, inner_product=IP_EUCLID):
        return PcgParams(precond, max_iter, tol, 1 if tol_rel else 0, lmin, lmax, degree,
                         precision, reliable_every, check_every, stencil_format, inner_product)

    def pcg(self, **kw):
        prm = self.pcg_params(**kw)
        res = PcgResult()
        self._chk(self.lib.les_pcg(self.h, C.byref(prm), C.byref(res)), "les_pcg")
        return res.as_dict()

    def pcg_trace(self, n):
        t = np.zeros(n)
        self._chk(self.lib.les_pcg_trace(self.h, t.ctypes.data_as(DP), n), "les_pcg_trace")
        return t

    # ---- stationary ----
    def pressure_bcs(self):
        self._chk(self.lib.les_pressure_bcs(self.h), "les_pressure_bcs")

    def jacobi_solve(self, dt, rho):
        s = C.c_int()
        self._chk(self.lib.les_jacobi_solve(self.h, dt, rho, C.byref(s)), "les_jacobi_solve")
        return s.value

    def sor_solve(self, dt, rho, omega=1.7, max_sweeps=200, ordering=SOLVER_SOR_LEX):
        s = C.c_int()
        self._chk(self.lib.les_sor_solve(self.h, ordering, dt, rho, omega, max_sweeps,
                                         C.byref(s)), "les_sor_solve")
        return s.value

    # ---- physics ----
    def sgs(self, Cs=0.18, apply_bcs=True, want_tau=False):
        tau = [np.zeros(self.N) for _ in range(6)] if want_tau else None
        ptrs = (C.c_void_p * 6)(*[_host_ptr(a, np.float64) for a in tau]) if tau else None
        self._chk(self.lib.les_sgs(self.h, Cs, 1 if apply_bcs else 0, ptrs, 0), "les_sgs")
        return tau

    def nut_bcs(self):
        self._chk(self.lib.les_nut_bcs(self.h), "les_nut_bcs")

    def momentum(self, dt, nu, apply_bcs=False):
        self._chk(self.lib.les_momentum(self.h, dt, nu, 1 if apply_bcs else 0), "les_momentum")

    def velocity_bcs(self):
        self._chk(self.lib.les_velocity_bcs(self.h), "les_velocity_bcs")

    def divergence_b(self, dt, rho):
        self._chk(self.lib.les_divergence_b(self.h, dt, rho), "les_divergence_b")

    def correct(self, dt, rho, apply_bcs=False):
        self._chk(self.lib.les_correct(self.h, dt, rho, 1 if apply_bcs else 0), "les_correct")

    def adaptive_dt(self, nu):
        r = C.c_double()
        self._chk(self.lib.les_adaptive_dt(self.h, nu, C.byref(r)), "les_adaptive_dt")
        return r.value

    def step(self, step=0, dt_fixed=0.0, nu=1e-3, rho=1.2, Cs=0.18, solver=SOLVER_PCG,
             omega=1.7, sor_sweeps=200, **pcg_kw):
        prm = StepParams(step, dt_fixed, nu, rho, Cs, solver, self.pcg_params(**pcg_kw), omega,
                         sor_sweeps)
        st = StepStats()
        self._chk(self.lib.les_step(self.h, C.byref(prm), C.byref(st)), "les_step")
        return {"dt": st.dt, "sweeps": st.sweeps, "speed_max": st.speed_max,
                "nu_t_max": st.nu_t_max, **st.pcg.as_dict()}

    # ---- differentiation (no reference counterpart) ----
    @classmethod
    def adj_params(cls, dt=1e-3, nsteps=20, checkpoint_every=5, nu=1e-3, rho=1.2, Cs=0.17,
                   **pcg_kw):
        """Parameters of les_rollout / les_adjoint.  The pressure solves default
        to Jacobi-PCG in the mirror-weighted inner product to 1e-13 relative."""
        kw = dict(precond=PC_JACOBI, max_iter=20000, tol=1e-13, tol_rel=True,
                  inner_product=IP_MIRROR)
        kw.update(pcg_kw)
        return AdjParams(nu, rho, Cs, dt, nsteps, checkpoint_every, cls.pcg_params(**kw))

    def rollout(self, prm):
        """Advances u, v, w by prm.nsteps steps of the fixed-dt step map and
        keeps checkpoints for adjoint()."""
        st = AdjStats()
        self._chk(self.lib.les_rollout(self.h, C.byref(prm), C.byref(st)), "les_rollout")
        return st.as_dict()

    def adjoint(self, prm, ubar, vbar, wbar):
        """dL/du^0 from dL/du^nsteps, in place (numpy or CUDA tensors)."""
        pu, k = _ptr(ubar)
        pv, _ = _ptr(vbar)
        pw, _ = _ptr(wbar)
        st = AdjStats()
        self._chk(self.lib.les_adjoint(self.h, C.byref(prm), pu, pv, pw, k, C.byref(st)),
                  "les_adjoint")
        return st.as_dict()

    # ---- diagnostics ----
    def kernel_timing(self, kernel_class):
        self._chk(self.lib.les_kernel_timing(self.h, kernel_class), "les_kernel_timing")

    def kernel_timing_read(self):
        n, ms, eb = C.c_int(), C.c_double(), C.c_int()
        self._chk(self.lib.les_kernel_timing_read(self.h, C.byref(n), C.byref(ms), C.byref(eb)),
                  "les_kernel_timing_read")
        return n.value, ms.value, eb.value

    def diag_copy(self, reps=1):
        self._chk(self.lib.les_diag_copy(self.h, reps), "les_diag_copy")
