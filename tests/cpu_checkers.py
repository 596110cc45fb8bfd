"""ctypes bindings of the two CPU checkers used by the tests.

* ``Oracle``  -- oracle/build/liblesoracle.so, the C restatement of the
  reference hot path (oracle/les_oracle.c).
* ``Reference`` -- oracle/_ref/libref.so, the unmodified reference solver
  compiled from its own sources with oracle/ref_harness.cpp.

Both expose the same Python methods so a test can run one against the other.
Only tests, the smoke check and the benchmark's CPU-baseline leg use this.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "build", "liblesoracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libref.so")

DP = C.POINTER(C.c_double)
U8P = C.POINTER(C.c_uint8)
IP = C.POINTER(C.c_int)

PC_NONE, PC_JACOBI, PC_IC0, PC_LEGACY_IC_A, PC_CHEBYSHEV, PC_MG = 0, 1, 2, 3, 4, 5


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    if a.dtype == np.uint8:
        return a.ctypes.data_as(U8P)
    assert a.dtype == np.float64, a.dtype
    return a.ctypes.data_as(DP)


def _p7(arrs):
    return (DP * 7)(*[_p(a) for a in arrs])


def _p6(arrs):
    return (DP * 6)(*[_p(a) for a in arrs])


def _p10(arrs):
    return (DP * 10)(*[_p(a) for a in arrs])


class OrcMesh(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double)]


class OrcMgParams(C.Structure):
    _fields_ = [("sweeps", C.c_int), ("coarse_sweeps", C.c_int), ("coarse_max", C.c_int),
                ("omega", C.c_double), ("scale", C.c_double)]


class OrcFields(C.Structure):
    _fields_ = [(n, DP) for n in ("u", "v", "w", "u_new", "v_new", "w_new", "p", "p_rhs", "nu_t")] + [
        ("solid", U8P)]


class OrcStepParams(C.Structure):
    _fields_ = [("nu", C.c_double), ("rho", C.c_double), ("Cs", C.c_double), ("step", C.c_int),
                ("dt_fixed", C.c_double), ("precond", C.c_int), ("degree", C.c_int),
                ("max_iter", C.c_int), ("lmin", C.c_double), ("lmax", C.c_double),
                ("tol", C.c_double), ("tol_rel", C.c_int), ("L", C.POINTER(DP)),
                ("A", C.POINTER(DP)), ("W", DP)]


class OrcStepStats(C.Structure):
    _fields_ = [("dt", C.c_double), ("iters", C.c_int), ("bnorm", C.c_double), ("res", C.c_double)]


def _dims(m):
    return m["nx"], m["ny"], m["nz"]


def _ncell(m):
    return m["nx"] * m["ny"] * m["nz"]


STATE_KEYS = ("u", "v", "w", "u_new", "v_new", "w_new", "p", "p_rhs", "nu_t")
IO_KEYS = ("u", "v", "w", "p", "nu_t")   # the fields a checkpoint holds (io.cpp:42-46)


class Oracle:
    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"oracle library not built: {path} (run make -C oracle oracle)")
        L = self.lib = C.CDLL(path)
        L.orc_dot.restype = C.c_double
        L.orc_norm.restype = C.c_double
        L.orc_estimate_lambda_max.restype = C.c_double
        L.orc_adaptive_dt.restype = C.c_double
        L.orc_cs_delta_sq.restype = C.c_double
        for f in ("orc_dot", "orc_norm", "orc_axpy", "orc_saxpy", "orc_copy", "orc_jacobi_precond"):
            pass
        L.orc_dot.argtypes = [DP, DP, C.c_long]
        L.orc_norm.argtypes = [DP, C.c_long]
        L.orc_jacobi_precond.argtypes = [DP, DP, DP, C.c_long]
        L.orc_cheb_apply.argtypes = [DP, DP, C.POINTER(DP), C.c_int, C.c_int, C.c_int,
                                     C.c_double, C.c_double, C.c_int]
        L.orc_pcg.argtypes = [C.POINTER(DP), DP, DP, C.c_int, C.c_int, C.c_int, C.c_int,
                              C.POINTER(DP), C.c_double, C.c_double, C.c_int, C.c_int,
                              C.c_double, DP]
        L.orc_calculate_divergence.argtypes = [DP, DP, DP, DP, U8P, C.POINTER(OrcMesh),
                                               C.c_double, C.c_double]
        L.orc_solve_pressure_jacobi.argtypes = [DP, DP, DP, DP, U8P, C.POINTER(OrcMesh),
                                                C.c_double, C.c_double]
        L.orc_solve_pressure_gs.argtypes = [DP, DP, DP, DP, U8P, C.POINTER(OrcMesh), C.c_double,
                                            C.c_double, C.c_int, C.c_double]
        L.orc_solve_pressure_rbsor.argtypes = L.orc_solve_pressure_gs.argtypes
        L.orc_adaptive_dt.argtypes = [DP, DP, DP, DP, C.POINTER(OrcMesh), C.c_double]
        L.orc_cs_delta_sq.argtypes = [C.POINTER(OrcMesh), C.c_double]
        L.orc_sgs.argtypes = [DP, C.POINTER(DP), DP, DP, DP, C.POINTER(OrcMesh), C.c_double]
        L.orc_momentum.argtypes = [DP, DP, DP, DP, DP, DP, DP, C.POINTER(OrcMesh), C.c_double,
                                   C.c_double]
        L.orc_correct.argtypes = [DP, DP, DP, DP, C.POINTER(OrcMesh), C.c_double, C.c_double]
        L.orc_estimate_lambda_max.argtypes = [C.POINTER(DP), C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_mg_default_params.restype = OrcMgParams
        L.orc_mg_levels.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)]
        L.orc_mg_coarsen.argtypes = [C.POINTER(DP), C.POINTER(DP), DP, C.c_int, C.c_int,
                                     C.c_int]
        L.orc_mg_apply.argtypes = [DP, DP, C.POINTER(DP), C.c_int, C.c_int, C.c_int,
                                   C.POINTER(OrcMgParams), DP]
        L.orc_pcg_mg.argtypes = [C.POINTER(DP), DP, DP, C.c_int, C.c_int, C.c_int,
                                 C.POINTER(OrcMgParams), C.c_int, C.c_double, DP, DP]
        L.orc_step.argtypes = [C.POINTER(OrcFields), C.POINTER(OrcMesh), C.POINTER(OrcStepParams),
                               C.POINTER(OrcStepStats)]

    name = "oracle"

    def set_threads(self, n):
        self.lib.orc_set_threads(int(n))

    # ---- time statistics (statistics.cpp:98-176) ----
    def stats_update(self, sums10, u, v, w, p):
        self.lib.orc_stats_update(_p10(sums10), _p(u), _p(v), _p(w), _p(p), C.c_long(len(u)))

    def stats_finalize(self, sums10, num_samples):
        out = [np.empty_like(sums10[0]) for _ in range(10)]
        self.lib.orc_stats_finalize(_p10(out), _p10(sums10), int(num_samples),
                                    C.c_long(len(sums10[0])))
        return out

    @staticmethod
    def _m(m):
        return C.byref(OrcMesh(m["nx"], m["ny"], m["nz"], m["dx"], m["dy"], m["dz"]))

    def cs_delta_sq(self, m, Cs):
        return self.lib.orc_cs_delta_sq(self._m(m), Cs)

    def dot(self, a, b):
        return self.lib.orc_dot(_p(a), _p(b), a.size)

    def build_stencil(self, solid, m):
        A = [np.zeros(_ncell(m)) for _ in range(7)]
        self.lib.orc_build_stencil(_p7(A), _p(solid), self._m(m))
        return A

    def matvec(self, p, A, m):
        q = np.zeros_like(p)
        self.lib.orc_matvec(_p(q), _p(p), _p7(A), *_dims(m))
        return q

    def divergence(self, un, vn, wn, solid, m, dt, rho):
        r = np.zeros(_ncell(m))
        self.lib.orc_calculate_divergence(_p(r), _p(un), _p(vn), _p(wn), _p(solid), self._m(m),
                                          dt, rho)
        return r

    def build_b(self, div, solid, m):
        b = np.zeros(_ncell(m))
        self.lib.orc_build_b_vector(_p(b), _p(div), _p(solid), self._m(m))
        return b

    def jacobi_precond(self, r, Ac):
        z = np.zeros_like(r)
        self.lib.orc_jacobi_precond(_p(z), _p(r), _p(Ac), r.size)
        return z

    def ic_build(self, A, m):
        L = [np.zeros(_ncell(m)) for _ in range(7)]
        self.lib.orc_ic_build(_p7(L), _p7(A), *_dims(m))
        return L

    def ic_apply(self, r, L, m):
        z = np.zeros_like(r)
        self.lib.orc_ic_apply(_p(z), _p(r), _p7(L), *_dims(m))
        return z

    def cheb_apply(self, r, A, m, lmin, lmax, degree):
        z = np.zeros_like(r)
        self.lib.orc_cheb_apply(_p(z), _p(r), _p7(A), *_dims(m), lmin, lmax, degree)
        return z

    def lambda_max(self, A, m, n_iter=20):
        return self.lib.orc_estimate_lambda_max(_p7(A), *_dims(m), n_iter)

    def pcg(self, A, b, x0, m, precond, L=None, lmin=0.0, lmax=0.0, degree=5, max_iter=1000,
            tol=1e-6):
        x = x0.copy()
        tr = np.zeros(max_iter + 1)
        Lp = _p7(L) if L is not None else _p7(A)
        it = self.lib.orc_pcg(_p7(A), _p(b), _p(x), *_dims(m), precond, Lp, lmin, lmax, degree,
                              max_iter, tol, _p(tr))
        return x, it, tr[: it + 1]

    # ---- multigrid preconditioner (no reference counterpart) ----
    def mg_params(self, **kw):
        p = self.lib.orc_mg_default_params()
        for k, v in kw.items():
            setattr(p, k, v)
        return p

    def mg_levels(self, m, coarse_max=4):
        dims = (C.c_int * 48)()
        nl = self.lib.orc_mg_levels(*_dims(m), coarse_max, dims)
        return [tuple(dims[3 * l:3 * l + 3]) for l in range(nl)]

    def mg_coarsen(self, A, dims, W=None):
        nx, ny, nz = dims
        nc = ((nx + 1) // 2) * ((ny + 1) // 2) * ((nz + 1) // 2)
        Ac = [np.zeros(nc) for _ in range(7)]
        self.lib.orc_mg_coarsen(_p7(Ac), _p7(A), _p(W), nx, ny, nz)
        return Ac

    def mg_apply(self, r, A, m, W=None, **kw):
        z = np.zeros_like(r)
        prm = self.mg_params(**kw)
        self.lib.orc_mg_apply(_p(z), _p(r), _p7(A), *_dims(m), C.byref(prm), _p(W))
        return z

    def pcg_mg(self, A, b, x0, m, max_iter=1000, tol=1e-6, W=None, **kw):
        x = x0.copy()
        tr = np.zeros(max_iter + 1)
        prm = self.mg_params(**kw)
        it = self.lib.orc_pcg_mg(_p7(A), _p(b), _p(x), *_dims(m), C.byref(prm), max_iter, tol,
                                 _p(tr), _p(W))
        return x, it, tr[: it + 1]

    def mirror_weights(self, A, m):
        """W = 1/2 per direction in which a row of A is mirrored."""
        W = np.zeros(_ncell(m))
        self.lib.orc_mirror_weights.argtypes = [DP, C.POINTER(DP), C.POINTER(OrcMesh)]
        self.lib.orc_mirror_weights(_p(W), _p7(A), self._m(m))
        return W

    def pcg_w(self, A, b, x0, m, precond, W, lmin=0.0, lmax=0.0, degree=5, max_iter=1000,
              tol=1e-6):
        """orc_pcg with r.z and p.q in the W-weighted inner product."""
        self.lib.orc_pcg_w.argtypes = [C.POINTER(DP), DP, DP, C.c_int, C.c_int, C.c_int, C.c_int,
                                       C.POINTER(DP), C.c_double, C.c_double, C.c_int, C.c_int,
                                       C.c_double, DP, DP]
        x = x0.copy()
        tr = np.zeros(max_iter + 1)
        it = self.lib.orc_pcg_w(_p7(A), _p(b), _p(x), *_dims(m), precond, _p7(A), lmin, lmax,
                                degree, max_iter, tol, _p(tr), _p(W))
        return x, it, tr[: it + 1]

    def pressure_bcs(self, p, solid, m):
        p = p.copy()
        self.lib.orc_apply_pressure_bcs(_p(p), _p(solid), self._m(m))
        return p

    def jacobi_solve(self, p, u, v, w, solid, m, dt, rho):
        p = p.copy()
        s = self.lib.orc_solve_pressure_jacobi(_p(p), _p(u), _p(v), _p(w), _p(solid), self._m(m),
                                               dt, rho)
        return p, s

    def gs_solve(self, p, u, v, w, solid, m, dt, rho, max_iter=200, omega=1.7):
        p = p.copy()
        s = self.lib.orc_solve_pressure_gs(_p(p), _p(u), _p(v), _p(w), _p(solid), self._m(m), dt,
                                           rho, max_iter, omega)
        return p, s

    def rbsor_solve(self, p, u, v, w, solid, m, dt, rho, max_iter=200, omega=1.7):
        p = p.copy()
        s = self.lib.orc_solve_pressure_rbsor(_p(p), _p(u), _p(v), _p(w), _p(solid), self._m(m),
                                              dt, rho, max_iter, omega)
        return p, s

    def adaptive_dt(self, u, v, w, nu_t, m, nu):
        return self.lib.orc_adaptive_dt(_p(u), _p(v), _p(w), _p(nu_t), self._m(m), nu)

    def sgs(self, u, v, w, m, Cs=0.18, nu_t=None, want_tau=False):
        nt = np.zeros(_ncell(m)) if nu_t is None else nu_t.copy()
        tau = [np.zeros(_ncell(m)) for _ in range(6)] if want_tau else None
        self.lib.orc_sgs(_p(nt), _p6(tau) if tau else None, _p(u), _p(v), _p(w), self._m(m), Cs)
        return (nt, tau) if want_tau else nt

    def nut_bcs(self, nu_t, solid, m):
        nt = nu_t.copy()
        self.lib.orc_nut_bcs(_p(nt), _p(solid), self._m(m))
        return nt

    def momentum(self, u, v, w, un, vn, wn, nu_t, m, dt, nu):
        un, vn, wn = un.copy(), vn.copy(), wn.copy()
        self.lib.orc_momentum(_p(un), _p(vn), _p(wn), _p(u), _p(v), _p(w), _p(nu_t), self._m(m),
                              dt, nu)
        # the reference swaps: Field.u <- new buffer, Field.u_new <- old u
        return un, vn, wn, u.copy(), v.copy(), w.copy()

    def velocity_bcs(self, u, v, w, solid, m):
        u, v, w = u.copy(), v.copy(), w.copy()
        self.lib.orc_velocity_bcs(_p(u), _p(v), _p(w), _p(solid), self._m(m))
        return u, v, w

    def correct(self, u, v, w, p, m, dt, rho):
        u, v, w = u.copy(), v.copy(), w.copy()
        self.lib.orc_correct(_p(u), _p(v), _p(w), _p(p), self._m(m), dt, rho)
        return u, v, w

    def run_steps(self, state, solid, m, A, nsteps, step0=0, precond=PC_LEGACY_IC_A, nu=1e-3,
                  rho=1.2, Cs=0.18, dt_fixed=0.0, lmin=0.0, lmax=0.0, degree=5, max_iter=1000,
                  tol=1e-6, tol_rel=False, L=None, W=None):
        st = {k: state[k].copy() for k in STATE_KEYS}
        solid = np.ascontiguousarray(solid)
        f = OrcFields(*[_p(st[k]) for k in STATE_KEYS], _p(solid))
        A7 = _p7(A)
        L7 = _p7(L) if L is not None else A7
        iters, dts = [], []
        for s in range(nsteps):
            prm = OrcStepParams(nu, rho, Cs, step0 + s, dt_fixed, precond, degree, max_iter, lmin,
                                lmax, tol, 1 if tol_rel else 0, L7, A7, _p(W))
            stats = OrcStepStats()
            self.lib.orc_step(C.byref(f), self._m(m), C.byref(prm), C.byref(stats))
            iters.append(stats.iters)
            dts.append(stats.dt)
        # orc_step swaps pointers; read back through the struct
        n = _ncell(m)
        out = {}
        for k in STATE_KEYS:
            ptr = getattr(f, k)
            out[k] = np.ctypeslib.as_array(ptr, shape=(n,)).copy()
        return out, iters, dts


class OrcAdjParams(C.Structure):
    _fields_ = [("nu", C.c_double), ("rho", C.c_double), ("Cs", C.c_double), ("dt", C.c_double),
                ("nsteps", C.c_int), ("precond", C.c_int), ("lmin", C.c_double),
                ("lmax", C.c_double), ("degree", C.c_int), ("max_iter", C.c_int),
                ("tol_rel", C.c_double), ("weighted", C.c_int)]


def _oracle_adjoint_bindings(L):
    L.orc_rollout.argtypes = [DP, DP, DP, DP, U8P, C.POINTER(OrcMesh), C.POINTER(DP),
                              C.POINTER(OrcAdjParams), DP]
    L.orc_adjoint.argtypes = [DP, DP, DP, DP, U8P, C.POINTER(OrcMesh), C.POINTER(DP),
                              C.POINTER(OrcAdjParams)]


def oracle_rollout(orc, u, v, w, solid, m, A, nsteps, dt=1e-3, nu=1e-3, rho=1.2, Cs=0.17,
                   precond=PC_JACOBI, lmin=0.0, lmax=0.0, degree=5, max_iter=20000,
                   tol_rel=1e-13, p=None, record=False, weighted=True):
    """nsteps LES steps with a fixed dt (orc_rollout); returns the final
    (u, v, w, p) and, with record, the trajectory u^n, n < nsteps."""
    _oracle_adjoint_bindings(orc.lib)
    N = u.size
    u, v, w = u.copy(), v.copy(), w.copy()
    p = np.zeros(N) if p is None else p.copy()
    prm = OrcAdjParams(nu, rho, Cs, dt, nsteps, precond, lmin, lmax, degree, max_iter, tol_rel,
                       1 if weighted else 0)
    traj = np.zeros(3 * N * nsteps) if record else None
    orc.lib.orc_rollout(_p(u), _p(v), _p(w), _p(p), _p(solid), orc._m(m), _p7(A), C.byref(prm),
                        _p(traj) if record else None)
    return u, v, w, p, traj


def oracle_adjoint(orc, bu, bv, bw, traj, solid, m, A, nsteps, dt=1e-3, nu=1e-3, rho=1.2,
                   Cs=0.17, precond=PC_JACOBI, lmin=0.0, lmax=0.0, degree=5, max_iter=20000,
                   tol_rel=1e-13, weighted=True):
    _oracle_adjoint_bindings(orc.lib)
    bu, bv, bw = bu.copy(), bv.copy(), bw.copy()
    prm = OrcAdjParams(nu, rho, Cs, dt, nsteps, precond, lmin, lmax, degree, max_iter, tol_rel,
                       1 if weighted else 0)
    orc.lib.orc_adjoint(_p(bu), _p(bv), _p(bw), _p(traj), _p(solid), orc._m(m), _p7(A),
                        C.byref(prm))
    return bu, bv, bw


class Reference:
    """The unmodified reference (src/*.cpp) through oracle/ref_harness.cpp."""

    name = "reference"

    def __init__(self, path=REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = self.lib = C.CDLL(path)
        L.ref_lambda_max.restype = C.c_double
        L.ref_adaptive_dt.restype = C.c_double
        L.ref_jacobi_precond.argtypes = [DP, DP, DP, C.c_long]
        L.ref_cheb_apply.argtypes = [DP, DP, C.POINTER(DP), C.c_int, C.c_int, C.c_int, C.c_double,
                                     C.c_double, C.c_int]
        L.ref_lambda_max.argtypes = [C.POINTER(DP), C.c_int, C.c_int, C.c_int, C.c_int]
        L.ref_pcg_legacy.argtypes = [C.POINTER(DP), DP, DP, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.c_double]
        L.ref_pcg_cheb.argtypes = [C.POINTER(DP), DP, DP, C.c_int, C.c_int, C.c_int, C.c_double,
                                   C.c_double, C.c_int, C.c_int, C.c_double]
        L.ref_pcg_harness.argtypes = [C.POINTER(DP), DP, DP, C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.POINTER(DP), C.c_double, C.c_double, C.c_int, C.c_int,
                                      C.c_double, DP]
        md = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double]
        L.ref_build_stencil.argtypes = [C.POINTER(DP), U8P] + md
        L.ref_divergence.argtypes = [DP, DP, DP, DP, U8P] + md + [C.c_double, C.c_double]
        L.ref_jacobi_solve.argtypes = [DP, DP, DP, DP, U8P] + md + [C.c_double, C.c_double]
        L.ref_gs_solve.argtypes = [DP, DP, DP, DP, U8P] + md + [C.c_double, C.c_double, C.c_int]
        L.ref_adaptive_dt.argtypes = [DP, DP, DP, DP] + md + [C.c_double]
        L.ref_sgs.argtypes = [DP, C.POINTER(DP), DP, DP, DP] + md
        L.ref_momentum.argtypes = [DP] * 7 + md + [C.c_double, C.c_double]
        L.ref_correct.argtypes = [DP, DP, DP, DP] + md + [C.c_double, C.c_double]
        io = [DP] * 5 + [U8P, C.c_int, C.c_int, C.c_int]
        L.ref_save_state.argtypes = [C.c_char_p] + io + [C.c_double, C.c_double, C.c_int]
        L.ref_load_state.argtypes = [C.c_char_p] + io + [DP, DP]
        L.ref_write_field_vtk.argtypes = [C.c_char_p] + io + [C.c_double] * 3 + [C.c_int]
        L.ref_statistics.argtypes = [C.c_char_p] + md + [C.c_int, C.c_int, IP] + [
            C.POINTER(DP)] * 4 + [C.POINTER(DP)]
        L.ref_run_steps.argtypes = [DP] * 9 + [U8P, C.POINTER(DP)] + md + [
            C.c_double, C.c_double, C.c_int, C.c_int, C.c_double, C.c_int, C.c_double,
            C.c_double, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, IP, DP, DP]

    def set_threads(self, n):
        self.lib.ref_set_threads(int(n))

    # ---- checkpoint, VTK and statistics writers (io.cpp, statistics.cpp),
    # run with `directory` as the working directory they write under ----
    def save_state(self, directory, f, m, time_simulated, dt, next_step):
        os.makedirs(os.path.join(directory, "checkpoint"), exist_ok=True)
        rc = self.lib.ref_save_state(os.fsencode(directory), *[_p(f[k]) for k in IO_KEYS],
                                     _p(f["solid"]), m["nx"], m["ny"], m["nz"],
                                     C.c_double(time_simulated), C.c_double(dt), int(next_step))
        assert rc == 0, directory

    def load_state(self, directory, f, m, time_simulated=0.0, dt=0.0):
        """Returns (step, time, dt, fields) as load_state leaves them, the
        fields starting from copies of f."""
        out = {k: np.array(f[k], copy=True) for k in IO_KEYS + ("solid",)}
        t, d = C.c_double(time_simulated), C.c_double(dt)
        step = self.lib.ref_load_state(os.fsencode(directory), *[_p(out[k]) for k in IO_KEYS],
                                       _p(out["solid"]), m["nx"], m["ny"], m["nz"], C.byref(t),
                                       C.byref(d))
        return step, t.value, d.value, out

    def write_field_vtk(self, directory, f, m, step):
        os.makedirs(os.path.join(directory, "output"), exist_ok=True)
        rc = self.lib.ref_write_field_vtk(os.fsencode(directory), *[_p(f[k]) for k in IO_KEYS],
                                          _p(f["solid"]), m["nx"], m["ny"], m["nz"], m["dx"],
                                          m["dy"], m["dz"], int(step))
        assert rc == 0, directory
        return os.path.join(directory, "output", "output_%d.vtk" % step)

    def statistics(self, directory, m, start_step, steps, states):
        """init_statistics + update_statistics per (step, state) +
        finalize_and_save into directory/output (when directory is given).
        Returns (num_samples, sums10)."""
        n = _ncell(m)
        sums = [np.empty(n) for _ in range(10)]
        k = len(steps)
        seq = {q: (DP * k)(*[_p(s[q]) for s in states]) for q in ("u", "v", "w", "p")}
        if directory is not None:
            os.makedirs(os.path.join(directory, "output"), exist_ok=True)
        ns = self.lib.ref_statistics(None if directory is None else os.fsencode(directory),
                                     m["nx"], m["ny"], m["nz"], m["dx"], m["dy"], m["dz"],
                                     int(start_step), k, (C.c_int * k)(*steps), seq["u"],
                                     seq["v"], seq["w"], seq["p"], _p10(sums))
        assert ns >= 0
        return ns, sums

    @staticmethod
    def _md(m):
        return (m["nx"], m["ny"], m["nz"], m["dx"], m["dy"], m["dz"])

    def build_stencil(self, solid, m):
        A = [np.zeros(_ncell(m)) for _ in range(7)]
        self.lib.ref_build_stencil(_p7(A), _p(solid), *self._md(m))
        return A

    def matvec(self, p, A, m):
        q = np.zeros_like(p)
        self.lib.ref_matvec(_p(q), _p(p), _p7(A), *_dims(m))
        return q

    def divergence(self, un, vn, wn, solid, m, dt, rho):
        r = np.zeros(_ncell(m))
        self.lib.ref_divergence(_p(r), _p(un), _p(vn), _p(wn), _p(solid), *self._md(m), dt, rho)
        return r

    def build_b(self, div, solid, m):
        b = np.zeros(_ncell(m))
        self.lib.ref_build_b(_p(b), _p(div), _p(solid), *_dims(m))
        return b

    def jacobi_precond(self, r, Ac):
        z = np.zeros_like(r)
        self.lib.ref_jacobi_precond(_p(z), _p(r), _p(Ac), r.size)
        return z

    def ic_build(self, A, m):
        L = [np.zeros(_ncell(m)) for _ in range(7)]
        self.lib.ref_ic_build(_p7(L), _p7(A), *_dims(m))
        return L

    def ic_apply(self, r, L, m):
        z = np.zeros_like(r)
        self.lib.ref_ic_apply(_p(z), _p(r), _p7(L), *_dims(m))
        return z

    def cheb_apply(self, r, A, m, lmin, lmax, degree):
        z = np.zeros_like(r)
        self.lib.ref_cheb_apply(_p(z), _p(r), _p7(A), *_dims(m), lmin, lmax, degree)
        return z

    def lambda_max(self, A, m, n_iter=20):
        return self.lib.ref_lambda_max(_p7(A), *_dims(m), n_iter)

    def pcg(self, A, b, x0, m, precond, L=None, lmin=0.0, lmax=0.0, degree=5, max_iter=1000,
            tol=1e-6):
        x = x0.copy()
        tr = np.zeros(max_iter + 1)
        Lp = _p7(L) if L is not None else _p7(A)
        it = self.lib.ref_pcg_harness(_p7(A), _p(b), _p(x), *_dims(m), precond, Lp, lmin, lmax,
                                      degree, max_iter, tol, _p(tr))
        return x, it, tr[: it + 1]

    def pcg_legacy(self, A, b, x0, m, max_iter=1000, tol=1e-6):
        x = x0.copy()
        it = self.lib.ref_pcg_legacy(_p7(A), _p(b), _p(x), *_dims(m), max_iter, tol)
        return x, it

    def pcg_cheb(self, A, b, x0, m, lmin, lmax, degree=5, max_iter=1000, tol=1e-6):
        x = x0.copy()
        it = self.lib.ref_pcg_cheb(_p7(A), _p(b), _p(x), *_dims(m), lmin, lmax, degree, max_iter,
                                   tol)
        return x, it

    def pressure_bcs(self, p, solid, m):
        p = p.copy()
        self.lib.ref_pressure_bcs(_p(p), _p(solid), *_dims(m))
        return p

    def jacobi_solve(self, p, u, v, w, solid, m, dt, rho):
        p = p.copy()
        self.lib.ref_jacobi_solve(_p(p), _p(u), _p(v), _p(w), _p(solid), *self._md(m), dt, rho)
        return p, None

    def gs_solve(self, p, u, v, w, solid, m, dt, rho, max_iter=200, omega=1.7):
        assert omega == 1.7
        p = p.copy()
        self.lib.ref_gs_solve(_p(p), _p(u), _p(v), _p(w), _p(solid), *self._md(m), dt, rho,
                              max_iter)
        return p, None

    def adaptive_dt(self, u, v, w, nu_t, m, nu):
        return self.lib.ref_adaptive_dt(_p(u), _p(v), _p(w), _p(nu_t), *self._md(m), nu)

    def sgs(self, u, v, w, m, Cs=0.18, nu_t=None, want_tau=False):
        assert Cs == 0.18, "the reference fixes Cs = 0.18 (main.cpp:425)"
        nt = np.zeros(_ncell(m)) if nu_t is None else nu_t.copy()
        tau = [np.zeros(_ncell(m)) for _ in range(6)] if want_tau else None
        self.lib.ref_sgs(_p(nt), _p6(tau) if tau else None, _p(u), _p(v), _p(w), *self._md(m))
        return (nt, tau) if want_tau else nt

    def nut_bcs(self, nu_t, solid, m):
        nt = nu_t.copy()
        self.lib.ref_nut_bcs(_p(nt), _p(solid), *_dims(m))
        return nt

    def momentum(self, u, v, w, un, vn, wn, nu_t, m, dt, nu):
        a = [x.copy() for x in (u, v, w, un, vn, wn)]
        self.lib.ref_momentum(*[_p(x) for x in a], _p(nu_t), *self._md(m), dt, nu)
        return tuple(a)

    def velocity_bcs(self, u, v, w, solid, m):
        u, v, w = u.copy(), v.copy(), w.copy()
        self.lib.ref_velocity_bcs(_p(u), _p(v), _p(w), _p(solid), *_dims(m))
        return u, v, w

    def correct(self, u, v, w, p, m, dt, rho):
        u, v, w = u.copy(), v.copy(), w.copy()
        self.lib.ref_correct(_p(u), _p(v), _p(w), _p(p), *self._md(m), dt, rho)
        return u, v, w

    def run_steps(self, state, solid, m, A, nsteps, step0=0, solver=0, nu=1e-3, rho=1.2,
                  dt_fixed=0.0, lmin=0.0, lmax=0.0, degree=5, max_iter=1000, tol=1e-6,
                  tol_rel=False, gs_iters=200, step_seconds=None):
        """The driver's step loop (main.cpp:115-144) from `state`; A None:
        the reference's build_stencil from `solid`.  With a list as
        step_seconds, each step's wall time (std::chrono around the step
        inside the harness) is appended to it."""
        st = {k: state[k].copy() for k in STATE_KEYS}
        iters = np.zeros(nsteps, dtype=np.int32)
        dts = np.zeros(nsteps)
        secs = np.zeros(nsteps)
        self.lib.ref_run_steps(*[_p(st[k]) for k in STATE_KEYS], _p(np.ascontiguousarray(solid)),
                               _p7(A) if A is not None else None, *self._md(m), nu, rho, step0,
                               nsteps, dt_fixed, solver, lmin, lmax, degree, max_iter, tol,
                               1 if tol_rel else 0, gs_iters, iters.ctypes.data_as(IP), _p(dts),
                               _p(secs))
        if step_seconds is not None:
            step_seconds.extend(float(x) for x in secs)
        return st, list(iters), list(dts)


def initial_state(u, v, w, n, solid=None, m=None, checker=None):
    """A Field as the reference driver starts it: zeros, then u,v,w set and
    apply_velocity_BCs (main.cpp:102) applied when a checker is given."""
    st = {k: np.zeros(n) for k in STATE_KEYS}
    st["u"], st["v"], st["w"] = u.copy(), v.copy(), w.copy()
    if checker is not None:
        st["u"], st["v"], st["w"] = checker.velocity_bcs(st["u"], st["v"], st["w"], solid, m)
    return st
