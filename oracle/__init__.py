"""ORACLE -- test infrastructure only.

ctypes binding of the CPU restatement of the reference's convection-diffusion
path (cd_oracle.c). Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it, as the checker and as the CPU baseline; the shipped
compute path (parmoon2022_amd) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle_cd.so")
_lib = None


class OSolver(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int), ("precond", ctypes.c_int),
                ("max_iter", ctypes.c_int), ("min_iter", ctypes.c_int),
                ("res_tol", ctypes.c_double), ("res_red", ctypes.c_double),
                ("div_factor", ctypes.c_double), ("restart", ctypes.c_int),
                ("n_pre", ctypes.c_int), ("n_post", ctypes.c_int),
                ("jacobi_damp", ctypes.c_double), ("mg_damp", ctypes.c_double),
                ("mixed", ctypes.c_int), ("smoother", ctypes.c_int), ("sor_omega", ctypes.c_double),
                ("cycle", ctypes.c_int), ("coarse_smoother", ctypes.c_int),
                ("coarse_residual", ctypes.c_double), ("coarse_max_iter", ctypes.c_int)]


class OStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("converged", ctypes.c_int),
                ("initial_residual", ctypes.c_double), ("final_residual", ctypes.c_double),
                ("seconds", ctypes.c_double)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.c_int] * 8 + [ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                                                      ctypes.c_double, ctypes.c_double]
        for name in ["orc_destroy", "orc_info", "orc_dofmap", "orc_pattern", "orc_cell_coords",
                     "orc_assemble", "orc_matrix", "orc_rhs", "orc_solution", "orc_set_solution",
                     "orc_apply", "orc_solve", "orc_vcycle", "orc_errors"]:
            getattr(L, name).restype = None if name != "orc_solve" else ctypes.c_int
        L.orc_create_adaptive.restype = ctypes.c_void_p
        L.orc_create_adaptive.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                                          ctypes.c_double, ctypes.c_double]
        L.orc_hanging.restype = ctypes.c_int
        L.orc_tcd_init.restype = ctypes.c_int
        L.orc_tcd_init.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.orc_tcd_step.restype = ctypes.c_int
        L.orc_tcd_config.restype = ctypes.c_int
        L.orc_afc_solve.restype = ctypes.c_int
        L.orc_tcd_config.argtypes = L.orc_tcd_init.argtypes
        L.orc_ddot.restype = ctypes.c_double
        L.orc_quad.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a, t=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(t))


def solver(desc):
    """OSolver from a parmoon2022_amd.SolverDesc-compatible object."""
    s = OSolver()
    for f, _ in OSolver._fields_:
        setattr(s, f, getattr(desc, f))
    return s


class Oracle:
    """Reference-order CPU problem: same constructor arguments as ProblemDesc."""

    def __init__(self, desc):
        if desc.geometry == 12:
            raise ValueError("the oracle builds default and structured meshes only, not mesh files")
        b = (ctypes.c_double * 3)(*list(desc.b))
        self.desc = desc
        adapt = int(getattr(desc, "adapt_steps", 0))
        if adapt > 0:
            # the ProblemDesc's indicator callback (a C function pointer) and context
            fn = ctypes.cast(desc.indicator, ctypes.c_void_p)
            p = lib().orc_create_adaptive(desc.geometry, desc.n_refine, adapt, fn, desc.indicator_ctx,
                                          desc.adapt_conforming, desc.order, desc.supg, desc.linear_quad,
                                          desc.example, desc.eps, b, desc.c, desc.f0)
            if not p:
                raise ValueError("adaptive meshes take order 1 or 2 on the 2D default meshes "
                                 "(the conforming closure on triangles)")
            self._p = ctypes.c_void_p(p)
        else:
            self._p = ctypes.c_void_p(lib().orc_create(desc.geometry, desc.structured_n, desc.n_refine,
                                                       desc.order, desc.supg,
                                                       desc.linear_quad, desc.mg_levels, desc.example,
                                                       desc.eps, b, desc.c, desc.f0))

    def hanging(self):
        """Rows (hanging DOF, 3 coupling DOFs, -1 beyond the count) and the
        coupling weights (0 beyond), in the sorted order."""
        n = lib().orc_hanging(self._p, None, None)
        out = np.zeros(4 * max(n, 1), dtype=np.int32)
        w = np.zeros(3 * max(n, 1))
        lib().orc_hanging(self._p, _p(out, ctypes.c_int), _p(w))
        return out[:4 * n].reshape(n, 4), w[:3 * n].reshape(n, 3)

    def __del__(self):
        if getattr(self, "_p", None):
            lib().orc_destroy(self._p)
            self._p = None

    def info(self, level=-1):
        a = (ctypes.c_longlong * 8)()
        lib().orc_info(self._p, level, a)
        keys = ["n_cells", "n_dof", "n_active", "nnz", "n_local", "n_quad", "n_levels", "n_dirichlet"]
        return dict(zip(keys, list(a)))

    def dofmap(self, level=-1):
        i = self.info(level)
        out = np.empty(i["n_cells"] * i["n_local"], dtype=np.int32)
        lib().orc_dofmap(self._p, level, _p(out, ctypes.c_int))
        return out.reshape(i["n_cells"], i["n_local"])

    def pattern(self, level=-1):
        i = self.info(level)
        rp = np.empty(i["n_dof"] + 1, dtype=np.int32)
        col = np.empty(i["nnz"], dtype=np.int32)
        lib().orc_pattern(self._p, level, _p(rp, ctypes.c_int), _p(col, ctypes.c_int))
        return rp, col

    def cell_coords(self, level=-1, nv=None):
        i = self.info(level)
        out = np.empty(i["n_cells"] * nv * 3)
        lib().orc_cell_coords(self._p, level, _p(out))
        return out.reshape(i["n_cells"], nv, 3)

    def assemble(self):
        lib().orc_assemble(self._p)

    def matrix_values(self, level=-1):
        v = np.empty(self.info(level)["nnz"])
        lib().orc_matrix(self._p, level, _p(v))
        return v

    def rhs(self, level=-1):
        v = np.empty(self.info(level)["n_dof"])
        lib().orc_rhs(self._p, level, _p(v))
        return v

    def solution(self):
        v = np.empty(self.info(-1)["n_dof"])
        lib().orc_solution(self._p, _p(v))
        return v

    def set_solution(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        lib().orc_set_solution(self._p, _p(v))

    def apply(self, x, level=-1, y=None, a=1.0):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros_like(x) if y is None else np.array(y, dtype=np.float64)
        lib().orc_apply(self._p, level, _p(x), _p(out), ctypes.c_double(a), 0 if y is None else 1)
        return out

    def solve(self, sdesc):
        s = solver(sdesc)
        st = OStats()
        rc = lib().orc_solve(self._p, ctypes.byref(s), ctypes.byref(st))
        if rc == 2:
            raise RuntimeError("oracle solver diverged")
        return st

    def vcycle(self, sdesc, v):
        s = solver(sdesc)
        v = np.ascontiguousarray(v, dtype=np.float64)
        z = np.zeros_like(v)
        lib().orc_vcycle(self._p, ctypes.byref(s), _p(v), _p(z))
        return z

    def set_matrix(self, v, level=-1):
        """Replace the level's matrix values (same pattern), e.g. by the
        device-assembled ones, so that later steps start from equal data."""
        v = np.ascontiguousarray(v, dtype=np.float64)
        lib().orc_set_matrix(self._p, level, _p(v))

    def sor_sweep(self, b, x, omega, flag, level=-1):
        b = np.ascontiguousarray(b, dtype=np.float64)
        out = np.array(x, dtype=np.float64)
        lib().orc_sor_sweep(self._p, level, _p(b), _p(out), ctypes.c_double(omega), int(flag))
        return out

    def errors(self):
        e = np.zeros(5)
        lib().orc_errors(self._p, _p(e))
        return e

    # time-dependent problems: the example's coefficients at the current time
    def tcd_init(self, scheme, bdf_first, tau, t0):
        """Assemble stiffness, mass and rhs at t0; the solution becomes the
        exact solution interpolated at the nodes."""
        if lib().orc_tcd_init(self._p, int(scheme), int(bdf_first), float(tau), float(t0)):
            raise ValueError("bad time discretization parameters")
        self.tcd_errors = np.zeros(6)
        self.tcd_tau = float(tau)

    def tcd_step(self, sdesc):
        """One time step; returns (OStats, new time)."""
        s = solver(sdesc)
        st = OStats()
        t = ctypes.c_double()
        rc = lib().orc_tcd_step(self._p, ctypes.byref(s), ctypes.byref(st), ctypes.byref(t))
        if rc == 1:
            raise RuntimeError("call tcd_init before tcd_step")
        if rc == 2:
            raise RuntimeError("oracle solver diverged")
        return st, t.value

    def afc_solve(self, sdesc, prm):
        """Algebraic flux correction (Kuzmin or BJK17 limiter, fixed point rhs
        or fixed point matrix scheme, dynamic damping) from the zero initial
        iterate; prm as parmoon2022_amd.afc_params. Returns (iterations, residual, rejected
        steps, initial residual)."""
        s = solver(sdesc)
        p = np.ascontiguousarray(prm, dtype=np.float64)
        out = np.zeros(4)
        rc = lib().orc_afc_solve(self._p, ctypes.byref(s), _p(p), _p(out))
        if rc == 3:
            raise ValueError("algebraic flux correction needs a single level")
        if rc == 4:
            raise ValueError("the BJK17 limiter needs P1 triangles or tetrahedra")
        if rc == 2:
            raise RuntimeError("oracle solver diverged")
        return int(out[0]), out[1], int(out[2]), out[3]

    def tcd_matrices(self, level=-1):
        """(stiffness, mass) values of the level's pattern."""
        nnz = self.info(level)["nnz"]
        a, m = np.empty(nnz), np.empty(nnz)
        lib().orc_tcd_matrices(self._p, level, _p(a), _p(m))
        return a, m

    def tcd_output(self):
        """Errors at the current time, accumulated over time as
        TimeConvectionDiffusion::output does (trapezoidal sums of the squared
        L2 and H1-semi errors, src/System/TimeConvectionDiffusion.C:938-963)."""
        e = self.errors()
        te, tau = self.tcd_errors, self.tcd_tau
        te[0] += (e[0] * e[0] + te[1] * te[1]) * tau * 0.5
        te[1] = e[0]
        te[2] += (e[1] * e[1] + te[3] * te[3]) * tau * 0.5
        te[3] = e[1]
        te[4] = max(te[4], e[4])
        te[5] = max(te[5], e[0])
        return e

    def tcd_get_errors(self):
        te = self.tcd_errors
        return np.array([np.sqrt(te[0]), np.sqrt(te[2]), te[4]])


def basis(shape, order, point):
    """Reference basis values/derivatives at a point: shape 0 tri, 1 quad, 2 tet, 3 hex."""
    out = np.zeros(10 * 27)
    p = np.ascontiguousarray(point, dtype=np.float64)
    lib().orc_basis(shape, order, _p(p), _p(out))
    return out


def quad_rule(shape, degree):
    w = np.zeros(64)
    x = np.zeros(64 * 3)
    n = lib().orc_quad(shape, degree, _p(w), _p(x))
    return w[:n], x[:3 * n].reshape(n, 3)


def sor_csr(rp, col, val, b, x, omega, flag):
    """TMatrix::sor_sweep on any CSR matrix; returns the new x."""
    n = len(rp) - 1
    out = np.array(x, dtype=np.float64)
    lib().orc_sor_csr(n, _p(np.ascontiguousarray(rp, np.int32), ctypes.c_int),
                      _p(np.ascontiguousarray(col, np.int32), ctypes.c_int),
                      _p(np.ascontiguousarray(val, np.float64)), _p(np.ascontiguousarray(b, np.float64)),
                      _p(out), ctypes.c_double(omega), int(flag))
    return out


def spmv_csr(n, n_active, rp, col, val, x, y, a=1.0):
    y = np.array(y, dtype=np.float64)
    lib().orc_spmv_csr(n, n_active, _p(np.ascontiguousarray(rp, np.int32), ctypes.c_int),
                       _p(np.ascontiguousarray(col, np.int32), ctypes.c_int),
                       _p(np.ascontiguousarray(val, np.float64)), _p(np.ascontiguousarray(x, np.float64)),
                       _p(y), ctypes.c_double(a))
    return y
