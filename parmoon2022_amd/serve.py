"""Request loop for convection-diffusion solves of mixed sizes on one GPU:
the C++ server of libpmcd.so (pmcd_server_*) seen from Python.

Continuous batching: requests wait in a FIFO queue; a worker admits the head
request as soon as a slot is free and either a resident system for its mesh
exists or its estimated device memory fits the budget. Admitted requests run
concurrently, each on its own HIP stream (one device-resident system per
running request), so their kernels overlap on the device; a finished request
frees its slot at once and the next one is admitted while the others keep
running.

Systems stay resident: a finished request parks its system (mesh hierarchy,
spaces, sparsity, transfer operators, device buffers) in a pool keyed by the
mesh, and a later request on the same mesh only replaces the PDE data,
reassembles and solves. Idle systems are destroyed, oldest first, when a new
request needs their memory.

The reference has no request loop: each problem is a new
ConvectionDiffusion<d> (src/System/ConvectionDiffusion.C:112-452) followed by
assemble(), solve() and output() (3DPrograms/CD3D_ParMooN.C:40-53); a served
request does the assemble and solve steps.
"""
import collections
import ctypes
import time

import numpy as np

from . import SolveStats, _check, lib, problem_desc, solver_desc

MESH_KEYS = ("geo_file", "refinement_n_initial_steps", "structured_n", "ANSATZ_ORDER",
             "space_discretization_type", "INTERNAL_PROBLEM_LINEAR", "multigrid_n_levels",
             "preconditioner", "solver_type")


def mesh_key(db):
    key = tuple(db.get(k) for k in MESH_KEYS)
    steps = int(db.get("refine_by_indicator_steps", 0))
    if steps > 0:
        f = db["refine_indicator"]
        key += (steps, tuple(f) if isinstance(f, (tuple, list)) else f)
    return key


def estimate_bytes(db):
    """Upper estimate of the device bytes of a system for `db` after a solve
    with its solver settings (pmcd_estimate_device_bytes): every per-level
    buffer, the assembly scratch and the solver's vectors, from the coarse
    mesh's counts and the regular-refinement recurrences."""
    out = ctypes.c_longlong()
    pd, sd = problem_desc(db), solver_desc(db)
    _check(lib().pmcd_estimate_device_bytes(ctypes.byref(pd), ctypes.byref(sd), ctypes.byref(out)))
    return out.value


class ServedResult(ctypes.Structure):
    _fields_ = [("id", ctypes.c_longlong), ("status", ctypes.c_int), ("reused", ctypes.c_int),
                ("arrival", ctypes.c_double), ("admitted", ctypes.c_double), ("finished", ctypes.c_double),
                ("assemble_ms", ctypes.c_double), ("n_dof", ctypes.c_longlong),
                ("reserved_bytes", ctypes.c_longlong), ("stats", SolveStats), ("error", ctypes.c_char * 256)]


class _Done:
    """Completion of one request: wait() blocks in the server (the GIL is
    released meanwhile) and, once the request has finished, copies its
    record and solution out of the server."""

    def __init__(self, req):
        self._req = req
        self._set = False

    def is_set(self):
        return self._set

    def wait(self, timeout=None):
        if self._set:
            return True
        done = ctypes.c_int(0)
        srv = self._req._server
        _check(lib().pmcd_server_wait(srv._h, self._req.id, -1.0 if timeout is None else float(timeout),
                                      ctypes.byref(done)))
        if done.value:
            self._req._collect()
            self._set = True
        return self._set


class Request:
    """One solve: `db` holds the problem and solver parameters (reference key
    names). Timestamps are time.perf_counter() values."""

    def __init__(self, server, rid, db, arrival):
        self._server = server
        self.id = rid
        self.db = dict(db)
        self.arrival = arrival
        self.admitted = None
        self.finished = None
        self.stats = None
        self.solution = None
        self.error = None
        self.reused = None
        self.assemble_ms = None
        self.n_dof = None
        self.reserved_bytes = None
        self.done = _Done(self)

    def _collect(self):
        srv = self._server
        r = ServedResult()
        _check(lib().pmcd_server_result(srv._h, self.id, ctypes.byref(r)))
        self.admitted = r.admitted + srv._offset
        self.finished = r.finished + srv._offset
        self.reused = bool(r.reused)
        self.assemble_ms = r.assemble_ms
        self.n_dof = r.n_dof
        self.reserved_bytes = r.reserved_bytes
        if r.status == 0:
            self.stats = r.stats
            if srv.keep_solution:
                u = np.empty(r.n_dof)
                _check(lib().pmcd_server_solution(srv._h, self.id, u.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                  ctypes.c_longlong(r.n_dof)))
                self.solution = u
        else:
            self.error = RuntimeError(r.error.decode())
        _check(lib().pmcd_server_release(srv._h, self.id))

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival


class Server:
    """Continuous-batching request loop (see the module docstring).

    max_slots: requests running at once; budget: device bytes the server may
    hold (default: 90% of the free memory at start); keep_solution: copy each
    solution back to the host (otherwise only the statistics are kept)."""

    def __init__(self, max_slots=8, budget=None, keep_solution=False):
        self.max_slots = int(max_slots)
        self.keep_solution = bool(keep_solution)
        self._h = ctypes.c_void_p()
        L = lib()
        L.pmcd_server_submit.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                         ctypes.POINTER(ctypes.c_longlong)]
        L.pmcd_server_wait.argtypes = [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_double,
                                       ctypes.POINTER(ctypes.c_int)]
        L.pmcd_server_result.argtypes = [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_void_p]
        L.pmcd_server_solution.argtypes = [ctypes.c_void_p, ctypes.c_longlong, ctypes.POINTER(ctypes.c_double),
                                           ctypes.c_longlong]
        L.pmcd_server_release.argtypes = [ctypes.c_void_p, ctypes.c_longlong]
        _check(L.pmcd_server_create(self.max_slots, ctypes.c_longlong(int(budget) if budget is not None else 0),
                                    int(self.keep_solution), ctypes.byref(self._h)))
        now = ctypes.c_double()
        _check(L.pmcd_server_now(self._h, ctypes.byref(now)))
        # perf_counter() = server clock + offset
        self._offset = time.perf_counter() - now.value
        self.budget = self.info()["budget"]

    def submit(self, db, arrival=None):
        if not self._h:
            raise RuntimeError("server is closed")
        pd, sd = problem_desc(db), solver_desc(db)
        rid = ctypes.c_longlong()
        t = time.perf_counter() if arrival is None else arrival
        _check(lib().pmcd_server_submit(self._h, ctypes.byref(pd), ctypes.byref(sd), t - self._offset,
                                        ctypes.byref(rid)))
        return Request(self, rid.value, db, t)

    def warm(self, dbs, per_mesh=1):
        """Create and park up to `per_mesh` systems per distinct mesh of `dbs`
        (never more than the mesh occurs), so that set-up cost stays outside a
        measured stream."""
        counts = collections.Counter(mesh_key(db) for db in dbs)
        first = {}
        for db in dbs:
            first.setdefault(mesh_key(db), db)
        for key, db in first.items():
            pd, sd = problem_desc(db), solver_desc(db)
            _check(lib().pmcd_server_warm(self._h, ctypes.byref(pd), ctypes.byref(sd), min(per_mesh, counts[key])))

    def info(self):
        a = (ctypes.c_longlong * 8)()
        _check(lib().pmcd_server_info(self._h, a))
        keys = ["queued", "running", "reserved_bytes", "idle_bytes", "idle_systems", "budget", "finished",
                "systems_created"]
        return dict(zip(keys, list(a)))

    def close(self):
        """Run every queued request to the end, then free the resident systems."""
        if self._h:
            _check(lib().pmcd_server_destroy(self._h))
            self._h = None

    def __del__(self):
        self.close()


def c5_requests(n_req, rate, seed=7, solver=None):
    """Served stream: Poisson arrivals at `rate` per second, half CD2D P2 on
    triangles (DoF (2n+1)^2), half CD3D Q1 on bricks (DoF (n+1)^3), target DoF
    LogNormal(ln 2e5, 1) clipped to [1.6e4, 4e6] and rounded to the nearest
    mesh of the uniform refinements, coefficients as in the batched workload
    (eps = 10^U(-6,0), b a random unit vector, f = 1, Dirichlet 0, SUPG).
    Returns (arrival offsets in seconds, parameter dictionaries)."""
    rng = np.random.default_rng(seed)
    solver = solver or {}
    t = np.cumsum(rng.exponential(1.0 / rate, n_req))
    out = []
    for _ in range(n_req):
        target = float(np.clip(rng.lognormal(np.log(2e5), 1.0), 1.6e4, 4e6))
        three_d = bool(rng.integers(0, 2))
        if three_d:
            ks = np.arange(4, 8)
            dofs = (2.0 ** ks + 1) ** 3
        else:
            ks = np.arange(5, 10)
            dofs = (2.0 ** (ks + 1) + 1) ** 2
        k = int(ks[np.argmin(np.abs(np.log(dofs) - np.log(target)))])
        eps = float(10.0 ** rng.uniform(-6, 0))
        g = rng.standard_normal(3)
        if not three_d:
            g[2] = 0.0
        b = g / np.linalg.norm(g)
        db = {"geo_file": "Default_UnitCube_Hexa" if three_d else "TwoTriangles",
              "ANSATZ_ORDER": 1 if three_d else 2, "refinement_n_initial_steps": k,
              "example": "const_rhs", "space_discretization_type": "supg", "diffusion_coefficient": eps,
              "convection": tuple(float(v) for v in b), "reaction": 0.0, "rhs_constant": 1.0,
              "solver_type": "iterative", "iterative_solver_type": "fgmres", "preconditioner": "multigrid",
              # hierarchy down to 33^2 (2D P2) or 9^3 (3D Q1) DoF: the dense
              # coarse solve stays small
              "multigrid_n_levels": k - 2 if three_d else k - 3,
              "multigrid_smoother": "sor", "sor_omega": 1.0, "multigrid_n_pre_smooth": 2,
              "multigrid_n_post_smooth": 2, "gmres_restart": 30, "residual_reduction": 1e-10,
              "residual_tolerance": 0.0, "max_n_iterations": 500}
        db.update(solver)
        out.append(db)
    return t, out
