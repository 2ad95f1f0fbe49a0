"""Request loop for convection-diffusion solves of mixed sizes on one GPU.

Continuous batching: requests wait in a FIFO queue; the dispatcher admits the
head request as soon as a slot is free and its estimated device memory fits
the free-HBM budget. Admitted requests run concurrently, each on its own HIP
stream (one device-resident system per running request), so their kernels
overlap on the device; a finished request frees its slot at once and the
next one is admitted while the others keep running.

Systems stay resident: a finished request returns its system (mesh
hierarchy, spaces, sparsity, transfer operators, device buffers) to a pool
keyed by the mesh, and a later request on the same mesh only replaces the
PDE data (pmcd_system_set_data), reassembles and solves. Idle systems are
destroyed, oldest first, when a new request needs their memory.

The reference has no request loop: each problem is a new
ConvectionDiffusion<d> (src/System/ConvectionDiffusion.C:112-452) followed by
assemble(), solve() and output(); a served request does exactly these steps.
"""
import collections
import threading
import time

import numpy as np

from . import ConvectionDiffusion, mem_info

MESH_KEYS = ("geo_file", "refinement_n_initial_steps", "structured_n", "ANSATZ_ORDER",
             "space_discretization_type", "INTERNAL_PROBLEM_LINEAR", "multigrid_n_levels",
             "preconditioner", "solver_type")


def mesh_key(db):
    return tuple(db.get(k) for k in MESH_KEYS)


def estimate_bytes(db):
    """Upper estimate of a system's device memory (all multigrid levels):
    CSR (values, columns, FP32 copy, slot map), local matrices, FGMRES basis
    and work vectors."""
    geo = db.get("geo_file")
    k = int(db.get("refinement_n_initial_steps", 0))
    p = int(db.get("ANSATZ_ORDER", 1))
    if str(geo).endswith(".mesh"):
        # counts of the file's coarse mesh; refinement multiplies cells by 8
        nvert, ncell, hexa = _mesh_file_counts(geo)
        cells = ncell * 8 ** k
        nl = {(False, 1): 4, (False, 2): 10, (True, 1): 8, (True, 2): 27}[(hexa, p)]
        n = int(nvert * 8 ** k * (1 if p == 1 else (8 if hexa else 7)))
        row = {4: 27, 10: 65, 8: 27, 27: 125}[nl]
        levels = 1.15
    elif str(geo).lower().endswith(".geo"):
        # 2D .GEO: counts line NEL NVT NMT NVE; refinement multiplies cells by 4
        with open(geo) as f:
            f.readline()
            f.readline()
            nel, nvt, _, nve = (int(v) for v in f.readline().split()[:4])
        cells = nel * 4 ** k
        tri = nve == 3
        nl = {(True, 1): 3, (True, 2): 6, (False, 1): 4, (False, 2): 9}[(tri, p)]
        # DoF per cell of a refined mesh (vertices, plus edge and cell nodes
        # for order 2), plus the boundary's share
        per_cell = {(True, 1): 0.5, (True, 2): 2.0, (False, 1): 1.0, (False, 2): 4.0}[(tri, p)]
        n = int(cells * per_cell) + nvt * p * 2 ** (k + 1)
        row = {3: 7, 6: 19, 4: 9, 9: 25}[nl]
        levels = 1.34
    elif geo in ("Default_UnitCube_Hexa", "Default_UnitCube_Tetra", "StructuredHexa", "StructuredTetra"):
        m = int(db.get("structured_n", 1)) * 2 ** k if geo.startswith("Structured") else 2 ** k
        n = (p * m + 1) ** 3
        cells = m ** 3 * (6 if "Tetra" in geo else 1)
        nl = {("Hexa", 1): 8, ("Hexa", 2): 27, ("Tetra", 1): 4, ("Tetra", 2): 10}[
            ("Tetra" if "Tetra" in geo else "Hexa", p)]
        row = {8: 27, 27: 125, 4: 27, 10: 65}[nl]
        levels = 1.15
    else:
        m = 2 ** k
        n = (p * m + 1) ** 2
        cells = m * m * (2 if "Triangles" in geo else 1)
        nl = {("T", 1): 3, ("T", 2): 6, ("Q", 1): 4, ("Q", 2): 9}[("T" if "Triangles" in geo else "Q", p)]
        row = {3: 7, 6: 19, 4: 9, 9: 25}[nl]
        levels = 1.34
    restart = int(db.get("gmres_restart", 20))
    nnz = n * row
    per_level = (nnz * (8 + 4 + 4) + cells * nl * nl * 9 + cells * nl * 16 + n * 8 * (2 * restart + 20))
    return int(per_level * levels) + (64 << 20)


def _mesh_file_counts(path):
    """(vertices, cells, hexahedral) of a .mesh file, from its section counts."""
    nvert = ncell = 0
    hexa = False
    with open(path) as f:
        tok = f.read().split()
    for i, t in enumerate(tok[:-1]):
        if t == "Vertices":
            nvert = int(tok[i + 1])
        elif t in ("Tetrahedra", "Hexahedra"):
            ncell = int(tok[i + 1])
            hexa = t == "Hexahedra"
    return nvert, ncell, hexa


class Request:
    """One solve: `db` holds the problem and solver parameters (reference key
    names). Timestamps are time.perf_counter() values."""

    def __init__(self, rid, db, arrival):
        self.id = rid
        self.db = dict(db)
        self.arrival = arrival
        self.admitted = None
        self.finished = None
        self.stats = None
        self.solution = None
        self.error = None
        self.done = threading.Event()

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival


class Server:
    """Continuous-batching request loop (see the module docstring).

    max_slots: requests running at once; budget: device bytes the server may
    hold (default: 90% of the free memory at start); keep_solution: copy each
    solution back to the host (otherwise only the statistics are kept)."""

    def __init__(self, max_slots=8, budget=None, keep_solution=False):
        free, _ = mem_info()
        self.budget = int(budget if budget is not None else 0.9 * free)
        self.max_slots = int(max_slots)
        self.keep_solution = keep_solution
        self._cv = threading.Condition()
        self._queue = collections.deque()
        self._running = 0
        self._reserved = 0          # bytes held by running requests
        self._idle = collections.OrderedDict()  # (key, serial) -> (system, bytes), oldest first
        self._idle_bytes = 0
        self._serial = 0
        self._closed = False
        self._next_id = 0
        self._threads = []
        self._dispatcher = threading.Thread(target=self._dispatch, daemon=True)
        self._dispatcher.start()

    # ---- public -------------------------------------------------------------
    def submit(self, db, arrival=None):
        with self._cv:
            if self._closed:
                raise RuntimeError("server is closed")
            req = Request(self._next_id, db, time.perf_counter() if arrival is None else arrival)
            self._next_id += 1
            self._queue.append(req)
            self._cv.notify_all()
        return req

    def warm(self, dbs, per_mesh=1):
        """Create and park up to `per_mesh` systems per distinct mesh of `dbs`
        (never more than the mesh occurs), so that set-up cost stays outside a
        measured stream."""
        counts = collections.Counter(mesh_key(db) for db in dbs)
        first = {}
        for db in dbs:
            first.setdefault(mesh_key(db), db)
        for key, db in first.items():
            for _ in range(min(per_mesh, counts[key])):
                self._warm_one(key, db)

    def _warm_one(self, key, db):
        need = estimate_bytes(db)
        with self._cv:
            if not self._make_room(need):
                return
            self._reserved += need
        cd = ConvectionDiffusion(db)
        with self._cv:
            self._reserved -= need
            self._park(key, cd, need)

    def close(self):
        with self._cv:
            self._closed = True
            self._cv.notify_all()
        self._dispatcher.join()
        for t in list(self._threads):
            t.join()
        with self._cv:
            for cd, _ in self._idle.values():
                cd.close()
            self._idle.clear()
            self._idle_bytes = 0

    # ---- internals ------------------------------------------------------------
    def _park(self, key, cd, nbytes):
        self._serial += 1
        self._idle[(key, self._serial)] = (cd, nbytes)
        self._idle_bytes += nbytes

    def _take_idle(self, key):
        for k in self._idle:
            if k[0] == key:
                cd, nbytes = self._idle.pop(k)
                self._idle_bytes -= nbytes
                return cd, nbytes
        return None, 0

    def _make_room(self, need):
        """Destroy idle systems, oldest first, until `need` more bytes fit the
        budget; returns whether they fit."""
        while self._reserved + self._idle_bytes + need > self.budget and self._idle:
            _, (cd, nbytes) = self._idle.popitem(last=False)
            self._idle_bytes -= nbytes
            cd.close()
        return self._reserved + self._idle_bytes + need <= self.budget

    def _dispatch(self):
        while True:
            with self._cv:
                while not self._closed and not self._queue:
                    self._cv.wait()
                if self._closed and not self._queue:
                    return
                req = self._queue[0]
                key = mesh_key(req.db)
                # admission: a free slot, and either a resident system for
                # this mesh or room for a new one (a request larger than the
                # whole budget runs once nothing else is running)
                while True:
                    if self._running < self.max_slots:
                        cd, nbytes = self._take_idle(key)
                        if cd is not None:
                            break
                        nbytes = estimate_bytes(req.db)
                        if self._make_room(nbytes) or self._running == 0:
                            break
                    self._cv.wait()
                self._queue.popleft()
                self._running += 1
                self._reserved += nbytes
                req.admitted = time.perf_counter()
            t = threading.Thread(target=self._run, args=(req, key, cd, nbytes), daemon=True)
            self._threads.append(t)
            t.start()

    def _run(self, req, key, cd, nbytes):
        try:
            if cd is None:
                cd = ConvectionDiffusion(req.db)
            else:
                cd.set_data(req.db)
            cd.assemble()
            req.stats = cd.solve(req.db)
            if self.keep_solution:
                req.solution = cd.solution()
        except Exception as ex:  # reported on the request
            req.error = ex
        req.finished = time.perf_counter()
        with self._cv:
            self._running -= 1
            self._reserved -= nbytes
            if cd is not None and req.error is None:
                self._park(key, cd, nbytes)
            elif cd is not None:
                cd.close()
            self._threads = [t for t in self._threads if t.is_alive() and t is not threading.current_thread()]
            self._cv.notify_all()
        req.done.set()


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
