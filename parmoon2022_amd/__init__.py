"""MI355X convection-diffusion assemble-and-solve path (ctypes binding of libpmcd.so).

The classes mirror the reference's interface for this path:
``ConvectionDiffusion`` (src/System/ConvectionDiffusion.C: constructor,
assemble(), solve(), output(), get_L2_error() ...), driven by a parameter
dictionary with the reference's parameter names (``ParameterDatabase`` keys of
Solver<>::default_solver_database, Multigrid::default_multigrid_database,
LocalAssembling::default_local_assembling_database).

The compute path is the HIP library only: if libpmcd.so is missing, every
entry point raises; there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpmcd.so")
_lib = None

GEOMETRIES = {"UnitSquare": 0, "TwoTriangles": 1, "TwoTriangles2": 2,
              "Default_UnitCube_Hexa": 3, "Default_UnitCube_Tetra": 4,
              "StructuredHexa": 10, "StructuredTetra": 11}
MESH_FILE = 12
EXAMPLES = {"sine_laplace_2d": 0, "sine_2d": 1, "sine_3d": 2, "test_p0": 3,
            "test_p1": 4, "test_p2": 5, "const_rhs": 6, "linear_space_time": 7, "quadratic_space_time": 8,
            "exp_sin_cos_2d": 9, "sharp_boundary_layer_2d": 10, "smooth_solution_3d": 11}
METHODS = {"direct": 0, "cg": 1, "fgmres": 2, "richardson": 3, "sor": 4, "ssor": 5}
PRECONDS = {"no_preconditioner": 0, "jacobi": 1, "multigrid": 2, "sor": 3, "ssor": 4}
SMOOTHERS = {"jacobi": 0, "sor": 1, "ssor": 2}
CYCLES = {"V": 0, "W": 1, "F": 2}
COARSE_SMOOTHERS = {"direct_solve": 0, "jacobi": 1, "sor": 2, "ssor": 3}


class ProblemDesc(ctypes.Structure):
    _fields_ = [("geometry", ctypes.c_int), ("n_refine", ctypes.c_int),
                ("structured_n", ctypes.c_int), ("order", ctypes.c_int),
                ("supg", ctypes.c_int), ("linear_quad", ctypes.c_int),
                ("mg_levels", ctypes.c_int), ("example", ctypes.c_int),
                ("eps", ctypes.c_double), ("b", ctypes.c_double * 3),
                ("c", ctypes.c_double), ("f0", ctypes.c_double), ("mesh_file", ctypes.c_char_p),
                ("adapt_steps", ctypes.c_int), ("indicator", ctypes.c_void_p), ("indicator_ctx", ctypes.c_void_p),
                ("adapt_conforming", ctypes.c_int)]


# double (*)(const double* x, void* ctx): indicator of the refinement by indicator
INDICATOR_FN = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.c_void_p)


class SolverDesc(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int), ("precond", ctypes.c_int),
                ("max_iter", ctypes.c_int), ("min_iter", ctypes.c_int),
                ("res_tol", ctypes.c_double), ("res_red", ctypes.c_double),
                ("div_factor", ctypes.c_double), ("restart", ctypes.c_int),
                ("n_pre", ctypes.c_int), ("n_post", ctypes.c_int),
                ("jacobi_damp", ctypes.c_double), ("mg_damp", ctypes.c_double),
                ("mixed", ctypes.c_int), ("smoother", ctypes.c_int), ("sor_omega", ctypes.c_double),
                ("cycle", ctypes.c_int), ("coarse_smoother", ctypes.c_int),
                ("coarse_residual", ctypes.c_double), ("coarse_max_iter", ctypes.c_int)]


class SolveStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("converged", ctypes.c_int),
                ("initial_residual", ctypes.c_double), ("final_residual", ctypes.c_double),
                ("seconds", ctypes.c_double)]


EXPORTS = [
    "pmcd_last_error", "pmcd_device_count", "pmcd_mem_info", "pmcd_system_set_data",
    "pmcd_host_create", "pmcd_host_destroy", "pmcd_host_info", "pmcd_host_dofmap",
    "pmcd_host_pattern", "pmcd_host_cell_coords", "pmcd_host_hanging", "pmcd_get_hanging", "pmcd_host_sweep_order", "pmcd_host_sell_layout", "pmcd_host_sell_footprint", "pmcd_sell_stats", "pmcd_apply_transfer",
    "pmcd_host_write_mesh", "pmcd_cycle_directions", "pmcd_vector_write", "pmcd_vector_read",
    "pmcd_vector_write_market", "pmcd_matrix_write_market", "pmcd_tcd_init", "pmcd_tcd_step", "pmcd_afc_solve",
    "pmcd_jacobi_sweep",
    "pmcd_system_create", "pmcd_system_destroy", "pmcd_system_info", "pmcd_get_dofmap",
    "pmcd_get_pattern", "pmcd_get_cell_coords", "pmcd_assemble", "pmcd_get_matrix",
    "pmcd_get_rhs", "pmcd_solve", "pmcd_get_solution", "pmcd_set_solution", "pmcd_errors",
    "pmcd_apply", "pmcd_sor_sweep", "pmcd_vcycle", "pmcd_time_kernel", "pmcd_set_tuning", "pmcd_get_tuning",
    "pmcd_probe_mfma",
    "pmcd_spmv_device", "pmcd_dot_device", "pmcd_dot_device_noalloc",
    "pmcd_assemble_into", "pmcd_assembler_create", "pmcd_assembler_destroy", "pmcd_assembler_assemble",
    "pmcd_assembler_assemble_device", "pmcd_solver_create", "pmcd_solver_destroy", "pmcd_solver_update_matrix",
    "pmcd_solver_solve", "pmcd_solver_is_using_multigrid", "pmcd_get_transfer", "pmcd_axpy_device",
    "pmcd_norm_device", "pmcd_jacobi_device", "pmcd_prolongate_device", "pmcd_restrict_device",
    "pmcd_batch_instances",
    "pmcd_batch_create", "pmcd_batch_destroy", "pmcd_batch_info", "pmcd_batch_assemble",
    "pmcd_batch_solve", "pmcd_batch_get_solution", "pmcd_batch_set_solution", "pmcd_batch_get_matrix",
    "pmcd_batch_get_rhs",
    "pmcd_system_device_bytes", "pmcd_estimate_device_bytes", "pmcd_estimate_counts", "pmcd_autotune",
    "pmcd_server_create", "pmcd_server_destroy", "pmcd_server_now", "pmcd_server_submit", "pmcd_server_warm",
    "pmcd_server_wait", "pmcd_server_result", "pmcd_server_solution", "pmcd_server_release", "pmcd_server_info",
]


def mem_info():
    """(free, total) device memory in bytes."""
    f, t = ctypes.c_ulonglong(0), ctypes.c_ulonglong(0)
    _check(lib().pmcd_mem_info(ctypes.byref(f), ctypes.byref(t)))
    return f.value, t.value


def set_sor_exact_limit(n):
    """SOR sweeps of systems created afterwards follow the exact level
    schedule (bit-identical to the serial sweep) when it has at most n
    classes, greedy multicolour order otherwise (default 24)."""
    _check(lib().pmcd_set_tuning(5, int(n)))


def set_vcycle_graphs(on):
    """Replay V-cycles from captured HIP graphs, or launch them kernel by
    kernel (default)."""
    _check(lib().pmcd_set_tuning(6, 1 if on else 0))


def get_vcycle_graphs():
    v = ctypes.c_int(0)
    _check(lib().pmcd_get_tuning(6, ctypes.byref(v)))
    return bool(v.value)


def write_vector(path, v):
    """BlockVector::write_to_stream: the length on its own line, then the raw
    doubles (the reference's checkpoint format)."""
    v = np.ascontiguousarray(v, dtype=np.float64)
    _check(lib().pmcd_vector_write(os.fsencode(path), ctypes.c_longlong(len(v)), _ptr(v, ctypes.c_double)))


def read_vector(path, n):
    """BlockVector::read_from_stream of a vector of length n."""
    out = np.empty(n, dtype=np.float64)
    _check(lib().pmcd_vector_read(os.fsencode(path), ctypes.c_longlong(n), _ptr(out, ctypes.c_double)))
    return out


def write_vector_market(path, v):
    """BlockVector::write (MatrixMarket array format)."""
    v = np.ascontiguousarray(v, dtype=np.float64)
    _check(lib().pmcd_vector_write_market(os.fsencode(path), ctypes.c_longlong(len(v)), _ptr(v, ctypes.c_double)))


def write_matrix_market(path, row_ptr, col, val, n_cols=None):
    """TMatrix::write (MatrixMarket coordinate format, 1-based)."""
    rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
    c = np.ascontiguousarray(col, dtype=np.int32)
    v = np.ascontiguousarray(val, dtype=np.float64)
    n = len(rp) - 1
    _check(lib().pmcd_matrix_write_market(os.fsencode(path), n, n if n_cols is None else int(n_cols),
                                          _ptr(rp, ctypes.c_int), _ptr(c, ctypes.c_int), _ptr(v, ctypes.c_double)))


def cycle_directions(cycle, n_levels):
    """Directions of one multigrid cycle ('V', 'W', 'F'): 0 down, 1 up, 2 end,
    3 the leading dummy (CycleControl)."""
    n = ctypes.c_int(0)
    out = np.empty(4096, dtype=np.int32)
    _check(lib().pmcd_cycle_directions(CYCLES[cycle], int(n_levels), _ptr(out, ctypes.c_int), len(out),
                                       ctypes.byref(n)))
    return out[:min(n.value, len(out))].copy()


def set_sell_unroll(u):
    """Entries per lane loaded per step by the sliced ELL kernels (4, 8, 16)."""
    _check(lib().pmcd_set_tuning(9, int(u)))


def get_sell_unroll():
    v = ctypes.c_int(0)
    _check(lib().pmcd_get_tuning(9, ctypes.byref(v)))
    return v.value


def set_cg_device(on):
    """CG with no or Jacobi preconditioning: scalars kept on the device, two
    fused kernels per iteration (default), or the loop that reads every
    scalar back to the host."""
    _check(lib().pmcd_set_tuning(7, 1 if on else 0))


def get_cg_device():
    v = ctypes.c_int(0)
    _check(lib().pmcd_get_tuning(7, ctypes.byref(v)))
    return bool(v.value)


def get_sor_exact_limit():
    v = ctypes.c_int(0)
    _check(lib().pmcd_get_tuning(5, ctypes.byref(v)))
    return v.value


def set_assembly_variant(variant):
    """0: reference operation order on vector ALUs; 1: per-point FP64 MFMA
    for Q2 hexahedra; 2 (default): reference tensors, one FP64-MFMA GEMM per
    level (every element but P1 simplices, which keep a thread per cell)."""
    _check(lib().pmcd_set_tuning(4, int(variant)))


def assembly_variant():
    v = ctypes.c_int(0)
    _check(lib().pmcd_get_tuning(4, ctypes.byref(v)))
    return v.value


def probe_mfma():
    out = np.zeros(256)
    _check(lib().pmcd_probe_mfma(_ptr(out, ctypes.c_double)))
    return out.reshape(64, 4)


def set_tuning(kernel, lanes_per_row, variant=0):
    """Launch shape of the 'spmv' or 'jacobi' kernels: lanes per row (4, 8,
    16, 32) and variant (0 plain, 1 streaming with XCD-contiguous rows, 2
    unrolled: all loads of a lane issued before the products, used where the
    longest row fits 16 entries per lane, 3 unrolled with non-temporal matrix
    loads, 4 sliced ELL: one lane per row, rows summed serially in column
    order; the lane count is not used)."""
    k = {"spmv": 0, "jacobi": 1}[kernel]
    _check(lib().pmcd_set_tuning(k, int(lanes_per_row)))
    _check(lib().pmcd_set_tuning(k + 2, int(variant)))

def tuning(kernel):
    """(lanes per row, variant) currently set for the 'spmv' or 'jacobi' kernels."""
    k = {"spmv": 0, "jacobi": 1}[kernel]
    out = []
    for key in (k, k + 2):
        v = ctypes.c_int(0)
        _check(lib().pmcd_get_tuning(key, ctypes.byref(v)))
        out.append(v.value)
    return tuple(out)


KERNELS = {"spmv_f64": 0, "jacobi_f64": 1, "spmv_f32": 2, "assemble_local": 3, "gather_rows": 4,
           "jacobi_f32": 5, "stream_copy": 6, "stream_copy4": 7, "stream_read": 8}


def lib():
    """Load libpmcd.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libpmcd.so not built: run __graft_entry__.build() "
                               "(python -m parmoon2022_amd.build)")
        L = ctypes.CDLL(LIB_PATH)
        L.pmcd_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().pmcd_last_error().decode())


def _ptr(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def problem_desc(db):
    """Build a ProblemDesc from a parameter dictionary (reference key names)."""
    d = ProblemDesc()
    geo = db.get("geo_file", "UnitSquare")
    if str(geo).endswith((".mesh", ".GEO", ".geo")):
        # the struct holds the path; keep the bytes alive with it
        d._mesh_path = os.fsencode(os.path.abspath(geo))
        d.geometry = MESH_FILE
        d.mesh_file = d._mesh_path
    else:
        d.geometry = GEOMETRIES[geo]
    d.n_refine = int(db.get("refinement_n_initial_steps", 0))
    d.structured_n = int(db.get("structured_n", 1))
    d.order = int(db.get("ANSATZ_ORDER", 1))
    d.supg = 1 if db.get("space_discretization_type", "galerkin") == "supg" else 0
    d.linear_quad = int(db.get("INTERNAL_PROBLEM_LINEAR", 0))
    use_mg = db.get("preconditioner") == "multigrid" and db.get("solver_type") == "iterative"
    d.mg_levels = int(db.get("multigrid_n_levels", 1)) if use_mg else 1
    d.example = EXAMPLES[db.get("example", "sine_laplace_2d")]
    if "diffusion_coefficient" in db or "reynolds_number" not in db:
        d.eps = float(db.get("diffusion_coefficient", 1.0))
    else:
        d.eps = 1.0 / float(db["reynolds_number"])  # examples with eps = 1 / RE_NR
    b = db.get("convection", (0.0, 0.0, 0.0))
    for k in range(3):
        d.b[k] = float(b[k]) if k < len(b) else 0.0
    d.c = float(db.get("reaction", 0.0))
    d.f0 = float(db.get("rhs_constant", 1.0))
    steps = int(db.get("refine_by_indicator_steps", 0))
    if steps > 0:
        # test/adaptive_hanging_test.c++:120-135: RefineByIndicator(indicator)
        # steps times; conforming_closure (the domain default, true) closes
        # the mesh with bisections, false leaves hanging vertices
        d.adapt_conforming = 1 if db.get("conforming_closure", True) else 0
        d.adapt_steps = steps
        d.indicator = ctypes.cast(indicator_callback(db["refine_indicator"]), ctypes.c_void_p)
        d.indicator_ctx = None
    return d


_INDICATORS = {}


def indicator_callback(f):
    """C callback for an indicator: a callable f(x, y, z) -> float, or the
    coefficients (a, b, c, e) of a x^2 + b y + c z + e. Callbacks are kept for
    the life of the process (a served request may call one later, on a server
    thread), and the same indicator gives the same callback, so requests on
    one adaptive mesh share resident systems."""
    key = tuple(float(v) for v in f) if isinstance(f, (tuple, list)) else f
    cb = _INDICATORS.get(key)
    if cb is None:
        if isinstance(key, tuple):
            a, b, c, e = key
            fn = lambda x, y, z: a * x * x + b * y + c * z + e  # noqa: E731
        else:
            fn = f
        cb = _INDICATORS[key] = INDICATOR_FN(lambda p, ctx: float(fn(p[0], p[1], p[2])))
    return cb


def solver_desc(db):
    """Build a SolverDesc from reference solver/multigrid parameter names."""
    s = SolverDesc()
    if db.get("solver_type", "direct") == "direct":
        s.method = 0
    else:
        s.method = METHODS[db.get("iterative_solver_type", "fgmres")]
    s.precond = PRECONDS[db.get("preconditioner", "no_preconditioner")]
    s.max_iter = int(db.get("max_n_iterations", 100))
    s.min_iter = int(db.get("min_n_iterations", 0))
    s.res_tol = float(db.get("residual_tolerance", 1e-8))
    s.res_red = float(db.get("residual_reduction", 0.0))
    s.div_factor = float(db.get("divergence_factor", 1e10))
    s.restart = int(db.get("gmres_restart", 20))
    s.n_pre = int(db.get("multigrid_n_pre_smooth", 1))
    s.n_post = int(db.get("multigrid_n_post_smooth", 1))
    s.jacobi_damp = float(db.get("jacobi_damp", 1.0))
    s.mg_damp = float(db.get("multigrid_correction_damp_factor", 1.0))
    s.mixed = int(db.get("multigrid_mixed_precision", 0))
    s.smoother = SMOOTHERS[db.get("multigrid_smoother", "jacobi")]
    s.sor_omega = float(db.get("sor_omega", 1.5))
    s.cycle = CYCLES[db.get("multigrid_cycle_type", "V")]
    s.coarse_smoother = COARSE_SMOOTHERS[db.get("multigrid_smoother_coarse", "direct_solve")]
    s.coarse_residual = float(db.get("multigrid_coarse_residual", 0.1))
    s.coarse_max_iter = int(db.get("multigrid_coarse_max_n_iterations", 10))
    return s


class HostSetup:
    """Host-only mesh/FE-space/sparsity setup (no GPU)."""

    def __init__(self, db):
        self._h = ctypes.c_void_p()
        self.desc = problem_desc(db)
        _check(lib().pmcd_host_create(ctypes.byref(self.desc), ctypes.byref(self._h)))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().pmcd_host_destroy(self._h)
            self._h = None

    def info(self, level=-1):
        a = (ctypes.c_longlong * 8)()
        _check(lib().pmcd_host_info(self._h, level, a))
        keys = ["n_cells", "n_dof", "n_active", "nnz", "n_local", "n_quad", "n_levels", "n_dirichlet"]
        return dict(zip(keys, list(a)))

    def dofmap(self, level=-1):
        i = self.info(level)
        out = np.empty(i["n_cells"] * i["n_local"], dtype=np.int32)
        _check(lib().pmcd_host_dofmap(self._h, level, _ptr(out, ctypes.c_int)))
        return out.reshape(i["n_cells"], i["n_local"])

    def pattern(self, level=-1):
        i = self.info(level)
        rp = np.empty(i["n_dof"] + 1, dtype=np.int32)
        col = np.empty(i["nnz"], dtype=np.int32)
        _check(lib().pmcd_host_pattern(self._h, level, _ptr(rp, ctypes.c_int), _ptr(col, ctypes.c_int)))
        return rp, col

    def sweep_order(self, level=-1, max_exact_classes=None):
        """Device SOR row classes: (rows, class_ptr, exact); the limit defaults
        to the library's current one (set_sor_exact_limit)."""
        if max_exact_classes is None:
            max_exact_classes = get_sor_exact_limit()
        n = self.info(level)["n_dof"]
        rows = np.empty(n, dtype=np.int32)
        ptr = np.empty(n + 1, dtype=np.int32)
        k, ex = ctypes.c_int(0), ctypes.c_int(0)
        _check(lib().pmcd_host_sweep_order(self._h, level, int(max_exact_classes), _ptr(rows, ctypes.c_int),
                                           _ptr(ptr, ctypes.c_int), ctypes.byref(k), ctypes.byref(ex)))
        return rows, ptr[:k.value + 1].copy(), bool(ex.value)

    def sell_layout(self, level=-1):
        """Sliced ELL layout of the level's active rows: (perm, len, slice_off),
        slot t of slice t // 64 holding row perm[t] (-1: padding) whose entry k
        sits at slice_off[t // 64] + 64 k + t % 64."""
        na = self.info(level)["n_active"]
        m = (na + 63) // 64 * 64 + 64
        perm = np.empty(m, dtype=np.int32)
        ln = np.empty(m, dtype=np.int32)
        off = np.empty(m // 64 + 1, dtype=np.int64)
        ns = ctypes.c_int(0)
        _check(lib().pmcd_host_sell_layout(self._h, level, _ptr(perm, ctypes.c_int), _ptr(ln, ctypes.c_int),
                                           _ptr(off, ctypes.c_longlong), ctypes.byref(ns)))
        k = ns.value
        return perm[:k].copy(), ln[:k].copy(), off[:k // 64 + 1].copy()

    def sell_footprint(self, level=-1):
        """Column footprints of the sliced ELL layout's 256-slot blocks:
        (fp_off, fp, lcol, fp_max), block b's distinct columns ascending in
        fp[fp_off[b]:fp_off[b+1]], lcol[e] the position of entry e's column
        in its block's list; None when a block is wider than the kernels take."""
        sizes = np.zeros(4, dtype=np.int64)
        fn = lib().pmcd_host_sell_footprint
        _check(fn(self._h, level, None, None, None, _ptr(sizes, ctypes.c_longlong)))
        if sizes[0] == 0:
            return None
        fp_off = np.empty(sizes[0], dtype=np.int32)
        fp = np.empty(max(1, sizes[1]), dtype=np.int32)
        lcol = np.empty(max(1, sizes[2]), dtype=np.uint16)
        _check(fn(self._h, level, _ptr(fp_off, ctypes.c_int), _ptr(fp, ctypes.c_int), _ptr(lcol, ctypes.c_ushort),
                  _ptr(sizes, ctypes.c_longlong)))
        return fp_off, fp[:sizes[1]], lcol[:sizes[2]], int(sizes[3])

    def write_mesh(self, path, level=-1):
        """The level's mesh as a 3D .mesh file (readable as geo_file)."""
        _check(lib().pmcd_host_write_mesh(self._h, level, os.fsencode(path)))

    def cell_coords(self, level=-1):
        i = self.info(level)
        if self.desc.geometry == MESH_FILE and self.desc.mesh_file.lower().endswith(b".geo"):
            nv = {3: 3, 6: 3, 4: 4, 9: 4}[i["n_local"]]  # triangles or quadrilaterals
        elif self.desc.geometry == MESH_FILE:  # tetrahedra (P1 4, P2 10 DOFs) or hexahedra (8, 27)
            nv = {4: 4, 10: 4, 8: 8, 27: 8}[i["n_local"]]
        else:
            nv = VERTICES_PER_CELL[self.desc.geometry]
        out = np.empty(i["n_cells"] * nv * 3, dtype=np.float64)
        _check(lib().pmcd_host_cell_coords(self._h, level, _ptr(out, ctypes.c_double)))
        return out.reshape(i["n_cells"], nv, 3)


VERTICES_PER_CELL = {0: 4, 1: 3, 2: 3, 3: 8, 4: 4, 10: 8, 11: 4}


class ConvectionDiffusion:
    """Device-resident CD problem: ConvectionDiffusion<d> of the reference."""

    def __init__(self, db):
        self.db = dict(db)
        self._s = ctypes.c_void_p()
        self.desc = problem_desc(self.db)
        _check(lib().pmcd_system_create(ctypes.byref(self.desc), ctypes.byref(self._s)))
        self.errors = np.zeros(5)
        self.stats = SolveStats()
        self.assemble_ms = 0.0

    def close(self):
        if getattr(self, "_s", None):
            lib().pmcd_system_destroy(self._s)
            self._s = None

    def __del__(self):
        self.close()

    def info(self, level=-1):
        a = (ctypes.c_longlong * 8)()
        _check(lib().pmcd_system_info(self._s, level, a))
        keys = ["n_cells", "n_dof", "n_active", "nnz", "n_local", "n_quad", "n_levels", "n_dirichlet"]
        return dict(zip(keys, list(a)))

    def assemble(self):
        ms = ctypes.c_double()
        _check(lib().pmcd_assemble(self._s, ctypes.byref(ms)))
        self.assemble_ms = ms.value

    def device_bytes(self):
        """(bytes, peak bytes) of the system's device buffers, each counted in
        whole 4 KiB pages."""
        b, p = ctypes.c_longlong(), ctypes.c_longlong()
        _check(lib().pmcd_system_device_bytes(self._s, ctypes.byref(b), ctypes.byref(p)))
        return b.value, p.value

    def set_data(self, db):
        """New PDE data (example, coefficients) on the same mesh hierarchy;
        reassemble afterwards."""
        desc = problem_desc(db)
        _check(lib().pmcd_system_set_data(self._s, ctypes.byref(desc)))
        self.db = dict(db)
        self.desc = desc

    def solve(self, db=None):
        sd = solver_desc(self.db if db is None else db)
        st = SolveStats()
        _check(lib().pmcd_solve(self._s, ctypes.byref(sd), ctypes.byref(st)))
        self.stats = st
        return st

    def output(self):
        e = np.zeros(5)
        _check(lib().pmcd_errors(self._s, _ptr(e, ctypes.c_double)))
        self.errors = e
        return e

    def get_L2_error(self):
        return self.errors[0]

    def get_H1_semi_error(self):
        return self.errors[1]

    def get_SD_error(self):
        return self.errors[2]

    def get_L_inf_error(self):
        return self.errors[4]

    def dofmap(self, level=-1):
        i = self.info(level)
        out = np.empty(i["n_cells"] * i["n_local"], dtype=np.int32)
        _check(lib().pmcd_get_dofmap(self._s, level, _ptr(out, ctypes.c_int)))
        return out.reshape(i["n_cells"], i["n_local"])

    def pattern(self, level=-1):
        i = self.info(level)
        rp = np.empty(i["n_dof"] + 1, dtype=np.int32)
        col = np.empty(i["nnz"], dtype=np.int32)
        _check(lib().pmcd_get_pattern(self._s, level, _ptr(rp, ctypes.c_int), _ptr(col, ctypes.c_int)))
        return rp, col

    def matrix_values(self, level=-1):
        i = self.info(level)
        v = np.empty(i["nnz"], dtype=np.float64)
        _check(lib().pmcd_get_matrix(self._s, level, _ptr(v, ctypes.c_double)))
        return v

    def rhs(self, level=-1):
        i = self.info(level)
        v = np.empty(i["n_dof"], dtype=np.float64)
        _check(lib().pmcd_get_rhs(self._s, level, _ptr(v, ctypes.c_double)))
        return v

    def solution(self):
        i = self.info(-1)
        v = np.empty(i["n_dof"], dtype=np.float64)
        _check(lib().pmcd_get_solution(self._s, _ptr(v, ctypes.c_double)))
        return v

    def set_solution(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        _check(lib().pmcd_set_solution(self._s, _ptr(v, ctypes.c_double)))

    def apply(self, x, level=-1, y=None, a=1.0):
        """BlockFEMatrix::apply (y None) or apply_scaled_add (y given)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        mode = 0 if y is None else 1
        out = np.zeros_like(x) if y is None else np.array(y, dtype=np.float64)
        _check(lib().pmcd_apply(self._s, level, _ptr(x, ctypes.c_double),
                                _ptr(out, ctypes.c_double), ctypes.c_double(a), mode))
        return out

    def transfer(self, level=-1):
        """Prolongation of a level from the next coarser one: (row_ptr, col,
        val, n_coarse)."""
        nf, nc, nz = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib().pmcd_get_transfer(self._s, level, ctypes.byref(nf), ctypes.byref(nc), ctypes.byref(nz),
                                       None, None, None))
        rp = np.zeros(nf.value + 1, np.int32)
        col = np.zeros(nz.value, np.int32)
        val = np.zeros(nz.value)
        _check(lib().pmcd_get_transfer(self._s, level, None, None, None, _ptr(rp, ctypes.c_int),
                                       _ptr(col, ctypes.c_int), _ptr(val, ctypes.c_double)))
        return rp, col, val, nc.value

    def write_solution(self, path):
        """The solution in the reference's checkpoint format (write_to_file)."""
        write_vector(path, self.solution())

    def read_solution(self, path):
        """Restart from a solution written by write_solution (read_from_file)."""
        self.set_solution(read_vector(path, self.info()["n_dof"]))

    def write_matrix(self, path, level=-1):
        """The level's matrix as MatrixMarket coordinates (TMatrix::write)."""
        rp, col = self.pattern(level)
        write_matrix_market(path, rp, col, self.matrix_values(level))

    def jacobi_sweep(self, b, x, omega=0.8, level=-1):
        """JacobiSmoother::smooth, one step: x + omega D^{-1} (b - A x)."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        out = np.array(x, dtype=np.float64)
        _check(lib().pmcd_jacobi_sweep(self._s, level, _ptr(b, ctypes.c_double), _ptr(out, ctypes.c_double),
                                       ctypes.c_double(omega)))
        return out

    def sor_sweep(self, b, x, omega=1.5, flag=2, level=-1):
        """TMatrix::sor_sweep on the level's matrix: returns (new x, exact) where
        exact tells whether the device order reproduces the serial sweep."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        out = np.array(x, dtype=np.float64)
        ex = ctypes.c_int(0)
        _check(lib().pmcd_sor_sweep(self._s, level, _ptr(b, ctypes.c_double), _ptr(out, ctypes.c_double),
                                    ctypes.c_double(omega), int(flag), ctypes.byref(ex)))
        return out, bool(ex.value)

    def apply_transfer(self, which, x, y=None, level=-1, a=1.0):
        """The V-cycle's transfer from/to `level`: which "prolong" returns
        y + a P x (x coarse, y fine; y zero when None), "restrict" returns R x
        (x fine); the kernels the V-cycle launches."""
        k = {"prolong": 0, "restrict": 1}[which]
        i_f, i_c = self.info(level), self.info(level - 1)
        ny = i_f["n_dof"] if k == 0 else i_c["n_dof"]
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros(ny) if y is None else np.array(y, dtype=np.float64)
        fn = lib().pmcd_apply_transfer
        fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double]
        _check(fn(self._s, level, k, x.ctypes.data, out.ctypes.data, a))
        return out

    def sell_stats(self, level=-1):
        """The level's sliced ELL layout as built: slots, layout entries,
        footprint blocks, footprint entries, widest footprint, launch mode."""
        out = np.zeros(6, dtype=np.int64)
        _check(lib().pmcd_sell_stats(self._s, level, _ptr(out, ctypes.c_longlong)))
        return dict(zip(("slots", "entries", "fp_blocks", "fp_entries", "fp_max", "mode"), (int(v) for v in out)))

    def time_kernel(self, name, level=-1, iters=20):
        """Average device time (ms) of one kernel from HIP events on the system stream."""
        ms = ctypes.c_double()
        _check(lib().pmcd_time_kernel(self._s, KERNELS[name], level, iters, ctypes.byref(ms)))
        return ms.value

    def vcycle(self, v, db=None):
        sd = solver_desc(self.db if db is None else db)
        v = np.ascontiguousarray(v, dtype=np.float64)
        z = np.zeros_like(v)
        _check(lib().pmcd_vcycle(self._s, ctypes.byref(sd), _ptr(v, ctypes.c_double),
                                 _ptr(z, ctypes.c_double)))
        return z


# ---- drop-in interface on caller-owned spaces, structures and matrices -------

SHAPES = {"triangle": 0, "quadrangle": 1, "tetrahedron": 2, "hexahedron": 3}


class FeSpaceDesc(ctypes.Structure):
    _fields_ = [("shape", ctypes.c_int), ("order", ctypes.c_int), ("n_cells", ctypes.c_int),
                ("n_dof", ctypes.c_int), ("n_active", ctypes.c_int),
                ("vertices", ctypes.POINTER(ctypes.c_double)), ("global_numbers", ctypes.POINTER(ctypes.c_int)),
                ("begin_index", ctypes.POINTER(ctypes.c_int)), ("n_hanging", ctypes.c_int),
                ("hanging", ctypes.POINTER(ctypes.c_int)), ("hanging_weights", ctypes.POINTER(ctypes.c_double))]


class StructureDesc(ctypes.Structure):
    _fields_ = [("n_rows", ctypes.c_int), ("row_ptr", ctypes.POINTER(ctypes.c_int)),
                ("col", ctypes.POINTER(ctypes.c_int))]


class CsrLevel(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("n_active", ctypes.c_int), ("row_ptr", ctypes.POINTER(ctypes.c_int)),
                ("col", ctypes.POINTER(ctypes.c_int)), ("entries", ctypes.POINTER(ctypes.c_double)),
                ("n_coarse", ctypes.c_int), ("p_row_ptr", ctypes.POINTER(ctypes.c_int)),
                ("p_col", ctypes.POINTER(ctypes.c_int)), ("p_entries", ctypes.POINTER(ctypes.c_double))]


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Assembler:
    """Assemble3D / Assemble2D with the convection-diffusion LocalAssembling
    (include/AssembleRoutines/Assemble3D.h:28-34) on a caller's FE space and
    structure: `shape` a key of SHAPES, `vertices` [n_cells, nv, 3] in the
    element's vertex order, `global_numbers` / `begin_index` the space's
    local-to-global map, `n_active` its active bound (hanging DOFs
    included), (`row_ptr`, `col`) the structure, `db` the PDE data and forms
    (reference keys), `hanging` an optional pair (n x 4 ints: hanging DOF and
    its coupling DOFs, -1 padded; n x 3 weights). assemble() adds into the
    caller's entries and rhs like the reference's scatter."""

    def __init__(self, shape, order, vertices, global_numbers, begin_index, n_dof, n_active, row_ptr, col, db,
                 hanging=None):
        self._keep = [_f64(vertices), _i32(global_numbers), _i32(begin_index), _i32(row_ptr), _i32(col)]
        v, g, b, rp, cl = self._keep
        nh, hp, hw = 0, None, None
        if hanging is not None and len(hanging[0]):
            hd, hwt = _i32(np.reshape(hanging[0], (-1, 4))), _f64(np.reshape(hanging[1], (-1, 3)))
            self._keep += [hd, hwt]
            nh, hp, hw = len(hd), _ptr(hd, ctypes.c_int), _ptr(hwt, ctypes.c_double)
        self.space = FeSpaceDesc(SHAPES[shape] if isinstance(shape, str) else int(shape), int(order),
                                 len(b) - 1, int(n_dof), int(n_active), _ptr(v, ctypes.c_double),
                                 _ptr(g, ctypes.c_int), _ptr(b, ctypes.c_int), nh, hp, hw)
        self.structure = StructureDesc(int(n_dof), _ptr(rp, ctypes.c_int), _ptr(cl, ctypes.c_int))
        self.desc = problem_desc(db)
        self.nnz = int(rp[-1])
        self._h = ctypes.c_void_p()
        _check(lib().pmcd_assembler_create(ctypes.byref(self.space), ctypes.byref(self.structure),
                                           ctypes.byref(self.desc), ctypes.byref(self._h)))

    def assemble(self, entries, rhs, dirichlet_values=None):
        """entries (float64, nnz) and rhs (float64, n_dof) are updated in place."""
        assert entries.dtype == np.float64 and entries.flags.c_contiguous and entries.size == self.nnz
        assert rhs.dtype == np.float64 and rhs.flags.c_contiguous and rhs.size == self.space.n_dof
        dv = None if dirichlet_values is None else _f64(dirichlet_values)
        _check(lib().pmcd_assembler_assemble(self._h, None if dv is None else _ptr(dv, ctypes.c_double),
                                             _ptr(entries, ctypes.c_double), _ptr(rhs, ctypes.c_double)))

    def assemble_device(self, entries_ptr, rhs_ptr, dirichlet_ptr=None, stream=None):
        """Device-pointer variant (integer addresses)."""
        fn = lib().pmcd_assembler_assemble_device
        fn.argtypes = [ctypes.c_void_p] * 5
        _check(fn(self._h, dirichlet_ptr, entries_ptr, rhs_ptr, stream))

    def close(self):
        if getattr(self, "_h", None):
            lib().pmcd_assembler_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def assemble_into(shape, order, vertices, global_numbers, begin_index, n_dof, n_active, row_ptr, col, db,
                  entries, rhs, dirichlet_values=None, hanging=None):
    """One-shot pmcd_assemble_into (see Assembler)."""
    a = Assembler(shape, order, vertices, global_numbers, begin_index, n_dof, n_active, row_ptr, col, db,
                  hanging)
    try:
        a.assemble(entries, rhs, dirichlet_values)
    finally:
        a.close()


class Solver:
    """Solver<BlockFEMatrix, BlockVector> on caller-owned matrices
    (include/Solver/Solver.h:45-94): Solver(db), update_matrix(levels),
    solve(b, x), is_using_multigrid(). `levels`, coarsest first: dicts with
    n_active, row_ptr, col, entries and, above the coarsest, the prolongation
    p_row_ptr, p_col, p_entries from the next coarser level."""

    def __init__(self, db):
        self.db = dict(db)
        self.sd = solver_desc(db)
        self._h = ctypes.c_void_p()
        _check(lib().pmcd_solver_create(ctypes.byref(self.sd), ctypes.byref(self._h)))
        self.stats = None

    def update_matrix(self, levels):
        arr = (CsrLevel * len(levels))()
        keep = []
        prev_n = 0
        for i, lv in enumerate(levels):
            rp, cl, ent = _i32(lv["row_ptr"]), _i32(lv["col"]), _f64(lv["entries"])
            keep += [rp, cl, ent]
            c = arr[i]
            c.n = len(rp) - 1
            c.n_active = int(lv["n_active"])
            c.row_ptr, c.col, c.entries = _ptr(rp, ctypes.c_int), _ptr(cl, ctypes.c_int), _ptr(ent, ctypes.c_double)
            if i > 0:
                prp, pcl, pent = _i32(lv["p_row_ptr"]), _i32(lv["p_col"]), _f64(lv["p_entries"])
                keep += [prp, pcl, pent]
                c.n_coarse = prev_n
                c.p_row_ptr, c.p_col = _ptr(prp, ctypes.c_int), _ptr(pcl, ctypes.c_int)
                c.p_entries = _ptr(pent, ctypes.c_double)
            prev_n = c.n
        _check(lib().pmcd_solver_update_matrix(self._h, len(levels), arr))

    def solve(self, b, x):
        """x: initial iterate on entry (float64, updated in place)."""
        b = _f64(b)
        assert x.dtype == np.float64 and x.flags.c_contiguous
        st = SolveStats()
        _check(lib().pmcd_solver_solve(self._h, _ptr(b, ctypes.c_double), _ptr(x, ctypes.c_double),
                                       ctypes.byref(st)))
        self.stats = st
        return st

    def is_using_multigrid(self):
        v = ctypes.c_int(0)
        _check(lib().pmcd_solver_is_using_multigrid(self._h, ctypes.byref(v)))
        return bool(v.value)

    def close(self):
        if getattr(self, "_h", None):
            lib().pmcd_solver_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class TcdDesc(ctypes.Structure):
    _fields_ = [("scheme", ctypes.c_int), ("bdf_first", ctypes.c_int), ("tau", ctypes.c_double),
                ("t0", ctypes.c_double)]


TIME_SCHEMES = {"backward_euler": 0, "crank_nicolson": 1, "bdf_two": 2}


AFC_PARAM_KEYS = [("afc_nonlinloop_epsilon", 1e-10), ("afc_nonlinloop_maxit", 1),
                  ("afc_nonlinloop_damping_factor", 1.0), ("afc_nonlinloop_damping_factor_max", 1.0),
                  ("afc_nonlinloop_damping_factor_min", 0.01), ("afc_nonlinloop_damping_factor_max_global", 1.0),
                  ("afc_nonlinloop_damping_factor_increase", 1.1), ("afc_nonlinloop_damping_factor_decrease", 0.5),
                  ("afc_nonlinloop_damping_factor_max_increase", 1.001),
                  ("afc_nonlinloop_damping_factor_max_decrease", 0.9),
                  ("afc_nonlinloop_damping_factor_min_tol", 1.001)]


def afc_params(db):
    """Parameters of the algebraic flux correction loop in the reference's
    names and defaults (AlgebraicFluxCorrection::default_afc_database,
    src/AssembleRoutines/AlgebraicFluxCorrection.C:1914-2056), as the 19
    doubles the C ABI takes: the 11 damping parameters, then the limiter
    (0 kuzmin, 1 BJK17; BJK17 for P1 triangles and tetrahedra), the iteration
    scheme (0 fixed_point_rhs, 1 fixed_point_matrix, the reference's default,
    2 newton), afc_fixed_point_matrix_weight, afc_fixed_point_derivative_weight,
    afc_change_method_threshold, afc_nonlinloop_anderson_acc (0/1),
    afc_nonlinloop_anderson_acc_vec and afc_nonlinloop_anderson_acc_start.
    Dynamic damping and the zero initial iterate."""
    fixed = {"afc_initial_iterate": "afc_zero", "afc_nonlinloop_damping_factor_constant": "no"}
    for k, v in fixed.items():
        if db.get(k, v) != v:
            raise ValueError("%s = %s is not built (only %s)" % (k, db.get(k), v))
    limiters = {"kuzmin": 0.0, "BJK17": 1.0}
    lim = db.get("afc_limiter", "kuzmin")
    if lim not in limiters:
        raise ValueError("afc_limiter = %s is not built (kuzmin, BJK17)" % lim)
    schemes = {"fixed_point_rhs": 0.0, "fixed_point_matrix": 1.0, "newton": 2.0}
    sch = db.get("afc_iteration_scheme", "fixed_point_matrix")
    if sch not in schemes:
        raise ValueError("afc_iteration_scheme = %s is not built (fixed_point_rhs, fixed_point_matrix, newton)" % sch)
    acc = db.get("afc_nonlinloop_anderson_acc", "no")
    if acc not in ("no", "yes"):
        raise ValueError("afc_nonlinloop_anderson_acc must be yes or no")
    return np.array([float(db.get(k, d)) for k, d in AFC_PARAM_KEYS] +
                    [limiters[lim], schemes[sch], float(db.get("afc_fixed_point_matrix_weight", 0.75)),
                     float(db.get("afc_fixed_point_derivative_weight", 0.1)),
                     float(db.get("afc_change_method_threshold", 1e-5)), 1.0 if acc == "yes" else 0.0,
                     float(db.get("afc_nonlinloop_anderson_acc_vec", 10)),
                     float(db.get("afc_nonlinloop_anderson_acc_start", 0))])


def tcd_desc(db):
    """TcdDesc from the reference's time parameters (time_discretization,
    bdf_two_first_step, time_step_length, time_start)."""
    t = TcdDesc()
    t.scheme = TIME_SCHEMES[db.get("time_discretization", "backward_euler")]
    t.bdf_first = TIME_SCHEMES[db.get("bdf_two_first_step", "backward_euler")]
    if t.bdf_first not in (0, 1):
        raise ValueError("bdf_two_first_step must be backward_euler or crank_nicolson")
    t.tau = float(db.get("time_step_length", 0.05))
    t.t0 = float(db.get("time_start", 0.0))
    return t


class TimeConvectionDiffusion(ConvectionDiffusion):
    """Time-dependent CD problem: TimeConvectionDiffusion<d> of the reference
    (src/System/TimeConvectionDiffusion.C). Driver loop as in the
    reference's programs:

        tcd = TimeConvectionDiffusion(db)
        tcd.assemble_initial_time()
        while tcd.current_time < db["time_end"] - 1e-10:
            tcd.step()      # set_time_disc_parameters, assemble, solve
            tcd.output()    # errors at the current time, L2(0,t;.) sums
    """

    def __init__(self, db):
        super().__init__(db)
        self.current_time = float(self.db.get("time_start", 0.0))
        self.current_step = 0
        # L2(0,t;L2)^2, L2 at t, L2(0,t;H1)^2, H1 at t, Linf(0,t;Linf), Linf(0,t;L2)
        self.time_errors = np.zeros(6)

    def assemble_initial_time(self):
        t = tcd_desc(self.db)
        _check(lib().pmcd_tcd_init(self._s, ctypes.byref(t)))
        self.current_time = t.t0
        self.current_step = 0

    def step(self, db=None):
        sd = solver_desc(self.db if db is None else db)
        st = SolveStats()
        tm = ctypes.c_double()
        _check(lib().pmcd_tcd_step(self._s, ctypes.byref(sd), ctypes.byref(st), ctypes.byref(tm)))
        self.current_time = tm.value
        self.current_step += 1
        self.stats = st
        return st

    def output(self):
        """Errors at the current time and their accumulation over time
        (TimeConvectionDiffusion::output: trapezoidal sums of the squared L2
        and H1-semi errors)."""
        e = super().output()
        tau = float(self.db.get("time_step_length", 0.05))
        te = self.time_errors
        te[0] += (e[0] * e[0] + te[1] * te[1]) * tau * 0.5
        te[1] = e[0]
        te[2] += (e[1] * e[1] + te[3] * te[3]) * tau * 0.5
        te[3] = e[1]
        te[4] = max(te[4], e[4])
        te[5] = max(te[5], e[0])
        return e

    def get_errors(self):
        """[L2(0,t;L2), L2(0,t;H1), Linf(0,t;Linf)] as get_errors() reports them."""
        te = self.time_errors
        return np.array([np.sqrt(te[0]), np.sqrt(te[2]), te[4]])


class ConvectionDiffusionAFC(ConvectionDiffusion):
    """ConvectionDiffusion_AFC<d> of the reference
    (src/System/ConvectionDiffusion_AFC.C) with the Kuzmin or BJK17 limiter,
    the fixed point rhs, fixed point matrix or Newton scheme, and optional
    Anderson acceleration (afc_params). `solve()` runs the whole nonlinear
    loop of the reference's programs (assemble(0), solve(0), solve(k) until
    converged)."""

    def __init__(self, db):
        db = dict(db)
        if db.get("space_discretization_type", "galerkin") != "galerkin":
            raise ValueError("algebraic flux correction uses the Galerkin discretization")
        super().__init__(db)
        self.afc = None

    def solve(self, db=None):
        d = self.db if db is None else db
        sd = solver_desc(d)
        prm = np.ascontiguousarray(afc_params(d))
        out = np.zeros(4)
        st = SolveStats()
        _check(lib().pmcd_afc_solve(self._s, ctypes.byref(sd), _ptr(prm, ctypes.c_double),
                                    _ptr(out, ctypes.c_double), ctypes.byref(st)))
        self.stats = st
        self.afc = {"iterations": int(out[0]), "residual": out[1], "rejected": int(out[2]),
                    "initial_residual": out[3]}
        return st


class Batch:
    """Many CD instances assembled and solved together (one device launch per
    step for all of them). `dbs` are per-instance parameter dictionaries with
    the same keys as ConvectionDiffusion's; ansatz order, discretisation and
    quadrature must agree, each instance is single-level."""

    def __init__(self, dbs):
        self.dbs = [dict(db) for db in dbs]
        descs = (ProblemDesc * len(self.dbs))(*[problem_desc(db) for db in self.dbs])
        self._b = ctypes.c_void_p()
        _check(lib().pmcd_batch_create(descs, len(self.dbs), ctypes.byref(self._b)))
        self.assemble_ms = 0.0

    def close(self):
        if getattr(self, "_b", None):
            lib().pmcd_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        self.close()

    def info(self, inst=-1):
        a = (ctypes.c_longlong * 5)()
        _check(lib().pmcd_batch_info(self._b, int(inst), a))
        keys = ["n_inst", "n_dof", "n_active", "nnz", "n_cells"]
        return dict(zip(keys, list(a)))

    def assemble(self):
        ms = ctypes.c_double()
        _check(lib().pmcd_batch_assemble(self._b, ctypes.byref(ms)))
        self.assemble_ms = ms.value

    def solve(self, db=None):
        """Flexible GMRES on every instance; returns one SolveStats per instance."""
        sd = solver_desc(self.dbs[0] if db is None else db)
        st = (SolveStats * len(self.dbs))()
        _check(lib().pmcd_batch_solve(self._b, ctypes.byref(sd), st))
        return list(st)

    def _vec(self, fn, inst, n):
        out = np.empty(n, dtype=np.float64)
        _check(fn(self._b, int(inst), _ptr(out, ctypes.c_double)))
        return out

    def solution(self, inst):
        return self._vec(lib().pmcd_batch_get_solution, inst, self.info(inst)["n_dof"])

    def set_solution(self, inst, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        _check(lib().pmcd_batch_set_solution(self._b, int(inst), _ptr(v, ctypes.c_double)))

    def rhs(self, inst):
        return self._vec(lib().pmcd_batch_get_rhs, inst, self.info(inst)["n_dof"])

    def matrix_values(self, inst):
        return self._vec(lib().pmcd_batch_get_matrix, inst, self.info(inst)["nnz"])


def c4_instances(n_inst=256, seed=20261202, n_range=(8, 40)):
    """Batched workload: unit cubes of N^3 cubes split into 6 tetrahedra each,
    N uniform in n_range, P1, SUPG, eps = 10^U(-6, 0), b a unit vector uniform
    on the sphere, f = 1, homogeneous Dirichlet data; drawn by
    std::mt19937_64 seeded `seed` in the library (pmcd_batch_instances)."""
    ns = np.zeros(n_inst, np.int32)
    epss = np.zeros(n_inst)
    bs = np.zeros(3 * n_inst)
    _check(lib().pmcd_batch_instances(ctypes.c_ulonglong(seed), n_inst, n_range[0], n_range[1],
                                      _ptr(ns, ctypes.c_int), _ptr(epss, ctypes.c_double),
                                      _ptr(bs, ctypes.c_double)))
    out = []
    for i in range(n_inst):
        n, eps, b = int(ns[i]), float(epss[i]), bs[3 * i:3 * i + 3]
        out.append({"geo_file": "StructuredTetra", "structured_n": n, "ANSATZ_ORDER": 1,
                    "refinement_n_initial_steps": 0, "example": "const_rhs", "space_discretization_type": "supg",
                    "diffusion_coefficient": eps, "convection": tuple(float(v) for v in b), "reaction": 0.0,
                    "rhs_constant": 1.0})
    return out


def autotune(cd, level=-1, iters=10, log=None, mixed=False):
    """Pick the launch shape of the SpMV and Jacobi kernels for this matrix
    (pmcd_autotune): every CSR variant at 8, 16 and 32 lanes per row and the
    sliced ELL kernels are timed with HIP events and the fastest is set; when
    the sliced ELL kernels win on `level`, the coarser levels on which they
    lose keep the CSR kernels (set_sell_min_rows). With `mixed` the Jacobi
    sweep is timed in FP32, the precision the V-cycle runs in (the system must
    have solved once, which makes its FP32 copies). Returns the timings
    {kernel: {shape: ms}}."""
    names = ("plain", "stream", "unrolled", "unrollednt")
    ms = np.zeros(26)
    chosen = (ctypes.c_int * 5)()
    _check(lib().pmcd_autotune(cd._s, int(level), int(iters), int(bool(mixed)), _ptr(ms, ctypes.c_double), chosen))
    out = {}
    for k, kernel in enumerate(("spmv", "jacobi")):
        res = {"%s%d" % (names[c // 3], (8, 16, 32)[c % 3]): float(ms[13 * k + c]) for c in range(12)}
        res["sell16"] = float(ms[13 * k + 12])
        out[kernel] = res
        if log:
            v, w = chosen[2 * k], chosen[2 * k + 1]
            log("autotune %s: %s -> %s%d" % (kernel, {a: round(b, 4) for a, b in res.items()},
                                             "sell" if v == 4 else names[v], w))
    if log and chosen[4]:
        log("autotune: sliced ELL on levels with at least %d active rows" % chosen[4])
    return out


def set_sell_window(slices):
    """Rows of a sliced ELL length-sorting window, in slices of 64 rows; read
    when a level's layout is first built."""
    _check(lib().pmcd_set_tuning(10, int(slices)))


def get_sell_window():
    v = ctypes.c_int(0)
    _check(lib().pmcd_get_tuning(10, ctypes.byref(v)))
    return v.value


def set_sell_min_rows(n):
    """Under variant 4, levels with fewer active rows keep the unrolled CSR kernels."""
    _check(lib().pmcd_set_tuning(11, int(n)))


def get_sell_min_rows():
    v = ctypes.c_int(0)
    _check(lib().pmcd_get_tuning(11, ctypes.byref(v)))
    return v.value


def set_sell_mode(mode):
    """Sliced ELL launch mode: bit 0 XCD-contiguous slices, bit 1 row-ordered
    output through LDS (layouts with 4-slice windows), bit 2 non-temporal
    matrix loads, bit 3 matrix loads pipelined one step ahead, bit 4 each
    block's x entries gathered once into LDS through its column footprint
    (with bit 1). Same bits in every mode."""
    _check(lib().pmcd_set_tuning(13, int(mode)))


def get_sell_mode():
    v = ctypes.c_int(0)
    _check(lib().pmcd_get_tuning(13, ctypes.byref(v)))
    return v.value


def set_tuning_value(key, value):
    """Raw tuning knob (pmcd_set_tuning key and value)."""
    _check(lib().pmcd_set_tuning(int(key), int(value)))


def tuning_value(key):
    v = ctypes.c_int(0)
    _check(lib().pmcd_get_tuning(int(key), ctypes.byref(v)))
    return v.value


def set_fgmres_lookahead(factor):
    """FGMRES queues its next iteration before reading the current Hessenberg
    column while the residual exceeds `factor` times the stopping threshold
    (0: never). Results do not depend on it."""
    _check(lib().pmcd_set_tuning(15, int(factor)))


def get_fgmres_lookahead():
    v = ctypes.c_int(0)
    _check(lib().pmcd_get_tuning(15, ctypes.byref(v)))
    return v.value


def set_sweep_block_rows(n):
    """Exact SOR sweeps whose largest class has at most n rows run as one
    single-workgroup launch per sweep (0: one launch per class always)."""
    _check(lib().pmcd_set_tuning(12, int(n)))


def get_sweep_block_rows():
    v = ctypes.c_int(0)
    _check(lib().pmcd_get_tuning(12, ctypes.byref(v)))
    return v.value


def gpu_available():
    n = ctypes.c_int(0)
    try:
        if lib().pmcd_device_count(ctypes.byref(n)) != 0:
            return False
    except RuntimeError:
        return False
    return n.value > 0
