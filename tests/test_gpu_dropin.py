"""Drop-in entries on caller-owned data, through the C ABI: assembly into a
caller's FE space and structure (Assemble3D / Assemble2D,
include/AssembleRoutines/Assemble3D.h:28-34), the solver on a caller's
matrices (Solver<>, include/Solver/Solver.h:45-94) and the device-pointer
building blocks. The caller's arrays are the host set-up's own (DOF map,
cells, pattern), so the results must be the bits of pmcd_assemble and
pmcd_solve on the same problem."""
import ctypes
import math

import numpy as np
import pytest

import parmoon2022_amd as pm
from hip_buffers import DeviceArray

pytestmark = pytest.mark.gpu

S3 = 1 / math.sqrt(3)
SHAPE = {"TwoTriangles": "triangle", "TwoTriangles2": "triangle", "UnitSquare": "quadrangle",
         "Default_UnitCube_Tetra": "tetrahedron", "Default_UnitCube_Hexa": "hexahedron"}

CASES = [
    {"geo_file": "Default_UnitCube_Hexa", "ANSATZ_ORDER": 2, "refinement_n_initial_steps": 2, "example": "sine_3d",
     "space_discretization_type": "supg", "diffusion_coefficient": 1e-3, "convection": (S3, S3, S3),
     "INTERNAL_PROBLEM_LINEAR": 1},
    {"geo_file": "Default_UnitCube_Hexa", "ANSATZ_ORDER": 1, "refinement_n_initial_steps": 3, "example": "sine_3d",
     "space_discretization_type": "supg", "diffusion_coefficient": 1e-2, "convection": (0.3, -0.5, 0.81)},
    {"geo_file": "Default_UnitCube_Tetra", "ANSATZ_ORDER": 2, "refinement_n_initial_steps": 1,
     "example": "const_rhs", "space_discretization_type": "supg", "diffusion_coefficient": 5e-2,
     "convection": (0.3, -0.5, 0.81), "reaction": 0.4},
    {"geo_file": "Default_UnitCube_Tetra", "ANSATZ_ORDER": 1, "refinement_n_initial_steps": 2,
     "example": "test_p1"},
    {"geo_file": "TwoTriangles", "ANSATZ_ORDER": 2, "refinement_n_initial_steps": 3, "example": "sine_2d",
     "space_discretization_type": "supg", "diffusion_coefficient": 1e-2, "convection": (1.0, 0.5, 0)},
    {"geo_file": "UnitSquare", "ANSATZ_ORDER": 1, "refinement_n_initial_steps": 4, "example": "sine_2d",
     "diffusion_coefficient": 1.0},
]


def caller_space(db, level=-1):
    """The host set-up's arrays as a caller would hold them: cells' vertices,
    GlobalNumbers / BeginIndex, n_dof, n_active, RowPtr / KCol."""
    h = pm.HostSetup(db)
    info = h.info(level)
    dof = h.dofmap(level)
    nl = info["n_local"]
    begin = np.arange(info["n_cells"] + 1, dtype=np.int32) * nl
    rp, col = h.pattern(level)
    return h.cell_coords(level), dof.reshape(-1), begin, info["n_dof"], info["n_active"], rp, col


@pytest.mark.parametrize("db", CASES, ids=lambda d: "%s-p%d-%s" % (d["geo_file"], d["ANSATZ_ORDER"], d["example"]))
def test_assemble_into_caller_space_equals_system(db):
    cd = pm.ConvectionDiffusion(db)
    cd.assemble()
    a_ref, f_ref = cd.matrix_values(), cd.rhs()
    cd.close()
    verts, gn, bi, n, na, rp, col = caller_space(db)
    entries = np.zeros(len(col))
    rhs = np.zeros(n)
    pm.assemble_into(SHAPE[db["geo_file"]], db["ANSATZ_ORDER"], verts, gn, bi, n, na, rp, col, db, entries, rhs,
                     dirichlet_values=f_ref[na:])
    assert np.array_equal(entries, a_ref)
    assert np.array_equal(rhs, f_ref)
    # Assemble3D adds into the caller's arrays: a second assembly doubles the
    # active rows and rhs; Dirichlet diagonals and values are set, not added
    asm = pm.Assembler(SHAPE[db["geo_file"]], db["ANSATZ_ORDER"], verts, gn, bi, n, na, rp, col, db)
    asm.assemble(entries, rhs, dirichlet_values=f_ref[na:])
    act = rp[na]
    assert np.array_equal(entries[:act], a_ref[:act] + a_ref[:act])
    assert np.array_equal(entries[act:], a_ref[act:])
    assert np.array_equal(rhs[:na], f_ref[:na] + f_ref[:na]) and np.array_equal(rhs[na:], f_ref[na:])
    # device arrays of the caller
    d_ent = DeviceArray(np.zeros(len(col)))
    d_rhs = DeviceArray(np.zeros(n))
    d_dir = DeviceArray(f_ref[na:].copy())
    asm.assemble_device(d_ent.ptr, d_rhs.ptr, d_dir.ptr)
    assert np.array_equal(d_ent.get(), a_ref) and np.array_equal(d_rhs.get(), f_ref)
    for b in (d_ent, d_rhs, d_dir):
        b.close()
    asm.close()


def test_assemble_into_rejects_a_pattern_without_a_coupling():
    db = CASES[0]
    verts, gn, bi, n, na, rp, col = caller_space(db)
    # drop the last entry of row 0 (a coupling of two DOFs of one cell)
    col2 = np.delete(col, rp[1] - 1)
    rp2 = rp.copy()
    rp2[1:] -= 1
    with pytest.raises(RuntimeError, match="not in the structure"):
        pm.assemble_into(SHAPE[db["geo_file"]], 2, verts, gn, bi, n, na, rp2, col2, db, np.zeros(len(col2)),
                         np.zeros(n))


HANGING_CASES = [("TwoTriangles", 1), ("TwoTriangles", 2), ("UnitSquare", 1), ("UnitSquare", 2)]


def _host_hanging(h):
    n = ctypes.c_int()
    pm._check(pm.lib().pmcd_host_hanging(h._h, -1, None, None, ctypes.byref(n)))
    out = np.zeros(4 * max(n.value, 1), dtype=np.int32)
    w = np.zeros(3 * max(n.value, 1))
    pm._check(pm.lib().pmcd_host_hanging(h._h, -1, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                         w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(n)))
    return out[:4 * n.value].reshape(-1, 4), w[:3 * n.value].reshape(-1, 3)


@pytest.mark.parametrize("geo,order", HANGING_CASES)
def test_assemble_into_hanging_space_equals_system(geo, order):
    """A caller's 2D space with hanging DOFs (eight refinements by indicator,
    hanging vertices): its hanging nodes handed over in DOF order
    (HangingNodeArray, not the sorted order the coupling needs) and the
    columns of every row in a shuffled order. The coupled entries and rhs
    are the bits of pmcd_assemble; a second assembly adds into the
    non-hanging active rows and leaves the constraint rows and zero hanging
    rhs entries as they are."""
    # sine_2d: Dirichlet boundary only (the drop-in adds no Neumann integrals)
    db = {"geo_file": geo, "ANSATZ_ORDER": order, "refinement_n_initial_steps": 0, "example": "sine_2d",
          "space_discretization_type": "supg", "diffusion_coefficient": 1e-2, "convection": (1.0, 0.5, 0),
          "refine_by_indicator_steps": 8, "refine_indicator": (0.5, 1.0, 0.0, -0.5), "conforming_closure": False}
    cd = pm.ConvectionDiffusion(db)
    cd.assemble()
    a_ref, f_ref = cd.matrix_values(), cd.rhs()
    cd.close()
    h = pm.HostSetup(db)
    hg, w = _host_hanging(h)
    assert len(hg) > 0
    order_dof = np.argsort(hg[:, 0], kind="stable")
    hg, w = hg[order_dof], w[order_dof]
    verts, gn, bi, n, na, rp, col = caller_space(db)
    rng = np.random.default_rng(7)
    perm = np.concatenate([rp[r] + rng.permutation(rp[r + 1] - rp[r]) for r in range(n)])
    col_c = col[perm]
    entries = np.zeros(len(col))
    rhs = np.zeros(n)
    pm.assemble_into(SHAPE[geo], order, verts, gn, bi, n, na, rp, col_c, db, entries, rhs,
                     dirichlet_values=f_ref[na:], hanging=(hg, w))
    assert np.array_equal(entries, a_ref[perm])
    assert np.array_equal(rhs, f_ref)
    first = na - len(hg)
    asm = pm.Assembler(SHAPE[geo], order, verts, gn, bi, n, na, rp, col_c, db, hanging=(hg, w))
    asm.assemble(entries, rhs, dirichlet_values=f_ref[na:])
    nh0, nh1 = rp[first], rp[na]
    twice = a_ref[perm].copy()
    twice[:nh0] += a_ref[perm][:nh0]
    assert np.array_equal(entries, twice)
    assert np.array_equal(rhs[:first], f_ref[:first] + f_ref[:first])
    assert np.array_equal(rhs[first:], f_ref[first:]) and not np.any(rhs[first:na])
    # device arrays of the caller
    d_ent = DeviceArray(np.zeros(len(col)))
    d_rhs = DeviceArray(np.full(n, 3.0))  # hanging entries are set, not added
    d_dir = DeviceArray(f_ref[na:].copy())
    asm.assemble_device(d_ent.ptr, d_rhs.ptr, d_dir.ptr)
    f_dev = f_ref.copy()
    f_dev[:first] += 3.0
    assert np.array_equal(d_ent.get(), a_ref[perm]) and np.array_equal(d_rhs.get(), f_dev)
    for b in (d_ent, d_rhs, d_dir):
        b.close()
    asm.close()
    assert nh1 > nh0


def test_assemble_into_rejects_bad_hanging_lists():
    db = {"geo_file": "UnitSquare", "ANSATZ_ORDER": 1, "refinement_n_initial_steps": 0,
          "example": "sine_laplace_2d", "refine_by_indicator_steps": 4, "refine_indicator": (0.5, 1.0, 0.0, -0.5),
          "conforming_closure": False}
    hg, w = _host_hanging(pm.HostSetup(db))
    verts, gn, bi, n, na, rp, col = caller_space(db)
    args = ("quadrangle", 1, verts, gn, bi, n, na, rp, col, db, np.zeros(len(col)), np.zeros(n))
    bad = hg.copy()
    bad[0, 0] = 0  # not one of the last n_hanging active DOFs
    with pytest.raises(RuntimeError, match="hanging DOF 0 not in"):
        pm.assemble_into(*args, hanging=(bad, w))
    bad = hg.copy()
    bad[0, 1:] = -1
    with pytest.raises(RuntimeError, match="without coupling DOFs"):
        pm.assemble_into(*args, hanging=(bad, w))
    # a structure without one of a hanging row's columns in the row of an
    # active coupling DOF (a column that no cell of that DOF couples)
    first = na - len(hg)
    cells = gn.reshape(len(bi) - 1, -1)
    drop = []
    for i in range(len(hg)):
        hrow = set(col[rp[hg[i, 0]]:rp[hg[i, 0] + 1]].tolist())
        for c in hg[i, 1:]:
            if c < 0 or c >= na:
                continue
            near = set(cells[np.any(cells == c, axis=1)].reshape(-1).tolist())
            drop = [q for q in range(rp[c], rp[c + 1]) if col[q] in hrow and col[q] not in near]
            if drop:
                break
        if drop:
            break
    assert drop and first < na
    col2 = np.delete(col, drop[0])
    rp2 = rp.copy()
    rp2[c + 1:] -= 1
    with pytest.raises(RuntimeError, match="could not be added to row %d" % c):
        pm.assemble_into("quadrangle", 1, verts, gn, bi, n, na, rp2, col2, db, np.zeros(len(col2)), np.zeros(n),
                         hanging=(hg, w))


def _mg_db(solver="fgmres", levels=3, mixed=0, smoother="jacobi"):
    return {"geo_file": "Default_UnitCube_Hexa", "ANSATZ_ORDER": 2, "refinement_n_initial_steps": 3,
            "example": "sine_3d", "space_discretization_type": "supg", "diffusion_coefficient": 1e-3,
            "convection": (S3, S3, S3), "INTERNAL_PROBLEM_LINEAR": 1, "solver_type": "iterative",
            "iterative_solver_type": solver, "preconditioner": "multigrid", "multigrid_n_levels": levels,
            "multigrid_n_pre_smooth": 2, "multigrid_n_post_smooth": 2, "jacobi_damp": 0.5,
            "multigrid_correction_damp_factor": 1.0, "gmres_restart": 30, "residual_reduction": 1e-10,
            "residual_tolerance": 0.0, "max_n_iterations": 300, "multigrid_mixed_precision": mixed,
            "multigrid_smoother": smoother, "sor_omega": 1.0}


@pytest.mark.parametrize("mixed,smoother", [(0, "jacobi"), (1, "jacobi"), (0, "ssor")])
def test_solver_on_caller_matrices_equals_system_solve(mixed, smoother):
    """Solver(db) + update_matrix(levels) + solve(b, x) on the system's own
    level matrices and prolongations: the iterations and bits of
    pmcd_solve; a second update with the same values changes nothing."""
    db = _mg_db(mixed=mixed, smoother=smoother)
    cd = pm.ConvectionDiffusion(db)
    cd.assemble()
    levels = []
    for l in range(cd.info()["n_levels"]):
        rp, col = cd.pattern(l)
        lv = {"n_active": cd.info(l)["n_active"], "row_ptr": rp, "col": col, "entries": cd.matrix_values(l)}
        if l > 0:
            prp, pcol, pval, _ = cd.transfer(l)
            lv.update({"p_row_ptr": prp, "p_col": pcol, "p_entries": pval})
        levels.append(lv)
    b = cd.rhs()
    x0 = cd.solution()  # zero with the Dirichlet values, as after assemble()
    st = cd.solve()
    u_ref = cd.solution()
    cd.close()
    s = pm.Solver(db)
    assert s.is_using_multigrid()
    s.update_matrix(levels)
    for _ in range(2):
        x = x0.copy()
        st2 = s.solve(b, x)
        assert st2.iterations == st.iterations and np.array_equal(x, u_ref)
        s.update_matrix(levels)
    s.close()


def test_solver_direct_on_caller_matrix():
    """The direct solver on one caller-owned level equals pmcd_solve's."""
    db = {"geo_file": "Default_UnitCube_Tetra", "ANSATZ_ORDER": 2, "refinement_n_initial_steps": 1,
          "example": "test_p2", "solver_type": "direct"}
    cd = pm.ConvectionDiffusion(db)
    cd.assemble()
    rp, col = cd.pattern()
    lv = [{"n_active": cd.info()["n_active"], "row_ptr": rp, "col": col, "entries": cd.matrix_values()}]
    b, x0 = cd.rhs(), cd.solution()
    cd.solve()
    u_ref = cd.solution()
    cd.close()
    s = pm.Solver(db)
    s.update_matrix(lv)
    x = x0.copy()
    s.solve(b, x)
    assert np.array_equal(x, u_ref)
    s.close()


def test_device_building_blocks(oracle_lib):
    """axpy, norm, Jacobi step, prolongation and restriction on caller device
    buffers against host arithmetic on the same data."""
    db = _mg_db(levels=2)
    cd = pm.ConvectionDiffusion(db)
    cd.assemble()
    n, na = cd.info()["n_dof"], cd.info()["n_active"]
    nc, nca = cd.info(0)["n_dof"], cd.info(0)["n_active"]
    rp, col = cd.pattern()
    val = cd.matrix_values()
    prp, pcol, pval, ncoarse = cd.transfer()
    cd.close()
    assert ncoarse == nc
    L = pm.lib()
    rng = np.random.default_rng(9)
    x, y, bvec = rng.standard_normal(n), rng.standard_normal(n), rng.standard_normal(n)
    xc = rng.standard_normal(nc)
    diag = np.array([val[rp[r]:rp[r + 1]][col[rp[r]:rp[r + 1]] == r][0] for r in range(n)])
    d = {k: DeviceArray(v) for k, v in dict(x=x, y=y, b=bvec, diag=diag, rp=rp.astype(np.int32),
                                              col=col.astype(np.int32), val=val, out=np.zeros(n), xc=xc,
                                              yc=np.zeros(nc), prp=prp, pcol=pcol, pval=pval).items()}
    vp = ctypes.c_void_p
    L.pmcd_axpy_device.argtypes = [ctypes.c_int, ctypes.c_double, vp, vp, vp]
    assert L.pmcd_axpy_device(n, 0.3, d["x"].ptr, d["y"].ptr, None) == 0
    assert np.array_equal(d["y"].get(), y + 0.3 * x)
    L.pmcd_norm_device.argtypes = [ctypes.c_int, vp, ctypes.POINTER(ctypes.c_double), vp]
    nrm = ctypes.c_double()
    assert L.pmcd_norm_device(n, d["x"].ptr, ctypes.byref(nrm), None) == 0
    assert abs(nrm.value - np.linalg.norm(x)) <= 1e-13 * np.linalg.norm(x)
    L.pmcd_jacobi_device.argtypes = [ctypes.c_int, ctypes.c_int] + [vp] * 7 + [ctypes.c_double, vp]
    assert L.pmcd_jacobi_device(n, na, d["rp"].ptr, d["col"].ptr, d["val"].ptr, d["diag"].ptr, d["b"].ptr,
                                d["x"].ptr, d["out"].ptr, 0.7, None) == 0
    r = oracle_lib.spmv_csr(n, na, rp, col, val, x, bvec, -1.0)
    ref = x + 0.7 * (r / diag)
    absr = oracle_lib.spmv_csr(n, na, rp, col, np.abs(val), np.abs(x), np.abs(bvec), 1.0)
    assert np.all(np.abs(d["out"].get() - ref) <= 0.7 * 4 * 125 * np.finfo(float).eps * absr / np.abs(diag) + 1e-300)
    # prolongation: fine += 0.5 P coarse on active rows
    L.pmcd_prolongate_device.argtypes = [ctypes.c_int, ctypes.c_int] + [vp] * 5 + [ctypes.c_double, vp]
    f0 = d["y"].get()
    assert L.pmcd_prolongate_device(n, na, d["prp"].ptr, d["pcol"].ptr, d["pval"].ptr, d["xc"].ptr, d["y"].ptr,
                                    0.5, None) == 0
    pf = np.array([np.dot(pval[prp[i]:prp[i + 1]], xc[pcol[prp[i]:prp[i + 1]]]) for i in range(n)])
    exp = f0.copy()
    exp[:na] += 0.5 * pf[:na]
    assert np.allclose(d["y"].get(), exp, rtol=1e-14, atol=1e-14)
    # restriction with R = P^T: coarse = R fine on active coarse rows, 0 on the tail
    import scipy.sparse as sp
    P = sp.csr_matrix((pval, pcol, prp), shape=(n, nc))
    R = P.T.tocsr()
    R.sort_indices()
    for k, v in dict(rrp=R.indptr, rcol=R.indices, rval=R.data).items():
        d[k] = DeviceArray(np.asarray(v, dtype=np.int32 if k != "rval" else np.float64))
    L.pmcd_restrict_device.argtypes = [ctypes.c_int, ctypes.c_int] + [vp] * 6
    assert L.pmcd_restrict_device(nc, nca, d["rrp"].ptr, d["rcol"].ptr, d["rval"].ptr, d["x"].ptr, d["yc"].ptr,
                                  None) == 0
    rc = R @ x
    rc[nca:] = 0
    assert np.allclose(d["yc"].get(), rc, rtol=1e-13, atol=1e-13)
    for b in d.values():
        b.close()
