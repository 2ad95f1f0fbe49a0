"""Time-dependent convection-diffusion on the device against the oracle and
against the reference's tcd3d solver fixtures.

Tolerances: system matrix and scheme rhs within 1e-12 relative per row (as the
stationary assembly), solution after each step within 1e-10 relative, the
fixtures' L2(0,t;L2) and L2(0,t;H1) errors within the reference's own bounds.
"""
import numpy as np
import pytest

import parmoon2022_amd as pm
from test_gpu_parity import row_rel_err
from test_oracle import TCD2D_GOLDEN, TCD2D_STEPS, TCD3D_CASES, TCD3D_SOLVERS, tcd2d_db, tcd3d_db

pytestmark = pytest.mark.gpu


def run_pair(oracle_lib, db, check_system=False, n_steps=None):
    """Steps the device problem and the oracle side by side; compares the
    solution after every step (and the system matrix and scheme rhs with
    check_system). Returns the device problem and the per-step worst
    relative solution difference."""
    cd = pm.TimeConvectionDiffusion(db)
    cd.assemble_initial_time()
    o = oracle_lib.Oracle(pm.problem_desc(db))
    t = pm.tcd_desc(db)
    o.tcd_init(t.scheme, t.bdf_first, t.tau, t.t0)
    np.testing.assert_allclose(cd.solution(), o.solution(), rtol=0, atol=1e-14)
    sd = pm.solver_desc(db)
    rp, _ = cd.pattern()
    diffs = []
    k = 0
    while cd.current_time < db["time_end"] - 1e-10 and (n_steps is None or k < n_steps):
        cd.step()
        cd.output()
        _, to = o.tcd_step(sd)
        o.tcd_output()
        assert abs(cd.current_time - to) < 1e-14
        if check_system:
            assert row_rel_err(rp, cd.matrix_values(), o.matrix_values()) < 1e-12
            ro = o.rhs()
            assert np.max(np.abs(cd.rhs() - ro)) <= 1e-12 * max(np.max(np.abs(ro)), 1.0)
        u, uo = cd.solution(), o.solution()
        diffs.append(np.max(np.abs(u - uo)) / max(np.max(np.abs(uo)), 1e-300))
        k += 1
    return cd, o, diffs


@pytest.mark.parametrize("solver", list(TCD3D_SOLVERS))
@pytest.mark.parametrize("geo", ["Default_UnitCube_Hexa", "Default_UnitCube_Tetra"])
@pytest.mark.parametrize("example,order,scheme", TCD3D_CASES)
def test_tcd3d_solvers_fixture_gpu(oracle_lib, geo, example, order, scheme, solver):
    """test/tcd3d_solvers_test.c++:170-218 on the device: ten steps of 0.1,
    L2(0,t;L2) and L2(0,t;H1) stay zero (direct 1e-13, iterative 1e-10)."""
    print("tcd3d fixture", geo, example, order, scheme, solver, flush=True)
    db = tcd3d_db(geo, example, order, scheme, solver)
    tol = TCD3D_SOLVERS[solver][1]
    cd = pm.TimeConvectionDiffusion(db)
    cd.assemble_initial_time()
    steps = 0
    while cd.current_time < db["time_end"] - 1e-10:
        cd.step()
        cd.output()
        steps += 1
        e = cd.get_errors()
        assert e[0] < tol and e[1] < tol, (cd.current_time, e)
    assert steps == 10 and abs(cd.current_time - 1.0) < 1e-12


@pytest.mark.parametrize("scheme,first,disc", [
    ("backward_euler", "backward_euler", "supg"),
    ("crank_nicolson", "backward_euler", "supg"),
    ("bdf_two", "backward_euler", "galerkin"),
    ("bdf_two", "crank_nicolson", "supg"),
])
@pytest.mark.parametrize("geo,example,order", [
    ("Default_UnitCube_Hexa", "quadratic_space_time", 2),
    ("Default_UnitCube_Tetra", "linear_space_time", 1),
])
def test_tcd_steps_match_oracle(oracle_lib, geo, example, order, scheme, first, disc):
    """System matrix factor A + M, the scheme's rhs and the solution after
    every step equal the oracle's (direct solver)."""
    print("tcd vs oracle", geo, example, order, scheme, first, disc, flush=True)
    db = tcd3d_db(geo, example, order, scheme, space_discretization_type=disc, bdf_two_first_step=first,
                  time_end=0.5)
    _, _, diffs = run_pair(oracle_lib, db, check_system=True)
    assert len(diffs) == 5 and max(diffs) < 1e-10, diffs


@pytest.mark.parametrize("scheme", ["crank_nicolson", "bdf_two"])
def test_tcd_multigrid_three_levels_matches_oracle(oracle_lib, scheme):
    """Three levels (every level carries its own A, M and system matrix),
    SUPG, FGMRES with W-cycle multigrid: the device steps follow the oracle's.
    The solver stops at a reduction of 1e-14, so both sides hold the discrete
    solution to rounding."""
    print("tcd multigrid", scheme, flush=True)
    db = tcd3d_db("Default_UnitCube_Hexa", "quadratic_space_time", 2, scheme, "multigrid",
                  refinement_n_initial_steps=3, multigrid_n_levels=3, space_discretization_type="supg",
                  time_step_length=0.05, residual_tolerance=0.0, residual_reduction=1e-14)
    cd, o, diffs = run_pair(oracle_lib, db, n_steps=4)
    assert max(diffs) < 1e-10, diffs
    for level in range(3):
        rp, _ = cd.pattern(level)
        assert row_rel_err(rp, cd.matrix_values(level), o.matrix_values(level)) < 1e-12, level


def test_tcd_crank_nicolson_exact_at_larger_size():
    """32^3 Q2 hexahedra (274,625 DoF), SUPG, Crank-Nicolson with FGMRES and
    4-level multigrid: the quadratic example lies in the space and the scheme
    is exact for it, so after each step the L2 error is at the level of the
    solver's reduction (1e-12), far below the size of u = t^2 (x^2 + y z)."""
    print("tcd 32^3 exactness", flush=True)
    db = tcd3d_db("Default_UnitCube_Hexa", "quadratic_space_time", 2, "crank_nicolson", "multigrid",
                  refinement_n_initial_steps=5, multigrid_n_levels=4, space_discretization_type="supg",
                  time_step_length=0.01, residual_tolerance=0.0, residual_reduction=1e-12,
                  multigrid_smoother="jacobi", jacobi_damp=0.8, multigrid_cycle_type="V",
                  max_n_iterations=200)
    cd = pm.TimeConvectionDiffusion(db)
    cd.assemble_initial_time()
    for _ in range(3):
        st = cd.step()
        e = cd.output()
        t = cd.current_time
        assert st.converged, (t, st.iterations)
        assert e[0] < 1e-9 * t * t and e[1] < 1e-8 * t * t, (t, e)


@pytest.mark.parametrize("disc", ["galerkin", "supg"])
def test_tcd2d_golden_gpu(oracle_lib, disc):
    """test/tcd2d_test.c++:47-68 on the device: L2 and H1-semi errors after
    steps 0-3 and 18-20 of Crank-Nicolson equal the reference's within 1e-10,
    and every step's solution equals the oracle's."""
    print("tcd2d golden", disc, flush=True)
    db = tcd2d_db(disc)
    cd = pm.TimeConvectionDiffusion(db)
    cd.assemble_initial_time()
    o = oracle_lib.Oracle(pm.problem_desc(db))
    t = pm.tcd_desc(db)
    o.tcd_init(t.scheme, t.bdf_first, t.tau, t.t0)
    sd = pm.solver_desc(db)
    got = {0: cd.output()}
    for m in range(1, 21):
        cd.step()
        got[m] = cd.output()
        o.tcd_step(sd)
        u, uo = cd.solution(), o.solution()
        assert np.max(np.abs(u - uo)) <= 1e-10 * np.max(np.abs(uo)), m
    l2, h1 = TCD2D_GOLDEN[disc]
    for i, m in enumerate(TCD2D_STEPS):
        assert abs(got[m][0] - l2[i]) < 1e-10 and abs(got[m][1] - h1[i]) < 1e-10, (m, got[m][:2], l2[i], h1[i])


def test_tcd_sliced_ell_jacobi_matches_oracle(oracle_lib):
    """With the sliced ELL Jacobi sweep on every level, the system matrix
    kernel writes the sliced ELL copy of each step's matrix itself: the
    device steps still follow the oracle's (W-cycle FGMRES, Jacobi
    smoothing)."""
    print("tcd sliced ell", flush=True)
    db = tcd3d_db("Default_UnitCube_Hexa", "quadratic_space_time", 2, "crank_nicolson", "multigrid",
                  refinement_n_initial_steps=3, multigrid_n_levels=3, space_discretization_type="supg",
                  time_step_length=0.05, residual_tolerance=0.0, residual_reduction=1e-14,
                  multigrid_smoother="jacobi", jacobi_damp=0.8)
    keep = pm.tuning("jacobi"), pm.get_sell_min_rows()
    try:
        pm.set_tuning("jacobi", 16, 4)
        pm.set_sell_min_rows(0)
        _, _, diffs = run_pair(oracle_lib, db, n_steps=4)
    finally:
        pm.set_tuning("jacobi", *keep[0])
        pm.set_sell_min_rows(keep[1])
    assert max(diffs) < 1e-10, diffs
