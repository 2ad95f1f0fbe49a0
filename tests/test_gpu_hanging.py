"""Meshes refined by indicator with hanging vertices (P1/P2 on TwoTriangles,
Q1/Q2 on UnitSquare, eight steps of 0.5 x^2 + y - 0.5) on the device:
assembly with the hanging-DOF coupling against the oracle (row-relative
1e-12, the constraint rows and hanging rhs entries exact), the reference's
golden errors of test/adaptive_hanging_test.c++ within its 1e-11 for the
direct and the FGMRES solvers, the solution against the oracle's, and served
requests; the conforming closure (P1/P2 on TwoTriangles) against the same
test's conforming goldens and the oracle's matrix.
"""
import numpy as np
import pytest

import parmoon2022_amd as pm
from test_oracle import CONFORMING_GOLDEN, HANGING_GOLDEN, hanging_db, hanging_constraint_residual
from test_gpu_parity import row_rel_err

pytestmark = pytest.mark.gpu

FGMRES = {"solver_type": "iterative", "iterative_solver_type": "fgmres", "max_n_iterations": 5000,
          "residual_tolerance": 1e-13, "gmres_restart": 50}


def _hanging(cd):
    import ctypes
    n = ctypes.c_int()
    pm._check(pm.lib().pmcd_get_hanging(cd._s, -1, None, None, ctypes.byref(n)))
    out = np.zeros(4 * max(n.value, 1), dtype=np.int32)
    w = np.zeros(3 * max(n.value, 1))
    pm._check(pm.lib().pmcd_get_hanging(cd._s, -1, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                        w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(n)))
    return out[:4 * n.value].reshape(-1, 4), w[:3 * n.value].reshape(-1, 3)


@pytest.mark.parametrize("variant", [0, 2], ids=["vector", "tensor"])
@pytest.mark.parametrize("geo,order", list(HANGING_GOLDEN))
def test_hanging_assembly_matches_oracle(oracle_lib, geo, order, variant):
    db = hanging_db(geo, order)
    print("hanging assembly", geo, order, "variant", variant, flush=True)
    old = pm.assembly_variant()
    pm.set_assembly_variant(variant)
    try:
        cd = pm.ConvectionDiffusion(db)
        cd.assemble()
    finally:
        pm.set_assembly_variant(old)
    o = oracle_lib.Oracle(pm.problem_desc(db))
    o.assemble()
    assert np.array_equal(cd.dofmap(), o.dofmap())
    hg, w = _hanging(cd)
    ohg, ow = o.hanging()
    assert np.array_equal(hg, ohg) and np.array_equal(w, ow)
    rp, col = cd.pattern()
    a, b = cd.matrix_values(), o.matrix_values()
    err = row_rel_err(rp, a, b)
    print("matrix: row-relative error %.3e, bit-identical fraction %.4f" % (err, np.mean(a == b)), flush=True)
    assert err <= 1e-12
    fa, fb = cd.rhs(), o.rhs()
    assert np.max(np.abs(fa - fb)) <= 1e-12 * np.max(np.abs(fb))
    # hanging rows: u_h - sum_k w_k u_{c_k} = 0, rhs 0 (exact)
    for (h, *cpl), wt in zip(hg, w):
        row = {int(c): v for c, v in zip(col[rp[h]:rp[h + 1]], a[rp[h]:rp[h + 1]])}
        want = {c: 0.0 for c in row}
        want[h] = 1.0
        for c, x in zip(cpl, wt):
            if c >= 0:
                want[c] = -x
        assert row == want, h
    assert np.all(fa[hg[:, 0]] == 0.0)
    cd.close()


@pytest.mark.parametrize("solver", ["direct", "fgmres_ssor", "fgmres_jacobi"])
@pytest.mark.parametrize("geo,order", list(HANGING_GOLDEN))
def test_hanging_golden_errors_on_gpu(oracle_lib, geo, order, solver):
    extra = {}
    if solver != "direct":
        extra = dict(FGMRES, preconditioner=solver.split("_")[1])
    db = hanging_db(geo, order, **extra)
    print("hanging golden", geo, order, solver, flush=True)
    cd = pm.ConvectionDiffusion(db)
    cd.assemble()
    st = cd.solve()
    assert st.converged
    e = cd.output()
    got = [e[0], e[1], e[2], e[4]]
    gold = HANGING_GOLDEN[(geo, order)]
    assert np.allclose(got, gold, rtol=0, atol=1e-11), (got, gold)
    u = cd.solution()
    hg, w = _hanging(cd)
    assert hanging_constraint_residual(u, hg, w) < 1e-12
    if solver == "direct" and order == 1:
        o = oracle_lib.Oracle(pm.problem_desc(db))
        o.assemble()
        o.solve(pm.solver_desc(db))
        ou = o.solution()
        assert np.max(np.abs(u - ou)) <= 1e-12 * np.max(np.abs(ou))
    cd.close()


def test_hanging_served_requests():
    """Requests on an adaptive mesh through the request server: the same
    solution as a system of their own; the second request reuses the
    resident system (same indicator, same callback)."""
    import parmoon2022_amd.serve as sv
    db = hanging_db("UnitSquare", **dict(FGMRES, preconditioner="ssor"))
    cd = pm.ConvectionDiffusion(db)
    cd.assemble()
    cd.solve()
    u = cd.solution()
    cd.close()
    srv = sv.Server(max_slots=1, keep_solution=True)
    reqs = [srv.submit(db), srv.submit(db)]
    for r in reqs:
        assert r.done.wait(120)
    info = srv.info()
    srv.close()
    assert info["systems_created"] == 1
    for r in reqs:
        assert r.error is None
        assert np.array_equal(r.solution, u)
    assert reqs[1].reused


@pytest.mark.parametrize("solver", ["direct", "fgmres_ssor"])
@pytest.mark.parametrize("order", [1, 2])
def test_conforming_closure_on_gpu(oracle_lib, order, solver):
    extra = dict(FGMRES, preconditioner="ssor") if solver != "direct" else {}
    db = hanging_db("TwoTriangles", order, conforming_closure=True, **extra)
    print("conforming closure", order, solver, flush=True)
    cd = pm.ConvectionDiffusion(db)
    cd.assemble()
    o = oracle_lib.Oracle(pm.problem_desc(db))
    o.assemble()
    assert np.array_equal(cd.dofmap(), o.dofmap())
    rp, col = cd.pattern()
    assert row_rel_err(rp, cd.matrix_values(), o.matrix_values()) <= 1e-12
    hg, _ = _hanging(cd)
    assert len(hg) == 0
    st = cd.solve()
    assert st.converged
    e = cd.output()
    got = [e[0], e[1], e[2], e[4]]
    assert np.allclose(got, CONFORMING_GOLDEN[order], rtol=0, atol=1e-11), (got, CONFORMING_GOLDEN[order])
    cd.close()


@pytest.mark.parametrize("flag", [0, 1, 2])
@pytest.mark.parametrize("geo,order", [("TwoTriangles", 1), ("UnitSquare", 2)])
def test_hanging_sor_sweep_equals_serial(oracle_lib, geo, order, flag):
    """The device SOR sweep on a coupled matrix (its pattern is not
    symmetric) equals the oracle's serial sweep bit for bit, and an
    SSOR-preconditioned FGMRES solve repeats bit for bit."""
    db = hanging_db(geo, order, **dict(FGMRES, preconditioner="ssor", sor_omega=1.0))
    print("hanging sor", geo, order, flag, flush=True)
    pm.set_sor_exact_limit(100000)  # read when the sweep order is first built
    try:
        cd = pm.ConvectionDiffusion(db)
        cd.assemble()
        rng = np.random.default_rng(flag)
        n = cd.info()["n_dof"]
        b, x0 = rng.standard_normal(n), rng.standard_normal(n)
        got, exact = cd.sor_sweep(b, x0.copy(), omega=1.5, flag=flag)
    finally:
        pm.set_sor_exact_limit(24)
    assert exact
    o = oracle_lib.Oracle(pm.problem_desc(db))
    o.assemble()
    o.set_matrix(cd.matrix_values())  # the same (device-assembled) matrix
    assert np.array_equal(got, o.sor_sweep(b, x0, 1.5, flag))
    st1 = cd.solve()
    u1 = cd.solution()
    cd.assemble()
    st2 = cd.solve()
    assert st1.converged and st1.iterations == st2.iterations and np.array_equal(u1, cd.solution())
    cd.close()
