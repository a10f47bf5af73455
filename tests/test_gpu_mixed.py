"""Mixed precision (FP32 inner PCG on MFMA, FP64 refinement) against the
FP64 oracle, the golden vectors and the library's own FP64 solve."""
import os

import numpy as np
import pytest

import oracle_api as O
from platepcg import Context, Plan, PlatePCGError, _lib
from platepcg import problem as P

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
E, NU = 210e9, 0.3
MIXED = _lib.PRECISION_MIXED


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def check_sweeps(st):
    n = min(st["sweeps"], _lib.MAX_SWEEPS)
    if st["sweeps"] <= _lib.MAX_SWEEPS:
        assert sum(st["sweepIterations"][:n]) == st["iterations"]
    # the last recorded sweep residual is the reported true residual
    if st["sweeps"] <= _lib.MAX_SWEEPS:
        assert st["sweepResidual"][n - 1] == pytest.approx(st["trueRelResidual"], rel=1e-12)


def test_mixed_c1_matches_fp64_oracle(ctx):
    w = P.workload("c1")
    plan = Plan(ctx, w["mesh"], w["E"], w["nu"], w["t"], w["mask"])
    op = O.OracleProblem(w["mesh"], w["E"], w["nu"], w["t"], w["mask"])
    F = plan.pressure_load(w["q0"])
    x64, x32, st = plan.solve(F, tolerance=1e-8, max_iterations=20000, precision=MIXED)
    assert st["converged"] and not st["breakdown"], st
    assert st["sweeps"] >= 2
    check_sweeps(st)
    Fd = F.astype(np.float64)
    assert st["trueRelResidual"] <= 1e-8
    assert op.true_rel_residual(x64, Fd) <= 1e-8
    xo, so = op.pcg(Fd, tol=1e-10)
    assert O.max_nodal_w_error(x64, xo) < 1e-6


def test_mixed_c2_matches_golden(ctx):
    g = np.load(os.path.join(ROOT, "tests", "golden", "c2_142.npz"))
    w = P.workload("c2")
    plan = Plan(ctx, w["mesh"], w["E"], w["nu"], w["t"], w["mask"])
    x64, x32, st = plan.solve(g["F"], tolerance=1e-8, max_iterations=40000, precision=MIXED)
    assert st["converged"], st
    check_sweeps(st)
    assert st["trueRelResidual"] <= 1e-8
    assert O.max_nodal_w_error(x64, g["x64"]) < 1e-6
    assert O.compare_max_w(x32, g["x_js"]) < 1e-4
    op = O.OracleProblem(w["mesh"], w["E"], w["nu"], w["t"], w["mask"])
    assert op.true_rel_residual(x64, g["F"].astype(np.float64)) <= 1e-8


def test_mixed_first_sweep_follows_oracle_pcg(ctx):
    """With one sweep the inner FP32 recurrence is PCG on the shared element
    matrix: its first iterates agree with the FP64 oracle's to FP32 accuracy."""
    w = P.workload("c1")
    plan = Plan(ctx, w["mesh"], w["E"], w["nu"], w["t"], w["mask"])
    op = O.OracleProblem(w["mesh"], w["E"], w["nu"], w["t"], w["mask"])
    F = plan.pressure_load(w["q0"])
    for k in (1, 5, 20):
        x64, _, st = plan.solve(F, tolerance=1e-8, max_iterations=k, precision=MIXED,
                                inner_tolerance=1e-12, max_refinements=1)
        assert st["iterations"] == k and st["sweeps"] == 1
        xo, _ = op.pcg(F.astype(np.float64), max_iter=k)
        assert rel(x64, xo) < 2e-4, k


def test_mixed_variable_thickness_matches_fp64(ctx):
    """Per-element thickness: the bending / shear MFMA split (t^3 and t
    scaling) against the FP64 solve and the oracle."""
    n = 48
    mesh = P.rect_mesh(n)
    mask = P.constrained_mask(mesh, [{"type": "pinned", "location": "all_edges"}])
    rng = np.random.default_rng(7)
    ii, jj = np.meshgrid(np.arange(n), np.arange(n))
    field = np.sin(ii / 7.0 + rng.uniform(0, 6)) * np.cos(jj / 5.0 + rng.uniform(0, 6))
    et = (0.08 * np.exp(0.4 * field)).astype(np.float32).ravel()
    plan = Plan(ctx, mesh, E, NU, 0.08, mask, element_t=et)
    op = O.OracleProblem(mesh, E, NU, 0.08, mask, element_t=et)
    F = plan.pressure_load(1e4, np.array([[4e4, 4.0, 6.5, 1.0]]))
    x64m, _, sm = plan.solve(F, tolerance=1e-8, max_iterations=20000, precision=MIXED)
    x64d, _, sd = plan.solve(F, tolerance=1e-8, max_iterations=20000)
    assert sm["converged"] and sd["converged"]
    assert op.true_rel_residual(x64m, F.astype(np.float64)) <= 1e-8
    assert O.max_nodal_w_error(x64m, x64d) < 1e-6


def test_mixed_reliable_update_converges(ctx):
    """Keeping the search direction across the FP64 residual replacement
    reaches the same answer."""
    w = P.workload("c1")
    plan = Plan(ctx, w["mesh"], w["E"], w["nu"], w["t"], w["mask"])
    F = plan.pressure_load(w["q0"])
    xr, _, ri = plan.solve(F, tolerance=1e-8, max_iterations=40000, precision=MIXED)
    xu, _, ru = plan.solve(F, tolerance=1e-8, max_iterations=40000, precision=MIXED,
                           inner_tolerance=0.1, flags=_lib.FLAG_RELIABLE_UPDATE)
    assert ru["converged"] and ru["trueRelResidual"] <= 1e-8, ru
    check_sweeps(ru)
    assert O.max_nodal_w_error(xu, xr) < 1e-6


def test_mixed_edge_cases(ctx):
    w = P.workload("c1")
    plan = Plan(ctx, w["mesh"], w["E"], w["nu"], w["t"], w["mask"])
    x64, _, st = plan.solve(np.zeros(plan.n_dof, np.float32), precision=MIXED)
    assert st["converged"] and st["iterations"] == 0 and not x64.any()
    # every DOF constrained: identity system, one sweep of one iteration
    mask = np.full(w["mesh"]["nodeCount"], 7, np.uint8)
    p2 = Plan(ctx, w["mesh"], E, NU, 0.1, mask)
    F = np.random.default_rng(5).standard_normal(p2.n_dof).astype(np.float32)
    x64, _, st = p2.solve(F, precision=MIXED)
    assert st["converged"] and rel(x64, F.astype(np.float64)) < 1e-7
    # a graded grid has no shared element matrix
    mesh = P.rect_mesh(16)
    nodes = mesh["nodes"].reshape(-1, 2).copy()
    xs = np.unique(nodes[:, 0])
    graded = {float(a): float(10.0 * (a / 10.0) ** 1.5) for a in xs}
    nodes[:, 0] = [graded[float(a)] for a in nodes[:, 0]]
    mesh["nodes"] = nodes.astype(np.float32).ravel()
    mask = P.constrained_mask(mesh, [{"type": "fixed", "location": "all_edges"}])
    p3 = Plan(ctx, mesh, E, NU, 0.1, mask)
    assert p3.structured
    with pytest.raises(PlatePCGError):
        p3.solve(p3.pressure_load(1e4), precision=MIXED)
    # triangles have no structured path
    tri = P.split_quads(P.rect_mesh(8))
    p4 = Plan(ctx, tri, E, NU, 0.1, P.constrained_mask(tri, [{"type": "fixed", "location": "all_edges"}]),
              P.colour_greedy(tri))
    with pytest.raises(PlatePCGError):
        p4.solve(p4.pressure_load(1e4), precision=MIXED)


def test_mixed_c3_full_size(ctx):
    """The 700 x 700 plate (1.47M DOF, t = 0.05): mixed solve to 1e-8 true
    relative residual, re-checked by the oracle's FP64 K.x, and the same
    displacements as the library's FP64 solve."""
    w = P.workload("c3")
    plan = Plan(ctx, w["mesh"], w["E"], w["nu"], w["t"], w["mask"])
    F = plan.pressure_load(w["q0"])
    x64m, _, sm = plan.solve(F, tolerance=1e-8, max_iterations=200000, precision=MIXED)
    print("c3 mixed", {k: sm[k] for k in ("iterations", "sweeps", "trueRelResidual", "deviceMs")})
    assert sm["converged"] and sm["trueRelResidual"] <= 1e-8, sm
    x64d, _, sd = plan.solve(F, tolerance=1e-8, max_iterations=200000)
    print("c3 fp64", {k: sd[k] for k in ("iterations", "sweeps", "trueRelResidual", "deviceMs")})
    assert sd["converged"]
    assert O.max_nodal_w_error(x64m, x64d) < 1e-6
    # The oracle's K.x (stored isoparametric Ke, another summation order) and
    # the library's (closed-form element operator) are both FP64, but at this
    # size |K| |x| / |F| ~ 1e8, so two FP64 evaluations of F - K x differ by
    # ~1e-9 relative: the oracle's check carries that margin.
    op = O.OracleProblem(w["mesh"], w["E"], w["nu"], w["t"], w["mask"])
    Fd = F.astype(np.float64)
    r_m, r_d = op.true_rel_residual(x64m, Fd), op.true_rel_residual(x64d, Fd)
    print("c3 oracle residuals", r_m, r_d)
    assert r_m <= 1.5e-8 and r_d <= 1.5e-8
