"""PCG preconditioned by one Galerkin multigrid V-cycle per iteration
(PP_PRECOND_MULTIGRID) on the GPU, through the C ABI: the same FP64 answers
as block-Jacobi PCG and the oracle, in a small fraction of the iterations.
The hierarchy coarsens 142 x 142 through odd counts (71, 36, 18)."""
import os

import numpy as np
import pytest

import oracle_api as O
from platepcg import Context, Plan, PlatePCGError, _lib
from platepcg import problem as P

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MG = _lib.PRECOND_MULTIGRID
E, NU = 210e9, 0.3


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def test_c2_matches_golden_in_few_iterations(ctx):
    g = np.load(os.path.join(ROOT, "tests", "golden", "c2_142.npz"))
    w = P.workload("c2")
    plan = Plan(ctx, w["mesh"], w["E"], w["nu"], w["t"], w["mask"])
    _, _, sb = plan.solve(g["F"], tolerance=1e-8, max_iterations=20000)
    its, xs = {}, {}
    for steps in (2, 6):
        x, _, st = plan.solve(g["F"], tolerance=1e-8, max_iterations=2000, precond=MG,
                              poly_degree=steps)
        assert st["converged"] and st["trueRelResidual"] <= 1e-8, (steps, st)
        assert O.max_nodal_w_error(x, g["x64"]) < 1e-6, steps
        its[steps], xs[steps] = st["iterations"], x
    print("iterations: block Jacobi", sb["iterations"], "multigrid", its)
    assert its[6] < its[2] < sb["iterations"] / 20
    # the hierarchy is kept for the plan: a second solve gives the same bits
    x2, _, st2 = plan.solve(g["F"], tolerance=1e-8, max_iterations=2000, precond=MG,
                            poly_degree=6)
    assert st2["iterations"] == its[6] and np.array_equal(x2, xs[6])
    plan.close()


@pytest.mark.parametrize("shape", [(71, 45, 7.1, 4.5), (64, 48, 10.0, 7.5)])
def test_plates_match_oracle(ctx, shape):
    """Variable thickness; 71 x 45 is odd from level 0 on (the last coarse
    element of each level one element wide)."""
    nx, ny, wd, ht = shape
    mesh = P.rect_mesh(nx, ny, width=wd, height=ht)
    mask = P.constrained_mask(mesh, [{"type": "fixed", "location": "all_edges"}])
    rng = np.random.default_rng(nx)
    et = (0.08 * np.exp(0.3 * rng.standard_normal(nx * ny))).astype(np.float32)
    plan = Plan(ctx, mesh, E, NU, 0.08, mask, element_t=et)
    op = O.OracleProblem(mesh, E, NU, 0.08, mask, element_t=et)
    F = plan.pressure_load(1e4, np.array([[3e4, 0.7 * wd, 0.4 * ht, 0.8]]))
    xo, so = op.pcg(F.astype(np.float64), tol=1e-10)
    x, _, st = plan.solve(F, tolerance=1e-8, max_iterations=2000, precond=MG)
    assert st["converged"] and st["trueRelResidual"] <= 1e-8, st
    assert O.max_nodal_w_error(x, xo) < 1e-6
    assert st["iterations"] < so["iterations"] / 10, (st["iterations"], so["iterations"])
    assert op.true_rel_residual(x, F.astype(np.float64)) <= 1e-8
    plan.close()


def test_unsupported_cases(ctx):
    w = P.workload("c1")
    plan = Plan(ctx, w["mesh"], w["E"], w["nu"], w["t"], w["mask"])
    F = plan.pressure_load(1e4)
    with pytest.raises(PlatePCGError):
        plan.solve(F, precond=MG, precision=_lib.PRECISION_MIXED)
    with pytest.raises(PlatePCGError):
        plan.solve(F, precond=MG, poly_degree=9)
    x, _, st = plan.solve(np.zeros(plan.n_dof, np.float32), precond=MG)
    assert st["converged"] and st["iterations"] == 0 and not x.any()
    plan.close()
    # too elongated to reach a small coarsest level
    mesh = P.rect_mesh(1500, 2, width=150.0, height=0.2)
    mask = P.constrained_mask(mesh, [{"type": "fixed", "location": "all_edges"}])
    plan = Plan(ctx, mesh, E, NU, 0.05, mask)
    with pytest.raises(PlatePCGError):
        plan.solve(plan.pressure_load(1e4), precond=MG)
    plan.close()
    tri = P.split_quads(P.rect_mesh(8))
    p4 = Plan(ctx, tri, E, NU, 0.1,
              P.constrained_mask(tri, [{"type": "fixed", "location": "all_edges"}]),
              P.colour_greedy(tri))
    with pytest.raises(PlatePCGError):
        p4.solve(p4.pressure_load(1e4), precond=MG)
    p4.close()


def test_restarts_near_the_fp64_floor(ctx):
    """Near the FP64 floor of the true residual the recurrence converges before
    F - K x does, so flexible PCG restarts from x with the true residual (the
    first tolerance of 1e-10 ... 1e-13 at which the 142^2 plate needs a
    restart; 1e-9 needs none).  The sweep statistics add up, and the solve
    reports convergence only when the FP64 true residual is at most the
    tolerance."""
    w = P.workload("c2")
    plan = Plan(ctx, w["mesh"], w["E"], w["nu"], w["t"], w["mask"])
    F = plan.pressure_load(w["q0"], w["patches"])
    for tol in (1e-10, 1e-11, 1e-12, 1e-13):
        x, _, st = plan.solve(F, tolerance=tol, max_iterations=2000, precond=MG)
        print("tol", tol, "restart sweeps", st["sweeps"], st["sweepIterations"],
              st["sweepResidual"], "converged", st["converged"], st["trueRelResidual"])
        if st["sweeps"] > 1:
            break
    assert st["sweeps"] > 1
    assert sum(st["sweepIterations"]) == st["iterations"]
    assert st["sweepResidual"][st["sweeps"] - 1] == st["trueRelResidual"]
    assert st["converged"] == (st["trueRelResidual"] <= tol)
    w_ref = O.OracleProblem(w["mesh"], w["E"], w["nu"], w["t"], w["mask"])
    assert w_ref.true_rel_residual(x, F.astype(np.float64)) < max(3 * tol, 1e-9)
    # without the FP64 check: one run, no true residual, convergence of the recurrence
    _, _, s1 = plan.solve(F, tolerance=tol, max_iterations=2000, precond=MG,
                          flags=_lib.FLAG_NO_TRUE_RESIDUAL)
    assert s1["sweeps"] == 1 and s1["trueRelResidual"] == 0.0 and s1["converged"]
    assert s1["iterations"] == st["sweepIterations"][0]
    plan.close()


@pytest.mark.parametrize("case", [("c1", 6), ("c2", 1), ("c2", 3), ("c2", 5), ("c2", 6),
                                  ("rect", 4), ("large", 5)])
def test_fused_launches_match_launch_per_operation(ctx, case):
    """The fused iteration (19 launches on the 142^2 plate: K p with the
    direction update, x/r update with the first smoothing step, residual with
    restriction and the coarse first step, prolongation with the first
    post-smoothing step, reductions finished by the last workgroup) computes
    the same V-cycle and PCG as one launch per operation: the same iteration
    count within one, and the same displacements to 1e-8 of max |w|.  Odd and
    even step counts (the buffer rotation ends in z either way), a plate with
    no stencil level (32^2: level 0 then the dense coarsest) and a
    rectangular odd-sized plate; at 400 x 400 (729 tiles) the reductions are
    finished by their own single-workgroup launch instead of by every
    consuming workgroup."""
    name, steps = case
    if name == "large":
        n = 400
        mesh = P.rect_mesh(n)
        mask = P.constrained_mask(mesh, [{"type": "fixed", "location": "all_edges"}])
        et = (0.1 * np.exp(0.4 * np.random.default_rng(3).standard_normal(n * n))).astype(np.float32)
        plan = Plan(ctx, mesh, E, NU, 0.1, mask, element_t=et)
        F = plan.pressure_load(1e4, np.array([[3e4, 4.0, 6.0, 1.0]]))
    elif name == "rect":
        mesh = P.rect_mesh(90, 37, width=10.0, height=4.0)
        mask = P.edge_mask(90, 37, ["fixed", "pinned", "fixed", "pinned"])
        plan = Plan(ctx, mesh, E, NU, 0.05, mask)
        F = plan.pressure_load(1e4, np.array([[3e4, 6.0, 1.5, 0.7]]))
    else:
        w = P.workload(name)
        plan = Plan(ctx, w["mesh"], w["E"], w["nu"], w["t"], w["mask"])
        F = plan.pressure_load(w["q0"], w["patches"])
    xf, _, sf = plan.solve(F, tolerance=1e-8, max_iterations=2000, precond=MG, poly_degree=steps)
    xu, _, su = plan.solve(F, tolerance=1e-8, max_iterations=2000, precond=MG, poly_degree=steps,
                           flags=_lib.FLAG_MG_UNFUSED)
    print(case, "fused", sf["iterations"], f"{sf['deviceMs']:.2f} ms", "unfused",
          su["iterations"], f"{su['deviceMs']:.2f} ms")
    assert sf["converged"] and su["converged"]
    assert sf["trueRelResidual"] <= 1e-8
    assert abs(sf["iterations"] - su["iterations"]) <= 1
    wmax = np.abs(xu[0::3]).max()
    assert np.abs(xf[0::3] - xu[0::3]).max() < 1e-8 * wmax
    plan.close()
