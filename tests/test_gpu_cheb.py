"""The Chebyshev-polynomial block-Jacobi preconditioner (PP_PRECOND_CHEBYSHEV)
on the GPU: the same FP64 answers as block-Jacobi PCG and the oracle, in a
fraction of the iterations, through the C ABI.  Degree 1 is block Jacobi
scaled by 1 / theta, which CG does not see: the same iteration count."""
import os

import numpy as np
import pytest

import oracle_api as O
from platepcg import Context, Plan, PlatePCGError, _lib
from platepcg import problem as P

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHEB = _lib.PRECOND_CHEBYSHEV
E, NU = 210e9, 0.3


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def test_c2_matches_golden_in_fewer_iterations(ctx):
    g = np.load(os.path.join(ROOT, "tests", "golden", "c2_142.npz"))
    w = P.workload("c2")
    plan = Plan(ctx, w["mesh"], w["E"], w["nu"], w["t"], w["mask"])
    _, _, sb = plan.solve(g["F"], tolerance=1e-8, max_iterations=20000)
    its = {}
    for deg in (1, 4, 8, 16):
        x64, _, st = plan.solve(g["F"], tolerance=1e-8, max_iterations=20000, precond=CHEB,
                                poly_degree=deg)
        assert st["converged"] and st["trueRelResidual"] <= 1e-8, (deg, st)
        assert O.max_nodal_w_error(x64, g["x64"]) < 1e-6, deg
        its[deg] = st["iterations"]
    print("iterations: block Jacobi", sb["iterations"], "Chebyshev", its)
    assert abs(its[1] - sb["iterations"]) <= 3 + sb["iterations"] // 100
    assert its[8] < sb["iterations"] / 4 and its[16] < its[8] < its[4]
    op = O.OracleProblem(w["mesh"], w["E"], w["nu"], w["t"], w["mask"])
    assert op.true_rel_residual(x64, g["F"].astype(np.float64)) <= 1e-8
    plan.close()


@pytest.mark.parametrize("shape", [(32, 32, 10.0, 10.0), (90, 37, 10.0, 4.0)])
def test_plates_match_oracle(ctx, shape):
    """C1 and a rectangular plate with a partial tile row and column (the
    neighbour exchange at plate edges and corners)."""
    nx, ny, wd, ht = shape
    mesh = P.rect_mesh(nx, ny, width=wd, height=ht)
    mask = P.constrained_mask(mesh, [{"type": "fixed", "location": "all_edges"}])
    rng = np.random.default_rng(nx)
    et = (0.08 * np.exp(0.3 * rng.standard_normal(nx * ny))).astype(np.float32)
    plan = Plan(ctx, mesh, E, NU, 0.08, mask, element_t=et)
    op = O.OracleProblem(mesh, E, NU, 0.08, mask, element_t=et)
    F = plan.pressure_load(1e4, np.array([[3e4, 0.7 * wd, 0.4 * ht, 0.8]]))
    xo, so = op.pcg(F.astype(np.float64), tol=1e-10)
    x, _, st = plan.solve(F, tolerance=1e-8, max_iterations=20000, precond=CHEB)
    assert st["converged"] and st["trueRelResidual"] <= 1e-8, st
    assert O.max_nodal_w_error(x, xo) < 1e-6
    assert st["iterations"] < so["iterations"] / 3, (st["iterations"], so["iterations"])
    # a second solve reuses the lambda_max estimate and gives the same bits
    x2, _, st2 = plan.solve(F, tolerance=1e-8, max_iterations=20000, precond=CHEB)
    assert st2["iterations"] == st["iterations"] and np.array_equal(x, x2)
    plan.close()


def test_unsupported_cases(ctx):
    # too many tiles to be co-resident
    mesh = P.rect_mesh(700)
    mask = P.constrained_mask(mesh, [{"type": "fixed", "location": "all_edges"}])
    plan = Plan(ctx, mesh, E, NU, 0.05, mask)
    with pytest.raises(PlatePCGError):
        plan.solve(plan.pressure_load(1e4), precond=CHEB)
    plan.close()
    w = P.workload("c1")
    plan = Plan(ctx, w["mesh"], w["E"], w["nu"], w["t"], w["mask"])
    F = plan.pressure_load(1e4)
    with pytest.raises(PlatePCGError):
        plan.solve(F, precond=CHEB, precision=_lib.PRECISION_MIXED)
    with pytest.raises(PlatePCGError):
        plan.solve(F, precond=CHEB, poly_degree=64)
    x, _, st = plan.solve(np.zeros(plan.n_dof, np.float32), precond=CHEB)
    assert st["converged"] and st["iterations"] == 0 and not x.any()
    plan.close()
    tri = P.split_quads(P.rect_mesh(8))
    p4 = Plan(ctx, tri, E, NU, 0.1,
              P.constrained_mask(tri, [{"type": "fixed", "location": "all_edges"}]),
              P.colour_greedy(tri))
    with pytest.raises(PlatePCGError):
        p4.solve(p4.pressure_load(1e4), precond=CHEB)
    p4.close()
