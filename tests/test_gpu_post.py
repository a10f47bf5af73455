"""Device post-processing (pp_moments, pp_max_displacement) against the FP64
oracle restatement of postprocess.ts and closed forms, through the C ABI.
Outputs are float32 (the reference's Float32Arrays): tolerance 2e-7 relative
per value plus 2e-7 of the largest |M|."""
import numpy as np
import pytest

import oracle_api as O
from platepcg import Context, Plan
from platepcg import problem as P

from test_post_oracle import D, E, NU, T, distorted, linear_field

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def close(dev, ref):
    for d, r in zip(dev, ref):
        np.testing.assert_allclose(d, r, rtol=2e-7, atol=2e-7 * np.abs(r).max())


def test_moments_of_solved_c1_plate(ctx):
    """Structured Q4 plan: device moments of the device solution (x=None)
    equal the oracle's of the same x, and the float32-storage restatement of
    the reference to 1e-4 of the largest moment."""
    w = P.workload("c1")
    plan = Plan(ctx, w["mesh"], w["E"], w["nu"], w["t"], w["mask"])
    x, _, st = plan.solve(plan.pressure_load(w["q0"]), tolerance=1e-8)
    op = O.OracleProblem(w["mesh"], w["E"], w["nu"], w["t"], w["mask"])
    dev = plan.moments()
    close(dev, op.moments(x))
    close(plan.moments(x), op.moments(x))
    for d, r in zip(dev, op.moments(x, ref32=True)):
        assert np.abs(d - r).max() <= 1e-4 * np.abs(r).max()
    # clamped square plate under uniform load: centre moments equal by symmetry
    c = (w["mesh"]["nodeCount"] - 1) // 2
    assert abs(dev[0][c] - dev[1][c]) <= 1e-5 * abs(dev[0][c])
    plan.close()


@pytest.mark.parametrize("kind", ["dkt", "distorted_q4"])
def test_moments_general_meshes(ctx, kind):
    """Coloured-path plans (split-quad triangles, a non-grid Q4 mesh): a
    random DOF vector, and the constant-curvature patch field."""
    base = P.rect_mesh(12, 9)
    mesh = P.split_quads(base) if kind == "dkt" else distorted(base)
    col = P.colour_greedy(mesh)
    mask = np.zeros(mesh["nodeCount"], np.uint8)
    plan = Plan(ctx, mesh, E, NU, T, mask, col)
    assert not plan.structured
    op = O.OracleProblem(mesh, E, NU, T, mask)
    x = np.random.default_rng(7).standard_normal(plan.n_dof) * 1e-3
    close(plan.moments(x), op.moments(x))
    a, b = (0.3e-3, -1.1e-3, 2.3e-3), (-0.7e-3, 1.9e-3, 0.4e-3)
    kx, ky, kxy = b[1], -a[2], b[2] - a[1]
    want = (D * (kx + NU * ky), D * (ky + NU * kx), D * (1 - NU) * kxy / 2)
    for g, v in zip(plan.moments(linear_field(mesh, a, b)), want):
        np.testing.assert_allclose(g, v, rtol=1e-6)
    plan.close()


def test_moments_per_element_thickness(ctx):
    n = 48
    mesh = P.rect_mesh(n)
    et = (0.1 * np.exp(0.5 * P.gaussian_field(n, n, 10.0, 10.0, 0.5, 5))).astype(np.float32).ravel()
    mask = P.constrained_mask(mesh, [{"type": "fixed", "location": "all_edges"}])
    plan = Plan(ctx, mesh, E, NU, 0.1, mask, element_t=et)
    op = O.OracleProblem(mesh, E, NU, 0.1, mask, element_t=et)
    x = np.random.default_rng(8).standard_normal(plan.n_dof) * 1e-3
    close(plan.moments(x), op.moments(x))
    plan.close()


def test_max_displacement(ctx):
    """The signed w of the largest |w|, lowest node among equal |w| (the
    strict '>' scan of postprocess.ts:304-311), and the node's coordinates."""
    w = P.workload("c1")
    plan = Plan(ctx, w["mesh"], w["E"], w["nu"], w["t"], w["mask"])
    x, _, _ = plan.solve(plan.pressure_load(w["q0"]), tolerance=1e-8)
    r = plan.max_displacement()
    i = int(np.argmax(np.abs(x[0::3])))  # first occurrence of the maximum
    assert r["nodeIndex"] == i and r["maxW"] == x[3 * i]
    xy = w["mesh"]["nodes"].reshape(-1, 2)
    assert (r["x"], r["y"]) == (float(xy[i, 0]), float(xy[i, 1]))
    y = np.zeros(plan.n_dof)
    y[3 * 5], y[3 * 9], y[3 * 700] = -3.0, 3.0, 3.0
    r = plan.max_displacement(y)
    assert r["nodeIndex"] == 5 and r["maxW"] == -3.0
    r = plan.max_displacement(np.zeros(plan.n_dof))
    assert r["nodeIndex"] == 0 and r["maxW"] == 0.0
    with pytest.raises(ValueError):
        plan.moments(np.zeros(5))
    plan.close()
