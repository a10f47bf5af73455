"""The CPU oracle of the learned preconditioner, pinned by the mathematics:
adjoint identities of its linear pieces and central differences of its
hand-derived gradient."""
import os
import sys

import numpy as np
import pytest

import oracle_api as O
from platepcg import problem as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import unet_ref as U  # noqa: E402

E, NU = 210e9, 0.3


def test_restrict_and_prolong_are_adjoint():
    """P = 4 R^T on vertex-centred grids: <R x, y> = <x, P y> / 4."""
    rng = np.random.default_rng(0)
    x = rng.standard_normal((17, 9, 3))
    y = rng.standard_normal((9, 5, 3))
    assert np.sum(U.restrict(x) * y) == pytest.approx(np.sum(x * U.prolong(y)) / 4, rel=1e-12)
    # constants are reproduced by P inside, and by R away from the boundary
    c = np.ones((9, 5, 2))
    np.testing.assert_allclose(U.prolong(c), 1.0)
    np.testing.assert_allclose(U.restrict(np.ones((17, 9, 1)))[1:-1, 1:-1], 1.0)


@pytest.mark.parametrize("dil", [1, 2])
def test_conv_backward_is_the_adjoint(dil):
    rng = np.random.default_rng(dil)
    x = rng.standard_normal((11, 7, 4))
    w = rng.standard_normal((5, 3, 3, 4))
    gy = rng.standard_normal((11, 7, 5))
    y0 = U.conv(x, w, np.zeros(5), dil)
    gw, gb, gx = U.conv_back(x, w, dil, gy)
    assert np.sum(y0 * gy) == pytest.approx(np.sum(x * gx), rel=1e-12)
    assert np.sum(y0 * gy) == pytest.approx(np.sum(w * gw), rel=1e-12)
    assert np.allclose(gb, gy.sum(axis=(0, 1)))
    # one tap against a direct sum
    ref = sum(x[h + dil, w_ + dil, 2] * gy[h, w_, 1] for h in range(11 - dil) for w_ in range(7 - dil))
    assert gw[1, 2, 2, 2] == pytest.approx(ref, rel=1e-12)


def small_problem(n=8, var_t=True):
    mesh = P.rect_mesh(n)
    mask = P.constrained_mask(mesh, [{"type": "fixed", "location": "all_edges"}])
    et = None
    if var_t:
        rng = np.random.default_rng(3)
        et = (0.1 * np.exp(0.4 * rng.standard_normal(n * n))).astype(np.float32)
    op = O.OracleProblem(mesh, E, NU, 0.1, mask, element_t=et)
    return mesh, mask, et, op


def init_params(shapes, seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    out = []
    for _, cin, cout, _, _ in shapes:
        out.append(rng.standard_normal(cout * 9 * cin) * np.sqrt(2.0 / (9 * cin)) * scale)
        out.append(0.1 * rng.standard_normal(cout))
    return np.concatenate(out).astype(np.float32)


@pytest.mark.parametrize("relu", [False, True])
def test_loss_gradient_matches_central_differences(relu):
    mesh, mask, et, op = small_problem(8)
    net = U.UNetOracle(8, 8, op.blk, mask, 0.1, et, quantize=False, relu=relu)
    p = init_params(net.shapes).astype(np.float64)
    e = U.training_error(mesh["nodes"], 8, 8, mask, seed=7)
    loss, g = net.loss_grad(p, e, op.apply_k, 0.4)
    assert 0 < loss and np.isfinite(g).all()
    rng = np.random.default_rng(11)
    for _ in range(3):
        v = rng.standard_normal(p.size)
        h = 1e-6 / np.linalg.norm(v) * np.linalg.norm(p)
        lp, _ = net.loss_grad(p + h * v, e, op.apply_k, 0.4)
        lm, _ = net.loss_grad(p - h * v, e, op.apply_k, 0.4)
        fd = (lp - lm) / (2 * h)
        assert fd == pytest.approx(float(g @ v), rel=1e-5, abs=1e-9 * abs(loss))


def test_forward_zero_network_is_block_jacobi():
    """With every weight and bias zero the output correction vanishes and
    M(r) is the block-Jacobi preconditioner (solver.ts:313-339)."""
    mesh, mask, et, op = small_problem(8, var_t=False)
    net = U.UNetOracle(8, 8, op.blk, mask, 0.1, None, quantize=False)
    p = np.zeros(sum(c * 9 * i + c for _, i, c, _, _ in net.shapes))
    r = np.random.default_rng(2).standard_normal(op.nn * 3)
    r[op.cdof.astype(bool)] = 0
    np.testing.assert_allclose(net.forward(p, r), op.precond(r), rtol=1e-14, atol=0)


def test_training_error_is_kirchhoff_consistent():
    """thx = dw/dy, thy = -dw/dx (the rigid-mode signs of element.ts) and
    zero on constrained DOFs."""
    n = 64
    mesh = P.rect_mesh(n)
    mask = P.constrained_mask(mesh, [{"type": "fixed", "location": "all_edges"}])
    e = U.training_error(mesh["nodes"], n, n, mask, seed=3).reshape(n + 1, n + 1, 3)
    h = 10.0 / n
    dwdy = (e[2:, 1:-1, 0] - e[:-2, 1:-1, 0]) / (2 * h)
    dwdx = (e[1:-1, 2:, 0] - e[1:-1, :-2, 0]) / (2 * h)
    scale = np.abs(e[..., 1]).max()
    assert np.abs(dwdy - e[1:-1, 1:-1, 1]).max() < 0.05 * scale
    assert np.abs(-dwdx - e[1:-1, 1:-1, 2]).max() < 0.05 * scale
    border = np.ones((n + 1, n + 1), bool)
    border[1:-1, 1:-1] = False
    assert not e[border].any()


def test_linear_network_is_linear():
    """Without activation and biases the preconditioner is a linear operator."""
    mesh, mask, et, op = small_problem(8)
    net = U.UNetOracle(8, 8, op.blk, mask, 0.1, et, quantize=False, relu=False)
    p = init_params(net.shapes, seed=4).astype(np.float64)
    rng = np.random.default_rng(5)
    a, b = rng.standard_normal(op.nn * 3), rng.standard_normal(op.nn * 3)
    a[op.cdof.astype(bool)] = 0
    b[op.cdof.astype(bool)] = 0
    lhs = net.precondition(p, 2.0 * a - 3.0 * b, op.apply_k, 0.4)
    rhs = 2.0 * net.precondition(p, a, op.apply_k, 0.4) - 3.0 * net.precondition(p, b, op.apply_k, 0.4)
    np.testing.assert_allclose(lhs, rhs, rtol=1e-10, atol=1e-12 * np.abs(rhs).max())
