"""The learned preconditioner on the GPU against its CPU oracle
(oracle/unet_ref.py), and what training it gains inside flexible PCG."""
import os
import sys

import numpy as np
import pytest

import oracle_api as O
from platepcg import Context, Net, Plan, PlatePCGError, _lib
from platepcg import problem as P

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import unet_ref as U  # noqa: E402

E, NU = 210e9, 0.3
LEARNED = _lib.PRECOND_LEARNED


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def plate(ctx, n, var_t=True, seed=3):
    mesh = P.rect_mesh(n)
    mask = P.constrained_mask(mesh, [{"type": "fixed", "location": "all_edges"}])
    et = None
    if var_t:
        rng = np.random.default_rng(seed)
        et = (0.1 * np.exp(0.4 * rng.standard_normal(n * n))).astype(np.float32)
    plan = Plan(ctx, mesh, E, NU, 0.1, mask, element_t=et)
    op = O.OracleProblem(mesh, E, NU, 0.1, mask, element_t=et)
    return mesh, mask, et, plan, op


def rand_params(shapes, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for _, cin, cout, _, _ in shapes:
        out.append(rng.standard_normal(cout * 9 * cin) * np.sqrt(2.0 / (9 * cin)))
        out.append(0.1 * rng.standard_normal(cout))
    return np.concatenate(out).astype(np.float32)


def test_training_error_matches_oracle(ctx):
    mesh, mask, _, plan, _ = plate(ctx, 32, var_t=False)
    net = Net(plan, seed=1)
    for seed in (1000, 1001, 4242):
        e = net.gen_error(seed)
        ref = U.training_error(mesh["nodes"], 32, 32, mask, seed)
        assert np.max(np.abs(e - ref)) <= 1e-12 * np.max(np.abs(ref)), seed
    net.close()


@pytest.mark.parametrize("n,relu", [(30, False), (32, False), (32, True), (130, False),
                                    (130, True)])
def test_forward_matches_oracle(ctx, n, relu):
    """M(r) on the GPU (BF16 MFMA, FP32 accumulate) against the oracle with
    the same BF16 storage points; n = 30 exercises the padded grid, n = 130
    several 64-pixel tiles per row on every level and partial row tiles."""
    mesh, mask, et, plan, op = plate(ctx, n)
    net = Net(plan, seed=1, relu=relu)
    ref = U.UNetOracle(n, n, op.blk, mask, 0.1, et, quantize=True, relu=relu)
    p = rand_params(ref.shapes, seed=n)
    net.set_params(p)
    np.testing.assert_array_equal(net.params(), p)
    e = U.training_error(mesh["nodes"], n, n, mask, seed=5)
    r = op.apply_k(e)
    m_gpu = net.apply(r)
    m_ref = ref.precondition(p, r, op.apply_k, net.omega)
    # the smoothing-only part (network output zero) is exact FP64 arithmetic
    zero = np.zeros_like(p)
    m_bj = U.UNetOracle(n, n, op.blk, mask, 0.1, et, quantize=False).precondition(
        zero, r, op.apply_k, net.omega)
    corr_gpu, corr_ref = m_gpu - m_bj, m_ref - m_bj
    err = np.max(np.abs(corr_gpu - corr_ref)) / np.max(np.abs(corr_ref))
    assert err < 2e-2, err
    # the network alone (pp_cnn_forward): N(v) = B v + the learned correction
    bv = np.einsum("nij,nj->ni", ref.blk, r.reshape(-1, 3)).reshape(-1)
    u_gpu, u_ref = net.forward(r), ref.forward(p, r)
    err = np.max(np.abs((u_gpu - bv) - (u_ref - bv))) / np.max(np.abs(u_ref - bv))
    assert err < 2e-2, err
    net.set_params(zero)
    np.testing.assert_allclose(net.apply(r), m_bj, rtol=1e-9, atol=1e-12 * np.abs(m_bj).max())
    np.testing.assert_allclose(net.forward(r), bv, rtol=1e-12, atol=1e-14 * np.abs(bv).max())
    net.close()


@pytest.mark.parametrize("n,relu", [(30, False), (130, False), (130, True), (257, False)])
def test_fused_first_layers_equal_separate_launches(ctx, n, relu):
    """The solver's forward pass runs e0a and e0b as one launch (the
    intermediate tensor kept in LDS); it must give the bits of separate
    launches, on padded grids, several tiles per row and a partial last row
    of tiles, with and without ReLU and biases."""
    mesh, mask, et, plan, op = plate(ctx, n)
    fused = Net(plan, seed=1, relu=relu)
    apart = Net(plan, seed=1, relu=relu, unfused_forward=True)
    p = rand_params(U.layer_shapes(), seed=n + 7)
    fused.set_params(p)
    apart.set_params(p)
    e = U.training_error(mesh["nodes"], n, n, mask, seed=3)
    r = op.apply_k(e)
    np.testing.assert_array_equal(fused.forward(r), apart.forward(r))
    np.testing.assert_array_equal(fused.apply(r), apart.apply(r))
    fused.close()
    apart.close()


@pytest.mark.parametrize("relu", [False, True])
def test_loss_and_gradient_match_oracle(ctx, relu):
    n = 32
    mesh, mask, et, plan, op = plate(ctx, n)
    net = Net(plan, seed=1, relu=relu)
    ref = U.UNetOracle(n, n, op.blk, mask, 0.1, et, quantize=True, relu=relu)
    p = rand_params(ref.shapes, seed=9)
    net.set_params(p)
    e = U.training_error(mesh["nodes"], n, n, mask, seed=17)
    loss, g = net.loss_grad(e)
    loss_ref, g_ref = ref.loss_grad(p, e, op.apply_k, net.omega)
    assert loss == pytest.approx(loss_ref, rel=1e-2)
    rel = np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
    cos = float(g @ g_ref) / (np.linalg.norm(g) * np.linalg.norm(g_ref))
    assert rel < 3e-2 and cos > 0.999, (rel, cos)
    off = 0
    for name, cin, cout, _, _ in ref.shapes:
        k = cout * 9 * cin + (cout if relu else 0)
        if not relu:
            assert not g[off + cout * 9 * cin:off + cout * 9 * cin + cout].any()
        a, b = g[off:off + k], g_ref[off:off + k]
        c = float(a @ b) / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-300)
        assert c > 0.99, (name, c)
        off += cout * 9 * cin + cout
    net.close()


def test_fcg_with_the_zero_network_converges(ctx):
    """With every parameter zero the preconditioner is block-Jacobi smoothing
    around a block-Jacobi correction taken with its energy-optimal step:
    flexible PCG (FCG(8), the default window) must reach 1e-8 with the
    block-Jacobi answer in fewer iterations than PCG."""
    n = 64
    mesh, mask, et, plan, op = plate(ctx, n, seed=8)
    F = plan.pressure_load(1e4, np.array([[3e4, 3.0, 6.0, 1.5]]))
    x_bj, _, st_bj = plan.solve(F, tolerance=1e-8, max_iterations=20000)
    net = Net(plan, seed=2)
    assert 0.05 < net.omega < 1.5
    net.set_params(np.zeros(net.n_params, np.float32))
    net.attach()
    x, _, st = plan.solve(F, tolerance=1e-8, max_iterations=20000, precond=LEARNED)
    print("iterations: block-Jacobi PCG", st_bj["iterations"], "FCG zero network", st["iterations"],
          "omega", net.omega)
    assert st["converged"] and st["trueRelResidual"] <= 1e-8, st
    assert O.max_nodal_w_error(x, x_bj) < 1e-6
    assert st["iterations"] < st_bj["iterations"]
    net.close()


def test_training_gains_inside_flexible_pcg(ctx):
    """Trained on solver-generated pairs (x* - x_k of block-Jacobi PCG on
    seeded pressure loads), the network lowers its energy-norm loss and then
    solves a new load case to 1e-8, with the FP64 block-Jacobi answer, in
    fewer iterations than block-Jacobi PCG, than the same preconditioner
    without the network (zero parameters) and than the untrained network;
    and FCG(8) beats FCG(1) with the trained network."""
    n = 64
    mesh, mask, et, plan, op = plate(ctx, n, seed=8)
    F = plan.pressure_load(1e4, np.array([[3e4, 3.0, 6.0, 1.5]]))
    x_bj, _, st_bj = plan.solve(F, tolerance=1e-8, max_iterations=20000)
    assert st_bj["converged"]
    its = {}
    for tag in ("zero", "untrained", "trained"):
        net = Net(plan, seed=2)
        if tag == "zero":
            net.set_params(np.zeros(net.n_params, np.float32))
        if tag == "trained":
            tr = net.train(pairs=48, snapshots=8, epochs=3, seed0=1000, lr=1e-3, lr_final=1e-5,
                           max_grad_norm=1.0, source=_lib.TRAIN_SOLVER)
            c = tr["loss_curve"]
            print("train %.1fs" % tr["seconds"], ["%.3g" % v for v in c[::8]])
            assert np.mean(c[-8:]) < 0.95 * np.mean(c[:8]), c
            trained = net.params()
        net.attach()
        x, _, st = plan.solve(F, tolerance=1e-8, max_iterations=5000, precond=LEARNED)
        its[tag] = st["iterations"] if st["converged"] else None
        if tag != "untrained":  # a random linear map need not be a preconditioner at all
            assert st["converged"] and st["trueRelResidual"] <= 1e-8, (tag, st)
            assert op.true_rel_residual(x, F.astype(np.float64)) <= 1e-8
            assert O.max_nodal_w_error(x, x_bj) < 1e-6
        if tag == "trained":
            net.close()
            net = Net(plan, seed=2, fcg_window=1)
            net.set_params(trained)
            net.attach()
            _, _, s1 = plan.solve(F, tolerance=1e-8, max_iterations=5000, precond=LEARNED)
            its["trained_fcg1"] = s1["iterations"] if s1["converged"] else None
        net.close()
    print("iterations: block-Jacobi PCG", st_bj["iterations"], its)
    assert its["trained"] < its["zero"] < st_bj["iterations"]
    assert its["untrained"] is None or its["trained"] < its["untrained"]
    assert its["trained_fcg1"] is None or its["trained"] < its["trained_fcg1"]


def test_learned_solve_edge_cases(ctx):
    mesh, mask, et, plan, op = plate(ctx, 16)
    F = plan.pressure_load(1e4)
    with pytest.raises(PlatePCGError):  # nothing attached
        plan.solve(F, precond=LEARNED)
    net = Net(plan, seed=3)
    net.attach()
    x, _, st = plan.solve(np.zeros(plan.n_dof, np.float32), precond=LEARNED)
    assert st["converged"] and st["iterations"] == 0 and not x.any()
    with pytest.raises(PlatePCGError):
        plan.solve(F, precond=LEARNED, precision=_lib.PRECISION_MIXED)
    # the network keeps its plan alive; either destruction order is safe
    plan.close()
    z = net.apply(op.apply_k(U.training_error(mesh["nodes"], 16, 16, mask, seed=1)))
    assert np.isfinite(z).all()
    net.close()
    # triangles have no structured grid
    tri = P.split_quads(P.rect_mesh(8))
    p4 = Plan(ctx, tri, E, NU, 0.1, P.constrained_mask(tri, [{"type": "fixed", "location": "all_edges"}]),
              P.colour_greedy(tri))
    with pytest.raises(PlatePCGError):
        Net(p4)
