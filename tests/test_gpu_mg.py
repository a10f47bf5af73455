"""The multigrid smoother of the learned preconditioner on the GPU against
its CPU oracle (oracle/mg_ref.py: sparse Galerkin triple products, built
independently of the library's coarse element matrices), through the C ABI;
and the learned preconditioner with it inside flexible PCG.  Coarse levels
store A as FP32 stencils: tolerances 2e-6 per level apply, 2e-5 per V-cycle
(relative to the largest entry)."""
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
import mg_ref as M  # noqa: E402
import unet_ref as U  # noqa: E402

E, NU = 210e9, 0.3
LEARNED = _lib.PRECOND_LEARNED


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def setup(ctx, nx, ny=None, var_t=True, seed=3, sweeps=0, coarse_sweeps=0, jacobi=False):
    ny = nx if ny is None else ny
    mesh = P.rect_mesh(nx, ny)
    mask = P.constrained_mask(mesh, [{"type": "fixed", "location": "all_edges"}])
    et = None
    if var_t:
        rng = np.random.default_rng(seed)
        et = (0.1 * np.exp(0.4 * rng.standard_normal(nx * ny))).astype(np.float32)
    plan = Plan(ctx, mesh, E, NU, 0.1, mask, element_t=et)
    op = O.OracleProblem(mesh, E, NU, 0.1, mask, element_t=et)
    net = Net(plan, seed=1, multigrid=True, mg_sweeps=sweeps, mg_coarse_sweeps=coarse_sweeps,
              mg_jacobi=jacobi)
    lv = net.mg_levels()
    K = M.assemble_q4(mesh["nodes"], nx, ny, E, NU, 0.1, et, O.ke_q4)
    free = (((mask[:, None] >> np.arange(3)) & 1) == 0).ravel()
    A0 = M.with_bc(K, free)
    ref = M.MgOracle(A0, nx, ny, free, sweeps=sweeps or (3 if jacobi else 6),
                     omegas=[w for _, _, w in lv[:-1]], blk0=op.blk,
                     coarse_sweeps=coarse_sweeps or 1, chebyshev=not jacobi)
    return dict(mesh=mesh, mask=mask, et=et, plan=plan, op=op, net=net, lv=lv, ref=ref, free=free)


def rel_err(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("nx,ny,var_t", [(32, 32, True), (64, 64, True), (48, 40, False),
                                          (50, 38, True)])
def test_levels_operators_and_vcycle_match_oracle(ctx, nx, ny, var_t):
    """Levels, their operators and one V-cycle against the oracle; 50 x 38
    coarsens through odd counts (25 x 19 -> 13 x 10), whose last coarse
    element is one level-1 element wide."""
    s = setup(ctx, nx, ny, var_t)
    net, ref, lv = s["net"], s["ref"], s["lv"]
    assert [(a, b) for a, b, _ in lv] == ref.dims
    assert lv[0][2] == pytest.approx(net.omega, rel=1e-15) and lv[-1][2] == 0.0
    # the device's power-iteration weights against the exact eigenvalue
    exact = M.MgOracle(ref.A[0], nx, ny, s["free"], omegas=None, blk0=s["op"].blk, factor=False)
    for l in range(len(lv) - 1):
        assert lv[l][2] == pytest.approx(exact.omega[l], rel=0.05), (l, lv[l][2], exact.omega[l])
    rng = np.random.default_rng(nx)
    for l in range(len(lv) - 1):
        # level 0: x vanishing on constrained DOFs, as every PCG vector does
        # (the reference's K x keeps their columns, replacing only the rows)
        x = rng.standard_normal(ref.A[l].shape[0]) * (s["free"] if l == 0 else 1.0)
        y = net.mg_apply(l, x)
        assert rel_err(y, ref.A[l] @ x) <= (1e-12 if l == 0 else 2e-6), l
    for k in range(2):
        b = rng.standard_normal(ref.A[0].shape[0]) * s["free"]
        assert rel_err(net.mg_cycle(b), ref.cycle(b)) <= 2e-5
    s["net"].close()
    s["plan"].close()


def test_sweep_count_and_unsupported_grids(ctx):
    # 2 damped block-Jacobi sweeps on every level, 64^2 (3 levels)
    t = setup(ctx, 64, sweeps=2, coarse_sweeps=2, jacobi=True)
    b = np.random.default_rng(5).standard_normal(t["ref"].A[0].shape[0]) * t["free"]
    assert rel_err(t["net"].mg_cycle(b), t["ref"].cycle(b)) <= 2e-5
    t["net"].close()
    t["plan"].close()
    s = setup(ctx, 32, sweeps=1)
    b = np.random.default_rng(4).standard_normal(s["ref"].A[0].shape[0]) * s["free"]
    assert rel_err(s["net"].mg_cycle(b), s["ref"].cycle(b)) <= 2e-5
    with pytest.raises(PlatePCGError):
        s["net"].mg_apply(len(s["lv"]) - 1, np.zeros(3 * 17 * 17))  # the coarsest has no stencil
    s["net"].close()
    no_mg = Net(s["plan"], seed=1)
    assert no_mg.mg_levels() == []
    with pytest.raises(PlatePCGError):
        no_mg.mg_cycle(b)
    no_mg.close()
    s["plan"].close()
    # 1500 x 2 coarsens to 750 x 1 and stops there (4,506 DOFs): too large a
    # coarsest level
    mesh = P.rect_mesh(1500, 2, width=150.0, height=0.2)
    mask = P.constrained_mask(mesh, [{"type": "fixed", "location": "all_edges"}])
    plan = Plan(ctx, mesh, E, NU, 0.1, mask)
    with pytest.raises(PlatePCGError):
        Net(plan, multigrid=True)
    plan.close()


@pytest.mark.parametrize("relu", [False, True])
def test_preconditioner_and_gradient_with_multigrid(ctx, relu):
    """M(r) and the loss gradient with the V-cycle as S, against the U-net
    oracle with the multigrid oracle as its smoother."""
    n = 32
    s = setup(ctx, n)
    net, op, ref, mesh, mask, et = s["net"], s["op"], s["ref"], s["mesh"], s["mask"], s["et"]
    net.close()
    net = Net(s["plan"], seed=1, relu=relu, multigrid=True)
    e = U.training_error(mesh["nodes"], n, n, mask, seed=5)
    r = op.apply_k(e)
    # zero network: FP64 smoothing and block-Jacobi correction only
    uq = U.UNetOracle(n, n, op.blk, mask, 0.1, et, quantize=True, relu=relu)
    zero = np.zeros(net.n_params, np.float32)
    net.set_params(zero)
    m0 = U.UNetOracle(n, n, op.blk, mask, 0.1, et, quantize=False).precondition(
        zero, r, op.apply_k, net.omega, smoother=ref.cycle)
    assert rel_err(net.apply(r), m0) <= 2e-5
    # a random network: the correction beyond the zero network, BF16 tolerance
    rng = np.random.default_rng(n)
    p = np.concatenate([np.concatenate([rng.standard_normal(co * 9 * ci) * np.sqrt(2.0 / (9 * ci)),
                                        0.1 * rng.standard_normal(co)])
                        for _, ci, co, _, _ in uq.shapes]).astype(np.float32)
    net.set_params(p)
    m_ref = uq.precondition(p, r, op.apply_k, net.omega, smoother=ref.cycle)
    err = np.abs((net.apply(r) - m0) - (m_ref - m0)).max() / np.abs(m_ref - m0).max()
    assert err < 2e-2, err
    loss, g = net.loss_grad(e)
    loss_ref, g_ref = uq.loss_grad(p, e, op.apply_k, net.omega, smoother=ref.cycle)
    assert loss == pytest.approx(loss_ref, rel=1e-2)
    cos = float(g @ g_ref) / (np.linalg.norm(g) * np.linalg.norm(g_ref))
    rel = np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
    assert rel < 3e-2 and cos > 0.999, (rel, cos)
    net.close()
    s["plan"].close()


def test_learned_multigrid_solve(ctx):
    """Flexible PCG with the multigrid-smoothed preconditioner (trained
    briefly) reaches 1e-8 with the FP64 block-Jacobi answer (0.01% bar on
    the nodal w) in a small fraction of block-Jacobi PCG's iterations."""
    n = 64
    s = setup(ctx, n, seed=8)
    plan, net, op = s["plan"], s["net"], s["op"]
    F = plan.pressure_load(1e4, np.array([[3e4, 3.0, 6.0, 1.5]]))
    x_bj, _, st_bj = plan.solve(F, tolerance=1e-8, max_iterations=20000)
    its = {}
    for tag in ("zero", "trained"):
        if tag == "zero":
            net.set_params(np.zeros(net.n_params, np.float32))
        else:
            net.close()
            net = Net(plan, seed=2, multigrid=True)
            tr = net.train(pairs=16, snapshots=8, epochs=2, seed0=1000, lr=1e-3, lr_final=1e-5,
                           max_grad_norm=1.0, source=_lib.TRAIN_SOLVER)
            print("train %.1fs loss %.4f -> %.4f" % (tr["seconds"], tr["loss_curve"][0],
                                                     tr["loss_curve"][-1]))
        net.attach()
        x, _, st = plan.solve(F, tolerance=1e-8, max_iterations=2000, precond=LEARNED)
        assert st["converged"] and st["trueRelResidual"] <= 1e-8, (tag, st)
        assert op.true_rel_residual(x, F.astype(np.float64)) <= 1e-8
        assert O.max_nodal_w_error(x, x_bj) < 1e-4
        its[tag] = st["iterations"]
    print("iterations: block-Jacobi PCG", st_bj["iterations"], its)
    assert max(its.values()) < 0.1 * st_bj["iterations"]
    net.close()
    plan.close()


def test_training_on_learned_iterates(ctx):
    """PP_TRAIN_LEARNED: pairs from flexible-PCG iterates with the network's
    own preconditioner.  The energy-norm loss falls over training and the
    trained preconditioner still solves to 1e-8 with the FP64 answer."""
    s = setup(ctx, 64, seed=8)
    plan, net, op = s["plan"], s["net"], s["op"]
    F = plan.pressure_load(1e4, np.array([[3e4, 3.0, 6.0, 1.5]]))
    x_bj, _, _ = plan.solve(F, tolerance=1e-8, max_iterations=20000)
    tr = net.train(pairs=4, snapshots=6, epochs=4, seed0=2000, lr=1e-3, lr_final=1e-5,
                   max_grad_norm=1.0, source=_lib.TRAIN_LEARNED)
    c = np.array(tr["loss_curve"])
    q = len(c) // 4
    print("loss %.4f -> %.4f" % (c[:q].mean(), c[-q:].mean()))
    assert np.isfinite(c).all() and c[-q:].mean() < c[:q].mean()
    net.attach()
    x, _, st = plan.solve(F, tolerance=1e-8, max_iterations=2000, precond=LEARNED)
    assert st["converged"] and st["trueRelResidual"] <= 1e-8, st
    assert O.max_nodal_w_error(x, x_bj) < 1e-4
    net.close()
    plan.close()


@pytest.mark.parametrize("case", ["pinned", "graded"])
def test_partial_constraints_and_graded_grid(ctx, case):
    """Pinned edges (only w constrained: partial exclusion masks per node)
    and a graded tensor grid (x_i = 10 (i / n)^1.5: every column its own
    element matrix): operators and V-cycle against the oracle, and a
    learned solve (zero network) to 1e-8 with the block-Jacobi answer."""
    n = 32
    mesh = P.rect_mesh(n)
    if case == "graded":
        nodes = mesh["nodes"].reshape(-1, 2).copy()
        xs = np.unique(nodes[:, 0])
        graded = {float(a): float(10.0 * (a / 10.0) ** 1.5) for a in xs}
        nodes[:, 0] = [graded[float(a)] for a in nodes[:, 0]]
        mesh["nodes"] = nodes.astype(np.float32).ravel()
    kind = "pinned" if case == "pinned" else "fixed"
    mask = P.constrained_mask(mesh, [{"type": kind, "location": "all_edges"}])
    plan = Plan(ctx, mesh, E, NU, 0.1, mask)
    assert plan.structured
    op = O.OracleProblem(mesh, E, NU, 0.1, mask)
    net = Net(plan, seed=1, multigrid=True)
    lv = net.mg_levels()
    K = M.assemble_q4(mesh["nodes"], n, n, E, NU, 0.1, None, O.ke_q4)
    free = (((mask[:, None] >> np.arange(3)) & 1) == 0).ravel()
    ref = M.MgOracle(M.with_bc(K, free), n, n, free, sweeps=6, coarse_sweeps=1,
                     omegas=[w for _, _, w in lv[:-1]], blk0=op.blk, chebyshev=True)
    rng = np.random.default_rng(11)
    for l in range(len(lv) - 1):
        x = rng.standard_normal(ref.A[l].shape[0]) * (free if l == 0 else 1.0)
        assert rel_err(net.mg_apply(l, x), ref.A[l] @ x) <= (1e-12 if l == 0 else 2e-6), l
    b = rng.standard_normal(ref.A[0].shape[0]) * free
    assert rel_err(net.mg_cycle(b), ref.cycle(b)) <= 2e-5
    F = plan.pressure_load(1e4)
    x_bj, _, _ = plan.solve(F, tolerance=1e-8, max_iterations=20000)
    net.set_params(np.zeros(net.n_params, np.float32))
    net.attach()
    x, _, st = plan.solve(F, tolerance=1e-8, max_iterations=2000, precond=LEARNED)
    assert st["converged"] and st["trueRelResidual"] <= 1e-8, st
    assert O.max_nodal_w_error(x, x_bj) < 1e-4
    net.close()
    plan.close()
