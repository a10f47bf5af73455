"""The 1024 x 1024 plate of the learned-preconditioner workload (3,151,875
DOF, log-normal thickness field, clamped, 10 kPa + three Gaussian patches)
at full size on the GPU: block-Jacobi PCG, multigrid PCG and the learned
preconditioner (trained briefly on solver-generated pairs) each reach the
1e-8 true relative residual, agree on the nodal displacements, and the FP64
oracle's own K x re-checks every residual."""
import numpy as np
import pytest

import oracle_api as O
from platepcg import Context, Net, Plan, _lib
from platepcg import problem as P

pytestmark = pytest.mark.gpu


def test_c5_full_size_three_solvers():
    w = P.workload("c5")
    ctx = Context(0)
    plan = Plan(ctx, w["mesh"], w["E"], w["nu"], w["t"], w["mask"], element_t=w["element_t"])
    assert plan.n_dof == 3_151_875
    F = plan.pressure_load(w["q0"], w["patches"])
    print("c5: block Jacobi", flush=True)
    xb, _, sb = plan.solve(F, tolerance=1e-8, max_iterations=200000)
    print("c5: multigrid", flush=True)
    xm, _, sm = plan.solve(F, tolerance=1e-8, max_iterations=5000,
                           precond=_lib.PRECOND_MULTIGRID)
    print("c5: learned (training 1 load case x 4 snapshots)", flush=True)
    net = Net(plan, seed=5, multigrid=True)
    net.train(pairs=1, snapshots=4, epochs=1, seed0=1000, lr=1e-3, lr_final=1e-5,
              max_grad_norm=1.0, source=_lib.TRAIN_SOLVER)
    net.attach()
    xl, _, sl = plan.solve(F, tolerance=1e-8, max_iterations=5000, precond=_lib.PRECOND_LEARNED)
    net.detach()
    for tag, st in (("block Jacobi", sb), ("multigrid", sm), ("learned", sl)):
        print(tag, {k: st[k] for k in ("iterations", "sweeps", "trueRelResidual", "deviceMs")})
        assert st["converged"] and st["trueRelResidual"] <= 1e-8, (tag, st)
    assert sm["iterations"] < sb["iterations"] / 50 and sl["iterations"] < sb["iterations"] / 50
    assert O.max_nodal_w_error(xm, xb) < 1e-6
    assert O.max_nodal_w_error(xl, xb) < 1e-6
    # The oracle's K.x (stored isoparametric element matrices, element-order
    # sums) and the library's (closed-form element operator, node-order sums)
    # are both FP64; on this plate the two evaluations of |F - K x| / |F|
    # differ by about 2e-10 (measured), so the oracle's check carries a 10%
    # margin.
    op = O.OracleProblem(w["mesh"], w["E"], w["nu"], w["t"], w["mask"],
                         element_t=w["element_t"])
    Fd = F.astype(np.float64)
    res = [op.true_rel_residual(x, Fd) for x in (xb, xm, xl)]
    print("oracle true residuals (block Jacobi, multigrid, learned):", res)
    assert max(res) <= 1.1e-8
    net.close()
    plan.close()
    ctx.close()
