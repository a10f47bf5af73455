"""Pins the multigrid oracle (oracle/mg_ref.py) by what the mathematics
fixes: rigid-body modes of a free plate stay in the kernel of every Galerkin
level (bilinear interpolation reproduces linear fields, so a wrong weight or
index breaks it), the V-cycle is a symmetric operator, its error propagation
I - V A contracts in the energy norm (with damped block-Jacobi or Chebyshev
smoothing), and as a preconditioner it converges to the oracle's own
block-Jacobi PCG solution in a fraction of the iterations.
CPU only."""
import os
import sys

import numpy as np
import pytest
import scipy.sparse.linalg as spl

import oracle_api as O
from platepcg import problem as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mg_ref as M  # noqa: E402

E, NU = 210e9, 0.3


def plate(n, ny=None, var_t=True, clamped=True, seed=3, h=None):
    """h: element size (float32 node coordinates are exactly uniform only
    for a dyadic size; default a 10 m square)."""
    ny = n if ny is None else ny
    mesh = P.rect_mesh(n, ny) if h is None else P.rect_mesh(n, ny, width=n * h, height=ny * h)
    et = None
    if var_t:
        rng = np.random.default_rng(seed)
        et = (0.1 * np.exp(0.4 * rng.standard_normal(n * ny))).astype(np.float32)
    sup = [{"type": "fixed", "location": "all_edges"}] if clamped else []
    mask = P.constrained_mask(mesh, sup)
    K = M.assemble_q4(mesh["nodes"], n, ny, E, NU, 0.1, et, O.ke_q4)
    free = np.repeat(mask == 0, 1)[:, None] & (((mask[:, None] >> np.arange(3)) & 1) == 0)
    free = free.ravel()
    return mesh, mask, et, K, free


def level_coords(nodes, nx, ny, l):
    """Coordinates of the level-l nodes: coarse node c of a side with n
    elements sits on node min(2c, n) of the level above."""
    xy = np.asarray(nodes, np.float64).reshape(ny + 1, nx + 1, 2)
    ix, iy = np.arange(nx + 1), np.arange(ny + 1)
    for _ in range(l):
        ix = ix[np.minimum(2 * np.arange((len(ix) - 1 + 1) // 2 + 1), len(ix) - 1)]
        iy = iy[np.minimum(2 * np.arange((len(iy) - 1 + 1) // 2 + 1), len(iy) - 1)]
    return xy[np.ix_(iy, ix)].reshape(-1, 2)


def test_hierarchy_shapes():
    assert M.hierarchy_dims(64, 64) == [(64, 64), (32, 32), (16, 16)]
    assert M.hierarchy_dims(32, 32) == [(32, 32), (16, 16)]
    assert M.hierarchy_dims(48, 40) == [(48, 40), (24, 20), (12, 10)]
    assert M.hierarchy_dims(1024, 1024)[-1] == (16, 16)
    # odd counts round up: the last coarse interval is one element wide
    assert M.hierarchy_dims(142, 142) == [(142, 142), (71, 71), (36, 36), (18, 18)]
    assert M.hierarchy_dims(50, 38) == [(50, 38), (25, 19), (13, 10)]
    assert M.interp_1d(7).shape == (8, 5)


@pytest.mark.parametrize("n,ny", [(64, 64), (50, 38)])
def test_rigid_modes_stay_in_the_kernel(n, ny):
    """Free plate: w = 1 and the two rigid rotations (w linear, the matching
    rotation constant) give A_l v = 0 on every level, also where odd element
    counts leave a last coarse interval one element wide (50 x 38 -> 25 x 19
    -> 13 x 10).  Elements of 0.25 m: index-space midpoints are then the
    physical ones in float32 coordinates."""
    mesh, mask, et, K, free = plate(n, ny, clamped=False, h=0.25)
    mg = M.MgOracle(K, n, ny, free, omegas=[1.0] * 3, factor=False)

    def modes(xy, signs):
        out = [np.tile([1.0, 0.0, 0.0], len(xy))]
        for (axis, dof), sign in zip(((0, 2), (1, 1)), signs):
            v = np.zeros(3 * len(xy))
            v[0::3] = xy[:, axis]
            v[dof::3] = sign
            out.append(v)
        return out

    # the sign convention of each rotation: the one K annihilates on level 0
    xy0 = level_coords(mesh["nodes"], n, ny, 0)
    scale0 = np.abs(K).max() * 10.0
    signs = []
    for k in range(2):
        ok = [s for s in (1.0, -1.0)
              if np.abs(K @ modes(xy0, [s, s])[1 + k]).max() <= 1e-9 * scale0]
        assert len(ok) == 1, "exactly one sign of each rotation is rigid"
        signs.append(ok[0])
    for l in range(len(mg.dims)):
        A = mg.A[l]
        scale = np.abs(A).max() * 10.0
        for v in modes(level_coords(mesh["nodes"], n, ny, l), signs):
            assert np.abs(A @ v).max() <= 1e-9 * scale, (l, np.abs(A @ v).max() / scale)
        # and a non-rigid field is not annihilated
        v = modes(level_coords(mesh["nodes"], n, ny, l), [-signs[0], -signs[1]])[1]
        assert np.abs(A @ v).max() > 1e-3 * scale


@pytest.mark.parametrize("chebyshev,n,ny", [(False, 64, 64), (True, 64, 64), (True, 50, 38)])
def test_vcycle_is_symmetric_and_contracts(chebyshev, n, ny):
    _, mask, _, K, free = plate(n, ny)
    A0 = M.with_bc(K, free)
    mg = M.MgOracle(A0, n, ny, free, sweeps=3, coarse_sweeps=1, chebyshev=chebyshev)
    assert len(mg.A) == 3
    rng = np.random.default_rng(0)
    x, y = rng.standard_normal((2, A0.shape[0]))
    a, b = y @ mg.cycle(x), x @ mg.cycle(y)
    assert abs(a - b) <= 1e-10 * abs(a)
    # positive definite on random vectors
    assert x @ mg.cycle(x) > 0
    # energy-norm contraction of e <- (I - V A) e
    e = rng.standard_normal(A0.shape[0]) * free
    en = lambda v: np.sqrt(v @ (A0 @ v))
    r0 = en(e)
    for _ in range(5):
        e = e - mg.cycle(A0 @ e)
    assert en(e) < 0.5 ** 5 * r0


def test_mg_pcg_matches_block_jacobi_solution():
    """V-cycle-preconditioned CG reaches the oracle's FP64 block-Jacobi PCG
    solution (the reference algorithm) on a clamped variable-thickness
    plate, in under a tenth of its iterations."""
    n = 64
    mesh, mask, et, K, free = plate(n)
    op = O.OracleProblem(mesh, E, NU, 0.1, mask, element_t=et)
    A0 = M.with_bc(K, free)
    x = np.random.default_rng(1).standard_normal(A0.shape[0]) * free
    np.testing.assert_allclose(A0 @ x, op.apply_k(x), rtol=1e-9, atol=1e-9 * np.abs(A0 @ x).max())
    F = op.pressure_load(10e3)
    xb, st = op.pcg(F, tol=1e-10)
    its = {}
    for cheb in (False, True):
        mg = M.MgOracle(A0, n, n, free, blk0=op.blk, sweeps=3, coarse_sweeps=1, chebyshev=cheb)
        it = [0]

        def cb(_):
            it[0] += 1

        Mop = spl.LinearOperator(A0.shape, matvec=mg.cycle)
        xm, info = spl.cg(A0, F, rtol=1e-10, atol=0.0, M=Mop, callback=cb, maxiter=2000)
        assert info == 0
        assert it[0] < 0.1 * st["iterations"], (it[0], st["iterations"])
        w, wm = xb[0::3], xm[0::3]
        assert np.abs(w - wm).max() <= 1e-7 * np.abs(w).max()
        its[cheb] = it[0]
    # the Chebyshev smoother (same number of operator applies) converges faster
    assert its[True] < its[False], its
