"""Parity of the HIP library with the CPU oracles, through the C ABI."""
import os

import numpy as np
import pytest

import oracle_api as O
from platepcg import Context, Plan, PlatePCGError, _lib
from platepcg import problem as P

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
E, NU = 210e9, 0.3


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def make(ctx, w, coloured=True, triangles=False):
    mesh = w["mesh"]
    if triangles:
        mesh = P.split_quads(mesh)
    col = None
    if coloured:
        col = P.colour_greedy(mesh) if triangles else P.colour_structured(mesh)
    plan = Plan(ctx, mesh, w["E"], w["nu"], w["t"], w["mask"], col)
    op = O.OracleProblem(mesh, w["E"], w["nu"], w["t"], w["mask"],
                         order=None if col is None else col["elements_by_colour"])
    return plan, op, mesh, col


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("name", ["c1", "c2"])
def test_apply_k_matches_oracle(ctx, name):
    plan, op, _, _ = make(ctx, P.workload(name))
    assert plan.structured
    x = np.random.default_rng(1).standard_normal(plan.n_dof)
    y_ref = op.apply_k(x)
    assert rel(plan.apply_k(x, _lib.PATH_STRUCTURED), y_ref) < 1e-12
    assert rel(plan.apply_k(x, _lib.PATH_COLOURED), y_ref) < 1e-12


def test_single_colour_apply_matches_oracle(ctx):
    """gpu/single_color_test.ts:175-237: one colour's K.x, constrained DOFs
    skipped in the scatter."""
    plan, op, mesh, col = make(ctx, P.workload("c1"))
    x = np.random.default_rng(2).standard_normal(plan.n_dof)
    nd = 12
    Ke = op.Ke.reshape(-1, nd, nd)
    el = mesh["elements"].reshape(-1, 4).astype(np.int64)
    off = col["colour_offsets"]
    for c in range(col["colour_count"]):
        y = np.zeros(plan.n_dof)
        for e in col["elements_by_colour"][off[c]:off[c + 1]]:
            dofs = (el[e][:, None] * 3 + np.arange(3)).ravel()
            ye = Ke[e] @ x[dofs]
            keep = op.cdof[dofs] == 0
            y[dofs[keep]] += ye[keep]
        assert rel(plan.apply_k_colour(c, x), y) < 1e-12


@pytest.mark.parametrize("name", ["c1", "c2"])
def test_block_inverse_and_precond_match_oracle(ctx, name):
    plan, op, _, _ = make(ctx, P.workload(name), coloured=False)
    b = plan.block_inverse().reshape(-1, 9)
    ob = op.blk.reshape(-1, 9)
    assert np.max(np.abs(b - ob) / np.abs(ob).max(axis=1, keepdims=True)) < 1e-10
    r = np.random.default_rng(3).standard_normal(plan.n_dof)
    assert rel(plan.apply_precond(r), op.precond(r)) < 1e-10


@pytest.mark.parametrize("name", ["c1", "c2"])
def test_pressure_load_matches_oracle(ctx, name):
    w = P.workload(name)
    plan, op, _, _ = make(ctx, w, coloured=False)
    F = plan.pressure_load(w["q0"], w["patches"])
    Fo = op.pressure_load(w["q0"], w["patches"]).astype(np.float32)
    np.testing.assert_allclose(F, Fo, rtol=2e-7, atol=0)


@pytest.mark.parametrize("path", [_lib.PATH_STRUCTURED, _lib.PATH_COLOURED])
def test_solve_c1_matches_oracles(ctx, path):
    w = P.workload("c1")
    plan, op, _, _ = make(ctx, w)
    F = plan.pressure_load(w["q0"])
    x64, x32, st = plan.solve(F, tolerance=1e-8, max_iterations=20000, path=path)
    assert st["converged"] and not st["breakdown"]
    xo, so = op.pcg(F.astype(np.float64), tol=1e-8)
    assert abs(st["iterations"] - so["iterations"]) <= 3
    assert st["trueRelResidual"] <= 1e-8
    assert op.true_rel_residual(x64, F.astype(np.float64)) <= 1e-8
    assert O.max_nodal_w_error(x64, xo) < 1e-6
    xj, js = O.run_js_solver(32, F, E, NU, 0.1, [{"type": "fixed", "location": "all_edges"}])
    assert O.compare_max_w(x32, xj) < 1e-4
    assert O.max_nodal_w_error(x32, xj) < 1e-3


@pytest.mark.parametrize("precond", [_lib.PRECOND_BLOCK_JACOBI, _lib.PRECOND_MULTIGRID])
def test_device_true_residual_against_exact(ctx, precond):
    """The device's FP64 evaluation of |F - K x| / |F|, which decides the
    1e-8 stop and the residual-replacement restarts, against the exact
    (rational) evaluation on the oracle's element matrices: within 1e-10 on
    a 32 x 32 plate with log-normal thickness (t^3 spanning about 20x), for
    the block-Jacobi and the multigrid solver."""
    n = 32
    mesh = P.rect_mesh(n)
    mask = P.constrained_mask(mesh, [{"type": "fixed", "location": "all_edges"}])
    rng = np.random.default_rng(11)
    et = (0.1 * np.exp(0.5 * rng.standard_normal(n * n))).astype(np.float32)
    plan = Plan(ctx, mesh, E, NU, 0.1, mask, element_t=et)
    op = O.OracleProblem(mesh, E, NU, 0.1, mask, element_t=et)
    F = plan.pressure_load(1e4, np.array([[3e4, 4.0, 6.0, 1.5]]))
    x64, _, st = plan.solve(F, tolerance=1e-8, max_iterations=20000, precond=precond)
    assert st["converged"]
    exact = op.exact_rel_residual(x64, F.astype(np.float64))
    print("device", st["trueRelResidual"], "exact", exact)
    assert exact <= 1.0001e-8
    assert abs(st["trueRelResidual"] - exact) < 1e-10, (st["trueRelResidual"], exact)


def test_solve_c2_matches_golden(ctx):
    """61,347 DOF: FP64 solve vs the FP64 oracle and the float32 solver.ts
    restatement (golden vectors)."""
    g = np.load(os.path.join(ROOT, "tests", "golden", "c2_142.npz"))
    w = P.workload("c2")
    plan = Plan(ctx, w["mesh"], w["E"], w["nu"], w["t"], w["mask"])
    F = plan.pressure_load(w["q0"], w["patches"])
    np.testing.assert_allclose(F, g["F"], rtol=2e-7, atol=0)
    x64, x32, st = plan.solve(g["F"], tolerance=1e-8, max_iterations=20000)
    assert st["converged"]
    assert st["trueRelResidual"] <= 1e-8, st
    assert O.max_nodal_w_error(x64, g["x64"]) < 1e-6
    assert O.compare_max_w(x32, g["x_js"]) < 1e-4
    assert O.max_nodal_w_error(x32, g["x_js"]) < 1e-3
    op = O.OracleProblem(w["mesh"], w["E"], w["nu"], w["t"], w["mask"])
    assert op.true_rel_residual(x64, g["F"].astype(np.float64)) <= 1e-8


def test_dkt_triangles_coloured_path(ctx):
    w = P.workload("c1")
    w["mesh"] = P.rect_mesh(16)
    w["mask"] = P.constrained_mask(w["mesh"], [{"type": "fixed", "location": "all_edges"}])
    plan, op, mesh, col = make(ctx, w, triangles=True)
    assert not plan.structured
    x = np.random.default_rng(4).standard_normal(plan.n_dof)
    assert rel(plan.apply_k(x), op.apply_k(x)) < 1e-12
    F = plan.pressure_load(1e4)
    np.testing.assert_allclose(F, op.pressure_load(1e4).astype(np.float32), rtol=2e-7)
    x64, x32, st = plan.solve(F, tolerance=1e-8, max_iterations=5000)
    xo, so = op.pcg(F.astype(np.float64), tol=1e-8)
    assert st["converged"] and abs(st["iterations"] - so["iterations"]) <= 2
    assert O.max_nodal_w_error(x64, xo) < 1e-6


def test_edge_cases(ctx):
    w = P.workload("c1")
    plan, op, _, _ = make(ctx, w)
    for path in (_lib.PATH_STRUCTURED, _lib.PATH_COLOURED):
        x64, _, st = plan.solve(np.zeros(plan.n_dof, np.float32), path=path)
        assert st["converged"] and st["iterations"] == 0 and not x64.any()
        F = plan.pressure_load(1e4)
        x64, _, st = plan.solve(F, max_iterations=1, path=path)
        assert not st["converged"] and st["iterations"] == 1
        xo, so = op.pcg(F.astype(np.float64), max_iter=1)
        assert rel(x64, xo) < 1e-12
    # every DOF constrained: identity system
    mask = np.full(w["mesh"]["nodeCount"], 7, np.uint8)
    p2 = Plan(ctx, w["mesh"], E, NU, 0.1, mask, P.colour_structured(w["mesh"]))
    F = np.random.default_rng(5).standard_normal(p2.n_dof).astype(np.float32)
    x64, _, st = p2.solve(F)
    assert st["converged"] and rel(x64, F.astype(np.float64)) < 1e-12


@pytest.mark.parametrize("n", [200, 600])
def test_large_grids_match_oracle_iterates(ctx, n):
    """Plates too big for one XCD run the device-wide kernels (resident at
    200 x 200, strided over tiles at 600 x 600): their first 30 PCG iterates
    equal the FP64 oracle's."""
    mesh = P.rect_mesh(n)
    mask = P.constrained_mask(mesh, [{"type": "pinned", "location": "all_edges"}])
    plan = Plan(ctx, mesh, E, NU, 0.05, mask)
    op = O.OracleProblem(mesh, E, NU, 0.05, mask)
    F = plan.pressure_load(1e4, np.array([[3e4, 3.0, 6.0, 1.5]]))
    x64, _, st = plan.solve(F, max_iterations=30)
    assert st["iterations"] == 30 and not st["converged"]
    xo, so = op.pcg(F.astype(np.float64), max_iter=30)
    assert rel(x64, xo) < 1e-9


def test_general_mesh_device_setup_and_loads(ctx):
    """Coloured path on a non-grid Q4 mesh with per-element thickness: the
    element matrices and block inverses built on the device, and the
    pressure load integrated on the device, against the oracle."""
    from test_post_oracle import distorted
    mesh = distorted(P.rect_mesh(20, 14))
    col = P.colour_greedy(mesh)
    mask = P.constrained_mask(mesh, [{"type": "pinned", "location": "all_edges"}])
    et = np.random.default_rng(2).uniform(0.05, 0.2, mesh["elementCount"]).astype(np.float32)
    plan = Plan(ctx, mesh, E, NU, 0.1, mask, col, element_t=et)
    assert not plan.structured
    op = O.OracleProblem(mesh, E, NU, 0.1, mask, element_t=et,
                         order=col["elements_by_colour"])
    b = plan.block_inverse().reshape(-1, 9)
    ob = op.blk.reshape(-1, 9)
    assert np.max(np.abs(b - ob) / np.abs(ob).max(axis=1, keepdims=True)) < 1e-10
    x = np.random.default_rng(3).standard_normal(plan.n_dof)
    assert rel(plan.apply_k(x), op.apply_k(x)) < 1e-12
    pt = np.array([[4e4, 3.0, 7.0, 1.2], [2e4, 6.5, 2.5, 0.8]])
    F = plan.pressure_load(1e4, pt)
    np.testing.assert_allclose(F, op.pressure_load(1e4, pt).astype(np.float32), rtol=2e-7,
                               atol=0)
    plan.close()


def test_rectangular_plate_full_solve(ctx):
    """nx != ny and dx != dy (90 x 37 elements on 10 m x 4 m): the resident
    single-reduction kernel, the mixed path and the moments against the
    oracle."""
    mesh = P.rect_mesh(90, 37, width=10.0, height=4.0)
    mask = P.constrained_mask(mesh, [{"type": "fixed", "location": "all_edges"}])
    plan = Plan(ctx, mesh, E, NU, 0.08, mask)
    assert plan.structured
    op = O.OracleProblem(mesh, E, NU, 0.08, mask)
    F = plan.pressure_load(1e4, np.array([[3e4, 7.0, 1.5, 0.8]]))
    np.testing.assert_allclose(F, op.pressure_load(1e4, np.array([[3e4, 7.0, 1.5, 0.8]])).astype(np.float32),
                               rtol=2e-7, atol=0)
    x64, _, st = plan.solve(F, tolerance=1e-8, max_iterations=20000)
    xo, so = op.pcg(F.astype(np.float64), tol=1e-8)
    assert st["converged"] and st["trueRelResidual"] <= 1e-8
    assert abs(st["iterations"] - so["iterations"]) <= 3 + so["iterations"] // 100
    assert O.max_nodal_w_error(x64, xo) < 1e-6
    xm, _, sm = plan.solve(F, tolerance=1e-8, max_iterations=200000, precision=_lib.PRECISION_MIXED)
    assert sm["converged"] and sm["trueRelResidual"] <= 1e-8
    assert O.max_nodal_w_error(xm, xo) < 1e-6
    for d, r in zip(plan.moments(x64), op.moments(x64)):
        np.testing.assert_allclose(d, r, rtol=2e-7, atol=2e-7 * np.abs(r).max())
    plan.close()


def test_rectangular_plate_strided_iterates(ctx):
    """700 x 300 elements on 10 m x 4.3 m (too large to be resident): the
    strided kernel's first 30 iterates equal the oracle's."""
    mesh = P.rect_mesh(700, 300, width=10.0, height=4.3)
    mask = P.constrained_mask(mesh, [{"type": "pinned", "location": "all_edges"}])
    plan = Plan(ctx, mesh, E, NU, 0.05, mask)
    op = O.OracleProblem(mesh, E, NU, 0.05, mask)
    F = plan.pressure_load(1e4, np.array([[3e4, 3.0, 2.0, 1.0]]))
    x64, _, st = plan.solve(F, max_iterations=30)
    assert st["iterations"] == 30 and not st["converged"]
    xo, so = op.pcg(F.astype(np.float64), max_iter=30)
    assert rel(x64, xo) < 1e-9
    plan.close()


def test_malformed_colour_offsets_are_rejected():
    """Colour offsets must start at 0, never decrease and end at the element
    count; anything else is refused before the offsets index anything."""
    ctx = Context(0)
    tri = P.split_quads(P.rect_mesh(8))
    mask = P.constrained_mask(tri, [{"type": "fixed", "location": "all_edges"}])
    good = P.colour_greedy(tri)
    ne = int(tri["elementCount"])
    for offs in ([0, 5000, 100] + [ne] * (good["colour_count"] - 2),
                 [1] + list(good["colour_offsets"][1:]),
                 list(good["colour_offsets"][:-1]) + [ne - 1]):
        bad = dict(good, colour_offsets=np.array(offs, np.uint32))
        with pytest.raises(PlatePCGError) as e:
            Plan(ctx, tri, 210e9, 0.3, 0.1, mask, bad)
        assert e.value.code == _lib.PP_ERR_INVALID
    Plan(ctx, tri, 210e9, 0.3, 0.1, mask, good).close()
    ctx.close()


@pytest.mark.parametrize("balance", [False, True])
def test_unstructured_triangulation_coloured_path(balance):
    """A genuinely unstructured mesh on the coloured path: a jittered Delaunay
    triangulation (irregular node valence, greedy colouring of
    coloring.ts:92-177 with 8+ colours, node -> element lists of every
    length).  K.x and each colour's apply equal the oracle's on the
    triangulator's own vertex order, where the triangle as written leaves
    some nodal blocks singular; there the block inverses agree wherever the
    nodal block is well conditioned.  With vertex orders rotated so that no
    node is only ever a third vertex, the block inverses, pressure load,
    moments, first 12 PCG iterates and the converged solve all equal the
    oracle's."""
    ctx = Context(0)
    mesh = O.unstructured_triangles(20, seed=11, jitter=0.3, balance=balance)
    mask = P.constrained_mask(mesh, [{"type": "fixed", "location": "all_edges"}])
    col = P.colour_greedy(mesh)
    plan = Plan(ctx, mesh, E, NU, 0.1, mask, col)
    assert not plan.structured and col["colour_count"] >= 8
    op = O.OracleProblem(mesh, E, NU, 0.1, mask, order=col["elements_by_colour"])
    rng = np.random.default_rng(12)
    x = rng.standard_normal(plan.n_dof)
    assert rel(plan.apply_k(x), op.apply_k(x)) < 1e-12
    Ke = op.Ke.reshape(-1, 9, 9)
    el = mesh["elements"].reshape(-1, 3).astype(np.int64)
    off = col["colour_offsets"]
    for c in range(col["colour_count"]):
        y = np.zeros(plan.n_dof)
        for e in col["elements_by_colour"][off[c]:off[c + 1]]:
            dofs = (el[e][:, None] * 3 + np.arange(3)).ravel()
            ye = Ke[e] @ x[dofs]
            keep = op.cdof[dofs] == 0
            y[dofs[keep]] += ye[keep]
        assert rel(plan.apply_k_colour(c, x), y) < 1e-12
    # nodal blocks before inversion, to tell the well-conditioned ones apart
    raw = np.zeros((mesh["nodeCount"], 3, 3))
    for i in range(3):
        np.add.at(raw, el[:, i], Ke[:, 3 * i:3 * i + 3, 3 * i:3 * i + 3])
    lam = np.linalg.eigvalsh(raw)
    good = (lam[:, 0] > 1e-9 * lam[:, 2]) | (mask != 0)
    assert good.all() == balance
    b = plan.block_inverse().reshape(-1, 9)[good]
    ob = op.blk.reshape(-1, 9)[good]
    assert np.max(np.abs(b - ob) / np.abs(ob).max(axis=1, keepdims=True)) < 1e-10
    if not balance:
        plan.close()
        ctx.close()
        return
    F = plan.pressure_load(1e4, np.array([[3e4, 4.0, 6.0, 1.5]]))
    Fo = op.pressure_load(1e4, np.array([[3e4, 4.0, 6.0, 1.5]]))
    np.testing.assert_allclose(F, Fo.astype(np.float32), rtol=2e-7, atol=0)
    for k in (1, 4, 12):
        xk, _, st = plan.solve(F, max_iterations=k)
        xo, so = op.pcg(F.astype(np.float64), max_iter=k)
        assert st["iterations"] == so["iterations"] == k
        assert rel(xk, xo) < 1e-9, k
    x64, _, st = plan.solve(F, tolerance=1e-8, max_iterations=5000)
    xo, so = op.pcg(F.astype(np.float64), tol=1e-8)
    assert st["converged"] and so["converged"]
    assert abs(st["iterations"] - so["iterations"]) <= 2
    assert op.true_rel_residual(x64, F.astype(np.float64)) <= 1e-8
    assert O.max_nodal_w_error(x64, xo) < 1e-6
    for got, want in zip(plan.moments(x64), op.moments(x64)):
        np.testing.assert_allclose(got, want, rtol=2e-7, atol=2e-7 * np.abs(want).max())
    plan.close()
    ctx.close()


def test_plate_with_holes_from_the_mesher():
    """A plate with a circular and a quadrilateral hole, meshed by the Node
    host's generateUnstructuredMesh (constrained Delaunay, Steiner lattice
    of mesher-unstructured.ts:97-146), greedy-coloured (coloring.ts:92-177),
    clamped on every boundary node (the outer edge and the holes' edges, as
    the reference's boundaryNodes give them), solved on the coloured DKT
    path: K.x, the pressure load, the converged solve and the moments equal
    the oracle's."""
    from test_mesher import circle, js_mesh, sampled_rect

    boundary = sampled_rect(0.0, 0.0, 10.0, 6.0, 0.5)
    holes = [circle(3.0, 3.0, 1.2, 40),
             np.array([6.5, 2.0, 8.5, 2.6, 8.0, 4.5, 6.2, 4.0], np.float32)]
    mesh = js_mesh(boundary, holes, 0.4)
    ctx = Context(0)
    mask = P.constrained_mask(mesh, [{"type": "fixed", "location": "all_edges"}])
    col = P.colour_greedy(mesh)
    plan = Plan(ctx, mesh, E, NU, 0.05, mask, col)
    assert not plan.structured
    op = O.OracleProblem(mesh, E, NU, 0.05, mask, order=col["elements_by_colour"])
    rng = np.random.default_rng(5)
    x = rng.standard_normal(plan.n_dof)
    assert rel(plan.apply_k(x), op.apply_k(x)) < 1e-12
    patch = np.array([[2e4, 5.0, 1.5, 1.0]])
    F = plan.pressure_load(1e4, patch)
    np.testing.assert_allclose(F, op.pressure_load(1e4, patch).astype(np.float32), rtol=2e-7,
                               atol=0)
    x64, _, st = plan.solve(F, tolerance=1e-8, max_iterations=20000)
    xo, so = op.pcg(F.astype(np.float64), tol=1e-8)
    print("mesh", mesh["nodeCount"], "nodes", mesh["elementCount"], "triangles,",
          col["colour_count"], "colours; iterations", st["iterations"], so["iterations"])
    assert st["converged"] and so["converged"]
    assert abs(st["iterations"] - so["iterations"]) <= 2
    assert op.true_rel_residual(x64, F.astype(np.float64)) <= 1e-8
    assert O.max_nodal_w_error(x64, xo) < 1e-6
    for got, want in zip(plan.moments(x64), op.moments(x64)):
        np.testing.assert_allclose(got, want, rtol=2e-7, atol=2e-7 * np.abs(want).max())
    plan.close()
    ctx.close()


@pytest.mark.parametrize("case", ["q4", "dkt", "unstructured", "dkt_strided"])
def test_coloured_persistent_kernel_equals_launch_loop(ctx, case):
    """The coloured path's persistent PCG kernel (K x gathered per node in
    colour order, one grid barrier and one all-reduce per iteration) against
    the launch-per-operation loop (a launch per colour, device scalars,
    PP_FLAG_COLOURED_HOST_LOOP): the same first iterates and the same
    converged answer, and both against the oracle.  dkt_strided has more
    nodes (160,801) than one pass of the grid covers, so every workgroup
    strides over several node blocks."""
    loop = _lib.FLAG_COLOURED_HOST_LOOP
    if case == "unstructured":
        mesh = O.unstructured_triangles(20, seed=11, jitter=0.3, balance=True)
    else:
        mesh = P.rect_mesh({"q4": 32, "dkt": 64, "dkt_strided": 400}[case])
        if case != "q4":
            mesh = P.split_quads(mesh)
    mask = P.constrained_mask(mesh, [{"type": "fixed", "location": "all_edges"}])
    col = P.colour_greedy(mesh) if case != "q4" else P.colour_structured(mesh)
    plan = Plan(ctx, mesh, E, NU, 0.1, mask, col)
    path = _lib.PATH_COLOURED
    F = plan.pressure_load(1e4, np.array([[3e4, 4.0, 6.0, 1.5]]))
    for k in (1, 5, 40):
        xa, _, sa = plan.solve(F, tolerance=1e-30, max_iterations=k, path=path,
                               flags=_lib.FLAG_NO_TRUE_RESIDUAL)
        xb, _, sb = plan.solve(F, tolerance=1e-30, max_iterations=k, path=path,
                               flags=loop | _lib.FLAG_NO_TRUE_RESIDUAL)
        assert sa["iterations"] == sb["iterations"] == k
        assert rel(xa, xb) < 1e-9, (k, rel(xa, xb))
    if case == "dkt_strided":
        plan.close()
        return
    xa, _, sa = plan.solve(F, tolerance=1e-8, max_iterations=20000, path=path)
    xb, _, sb = plan.solve(F, tolerance=1e-8, max_iterations=20000, path=path, flags=loop)
    print(case, mesh["nodeCount"], "nodes: iterations", sa["iterations"], sb["iterations"],
          "ms", sa["deviceMs"], sb["deviceMs"])
    assert sa["converged"] and sb["converged"]
    assert abs(sa["iterations"] - sb["iterations"]) <= 2
    assert sa["trueRelResidual"] <= 1e-8
    assert O.max_nodal_w_error(xa, xb) < 1e-6
    op = O.OracleProblem(mesh, E, NU, 0.1, mask, order=col["elements_by_colour"])
    xo, so = op.pcg(F.astype(np.float64), tol=1e-8)
    assert abs(sa["iterations"] - so["iterations"]) <= 2
    assert O.max_nodal_w_error(xa, xo) < 1e-6
    plan.close()
