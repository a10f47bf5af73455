"""Pins the oracle's nodal bending moments (postprocess.ts:40-270) with closed
forms: constant-curvature patch tests on rectangular, distorted-quad and
triangle meshes (linear rotation fields give the same exact moments at every
node), a quadratic field whose node averages are known in closed form (which
elements a node averages and how many), and the degenerate-triangle rule.
CPU only."""
import numpy as np
import pytest

import oracle_api as O
from platepcg import problem as P

E, NU, T = 210e9, 0.3, 0.1
D = E * T ** 3 / (12 * (1 - NU ** 2))


def distorted(mesh, frac=0.2, seed=3):
    """Interior nodes moved by up to frac of the spacing (quads stay convex)."""
    m = dict(mesh)
    nx, ny = mesh["gridNx"], mesh["gridNy"]
    xy = mesh["nodes"].reshape(-1, 2).astype(np.float64).copy()
    h = min(10.0 / nx, 10.0 / ny)
    rng = np.random.default_rng(seed)
    for j in range(1, ny):
        for i in range(1, nx):
            xy[j * (nx + 1) + i] += rng.uniform(-frac, frac, 2) * h
    m["nodes"] = xy.astype(np.float32).ravel()
    m["gridNx"] = m["gridNy"] = 0
    return m


def linear_field(mesh, a, b):
    """w arbitrary, thx = a0 + a1 x + a2 y, thy = b0 + b1 x + b2 y."""
    xy = mesh["nodes"].reshape(-1, 2).astype(np.float64)
    x = np.zeros(3 * len(xy))
    x[0::3] = np.sin(xy[:, 0]) * np.cos(xy[:, 1])
    x[1::3] = a[0] + a[1] * xy[:, 0] + a[2] * xy[:, 1]
    x[2::3] = b[0] + b[1] * xy[:, 0] + b[2] * xy[:, 1]
    return x


def oracle(mesh, element_t=None):
    return O.OracleProblem(mesh, E, NU, T, np.zeros(mesh["nodeCount"], np.uint8),
                           element_t=element_t)


MESHES = {
    "rect": lambda: P.rect_mesh(6, 4),
    "distorted_q4": lambda: distorted(P.rect_mesh(7, 5)),
    "triangles": lambda: P.split_quads(P.rect_mesh(5, 6)),
    "distorted_triangles": lambda: P.split_quads(distorted(P.rect_mesh(6, 6))),
}


@pytest.mark.parametrize("name", list(MESHES))
@pytest.mark.parametrize("ref32", [False, True])
def test_constant_curvature_patch(name, ref32):
    """k_x = d thy/dx = b1, k_y = -d thx/dy = -a2, k_xy = d thy/dy - d thx/dx
    = b2 - a1 exactly for a linear field on any Q4 or triangle; every node
    then carries M = D (k_x + nu k_y, k_y + nu k_x, (1 - nu) k_xy / 2)."""
    mesh = MESHES[name]()
    a, b = (0.3e-3, -1.1e-3, 2.3e-3), (-0.7e-3, 1.9e-3, 0.4e-3)
    kx, ky, kxy = b[1], -a[2], b[2] - a[1]
    want = (D * (kx + NU * ky), D * (ky + NU * kx), D * (1 - NU) * kxy / 2)
    got = oracle(mesh).moments(linear_field(mesh, a, b), ref32=ref32)
    tol = 2e-5 if ref32 else 1e-10
    for g, w in zip(got, want):
        np.testing.assert_allclose(g, w, rtol=tol, atol=tol * abs(w))


def test_quadratic_field_node_averages():
    """thy = x^2 on a uniform nx x ny grid (h = 10/nx): an element's k_x at
    its centroid is (x1^2 - x0^2)/h = x0 + x1 = 2 x_c, so M_x = 2 D x_c,
    M_y = 2 nu D x_c, M_xy = 0 per element.  A node at x_i averages the
    elements on both sides: 2 x_i inside, h on the left edge, 2L - h on the
    right edge -- which pins the node-to-element averaging and the counts."""
    nx, ny = 8, 3
    mesh = P.rect_mesh(nx, ny)
    xy = mesh["nodes"].reshape(-1, 2).astype(np.float64)
    x = np.zeros(3 * len(xy))
    x[2::3] = xy[:, 0] ** 2
    Mx, My, Mxy = oracle(mesh).moments(x)
    L, h = 10.0, 10.0 / nx
    i = np.tile(np.arange(nx + 1), ny + 1)
    xs = xy[:, 0]
    k = np.where(i == 0, h, np.where(i == nx, 2 * L - h, 2 * xs))
    np.testing.assert_allclose(Mx, D * k, rtol=1e-12)
    np.testing.assert_allclose(My, NU * D * k, rtol=1e-12)
    assert np.abs(Mxy).max() <= 1e-12 * np.abs(Mx).max()


def test_per_element_thickness_scales_by_t_cubed():
    mesh = P.rect_mesh(4, 4)
    x = linear_field(mesh, (0, 0, -1e-3), (0, 2e-3, 0))
    et = np.full(16, 0.2)
    a = oracle(mesh).moments(x)
    b = oracle(mesh, element_t=et).moments(x)
    for u, v in zip(a, b):
        np.testing.assert_allclose(v, u * (0.2 / T) ** 3, rtol=1e-12, atol=1e-9)


def test_degenerate_triangle_is_skipped():
    """postprocess.ts:202: a triangle with |2A| < 1e-15 adds nothing and is
    not counted; a node that only it touches keeps 0."""
    nodes = np.array([0, 0, 1, 0, 0, 1, 2, 0], np.float32)
    mesh = {"nodes": nodes, "elements": np.array([0, 1, 2, 0, 1, 3], np.uint32),
            "nodeCount": 4, "elementCount": 2, "nodesPerElement": 3}
    # thx = 3x + 4y, thy = x + 2y: k_x = 1, k_y = -4, k_xy = 2 - 3 = -1
    x = np.zeros(12)
    xy = nodes.reshape(-1, 2).astype(np.float64)
    x[1::3] = 3 * xy[:, 0] + 4 * xy[:, 1]
    x[2::3] = xy[:, 0] + 2 * xy[:, 1]
    Mx, My, Mxy = O.OracleProblem(mesh, E, NU, T, np.zeros(4, np.uint8)).moments(x)
    want = (D * (1 - 4 * NU), D * (-4 + NU), D * (1 - NU) * -1 / 2)
    for g, w in zip((Mx, My, Mxy), want):
        np.testing.assert_allclose(g[:3], w, rtol=1e-12)
        assert g[3] == 0.0


def test_reference_precision_close_to_fp64():
    """The float32 storage of the reference (ref32) moves the moments of a
    solved plate by far less than the 1e-4 display tolerance."""
    w = P.workload("c1")
    op = O.OracleProblem(w["mesh"], w["E"], w["nu"], w["t"], w["mask"])
    F = op.pressure_load(w["q0"])
    x, _ = op.pcg(F, tol=1e-10)
    for a, b in zip(op.moments(x), op.moments(x, ref32=True)):
        assert np.abs(a - b).max() <= 1e-4 * np.abs(a).max()
