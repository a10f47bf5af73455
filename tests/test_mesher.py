"""The unstructured mesher of the Node host (gpu-gauntlet_amd/js/mesher.js,
generateUnstructuredMesh of mesher-unstructured.ts:36-78 with its own
constrained Delaunay triangulation in place of poly2tri), checked by what
fixes a constrained Delaunay triangulation, in exact rational arithmetic
here, independently of the mesher's own predicates: every triangle
counter-clockwise, the areas summing to the domain's, every boundary and hole
segment an edge, no triangle outside the boundary or inside a hole, no point
strictly inside the circumcircle across any edge that is not a segment, and
the node set the reference's Steiner lattice rule (mesher-unstructured.ts:
97-146) gives.  On a square the triangulation equals scipy's (Qhull)
Delaunay triangulation of the same points up to exactly cocircular ties.
CPU only."""
import json
import os
import subprocess
from fractions import Fraction

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MESHER = os.path.join(ROOT, "gpu-gauntlet_amd", "js", "mesher.js")


def js_mesh(boundary, holes, h, balance=True):
    """generateUnstructuredMesh through Node; arrays as numpy."""
    geo = {"boundary": [float(v) for v in boundary],
           "holes": [[float(v) for v in q] for q in holes]}
    code = (f"const M = require({json.dumps(MESHER)});"
            f"const G = {json.dumps(geo)};"
            "const m = M.generateUnstructuredMesh({boundary: new Float32Array(G.boundary),"
            " holes: G.holes.map((a) => new Float32Array(a))},"
            f" {{targetEdgeLength: {h}, balanceVertexOrder: {json.dumps(bool(balance))}}});"
            "console.log(JSON.stringify({nodes: Array.from(m.nodes), elements: Array.from(m.elements),"
            " boundaryNodes: Array.from(m.boundaryNodes), nodeCount: m.nodeCount,"
            " elementCount: m.elementCount, nodesPerElement: m.nodesPerElement}));")
    out = subprocess.run(["node", "-e", code], check=True, capture_output=True, text=True,
                         timeout=300).stdout
    m = json.loads(out)
    return {"nodes": np.array(m["nodes"], np.float32), "elements": np.array(m["elements"], np.uint32),
            "boundaryNodes": np.array(m["boundaryNodes"], np.uint32), "nodeCount": m["nodeCount"],
            "elementCount": m["elementCount"], "nodesPerElement": m["nodesPerElement"],
            "gridNx": 0, "gridNy": 0}


def circle(cx, cy, r, n):
    t = 2 * np.pi * np.arange(n) / n
    return np.stack([cx + r * np.cos(t), cy + r * np.sin(t)], 1).astype(np.float32).ravel()


def sampled_rect(x0, y0, x1, y1, h):
    """Counter-clockwise rectangle with vertices every h along its sides."""
    nx, ny = int(round((x1 - x0) / h)), int(round((y1 - y0) / h))
    xs, ys = np.linspace(x0, x1, nx + 1), np.linspace(y0, y1, ny + 1)
    pts = [(x, y0) for x in xs[:-1]] + [(x1, y) for y in ys[:-1]] + \
        [(x, y1) for x in xs[:0:-1]] + [(x0, y) for y in ys[:0:-1]]
    return np.array(pts, np.float32).ravel()


def exact(p):
    return [(Fraction(float(x)), Fraction(float(y))) for x, y in p]


def orient(a, b, c):
    return (a[0] - c[0]) * (b[1] - c[1]) - (a[1] - c[1]) * (b[0] - c[0])


def incircle(a, b, c, d):
    ax, ay, bx, by = a[0] - d[0], a[1] - d[1], b[0] - d[0], b[1] - d[1]
    cx, cy = c[0] - d[0], c[1] - d[1]
    return ((ax * ax + ay * ay) * (bx * cy - cx * by) + (bx * bx + by * by) * (cx * ay - ax * cy)
            + (cx * cx + cy * cy) * (ax * by - bx * ay))


def inside(x, y, poly):
    """Even-odd ray crossing (mesher-utils.ts:145-165)."""
    q = np.asarray(poly, np.float64).reshape(-1, 2)
    c = False
    j = len(q) - 1
    for i in range(len(q)):
        (xi, yi), (xj, yj) = q[i], q[j]
        if (yi > y) != (yj > y) and x < (xj - xi) * (y - yi) / (yj - yi) + xi:
            c = not c
        j = i
    return c


def on_segment(px, py, x1, y1, x2, y2, tol):
    dx, dy = x2 - x1, y2 - y1
    ln = np.hypot(dx, dy)
    if ln < tol:
        return False
    t = ((px - x1) * dx + (py - y1) * dy) / (ln * ln)
    if t < -tol / ln or t > 1 + tol / ln:
        return False
    return np.hypot(px - (x1 + t * dx), py - (y1 + t * dy)) < tol


def on_polygon(x, y, poly, tol):
    q = np.asarray(poly, np.float64).reshape(-1, 2)
    return any(on_segment(x, y, *q[i], *q[(i + 1) % len(q)], tol) for i in range(len(q)))


def steiner(boundary, holes, h):
    """The reference's lattice rule, restated: spacing h, odd rows shifted by
    h / 2, inside the boundary, outside the holes, h / 10 clear of segments."""
    b = np.asarray(boundary, np.float64).reshape(-1, 2)
    mnx, mny, mxx, mxy = b[:, 0].min(), b[:, 1].min(), b[:, 0].max(), b[:, 1].max()
    m = 0.1 * h
    out = []
    for j in range(1, int(np.ceil((mxy - mny) / h))):
        off = 0.0 if j % 2 == 0 else 0.5 * h
        for i in range(1, int(np.ceil((mxx - mnx) / h))):
            x, y = mnx + i * h + off, mny + j * h
            if x < mnx + m or x > mxx - m or y < mny + m or y > mxy - m:
                continue
            if not inside(x, y, boundary) or any(inside(x, y, q) for q in holes):
                continue
            if on_polygon(x, y, boundary, m) or any(on_polygon(x, y, q, m) for q in holes):
                continue
            out.append((np.float32(x), np.float32(y)))
    return out


def polygon_area(poly):
    q = np.asarray(poly, np.float64).reshape(-1, 2)
    return 0.5 * abs(np.sum(q[:, 0] * np.roll(q[:, 1], -1) - np.roll(q[:, 0], -1) * q[:, 1]))


def test_square_matches_delaunay_up_to_cocircular_ties():
    from scipy.spatial import Delaunay

    m = js_mesh([0, 0, 10, 0, 10, 10, 0, 10], [], 0.5, balance=False)
    pts = m["nodes"].reshape(-1, 2).astype(np.float64)
    mine = {tuple(sorted(t)) for t in m["elements"].reshape(-1, 3).tolist()}
    qhull = {tuple(sorted(t)) for t in Delaunay(pts).simplices.tolist()}
    assert len(mine) == len(qhull) == m["elementCount"]
    diff = mine ^ qhull
    assert len(diff) <= 0.01 * len(mine)
    P = exact(pts)
    # every difference is the other diagonal of four exactly cocircular points
    for t in diff:
        quad = [s for s in diff if len(set(s) & set(t)) == 2]
        verts = sorted(set(t).union(*[set(s) for s in quad]))
        assert len(verts) == 4
        assert incircle(P[verts[0]], P[verts[1]], P[verts[2]], P[verts[3]]) == 0


def test_plate_with_holes_is_constrained_delaunay():
    h = 0.4
    boundary = sampled_rect(0.0, 0.0, 10.0, 6.0, 0.5)[::-1].reshape(-1, 2)[:, ::-1].ravel()  # clockwise given
    holes = [circle(3.0, 3.0, 1.2, 40),
             np.array([6.5, 2.0, 8.5, 2.6, 8.0, 4.5, 6.2, 4.0], np.float32)]
    m = js_mesh(boundary, holes, h)
    nodes = m["nodes"].reshape(-1, 2)
    tri = m["elements"].reshape(-1, 3).astype(np.int64)
    P = exact(nodes)
    # counter-clockwise, no degenerate triangle
    assert all(orient(P[a], P[b], P[c]) > 0 for a, b, c in tri)
    # areas: the domain's (boundary minus holes)
    area = sum(orient(P[a], P[b], P[c]) for a, b, c in tri) / 2
    want = polygon_area(boundary) - sum(polygon_area(q) for q in holes)
    assert abs(float(area) - want) < 1e-9 * want
    # every boundary and hole segment is an edge; the other edges are
    # constrained-Delaunay: no vertex strictly inside a circumcircle across
    # them; every interior edge has two triangles
    index = {tuple(np.float32(v)): i for i, v in enumerate(nodes)}
    segs = set()
    for poly in [boundary] + holes:
        q = np.asarray(poly, np.float32).reshape(-1, 2)
        ids = [index[tuple(v)] for v in q]
        segs |= {frozenset((ids[i], ids[(i + 1) % len(ids)])) for i in range(len(ids))}
    opp = {}
    for a, b, c in tri:
        for u, w, x in ((a, b, c), (b, c, a), (c, a, b)):
            opp.setdefault(frozenset((u, w)), []).append((u, w, x))
    assert segs <= set(opp)
    for e, sides in opp.items():
        if e in segs:
            assert len(sides) == 1, e
            continue
        assert len(sides) == 2, e
        (u, w, x), (_, _, y) = sides
        assert incircle(P[u], P[w], P[x], P[y]) <= 0, e
    # no triangle outside the boundary or inside a hole
    for a, b, c in tri:
        gx, gy = nodes[[a, b, c]].astype(np.float64).mean(0)
        assert inside(gx, gy, boundary) and not any(inside(gx, gy, q) for q in holes)
    # nodes: the polygon vertices and the reference's Steiner points
    verts = {tuple(v) for poly in [boundary] + holes for v in np.asarray(poly, np.float32).reshape(-1, 2)}
    want_nodes = verts | set(steiner(boundary, holes, h))
    assert {tuple(v) for v in nodes} == want_nodes
    # boundary nodes: on a segment within 1e-6 (mesher-unstructured.ts:259-286)
    bn = [i for i, (x, y) in enumerate(nodes.astype(np.float64))
          if on_polygon(x, y, boundary, 1e-6) or any(on_polygon(x, y, q, 1e-6) for q in holes)]
    assert m["boundaryNodes"].tolist() == bn
    # no node is only ever the third vertex of its triangles
    firsts = set(tri[:, 0]) | set(tri[:, 1])
    assert firsts == set(range(m["nodeCount"]))


def test_vertex_order_balancing_keeps_the_triangles():
    boundary = [0, 0, 8, 0, 8, 8, 0, 8]
    holes = [circle(4.0, 4.0, 1.5, 32)]
    a = js_mesh(boundary, holes, 0.5, balance=False)
    b = js_mesh(boundary, holes, 0.5, balance=True)
    np.testing.assert_array_equal(a["nodes"], b["nodes"])
    ta, tb = a["elements"].reshape(-1, 3), b["elements"].reshape(-1, 3)
    # the same triangles, each rotated cyclically at most
    for s, t in zip(ta.tolist(), tb.tolist()):
        assert t in (s, s[1:] + s[:1], s[2:] + s[:2])
