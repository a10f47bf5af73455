"""Test-side access to the CPU oracles under oracle/ (never used by the product).

- the FP64 C oracle (oracle/plate_oracle.c) through ctypes;
- the float32 JavaScript restatement of solver.ts (oracle/solver_f32.js)
  run in Node as a child process.
"""
import ctypes as C
import json
import os
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
LIB = os.path.join(ORACLE, "_build", "liborc.so")
JS = os.path.join(ORACLE, "solver_f32.js")

_lib = None


class PCGResult(C.Structure):
    _fields_ = [("iterations", C.c_int64), ("final_residual", C.c_double), ("converged", C.c_int),
                ("breakdown", C.c_int)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.check_call(["make", "-s", "-C", ORACLE])
        _lib = C.CDLL(LIB)
    return _lib


def _p(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def ke_q4(xy, E, nu, t):
    Ke = np.zeros(144)
    xy = np.ascontiguousarray(xy, np.float64)
    lib().orc_ke_q4(_p(xy), C.c_double(E), C.c_double(nu), C.c_double(t), _p(Ke))
    return Ke.reshape(12, 12)


def ke_dkt(xy, E, nu, t):
    Ke = np.zeros(81)
    xy = np.ascontiguousarray(xy, np.float64)
    lib().orc_ke_dkt(_p(xy), C.c_double(E), C.c_double(nu), C.c_double(t), _p(Ke))
    return Ke.reshape(9, 9)


class OracleProblem:
    """FP64 oracle state for one mesh/material/support set."""

    def __init__(self, mesh, E, nu, t, mask, element_t=None, order=None):
        self.nodes = np.ascontiguousarray(mesh["nodes"], np.float32)
        self.elems = np.ascontiguousarray(mesh["elements"], np.uint32)
        self.nn = int(mesh["nodeCount"])
        self.ne = int(mesh["elementCount"])
        self.npe = int(mesh.get("nodesPerElement", 4))
        self.nd = 3 * self.npe
        self.E, self.nu, self.t = E, nu, t
        self.et = None if element_t is None else np.ascontiguousarray(element_t, np.float64)
        self.order = None if order is None else np.ascontiguousarray(order, np.uint32)
        self.cdof = np.zeros(3 * self.nn, np.uint8)
        for d in range(3):
            self.cdof[d::3] = (np.asarray(mask) >> d) & 1
        self.Ke = np.zeros(self.ne * self.nd * self.nd)
        lib().orc_all_ke(_p(self.nodes), _p(self.elems), C.c_int64(self.nn), C.c_int64(self.ne),
                         C.c_int(self.npe), C.c_double(E), C.c_double(nu), C.c_double(t),
                         _p(self.et), _p(self.Ke))
        self.blk = np.zeros(9 * self.nn)
        lib().orc_block_inverse(_p(self.Ke), _p(self.elems), C.c_int64(self.nn), C.c_int64(self.ne),
                                C.c_int(self.npe), _p(self.cdof), _p(self.blk))

    def apply_k(self, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(3 * self.nn)
        lib().orc_apply_k(_p(self.Ke), _p(self.elems), C.c_int64(self.nn), C.c_int64(self.ne),
                          C.c_int(self.npe), _p(self.order), _p(self.cdof), _p(x), _p(y))
        return y

    def precond(self, r):
        r = np.ascontiguousarray(r, np.float64)
        z = np.zeros(3 * self.nn)
        lib().orc_apply_block(_p(self.blk), C.c_int64(self.nn), _p(r), _p(z))
        return z

    def pcg(self, b, tol=1e-8, max_iter=100000):
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros(3 * self.nn)
        res = PCGResult()
        lib().orc_pcg(_p(self.Ke), _p(self.elems), C.c_int64(self.nn), C.c_int64(self.ne),
                      C.c_int(self.npe), _p(self.order), _p(self.cdof), _p(self.blk), _p(b),
                      C.c_double(tol), C.c_int64(max_iter), _p(x), C.byref(res))
        return x, {"iterations": res.iterations, "finalResidual": res.final_residual,
                   "converged": bool(res.converged), "breakdown": bool(res.breakdown)}

    def pressure_load(self, q0, patches=None):
        pt = np.ascontiguousarray(np.zeros((0, 4)) if patches is None else patches, np.float64)
        F = np.zeros(3 * self.nn)
        lib().orc_pressure_load(_p(self.nodes), _p(self.elems), C.c_int64(self.nn),
                                C.c_int64(self.ne), C.c_int(self.npe), C.c_double(q0),
                                _p(pt) if pt.size else None, C.c_int(pt.shape[0]), _p(self.cdof),
                                _p(F))
        return F

    def moments(self, x, ref32=False):
        """Nodal (Mx, My, Mxy) of x by postprocess.ts (ref32: in its float32
        storage as well)."""
        x = np.ascontiguousarray(x, np.float64)
        out = [np.zeros(self.nn) for _ in range(3)]
        lib().orc_moments(_p(self.nodes), _p(self.elems), C.c_int64(self.nn), C.c_int64(self.ne),
                          C.c_int(self.npe), C.c_double(self.E), C.c_double(self.nu),
                          C.c_double(self.t), _p(self.et), _p(x), C.c_int(int(ref32)),
                          *[_p(a) for a in out])
        return tuple(out)

    def true_rel_residual(self, x, F):
        return float(np.linalg.norm(F - self.apply_k(x)) / np.linalg.norm(F))

    def exact_rel_residual(self, x, F):
        """|F - K x| / |F| with every product and sum exact (rationals): the
        FP64 element matrices, x and F taken as the exact values they hold,
        the constraint rows as in apply_k (y = x on constrained DOFs).  Only
        the final square root and division round."""
        from fractions import Fraction
        nd = self.nd
        Ke = self.Ke.reshape(self.ne, nd, nd)
        el = self.elems.reshape(self.ne, self.npe)
        xf = [Fraction(float(v)) for v in np.asarray(x, np.float64)]
        y = [Fraction(0)] * (3 * self.nn)
        for e in range(self.ne):
            dof = [3 * int(n) + d for n in el[e] for d in range(3)]
            xl = [xf[j] for j in dof]
            K = [[Fraction(float(v)) for v in row] for row in Ke[e]]
            for a in range(nd):
                y[dof[a]] += sum(K[a][b] * xl[b] for b in range(nd))
        num = Fraction(0)
        den = Fraction(0)
        for i, f in enumerate(np.asarray(F, np.float64)):
            fi = Fraction(float(f))
            ri = fi - (xf[i] if self.cdof[i] else y[i])
            num += ri * ri
            den += fi * fi
        return float(np.sqrt(float(num / den)))


def test_system():
    x = np.zeros(10)
    it = lib().orc_test_system(_p(x))
    return x, it


def run_js_solver(nx, F32, E, nu, t, supports, tol=1e-8, max_iter=20000, triangles=False,
                  width=10.0):
    """Runs oracle/solver_f32.js on a rectangular nx x nx plate with load F32."""
    with tempfile.TemporaryDirectory() as d:
        fp, op, sp = (os.path.join(d, n) for n in ("F.bin", "x.bin", "spec.json"))
        np.ascontiguousarray(F32, np.float32).tofile(fp)
        spec = dict(nx=nx, ny=nx, width=width, height=width, E=E, nu=nu, t=t, triangles=triangles,
                    supports=supports, F_path=fp, tolerance=tol, maxIterations=max_iter,
                    out_path=op)
        with open(sp, "w") as f:
            json.dump(spec, f)
        out = subprocess.run(["node", JS, sp], check=True, capture_output=True, text=True)
        stats = json.loads(out.stdout.strip().splitlines()[-1])
        x = np.fromfile(op, np.float32)
    return x, stats


def run_js(code):
    """Evaluates a Node snippet with the oracle module bound to `O`; returns parsed JSON."""
    prog = f"const O = require({json.dumps(JS)});\n" + code
    out = subprocess.run(["node", "-e", prog], check=True, capture_output=True, text=True)
    return json.loads(out.stdout)


def compare_max_w(x_test, x_ref):
    """The harness metric (benchmark/page.tsx:121-168): signed w at the node of
    largest |w| in each solution, relative error."""
    wt, wr = np.asarray(x_test)[0::3], np.asarray(x_ref)[0::3]
    a, b = wt[np.argmax(np.abs(wt))], wr[np.argmax(np.abs(wr))]
    return abs(float(a) - float(b)) / abs(float(b))


def max_nodal_w_error(x_test, x_ref, floor=0.01):
    """max over nodes with |w_r| > floor * max|w_r| of |w_t - w_r| / |w_r|."""
    wt, wr = np.asarray(x_test, np.float64)[0::3], np.asarray(x_ref, np.float64)[0::3]
    sel = np.abs(wr) > floor * np.abs(wr).max()
    return float(np.max(np.abs(wt[sel] - wr[sel]) / np.abs(wr[sel])))


def unstructured_triangles(n, seed, width=10.0, jitter=0.35, balance=False):
    """A genuinely unstructured triangle mesh of the width x width square for
    the coloured path: n + 1 points per edge, interior points on an n x n grid
    each moved by up to `jitter` of the spacing in both directions, Delaunay-
    triangulated (scipy), every triangle counter-clockwise.  boundaryNodes are
    the points on the edges (their order does not matter: solver.ts:92-134
    puts them in a Set).

    The triangle of element.ts:642-943 as written gives its THIRD vertex a
    rank-2 diagonal 3x3 block, so a node that is the third vertex of every
    triangle around it has a singular nodal block and the assembled stiffness
    a zero-energy mode there.  balance=True rotates vertex orders (cyclically,
    so every triangle stays counter-clockwise) until each node is the first or
    second vertex of at least one triangle; the mesh is otherwise the same."""
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(seed)
    h = width / n
    s = np.arange(n + 1) * h
    edge = [(x, 0.0) for x in s] + [(width, y) for y in s[1:]] + \
        [(x, width) for x in s[-2::-1]] + [(0.0, y) for y in s[-2:0:-1]]
    gx, gy = np.meshgrid(s[1:-1], s[1:-1])
    inner = np.stack([gx.ravel(), gy.ravel()], 1)
    inner = inner + rng.uniform(-jitter * h, jitter * h, inner.shape)
    pts = np.vstack([np.array(edge), inner]).astype(np.float32)
    tri = Delaunay(pts.astype(np.float64)).simplices.astype(np.int64)
    p = pts.astype(np.float64)[tri]
    cross = (p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1]) - \
        (p[:, 1, 1] - p[:, 0, 1]) * (p[:, 2, 0] - p[:, 0, 0])
    tri[cross < 0] = tri[cross < 0][:, [0, 2, 1]]
    if balance:
        tri = _balance_third_vertices(tri, len(pts))
    return {"nodes": pts.ravel(), "elements": tri.astype(np.uint32).ravel(),
            "boundaryNodes": np.arange(len(edge), dtype=np.uint32), "nodeCount": len(pts),
            "elementCount": len(tri), "nodesPerElement": 3, "gridNx": 0, "gridNy": 0}


def _balance_third_vertices(tri, n_nodes):
    """Cyclic vertex rotations so that every node is the first or second
    vertex of some triangle (greedy, deterministic)."""
    tri = tri.copy()
    inc = [[] for _ in range(n_nodes)]
    for e, t in enumerate(tri):
        for v in t:
            inc[v].append(e)

    def covered(v):
        return any(tri[e][2] != v for e in inc[v])

    for _ in range(100):
        bare = [v for v in range(n_nodes) if not covered(v)]
        if not bare:
            return tri
        for v in bare:
            if covered(v):
                continue
            # rotate v to the front of a triangle around it, preferring one
            # whose vertex that becomes third is still covered elsewhere
            for e in inc[v]:
                k = list(tri[e]).index(v)
                new = np.roll(tri[e], -k)
                old = tri[e].copy()
                tri[e] = new
                if covered(new[2]):
                    break
                tri[e] = old
            else:
                e = inc[v][0]
                tri[e] = np.roll(tri[e], -list(tri[e]).index(v))
    raise RuntimeError("could not balance third vertices")
