'use strict';
// Unstructured triangle meshes of polygonal plates with holes, for the
// coloured (DKT) path: generateUnstructuredMesh keeps the call shape and the
// PlateMesh result of the reference's mesher-unstructured.ts:36-78 (outer
// boundary and holes as flat [x, y, ...] arrays, optional Steiner points on a
// staggered lattice of the target edge length, mesher-unstructured.ts:97-146),
// with a constrained Delaunay triangulation of this file's own in place of
// the poly2tri sweep the reference calls:
//
//   1. Bowyer-Watson insertion of every point into a triangle that encloses
//      them (visibility walk to the containing triangle, cavity of the
//      triangles whose circumcircle holds the point, fan from the point);
//   2. each boundary and hole segment recovered by flipping the edges it
//      crosses (Sloan's method), then Lawson flips restore the Delaunay
//      property of every edge that is not a segment;
//   3. the triangles outside the boundary or inside a hole removed.
//
// Orientation and in-circle tests are evaluated in double precision with a
// forward error bound and, when the bound does not settle the sign, exactly
// in BigInt arithmetic on the coordinates' binary expansions, so nearly
// cocircular points (polygonal circles) cannot corrupt the triangulation.
// Steiner points are rounded to float32 before triangulating, so the stored
// Float32Array mesh is exactly the one triangulated.

// --------------------------------------------------------------- utilities
// (the polygon helpers of mesher-utils.ts:22-211, restated)

function signedArea2(v) {
  let a = 0;
  const n = v.length / 2;
  for (let i = 0; i < n; i++) {
    const j = (i + 1) % n;
    a += v[2 * i] * v[2 * j + 1] - v[2 * j] * v[2 * i + 1];
  }
  return a;
}

// the polygon with the requested winding (clockwise or counter-clockwise)
function ensureWindingOrder(v, clockwise) {
  const ccw = signedArea2(v) > 0;
  if (clockwise !== ccw) return v;
  const n = v.length / 2;
  const out = new Float32Array(v.length);
  for (let i = 0; i < n; i++) {
    out[2 * i] = v[2 * (n - 1 - i)];
    out[2 * i + 1] = v[2 * (n - 1 - i) + 1];
  }
  return out;
}

// even-odd ray crossing
function isInsidePolygon(x, y, poly) {
  let inside = false;
  const n = poly.length / 2;
  for (let i = 0, j = n - 1; i < n; j = i++) {
    const xi = poly[2 * i], yi = poly[2 * i + 1], xj = poly[2 * j], yj = poly[2 * j + 1];
    if ((yi > y) !== (yj > y) && x < ((xj - xi) * (y - yi)) / (yj - yi) + xi) inside = !inside;
  }
  return inside;
}

// within `tol` of the segment, with its projection inside it up to tol
function isPointOnSegment(px, py, x1, y1, x2, y2, tol) {
  const dx = x2 - x1, dy = y2 - y1;
  const len = Math.sqrt(dx * dx + dy * dy);
  if (len < tol) return false;
  const t = ((px - x1) * dx + (py - y1) * dy) / (len * len);
  if (t < -tol / len || t > 1 + tol / len) return false;
  const qx = x1 + t * dx, qy = y1 + t * dy;
  return Math.sqrt((px - qx) ** 2 + (py - qy) ** 2) < tol;
}

function isOnPolygonBoundary(x, y, poly, tol) {
  const n = poly.length / 2;
  for (let i = 0; i < n; i++) {
    const j = (i + 1) % n;
    if (isPointOnSegment(x, y, poly[2 * i], poly[2 * i + 1], poly[2 * j], poly[2 * j + 1], tol))
      return true;
  }
  return false;
}

function boundingBox(v) {
  let minX = Infinity, minY = Infinity, maxX = -Infinity, maxY = -Infinity;
  for (let i = 0; i < v.length; i += 2) {
    minX = Math.min(minX, v[i]);
    minY = Math.min(minY, v[i + 1]);
    maxX = Math.max(maxX, v[i]);
    maxY = Math.max(maxY, v[i + 1]);
  }
  return { minX, minY, maxX, maxY };
}

// Steiner points: a lattice of spacing h, odd rows shifted by h / 2, inside
// the boundary, outside the holes and at least h / 10 from every segment
// (mesher-unstructured.ts:97-146)
function steinerPoints(boundary, holes, h) {
  const pts = [];
  const b = boundingBox(boundary);
  const margin = h * 0.1;
  const nx = Math.ceil((b.maxX - b.minX) / h), ny = Math.ceil((b.maxY - b.minY) / h);
  for (let j = 1; j < ny; j++) {
    const off = j % 2 === 0 ? 0 : h * 0.5;
    for (let i = 1; i < nx; i++) {
      const x = b.minX + i * h + off, y = b.minY + j * h;
      if (x < b.minX + margin || x > b.maxX - margin || y < b.minY + margin || y > b.maxY - margin)
        continue;
      if (!isInsidePolygon(x, y, boundary) || holes.some((q) => isInsidePolygon(x, y, q))) continue;
      if (isOnPolygonBoundary(x, y, boundary, margin) ||
          holes.some((q) => isOnPolygonBoundary(x, y, q, margin)))
        continue;
      pts.push(x, y);
    }
  }
  return pts;
}

// ------------------------------------------------------------- predicates

const EPS = 1.1102230246251565e-16;  // 2^-53
const ORIENT_BOUND = (3 + 16 * EPS) * EPS * 2;
const INCIRCLE_BOUND = (10 + 96 * EPS) * EPS * 2;

const f64 = new Float64Array(1);
const u32 = new Uint32Array(f64.buffer);

// x = m 2^e exactly (m an integer BigInt), for finite doubles
function decompose(x) {
  if (x === 0) return [0n, 0];
  f64[0] = x;
  const hi = u32[1], lo = u32[0];
  const bexp = (hi >>> 20) & 0x7ff;
  let m = (BigInt(hi & 0xfffff) << 32n) | BigInt(lo);
  let e;
  if (bexp === 0) e = -1074;
  else {
    m |= 1n << 52n;
    e = bexp - 1075;
  }
  if (hi >>> 31) m = -m;
  return [m, e];
}

// Exact coordinates: every coordinate times 2^shift, as BigInt.
class ExactCoords {
  constructor(xy) {
    let emin = 0;
    const parts = new Array(xy.length);
    for (let i = 0; i < xy.length; i++) {
      parts[i] = decompose(xy[i]);
      if (parts[i][0] !== 0n) emin = Math.min(emin, parts[i][1]);
    }
    this.v = parts.map(([m, e]) => (m === 0n ? 0n : m << BigInt(e - emin)));
  }
}

class Predicates {
  constructor(xy) {
    this.xy = xy;
    this.exact = null;  // built on the first call that needs it
  }
  big() {
    if (!this.exact) this.exact = new ExactCoords(this.xy).v;
    return this.exact;
  }
  // > 0: c left of a -> b (counter-clockwise a, b, c)
  orient(a, b, c) {
    const p = this.xy;
    const l = (p[2 * a] - p[2 * c]) * (p[2 * b + 1] - p[2 * c + 1]);
    const r = (p[2 * a + 1] - p[2 * c + 1]) * (p[2 * b] - p[2 * c]);
    const d = l - r;
    if (Math.abs(d) > ORIENT_BOUND * (Math.abs(l) + Math.abs(r))) return Math.sign(d);
    const q = this.big();
    const e = (q[2 * a] - q[2 * c]) * (q[2 * b + 1] - q[2 * c + 1]) -
              (q[2 * a + 1] - q[2 * c + 1]) * (q[2 * b] - q[2 * c]);
    return e > 0n ? 1 : e < 0n ? -1 : 0;
  }
  // > 0: d strictly inside the circumcircle of the counter-clockwise a, b, c
  incircle(a, b, c, d) {
    const p = this.xy;
    const dx = p[2 * d], dy = p[2 * d + 1];
    const adx = p[2 * a] - dx, ady = p[2 * a + 1] - dy;
    const bdx = p[2 * b] - dx, bdy = p[2 * b + 1] - dy;
    const cdx = p[2 * c] - dx, cdy = p[2 * c + 1] - dy;
    const al = adx * adx + ady * ady, bl = bdx * bdx + bdy * bdy, cl = cdx * cdx + cdy * cdy;
    const t1 = bdx * cdy - cdx * bdy, t2 = cdx * ady - adx * cdy, t3 = adx * bdy - bdx * ady;
    const det = al * t1 + bl * t2 + cl * t3;
    const perm = (Math.abs(bdx * cdy) + Math.abs(cdx * bdy)) * al +
                 (Math.abs(cdx * ady) + Math.abs(adx * cdy)) * bl +
                 (Math.abs(adx * bdy) + Math.abs(bdx * ady)) * cl;
    if (Math.abs(det) > INCIRCLE_BOUND * perm) return Math.sign(det);
    const q = this.big();
    const X = q[2 * d], Y = q[2 * d + 1];
    const ax = q[2 * a] - X, ay = q[2 * a + 1] - Y, bx = q[2 * b] - X, by = q[2 * b + 1] - Y;
    const cx = q[2 * c] - X, cy = q[2 * c + 1] - Y;
    const e = (ax * ax + ay * ay) * (bx * cy - cx * by) + (bx * bx + by * by) * (cx * ay - ax * cy) +
              (cx * cx + cy * cy) * (ax * by - bx * ay);
    return e > 0n ? 1 : e < 0n ? -1 : 0;
  }
}

// ------------------------------------------------------- triangulation

// Triangle t: vertices V[3t + k] counter-clockwise; N[3t + k] is the
// triangle across the edge opposite vertex k, i.e. the edge
// (V[3t + (k + 1) % 3], V[3t + (k + 2) % 3]), or -1.
class Triangulation {
  constructor(xy, npts) {
    this.pr = new Predicates(xy);
    this.V = [];
    this.N = [];
    this.alive = [];
    this.vt = new Int32Array(xy.length / 2).fill(-1);  // a live triangle at each vertex
    this.npts = npts;
    this.last = 0;
  }
  add(a, b, c) {
    const t = this.alive.length;
    this.V.push(a, b, c);
    this.N.push(-1, -1, -1);
    this.alive.push(true);
    this.vt[a] = this.vt[b] = this.vt[c] = t;
    return t;
  }
  // k with V[3t + k] == v
  slot(t, v) {
    const V = this.V;
    return V[3 * t] === v ? 0 : V[3 * t + 1] === v ? 1 : 2;
  }
  // the triangle holding p (inside or on its boundary), by a visibility walk
  locate(p) {
    const { V, N, pr } = this;
    let t = this.last;
    if (!this.alive[t]) t = this.alive.lastIndexOf(true);
    for (let steps = 0; steps < 4 * this.alive.length + 16; steps++) {
      let moved = false;
      const k0 = steps % 3;  // rotate the starting edge so the walk cannot cycle
      for (let i = 0; i < 3; i++) {
        const k = (k0 + i) % 3;
        const a = V[3 * t + (k + 1) % 3], b = V[3 * t + (k + 2) % 3];
        if (pr.orient(a, b, p) < 0 && N[3 * t + k] >= 0) {
          t = N[3 * t + k];
          moved = true;
          break;
        }
      }
      if (!moved) return t;
    }
    throw new Error('mesher: point location did not terminate');
  }
  insert(p) {
    const { V, N, pr } = this;
    const t0 = this.locate(p);
    // cavity: triangles connected to t0 whose circumcircle holds p
    const cav = new Set([t0]);
    const stack = [t0];
    while (stack.length) {
      const t = stack.pop();
      for (let k = 0; k < 3; k++) {
        const u = N[3 * t + k];
        if (u < 0 || cav.has(u)) continue;
        if (pr.incircle(V[3 * u], V[3 * u + 1], V[3 * u + 2], p) > 0) {
          cav.add(u);
          stack.push(u);
        }
      }
    }
    // boundary edges of the cavity, each fanned to p
    const edges = [];
    for (const t of cav) {
      for (let k = 0; k < 3; k++) {
        const u = N[3 * t + k];
        if (u >= 0 && cav.has(u)) continue;
        edges.push([V[3 * t + (k + 1) % 3], V[3 * t + (k + 2) % 3], u]);
      }
    }
    for (const t of cav) this.alive[t] = false;
    const byStart = new Map();  // a -> new triangle (a, b, p)
    const made = [];
    for (const [a, b, u] of edges) {
      if (pr.orient(a, b, p) <= 0) throw new Error('mesher: cavity not star-shaped');
      const t = this.add(a, b, p);
      N[3 * t + 2] = u;
      if (u >= 0) {
        // u's slot across edge (b, a)
        for (let k = 0; k < 3; k++) {
          if (V[3 * u + (k + 1) % 3] === b && V[3 * u + (k + 2) % 3] === a) N[3 * u + k] = t;
        }
      }
      byStart.set(a, t);
      made.push(t);
    }
    for (const t of made) {
      const a = V[3 * t], b = V[3 * t + 1];
      // edge (b, p) is opposite a: shared with the new triangle starting at b
      N[3 * t] = byStart.get(b);
      // edge (p, a) is opposite b: shared with the new triangle ending at a,
      // i.e. whose (b', p) edge is (a, p): the one starting at the vertex before a
      const s = made.find((w) => V[3 * w + 1] === a);
      N[3 * t + 1] = s;
    }
    this.last = made[0];
  }
  // the live triangle with directed edge u -> w, and the slot of its third vertex
  findEdge(u, w) {
    const { V, N } = this;
    const start = this.vt[u];
    if (start < 0 || !this.alive[start]) return null;
    // rotate around u in both directions
    for (const dir of [0, 1]) {
      let t = start;
      for (let guard = 0; guard < (1 << 22) && t >= 0; guard++) {
        const k = this.slot(t, u);
        if (V[3 * t + (k + 1) % 3] === w) return [t, (k + 2) % 3];
        // next triangle around u: across the edge (u, V[k+1]) (opposite k+2) or
        // (V[k+2], u) (opposite k+1)
        t = dir === 0 ? N[3 * t + (k + 2) % 3] : N[3 * t + (k + 1) % 3];
        if (t === start) break;
      }
    }
    return null;
  }
  // flip the edge shared by t (third vertex slot kt) and its neighbour
  flip(t, kt) {
    const { V, N } = this;
    const u = V[3 * t + (kt + 1) % 3], w = V[3 * t + (kt + 2) % 3], x = V[3 * t + kt];
    const s = N[3 * t + kt];
    const ks = this.slot(s, V[3 * s] !== u && V[3 * s] !== w ? V[3 * s]
      : V[3 * s + 1] !== u && V[3 * s + 1] !== w ? V[3 * s + 1] : V[3 * s + 2]);
    const y = V[3 * s + ks];
    // outer neighbours: across (w, x), (x, u) in t; across (u, y), (y, w) in s
    const nWX = N[3 * t + (kt + 1) % 3], nXU = N[3 * t + (kt + 2) % 3];
    const nUY = N[3 * s + (ks + 1) % 3], nYW = N[3 * s + (ks + 2) % 3];
    // t becomes (x, u, y), s becomes (y, w, x)
    V[3 * t] = x; V[3 * t + 1] = u; V[3 * t + 2] = y;
    V[3 * s] = y; V[3 * s + 1] = w; V[3 * s + 2] = x;
    N[3 * t] = nUY; N[3 * t + 1] = s; N[3 * t + 2] = nXU;
    N[3 * s] = nWX; N[3 * s + 1] = t; N[3 * s + 2] = nYW;
    const relink = (n, from, to) => {
      if (n < 0) return;
      for (let k = 0; k < 3; k++) if (N[3 * n + k] === from) N[3 * n + k] = to;
    };
    relink(nUY, s, t);
    relink(nWX, t, s);
    this.vt[x] = t; this.vt[u] = t; this.vt[y] = s; this.vt[w] = s;
  }
}

const key = (a, b) => (a < b ? a * 4294967296 + b : b * 4294967296 + a);

// Sloan's recovery of segment (a, b), then Lawson flips of the new edges.
function recoverSegment(T, a, b, fixed) {
  const { V, N, pr } = T;
  if (T.findEdge(a, b) || T.findEdge(b, a)) {
    fixed.add(key(a, b));
    return;
  }
  // the edges a -> b crosses, walking from a
  const crosses = (u, w) =>
    pr.orient(a, b, u) * pr.orient(a, b, w) < 0 && pr.orient(u, w, a) * pr.orient(u, w, b) < 0;
  const queue = [];
  {
    // triangle at a whose opposite edge the segment crosses
    let t = -1;
    for (let i = 0; i < T.alive.length; i++) {
      if (!T.alive[i]) continue;
      const k = V[3 * i] === a ? 0 : V[3 * i + 1] === a ? 1 : V[3 * i + 2] === a ? 2 : -1;
      if (k < 0) continue;
      const u = V[3 * i + (k + 1) % 3], w = V[3 * i + (k + 2) % 3];
      if (crosses(u, w)) { t = i; queue.push([u, w]); break; }
    }
    if (t < 0) throw new Error('mesher: segment start not found');
    let [u, w] = queue[0];
    for (let guard = 0; guard < 1 << 20; guard++) {
      // across (u, w) from t
      const k = T.slot(t, V[3 * t] !== u && V[3 * t] !== w ? V[3 * t]
        : V[3 * t + 1] !== u && V[3 * t + 1] !== w ? V[3 * t + 1] : V[3 * t + 2]);
      const s = N[3 * t + k];
      const ks = (() => { for (let j = 0; j < 3; j++) { const v = V[3 * s + j]; if (v !== u && v !== w) return j; } return -1; })();
      const y = V[3 * s + ks];
      if (y === b) break;
      if (crosses(u, y)) { queue.push([u, y]); w = y; } else { queue.push([y, w]); u = y; }
      t = s;
    }
  }
  const fresh = [];
  for (let guard = 0; queue.length && guard < 1 << 22; guard++) {
    const [u, w] = queue.shift();
    const f = T.findEdge(u, w);
    if (!f) continue;  // already flipped away
    const [t, kt] = f;
    const x = V[3 * t + kt];
    const s = N[3 * t + kt];
    let y = -1;
    for (let j = 0; j < 3; j++) { const v = V[3 * s + j]; if (v !== u && v !== w) y = v; }
    // strictly convex quad u, y, w, x: flip
    if (pr.orient(x, u, y) > 0 && pr.orient(y, w, x) > 0) {
      T.flip(t, kt);
      if (crosses(x, y)) queue.push([x, y]);
      else fresh.push([x, y]);
    } else {
      queue.push([u, w]);
    }
  }
  if (queue.length) throw new Error('mesher: segment recovery did not finish');
  fixed.add(key(a, b));
  lawson(T, fresh, fixed);
}

function lawson(T, edges, fixed) {
  const { V, N, pr } = T;
  const stack = edges.slice();
  for (let guard = 0; stack.length && guard < 1 << 24; guard++) {
    const [u, w] = stack.pop();
    if (fixed.has(key(u, w))) continue;
    const f = T.findEdge(u, w) || T.findEdge(w, u);
    if (!f) continue;
    const [t, kt] = f;
    const s = N[3 * t + kt];
    if (s < 0) continue;
    const a = V[3 * t + (kt + 1) % 3], c = V[3 * t + (kt + 2) % 3], x = V[3 * t + kt];
    let y = -1;
    for (let j = 0; j < 3; j++) { const v = V[3 * s + j]; if (v !== a && v !== c) y = v; }
    if (pr.incircle(a, c, x, y) > 0 && pr.orient(x, a, y) > 0 && pr.orient(y, c, x) > 0) {
      T.flip(t, kt);
      stack.push([x, a], [a, y], [y, c], [c, x]);
    }
  }
}

// Cyclic vertex rotations (every triangle stays counter-clockwise) so that
// each node is the first or second vertex of some triangle: the DKT element
// as written (element.ts:642-943) gives its third vertex a rank-2 nodal
// block, so a node that is only ever a third vertex leaves the stiffness
// singular.  Greedy and deterministic.
function balanceThirdVertices(tri, nNodes) {
  const ne = tri.length / 3;
  const inc = Array.from({ length: nNodes }, () => []);
  for (let e = 0; e < ne; e++) for (let k = 0; k < 3; k++) inc[tri[3 * e + k]].push(e);
  const covered = (v) => inc[v].some((e) => tri[3 * e + 2] !== v);
  const rot = (e, k) => {
    const a = tri[3 * e], b = tri[3 * e + 1], c = tri[3 * e + 2];
    const r = [a, b, c];
    tri[3 * e] = r[k]; tri[3 * e + 1] = r[(k + 1) % 3]; tri[3 * e + 2] = r[(k + 2) % 3];
  };
  for (let pass = 0; pass < 100; pass++) {
    const bare = [];
    for (let v = 0; v < nNodes; v++) if (inc[v].length && !covered(v)) bare.push(v);
    if (!bare.length) return tri;
    for (const v of bare) {
      if (covered(v)) continue;
      for (const e of inc[v]) {
        const old = [tri[3 * e], tri[3 * e + 1], tri[3 * e + 2]];
        rot(e, old.indexOf(v));
        if (covered(tri[3 * e + 2])) break;
        tri[3 * e] = old[0]; tri[3 * e + 1] = old[1]; tri[3 * e + 2] = old[2];
      }
    }
  }
  return tri;
}

/**
 * Triangular mesh of a polygonal plate with holes (mesher-unstructured.ts:36-78).
 * geometry: { boundary: Float32Array [x0, y0, ...], holes: Float32Array[] }
 * options: { targetEdgeLength?: number (Steiner lattice spacing),
 *            balanceVertexOrder?: boolean (default true: rotate vertex orders so
 *            no node is only ever a third vertex; false keeps the triangulation's) }
 * Returns a PlateMesh with nodesPerElement 3 and gridNx = gridNy = 0.
 */
function generateUnstructuredMesh(geometry, options = {}) {
  const boundary = ensureWindingOrder(geometry.boundary, false);
  const holes = (geometry.holes || []).map((h) => ensureWindingOrder(h, true));
  const xy = [];
  const segs = [];
  const seen = new Map();
  const addPoint = (x, y) => {
    const k = `${x},${y}`;
    if (seen.has(k)) return seen.get(k);
    const i = xy.length / 2;
    xy.push(x, y);
    seen.set(k, i);
    return i;
  };
  for (const poly of [boundary, ...holes]) {
    const n = poly.length / 2;
    const ids = [];
    for (let i = 0; i < n; i++) ids.push(addPoint(poly[2 * i], poly[2 * i + 1]));
    for (let i = 0; i < n; i++) if (ids[i] !== ids[(i + 1) % n]) segs.push([ids[i], ids[(i + 1) % n]]);
  }
  if (options.targetEdgeLength) {
    const sp = steinerPoints(boundary, holes, options.targetEdgeLength);
    for (let i = 0; i < sp.length; i += 2) addPoint(Math.fround(sp[i]), Math.fround(sp[i + 1]));
  }
  const npts = xy.length / 2;
  if (npts < 3) throw new Error('mesher: fewer than three points');
  // enclosing triangle: far outside the bounding box, on powers of two
  const b = boundingBox(xy);
  const span = Math.max(b.maxX - b.minX, b.maxY - b.minY, 1e-30);
  const big = Math.pow(2, Math.ceil(Math.log2(span)) + 6);
  const cx = Math.fround((b.minX + b.maxX) / 2), cy = Math.fround((b.minY + b.maxY) / 2);
  xy.push(cx - 2 * big, cy - big, cx + 2 * big, cy - big, cx, cy + 2 * big);
  const T = new Triangulation(Float64Array.from(xy), npts);
  T.add(npts, npts + 1, npts + 2);
  // spatially coherent insertion order (rows of cells of the target size)
  const order = Array.from({ length: npts }, (_, i) => i);
  const cell = span / Math.max(1, Math.sqrt(npts));
  order.sort((p, q) => {
    const rp = Math.floor((xy[2 * p + 1] - b.minY) / cell), rq = Math.floor((xy[2 * q + 1] - b.minY) / cell);
    if (rp !== rq) return rp - rq;
    const sgn = rp % 2 ? -1 : 1;
    return sgn * (xy[2 * p] - xy[2 * q]) || p - q;
  });
  for (const p of order) T.insert(p);
  const fixed = new Set();
  for (const [a, c] of segs) recoverSegment(T, a, c, fixed);
  // keep the triangles inside the boundary and outside every hole
  const keep = [];
  for (let t = 0; t < T.alive.length; t++) {
    if (!T.alive[t]) continue;
    const a = T.V[3 * t], c = T.V[3 * t + 1], d = T.V[3 * t + 2];
    if (a >= npts || c >= npts || d >= npts) continue;
    const gx = (xy[2 * a] + xy[2 * c] + xy[2 * d]) / 3, gy = (xy[2 * a + 1] + xy[2 * c + 1] + xy[2 * d + 1]) / 3;
    if (!isInsidePolygon(gx, gy, boundary) || holes.some((h) => isInsidePolygon(gx, gy, h))) continue;
    keep.push(a, c, d);
  }
  // renumber the points in use, in input order
  const used = new Int32Array(npts).fill(-1);
  for (const v of keep) used[v] = 0;
  let nn = 0;
  for (let i = 0; i < npts; i++) if (used[i] === 0) used[i] = nn++;
  const nodes = new Float32Array(2 * nn);
  for (let i = 0; i < npts; i++) {
    if (used[i] < 0) continue;
    nodes[2 * used[i]] = xy[2 * i];
    nodes[2 * used[i] + 1] = xy[2 * i + 1];
  }
  let elements = Uint32Array.from(keep, (v) => used[v]);
  if (options.balanceVertexOrder !== false) elements = balanceThirdVertices(elements, nn);
  // boundary nodes: on a boundary or hole segment (mesher-unstructured.ts:259-286)
  const bn = [];
  for (let i = 0; i < nn; i++) {
    const x = nodes[2 * i], y = nodes[2 * i + 1];
    if (isOnPolygonBoundary(x, y, boundary, 1e-6) || holes.some((h) => isOnPolygonBoundary(x, y, h, 1e-6)))
      bn.push(i);
  }
  return {
    nodes,
    elements,
    boundaryNodes: Uint32Array.from(bn),
    nodeCount: nn,
    elementCount: elements.length / 3,
    nodesPerElement: 3,
    gridNx: 0,
    gridNy: 0,
  };
}

// a polygonal circle (mesher-utils.ts:122-137)
function generateCirclePolygon(cx, cy, r, segments = 64) {
  const v = new Float32Array(2 * segments);
  for (let i = 0; i < segments; i++) {
    const th = (2 * Math.PI * i) / segments;
    v[2 * i] = cx + r * Math.cos(th);
    v[2 * i + 1] = cy + r * Math.sin(th);
  }
  return v;
}

module.exports = {
  generateUnstructuredMesh,
  generateCirclePolygon,
  ensureWindingOrder,
  isInsidePolygon,
  isPointOnSegment,
};
