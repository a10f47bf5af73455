'use strict';
// Concurrent solves through the Node host: two solveGPU calls in flight at
// once, two solve() calls on one PlateSolver at once, and destroy() while a
// solve is still running.  Each answer is compared with the same solve run
// alone (the FP64 path is deterministic, so they must be bit-identical) and
// with the float32 solver.ts restatement.  Prints one JSON line.  Run by
// tests/test_node_host.py.
const path = require('path');
const host = require(path.join(__dirname, '..', '..', 'gpu-gauntlet_amd', 'js', 'index.js'));
const O = require(path.join(__dirname, '..', '..', 'oracle', 'solver_f32.js'));

function maxWRel(a, b) {
  let ia = 0, ib = 0;
  for (let n = 0; 3 * n < a.length; n++) {
    if (Math.abs(a[3 * n]) > Math.abs(a[3 * ia])) ia = n;
    if (Math.abs(b[3 * n]) > Math.abs(b[3 * ib])) ib = n;
  }
  return Math.abs(a[3 * ia] - b[3 * ib]) / Math.abs(b[3 * ib]);
}

function same(a, b) {
  if (a.length !== b.length) return false;
  for (let i = 0; i < a.length; i++) if (a[i] !== b[i]) return false;
  return true;
}

async function main() {
  const mesh = O.generateRectangularMesh(10, 10, 10 / 32);
  const material = { E: 210e9, nu: 0.3, t: 0.1 };
  const supports = [{ type: 'fixed', location: 'all_edges' }];
  const coloring = O.computeElementColoring(mesh);
  const constrained = O.identifyConstrainedDOFs(mesh, supports);
  const opts = { tolerance: 1e-8, maxIterations: 20000 };
  const solver = new host.PlateSolver(mesh, material, coloring, constrained);
  const F1 = solver.pressureLoad(1e4, []);
  const F2 = solver.pressureLoad(2e3, [{ A: 5e4, cx: 3.0, cy: 6.5, sigma: 1.0 }]);
  // each load alone
  const a1 = await solver.solve(F1, opts);
  const a2 = await solver.solve(F2, opts);
  // two solveGPU calls (each with its own plan on the shared context) at once
  const [g1, g2] = await Promise.all([
    host.solveGPU(mesh, material, coloring, F1, constrained, opts),
    host.solveGPU(mesh, material, coloring, F2, constrained, opts),
  ]);
  // two solve() calls on one plan at once
  const [s2, s1] = await Promise.all([solver.solve(F2, opts), solver.solve(F1, opts)]);
  // destroy() while a solve is in flight: the solve still resolves
  const late = new host.PlateSolver(mesh, material, coloring, constrained);
  const pending = late.solve(F2, opts);
  late.destroy();
  const d2 = await pending;
  solver.destroy();
  host.destroyContext();
  const c1 = O.solveWithF(mesh, material, supports, F1, 1e-8, 20000);
  const c2 = O.solveWithF(mesh, material, supports, F2, 1e-8, 20000);
  console.log(JSON.stringify({
    converged: [g1, g2, s1, s2, d2].every((r) => r.converged),
    solveGpuSame: same(g1.solution64, a1.solution64) && same(g2.solution64, a2.solution64),
    planSame: same(s1.solution64, a1.solution64) && same(s2.solution64, a2.solution64),
    destroyedSame: same(d2.solution64, a2.solution64),
    deviceMsPositive: [g1, g2, s1, s2, d2].every((r) => r.deviceMs > 0),
    maxWRelError: Math.max(maxWRel(g1.solution, c1.solution), maxWRel(g2.solution, c2.solution),
      maxWRel(s1.solution, c1.solution), maxWRel(s2.solution, c2.solution)),
    loadsDiffer: !same(a1.solution64, a2.solution64),
  }));
}

main().catch(function (e) { console.error(e); process.exit(1); });
