'use strict';
// Drives solveGPU (the Node host over the N-API addon) on the 32 x 32 clamped
// plate and compares it with the float32 solver.ts restatement.  Prints one
// JSON line.  Run by tests/test_node_host.py.
const path = require('path');
const host = require(path.join(__dirname, '..', '..', 'gpu-gauntlet_amd', 'js', 'index.js'));
const O = require(path.join(__dirname, '..', '..', 'oracle', 'solver_f32.js'));

async function main() {
  const mesh = O.generateRectangularMesh(10, 10, 10 / 32);
  const material = { E: 210e9, nu: 0.3, t: 0.1 };
  const supports = [{ type: 'fixed', location: 'all_edges' }];
  const coloring = O.computeElementColoring(mesh);
  const constrained = O.identifyConstrainedDOFs(mesh, supports);
  const solver = new host.PlateSolver(mesh, material, coloring, constrained);
  const F = solver.pressureLoad(1e4, []);
  solver.destroy();
  const gpu = await host.solveGPU(mesh, material, coloring, F, constrained,
    { tolerance: 1e-8, maxIterations: 20000 });
  const cpu = O.solveWithF(mesh, material, supports, F, 1e-8, 20000);
  let gi = 0, ci = 0;
  for (let n = 0; n < mesh.nodeCount; n++) {
    if (Math.abs(gpu.solution[3 * n]) > Math.abs(gpu.solution[3 * gi])) gi = n;
    if (Math.abs(cpu.solution[3 * n]) > Math.abs(cpu.solution[3 * ci])) ci = n;
  }
  const rel = Math.abs(gpu.solution[3 * gi] - cpu.solution[3 * ci]) / Math.abs(cpu.solution[3 * ci]);
  const cheb = await host.solveGPU(mesh, material, coloring, F, constrained,
    { tolerance: 1e-8, maxIterations: 20000, preconditioner: 'chebyshev', polyDegree: 6 });
  const relCheb = Math.abs(cheb.solution[3 * gi] - cpu.solution[3 * ci]) / Math.abs(cpu.solution[3 * ci]);
  const mg = await host.solveGPU(mesh, material, coloring, F, constrained,
    { tolerance: 1e-8, maxIterations: 2000, preconditioner: 'multigrid' });
  const relMg = Math.abs(mg.solution[3 * gi] - cpu.solution[3 * ci]) / Math.abs(cpu.solution[3 * ci]);
  let forceCpuRejected = false;
  try {
    await host.solveGPU(mesh, material, coloring, F, constrained, { forceCPU: true });
  } catch (e) { forceCpuRejected = true; }
  host.destroyContext();
  console.log(JSON.stringify({
    usedGPU: gpu.usedGPU, converged: gpu.converged, iterations: gpu.iterations,
    cpuIterations: cpu.iterations, trueRelResidual: gpu.trueRelResidual, maxWRelError: rel,
    forceCpuRejected: forceCpuRejected, chebConverged: cheb.converged,
    chebIterations: cheb.iterations, chebTrueRelResidual: cheb.trueRelResidual,
    chebMaxWRelError: relCheb, mgConverged: mg.converged, mgIterations: mg.iterations,
    mgTrueRelResidual: mg.trueRelResidual, mgMaxWRelError: relMg,
  }));
}

main().catch(function (e) { console.error(e); process.exit(1); });
