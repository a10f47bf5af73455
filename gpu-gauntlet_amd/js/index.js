'use strict';
/*
 * Node host of the MI355X plate solver, with the reference's solver call
 * shape (src/lib/plate/gpu/solver.ts:80-104, exported through
 * src/lib/plate/gpu/index.ts:43 and src/lib/plate/index.ts:108-111):
 *
 *   solveGPU(mesh, material, coloring, F, constrainedDOFs, options)
 *     -> Promise<GPUSolveResult>
 *
 * It converts the reference types (PlateMesh, PlateMaterial, ElementColoring,
 * Set<number>) into the C ABI's plain arrays and calls the HIP library through
 * the N-API addon (lib/platepcg.node).  There is no CPU fallback: when the
 * addon cannot load or the GPU call fails the Promise rejects (the reference
 * silently fell back to its CPU solver, gpu/solver.ts:88-103).  Types are in
 * index.d.ts.  ES2017 for Node 12.
 */
const path = require('path');

const PRECISION = { fp64: 0, mixed: 1 };
const PRECOND = { blockJacobi: 0, learned: 1, chebyshev: 2, multigrid: 3 };
const PATH = { auto: 0, structured: 1, coloured: 2 };
const FIX = { w: 1, thx: 2, thy: 4 };

let addon = null;
let cachedContext = null;

function loadAddon() {
  if (!addon) addon = require(path.join(__dirname, '..', 'lib', 'platepcg.node'));
  return addon;
}

// Capability probe; replaces isWebGPUAvailable (gpu/context.ts:19-21).
function isGPUAvailable() {
  try {
    return loadAddon().deviceCount() > 0;
  } catch (e) {
    return false;
  }
}

// Device context, created once per process (gpu/context.ts:31-106 caches
// its GPUDevice the same way).
function getContext() {
  if (!cachedContext) cachedContext = loadAddon().createContext(0);
  return cachedContext;
}

function destroyContext() {
  if (cachedContext) loadAddon().destroyContext(cachedContext);
  cachedContext = null;
}

// Set<number> of constrained DOFs -> PP_FIX_* bits per node (what
// buffers.ts:225-235 does with a u32 per DOF).
function constrainedMask(nodeCount, constrainedDOFs) {
  const mask = new Uint8Array(nodeCount);
  for (const dof of constrainedDOFs) {
    const n = Math.floor(dof / 3);
    if (n < 0 || n >= nodeCount) throw new RangeError('constrained DOF ' + dof + ' out of range');
    mask[n] |= 1 << (dof % 3);
  }
  return mask;
}

// ElementColoring {colors: Uint32Array[]} -> concatenated list + offsets
// (gpu/solver.ts:143-150 computes the same offsets).
function packColouring(coloring) {
  if (!coloring || !coloring.colors) return null;
  let total = 0;
  for (const c of coloring.colors) total += c.length;
  const order = new Uint32Array(total);
  const offsets = new Uint32Array(coloring.colors.length + 1);
  let o = 0;
  coloring.colors.forEach(function (c, i) {
    offsets[i] = o;
    order.set(c, o);
    o += c.length;
  });
  offsets[coloring.colors.length] = o;
  return { elementsByColour: order, colourOffsets: offsets };
}

function meshArgs(mesh) {
  return {
    nodes: mesh.nodes, elements: mesh.elements, nodeCount: mesh.nodeCount,
    elementCount: mesh.elementCount,
    nodesPerElement: mesh.nodesPerElement !== undefined && mesh.nodesPerElement !== null
      ? mesh.nodesPerElement : 4,
    gridNx: mesh.gridNx || 0, gridNy: mesh.gridNy || 0,
  };
}

// A solver plan: set-up once (block inverses in FP64 on the device), solve
// many load cases.
class PlateSolver {
  constructor(mesh, material, coloring, constrainedDOFs) {
    this.lib = loadAddon();
    this.nodeCount = mesh.nodeCount;
    const mask = constrainedMask(mesh.nodeCount, constrainedDOFs);
    this.handle = this.lib.createPlan(getContext(), meshArgs(mesh), material,
      packColouring(coloring), mask);
    this.structured = this.lib.isStructured(this.handle);
  }

  solve(F, options) {
    const o = Object.assign({ tolerance: 1e-6, maxIterations: 1000 }, options || {});
    if (!(F instanceof Float32Array) || F.length !== 3 * this.nodeCount) {
      return Promise.reject(new TypeError('F must be a Float32Array of 3 * nodeCount'));
    }
    return this.lib.solve(this.handle, F, {
      tolerance: o.tolerance, maxIterations: o.maxIterations,
      precision: PRECISION[o.precision || 'fp64'], precond: PRECOND[o.preconditioner || 'blockJacobi'],
      path: PATH[o.path || 'auto'], innerTolerance: o.innerTolerance || 0,
      maxRefinements: o.maxRefinements || 0, polyDegree: o.polyDegree || 0,
    });
  }

  applyK(x, pathName) {
    const y = new Float64Array(x.length);
    this.lib.applyK(this.handle, x, y, PATH[pathName || 'auto']);
    return y;
  }

  // Consistent nodal load of q0 + Gaussian patches [{A, cx, cy, sigma}].
  pressureLoad(q0, patches) {
    const p = new Float64Array(4 * (patches ? patches.length : 0));
    (patches || []).forEach(function (g, i) { p.set([g.A, g.cx, g.cy, g.sigma], 4 * i); });
    const F = new Float32Array(3 * this.nodeCount);
    this.lib.pressureLoad(this.handle, q0, p, F);
    return F;
  }

  // Nodal bending moments (computeMoments, postprocess.ts:40-270) of a
  // Float64Array DOF vector, or of the last solve's solution on the device.
  computeMoments(solution64) {
    const Mx = new Float32Array(this.nodeCount), My = new Float32Array(this.nodeCount);
    const Mxy = new Float32Array(this.nodeCount);
    this.lib.moments(this.handle, solution64 || null, Mx, My, Mxy);
    return { Mx, My, Mxy };
  }

  // Largest |w| (findMaxDisplacement, postprocess.ts:296-319): signed maxW,
  // nodeIndex (lowest among equals), x, y.
  findMaxDisplacement(solution64) {
    return this.lib.maxDisplacement(this.handle, solution64 || null);
  }

  destroy() {
    if (this.handle) this.lib.destroyPlan(this.handle);
    this.handle = null;
  }
}

// w of every node from the DOF vector (extractVerticalDisplacements,
// postprocess.ts:279-288).
function extractVerticalDisplacements(displacements, nodeCount) {
  const w = new Float32Array(nodeCount);
  for (let i = 0; i < nodeCount; i++) w[i] = displacements[3 * i];
  return w;
}

// The solve and post-processing steps of solvePlate (solver.ts:497-574) on
// the device, for a caller that has the mesh, colouring, load vector and
// constrained DOFs: the PlateResult fields displacements, w, Mx, My, Mxy
// and solverInfo (iterations, finalResidual, converged, solveTimeMs,
// postprocessTimeMs, totalTimeMs), plus solution64 / trueRelResidual.
// Defaults follow solvePlate: tolerance 1e-6, 1000 iterations.
async function solvePlateGPU(mesh, material, coloring, F, constrainedDOFs, options) {
  const t0 = process.hrtime.bigint();
  const ms = (a, b) => Number(b - a) / 1e6;
  const solver = new PlateSolver(mesh, material, coloring, constrainedDOFs);
  try {
    const t1 = process.hrtime.bigint();
    const r = await solver.solve(F, Object.assign({ tolerance: 1e-6, maxIterations: 1000 }, options || {}));
    const t2 = process.hrtime.bigint();
    const { Mx, My, Mxy } = solver.computeMoments();
    const t3 = process.hrtime.bigint();
    return {
      mesh, displacements: r.solution, solution64: r.solution64,
      w: extractVerticalDisplacements(r.solution, mesh.nodeCount), Mx, My, Mxy,
      trueRelResidual: r.trueRelResidual,
      solverInfo: {
        iterations: r.iterations, finalResidual: r.finalResidual, converged: r.converged,
        solveTimeMs: ms(t1, t2), postprocessTimeMs: ms(t2, t3), totalTimeMs: ms(t0, process.hrtime.bigint()),
      },
    };
  } finally {
    solver.destroy();
  }
}

// solveGPU with the reference's signature and result shape
// (GPUSolveResult, gpu/solver.ts:47-60), plus solution64 / trueRelResidual /
// deviceMs.  Defaults follow the reference: tolerance 1e-6, 1000 iterations
// (gpu/solver.ts:121-122).
async function solveGPU(mesh, material, coloring, F, constrainedDOFs, options) {
  const opts = options || {};
  if (opts.forceCPU) throw new Error('forceCPU: this library has no CPU solver');
  const t0 = process.hrtime.bigint();
  const solver = new PlateSolver(mesh, material, coloring, constrainedDOFs);
  try {
    const r = await solver.solve(F, opts);
    r.gpuTimeMs = Number(process.hrtime.bigint() - t0) / 1e6;
    r.usedGPU = true;
    return r;
  } finally {
    solver.destroy();
  }
}

// the unstructured mesher (mesher-unstructured.ts:36-78) for the coloured path
const { generateUnstructuredMesh, generateCirclePolygon } = require('./mesher');

module.exports = {
  solveGPU, solvePlateGPU, extractVerticalDisplacements, isGPUAvailable, PlateSolver,
  destroyContext, constrainedMask, packColouring, generateUnstructuredMesh, generateCirclePolygon,
  PRECISION, PRECOND, PATH, FIX,
};
