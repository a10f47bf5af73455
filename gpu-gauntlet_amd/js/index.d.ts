// TypeScript declarations of the Node host (index.js).  The data types are
// the reference's (src/lib/plate/types.ts:14-102) so a caller of
// solveGPU (src/lib/plate/gpu/solver.ts:80-87) compiles unchanged.

export interface PlateMaterial {
  E: number;
  nu: number;
  t: number;
  /** Extension: per-element thickness (no reference counterpart). */
  elementT?: Float32Array | null;
}

export interface PlateMesh {
  nodes: Float32Array;
  elements: Uint32Array;
  boundaryNodes: Uint32Array;
  nodeCount: number;
  elementCount: number;
  nodesPerElement?: number;
  gridNx: number;
  gridNy: number;
}

export interface ElementColoring {
  colors: Uint32Array[];
  elementColors: Uint8Array;
  colorCount: number;
}

/** GPUSolverOptions (gpu/solver.ts:35-42) plus this library's choices. */
export interface GPUSolverOptions {
  tolerance?: number;
  maxIterations?: number;
  /** Rejected: this library has no CPU solver. */
  forceCPU?: boolean;
  precision?: 'fp64' | 'mixed';
  /** chebyshev: block Jacobi inside a Chebyshev polynomial of degree polyDegree in B K
   *  (structured FP64 path, plates of up to 512 tiles of 18 x 18 nodes);
   *  multigrid: one Galerkin multigrid V-cycle per PCG iteration (structured FP64 path) */
  preconditioner?: 'blockJacobi' | 'learned' | 'chebyshev' | 'multigrid';
  /** chebyshev: polynomial degree (default 8, at most 32); multigrid: smoothing steps on the
   *  finest level (default 6 below 100,000 nodes, else 5; at most 8) */
  polyDegree?: number;
  path?: 'auto' | 'structured' | 'coloured';
  /** mixed: a refinement sweep ends when the FP32 residual has fallen by this factor (default 1e-3) */
  innerTolerance?: number;
  /** mixed: most FP64 refinement sweeps (default 200) */
  maxRefinements?: number;
}

/** GPUSolveResult (gpu/solver.ts:47-60) plus the FP64 fields. */
export interface GPUSolveResult {
  solution: Float32Array;
  iterations: number;
  finalResidual: number;
  converged: boolean;
  gpuTimeMs: number;
  usedGPU: boolean;
  solution64: Float64Array;
  trueRelResidual: number;
  deviceMs: number;
  breakdown: boolean;
  sweeps: number;
}

export declare function solveGPU(
  mesh: PlateMesh,
  material: PlateMaterial,
  coloring: ElementColoring,
  F: Float32Array,
  constrainedDOFs: Set<number>,
  options?: GPUSolverOptions
): Promise<GPUSolveResult>;

export declare function isGPUAvailable(): boolean;
export declare function destroyContext(): void;

export interface GaussianPatch { A: number; cx: number; cy: number; sigma: number; }

export declare class PlateSolver {
  constructor(mesh: PlateMesh, material: PlateMaterial, coloring: ElementColoring | null,
              constrainedDOFs: Set<number>);
  readonly structured: boolean;
  solve(F: Float32Array, options?: GPUSolverOptions): Promise<Omit<GPUSolveResult, 'gpuTimeMs' | 'usedGPU'>>;
  applyK(x: Float64Array, path?: 'auto' | 'structured' | 'coloured'): Float64Array;
  pressureLoad(q0: number, patches?: GaussianPatch[]): Float32Array;
  /** computeMoments (postprocess.ts:40-270) of x, or of the last solution on the device */
  computeMoments(solution64?: Float64Array): { Mx: Float32Array; My: Float32Array; Mxy: Float32Array };
  /** findMaxDisplacement (postprocess.ts:296-319) */
  findMaxDisplacement(solution64?: Float64Array): { maxW: number; nodeIndex: number; x: number; y: number };
  destroy(): void;
}

/** extractVerticalDisplacements (postprocess.ts:279-288) */
export declare function extractVerticalDisplacements(displacements: Float32Array | Float64Array,
                                                     nodeCount: number): Float32Array;

/** PlateResult (types.ts:119-138) fields produced by the device solve and
 * post-processing, plus the FP64 solution and true residual. */
export interface PlateResultGPU {
  mesh: PlateMesh;
  displacements: Float32Array;
  solution64: Float64Array;
  w: Float32Array;
  Mx: Float32Array;
  My: Float32Array;
  Mxy: Float32Array;
  trueRelResidual: number;
  solverInfo: {
    iterations: number; finalResidual: number; converged: boolean;
    solveTimeMs: number; postprocessTimeMs: number; totalTimeMs: number;
  };
}

/** The solve + post-processing steps of solvePlate (solver.ts:497-574) on the device. */
export declare function solvePlateGPU(
  mesh: PlateMesh,
  material: PlateMaterial,
  coloring: ElementColoring | null,
  F: Float32Array,
  constrainedDOFs: Set<number>,
  options?: GPUSolverOptions
): Promise<PlateResultGPU>;

/** PlateGeometry (types.ts): outer boundary and holes as flat [x0, y0, x1, y1, ...]. */
export interface PlateGeometry {
  boundary: Float32Array;
  holes: Float32Array[];
}

export interface UnstructuredMeshOptions {
  /** Spacing of the Steiner lattice (mesher-unstructured.ts:97-146); none: polygon vertices only. */
  targetEdgeLength?: number;
  /** Extension: rotate each triangle's vertex order (still counter-clockwise) so that no node is
   *  only ever a third vertex, which the triangle as written would leave singular (default true). */
  balanceVertexOrder?: boolean;
}

/** generateUnstructuredMesh (mesher-unstructured.ts:36-78): constrained Delaunay triangles of the
 *  boundary minus the holes (this library's triangulator in place of poly2tri). */
export declare function generateUnstructuredMesh(
  geometry: PlateGeometry,
  options?: UnstructuredMeshOptions
): PlateMesh;

/** generateCirclePolygon (mesher-utils.ts:122-137). */
export declare function generateCirclePolygon(
  cx: number, cy: number, radius: number, segments?: number
): Float32Array;
