'use strict';
/*
 * Float32-storage restatement of the reference's CPU solve path.
 *
 * TEST INFRASTRUCTURE ONLY: the parity target the GPU answer is compared
 * with and the CPU baseline that is timed beside it.  Nothing in the product
 * loads this file.
 *
 * It follows cseifert512/gpu-gauntlet src/lib/plate line by line in
 * behaviour, keeping the reference's storage types (Float32Array vectors,
 * Ke, block inverses; JS-double scalars and dot sums) so its rounding matches
 * solver.ts:
 *   generateRectangularMesh   mesher.ts:364-439
 *   computeElementColoring    coloring.ts:38-81
 *   computeElementColoringGreedy coloring.ts:92-177
 *   identifyConstrainedDOFs   solver.ts:92-134
 *   computeElementStiffness   element.ts:258-483 (+ :56-240)
 *   computeDKTStiffness       element.ts:642-943
 *   computeBlockDiagonal      solver.ts:188-225
 *   invertBlockDiagonal       solver.ts:231-304
 *   applyGlobalK (+ Ke memo)  solver.ts:47-83, 422-483
 *   solvePCG / solveTestSystem pcg.ts:26-88, 98-247
 *   buildLoadVector / applyBCsToRHS solver.ts:363-407
 * Written in ES2017 so Node 12 runs it (the reference's `??` is spelled out).
 */

const DOFS_PER_NODE = 3;

// ---------------------------------------------------------------- mesher

function generateRectangularMesh(width, height, targetSize) {
  const nx = Math.max(2, Math.ceil(width / targetSize));
  const ny = Math.max(2, Math.ceil(height / targetSize));
  return rectangularMeshN(width, height, nx, ny);
}

// Same construction as mesher.ts:364-439 with nx, ny given directly (the
// reference derives them from a target size, which cannot express every n).
function rectangularMeshN(width, height, nx, ny) {
  const dx = width / nx;
  const dy = height / ny;
  const nodesX = nx + 1;
  const nodesY = ny + 1;
  const nodeCount = nodesX * nodesY;
  const elementCount = nx * ny;
  const nodes = new Float32Array(nodeCount * 2);
  for (let j = 0; j <= ny; j++) {
    for (let i = 0; i <= nx; i++) {
      const idx = j * nodesX + i;
      nodes[idx * 2] = i * dx;
      nodes[idx * 2 + 1] = j * dy;
    }
  }
  const elements = new Uint32Array(elementCount * 4);
  let e = 0;
  for (let j = 0; j < ny; j++) {
    for (let i = 0; i < nx; i++) {
      elements[e * 4] = j * nodesX + i;
      elements[e * 4 + 1] = j * nodesX + (i + 1);
      elements[e * 4 + 2] = (j + 1) * nodesX + (i + 1);
      elements[e * 4 + 3] = (j + 1) * nodesX + i;
      e++;
    }
  }
  const b = [];
  for (let i = 0; i <= nx; i++) b.push(i);
  for (let j = 1; j <= ny; j++) b.push(j * nodesX + nx);
  for (let i = nx - 1; i >= 0; i--) b.push(ny * nodesX + i);
  for (let j = ny - 1; j >= 1; j--) b.push(j * nodesX);
  return {
    nodes: nodes, elements: elements, boundaryNodes: new Uint32Array(b),
    nodeCount: nodeCount, elementCount: elementCount, nodesPerElement: 4,
    gridNx: nx, gridNy: ny,
  };
}

function npeOf(mesh) {
  return mesh.nodesPerElement !== undefined && mesh.nodesPerElement !== null
    ? mesh.nodesPerElement : 4;
}

function getElementCoords(mesh, e) {
  const npe = npeOf(mesh);
  const c = new Float32Array(npe * 2);
  for (let i = 0; i < npe; i++) {
    const n = mesh.elements[e * npe + i];
    c[i * 2] = mesh.nodes[n * 2];
    c[i * 2 + 1] = mesh.nodes[n * 2 + 1];
  }
  return c;
}

function getElementNodeIndices(mesh, e) {
  const npe = npeOf(mesh);
  const out = [];
  for (let i = 0; i < npe; i++) out.push(mesh.elements[e * npe + i]);
  return out;
}

// -------------------------------------------------------------- colouring

function computeElementColoring(mesh) {
  const elementCount = mesh.elementCount, gridNx = mesh.gridNx;
  const elementColors = new Uint8Array(elementCount);
  const counts = [0, 0, 0, 0];
  for (let e = 0; e < elementCount; e++) {
    const i = e % gridNx;
    const j = Math.floor(e / gridNx);
    const c = (i % 2) * 2 + (j % 2);
    elementColors[e] = c;
    counts[c]++;
  }
  const colors = [];
  const pos = [0, 0, 0, 0];
  for (let c = 0; c < 4; c++) colors.push(new Uint32Array(counts[c]));
  for (let e = 0; e < elementCount; e++) {
    const c = elementColors[e];
    colors[c][pos[c]++] = e;
  }
  return { colors: colors, elementColors: elementColors, colorCount: 4 };
}

function computeElementColoringGreedy(mesh) {
  const elementCount = mesh.elementCount, elements = mesh.elements;
  const npe = npeOf(mesh);
  const nodeToElements = [];
  for (let e = 0; e < elementCount; e++) {
    for (let i = 0; i < npe; i++) {
      const n = elements[e * npe + i];
      if (!nodeToElements[n]) nodeToElements[n] = [];
      nodeToElements[n].push(e);
    }
  }
  const adj = [];
  for (let e = 0; e < elementCount; e++) adj.push(new Set());
  for (let e = 0; e < elementCount; e++) {
    for (let i = 0; i < npe; i++) {
      const n = elements[e * npe + i];
      for (const nb of nodeToElements[n]) if (nb !== e) adj[e].add(nb);
    }
  }
  const elementColors = new Uint8Array(elementCount);
  elementColors.fill(255);
  let maxColor = 0;
  for (let e = 0; e < elementCount; e++) {
    const used = new Set();
    for (const nb of adj[e]) if (elementColors[nb] !== 255) used.add(elementColors[nb]);
    let c = 0;
    while (used.has(c)) c++;
    elementColors[e] = c;
    maxColor = Math.max(maxColor, c);
  }
  const colorCount = maxColor + 1;
  const counts = new Array(colorCount).fill(0);
  for (let e = 0; e < elementCount; e++) counts[elementColors[e]]++;
  const colors = [];
  const pos = new Array(colorCount).fill(0);
  for (let c = 0; c < colorCount; c++) colors.push(new Uint32Array(counts[c]));
  for (let e = 0; e < elementCount; e++) {
    const c = elementColors[e];
    colors[c][pos[c]++] = e;
  }
  return { colors: colors, elementColors: elementColors, colorCount: colorCount };
}

// -------------------------------------------------------------- supports

function identifyConstrainedDOFs(mesh, supports) {
  const constrained = new Set();
  for (const s of supports) {
    const nodes = s.location === 'all_edges' ? Array.from(mesh.boundaryNodes) : [s.location];
    for (const n of nodes) {
      const b = n * DOFS_PER_NODE;
      if (s.type === 'pinned' || s.type === 'roller') {
        constrained.add(b);
      } else if (s.type === 'fixed') {
        constrained.add(b); constrained.add(b + 1); constrained.add(b + 2);
      }
    }
  }
  return constrained;
}

// ------------------------------------------------------------- elements

const GP = 0.5773502691896257;
const GAUSS_2X2 = [[-GP, -GP], [GP, -GP], [GP, GP], [-GP, GP]];

function dN(xi, eta) {
  const xm = 1 - xi, xp = 1 + xi, em = 1 - eta, ep = 1 + eta;
  return {
    dNdXi: [-em / 4, em / 4, ep / 4, -ep / 4],
    dNdEta: [-xm / 4, -xp / 4, xp / 4, xm / 4],
  };
}

function physical(dNdXi, dNdEta, xy) {
  let J11 = 0, J12 = 0, J21 = 0, J22 = 0;
  for (let i = 0; i < 4; i++) {
    J11 += dNdXi[i] * xy[i * 2]; J12 += dNdXi[i] * xy[i * 2 + 1];
    J21 += dNdEta[i] * xy[i * 2]; J22 += dNdEta[i] * xy[i * 2 + 1];
  }
  const detJ = J11 * J22 - J12 * J21;
  const inv = 1.0 / detJ;
  const dNdx = [0, 0, 0, 0], dNdy = [0, 0, 0, 0];
  for (let i = 0; i < 4; i++) {
    dNdx[i] = inv * (J22 * dNdXi[i] - J12 * dNdEta[i]);
    dNdy[i] = inv * (-J21 * dNdXi[i] + J11 * dNdEta[i]);
  }
  return { dNdx: dNdx, dNdy: dNdy, detJ: detJ };
}

function flexuralRigidity(m) {
  return (m.E * Math.pow(m.t, 3)) / (12 * (1 - Math.pow(m.nu, 2)));
}

function computeElementStiffness(xy, material) {
  const Ke = new Float32Array(144);
  const D = flexuralRigidity(material);
  const nu = material.nu;
  const Db = new Float32Array([D, D * nu, 0, D * nu, D, 0, 0, 0, (D * (1 - nu)) / 2]);
  const G = material.E / (2 * (1 + nu));
  const Ds = new Float32Array([(5.0 / 6.0) * G * material.t, 0, 0, (5.0 / 6.0) * G * material.t]);
  const D11 = Db[0], D12 = Db[1], D22 = Db[4], D33 = Db[8];
  for (let gp = 0; gp < 4; gp++) {
    const d = dN(GAUSS_2X2[gp][0], GAUSS_2X2[gp][1]);
    const ph = physical(d.dNdXi, d.dNdEta, xy);
    const f = 1.0 * Math.abs(ph.detJ);
    for (let i = 0; i < 4; i++) {
      for (let j = 0; j < 4; j++) {
        const txi = i * 3 + 1, tyi = i * 3 + 2, txj = j * 3 + 1, tyj = j * 3 + 2;
        const xi_ = ph.dNdx[i], yi_ = ph.dNdy[i], xj_ = ph.dNdx[j], yj_ = ph.dNdy[j];
        const Ktt = yi_ * D22 * yj_ + xi_ * D33 * xj_;
        const Kyy = xi_ * D11 * xj_ + yi_ * D33 * yj_;
        const Kty = -yi_ * D12 * xj_ - xi_ * D33 * yj_;
        const Kyt = -xi_ * D12 * yj_ - yi_ * D33 * xj_;
        Ke[txi * 12 + txj] += f * Ktt;
        Ke[tyi * 12 + tyj] += f * Kyy;
        Ke[txi * 12 + tyj] += f * Kty;
        Ke[tyi * 12 + txj] += f * Kyt;
      }
    }
  }
  {
    const N = [0.25, 0.25, 0.25, 0.25];
    const d = dN(0, 0);
    const ph = physical(d.dNdXi, d.dNdEta, xy);
    const f = 4.0 * Math.abs(ph.detJ);
    const Ds11 = Ds[0], Ds22 = Ds[3];
    for (let i = 0; i < 4; i++) {
      for (let j = 0; j < 4; j++) {
        const wi = i * 3, txi = wi + 1, tyi = wi + 2, wj = j * 3, txj = wj + 1, tyj = wj + 2;
        const Ni = N[i], Nj = N[j];
        const xi_ = ph.dNdx[i], yi_ = ph.dNdy[i], xj_ = ph.dNdx[j], yj_ = ph.dNdy[j];
        Ke[wi * 12 + wj] += f * (xi_ * Ds11 * xj_ + yi_ * Ds22 * yj_);
        Ke[wi * 12 + txj] += f * (yi_ * Ds22 * (-Nj));
        Ke[wi * 12 + tyj] += f * (xi_ * Ds11 * Nj);
        Ke[txi * 12 + wj] += f * ((-Ni) * Ds22 * yj_);
        Ke[txi * 12 + txj] += f * ((-Ni) * Ds22 * (-Nj));
        Ke[txi * 12 + tyj] += f * 0;
        Ke[tyi * 12 + wj] += f * (Ni * Ds11 * xj_);
        Ke[tyi * 12 + txj] += f * 0;
        Ke[tyi * 12 + tyj] += f * (Ni * Ds11 * Nj);
      }
    }
  }
  return Ke;
}

const TRI_GP = [[1 / 6, 1 / 6, 2 / 3], [2 / 3, 1 / 6, 1 / 6], [1 / 6, 2 / 3, 1 / 6]];
const TRI_W = [1 / 3, 1 / 3, 1 / 3];

function dktBx(L1, L2, P4, P5, P6, q4, q5, q6, c4, c5, c6) {
  const a = new Float32Array(9), b = new Float32Array(9);
  a[0] = P6 * (1 - 2 * L1) + (P5 - P6) * L2;
  a[1] = q6 * (1 - 2 * L1) - (q5 + q6) * L2;
  a[2] = -4 + 6 * (L1 + L2) + (c5 - c6) * L2 + c6 * (1 - 2 * L1);
  b[0] = -P6 * (1 - 2 * L2) + (P4 - P5) * L1 - P5 * (1 - 2 * L2);
  b[1] = q6 * (1 - 2 * L2) - (q4 + q5) * L1 + q5 * (1 - 2 * L2);
  b[2] = -4 + 6 * (L1 + L2) + (c4 - c5) * L1 + c5 * (1 - 2 * L2);
  a[3] = P4 * L2 + P6 * (1 - 2 * L1);
  a[4] = q4 * L2 - q6 * (1 - 2 * L1);
  a[5] = c4 * L2 - c6 * (1 - 2 * L1);
  b[3] = P4 * (1 - 2 * L2) + P5 * L1;
  b[4] = q4 * (1 - 2 * L2) - q5 * L1;
  b[5] = c4 * (1 - 2 * L2) - c5 * L1;
  a[6] = -P5 * L2 - P6 * (1 - 2 * L1);
  a[7] = q5 * L2 + q6 * (1 - 2 * L1);
  a[8] = -c5 * L2 + c6 * (1 - 2 * L1);
  b[6] = -P4 * L1 - P5 * (1 - 2 * L2);
  b[7] = q4 * L1 + q5 * (1 - 2 * L2);
  b[8] = -c4 * L1 + c5 * (1 - 2 * L2);
  return { L1: a, L2: b };
}

function dktBy(L1, L2, t4, t5, t6, r4, r5, r6, e4, e5, e6, b4, b5, b6) {
  const a = new Float32Array(9), b = new Float32Array(9);
  a[0] = t6 * (1 - 2 * L1) + (t5 - t6) * L2;
  a[1] = 1 + r6 * (1 - 2 * L1) - (r5 + r6) * L2;
  a[2] = -e6 * (1 - 2 * L1) + (b5 - b6) * L2;
  b[0] = -t6 * (1 - 2 * L2) + (t4 - t5) * L1 - t5 * (1 - 2 * L2);
  b[1] = 1 + r6 * (1 - 2 * L2) - (r4 + r5) * L1 + r5 * (1 - 2 * L2);
  b[2] = e6 * (1 - 2 * L2) + (b4 - b5) * L1 - b5 * (1 - 2 * L2);
  a[3] = t4 * L2 + t6 * (1 - 2 * L1);
  a[4] = -1 + r4 * L2 - r6 * (1 - 2 * L1);
  a[5] = -e4 * L2 - b6 * (1 - 2 * L1);
  b[3] = t4 * (1 - 2 * L2) + t5 * L1;
  b[4] = -1 + r4 * (1 - 2 * L2) - r5 * L1;
  b[5] = -e4 * (1 - 2 * L2) + b5 * L1;
  a[6] = -t5 * L2 - t6 * (1 - 2 * L1);
  a[7] = r5 * L2 + r6 * (1 - 2 * L1);
  a[8] = e5 * L2 + b6 * (1 - 2 * L1);
  b[6] = -t4 * L1 - t5 * (1 - 2 * L2);
  b[7] = r4 * L1 + r5 * (1 - 2 * L2);
  b[8] = e4 * L1 - b5 * (1 - 2 * L2);
  return { L1: a, L2: b };
}

function computeDKTStiffness(xy, material) {
  const Ke = new Float32Array(81);
  const x1 = xy[0], y1 = xy[1], x2 = xy[2], y2 = xy[3], x3 = xy[4], y3 = xy[5];
  const x12 = x1 - x2, y12 = y1 - y2, x23 = x2 - x3, y23 = y2 - y3, x31 = x3 - x1, y31 = y3 - y1;
  const L12 = x12 * x12 + y12 * y12, L23 = x23 * x23 + y23 * y23, L31 = x31 * x31 + y31 * y31;
  const twoA = x12 * y31 - x31 * y12;
  const area = Math.abs(twoA) / 2;
  if (area < 1e-15) return Ke;
  const D = flexuralRigidity(material);
  const nu = material.nu;
  const D11 = D, D12 = D * nu, D22 = D, D33 = (D * (1 - nu)) / 2;
  const a4 = -x23 / L23, b4 = (3 * x23 * y23) / (4 * L23), c4 = (x23 * x23 - 2 * y23 * y23) / (4 * L23);
  const d4 = -y23 / L23, e4 = (y23 * y23 - 2 * x23 * x23) / (4 * L23);
  const a5 = -x31 / L31, b5 = (3 * x31 * y31) / (4 * L31), c5 = (x31 * x31 - 2 * y31 * y31) / (4 * L31);
  const d5 = -y31 / L31, e5 = (y31 * y31 - 2 * x31 * x31) / (4 * L31);
  const a6 = -x12 / L12, b6 = (3 * x12 * y12) / (4 * L12), c6 = (x12 * x12 - 2 * y12 * y12) / (4 * L12);
  const d6 = -y12 / L12, e6 = (y12 * y12 - 2 * x12 * x12) / (4 * L12);
  const P4 = -6 * a4, P5 = -6 * a5, P6 = -6 * a6;
  const t4 = -6 * d4, t5 = -6 * d5, t6 = -6 * d6;
  const q4 = 3 * a4, q5 = 3 * a5, q6 = 3 * a6;
  const r4 = 3 * d4, r5 = 3 * d5, r6 = 3 * d6;
  for (let gp = 0; gp < 3; gp++) {
    const L1 = TRI_GP[gp][0], L2 = TRI_GP[gp][1];
    const w = TRI_W[gp];
    const Bx = dktBx(L1, L2, P4, P5, P6, q4, q5, q6, c4, c5, c6);
    const By = dktBy(L1, L2, t4, t5, t6, r4, r5, r6, e4, e5, e6, b4, b5, b6);
    const inv2A = 1 / twoA;
    const B = new Float32Array(27);
    for (let j = 0; j < 9; j++) {
      B[j] = inv2A * ((-y23) * Bx.L1[j] + (-y31) * Bx.L2[j]);
      B[9 + j] = inv2A * (x23 * By.L1[j] + x31 * By.L2[j]);
      B[18 + j] = inv2A * (x23 * Bx.L1[j] + x31 * Bx.L2[j] + (-y23) * By.L1[j] + (-y31) * By.L2[j]);
    }
    const f = w * area;
    const DB = new Float32Array(27);
    for (let j = 0; j < 9; j++) {
      DB[j] = D11 * B[j] + D12 * B[9 + j];
      DB[9 + j] = D12 * B[j] + D22 * B[9 + j];
      DB[18 + j] = D33 * B[18 + j];
    }
    for (let i = 0; i < 9; i++) {
      for (let j = 0; j < 9; j++) {
        let s = 0;
        for (let k = 0; k < 3; k++) s += B[k * 9 + i] * DB[k * 9 + j];
        Ke[i * 9 + j] += f * s;
      }
    }
  }
  return Ke;
}

function elementKe(mesh, material, e) {
  const xy = getElementCoords(mesh, e);
  return npeOf(mesh) === 3 ? computeDKTStiffness(xy, material) : computeElementStiffness(xy, material);
}

// ------------------------------------------------------ preconditioner

function computeBlockDiagonal(mesh, material) {
  const blk = new Float32Array(mesh.nodeCount * 9);
  const npe = npeOf(mesh);
  const nd = npe * 3;
  for (let e = 0; e < mesh.elementCount; e++) {
    const Ke = elementKe(mesh, material, e);
    const nodes = getElementNodeIndices(mesh, e);
    for (let i = 0; i < npe; i++) {
      const off = nodes[i] * 9, lb = i * 3;
      for (let r = 0; r < 3; r++)
        for (let c = 0; c < 3; c++) blk[off + r * 3 + c] += Ke[(lb + r) * nd + lb + c];
    }
  }
  return blk;
}

function invertBlockDiagonal(b, constrained) {
  const nodeCount = b.length / 9;
  for (let n = 0; n < nodeCount; n++) {
    const o = n * 9, d = n * 3;
    const cw = constrained.has(d), cx = constrained.has(d + 1), cy = constrained.has(d + 2);
    if (cw && cx && cy) {
      b[o] = 1; b[o + 1] = 0; b[o + 2] = 0; b[o + 3] = 0; b[o + 4] = 1; b[o + 5] = 0;
      b[o + 6] = 0; b[o + 7] = 0; b[o + 8] = 1;
      continue;
    }
    if (cw) { b[o] = 1; b[o + 1] = 0; b[o + 2] = 0; b[o + 3] = 0; b[o + 6] = 0; }
    if (cx) { b[o + 4] = 1; b[o + 1] = 0; b[o + 7] = 0; b[o + 3] = 0; b[o + 5] = 0; }
    if (cy) { b[o + 8] = 1; b[o + 2] = 0; b[o + 5] = 0; b[o + 6] = 0; b[o + 7] = 0; }
    const a00 = b[o], a01 = b[o + 1], a02 = b[o + 2], a10 = b[o + 3], a11 = b[o + 4];
    const a12 = b[o + 5], a20 = b[o + 6], a21 = b[o + 7], a22 = b[o + 8];
    const det = a00 * (a11 * a22 - a12 * a21) - a01 * (a10 * a22 - a12 * a20) + a02 * (a10 * a21 - a11 * a20);
    if (Math.abs(det) < 1e-30) {
      b[o] = 1; b[o + 1] = 0; b[o + 2] = 0; b[o + 3] = 0; b[o + 4] = 1; b[o + 5] = 0;
      b[o + 6] = 0; b[o + 7] = 0; b[o + 8] = 1;
      continue;
    }
    const id = 1.0 / det;
    b[o] = id * (a11 * a22 - a12 * a21);
    b[o + 1] = id * (a02 * a21 - a01 * a22);
    b[o + 2] = id * (a01 * a12 - a02 * a11);
    b[o + 3] = id * (a12 * a20 - a10 * a22);
    b[o + 4] = id * (a00 * a22 - a02 * a20);
    b[o + 5] = id * (a02 * a10 - a00 * a12);
    b[o + 6] = id * (a10 * a21 - a11 * a20);
    b[o + 7] = id * (a01 * a20 - a00 * a21);
    b[o + 8] = id * (a00 * a11 - a01 * a10);
  }
}

// ---------------------------------------------------------------- K . x

let cachedKe = null, cachedMesh = null, cachedMaterial = null;

function getOrComputeKe(mesh, material, e) {
  if (cachedMesh !== mesh || cachedMaterial !== material) {
    cachedKe = new Map(); cachedMesh = mesh; cachedMaterial = material;
  }
  if (!cachedKe.has(e)) cachedKe.set(e, elementKe(mesh, material, e));
  return cachedKe.get(e);
}

function clearKeCache() { cachedKe = null; cachedMesh = null; cachedMaterial = null; }

function applyGlobalK(mesh, material, coloring, x, y, constrained) {
  y.fill(0);
  const npe = npeOf(mesh);
  const nd = npe * DOFS_PER_NODE;
  const xl = new Float32Array(nd), yl = new Float32Array(nd);
  for (const idx of coloring.colors) {
    for (let i = 0; i < idx.length; i++) {
      const e = idx[i];
      const nodes = getElementNodeIndices(mesh, e);
      const dofs = [];
      for (let n = 0; n < npe; n++) {
        const b = nodes[n] * DOFS_PER_NODE;
        dofs.push(b, b + 1, b + 2);
      }
      for (let j = 0; j < nd; j++) xl[j] = x[dofs[j]];
      const Ke = getOrComputeKe(mesh, material, e);
      yl.fill(0);
      for (let r = 0; r < nd; r++) {
        let s = 0;
        for (let c = 0; c < nd; c++) s += Ke[r * nd + c] * xl[c];
        yl[r] = s;
      }
      for (let j = 0; j < nd; j++) y[dofs[j]] += yl[j];
    }
  }
  for (const d of constrained) y[d] = x[d];
}

// ------------------------------------------------------------------ PCG

function dot(a, b) { let s = 0; for (let i = 0; i < a.length; i++) s += a[i] * b[i]; return s; }
function norm(a) { return Math.sqrt(dot(a, a)); }
function axpy(alpha, x, y) { for (let i = 0; i < x.length; i++) y[i] += alpha * x[i]; }

function applyBlockPrecond(bi, r, z) {
  const nodeCount = bi.length / 9;
  for (let n = 0; n < nodeCount; n++) {
    const o = n * 9, d = n * 3;
    const r0 = r[d], r1 = r[d + 1], r2 = r[d + 2];
    z[d] = bi[o] * r0 + bi[o + 1] * r1 + bi[o + 2] * r2;
    z[d + 1] = bi[o + 3] * r0 + bi[o + 4] * r1 + bi[o + 5] * r2;
    z[d + 2] = bi[o + 6] * r0 + bi[o + 7] * r1 + bi[o + 8] * r2;
  }
}

function precond(opts, r, z) {
  const n = r.length;
  if (opts.blockPreconditioner) applyBlockPrecond(opts.blockPreconditioner, r, z);
  else if (opts.preconditioner) {
    const M = opts.preconditioner;
    for (let i = 0; i < n; i++) z[i] = M[i] > 1e-30 ? r[i] / M[i] : r[i];
  } else z.set(r);
}

function solvePCG(applyK, b, opts) {
  const n = b.length;
  const tol = opts.tolerance, maxIt = opts.maxIterations;
  const x = new Float32Array(n), r = new Float32Array(n), z = new Float32Array(n);
  const p = new Float32Array(n), Ap = new Float32Array(n);
  r.set(b);
  const bn = norm(b);
  if (bn < 1e-30) return { solution: x, iterations: 0, finalResidual: 0, converged: true };
  precond(opts, r, z);
  p.set(z);
  let rz = dot(r, z);
  for (let it = 0; it < maxIt; it++) {
    applyK(p, Ap);
    const pAp = dot(p, Ap);
    if (Math.abs(pAp) < 1e-30) {
      return { solution: x, iterations: it + 1, finalResidual: norm(r), converged: false };
    }
    const alpha = rz / pAp;
    axpy(alpha, p, x);
    axpy(-alpha, Ap, r);
    const rn = norm(r);
    if (rn < tol || rn / bn < tol) {
      return { solution: x, iterations: it + 1, finalResidual: rn, converged: true };
    }
    precond(opts, r, z);
    const rzn = dot(r, z);
    const beta = rzn / rz;
    for (let i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
    rz = rzn;
  }
  return { solution: x, iterations: maxIt, finalResidual: norm(r), converged: false };
}

function solveTestSystem() {
  const n = 10;
  const applyA = function (x, y) {
    y.fill(0);
    for (let i = 0; i < n; i++) {
      y[i] = 2 * x[i];
      if (i > 0) y[i] -= x[i - 1];
      if (i < n - 1) y[i] -= x[i + 1];
    }
  };
  const b = new Float32Array(n);
  b[0] = 1; b[n - 1] = 1;
  const M = new Float32Array(n).fill(2);
  return solvePCG(applyA, b, { tolerance: 1e-10, maxIterations: 100, preconditioner: M });
}

// ---------------------------------------------------------------- loads

function buildLoadVector(mesh, loads) {
  const F = new Float32Array(mesh.nodeCount * DOFS_PER_NODE);
  for (const load of loads) {
    const px = load.position[0], py = load.position[1];
    let nearest = 0, minDist = Infinity;
    for (let i = 0; i < mesh.nodeCount; i++) {
      const nx = mesh.nodes[i * 2], ny = mesh.nodes[i * 2 + 1];
      const d = Math.sqrt(Math.pow(nx - px, 2) + Math.pow(ny - py, 2));
      if (d < minDist) { minDist = d; nearest = i; }
    }
    F[nearest * DOFS_PER_NODE] += load.magnitude;
  }
  return F;
}

function applyBCsToRHS(F, constrained) { for (const d of constrained) F[d] = 0.0; }

// The solve of solver.ts:497-543 (steps 2-7) on a given mesh and a given F.
function solveWithF(mesh, material, supports, F, tolerance, maxIterations) {
  const t0 = Date.now();
  const coloring = npeOf(mesh) === 3 || mesh.gridNx === 0
    ? computeElementColoringGreedy(mesh) : computeElementColoring(mesh);
  const constrained = identifyConstrainedDOFs(mesh, supports);
  const tA = Date.now();
  const blk = computeBlockDiagonal(mesh, material);
  invertBlockDiagonal(blk, constrained);
  const assemblyTimeMs = Date.now() - tA;
  const Fc = new Float32Array(F);
  applyBCsToRHS(Fc, constrained);
  const applyK = function (x, y) { applyGlobalK(mesh, material, coloring, x, y, constrained); };
  const tS = process.hrtime.bigint();
  const res = solvePCG(applyK, Fc, { tolerance: tolerance, maxIterations: maxIterations, blockPreconditioner: blk });
  const solveTimeMs = Number(process.hrtime.bigint() - tS) / 1e6;
  clearKeCache();
  res.solveTimeMs = solveTimeMs;
  res.assemblyTimeMs = assemblyTimeMs;
  res.totalTimeMs = Date.now() - t0;
  return res;
}

// Split every quad (n0,n1,n2,n3) of a structured mesh into the triangles
// (n0,n1,n2) and (n0,n2,n3).  The reference has no such mesher; the batched
// workload defines its DKT plates this way.
function splitQuadMesh(q) {
  const tri = new Uint32Array(q.elementCount * 6);
  for (let e = 0; e < q.elementCount; e++) {
    const n0 = q.elements[e * 4], n1 = q.elements[e * 4 + 1];
    const n2 = q.elements[e * 4 + 2], n3 = q.elements[e * 4 + 3];
    tri.set([n0, n1, n2, n0, n2, n3], e * 6);
  }
  return {
    nodes: q.nodes, elements: tri, boundaryNodes: q.boundaryNodes, nodeCount: q.nodeCount,
    elementCount: q.elementCount * 2, nodesPerElement: 3, gridNx: q.gridNx, gridNy: q.gridNy,
  };
}

// Command line: node solver_f32.js spec.json
//   spec: {nx, ny, width, height, E, nu, t, triangles, supports, F_path,
//          tolerance, maxIterations, out_path}
// F_path holds the float32 load vector; the float32 solution goes to out_path
// and one JSON line of statistics goes to stdout.
if (require.main === module) {
  const fs = require('fs');
  const spec = JSON.parse(fs.readFileSync(process.argv[2], 'utf8'));
  let mesh = rectangularMeshN(spec.width, spec.height, spec.nx, spec.ny);
  if (spec.triangles) mesh = splitQuadMesh(mesh);
  const buf = fs.readFileSync(spec.F_path);
  const F = new Float32Array(buf.buffer.slice(buf.byteOffset, buf.byteOffset + buf.byteLength));
  const material = { E: spec.E, nu: spec.nu, t: spec.t };
  const res = solveWithF(mesh, material, spec.supports, F, spec.tolerance, spec.maxIterations);
  if (spec.out_path) fs.writeFileSync(spec.out_path, Buffer.from(res.solution.buffer));
  const os = require('os');
  process.stdout.write(JSON.stringify({
    iterations: res.iterations, finalResidual: res.finalResidual, converged: res.converged,
    solveTimeMs: res.solveTimeMs, assemblyTimeMs: res.assemblyTimeMs, totalTimeMs: res.totalTimeMs,
    node: process.version, cpus: os.cpus().length, cpuModel: os.cpus()[0].model,
  }) + '\n');
}

module.exports = {
  splitQuadMesh,
  generateRectangularMesh, rectangularMeshN, getElementCoords, getElementNodeIndices,
  computeElementColoring, computeElementColoringGreedy, identifyConstrainedDOFs,
  computeElementStiffness, computeDKTStiffness, computeBlockDiagonal, invertBlockDiagonal,
  applyGlobalK, clearKeCache, solvePCG, solveTestSystem, buildLoadVector, applyBCsToRHS,
  applyBlockPrecond, solveWithF, dot, norm,
};
