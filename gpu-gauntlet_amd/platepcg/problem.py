"""Problem set-up on the host: meshes, colourings, supports and the workloads.

These mirror the reference's set-up layer (src/lib/plate) that the caller of
solveGPU runs before the solve:
  rect_mesh            mesher.ts:364-439 (generateRectangularMesh, n given)
  split_quads          no reference counterpart (triangles for DKT plates)
  colour_structured    coloring.ts:38-81 (computeElementColoring)
  colour_greedy        coloring.ts:92-177 (computeElementColoringGreedy)
  constrained_mask     solver.ts:92-134 (identifyConstrainedDOFs), as a
                       per-node bit mask (PP_FIX_*)
and define the synthetic workloads (uniform and Gaussian-patch pressures,
random per-plate materials, a log-normal thickness field) with a SplitMix64
generator so every input is reproducible from its seed.
"""
import math

import numpy as np

from ._lib import PP_FIX_THX, PP_FIX_THY, PP_FIX_W

E_STEEL = 210e9
NU_STEEL = 0.3


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014); uniform() gives 53-bit [0,1)."""

    MASK = (1 << 64) - 1

    def __init__(self, seed):
        self.state = seed & self.MASK

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & self.MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))

    def normal(self):
        # Box-Muller on two uniforms
        u1 = max(self.uniform(), 1e-300)
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


def splitmix_normals(seed, n):
    """n standard normals from the SplitMix64 stream of `seed` (Box-Muller on
    consecutive pairs of 53-bit uniforms), vectorised; element i equals what
    SplitMix64(seed) would draw as its i-th normal()."""
    gamma = np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + np.arange(1, 2 * n + 1, dtype=np.uint64) * gamma
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
    u1 = np.maximum(u[0::2], 1e-300)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u[1::2])


def gaussian_field(nx, ny, width, height, corr_len, seed):
    """Unit-variance Gaussian random field on the element centres of an
    nx x ny grid, squared-exponential covariance exp(-d^2 / (2 l^2)) with
    l = corr_len, by spectral synthesis: white noise on a grid padded to
    twice the size (no wrap-around correlation), filtered by the square root
    of the spectral density exp(-l^2 |k|^2 / 2), cropped and normalised to
    unit sample variance."""
    hx, hy = width / nx, height / ny
    Nx, Ny = 2 * nx, 2 * ny
    w = splitmix_normals(seed, Nx * Ny).reshape(Ny, Nx)
    kx = 2 * np.pi * np.fft.fftfreq(Nx, d=hx)
    ky = 2 * np.pi * np.fft.fftfreq(Ny, d=hy)
    K2 = ky[:, None] ** 2 + kx[None, :] ** 2
    g = np.real(np.fft.ifft2(np.fft.fft2(w) * np.exp(-0.25 * corr_len ** 2 * K2)))[:ny, :nx]
    g -= g.mean()
    return g / g.std()


def rect_mesh(nx, ny=None, width=10.0, height=None):
    """Structured quad mesh; node coordinates are float32(i * (width / nx))."""
    ny = nx if ny is None else ny
    height = width if height is None else height
    dx, dy = width / nx, height / ny
    W = nx + 1
    xs = (np.arange(nx + 1, dtype=np.float64) * dx).astype(np.float32)
    ys = (np.arange(ny + 1, dtype=np.float64) * dy).astype(np.float32)
    nodes = np.empty(((ny + 1) * W, 2), np.float32)
    nodes[:, 0] = np.tile(xs, ny + 1)
    nodes[:, 1] = np.repeat(ys, W)
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    n0 = (j * W + i).ravel()
    elements = np.stack([n0, n0 + 1, n0 + W + 1, n0 + W], axis=1).astype(np.uint32)
    b = list(range(nx + 1))
    b += [jj * W + nx for jj in range(1, ny + 1)]
    b += [ny * W + ii for ii in range(nx - 1, -1, -1)]
    b += [jj * W for jj in range(ny - 1, 0, -1)]
    return {
        "nodes": nodes.ravel(), "elements": elements.ravel(),
        "boundaryNodes": np.array(b, np.uint32), "nodeCount": (nx + 1) * (ny + 1),
        "elementCount": nx * ny, "nodesPerElement": 4, "gridNx": nx, "gridNy": ny,
    }


def split_quads(mesh):
    """Each quad (n0,n1,n2,n3) -> triangles (n0,n1,n2), (n0,n2,n3)."""
    q = mesh["elements"].reshape(-1, 4)
    tri = np.stack([q[:, [0, 1, 2]], q[:, [0, 2, 3]]], axis=1).reshape(-1)
    out = dict(mesh)
    out.update(elements=tri.astype(np.uint32), elementCount=2 * mesh["elementCount"],
               nodesPerElement=3)
    return out


def colour_structured(mesh):
    """c = (i % 2) * 2 + (j % 2) with i = e % nx, j = e // nx."""
    e = np.arange(mesh["elementCount"])
    i, j = e % mesh["gridNx"], e // mesh["gridNx"]
    colors = ((i % 2) * 2 + (j % 2)).astype(np.uint8)
    return _pack_colours(colors, 4)


def colour_split_quads(mesh):
    """Split-quad triangles (2 per cell, interleaved): 8 colours
    c = ((i % 2) * 2 + (j % 2)) * 2 + triangle; triangles of one colour
    share no node (same-parity cells are two cells apart)."""
    e = np.arange(mesh["elementCount"])
    cell, tri = e // 2, e % 2
    i, j = cell % mesh["gridNx"], cell // mesh["gridNx"]
    colors = (((i % 2) * 2 + (j % 2)) * 2 + tri).astype(np.uint8)
    return _pack_colours(colors, 8)


def colour_greedy(mesh):
    """Smallest colour not used by an element sharing a node, in element order."""
    npe = mesh["nodesPerElement"]
    el = mesh["elements"].reshape(-1, npe)
    node_elems = [[] for _ in range(mesh["nodeCount"])]
    for e, row in enumerate(el):
        for n in row:
            node_elems[n].append(e)
    colors = np.full(len(el), 255, np.uint8)
    for e, row in enumerate(el):
        used = set()
        for n in row:
            for nb in node_elems[n]:
                if nb != e and colors[nb] != 255:
                    used.add(int(colors[nb]))
        c = 0
        while c in used:
            c += 1
        colors[e] = c
    return _pack_colours(colors, int(colors.max()) + 1)


def _pack_colours(colors, count):
    order = np.argsort(colors, kind="stable").astype(np.uint32)
    counts = np.bincount(colors, minlength=count)
    offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32)
    return {"elements_by_colour": order, "colour_offsets": offsets, "colour_count": count,
            "elementColors": colors}


def constrained_mask(mesh, supports):
    """supports: list of {"type": "fixed"|"pinned"|"roller", "location": node|"all_edges"}."""
    mask = np.zeros(mesh["nodeCount"], np.uint8)
    for s in supports:
        nodes = mesh["boundaryNodes"] if s["location"] == "all_edges" else [s["location"]]
        bits = PP_FIX_W | PP_FIX_THX | PP_FIX_THY if s["type"] == "fixed" else PP_FIX_W
        mask[np.asarray(nodes, np.int64)] |= bits
    return mask


def edge_mask(nx, ny, edge_types):
    """Per-edge supports (bottom, right, top, left), each "fixed" or "pinned"."""
    W = nx + 1
    mask = np.zeros((ny + 1) * W, np.uint8)
    edges = [np.arange(nx + 1), np.arange(ny + 1) * W + nx, ny * W + np.arange(nx + 1),
             np.arange(ny + 1) * W]
    for nodes, kind in zip(edges, edge_types):
        mask[nodes] |= (PP_FIX_W | PP_FIX_THX | PP_FIX_THY) if kind == "fixed" else PP_FIX_W
    return mask


def mask_to_dofs(mask):
    """Per-node bits -> sorted constrained DOF indices (the reference's Set<number>)."""
    dofs = []
    for d, bit in enumerate((PP_FIX_W, PP_FIX_THX, PP_FIX_THY)):
        dofs.append(np.nonzero(mask & bit)[0] * 3 + d)
    return np.sort(np.concatenate(dofs)).astype(np.int64)


def dofs_to_mask(n_nodes, dofs):
    mask = np.zeros(n_nodes, np.uint8)
    for dof in dofs:
        mask[dof // 3] |= 1 << (dof % 3)
    return mask


# ------------------------------------------------------------- workloads

def batch_plates(count=256, seed=4, n_min=16, n_max=200):
    """The batched workload: `count` 10 m x 10 m plates from one SplitMix64
    stream.  Per plate, in draw order: side n = floor(exp(U(ln n_min,
    ln(n_max + 1)))) clamped to [n_min, n_max]; t ~ U[0.01, 0.3] m;
    nu ~ U[0.25, 0.35]; each edge (bottom, right, top, left) fixed or pinned
    with probability 1/2; k in {1, 2, 3} Gaussian patches (peak U[10, 50] kPa,
    sigma 1 m, centre uniform) on 10 kPa.  Even plates are Q4, odd plates DKT
    (each quad split into (n0,n1,n2), (n0,n2,n3))."""
    rng = SplitMix64(seed)
    plates = []
    for i in range(count):
        n = int(math.floor(math.exp(rng.uniform(math.log(n_min), math.log(n_max + 1)))))
        n = min(max(n, n_min), n_max)
        t = rng.uniform(0.01, 0.3)
        nu = rng.uniform(0.25, 0.35)
        edges = ["fixed" if rng.uniform() < 0.5 else "pinned" for _ in range(4)]
        k = 1 + int(rng.uniform() * 3)
        patches = np.array([[rng.uniform(10e3, 50e3), rng.uniform(0, 10), rng.uniform(0, 10), 1.0]
                            for _ in range(k)])
        mesh = rect_mesh(n)
        if i % 2:
            mesh = split_quads(mesh)
        plates.append(dict(mesh=mesh, E=E_STEEL, nu=nu, t=t, mask=edge_mask(n, n, edges),
                           edges=edges, q0=10e3, patches=patches, kind="dkt" if i % 2 else "q4",
                           n=n))
    return plates


def workload(name, seed=None, element_t=None):
    """The named synthetic workloads.  Returns a dict with mesh, material,
    mask, q0 and patches (A, cx, cy, sigma per Gaussian patch).  For "c5",
    element_t may supply the thickness field already made (on the device by
    platepcg.lognormal_thickness) instead of the host synthesis here."""
    if name == "c1":
        mesh = rect_mesh(32)
        return dict(name=name, mesh=mesh, E=E_STEEL, nu=NU_STEEL, t=0.1,
                    mask=constrained_mask(mesh, [{"type": "fixed", "location": "all_edges"}]),
                    q0=10e3, patches=np.zeros((0, 4)))
    if name == "c2":
        rng = SplitMix64(1 if seed is None else seed)
        cx, cy = rng.uniform(0, 10), rng.uniform(0, 10)
        mesh = rect_mesh(142)
        return dict(name=name, mesh=mesh, E=E_STEEL, nu=NU_STEEL, t=0.1,
                    mask=constrained_mask(mesh, [{"type": "fixed", "location": "all_edges"}]),
                    q0=10e3, patches=np.array([[50e3, cx, cy, 1.0]]))
    if name == "c5":
        # 1024 x 1024, t_e = 0.1 exp(0.5 g(x_e)), g unit-variance, correlation
        # length 0.5 m (seed 5); 10 kPa + three Gaussian patches
        n = 1024
        mesh = rect_mesh(n)
        if element_t is None:
            g = gaussian_field(n, n, 10.0, 10.0, 0.5, 5 if seed is None else seed)
            element_t = (0.1 * np.exp(0.5 * g)).astype(np.float32).ravel()
        rng = SplitMix64(1005 if seed is None else seed + 1000)
        patches = np.array([[rng.uniform(10e3, 50e3), rng.uniform(0, 10), rng.uniform(0, 10), 1.0]
                            for _ in range(3)])
        return dict(name=name, mesh=mesh, E=E_STEEL, nu=NU_STEEL, t=0.1,
                    element_t=np.ascontiguousarray(element_t, np.float32).ravel(),
                    mask=constrained_mask(mesh, [{"type": "fixed", "location": "all_edges"}]),
                    q0=10e3, patches=patches)
    if name == "c3":
        mesh = rect_mesh(700)
        return dict(name=name, mesh=mesh, E=E_STEEL, nu=NU_STEEL, t=0.05,
                    mask=constrained_mask(mesh, [{"type": "fixed", "location": "all_edges"}]),
                    q0=10e3, patches=np.zeros((0, 4)))
    raise KeyError(name)
