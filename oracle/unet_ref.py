"""CPU oracle of the learned preconditioner (numpy).

TEST INFRASTRUCTURE ONLY: the checker the HIP network is compared against.
Only tests/ may import it; the product never does.

It restates the library's U-net preconditioner and its training loss:

  M(r) = z + s * o / c,  z = B^-1 r (the 3x3 nodal block-Jacobi inverse of
  solver.ts:231-304 applied per node), c = (c_w, c_t, c_t) channel scales,
  s = sqrt(sum_n |c z_n|^2 / n_nodes), o = U(x), x = (c z / s, log(t / t0) c_w z_w / s)

U is a 3-level U-net on the node grid padded to 4k+1 per side: 3x3 convs
(zero padding, dilation d), ReLU, full-weighting restriction R and bilinear
prolongation P = 4 R^T between the vertex-centred levels, skip concatenation
[P(deeper) | skip].  Loss of a pair (e, r = K e): (M(r) - e)^T K (M(r) - e) /
e^T K e, M the network between two damped block-Jacobi smoothing steps
(precondition below).  The gradient is derived by hand (reverse mode through every op);
with quantize=False everything is float64 and is checked against central
differences in the tests, so the backward pass is pinned by the mathematics.
quantize=True rounds to bfloat16 at the points where the GPU stores BF16
(inputs, weights, every activation and activation gradient), so GPU and
oracle agree to accumulation-order effects.

Parameter layout (as the library's): per conv layer, weights [cout][3][3][cin]
(tap = 3 (dy + 1) + (dx + 1)) then bias [cout]; layers e0a e0b e1a e1b ba bb
d1a d1b d0a d0b head.
"""
import numpy as np


def bf16(a):
    """Round to bfloat16 (nearest even), returned as float64."""
    a32 = np.ascontiguousarray(a, np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def layer_shapes(c=(16, 32, 64), bdil=1):
    C0, C1, C2 = c
    return [("e0a", 4, C0, 1, 1), ("e0b", C0, C0, 1, 1), ("e1a", C0, C1, 1, 1),
            ("e1b", C1, C1, 1, 1), ("ba", C1, C2, 1, 1), ("bb", C2, C2, 1, bdil),
            ("d1a", C2 + C1, C1, 1, 1), ("d1b", C1, C1, 1, 1), ("d0a", C1 + C0, C0, 1, 1),
            ("d0b", C0, C0, 1, 1), ("head", C0, 3, 0, 1)]


def unpack(params, shapes):
    out, off = {}, 0
    for name, cin, cout, relu, dil in shapes:
        n = cout * 9 * cin
        out[name] = (params[off:off + n].reshape(cout, 3, 3, cin).astype(np.float64),
                     params[off + n:off + n + cout].astype(np.float64), relu, dil)
        off += n + cout
    assert off == params.size
    return out


def pack(grads, shapes):
    return np.concatenate([np.concatenate([grads[n][0].ravel(), grads[n][1].ravel()])
                           for n, *_ in shapes])


def _shift(x, dy, dx):
    """y[h, w] = x[h + dy, w + dx], zero outside."""
    H, W = x.shape[:2]
    y = np.zeros_like(x)
    ys, ye = max(0, -dy), min(H, H - dy)
    xs, xe = max(0, -dx), min(W, W - dx)
    if ys < ye and xs < xe:
        y[ys:ye, xs:xe] = x[ys + dy:ye + dy, xs + dx:xe + dx]
    return y


def conv(x, w, b, dil):
    H, W, _ = x.shape
    y = np.zeros((H, W, w.shape[0])) + b
    for ky in range(3):
        for kx in range(3):
            y += _shift(x, (ky - 1) * dil, (kx - 1) * dil) @ w[:, ky, kx, :].T
    return y


def conv_back(x, w, dil, gy, need_gx=True):
    """Gradients of conv in w, b and (optionally) x."""
    gw = np.zeros_like(w)
    gx = np.zeros_like(x) if need_gx else None
    for ky in range(3):
        for kx in range(3):
            dy, dx = (ky - 1) * dil, (kx - 1) * dil
            xs = _shift(x, dy, dx)
            gw[:, ky, kx, :] = gy.reshape(-1, gy.shape[2]).T @ xs.reshape(-1, xs.shape[2])
            if need_gx:
                gx += _shift(gy @ w[:, ky, kx, :], -dy, -dx)
    return gw, gy.sum(axis=(0, 1)), gx


def restrict(x):
    """Full weighting, fine (2k+1) -> coarse (k+1), zero outside."""
    Hf, Wf, C = x.shape
    Hc, Wc = (Hf + 1) // 2, (Wf + 1) // 2
    wts = {-1: 1.0, 0: 2.0, 1: 1.0}
    y = np.zeros((Hc, Wc, C))
    xp = np.zeros((Hf + 2, Wf + 2, C))
    xp[1:-1, 1:-1] = x
    for b in (-1, 0, 1):
        for a in (-1, 0, 1):
            y += wts[a] * wts[b] * xp[1 + b:1 + b + 2 * Hc - 1:2, 1 + a:1 + a + 2 * Wc - 1:2]
    return y / 16.0


def prolong(x):
    """Bilinear, coarse (k+1) -> fine (2k+1)."""
    Hc, Wc, C = x.shape
    y = np.zeros((2 * Hc - 1, 2 * Wc - 1, C))
    y[::2, ::2] = x
    y[1::2, ::2] = 0.5 * (x[:-1] + x[1:])
    y[::2, 1::2] = 0.5 * (x[:, :-1] + x[:, 1:])
    y[1::2, 1::2] = 0.25 * (x[:-1, :-1] + x[1:, :-1] + x[:-1, 1:] + x[1:, 1:])
    return y


class UNetOracle:
    """The preconditioner on one structured plate.

    nx, ny: elements per side; blk: 9 doubles per node (block-Jacobi inverse);
    cmask: PP_FIX bits per node; elem_t: per-element thickness or None.
    """

    def __init__(self, nx, ny, blk, cmask, t0, elem_t=None, channels=(16, 32, 64), bdil=1,
                 rotation_scale=0.3, quantize=True, relu=True):
        self.nx, self.ny = nx, ny
        self.relu = relu  # False: no activation and no biases (a linear operator)
        self.blk = np.asarray(blk, np.float64).reshape(-1, 3, 3)
        self.cmask = np.asarray(cmask, np.uint8)
        self.shapes = layer_shapes(channels, bdil)
        self.c = np.array([1.0, rotation_scale, rotation_scale])
        self.q = bf16 if quantize else (lambda a: np.asarray(a, np.float64))
        # float32 storage points of the library (inputs, head output)
        self.f32 = ((lambda a: np.asarray(a, np.float32).astype(np.float64)) if quantize
                    else (lambda a: np.asarray(a, np.float64)))
        self.W0 = (nx + 3) // 4 * 4 + 1
        self.H0 = (ny + 3) // 4 * 4 + 1
        et = np.full(nx * ny, t0) if elem_t is None else np.asarray(elem_t, np.float64)
        lt = np.log(et / t0).reshape(ny, nx)
        acc = np.zeros((ny + 1, nx + 1))
        cnt = np.zeros((ny + 1, nx + 1))
        for b in (0, 1):
            for a in (0, 1):
                acc[b:b + ny, a:a + nx] += lt
                cnt[b:b + ny, a:a + nx] += 1
        self.logt = (acc / cnt).astype(np.float32).astype(np.float64)
        free = np.zeros((ny + 1, nx + 1, 3), bool)
        for d in range(3):
            free[..., d] = ((self.cmask.reshape(ny + 1, nx + 1) >> d) & 1) == 0
        self.free = free

    # ---- forward
    def forward(self, params, r):
        q = self.q
        P = unpack(self.f32(params), self.shapes)
        Wq = {k: (q(w), b, relu, dil) for k, (w, b, relu, dil) in P.items()}
        ny, nx = self.ny, self.nx
        z = np.einsum("nij,nj->ni", self.blk, np.asarray(r, np.float64).reshape(-1, 3))
        cz = z * self.c
        s = np.sqrt(np.sum(cz * cz) / z.shape[0])
        s = s if s > 0 else 1.0
        x = np.zeros((self.H0, self.W0, 4))
        x[:ny + 1, :nx + 1, :3] = self.f32(cz / s).reshape(ny + 1, nx + 1, 3)
        x[:ny + 1, :nx + 1, 3] = self.logt * x[:ny + 1, :nx + 1, 0]
        x = q(x)
        t = {"x": x}

        def cv(name, inp):
            w, b, relu, dil = Wq[name]
            y = conv(inp, w, b if self.relu else 0.0 * b, dil)
            if relu and self.relu:
                y = np.maximum(y, 0.0)
            return y

        t["a0"] = q(cv("e0a", x))
        t["s0"] = q(cv("e0b", t["a0"]))
        t["r1"] = q(restrict(t["s0"]))
        t["a1"] = q(cv("e1a", t["r1"]))
        t["s1"] = q(cv("e1b", t["a1"]))
        t["r2"] = q(restrict(t["s1"]))
        t["a2"] = q(cv("ba", t["r2"]))
        t["b"] = q(cv("bb", t["a2"]))
        t["cat1"] = np.concatenate([q(prolong(t["b"])), t["s1"]], axis=2)
        t["c1"] = q(cv("d1a", t["cat1"]))
        t["d1"] = q(cv("d1b", t["c1"]))
        t["cat0"] = np.concatenate([q(prolong(t["d1"])), t["s0"]], axis=2)
        t["c0"] = q(cv("d0a", t["cat0"]))
        t["c0b"] = q(cv("d0b", t["c0"]))
        o = self.f32(cv("head", t["c0b"]))
        t["o"] = o
        M = z.copy().reshape(ny + 1, nx + 1, 3)
        M += np.where(self.free, s * o[:ny + 1, :nx + 1, :3] / self.c, 0.0)
        self.cache = (Wq, t, s, z)
        return M.reshape(-1)

    def smooth(self, v, omega):
        """omega B v (block-Jacobi, per node)."""
        return omega * np.einsum("nij,nj->ni", self.blk, v.reshape(-1, 3)).reshape(-1)

    def precondition(self, params, r, apply_k, omega, smoother=None):
        """z0 = S r; z1 = z0 + N(r - K z0); z = z1 + S (r - K z1), S = w B or
        the given symmetric smoother (v -> S v, e.g. a multigrid V-cycle)."""
        S = smoother if smoother is not None else (lambda v: self.smooth(v, omega))
        r = np.asarray(r, np.float64)
        z0 = S(r)
        u = self.forward(params, r - apply_k(z0))
        z1 = z0 + u
        return z1 + S(r - apply_k(z1))

    # ---- loss and gradient
    def loss_grad(self, params, e, apply_k, omega, smoother=None):
        """Loss e_f^T K e_f / e^T K e of the pair (e, r = K e), e_f = e - M(r)
        with the smoothed preconditioner above, and its gradient in the flat
        parameters.  With E = I - S K (S = w B or the symmetric smoother
        given), e_f = E (e - z0 - u) for the network output u, so dL/du =
        -2 E^T K e_f / e^T K e, E^T K e_f = K e_f - K S K e_f.  apply_k: the
        FP64 operator (with the reference's BC rows)."""
        q = self.q
        S = smoother if smoother is not None else (lambda v: self.smooth(v, omega))
        e = np.asarray(e, np.float64)
        r = apply_k(e)
        ef = e - self.precondition(params, r, apply_k, omega, smoother)
        q1 = apply_k(ef)
        eke = float(e @ r)
        loss = float(ef @ q1) / eke
        Wq, t, s, z = self.cache
        ny, nx = self.ny, self.nx
        g = q1 - apply_k(S(q1))
        gM = (-2.0 / eke) * g.reshape(ny + 1, nx + 1, 3)
        go = np.zeros((self.H0, self.W0, 4))
        go[:ny + 1, :nx + 1, :3] = np.where(self.free, gM * s / self.c, 0.0)
        go = q(go)
        G = {}

        def back(name, inp, gy, need_gx=True):
            w, _, _, dil = Wq[name]
            gw, gb, gx = conv_back(inp, w, dil, gy[..., :w.shape[0]], need_gx)
            G[name] = (gw, gb if self.relu else 0.0 * gb)
            return gx

        if self.relu:
            relu = lambda y: (y > 0).astype(np.float64)
        else:
            relu = lambda y: 1.0
        C1 = t["d1"].shape[2]
        C2 = t["b"].shape[2]
        g = q(back("head", t["c0b"], go) * relu(t["c0b"]))
        g = q(back("d0b", t["c0"], g) * relu(t["c0"]))
        gcat0 = back("d0a", t["cat0"], g)
        gs0 = q(gcat0[..., C1:] * relu(t["s0"]))
        gd1 = q(4.0 * restrict(q(gcat0[..., :C1])) * relu(t["d1"]))
        g = q(back("d1b", t["c1"], gd1) * relu(t["c1"]))
        gcat1 = back("d1a", t["cat1"], g)
        gs1 = q(gcat1[..., C2:] * relu(t["s1"]))
        gb_ = q(4.0 * restrict(q(gcat1[..., :C2])) * relu(t["b"]))
        g = q(back("bb", t["a2"], gb_) * relu(t["a2"]))
        gr2 = q(back("ba", t["r2"], g))
        gs1 = q((0.25 * prolong(gr2) + gs1) * relu(t["s1"]))
        g = q(back("e1b", t["a1"], gs1) * relu(t["a1"]))
        gr1 = q(back("e1a", t["r1"], g))
        gs0 = q((0.25 * prolong(gr1) + gs0) * relu(t["s0"]))
        g = q(back("e0b", t["a0"], gs0) * relu(t["a0"]))
        back("e0a", t["x"], g, need_gx=False)
        return loss, pack(G, self.shapes)


# ---------------------------------------------------------------- data

def _splitmix(state):
    state = (state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return state, z ^ (z >> 31)


class _Rng:
    def __init__(self, seed):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def uni(self):
        self.s, z = _splitmix(self.s)
        return (z >> 11) * (1.0 / 9007199254740992.0)

    def gauss(self):
        u1 = max(self.uni(), 1e-300)
        u2 = self.uni()
        return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def training_error(nodes, nx, ny, cmask, seed, n_bumps=32):
    """The training error field of one seed: 32 Gaussian bumps in w
    (sigma log-uniform in [2h, L/2], amplitude N(0,1) sigma), windowed by
    sin^2 sin^2 over the plate, rotations thx = dw/dy, thy = -dw/dx, zero on
    constrained DOFs."""
    xy = np.asarray(nodes, np.float32).astype(np.float64).reshape(-1, 2)
    xs = xy[:nx + 1, 0]
    ys = xy[::nx + 1, 1][:ny + 1]
    x0, x1, y0, y1 = xs[0], xs[-1], ys[0], ys[-1]
    L = max(x1 - x0, y1 - y0)
    h = min((x1 - x0) / nx, (y1 - y0) / ny)
    rng = _Rng(seed)
    bumps = []
    for _ in range(n_bumps):
        sig = np.exp(np.log(2 * h) + rng.uni() * (np.log(0.5 * L) - np.log(2 * h)))
        a = rng.gauss() * sig
        cx = x0 + rng.uni() * (x1 - x0)
        cy = y0 + rng.uni() * (y1 - y0)
        bumps.append((a, cx, cy, sig))
    X, Y = np.meshgrid(xs, ys)
    f = np.zeros_like(X)
    fx = np.zeros_like(X)
    fy = np.zeros_like(X)
    for a, cx, cy, sig in bumps:
        g = a * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2 * sig * sig))
        f += g
        fx -= g * (X - cx) / (sig * sig)
        fy -= g * (Y - cy) / (sig * sig)
    su, sv = np.sin(np.pi * (X - x0) / (x1 - x0)), np.sin(np.pi * (Y - y0) / (y1 - y0))
    cu, cv = np.cos(np.pi * (X - x0) / (x1 - x0)), np.cos(np.pi * (Y - y0) / (y1 - y0))
    ph = su * su * sv * sv
    phx = 2 * su * cu * (np.pi / (x1 - x0)) * sv * sv
    phy = su * su * 2 * sv * cv * (np.pi / (y1 - y0))
    w = ph * f
    wx = phx * f + ph * fx
    wy = phy * f + ph * fy
    e = np.stack([w, wy, -wx], axis=2).reshape(-1, 3)
    cm = np.asarray(cmask, np.uint8)
    for d in range(3):
        e[(cm >> d) & 1 == 1, d] = 0.0
    return e.reshape(-1)
