"""CPU oracle of the Galerkin multigrid V-cycle the learned preconditioner can
use as its smoother (scipy.sparse, FP64).

TEST INFRASTRUCTURE ONLY: the checker the HIP multigrid is compared against.
Only tests/ may import it; the product never does.

Built independently of the library's construction: level 0 is the assembled
FP64 stiffness K of the structured Q4 plate (the oracle's own element matrix,
element.ts:258-483, per element) with the reference's constraint rows
(solver.ts:479-482: identity on constrained DOFs); P_l is the global
bilinear interpolation matrix (index space, per DOF component) times the
0/1 mask of the DOFs kept on level l; A_{l+1} = P_l^T A_l P_l is formed as a
sparse triple product (not from coarse element matrices, as the library
does); a coarse DOF with a zero diagonal gets an identity row and is masked on
its level.  The coarsest level is solved exactly.  The cycle, with nu sweeps
of damped block Jacobi (w_l B_l, B_l the inverse of the 3x3 diagonal blocks)
before and after the coarse correction, is the library's:

  x = S_nu(b, 0);  b_c = P^T (b - A x);  x += P V_{l+1}(b_c);  x = S_nu(b, x),

S_nu nu damped block-Jacobi sweeps x += w B (b - A x) or nu Chebyshev steps
(smooth below).

Coarsening rule (the library's): n -> (n + 1) / 2 elements per side (coarse
node c on fine node min(2c, n), so an odd count ends in a coarse interval one
fine element wide) while both counts are at least 2, stopping at the first
level (below 0) with at most 1,200 DOFs.
"""
import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp

DENSE_STOP = 1200


def assemble_q4(nodes, nx, ny, E, nu, t0, elem_t, ke_q4):
    """K (no constraint rows) of the structured grid; ke_q4(xy, E, nu, t)
    returns the 12 x 12 element matrix."""
    W = nx + 1
    xy = np.asarray(nodes, np.float64).reshape(-1, 2)
    rows, cols, vals = [], [], []
    for j in range(ny):
        for i in range(nx):
            n = [j * W + i, j * W + i + 1, (j + 1) * W + i + 1, (j + 1) * W + i]
            t = t0 if elem_t is None else float(elem_t[j * nx + i])
            Ke = ke_q4(xy[n], E, nu, t)
            d = np.array([3 * a + c for a in n for c in range(3)])
            rows.append(np.repeat(d, 12))
            cols.append(np.tile(d, 12))
            vals.append(Ke.ravel())
    nd = 3 * W * (ny + 1)
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(nd, nd))


def with_bc(K, free):
    """D K D + (I - D), D = diag(free)."""
    D = sp.diags(free.astype(np.float64))
    return (D @ K @ D + sp.diags(1.0 - free)).tocsr()


def interp_1d(n):
    """(n + 1) x ((n + 1) // 2 + 1): fine node i on coarse node i / 2 when i
    is even, on the last coarse node when i = n is odd, else half of each
    neighbouring coarse node."""
    m = (n + 1) // 2
    r, c, v = [], [], []
    for i in range(n + 1):
        if i % 2 == 0:
            r.append(i); c.append(i // 2); v.append(1.0)
        elif i == n:
            r.append(i); c.append(m); v.append(1.0)
        else:
            r += [i, i]; c += [(i - 1) // 2, (i + 1) // 2]; v += [0.5, 0.5]
    return sp.csr_matrix((v, (r, c)), shape=(n + 1, m + 1))


def prolongation(nx, ny, keep):
    """Masked bilinear P of a level with nx x ny elements; node index
    j (nx + 1) + i, DOF 3 node + d."""
    P2 = sp.kron(interp_1d(ny), interp_1d(nx))
    return (sp.diags(keep.astype(np.float64)) @ sp.kron(P2, sp.eye(3))).tocsr()


def block_inverse(A, keep):
    """Inverse 3x3 diagonal blocks (excluded DOFs: identity rows and columns)."""
    n = A.shape[0] // 3
    idx = np.arange(n) * 3
    blk = np.zeros((n, 3, 3))
    for a in range(3):
        for b in range(3):
            blk[:, a, b] = np.asarray(A[idx + a, idx + b]).ravel()
    k = keep.reshape(n, 3)
    for a in range(3):
        off = ~k[:, a]
        blk[off, a, :] = 0.0
        blk[off, :, a] = 0.0
        blk[off, a, a] = 1.0
    return np.linalg.inv(blk)


def hierarchy_dims(nx, ny):
    dims = [(nx, ny)]
    while True:
        a, b = dims[-1]
        if len(dims) > 1 and 3 * (a + 1) * (b + 1) <= DENSE_STOP:
            break
        if a < 2 or b < 2:
            break
        dims.append(((a + 1) // 2, (b + 1) // 2))
    return dims


class MgOracle:
    """A0: level-0 operator with constraint rows; keep0: bool per DOF (not
    constrained); blk0: level-0 block inverse (n_nodes x 3 x 3) or None;
    omegas: the smoothing weight of every level but the coarsest (None:
    4 / (3 * 1.05 * lambda_max(B A)) from the exact eigenvalue); factor:
    factor the coarsest level for the cycle."""

    def __init__(self, A0, nx, ny, keep0, sweeps=2, omegas=None, blk0=None, factor=True,
                 coarse_sweeps=None, chebyshev=False):
        self.nu = sweeps  # on level 0; coarse_sweeps (None: the same) on the others
        self.nuc = sweeps if coarse_sweeps is None else coarse_sweeps
        self.chebyshev = chebyshev
        self.dims = hierarchy_dims(nx, ny)
        self.A, self.P, self.B, self.keep = [A0], [], [], [keep0]
        for l in range(len(self.dims) - 1):
            a, b = self.dims[l]
            P = prolongation(a, b, self.keep[l])
            Ac = (P.T @ self.A[l] @ P).tocsr()
            dg = Ac.diagonal()
            dead = dg == 0.0
            Ac = (Ac + sp.diags(dead.astype(np.float64))).tocsr()
            self.P.append(P)
            self.A.append(Ac)
            self.keep.append(~dead)
        for l in range(len(self.dims) - 1):
            if l == 0 and blk0 is not None:
                self.B.append(np.asarray(blk0, np.float64).reshape(-1, 3, 3))
            else:
                self.B.append(block_inverse(self.A[l], self.keep[l]))
        # factor=False: operators only (a free plate's levels are singular)
        self.coarse = sla.cho_factor(self.A[-1].toarray()) if factor else None
        if omegas is None:
            omegas = [4.0 / (3.0 * 1.05 * self.lambda_max(l)) for l in range(len(self.B))]
        self.omega = list(omegas)

    def bj(self, l, v):
        return np.einsum("nij,nj->ni", self.B[l], v.reshape(-1, 3)).reshape(-1)

    def lambda_max(self, l):
        from scipy.sparse.linalg import LinearOperator, eigs
        n = self.A[l].shape[0]
        op = LinearOperator((n, n), matvec=lambda v: self.bj(l, self.A[l] @ v))
        return float(np.max(np.real(eigs(op, k=1, which="LR", tol=1e-8)[0])))

    def smooth(self, l, b, x, nu):
        """nu damped block-Jacobi sweeps, or nu steps of the Chebyshev
        semi-iteration for B A x = B b on [hi / 16, hi], hi = 1.05 lambda_max
        with lambda_max = 4 / (3 * 1.05 w) from the level's weight w."""
        A, w = self.A[l], self.omega[l]
        if not self.chebyshev:
            for _ in range(nu):
                x = x + w * self.bj(l, b - A @ x)
            return x
        hi = 1.05 * 4.0 / (3.0 * 1.05 * w)
        lo = hi / 16.0
        theta, delta = (hi + lo) / 2, (hi - lo) / 2
        sigma = theta / delta
        rho = 1.0 / sigma
        d = self.bj(l, b - A @ x) / theta
        x = x + d
        for _ in range(nu - 1):
            rn = 1.0 / (2.0 * sigma - rho)
            d = rn * rho * d + 2.0 * rn / delta * self.bj(l, b - A @ x)
            rho = rn
            x = x + d
        return x

    def cycle(self, b, l=0):
        if l == len(self.A) - 1:
            return sla.cho_solve(self.coarse, b)
        A = self.A[l]
        nu = self.nu if l == 0 else self.nuc
        x = self.smooth(l, b, np.zeros_like(b), nu)
        bc = self.P[l].T @ (b - A @ x)
        x = x + self.P[l] @ self.cycle(bc, l + 1)
        return self.smooth(l, b, x, nu)
