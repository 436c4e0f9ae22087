"""Independent matrix-algebra constructions used to pin the oracle.

Nothing here calls the oracle's stencil, sweep or transfer loops.  Every
operator is rebuilt from the grid geometry (cell widths and positions) as a
scipy sparse matrix, the way a textbook would state it:

* the 7-point operator as a Kronecker sum of 1-D cell-centred
  finite-volume second differences with zero Dirichlet data at the faces;
* red/black SOR as the matrix splitting x <- x + (D/w + L)^-1 (b - A x) of
  the system permuted red-first;
* restriction as the Kronecker product of 1-D volume-weighted averages;
* prolongation as the Kronecker product of 1-D linear interpolation between
  coarse nodes (children's shared face, or the domain face with value 0);
* the V-cycle as the composition of those matrices.

Vectors are flattened interiors in C order of (nz, ny, nx) arrays.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def interior(a):
    return a[1:-1, 1:-1, 1:-1]


def flat(a):
    return np.ascontiguousarray(interior(a)).reshape(-1)


def unflat(v, nx, ny, nz):
    out = np.zeros((nz + 2, ny + 2, nx + 2))
    out[1:-1, 1:-1, 1:-1] = v.reshape(nz, ny, nx)
    return out


def widths_uniform(n, l):
    return np.full(n, l / n)


def widths_stretched(n, l, a):
    """Symmetric geometric stretch: first width w1 with ratio a towards the
    middle, mirrored; widths sum to l."""
    h = n // 2
    w = np.empty(n)
    w1 = 0.5 * l * (a - 1.0) / (a ** h - 1.0)
    for i in range(h):
        w[i] = w1 * a ** i
        w[n - 1 - i] = w[i]
    return w


def lap1d(w):
    """1-D cell-centred second difference (u_{i+1}-u_i)/d_{i+1} - (u_i-u_{i-1})/d_i,
    divided by the cell width, with u = 0 on both faces.  d_i is the distance
    between the centres of cells i-1 and i; at a face, centre to face."""
    n = len(w)
    faces = np.concatenate([[0.0], np.cumsum(w)])
    c = 0.5 * (faces[:-1] + faces[1:])
    nodes = np.concatenate([[faces[0]], c, [faces[-1]]])
    d = np.diff(nodes)                       # length n+1: d[0] left face gap, d[n] right
    main = -(1.0 / d[:-1] + 1.0 / d[1:]) / w
    lower = (1.0 / d[1:-1]) / w[1:]          # coupling of i to i-1
    upper = (1.0 / d[1:-1]) / w[:-1]         # coupling of i to i+1
    return sp.diags([lower, main, upper], [-1, 0, 1], shape=(n, n), format="csr")


def laplacian(wx, wy, wz):
    Tx, Ty, Tz = lap1d(wx), lap1d(wy), lap1d(wz)
    Ix, Iy, Iz = (sp.identity(len(v), format="csr") for v in (wx, wy, wz))
    return (sp.kron(sp.kron(Tz, Iy), Ix) + sp.kron(sp.kron(Iz, Ty), Ix) +
            sp.kron(sp.kron(Iz, Iy), Tx)).tocsr()


def red_mask(nx, ny, nz):
    """First-swept colour: 1-based i+j+k odd."""
    k, j, i = np.meshgrid(np.arange(1, nz + 1), np.arange(1, ny + 1), np.arange(1, nx + 1),
                          indexing="ij")
    return ((i + j + k) % 2 == 1).reshape(-1)


def sor_rb(A, b, x, red, omega, iters):
    """Red/black SOR as the textbook splitting on the red-first permutation."""
    perm = np.concatenate([np.nonzero(red)[0], np.nonzero(~red)[0]])
    Ap = A[perm][:, perm].tocsr()
    D = sp.diags(Ap.diagonal())
    Lo = sp.tril(Ap, k=-1)
    M = (D / omega + Lo).tocsr()
    xp = x[perm].copy()
    bp = b[perm]
    for _ in range(iters):
        xp = xp + spla.spsolve_triangular(M, bp - Ap @ xp, lower=True)
    out = np.empty_like(x)
    out[perm] = xp
    return out


def restrict1d(wf, coarsen):
    """Volume-weighted average of the two children (or identity)."""
    nf = len(wf)
    if not coarsen:
        return sp.identity(nf, format="csr")
    nc = nf // 2
    rows, cols, vals = [], [], []
    for c in range(nc):
        a, b = wf[2 * c], wf[2 * c + 1]
        rows += [c, c]
        cols += [2 * c, 2 * c + 1]
        vals += [a / (a + b), b / (a + b)]
    return sp.csr_matrix((vals, (rows, cols)), shape=(nc, nf))


def prolong1d(wf, coarsen):
    """Linear interpolation to fine centres between coarse nodes placed at the
    face shared by the two children, with a zero-valued node at each domain
    face."""
    nf = len(wf)
    if not coarsen:
        return sp.identity(nf, format="csr")
    nc = nf // 2
    faces = np.concatenate([[0.0], np.cumsum(wf)])
    centres = 0.5 * (faces[:-1] + faces[1:])
    # children 2c-1, 2c (1-based) share the face after 2c-1 cells
    node_pos = np.concatenate([[faces[0]], faces[1:-1:2], [faces[-1]]])   # nc+2 nodes
    rows, cols, vals = [], [], []
    for f in range(nf):
        x = centres[f]
        m = np.searchsorted(node_pos, x) - 1          # node_pos[m] <= x < node_pos[m+1]
        x0, x1 = node_pos[m], node_pos[m + 1]
        t = (x - x0) / (x1 - x0)
        for node, wgt in ((m, 1.0 - t), (m + 1, t)):
            if 1 <= node <= nc:                      # nodes 0 and nc+1 carry zero
                rows.append(f)
                cols.append(node - 1)
                vals.append(wgt)
    return sp.csr_matrix((vals, (rows, cols)), shape=(nf, nc))


def kron3(Mz, My, Mx):
    return sp.kron(sp.kron(Mz, My), Mx).tocsr()


class MatrixMG:
    """V-cycle built from the matrices above (reference schedule of
    src/multigrid.f90:1907-1976, stated as linear algebra)."""

    def __init__(self, wx, wy, wz, nlevel):
        self.A, self.R, self.P, self.red, self.dims = [], [], [], [], []
        w = [np.asarray(wx, float), np.asarray(wy, float), np.asarray(wz, float)]
        for l in range(nlevel + 1):
            nx, ny, nz = (len(v) for v in w)
            self.dims.append((nx, ny, nz))
            self.A.append(laplacian(*w))
            self.red.append(red_mask(nx, ny, nz))
            if l == nlevel:
                break
            co = [len(v) % 2 == 0 for v in w]
            self.R.append(kron3(restrict1d(w[2], co[2]), restrict1d(w[1], co[1]),
                                restrict1d(w[0], co[0])))
            self.P.append(kron3(prolong1d(w[2], co[2]), prolong1d(w[1], co[1]),
                                prolong1d(w[0], co[0])))
            w = [v.reshape(-1, 2).sum(axis=1) if c else v for v, c in zip(w, co)]

    def coarse_solve(self, b, tol, omega):
        A, red = self.A[-1], self.red[-1]
        if A.shape[0] == 1:
            return b / A.diagonal()
        x = np.zeros_like(b)
        r0 = np.dot(b, b)
        for _ in range(1000):
            x = sor_rb(A, b, x, red, omega, 1)
            r = b - A @ x
            if np.sqrt(np.dot(r, r) / r0) <= tol:
                break
        return x

    def vcycle(self, x, b, nu, tol, omega, ncycle):
        L = len(self.A) - 1
        r = b - self.A[0] @ x
        res0 = np.dot(r, r)
        hist = []
        for _ in range(ncycle):
            x = sor_rb(self.A[0], b, x, self.red[0], omega, nu)
            bs = [b, self.R[0] @ (b - self.A[0] @ x)]
            xs = [x]
            for l in range(1, L):
                xl = sor_rb(self.A[l], bs[l], np.zeros_like(bs[l]), self.red[l], omega, nu)
                xs.append(xl)
                bs.append(self.R[l] @ (bs[l] - self.A[l] @ xl))
            xs.append(self.coarse_solve(bs[L], tol, omega))
            for l in range(L - 1, 0, -1):
                xs[l] = xs[l] + self.P[l] @ xs[l + 1]
                xs[l] = sor_rb(self.A[l], bs[l], xs[l], self.red[l], omega, nu)
            x = xs[0] + self.P[0] @ xs[1]
            x = sor_rb(self.A[0], b, x, self.red[0], omega, nu)
            r = b - self.A[0] @ x
            rho = np.sqrt(np.dot(r, r) / res0)
            hist.append(rho)
            if rho < tol:
                break
        return x, np.array(hist)
