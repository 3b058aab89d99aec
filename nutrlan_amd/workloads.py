"""Seeded synthetic operators and start vectors for the eigensolver workloads.

All generators use numpy's PCG64 (`np.random.default_rng(seed)`) and return
CSR arrays (row_ptr int32, col_idx int32 sorted within each row, values) so
the same operator can be handed to the GPU solver, the test oracle and the
compiled reference.

Workloads (sizes as named):
  diag_csr(n)                diag(1, 2, ..., n)
  laplace3d_csr(L, seed)     periodic 7-point Laplacian on L^3: off-diagonal
                             -1, diagonal 6 + U[0,1)
  anderson3d_csr(L, ...)     complex Hermitian periodic tight binding on L^3:
                             hopping -1 with Peierls phase exp(2 pi i phi x)
                             on y-bonds, on-site U[-W/2, W/2]
  random_sym_csr(n, ...)     symmetric random pattern, row degree ~Poisson(d),
                             values N(0,1)/4 mirrored, diagonal U[-1,1]
  anderson2d_csr(L, W, ...)  periodic L x L tight binding, hopping -1,
                             on-site U[-W/2, W/2]
"""
import numpy as np


def _coo_to_csr(n, rows, cols, vals, dtype):
    """CSR of a COO matrix: rows and columns sorted, duplicates summed in
    their input order (one stable sort of row * n + col)."""
    key = rows.astype(np.int64) * n + cols
    order = np.argsort(key, kind="stable")
    key = key[order]
    vals = vals[order]
    if len(key):
        start = np.flatnonzero(np.concatenate([[True], key[1:] != key[:-1]]))
        vals = np.add.reduceat(vals, start)
        key = key[start]
    rows = key // n
    cols = key - rows * n
    rp = np.zeros(n + 1, dtype=np.int64)
    rp[1:] = np.cumsum(np.bincount(rows, minlength=n))
    return rp.astype(np.int32), cols.astype(np.int32), vals.astype(dtype)


def diag_csr(n, dtype=np.float64):
    rp = np.arange(n + 1, dtype=np.int32)
    ci = np.arange(n, dtype=np.int32)
    v = np.arange(1, n + 1, dtype=dtype)
    return rp, ci, v


def _stencil_csr(n, nbr_cols, nbr_vals, diag_vals, dtype):
    """Assemble CSR for an operator with a fixed number of off-diagonal
    neighbours per row plus a diagonal; columns sorted per row."""
    k = nbr_cols.shape[1]
    cols = np.concatenate([nbr_cols, np.arange(n, dtype=np.int64)[:, None]], axis=1)
    vals = np.concatenate([nbr_vals, diag_vals[:, None]], axis=1).astype(dtype)
    order = np.argsort(cols, axis=1, kind="stable")
    cols = np.take_along_axis(cols, order, axis=1)
    vals = np.take_along_axis(vals, order, axis=1)
    rp = (np.arange(n + 1, dtype=np.int64) * (k + 1)).astype(np.int32)
    return rp, cols.reshape(-1).astype(np.int32), vals.reshape(-1)


def laplace3d_csr(L, seed=0, perturb=True, dtype=np.float64):
    """Periodic 3D 7-point Laplacian, off-diagonal -1, diagonal 6 + U[0,1).

    Needs L >= 3 so the six neighbours of a site are distinct.
    """
    n = L ** 3
    idx = np.arange(n, dtype=np.int64)
    x = idx % L
    y = (idx // L) % L
    z = idx // (L * L)

    def site(xx, yy, zz):
        return (xx % L) + L * ((yy % L) + L * (zz % L))

    nb = np.stack([site(x - 1, y, z), site(x + 1, y, z), site(x, y - 1, z),
                   site(x, y + 1, z), site(x, y, z - 1), site(x, y, z + 1)], axis=1)
    nv = -np.ones_like(nb, dtype=np.float64)
    rng = np.random.default_rng(seed)
    d = 6.0 + (rng.random(n) if perturb else np.zeros(n))
    return _stencil_csr(n, nb, nv, d, dtype)


def anderson3d_csr(L, phi=1.0 / 16.0, W=4.0, seed=0):
    """Complex Hermitian 3D Anderson model with a Peierls phase on y-bonds.

    H[(x,y+1,z),(x,y,z)] = -exp(2 pi i phi x), its conjugate on the reverse
    bond, -1 on x- and z-bonds, on-site U[-W/2, W/2] (W=4: U[-2,2]).
    """
    n = L ** 3
    idx = np.arange(n, dtype=np.int64)
    x = idx % L
    y = (idx // L) % L
    z = idx // (L * L)

    def site(xx, yy, zz):
        return (xx % L) + L * ((yy % L) + L * (zz % L))

    ph = np.exp(2j * np.pi * phi * x)
    nb = np.stack([site(x - 1, y, z), site(x + 1, y, z), site(x, y - 1, z),
                   site(x, y + 1, z), site(x, y, z - 1), site(x, y, z + 1)], axis=1)
    # row r=(x,y,z): entry to (x,y-1,z) is H[r, r-y] = -exp(2 pi i phi x)
    # (the bond (x,y-1)->(x,y)); entry to (x,y+1,z) is its conjugate.
    nv = np.stack([-np.ones(n), -np.ones(n), -ph, -np.conj(ph), -np.ones(n), -np.ones(n)],
                  axis=1).astype(np.complex128)
    rng = np.random.default_rng(seed)
    d = (rng.random(n) - 0.5) * W
    return _stencil_csr(n, nb, nv, d.astype(np.complex128), np.complex128)


def anderson2d_csr(L, W, seed=0, dtype=np.float64):
    """Periodic L x L Anderson model: hopping -1, on-site U[-W/2, W/2]."""
    n = L * L
    idx = np.arange(n, dtype=np.int64)
    x = idx % L
    y = idx // L

    def site(xx, yy):
        return (xx % L) + L * (yy % L)

    nb = np.stack([site(x - 1, y), site(x + 1, y), site(x, y - 1), site(x, y + 1)], axis=1)
    nv = -np.ones_like(nb, dtype=np.float64)
    rng = np.random.default_rng(seed)
    d = (rng.random(n) - 0.5) * W
    return _stencil_csr(n, nb, nv, d, dtype), d


def random_sym_csr(n, degree=16.0, seed=0, dtype=np.float64):
    """Symmetric random sparse matrix: off-diagonal pattern symmetric with row
    degree ~Poisson(degree), values N(0,1)/4 mirrored, diagonal U[-1,1]."""
    rng = np.random.default_rng(seed)
    # each row draws Poisson(degree/2) partners; mirroring doubles the degree
    cnt = rng.poisson(degree / 2.0, size=n)
    m = int(cnt.sum())
    r = np.repeat(np.arange(n, dtype=np.int64), cnt)
    c = rng.integers(0, n - 1, size=m, dtype=np.int64)
    c = c + (c >= r)  # never the diagonal
    v = rng.standard_normal(m) / 4.0
    rows = np.concatenate([r, c, np.arange(n, dtype=np.int64)])
    cols = np.concatenate([c, r, np.arange(n, dtype=np.int64)])
    vals = np.concatenate([v, v, rng.uniform(-1.0, 1.0, size=n)])
    return _coo_to_csr(n, rows, cols, vals, dtype)


def start_vector(n, seed=0, complex_=False):
    """Start vector x_i ~ U[0,1) (complex: independent U[0,1) parts)."""
    rng = np.random.default_rng(seed)
    if complex_:
        return rng.random(n) + 1j * rng.random(n)
    return rng.random(n)


def csr_matvec(rp, ci, v):
    """Returns a numpy y = A x closure over a CSR matrix (for tests)."""
    import scipy.sparse as sp

    n = len(rp) - 1
    A = sp.csr_matrix((v, ci, rp), shape=(n, n))
    return lambda x: A @ x


def csr_dense(rp, ci, v):
    import scipy.sparse as sp

    n = len(rp) - 1
    return sp.csr_matrix((v, ci, rp), shape=(n, n)).toarray()
