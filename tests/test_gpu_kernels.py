"""Kernel-level checks of the device layer against numpy FP64 on the same
data (each kernel replaces a reference BLAS/communication group)."""
import numpy as np
import pytest

from nutrlan_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("complex_", [False, True])
@pytest.mark.parametrize("n,m", [(1000, 1), (1000, 7), (4097, 16), (20000, 33), (70001, 64)])
def test_proj_update_cgs(gpu_lib, n, m, complex_):
    T = gpu_lib
    rng = np.random.default_rng(n + m)
    dt = np.complex128 if complex_ else np.float64
    V = rng.standard_normal((n, m + 1))
    if complex_:
        V = V + 1j * rng.standard_normal((n, m + 1))
    print(f"proj/update n={n} m={m} complex={complex_}", flush=True)
    c = T.Ctx(n, m + 2, complex_)
    for j in range(m + 1):
        c.set(j, V[:, j])
    r = V[:, m].astype(dt)
    h = c.proj(0, m, m)
    href = V[:, :m].conj().T @ r
    assert np.abs(h - href).max() <= 1e-12 * np.abs(href).max() * np.sqrt(n)
    hh, rr2, h2 = c.cgs_pass(0, m, m)
    rnew = r - V[:, :m] @ href
    got = c.get(m)
    assert np.abs(got - rnew).max() <= 1e-11 * np.abs(r).max() * m
    assert abs(rr2 - np.vdot(rnew, rnew).real) <= 1e-11 * abs(np.vdot(rnew, rnew))
    assert abs(hh - np.vdot(href, href).real) <= 1e-11 * abs(np.vdot(href, href))
    # explicit update with a host h
    c.set(m, r)
    hx = (rng.standard_normal(m) + (1j * rng.standard_normal(m) if complex_ else 0)).astype(dt)
    rr2b = c.update(0, m, hx, m)
    ref = r - V[:, :m] @ hx
    assert np.abs(c.get(m) - ref).max() <= 1e-11 * np.abs(ref).max() * m
    assert abs(rr2b - np.vdot(ref, ref).real) <= 1e-11 * np.vdot(ref, ref).real
    c.close()


@pytest.mark.parametrize("complex_", [False, True])
@pytest.mark.parametrize("n,lck,jl,k", [(100, 0, 5, 3), (1000, 2, 17, 9), (5000, 0, 64, 33),
                                         (33333, 3, 40, 40), (4099, 0, 128, 128)])
def test_ritz_mfma_inplace(gpu_lib, n, lck, jl, k, complex_):
    T = gpu_lib
    rng = np.random.default_rng(jl * 7 + k)
    ncols = lck + jl + 1
    V = rng.standard_normal((n, ncols))
    if complex_:
        V = V + 1j * rng.standard_normal((n, ncols))
    Y = rng.standard_normal((jl, k))
    print(f"ritz n={n} lck={lck} jl={jl} k={k} complex={complex_}", flush=True)
    c = T.Ctx(n, ncols, complex_)
    for j in range(ncols):
        c.set(j, V[:, j])
    c.ritz(lck, jl, k, Y)
    ref = V[:, lck:lck + jl] @ Y
    for j in range(k):
        got = c.get(lck + j)
        assert np.abs(got - ref[:, j]).max() <= 1e-12 * np.abs(ref).max() * np.sqrt(jl), j
    for j in list(range(lck)) + list(range(lck + k, ncols)):
        assert np.array_equal(c.get(j), V[:, j]), j  # untouched columns
    c.close()


@pytest.mark.parametrize("builder", ["lap", "rand", "diag", "anderson3d"])
def test_spmv_alpha(gpu_lib, builder):
    T = gpu_lib
    cplx = builder == "anderson3d"
    if builder == "lap":
        rp, ci, v = W.laplace3d_csr(20, seed=1)
    elif builder == "rand":
        rp, ci, v = W.random_sym_csr(50000, seed=2)
    elif builder == "diag":
        rp, ci, v = W.diag_csr(12345)
    else:
        rp, ci, v = W.anderson3d_csr(10, seed=3)
    n = len(rp) - 1
    x = W.start_vector(n, seed=5, complex_=cplx)
    A = T.Csr(rp, ci, v, complex_=cplx)
    c = T.Ctx(n, 2, cplx)
    c.set(0, x)
    a = c.spmv(A, 0, 1)
    y = W.csr_matvec(rp, ci, v)(x)
    assert np.abs(c.get(1) - y).max() <= 1e-13 * np.abs(y).max()
    assert abs(a - np.vdot(x, y).real) <= 1e-12 * abs(np.vdot(x, y))
    c.close()
    A.close()


def _arrow_csr(n, seed):
    """Symmetric arrowhead: dense first row/column plus a diagonal.  Its one
    long row sends the SpMV to the row-block (CSR-stream) kernel."""
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(n)
    d = rng.standard_normal(n)
    rows = [np.arange(n)] + [np.array([0, i]) for i in range(1, n)]
    vals = [np.concatenate([[d[0]], a[1:]])] + [np.array([a[i], d[i]]) for i in range(1, n)]
    rp = np.zeros(n + 1, dtype=np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    return rp, np.concatenate(rows).astype(np.int32), np.concatenate(vals)


@pytest.mark.parametrize("builder", ["lap", "arrow", "anderson3d"])
def test_fused_step_sequence(gpu_lib, builder):
    """Three trlhip_step_cgs steps with the normalising scale-copy deferred
    into the next step's SpMV (vector-CSR kernels) or applied before it
    (CSR-stream kernel), against the same recurrence in numpy."""
    T = gpu_lib
    cplx = builder == "anderson3d"
    if builder == "lap":
        rp, ci, v = W.laplace3d_csr(16, seed=2)
    elif builder == "arrow":
        rp, ci, v = _arrow_csr(3000, seed=3)
    else:
        rp, ci, v = W.anderson3d_csr(12, seed=4)
    n = len(rp) - 1
    f = W.csr_matvec(rp, ci, v)
    x = W.start_vector(n, seed=6, complex_=cplx)
    x = x / np.linalg.norm(x)
    A = T.Csr(rp, ci, v, complex_=cplx)
    ncols = 6
    c = T.Ctx(n, ncols, cplx)
    c.set(0, x)
    Q = [x]
    beta_prev = 0.0
    scale = 1.0
    for j in range(1, 4):
        qa, qb = j - 1, (j - 2 if j >= 2 else -1)
        first = 1 if j == 1 else 0
        out = c.step_cgs(A, first, qa, qb, beta_prev, None, j, j)
        # numpy restatement of the step
        q = Q[qa]
        y = f(q)
        alpha = np.vdot(q, y).real
        r = y - alpha * q - (beta_prev * Q[qb] if qb >= 0 else 0.0)
        Vm = np.stack(Q[:j], axis=1)
        h = Vm.conj().T @ r
        s = r - Vm @ h
        ref = [alpha, np.vdot(r, r).real, np.vdot(h, h).real, np.vdot(s, s).real]
        scl = max(1.0, abs(alpha))
        assert abs(out[0] - ref[0]) <= 1e-12 * scl, (j, out, ref)
        assert abs(out[1] - ref[1]) <= 1e-11 * max(ref[1], 1.0)
        assert abs(out[3] - ref[3]) <= 1e-9 * max(ref[1], 1.0)
        assert np.abs(c.get(j) - r).max() <= 1e-12 * max(np.abs(r).max(), 1.0)
        beta = np.sqrt(out[3])
        scale = 1.0 / beta
        c.scale_copy(j, ncols, scale)   # deferred into the next step's SpMV
        Q.append(s * scale)
        beta_prev = beta
    # the last deferred copy is applied by any other entry point (col_get)
    assert np.abs(c.get(3) - Q[3]).max() <= 1e-10
    c.close()
    A.close()


@pytest.mark.parametrize("n,k,ncols,avoid", [(3000, 5, 12, 5), (20000, 8, 10, 9), (4096, 7, 8, 7)])
def test_refine_rayleigh_quotients(gpu_lib, n, k, ncols, avoid):
    """trlhip_refine: rho_i = v'Av / v'v and ||Av - rho v|| / ||v|| for
    columns 0..k-1, against numpy; the Ritz columns and the kept column
    `avoid` are left as they were.  (20000, 8, 10, 9) has room for two
    pairs at a time, so it runs in chunks; (4096, 7, 8, 7) for one."""
    T = gpu_lib
    rng = np.random.default_rng(n + k)
    if n == 4096:
        rp, ci, v = W.anderson2d_csr(64, 5.0, seed=3)[0]
    else:
        rp, ci, v = W.random_sym_csr(n, seed=k)
    import scipy.sparse as sp
    A = sp.csr_matrix((v, ci, rp), shape=(n, n))
    X = rng.standard_normal((n, k))
    X[:, 1] = 0.0  # a zero column gives NaN
    keep = rng.standard_normal(n)
    print(f"refine n={n} k={k} ncols={ncols} avoid={avoid}", flush=True)
    c = T.Ctx(n, ncols)
    for j in range(k):
        c.set(j, X[:, j])
    c.set(avoid, keep)
    Ad = T.Csr(rp, ci, v)
    rho, res = c.refine(Ad, k, avoid)
    for j in range(k):
        x = X[:, j]
        if j == 1:
            assert np.isnan(rho[j]) and np.isnan(res[j])
            continue
        y = A @ x
        r0 = x @ y / (x @ x)
        r1 = np.linalg.norm(y - r0 * x) / np.linalg.norm(x)
        assert abs(rho[j] - r0) <= 1e-13 * abs(y).max() * np.sqrt(n), (j, rho[j], r0)
        assert abs(res[j] - r1) <= 1e-12 * r1, (j, res[j], r1)
        assert np.array_equal(c.get(j), x)
    assert np.array_equal(c.get(avoid), keep)
    Ad.close()
    c.close()


@pytest.mark.parametrize("complex_", [False, True])
def test_cols_get_matches_col_get(gpu_lib, complex_):
    """trlhip_cols_get (the returned Ritz vectors in one transfer) returns
    exactly what trlhip_col_get returns column by column."""
    T = gpu_lib
    n, ncols = 1000, 9
    rng = np.random.default_rng(5)
    c = T.Ctx(n, ncols, complex_=complex_)
    cols = []
    for j in range(ncols):
        x = rng.standard_normal(n) + (1j * rng.standard_normal(n) if complex_ else 0.0)
        c.set(j, x)
        cols.append(x)
    got = c.cols_get(2, 5)
    for i in range(5):
        assert np.array_equal(got[:, i], c.get(2 + i))
        assert np.array_equal(got[:, i], cols[2 + i])
    c.close()


@pytest.mark.parametrize("n,k", [(4096, 8), (20000, 5)])
def test_polish_residuals_and_gram(gpu_lib, n, k):
    """The polish's two device passes against numpy: trlhip_residuals puts
    A v_i - rho_i v_i in columns k+i with its squared norm; trlhip_gram
    returns Q'Q and Q'AQ over the first m = 2k columns (A Q_j formed in
    column m), leaving those columns as they were."""
    T = gpu_lib
    rng = np.random.default_rng(n + 17 * k)
    if n == 4096:
        rp, ci, v = W.anderson2d_csr(64, 7.0, seed=5)[0]
    else:
        rp, ci, v = W.random_sym_csr(n, seed=k)
    import scipy.sparse as sp
    A = sp.csr_matrix((v, ci, rp), shape=(n, n))
    X = rng.standard_normal((n, k))
    rho = rng.standard_normal(k)
    m = 2 * k
    print(f"polish passes n={n} k={k}", flush=True)
    c = T.Ctx(n, m + 2)
    for j in range(k):
        c.set(j, X[:, j])
    Ad = T.Csr(rp, ci, v)
    rr2 = c.residuals(Ad, rho)
    R = np.empty((n, k))
    scale = abs(A).max() * 6
    for i in range(k):
        ref = A @ X[:, i] - rho[i] * X[:, i]
        R[:, i] = c.get(k + i)
        assert np.allclose(R[:, i], ref, rtol=0, atol=1e-13 * scale * (1 + abs(rho[i])))
        assert abs(rr2[i] - ref @ ref) <= 1e-12 * (ref @ ref), (i, rr2[i], ref @ ref)
    Q = np.hstack([X, R])
    G, H = c.gram(Ad, m, m)
    Gr, Hr = Q.T @ Q, Q.T @ (A @ Q)
    assert np.allclose(G, Gr, rtol=0, atol=1e-12 * abs(Gr).max())
    assert np.allclose(H, Hr, rtol=0, atol=1e-12 * abs(Hr).max())
    for j in range(m):
        assert np.array_equal(c.get(j), Q[:, j])
    Ad.close()
    c.close()


@pytest.mark.parametrize("builder", ["lap", "rand", "anderson3d"])
def test_autotune_spmv(gpu_lib, builder):
    """trlhip_csr_autotune picks one of the SpMV kernels (0 = CSR-stream,
    else vector lanes per row); the chosen kernel's y = A x and x'y match
    numpy, on real and complex matrices."""
    T = gpu_lib
    cplx = builder == "anderson3d"
    if builder == "lap":
        rp, ci, v = W.laplace3d_csr(24, seed=1)
    elif builder == "rand":
        rp, ci, v = W.random_sym_csr(60000, seed=3)
    else:
        rp, ci, v = W.anderson3d_csr(12, seed=4)
    n = len(rp) - 1
    A = T.Csr(rp, ci, v, complex_=cplx)
    lpr, ms = A.autotune(reps=4)
    print(f"autotune {builder}: lanes per row {lpr}, {1e3 * ms:.1f} us per SpMV", flush=True)
    assert lpr in (0, 2, 4, 8, 16, 32) and ms > 0
    x = W.start_vector(n, seed=6, complex_=cplx)
    c = T.Ctx(n, 2, cplx)
    c.set(0, x)
    a = c.spmv(A, 0, 1)
    y = W.csr_matvec(rp, ci, v)(x)
    assert np.abs(c.get(1) - y).max() <= 1e-13 * np.abs(y).max()
    assert abs(a - np.vdot(x, y).real) <= 1e-12 * abs(np.vdot(x, y))
    c.close()
    A.close()


def test_autotuned_solve_matches_oracle(gpu_lib):
    """A solve on an autotuned operator (vector kernel forced by timing or
    not) keeps the eigenvalue bar against the oracle."""
    from oracle import trlan_oracle as O

    T = gpu_lib
    rp, ci, v = W.random_sym_csr(20000, seed=8)
    n = len(rp) - 1
    x0 = W.start_vector(n, seed=9)
    A = T.Csr(rp, ci, v)
    lpr, _ = A.autotune(reps=3)
    g = T.solve_csr(A, 1, 8, 40, 1e-8, 1, 20000, 8, x0)
    A.close()
    info = O.init_info(n, 40, 1, 8, 1e-8, 1, 20000)
    O.set_iguess(info, 0, 1)
    o = O.trlan(W.csr_matvec(rp, ci, v), info, n, 8, x0)
    anrm = max(abs(info.anrm), 1e-300)
    assert g["nec"] == 8 and info.nec == 8
    err = np.abs(np.sort(g["eval"][:8]) - np.sort(o["eval"][:8])).max() / anrm
    print(f"autotuned (lanes {lpr}) vs oracle: max |dlambda|/||A|| {err:.1e}", flush=True)
    assert err <= 1e-10


def _csr_pair(T, rp, ci, v, monkeypatch, mixed=False):
    """The same matrix as a packed operator (default) and an unpacked one
    (TRLHIP_CSR_PACK=0 at creation)."""
    A1 = T.Csr(rp, ci, v, mixed=mixed)
    monkeypatch.setenv("TRLHIP_CSR_PACK", "0")
    A0 = T.Csr(rp, ci, v, mixed=mixed)
    monkeypatch.delenv("TRLHIP_CSR_PACK")
    return A1, A0


def _band_csr(n, half, seed):
    """Symmetric band matrix: random diagonal, off-diagonals -1/d at
    distance d <= half (rows of 2*half + 1 nonzeros)."""
    rng = np.random.default_rng(seed)
    rows, cols, vals = [], [], []
    for d in range(-half, half + 1):
        i = np.arange(max(0, -d), min(n, n - d))
        rows.append(i)
        cols.append(i + d)
        vals.append(rng.standard_normal(len(i)) if d == 0 else np.full(len(i), -1.0 / abs(d)))
    r, c, v = np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)
    o = np.lexsort((c, r))
    r, c, v = r[o], c[o], v[o]
    rp = np.zeros(n + 1, dtype=np.int64)
    np.add.at(rp, r + 1, 1)
    return np.cumsum(rp).astype(np.int32), c.astype(np.int32), v


@pytest.mark.parametrize("builder", ["lap", "and2d", "diag", "band"])
def test_packed_spmv_bitwise(gpu_lib, builder, monkeypatch):
    """Matrices with few distinct off-diagonal values get packed nonzeros
    (short and long rows, a diagonal-only matrix); the packed CSR-stream
    SpMV reproduces the unpacked one bit for bit (y and the alpha partial
    sums), and both match numpy."""
    T = gpu_lib
    if builder == "lap":
        rp, ci, v = W.laplace3d_csr(20, seed=2)
    elif builder == "and2d":
        (rp, ci, v), _ = W.anderson2d_csr(150, 7.5, seed=3)
    elif builder == "band":
        rp, ci, v = _band_csr(20000, 12, seed=4)
    else:
        rp, ci, v = W.diag_csr(5000)
    n = len(rp) - 1
    A1, A0 = _csr_pair(T, rp, ci, v, monkeypatch)
    print(f"packed spmv {builder}: n={n} packed={A1.packed}", flush=True)
    assert A1.packed and not A0.packed
    x = W.start_vector(n, seed=11)
    c = T.Ctx(n, 3, False)
    c.set(0, x)
    a1 = c.spmv(A1, 0, 1)
    a0 = c.spmv(A0, 0, 2)
    y1, y0 = c.get(1), c.get(2)
    assert np.array_equal(y1, y0) and a1 == a0
    y = W.csr_matvec(rp, ci, v)(x)
    assert np.abs(y1 - y).max() <= 1e-13 * np.abs(y).max()
    c.close()
    A1.close()
    A0.close()


def test_packing_declined(gpu_lib):
    """More than 255 distinct off-diagonal values: no packed copy."""
    T = gpu_lib
    rp, ci, v = W.random_sym_csr(3000, seed=5)
    A = T.Csr(rp, ci, v)
    assert not A.packed
    A.close()


@pytest.mark.parametrize("mixed", [False, True])
def test_packed_solves_bitwise(gpu_lib, mixed, monkeypatch):
    """Whole solves on packed and unpacked operators agree bit for bit:
    single (trlan) and batched (trlan_batch, kb_spmv_pk against
    kb_spmv_stream), plain FP64 and mixed precision."""
    T = gpu_lib
    rp, ci, v = W.laplace3d_csr(16, seed=4)
    n = len(rp) - 1
    x0 = W.start_vector(n, seed=12)
    A1, A0 = _csr_pair(T, rp, ci, v, monkeypatch, mixed=mixed)
    print(f"packed solves mixed={mixed} n={n}", flush=True)
    g1 = T.solve_csr(A1, -1, 6, 30, 1e-9, 1, 20000, 6, x0)
    g0 = T.solve_csr(A0, -1, 6, 30, 1e-9, 1, 20000, 6, x0)
    assert g1["nec"] == g0["nec"] == 6
    assert g1["matvec"] == g0["matvec"] and g1["nloop"] == g0["nloop"]
    assert np.array_equal(g1["eval"][:6], g0["eval"][:6])
    lat = [W.anderson2d_csr(L, 6.0, seed=20 + L)[0] for L in (48, 64, 70)]
    ops1, ops0 = [], []
    for (r, c_, vv) in lat:
        a1, a0 = _csr_pair(T, r, c_, vv, monkeypatch, mixed=mixed)
        assert a1.packed
        ops1.append(a1)
        ops0.append(a0)
    probs = lambda ops: [(A, -1, 4, 24, 1e-8, 1, 20000, 4,
                          W.start_vector(len(lat[i][0]) - 1, seed=30 + i)) for i, A in enumerate(ops)]
    rc1, r1, _ = T.solve_batch(probs(ops1))
    rc0, r0, _ = T.solve_batch(probs(ops0))
    assert rc1 == 0 and rc0 == 0
    for a, b in zip(r1, r0):
        assert a["matvec"] == b["matvec"] and a["nec"] == b["nec"] == 4
        assert np.array_equal(a["eval"][:4], b["eval"][:4])
    for A in ops1 + ops0 + [A1, A0]:
        A.close()


def test_packing_declined_far_column_and_double_diagonal(gpu_lib):
    """Packing needs every column within 2^23 rows of its row block and at
    most one diagonal entry per row; a matrix breaking either keeps the
    unpacked SpMV, whose result still matches numpy."""
    T = gpu_lib
    # n > 2^23: diagonal plus one symmetric pair (0, n-1), 2^23 + 511 rows apart
    n = (1 << 23) + 512
    rows = np.concatenate([np.arange(n), [0, n - 1]])
    cols = np.concatenate([np.arange(n), [n - 1, 0]])
    vals = np.concatenate([np.full(n, 2.0), [-1.0, -1.0]])
    o = np.lexsort((cols, rows))
    rows, cols, vals = rows[o], cols[o], vals[o]
    rp = np.zeros(n + 1, dtype=np.int64)
    np.add.at(rp, rows + 1, 1)
    rp = np.cumsum(rp).astype(np.int32)
    A = T.Csr(rp, cols.astype(np.int32), vals)
    print(f"far column: n={n} packed={A.packed}", flush=True)
    assert not A.packed
    x = W.start_vector(n, seed=13)
    c = T.Ctx(n, 2, False)
    c.set(0, x)
    c.spmv(A, 0, 1)
    y = 2.0 * x
    y[0] -= x[n - 1]
    y[n - 1] -= x[0]
    assert np.abs(c.get(1) - y).max() <= 1e-14 * np.abs(y).max()
    c.close()
    A.close()
    # a row with two diagonal entries (duplicates are summed)
    rp2 = np.array([0, 2, 3, 4], dtype=np.int32)
    ci2 = np.array([0, 0, 1, 2], dtype=np.int32)
    v2 = np.array([1.5, 0.25, 3.0, 4.0])
    B = T.Csr(rp2, ci2, v2)
    assert not B.packed
    c = T.Ctx(3, 2, False)
    x3 = np.array([1.0, -2.0, 0.5])
    c.set(0, x3)
    c.spmv(B, 0, 1)
    assert np.allclose(c.get(1), [1.75 * 1.0, 3.0 * -2.0, 4.0 * 0.5], rtol=0, atol=1e-15)
    c.close()
    B.close()


@pytest.mark.parametrize("cb", [4, 16])
def test_projection_tiles_keep_batch_equal_single(gpu_lib, cb):
    """With other projection tiles (process-wide), batched solves still equal
    their single solves bit for bit, plain and mixed, and the eigenvalues
    stay within 1e-10 ||A|| of the default tiles' solve."""
    T = gpu_lib
    lat = [W.anderson2d_csr(L, 7.0, seed=60 + L)[0] for L in (40, 56)]
    ops = [T.Csr(*c, mixed=(i == 1)) for i, c in enumerate(lat)]
    probs = [(A, -1, 6, 28, 1e-8, 1, 20000, 6, W.start_vector(A.n, seed=70 + i))
             for i, A in enumerate(ops)]
    ref = [T.solve_csr(*p) for p in probs]
    try:
        T.proj_tiles(cb, cb)
        rc, res, _ = T.solve_batch(probs)
        assert rc == 0
        for p, g, r0 in zip(probs, res, ref):
            s = T.solve_csr(*p)
            assert g["matvec"] == s["matvec"] and np.array_equal(g["eval"], s["eval"])
            assert np.array_equal(g["evec"], s["evec"])
            assert g["nec"] == r0["nec"] == 6
            assert np.abs(g["eval"][:6] - r0["eval"][:6]).max() <= 1e-10 * r0["anrm"]
    finally:
        T.proj_tiles()
    for A in ops:
        A.close()


def test_projection_autotune_and_cache(gpu_lib, tmp_path):
    """The projection-tile autotuner picks one of 4, 8, 16 per precision and
    sets it; the JSON cache replays that choice and the SpMV choice."""
    T = gpu_lib
    try:
        (c64, c32), (m64, m32) = T.proj_autotune(1 << 18, 24, reps=3)
        print(f"proj tiles: fp64 {c64} ({1e3 * m64:.1f} us), fp32 {c32} ({1e3 * m32:.1f} us)", flush=True)
        assert c64 in (4, 8, 16) and c32 in (4, 8, 16) and m64 > 0 and m32 > 0
        assert T.proj_tiles_get() == (c64, c32)
        rp, ci, v = W.laplace3d_csr(16, seed=3)
        A = T.Csr(rp, ci, v)
        path = str(tmp_path / "autotune.json")
        cache = T.autotune_cached(path, 1 << 16, 20, [A], reps=2)
        key = f"{A.n}:{A.nnz}"
        assert key in cache["spmv"] and "65536x20" in cache["proj"]
        T.proj_tiles()
        again = T.autotune_cached(path, 1 << 16, 20, [A], reps=2)
        assert again == cache and T.proj_tiles_get() == tuple(cache["proj"]["65536x20"])
        A.close()
    finally:
        T.proj_tiles()


# ------------------------------------------------------------ stencils

def _csr_three(T, rp, ci, v, monkeypatch, mixed=False):
    """The same matrix with the stencil SpMV (default), with packed
    nonzeros only (TRLHIP_CSR_STENCIL=0) and unpacked (TRLHIP_CSR_PACK=0)."""
    S = T.Csr(rp, ci, v, mixed=mixed)
    monkeypatch.setenv("TRLHIP_CSR_STENCIL", "0")
    P = T.Csr(rp, ci, v, mixed=mixed)
    monkeypatch.delenv("TRLHIP_CSR_STENCIL")
    monkeypatch.setenv("TRLHIP_CSR_PACK", "0")
    U = T.Csr(rp, ci, v, mixed=mixed)
    monkeypatch.delenv("TRLHIP_CSR_PACK")
    return S, P, U


@pytest.mark.parametrize("builder", ["lap3d", "and2d", "lap3d_odd"])
def test_stencil_spmv_bitwise(gpu_lib, builder, monkeypatch):
    """A periodic 2D/3D stencil's SpMV, which forms its columns from the
    lattice, equals the packed and the unpacked CSR-stream SpMV bit for bit
    (y and alpha), and numpy to rounding."""
    T = gpu_lib
    if builder == "lap3d":
        rp, ci, v = W.laplace3d_csr(20, seed=2)
        want = (3, 20)
    elif builder == "lap3d_odd":
        rp, ci, v = W.laplace3d_csr(3, seed=5)
        want = (3, 3)
    else:
        (rp, ci, v), _ = W.anderson2d_csr(150, 7.5, seed=3)
        want = (2, 150)
    n = len(rp) - 1
    S, P, U = _csr_three(T, rp, ci, v, monkeypatch)
    print(f"stencil spmv {builder}: n={n} stencil={S.stencil}", flush=True)
    assert S.stencil == want and P.stencil == (0, 0) and U.stencil == (0, 0)
    assert P.packed and not U.packed
    x = W.start_vector(n, seed=11)
    c = T.Ctx(n, 4, False)
    c.set(0, x)
    a = [c.spmv(A, 0, j) for j, A in zip((1, 2, 3), (S, P, U))]
    ys = [c.get(j) for j in (1, 2, 3)]
    assert np.array_equal(ys[0], ys[1]) and np.array_equal(ys[0], ys[2])
    assert a[0] == a[1] == a[2]
    y = W.csr_matvec(rp, ci, v)(x)
    assert np.abs(ys[0] - y).max() <= 1e-13 * np.abs(y).max()
    c.close()
    for A in (S, P, U):
        A.close()


@pytest.mark.parametrize("mixed", [False, True])
def test_stencil_solves_bitwise(gpu_lib, mixed, monkeypatch):
    """Whole solves with the stencil SpMV equal those with the packed
    CSR-stream SpMV bit for bit: single (trlan) and batched (trlan_batch,
    kb_spmv_st against kb_spmv_pk), plain FP64 and mixed precision."""
    T = gpu_lib
    rp, ci, v = W.laplace3d_csr(16, seed=4)
    n = len(rp) - 1
    x0 = W.start_vector(n, seed=12)
    S, P, U = _csr_three(T, rp, ci, v, monkeypatch, mixed=mixed)
    print(f"stencil solves mixed={mixed} n={n}", flush=True)
    gs = T.solve_csr(S, -1, 6, 30, 1e-9, 1, 20000, 6, x0)
    gp = T.solve_csr(P, -1, 6, 30, 1e-9, 1, 20000, 6, x0)
    assert gs["nec"] == gp["nec"] == 6
    assert (gs["matvec"], gs["nloop"]) == (gp["matvec"], gp["nloop"])
    assert np.array_equal(gs["eval"][:6], gp["eval"][:6])
    assert np.array_equal(gs["evec"][:, :6], gp["evec"][:, :6])
    lat = [W.anderson2d_csr(L, 6.0, seed=20 + L)[0] for L in (48, 64, 70)]
    opsS, opsP = [], []
    for (r, c_, vv) in lat:
        a, b, u = _csr_three(T, r, c_, vv, monkeypatch, mixed=mixed)
        u.close()
        assert a.stencil[0] == 2 and b.stencil == (0, 0)
        opsS.append(a)
        opsP.append(b)
    probs = lambda ops: [(A, -1, 4, 24, 1e-8, 1, 20000, 4,
                          W.start_vector(len(lat[i][0]) - 1, seed=30 + i)) for i, A in enumerate(ops)]
    rcS, rS, _ = T.solve_batch(probs(opsS))
    rcP, rP, _ = T.solve_batch(probs(opsP))
    assert rcS == 0 and rcP == 0
    for a, b in zip(rS, rP):
        assert a["matvec"] == b["matvec"] and a["nec"] == b["nec"] == 4
        assert np.array_equal(a["eval"][:4], b["eval"][:4])
    for A in opsS + opsP + [S, P, U]:
        A.close()


def test_stencil_detection_declines(gpu_lib):
    """Only periodic lattices with every off-diagonal value equal (bit for
    bit) and ascending columns use the stencil SpMV; others keep the CSR
    kernels."""
    T = gpu_lib
    rp, ci, v = W.random_sym_csr(3000, seed=5)
    A = T.Csr(rp, ci, v)
    assert A.stencil == (0, 0)
    A.close()
    (rp, ci, v), _ = W.anderson2d_csr(32, 4.0, seed=1)
    off = np.flatnonzero(ci != np.repeat(np.arange(32 * 32), 5))
    v2 = v.copy()
    v2[off[7]] = np.nextafter(v2[off[7]], 0.0)  # one off-diagonal value differs
    A = T.Csr(rp, ci, v2)
    assert A.stencil == (0, 0) and A.packed
    A.close()
    ci2 = ci.copy()
    ci2[[0, 1]] = ci2[[1, 0]]  # row 0's columns out of order
    v3 = v.copy()
    v3[[0, 1]] = v3[[1, 0]]
    A = T.Csr(rp, ci2, v3)
    assert A.stencil == (0, 0)
    A.close()
    A = T.Csr(rp, ci, v)
    assert A.stencil == (2, 32)
    A.close()


# ------------------------------------------------------- column panels

def _paneled(T, rp, ci, v, monkeypatch, cols=4096, mixed=False):
    """The operator with its nonzeros split into column panels of `cols`
    columns (forced at this size by TRLHIP_CSR_PANEL=1)."""
    monkeypatch.setenv("TRLHIP_CSR_PANEL", "1")
    monkeypatch.setenv("TRLHIP_CSR_PANEL_COLS", str(cols))
    A = T.Csr(rp, ci, v, mixed=mixed)
    monkeypatch.delenv("TRLHIP_CSR_PANEL")
    monkeypatch.delenv("TRLHIP_CSR_PANEL_COLS")
    return A


@pytest.mark.parametrize("n,cols", [(50000, 4096), (30001, 64), (5000, 1 << 18)])
def test_panel_spmv(gpu_lib, n, cols, monkeypatch):
    """The column-panel SpMV (many panels, panels of one block of columns,
    a single panel) against numpy, its alpha against numpy, the same bits
    run to run, and against the CSR-stream kernel to rounding."""
    T = gpu_lib
    rp, ci, v = W.random_sym_csr(n, seed=7)
    A = _paneled(T, rp, ci, v, monkeypatch, cols)
    B = T.Csr(rp, ci, v)
    print(f"panel spmv n={n} cols={cols} paneled={A.paneled}", flush=True)
    assert A.paneled and not B.paneled
    x = W.start_vector(n, seed=8) - 0.5
    c = T.Ctx(n, 4)
    c.set(0, x)
    a1 = c.spmv(A, 0, 1)
    a2 = c.spmv(A, 0, 2)
    a3 = c.spmv(B, 0, 3)
    y1, y2, y3 = c.get(1), c.get(2), c.get(3)
    assert np.array_equal(y1, y2) and a1 == a2
    y = W.csr_matvec(rp, ci, v)(x)
    scale = np.abs(y).max()
    assert np.abs(y1 - y).max() <= 1e-13 * scale
    assert np.abs(y1 - y3).max() <= 1e-13 * scale
    assert abs(a1 - x @ y) <= 1e-12 * abs(x) @ abs(y)
    c.close()
    A.close()
    B.close()


def test_panel_fused_steps_and_solve(gpu_lib, monkeypatch):
    """Lanczos steps whose normalisation is fused into the column-panel
    SpMV, against numpy; and a whole solve on a paneled operator against
    the FP64 oracle (eigenvalues within 1e-10 ||A||, true residuals)."""
    T = gpu_lib
    from oracle import trlan_oracle as O

    n = 20000
    rp, ci, v = W.random_sym_csr(n, seed=9)
    f = W.csr_matvec(rp, ci, v)
    A = _paneled(T, rp, ci, v, monkeypatch, 2048)
    assert A.paneled
    x = W.start_vector(n, seed=6)
    x = x / np.linalg.norm(x)
    ncols = 6
    c = T.Ctx(n, ncols)
    c.set(0, x)
    Q = [x]
    beta_prev = 0.0
    for j in range(1, 4):
        qa, qb = j - 1, (j - 2 if j >= 2 else -1)
        out = c.step_cgs(A, 1 if j == 1 else 0, qa, qb, beta_prev, None, j, j)
        q = Q[qa]
        y = f(q)
        alpha = q @ y
        r = y - alpha * q - (beta_prev * Q[qb] if qb >= 0 else 0.0)
        Vm = np.stack(Q[:j], axis=1)
        s = r - Vm @ (Vm.T @ r)
        assert abs(out[0] - alpha) <= 1e-12 * max(1.0, abs(alpha))
        assert np.abs(c.get(j) - r).max() <= 1e-12 * max(np.abs(r).max(), 1.0)
        beta = np.sqrt(out[3])
        c.scale_copy(j, ncols, 1.0 / beta)
        Q.append(s / beta)
        beta_prev = beta
    assert np.abs(c.get(3) - Q[3]).max() <= 1e-10
    c.close()
    x0 = W.start_vector(n, seed=10)
    g = T.solve_csr(A, 1, 12, 48, 1e-8, 1, 20000, 12, x0)
    A.close()
    info = O.init_info(n, 48, 1, 12, 1e-8, 1, 20000)
    O.set_iguess(info, 0, 1)
    o = O.trlan(f, info, n, 12, x0)
    assert g["stat"] == 0 and g["nec"] == 12 == info.nec
    assert np.abs(g["eval"][:12] - o["eval"][:12]).max() <= 1e-10 * info.anrm
    for i in range(12):
        xx = g["evec"][:, i]
        assert np.linalg.norm(f(xx) - g["eval"][i] * xx) <= 1.01e-8 * g["anrm"]
    print(f"paneled solve: matvec {g['matvec']} oracle {info.matvec}", flush=True)
