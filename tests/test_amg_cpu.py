"""The AMG setup (the role of HYPRE BoomerAMG, poisson_hypre.c:34-97: PMIS
coarsening, FF1 interpolation, Galerkin coarse operators) run on the host
through the C ABI, pinned by what the algorithms fix:

- on the 1-D Laplacian, an F-point between two C-points interpolates with
  weights 1/2, 1/2 (the extended+i formula is exact for linear functions);
- interpolation rows of zero-row-sum operators sum to one (constants are
  interpolated exactly) and C-point rows are unit rows;
- every point is C or F, and where the strength relation is symmetric no two
  C-points are strongly connected (PMIS picks independent sets);
- each coarse operator equals P^T A P (computed here with scipy.sparse) and
  keeps zero row sums;
- the hierarchy of a curvilinear channel operator coarsens to the dense level
  with bounded operator complexity.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from helpers import cases, vfs
import pyoracle as po


def csr_of(ptr, col, val, shape):
    return sp.csr_matrix((val, col, ptr), shape=shape)


def laplace_1d(n):
    A = sp.diags([-np.ones(n - 1), 2 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1], format="csr")
    return A


def laplace_3d_neumann(nx, ny, nz):
    def d1(n):
        m = sp.diags([-np.ones(n - 1), 2 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1], format="lil")
        m[0, 0] = m[n - 1, n - 1] = 1
        return m.tocsr()

    I = lambda n: sp.identity(n, format="csr")
    return (sp.kron(sp.kron(I(nz), I(ny)), d1(nx)) + sp.kron(sp.kron(I(nz), d1(ny)), I(nx)) +
            sp.kron(sp.kron(d1(nz), I(ny)), I(nx))).tocsr()


def host(A, theta=0.75, max_coarse=0):
    A = sp.csr_matrix(A)
    A.sort_indices()
    return vfs.AmgHost(A.indptr, A.indices, A.data, theta=theta, max_coarse=max_coarse)


def strong_pairs(A, theta):
    A = sp.csr_matrix(A)
    S = []
    for i in range(A.shape[0]):
        cols, vals = A.indices[A.indptr[i]:A.indptr[i + 1]], A.data[A.indptr[i]:A.indptr[i + 1]]
        d = vals[cols == i][0]
        s = -np.sign(d) * vals
        off = cols != i
        mx = s[off].max() if off.any() else 0
        S.append(set(cols[off & (s >= theta * mx)]) if mx > 0 else set())
    return S


def test_1d_linear_interpolation():
    h = host(laplace_1d(101), max_coarse=10)
    cf = h.splitting(0)
    ptr, col, val, shape = h.matrix(0, 1)
    P = csr_of(ptr, col, val, shape).toarray()
    checked = 0
    for i in range(1, 100):
        if cf[i] == -1 and cf[i - 1] == 1 and cf[i + 1] == 1:
            row = P[i][P[i] != 0]
            assert np.allclose(row, [0.5, 0.5], rtol=0, atol=1e-15)
            checked += 1
    assert checked > 10


def test_3d_neumann_hierarchy_properties():
    A = laplace_3d_neumann(12, 10, 8)
    theta = 0.25
    h = host(A, theta=theta, max_coarse=40)
    assert h.levels >= 3
    Al = A
    for lev in range(h.levels - 1):
        cf = h.splitting(lev)
        assert set(np.unique(cf)) <= {-1, 1}
        C = np.flatnonzero(cf == 1)
        if lev == 0:
            # the strength relation of the finest level is symmetric (every off-diagonal is -1), so the
            # C-points PMIS selects are pairwise not strongly connected
            S = strong_pairs(Al, theta)
            Cset = set(C)
            for i in C:
                assert not (S[i] & Cset), "two strongly connected C-points"
        P = csr_of(*h.matrix(lev, 1))
        assert np.allclose(P @ np.ones(P.shape[1]), 1.0, rtol=0, atol=1e-12)
        Pd = P.toarray()
        for ci, i in enumerate(C):
            assert np.array_equal(np.flatnonzero(Pd[i]), [ci]) and Pd[i, ci] == 1.0
        Ac = csr_of(*h.matrix(lev + 1, 0))
        ref = (P.T @ Al @ P).toarray()
        assert np.allclose(Ac.toarray(), ref, rtol=0, atol=1e-12 * np.abs(ref).max())
        assert np.allclose(Ac @ np.ones(Ac.shape[0]), 0.0, rtol=0, atol=1e-11)
        Al = Ac
    assert Al.shape[0] <= 40


def channel_matrix(nx=18, ny=12, nz=18):
    """The curvilinear channel's Poisson operator (oracle), probed into CSR
    over its rows with 27 colour vectors (periodic counts are multiples of 3,
    so the colouring is consistent across the wrap)."""
    c = cases.channel(nx, ny, nz)
    p = po.make_params(*c.da, per=c.per, bctype=c.bctype, ren=c.ren)
    o = po.Oracle(p, c.coor)
    mz, my, mx = o.shape
    K, J, I = np.meshgrid(np.arange(mz), np.arange(my), np.arange(mx), indexing="ij")
    rows = (K >= 1) & (K < mz - 1) & (J >= 1) & (J < my - 1) & (I >= 1) & (I < mx - 1)
    idx = -np.ones(o.shape, int)
    idx[rows] = np.arange(rows.sum())
    n = int(rows.sum())
    M = sp.lil_matrix((n, n))
    colour = (I % 3) + 3 * (J % 3) + 9 * (K % 3)
    for q in range(27):
        x = ((colour == q) & rows).astype(float)
        y = o.poisson_apply(x)
        # each row sees at most one probed column per colour: its 3x3x3 neighbour of that colour
        for (k, j, i) in zip(*np.nonzero(rows & (y != 0))):
            for dk in (-1, 0, 1):
                for dj in (-1, 0, 1):
                    for di in (-1, 0, 1):
                        kk, jj, ii = (k + dk - 1) % (mz - 2) + 1, j + dj, (i + di - 1) % (mx - 2) + 1
                        if 1 <= jj < my - 1 and colour[kk, jj, ii] == q and rows[kk, jj, ii]:
                            M[idx[k, j, i], idx[kk, jj, ii]] = y[k, j, i]
    return M.tocsr(), o, rows, idx


def test_channel_operator_hierarchy():
    M, o, rows, idx = channel_matrix()
    # the probed matrix reproduces the oracle's operator
    x = np.random.default_rng(2).standard_normal(o.shape) * rows
    y = o.poisson_apply(x)
    assert np.allclose(M @ x[rows], y[rows], rtol=0, atol=1e-12 * np.abs(y).max())
    h = host(M, theta=0.75)
    assert h.levels >= 2
    nnz = sum(h.matrix(lev, 0)[1].size for lev in range(h.levels))
    assert nnz / M.nnz < 4.0
    assert h.matrix(h.levels - 1, 0)[3][0] <= 512


def test_smoother_colouring_is_independent():
    """Rows of one colour share no matrix entry (so one launch per colour is a
    Gauss-Seidel sweep whatever the thread order), every row has one colour,
    and the 19-point level-0 operator needs no more colours than the 2x2x2
    parity colouring of a 27-point stencil."""
    M, o, rows, idx = channel_matrix()
    h = host(M, theta=0.75)
    for lev in range(h.levels):
        A = csr_of(*h.matrix(lev, 0)).tocsr()
        r, off = h.colouring(lev)
        assert np.array_equal(np.sort(r), np.arange(A.shape[0]))
        assert off[0] == 0 and off[-1] == A.shape[0] and np.all(np.diff(off) > 0)
        colour = np.empty(A.shape[0], int)
        for c in range(len(off) - 1):
            colour[r[off[c]:off[c + 1]]] = c
        Ac = A.tocoo()
        off_diag = Ac.row != Ac.col
        assert not np.any(colour[Ac.row[off_diag]] == colour[Ac.col[off_diag]])
        if lev == 0:
            assert len(off) - 1 <= 8
