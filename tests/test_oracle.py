"""The oracle pinned: each tile routine against what the mathematics fixes
(closed forms, identities, brute force), the clamp rule, and the whole CPU
factor+solve against the reference's FP64 solutions and scipy."""
import numpy as np
import pytest

import oracle
import sogl
import workloads as W

rng = np.random.default_rng(12345)


def dd_tile(scale=1.0):
    a = rng.standard_normal((64, 64)) * scale
    a[np.diag_indices(64)] = np.abs(a).sum(1) + 1.0
    return a


def test_gemm_matches_numpy_and_sign():
    a, b, c = rng.standard_normal((3, 64, 64))
    assert np.allclose(oracle.gemm(a, b, c, 1), c + a @ b, rtol=0, atol=1e-12)
    assert np.allclose(oracle.gemm(a, b, c, -1), c - a @ b, rtol=0, atol=1e-12)
    # operand order matters: a transposed operand must fail this
    assert not np.allclose(oracle.gemm(a, b), b @ a)


def test_sub_copy_neg_exact():
    a, b = rng.standard_normal((2, 64, 64))
    assert np.array_equal(oracle.sub(a, b), a - b)
    assert np.array_equal(oracle.sub(a, None), a)
    assert np.array_equal(oracle.sub(None, b), -b)


def test_lu_reconstructs_and_is_triangular():
    a = dd_tile()
    l, u = oracle.lu(a)
    assert np.allclose(l @ u, a, rtol=0, atol=1e-12 * np.abs(a).max())
    assert np.array_equal(np.diag(l), np.ones(64))
    assert np.all(np.triu(l, 1) == 0) and np.all(np.tril(u, -1) == 0)
    # against a textbook no-pivot elimination (brute force)
    m = a.copy()
    for k in range(64):
        m[k + 1:, k] /= m[k, k]
        m[k + 1:, k + 1:] -= np.outer(m[k + 1:, k], m[k, k + 1:])
    assert np.allclose(np.triu(m), u, rtol=1e-13, atol=1e-13)


def test_lu_pivot_clamp():
    a = np.eye(64)
    a[5, 5] = 0.0          # zero pivot -> +1e-9
    a[9, 9] = -3e-10       # tiny negative -> -1e-9
    l, u = oracle.lu(a)
    assert u[5, 5] == 1e-9 and u[9, 9] == -1e-9


def test_triangular_inverses():
    a = dd_tile()
    l, u = oracle.lu(a)
    yl, yu = oracle.inv_lower(l), oracle.inv_upper(u)
    assert np.allclose(l @ yl, np.eye(64), atol=1e-13)
    assert np.allclose(u @ yu, np.eye(64), atol=1e-13)
    assert np.all(np.triu(yl, 1) == 0) and np.all(np.tril(yu, -1) == 0)
    # general (non-unit) diagonal of L, as inv_lower divides by l_ii
    g = np.tril(dd_tile())
    assert np.allclose(g @ oracle.inv_lower(g), np.eye(64), atol=1e-13)


def spd_tile():
    m = rng.standard_normal((64, 64))
    return m @ m.T + 64 * np.eye(64)


def test_gemm_t_matches_numpy():
    a, b, c = rng.standard_normal((3, 64, 64))
    assert np.allclose(oracle.gemm_t(a, b, c), c + a @ b.T, rtol=0, atol=1e-12)
    assert np.allclose(oracle.gemm_t(a, a), a @ a.T, rtol=0, atol=1e-12)
    assert not np.allclose(oracle.gemm_t(a, b), a @ b)  # the transpose is there


def test_llt_matches_cholesky_and_clamp():
    a = spd_tile()
    l = oracle.llt(a)
    assert np.all(np.triu(l, 1) == 0)
    assert np.allclose(l, np.linalg.cholesky(a), rtol=0, atol=1e-12 * np.abs(l).max())
    assert np.allclose(l @ l.T, a, rtol=0, atol=1e-12 * np.abs(a).max())
    # the upper triangle is never read
    b = a.copy()
    b[np.triu_indices(64, 1)] = 7.0
    assert np.array_equal(oracle.llt(b), l)
    # a non-positive pivot is clamped to 1e-20 before the square root
    z = np.eye(64)
    z[3, 3] = -2.0
    assert oracle.llt(z)[3, 3] == np.sqrt(1e-20)


SYM = {
    "sym_poisson2d_64": lambda: W.symmetric_lower(W.poisson2d(64)),
    "sym_diffusion3d_16": lambda: W.symmetric_lower(W.diffusion_varcoef(16, 3, 21)),
    "sym_diffusion2d_100": lambda: W.symmetric_lower(W.diffusion_varcoef(100, 2, 22)),
}


@pytest.mark.parametrize("name", list(SYM))
def test_llt_factor_solve_vs_reference(name, golden):
    """Symmetric input: L L^T factor and the L, L^T substitutions against the
    reference's own symmetric solve."""
    g = golden(name)
    p = SYM[name]()
    pl = sogl.Plan(p.n, p.row, p.col, p.val, p.b, lib="libsogl_host.so", symmetric=True)
    xp, yp = oracle.factor_solve(pl, pl.permute_rhs(p.b), threads=8)
    x = pl.unpermute(xp)
    ref = g["xorig"]
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 1e-12
    assert np.linalg.norm(yp - g["y"]) / np.linalg.norm(g["y"]) < 1e-12
    r, c, v = W.full_matrix(p)
    assert sogl.oracle_residual(p.n, r, c, v, x, p.b) < 1e-15


CASES = {
    "c1": lambda: W.config("c1"),
    "disconnected_40": lambda: W.edge_disconnected(40),
    "zerodiag_48": lambda: W.edge_zero_diag(48),
    "unsym_3000": lambda: W.edge_unsym_pattern(3000),
    "convdiff2d_128": lambda: W.convdiff2d(128),
    "diffusion3d_20": lambda: W.diffusion_varcoef(20, 3, 5),
    "random_band_2048": lambda: W.random_band(2048, nrhs=1),
}


@pytest.mark.parametrize("name", list(CASES))
def test_factor_solve_vs_reference(name, golden):
    g = golden(name)
    p = CASES[name]()
    b = p.b if p.b.ndim == 1 else p.b[:, 0]
    pl = sogl.Plan(p.n, p.row, p.col, p.val, b, lib="libsogl_host.so")
    xp, yp = oracle.factor_solve(pl, pl.permute_rhs(b), threads=8)
    x = pl.unpermute(xp)
    ref = g["xorig"]
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 1e-12
    assert np.linalg.norm(yp - g["y"]) / np.linalg.norm(g["y"]) < 1e-12
    assert sogl.oracle_residual(p.n, p.row, p.col, p.val, x, b) < 1e-15


def test_factor_solve_vs_scipy():
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    p = W.convdiff2d(96)
    pl = sogl.Plan(p.n, p.row, p.col, p.val, p.b, lib="libsogl_host.so")
    xp, _ = oracle.factor_solve(pl, pl.permute_rhs(p.b), threads=8)
    x = pl.unpermute(xp)
    A = sp.csc_matrix((p.val, (p.row, p.col)), shape=(p.n, p.n))
    xs = sla.spsolve(A, p.b)
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-12


def test_planted_mistake_is_caught():
    """A dropped Schur term (sub as copy) must break agreement with the reference."""
    p = W.config("c1")
    pl = sogl.Plan(p.n, p.row, p.col, p.val, p.b, lib="libsogl_host.so")
    ops = pl.ops(0)
    sub = np.nonzero((ops[:, 2] == 4) & (ops[:, 0] > 0) & (ops[:, 1] > 0))[0][0]
    bad = ops.copy()
    bad[sub, 0] = 0  # result = src2 - src  ->  result = src2

    class P:
        pass
    q = P()
    q.ops = lambda level=0: bad
    q.inputs, q.tile_stages, q.factor_tiles = pl.inputs, pl.tile_stages, pl.factor_tiles
    q.info, q.N = pl.info, pl.N
    xp, _ = oracle.factor_solve(q, pl.permute_rhs(p.b))
    x = pl.unpermute(xp)
    assert sogl.oracle_residual(p.n, p.row, p.col, p.val, x, p.b) > 1e-10
