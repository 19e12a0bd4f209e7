"""GPU parity: each HIP tile kernel against the oracle through the C ABI, and the
whole factor + solve against the reference's FP64 solutions (golden fixtures) and
the oracle.  Tolerances: FP64 paths 1e-12 relative in x (1e-13 for single tiles),
FP32 tiles 2e-5 relative, FP32/bf16 factors after FP64 refinement: oracle
residual <= 1e-12."""
import numpy as np
import pytest

import oracle
import sogl
import workloads as W

pytestmark = pytest.mark.gpu

rng = np.random.default_rng(7)


@pytest.fixture(scope="module")
def dev():
    d = sogl.Device(0)
    yield d
    d.close()


def dd_tile(scale=1.0):
    a = rng.standard_normal((64, 64)) * scale
    a[np.diag_indices(64)] = np.abs(a).sum(1) + 1.0
    return a


def op(src, src2, code, res, res2=0, stage=2, seq=0, group=0):
    return [src, src2, code, res, res2, stage, seq, group]


def run_ops(dev, n_ids, ops, inputs, out_id, precision=sogl.FP64, out2=None):
    ids = np.arange(1, len(inputs) + 1, dtype=np.int32)
    tiles = np.stack(inputs).astype(np.float64)
    lt = np.array([[0, 0, out_id]], np.int32)
    ut = np.array([[0, 0, out2 if out2 else out_id]], np.int32)
    fid = dev.load_ops(n_ids, np.array(ops, np.int32), ids, tiles, lt, ut, 1, precision)
    dev.factor(fid, sogl.EXEC_PERSISTENT)
    o1 = dev.factor_tiles(fid, 0, 1)[0]
    o2 = dev.factor_tiles(fid, 1, 1)[0]
    dev.release(fid)
    return o1, o2


@pytest.mark.parametrize("precision,tol", [(sogl.FP64, 1e-13), (sogl.FP32, 2e-5), (sogl.BF16, 2e-2)])
@pytest.mark.parametrize("npairs", [1, 5])
def test_gemm_tile(dev, precision, tol, npairs):
    mats = [rng.standard_normal((64, 64)) for _ in range(2 * npairs)]
    r = 2 * npairs + 1
    ops = [op(2 * s + 1, 2 * s + 2, 8 if s % 2 == 0 else 9, r) for s in range(npairs)]
    c, _ = run_ops(dev, r + 1, ops, mats, r, precision)
    want = np.zeros((64, 64))
    for s in range(npairs):
        want = oracle.gemm(mats[2 * s], mats[2 * s + 1], want, 1 if s % 2 == 0 else -1)
    err = np.abs(c - want).max() / np.abs(want).max()
    assert err < tol, err


@pytest.mark.parametrize("precision,tol", [(sogl.FP64, 1e-13), (sogl.FP32, 2e-5)])
def test_gemm_task_windows(dev, precision, tol):
    """One GEMM task of 70 pairs: more than the 64-pair mask ring and three flag
    windows of at most 32 pairs.  Eight of its first operands are made by other GEMM
    tasks of the same run, so the task waits for them after its ready windows; every
    fifth second operand is a zero tile (a product the block masks skip, with no
    barrier) and others have zero 16x16 blocks (loaded as zeros).  Two runs must agree
    bit for bit and match the oracle's sum."""
    g = np.random.default_rng(77)
    ins = []

    def add(t):
        ins.append(t)
        return len(ins)  # input ids are 1..len(ins)

    made, pairs = {}, []
    for k in range(70):
        y = g.standard_normal((64, 64))
        if k % 5 == 3:
            y = np.zeros((64, 64))
        elif k % 4 == 1:
            y[16:32, :] = 0
            y[:, 48:] = 0
        if k % 9 == 4:
            made[k] = (add(g.standard_normal((64, 64))), add(g.standard_normal((64, 64))))
            xid, x = None, None
        else:
            x = g.standard_normal((64, 64))
            xid = add(x)
        pairs.append((k, xid, add(y), x, y))
    nid = len(ins) + 1
    ops, xs = [], {}
    for k, (uid, vid) in made.items():
        ops.append(op(uid, vid, 8, nid, stage=1))
        xs[k] = (nid, oracle.gemm(ins[uid - 1], ins[vid - 1], np.zeros((64, 64)), 1))
        nid += 1
    r = nid
    want = np.zeros((64, 64))
    for k, xid, yid, x, y in pairs:
        if xid is None:
            xid, x = xs[k]
        sign = 1 if k % 2 == 0 else -1
        ops.append(op(xid, yid, 8 if sign > 0 else 9, r, stage=2))
        want = oracle.gemm(x, y, want, sign)
    c1, _ = run_ops(dev, r + 1, ops, ins, r, precision)
    c2, _ = run_ops(dev, r + 1, ops, ins, r, precision)
    assert np.array_equal(c1, c2)
    err = np.abs(c1 - want).max() / np.abs(want).max()
    assert err < tol, err


@pytest.mark.parametrize("precision,tol", [(sogl.FP64, 1e-13), (sogl.FP32, 2e-5)])
def test_lu_tile(dev, precision, tol):
    a = dd_tile()
    l, u = run_ops(dev, 4, [op(1, 0, 1, 2, 3)], [a], 2, precision, out2=3)
    lo, uo = oracle.lu(a)
    assert np.abs(l - lo).max() < tol * 10
    assert np.abs(u - uo).max() / np.abs(uo).max() < tol
    assert np.array_equal(np.diag(l), np.ones(64))


def test_lu_tile_clamp(dev):
    a = np.eye(64)
    a[5, 5] = 0.0
    a[9, 9] = -3e-10
    l, u = run_ops(dev, 4, [op(1, 0, 1, 2, 3)], [a], 2, out2=3)
    assert u[5, 5] == 1e-9 and u[9, 9] == -1e-9


def test_lu_tile_fp32_growth_check(dev):
    """FP32 lu tiles report the pivot-free growth max|U| / max|A|; above the limit
    (1e4) the tile is factored again in FP64 (L, U and both inverses)."""
    a = dd_tile()
    fid = dev.load_ops(4, np.array([op(1, 0, 1, 2, 3)], np.int32), np.array([1], np.int32), np.stack([a]),
                       np.array([[0, 0, 2]], np.int32), np.array([[0, 0, 3]], np.int32), 1, sogl.FP32)
    dev.factor(fid, sogl.EXEC_PERSISTENT)
    st = dev.stats(fid)
    dev.release(fid)
    lo, uo = oracle.lu(a)
    assert st.growth_tiles == 0
    assert abs(st.max_growth - np.abs(uo).max() / np.abs(a).max()) < 1e-5 * st.max_growth
    # a 1e-6 leading pivot (above the 1e-9 clamp): u ~ 1e6 against max |a| ~ 64
    g = dd_tile()
    g[0, 0] = 1e-6
    g[0, 1:] = 1.0
    g[1:, 0] = 1.0
    ops = [op(1, 0, 1, 2, 3, stage=2, seq=0), op(2, 0, 2, 4, stage=3, seq=1), op(3, 0, 3, 5, stage=3, seq=2)]
    fid = dev.load_ops(7, np.array(ops, np.int32), np.array([1, 6], np.int32), np.stack([g, np.eye(64)]),
                       np.array([[0, 0, 2], [1, 0, 4], [1, 1, 6]], np.int32),
                       np.array([[0, 0, 3], [0, 1, 5], [1, 1, 6]], np.int32), 2, sogl.FP32)
    dev.factor(fid, sogl.EXEC_PERSISTENT)
    st = dev.stats(fid)
    l, x, _ = dev.factor_tiles(fid, 0, 3)
    u, z, _ = dev.factor_tiles(fid, 1, 3)
    dev.release(fid)
    lo, uo = oracle.lu(g)
    go = np.abs(uo).max() / np.abs(g).max()
    assert go > 1e4 and st.growth_tiles == 1 and abs(st.max_growth - go) < 1e-4 * go
    for got, want in ((l, lo), (u, uo), (x, oracle.inv_lower(lo)), (z, oracle.inv_upper(uo))):
        assert np.abs(got - want).max() <= 1e-6 * np.abs(want).max()


@pytest.mark.parametrize("precision", [sogl.FP64, sogl.FP32])
def test_upper_inverse_check(dev, precision):
    """The reference's check of each upper inverse (inv_check_diag): a well-conditioned
    tile passes; a tile holding a NaN fails on the diagonal of U U^-1 with the
    reference's max of 10, at its block row, and the factor run still completes."""
    ops = [op(1, 0, 1, 2, 3, stage=2, seq=0), op(3, 0, 3, 5, stage=3, seq=1)]
    out = []
    for bad in (False, True):
        a = dd_tile()
        if bad:
            a[5, 9] = np.nan
        fid = dev.load_ops(6, np.array(ops, np.int32), np.array([1], np.int32), a[None],
                           np.array([[0, 0, 2]], np.int32), np.array([[0, 0, 3]], np.int32), 1, precision)
        dev.factor(fid)
        out.append((dev.uinv_check(fid), dev.stats(fid).uinv_out_of_tol))
        dev.release(fid)
    (n0, rows0), s0 = out[0]
    (n1, rows1), s1 = out[1]
    assert n0 == 0 and rows0 == [] and s0 == 0
    assert n1 == 1 and s1 == 1
    blk, kind, mn, mx = rows1[0]
    assert blk == 0 and kind == "diag" and mx == 10.0 and mn <= 1.0


@pytest.mark.parametrize("mode", [sogl.EXEC_PERSISTENT, sogl.EXEC_LEVELS])
@pytest.mark.parametrize("precision,tol", [(sogl.FP64, 1e-13), (sogl.FP32, 2e-5)])
def test_lu_tile_fused_inverses(dev, precision, tol, mode):
    """lu followed by lowerInv / upperInv of its outputs: the inverses are made inside
    the lu task from the inverses of its 16 x 16 diagonal blocks; every block of both
    inverses against the oracle's inv_lower / inv_upper of the oracle's factors.
    X and Z are read back as extra factor tiles of a 2 x 2 block map (an identity input
    tile fills its second diagonal block)."""
    a = dd_tile()
    ops = [op(1, 0, 1, 2, 3, stage=2, seq=0), op(2, 0, 2, 4, stage=3, seq=1), op(3, 0, 3, 5, stage=3, seq=2)]
    fid = dev.load_ops(7, np.array(ops, np.int32), np.array([1, 6], np.int32), np.stack([a, np.eye(64)]),
                       np.array([[0, 0, 2], [1, 0, 4], [1, 1, 6]], np.int32),
                       np.array([[0, 0, 3], [0, 1, 5], [1, 1, 6]], np.int32), 2, precision)
    dev.factor(fid, mode)
    l, x, _ = dev.factor_tiles(fid, 0, 3)
    u, z, _ = dev.factor_tiles(fid, 1, 3)
    dev.release(fid)
    lo, uo = oracle.lu(a)
    xo, zo = oracle.inv_lower(lo), oracle.inv_upper(uo)
    assert np.abs(l - lo).max() < tol * 10
    assert np.abs(u - uo).max() / np.abs(uo).max() < tol
    for name, got, want in (("L^-1", x, xo), ("U^-1", z, zo)):
        for bi in range(4):
            for bj in range(4):
                g = got[16 * bi:16 * bi + 16, 16 * bj:16 * bj + 16]
                w = want[16 * bi:16 * bi + 16, 16 * bj:16 * bj + 16]
                assert np.abs(g - w).max() <= tol * np.abs(want).max(), (name, bi, bj)
    assert np.all(np.triu(x, 1) == 0) and np.all(np.tril(z, -1) == 0)


@pytest.mark.parametrize("lower", [True, False])
@pytest.mark.parametrize("precision,tol", [(sogl.FP64, 1e-13), (sogl.FP32, 2e-5)])
def test_triangular_inverse_tile(dev, lower, precision, tol):
    t = np.tril(dd_tile()) if lower else np.triu(dd_tile())
    y, _ = run_ops(dev, 3, [op(1, 0, 2 if lower else 3, 2)], [t], 2, precision)
    want = oracle.inv_lower(t) if lower else oracle.inv_upper(t)
    assert np.abs(y - want).max() / np.abs(want).max() < tol


def test_sub_copy_neg_tiles_exact(dev):
    a, b = rng.standard_normal((2, 64, 64))
    r, _ = run_ops(dev, 4, [op(2, 1, 4, 3)], [a, b], 3)
    assert np.array_equal(r, a - b)  # sub(src=b, src2=a) = a - b
    r, _ = run_ops(dev, 4, [op(0, 1, 4, 3)], [a, b], 3)
    assert np.array_equal(r, a)
    r, _ = run_ops(dev, 4, [op(2, 0, 4, 3)], [a, b], 3)
    assert np.array_equal(r, -b)


# ---- L L^T (symmetric input)
LLT, MULT, LINV, SUBOP = 10, 11, 2, 4


def spd_tile(scale=1.0):
    m = rng.standard_normal((64, 64))
    return (m @ m.T + 64 * np.eye(64)) * scale


def run_llt(dev, n_ids, ops, inputs, ltiles, block_rows, precision=sogl.FP64, b=None):
    """Load an L L^T op list, factor, return the L tiles and (with b) x and y."""
    ids = np.arange(1, len(inputs) + 1, dtype=np.int32)
    fid = dev.load_ops(n_ids, np.array(ops, np.int32), ids, np.stack(inputs).astype(np.float64),
                       np.array(ltiles, np.int32), np.zeros((0, 3), np.int32), block_rows, precision)
    dev.factor(fid, sogl.EXEC_PERSISTENT)
    lt = dev.factor_tiles(fid, 0, len(ltiles))
    out = (lt,)
    if b is not None:
        x, y = dev.solve(fid, b, want_y=True)
        out = (lt, x, y)
    dev.release(fid)
    return out


@pytest.mark.parametrize("precision,tol", [(sogl.FP64, 1e-13), (sogl.FP32, 2e-5)])
def test_llt_tile(dev, precision, tol):
    a = spd_tile()
    b = rng.standard_normal(64)
    (l,), x, y = run_llt(dev, 3, [op(1, 0, LLT, 2)], [a], [[0, 0, 2]], 1, precision, b)
    want = oracle.llt(a)
    assert np.all(np.triu(l, 1) == 0)
    assert np.abs(l - want).max() / np.abs(want).max() < tol
    # L y = b, L^T x = y through the inverse and transposed tiles; y as the reference
    # leaves it: diag(L) y
    xs = np.linalg.solve(a, b)
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 50 * tol
    yl = np.linalg.solve(want, b)
    assert np.linalg.norm(y - np.diag(want) * yl) / np.linalg.norm(yl) < 50 * tol


@pytest.mark.parametrize("precision,tol", [(sogl.FP64, 1e-13), (sogl.FP32, 2e-5)])
def test_llt_tile_fused_inverse(dev, precision, tol):
    """llt followed by lowerInv of its factor: L^-1 made inside the llt task (rows of
    L_u^-1 scaled by 1 / d), every block against the inverse of the oracle's factor."""
    a = spd_tile()
    ops = [op(1, 0, LLT, 2, stage=2, seq=0), op(2, 0, LINV, 4, stage=3, seq=1)]
    fid = dev.load_ops(6, np.array(ops, np.int32), np.array([1, 5], np.int32), np.stack([a, np.eye(64)]),
                       np.array([[0, 0, 2], [1, 0, 4], [1, 1, 5]], np.int32), np.zeros((0, 3), np.int32), 2,
                       precision)
    dev.factor(fid, sogl.EXEC_PERSISTENT)
    l, x, _ = dev.factor_tiles(fid, 0, 3)
    dev.release(fid)
    want = oracle.llt(a)
    xo = oracle.inv_lower(want)
    assert np.abs(l - want).max() / np.abs(want).max() < tol
    for bi in range(4):
        for bj in range(4):
            g = x[16 * bi:16 * bi + 16, 16 * bj:16 * bj + 16]
            w = xo[16 * bi:16 * bi + 16, 16 * bj:16 * bj + 16]
            assert np.abs(g - w).max() <= tol * np.abs(xo).max(), (bi, bj)


def test_llt_tile_clamp(dev):
    z = np.eye(64)
    z[3, 3] = -2.0
    l = run_llt(dev, 3, [op(1, 0, LLT, 2)], [z], [[0, 0, 2]], 1)[0][0]
    assert l[3, 3] == np.sqrt(1e-20) and l[4, 4] == 1.0


@pytest.mark.parametrize("precision,tol", [(sogl.FP64, 1e-12), (sogl.FP32, 2e-5), (sogl.BF16, 2e-2)])
def test_llt_two_by_two_blocks(dev, precision, tol):
    """The L L^T recursion one level up, on 2 x 2 tiles: L11 = llt(A11), X = L11^-1,
    L21 = A21 X^T (mult), S = L21 L21^T (mult, self product), D = A22 - S (folded into
    the GEMM), L22 = llt(D); against the Cholesky factor of the 128 x 128 matrix."""
    m = rng.standard_normal((128, 128))
    a = m @ m.T + 128 * np.eye(128)
    a11, a21, a22 = a[:64, :64], a[64:, :64], a[64:, 64:]
    ops = [op(1, 0, LLT, 4, stage=2, seq=0), op(4, 0, LINV, 5, stage=3, seq=1),
           op(2, 5, MULT, 6, stage=4, seq=2), op(6, 6, MULT, 7, stage=5, seq=3),
           op(7, 3, SUBOP, 8, stage=6, seq=4), op(8, 0, LLT, 9, stage=7, seq=5)]
    b = rng.standard_normal(128)
    lt, x, y = run_llt(dev, 10, ops, [a11, a21, a22], [[0, 0, 4], [1, 0, 6], [1, 1, 9]], 2, precision, b)
    L = np.zeros((128, 128))
    L[:64, :64], L[64:, :64], L[64:, 64:] = lt[0], lt[1], lt[2]
    want = np.linalg.cholesky(a)
    assert np.abs(L - want).max() / np.abs(want).max() < tol
    xs = np.linalg.solve(a, b)
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 100 * tol


SYM = {
    "sym_poisson2d_64": lambda: W.symmetric_lower(W.poisson2d(64)),
    "sym_diffusion3d_16": lambda: W.symmetric_lower(W.diffusion_varcoef(16, 3, 21)),
    "sym_diffusion2d_100": lambda: W.symmetric_lower(W.diffusion_varcoef(100, 2, 22)),
}


@pytest.mark.parametrize("name", list(SYM))
def test_llt_solve_vs_reference(dev, name, golden):
    """Symmetric input end to end in FP64 against the reference's own symmetric solve:
    x, the right-hand side as the reference leaves it, the residual, the oracle."""
    g = golden(name)
    p = SYM[name]()
    pl = sogl.Plan(p.n, p.row, p.col, p.val, p.b, symmetric=True)
    fid = dev.load(pl, sogl.FP64)
    dev.factor(fid)
    xp, yp = dev.solve(fid, pl.permute_rhs(p.b), want_y=True)
    dev.release(fid)
    x = pl.unpermute(xp)
    ref = g["xorig"]
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 1e-12
    assert np.linalg.norm(yp - g["y"]) / np.linalg.norm(g["y"]) < 1e-12
    r, c, v = W.full_matrix(p)
    assert sogl.oracle_residual(p.n, r, c, v, x, p.b) < 1e-15
    xo, _ = oracle.factor_solve(pl, pl.permute_rhs(p.b), threads=8)
    assert np.linalg.norm(xp - xo) / np.linalg.norm(xo) < 1e-12


def test_llt_executor_modes_and_refinement(dev, golden):
    p = SYM["sym_diffusion2d_100"]()
    ref = golden("sym_diffusion2d_100")["xorig"]
    pl = sogl.Plan(p.n, p.row, p.col, p.val, p.b, symmetric=True)
    xs = []
    for mode in (sogl.EXEC_PERSISTENT, sogl.EXEC_LEVELS, sogl.EXEC_GRAPH, sogl.EXEC_PER_OP):
        dev.set_pool_reuse(sogl.REUSE_OFF if mode == sogl.EXEC_PER_OP else sogl.REUSE_ON)
        fid = dev.load(pl, sogl.FP64)
        dev.set_pool_reuse(sogl.REUSE_AUTO)
        dev.factor(fid, mode)
        xs.append(dev.solve(fid, pl.permute_rhs(p.b)))
        dev.release(fid)
    assert np.array_equal(xs[0], xs[1]) and np.array_equal(xs[0], xs[2])
    assert np.linalg.norm(xs[0] - xs[3]) / np.linalg.norm(xs[3]) < 1e-13
    r, c, v = W.full_matrix(p)
    for prec in (sogl.FP32, sogl.BF16):
        fid = dev.load(pl, prec)
        dev.factor(fid)
        dev.attach_refine(fid, pl)
        x, it, res, rc = dev.solve_refined(fid, p.b, tol=1e-12, max_iter=60)
        dev.release(fid)
        assert rc == 0, (prec, it, res)
        assert sogl.oracle_residual(p.n, r, c, v, x, p.b) <= 1e-12
        # forward error <= cond(A) x residual; log-normal coefficients give cond ~ 1e5
        assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 1e-6


@pytest.mark.slow
def test_llt_full_size(dev, golden):
    """3D 48^3 variable-coefficient diffusion, symmetric input (n=110,592): FP64 vs the
    reference's symmetric solve (sample), FP32 factor + FP64 refinement to 1e-12."""
    g = golden("sym_diffusion3d_48")
    p = W.symmetric_lower(W.diffusion_varcoef(48, 3, 23))
    pl = sogl.Plan(p.n, p.row, p.col, p.val, p.b, symmetric=True)
    fid = dev.load(pl, sogl.FP64)
    dev.factor(fid)
    x = pl.unpermute(dev.solve(fid, pl.permute_rhs(p.b)))
    dev.release(fid)
    idx, ref = g["xorig_idx"], g["xorig_sample"]
    assert np.linalg.norm(x[idx] - ref) / np.linalg.norm(ref) < 1e-12
    assert abs(np.linalg.norm(x) - float(g["xorig_norm"])) / float(g["xorig_norm"]) < 1e-12
    r, c, v = W.full_matrix(p)
    assert sogl.oracle_residual(p.n, r, c, v, x, p.b) < 1e-15
    fid = dev.load(pl, sogl.FP32)
    dev.factor(fid)
    dev.attach_refine(fid, pl)
    x32, it, res, rc = dev.solve_refined(fid, p.b, tol=1e-12, max_iter=30)
    dev.release(fid)
    assert rc == 0 and res <= 1e-12
    # forward error <= cond(A) x residual (log-normal coefficients: cond ~ 1e5)
    assert np.linalg.norm(x32[idx] - ref) / np.linalg.norm(ref) < 1e-6


CASES = {
    "c1": lambda: W.config("c1"),
    "disconnected_40": lambda: W.edge_disconnected(40),
    "zerodiag_48": lambda: W.edge_zero_diag(48),
    "unsym_3000": lambda: W.edge_unsym_pattern(3000),
    "convdiff2d_128": lambda: W.convdiff2d(128),
    "diffusion3d_20": lambda: W.diffusion_varcoef(20, 3, 5),
    "random_band_2048": lambda: W.random_band(2048, nrhs=1),
}


def solve_fp64(dev, p, mode=sogl.EXEC_PERSISTENT, b=None):
    b = (p.b if p.b.ndim == 1 else p.b[:, 0]) if b is None else b
    pl = sogl.Plan(p.n, p.row, p.col, p.val, b)
    dev.set_pool_reuse(sogl.REUSE_OFF if mode == sogl.EXEC_PER_OP else sogl.REUSE_ON)
    fid = dev.load(pl, sogl.FP64)
    dev.set_pool_reuse(sogl.REUSE_AUTO)
    dev.factor(fid, mode)
    bp = pl.permute_rhs(b)
    xp, yp = dev.solve(fid, bp, want_y=True)
    dev.release(fid)
    return pl, pl.unpermute(xp), xp, yp


@pytest.mark.parametrize("name", list(CASES))
def test_solve_vs_reference(dev, name, golden):
    g = golden(name)
    p = CASES[name]()
    b = p.b if p.b.ndim == 1 else p.b[:, 0]
    pl, x, xp, yp = solve_fp64(dev, p)
    ref = g["xorig"]
    err = np.linalg.norm(x - ref) / np.linalg.norm(ref)
    assert err < 1e-12, err
    assert np.linalg.norm(yp - g["y"]) / np.linalg.norm(g["y"]) < 1e-12
    assert sogl.oracle_residual(p.n, p.row, p.col, p.val, x, b) < 1e-15
    xo, _ = oracle.factor_solve(pl, pl.permute_rhs(b), threads=8)
    assert np.linalg.norm(xp - xo) / np.linalg.norm(xo) < 1e-12


@pytest.mark.parametrize("name", ["c1", "diffusion3d_20"])
def test_executor_modes(dev, name):
    """persistent and per-level executors run the same tasks with the same pair order:
    identical bits; the per-op executor sums each tile's products in plan order."""
    p = CASES[name]()
    xs = [solve_fp64(dev, p, mode)[2] for mode in (sogl.EXEC_PERSISTENT, sogl.EXEC_LEVELS,
                                                     sogl.EXEC_PER_OP)]
    assert np.array_equal(xs[0], xs[1])
    assert np.linalg.norm(xs[0] - xs[2]) / np.linalg.norm(xs[2]) < 1e-13


@pytest.mark.parametrize("name", ["c1", "convdiff2d_128", "random_band_2048"])
def test_block_masks(dev, name, golden):
    """16x16 block masks: a masked factor (zero blocks skipped in products and
    substitutions) against a dense one and the reference x.  The automatic policy keeps
    the masks after the first run for FP64 factors, and for FP32 factors only when a
    quarter of the factor's blocks are zero."""
    p = CASES[name]()
    b = p.b if p.b.ndim == 1 else p.b[:, 0]
    pl = sogl.Plan(p.n, p.row, p.col, p.val, b)
    ref = golden(name)["xorig"]
    xs, zs = {}, None
    for mode in (sogl.MASKS_ON, sogl.MASKS_OFF):
        fid = dev.load(pl, sogl.FP64)
        dev.set_masks(fid, mode)
        dev.factor(fid)
        st = dev.stats(fid)
        if mode == sogl.MASKS_ON:
            zs = st.zero_blocks
            assert 0.0 <= zs <= 1.0 and st.block_masks == 1
        else:
            assert st.block_masks == 0
        xs[mode] = pl.unpermute(dev.solve(fid, pl.permute_rhs(b)))
        dev.release(fid)
    for x in xs.values():
        assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 1e-12
    assert np.linalg.norm(xs[sogl.MASKS_ON] - xs[sogl.MASKS_OFF]) / np.linalg.norm(ref) < 1e-13
    fid = dev.load(pl, sogl.FP64)  # automatic (the default)
    dev.factor(fid)
    st = dev.stats(fid)
    assert abs(st.zero_blocks - zs) < 1e-12 and st.block_masks == 1
    dev.factor(fid)
    x2 = pl.unpermute(dev.solve(fid, pl.permute_rhs(b)))
    dev.release(fid)
    assert np.linalg.norm(x2 - ref) / np.linalg.norm(ref) < 1e-12
    fid = dev.load(pl, sogl.FP32)
    dev.factor(fid)
    st = dev.stats(fid)
    dev.release(fid)
    assert 0.0 <= st.zero_blocks <= 1.0 and st.block_masks == (1 if st.zero_blocks >= 0.25 else 0)


def _factor_tiles(dev, pl, run):
    fid = dev.load(pl, sogl.FP64)
    run(fid)
    lt = dev.factor_tiles(fid, 0, pl.info.n_l_tiles)
    ut = dev.factor_tiles(fid, 1, pl.info.n_u_tiles)
    return fid, lt, ut


@pytest.mark.parametrize("name", ["c1", "repeated_1200"])
@pytest.mark.parametrize("precision", [sogl.FP64, sogl.FP32])
def test_plan_cells_match_dense_inputs(dev, name, precision):
    """A plan loads its input tiles as cells scattered on the device; the same tiles
    given densely through the factor description must give the same factor bits."""
    p = W.edge_repeated_entries(1200) if name == "repeated_1200" else CASES[name]()
    pl = sogl.Plan(p.n, p.row, p.col, p.val, p.b)
    ids, _, _, tiles = pl.inputs()
    f1 = dev.load(pl, precision)
    f2 = dev.load_ops(pl.info.storage_count, pl.ops(0), ids, tiles, np.stack(pl.factor_tiles(0), 1),
                      np.stack(pl.factor_tiles(1), 1), pl.info.block_rows, precision)
    out = []
    for fid in (f1, f2):
        dev.factor(fid)
        out.append((dev.factor_tiles(fid, 0, pl.info.n_l_tiles), dev.factor_tiles(fid, 1, pl.info.n_u_tiles)))
        dev.release(fid)
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("name", ["c1", "diffusion3d_20"])
def test_level_graph_replays(dev, name):
    """The captured level schedule gives the levels executor's bits on every replay,
    with eager work (a solve on the same stream) between the replays."""
    p = CASES[name]()
    pl = sogl.Plan(p.n, p.row, p.col, p.val, p.b)
    fid0, l0, u0 = _factor_tiles(dev, pl, lambda f: dev.factor(f, sogl.EXEC_LEVELS))
    dev.release(fid0)
    fid = dev.load(pl, sogl.FP64)
    bp = pl.permute_rhs(p.b)
    for _ in range(3):
        dev.factor(fid, sogl.EXEC_GRAPH)
        assert np.array_equal(dev.factor_tiles(fid, 0, pl.info.n_l_tiles), l0)
        assert np.array_equal(dev.factor_tiles(fid, 1, pl.info.n_u_tiles), u0)
        dev.solve(fid, bp)
    dev.release(fid)


@pytest.mark.parametrize("chunk,wg", [(2, 0), (4, 0), (1, 1)])
def test_persistent_queue_settings(dev, chunk, wg):
    """Chunked fetches and one workgroup per CU change only the schedule: same bits."""
    p = W.diffusion_varcoef(20, 3, 5)
    pl = sogl.Plan(p.n, p.row, p.col, p.val, p.b)
    fid0, l0, u0 = _factor_tiles(dev, pl, lambda f: dev.factor(f, sogl.EXEC_PERSISTENT))
    dev.release(fid0)
    fid, l1, u1 = _factor_tiles(dev, pl, lambda f: dev.factor_config(f, sogl.EXEC_PERSISTENT, chunk, wg))
    dev.release(fid)
    assert np.array_equal(l0, l1) and np.array_equal(u0, u1)


def test_autotune_and_auto_mode(dev):
    p = W.convdiff2d(128)
    pl = sogl.Plan(p.n, p.row, p.col, p.val, p.b)
    fid0, l0, u0 = _factor_tiles(dev, pl, lambda f: dev.factor(f, sogl.EXEC_LEVELS))
    dev.release(fid0)
    fid = dev.load(pl, sogl.FP64)
    best, times = dev.autotune(fid, 1)
    assert 0 <= best < len(sogl.TUNE_CANDIDATES) and np.all(times > 0)
    assert times[best] == times.min()
    dev.factor(fid, sogl.EXEC_AUTO)
    assert np.array_equal(dev.factor_tiles(fid, 0, pl.info.n_l_tiles), l0)
    assert np.array_equal(dev.factor_tiles(fid, 1, pl.info.n_u_tiles), u0)
    dev.release(fid)


REUSE_CASES = {
    "lu": lambda: W.convdiff2d(128),
    "llt": lambda: W.symmetric_lower(W.diffusion_varcoef(16, 3, 21)),
    "c1": lambda: W.config("c1"),
    "c2": lambda: W.config("c2"),
}


@pytest.mark.parametrize("precision", [sogl.FP64, sogl.FP32, sogl.BF16])
@pytest.mark.parametrize("case", sorted(REUSE_CASES))
def test_pool_slot_reuse_keeps_bits(dev, precision, case):
    """Pool slot reuse (a result takes over the slot of a value whose readers are done,
    after waiting for them, and reads a slot's later values with coherent loads) leaves
    the factor bit for bit as a factor loaded without reuse gives it: the persistent
    executor at one and two workgroups per CU, the levels executor and its captured
    graph, each run twice (the inputs scattered into their reused slots again before
    every run), and the solve after them.  The pool is smaller; the per-op executor,
    which accumulates in the results' slots, is refused."""
    p = REUSE_CASES[case]()
    pl = sogl.Plan(p.n, p.row, p.col, p.val, p.b, symmetric=p.symmetric)
    n_u = 0 if p.symmetric else pl.info.n_u_tiles
    bp = pl.permute_rhs(p.b if p.b.ndim == 1 else p.b[:, 0])

    def tiles(fid):
        return dev.factor_tiles(fid, 0, pl.info.n_l_tiles), dev.factor_tiles(fid, 1, n_u) if n_u else None

    dev.set_pool_reuse(False)
    try:
        f0 = dev.load(pl, precision)
        dev.set_pool_reuse(True)
        fid = dev.load(pl, precision)
    finally:
        dev.set_pool_reuse(sogl.REUSE_AUTO)
    dev.set_masks(f0, sogl.MASKS_ON)
    dev.factor(f0, sogl.EXEC_LEVELS)
    l0, u0 = tiles(f0)
    x0 = dev.solve(f0, bp)
    s0 = dev.stats(f0)
    dev.release(f0)
    assert s0.pool_tiles == s0.value_tiles
    dev.set_masks(fid, sogl.MASKS_ON)
    st = dev.stats(fid)
    assert st.value_tiles == s0.value_tiles and st.pool_tiles < st.value_tiles, (st.pool_tiles, st.value_tiles)
    assert st.pool_bytes < s0.pool_bytes
    for mode, wg in ((sogl.EXEC_PERSISTENT, 1), (sogl.EXEC_PERSISTENT, 0), (sogl.EXEC_LEVELS, 0),
                     (sogl.EXEC_GRAPH, 0)):
        for rep in range(2):
            print(f"reuse {case} prec {precision} mode {mode} wg {wg} run {rep}", flush=True)
            dev.factor_config(fid, mode, 1, wg)
            l1, u1 = tiles(fid)
            assert np.array_equal(l1, l0), (mode, wg, rep)
            if n_u:
                assert np.array_equal(u1, u0), (mode, wg, rep)
    assert np.array_equal(dev.solve(fid, bp), x0)
    with pytest.raises(sogl.SoglError):
        dev.factor(fid, sogl.EXEC_PER_OP)
    dev.release(fid)


PERSISTENT_CASES = {
    "lu": lambda: W.convdiff2d(128),
    "llt": lambda: W.symmetric_lower(W.diffusion_varcoef(16, 3, 21)),
}


@pytest.mark.parametrize("precision", [sogl.FP64, sogl.FP32])
@pytest.mark.parametrize("case", sorted(PERSISTENT_CASES))
def test_one_workgroup_per_cu_kernel(dev, precision, case):
    """The persistent kernel built for one workgroup per CU (its FP64 GEMMs stage the
    next pair as well, transposed pairs of an L L^T factor included) and the one built
    for two, each with dynamic chaining on and off, give the factor bits of the levels
    executor, masks on; with chaining on, some tasks are claimed."""
    p = PERSISTENT_CASES[case]()
    pl = sogl.Plan(p.n, p.row, p.col, p.val, p.b, symmetric=p.symmetric)
    out, claimed = [], []
    runs = [(sogl.EXEC_LEVELS, 0, True)] + [(sogl.EXEC_PERSISTENT, wg, ch) for wg in (1, 0) for ch in (True, False)]
    for mode, wg, chain in runs:
        fid = dev.load(pl, precision)
        dev.set_masks(fid, sogl.MASKS_ON)
        dev.set_chaining(fid, chain)
        for _ in range(2):
            dev.factor_config(fid, mode, 1, wg)
            st = dev.stats(fid)
            # every task ran exactly once; a claimed task was dropped by its queue workgroup
            # or never taken from the queue
            assert st.tasks_run == st.n_tasks, (mode, wg, chain, st.tasks_run, st.n_tasks)
            assert st.dropped_tasks <= st.claimed_tasks
        claimed.append(st.claimed_tasks)
        n_u = 0 if p.symmetric else pl.info.n_u_tiles
        out.append((dev.factor_tiles(fid, 0, pl.info.n_l_tiles), dev.factor_tiles(fid, 1, n_u) if n_u else None))
        dev.release(fid)
    for k in range(1, len(runs)):
        assert np.array_equal(out[0][0], out[k][0]), runs[k]
        if out[0][1] is not None:
            assert np.array_equal(out[0][1], out[k][1]), runs[k]
    assert claimed[0] == 0 and claimed[2] == 0 and claimed[4] == 0  # levels, chaining off
    assert claimed[1] > 0 or claimed[3] > 0, claimed


@pytest.mark.parametrize("precision", [sogl.FP64, sogl.FP32])
@pytest.mark.parametrize("case", sorted(PERSISTENT_CASES))
def test_task_links_keep_bits(dev, precision, case):
    """Task links of the persistent executor: an lu task that runs its panel GEMM itself
    from the inverse it keeps in LDS (FP32 LU factors), after-links (the workgroup that
    made a critical tile runs the one-pair GEMM that reads it next, not queued), both
    and none: every task runs once, and the factor bits and x equal the levels
    executor's, at one and two workgroups per CU."""
    p = PERSISTENT_CASES[case]()
    pl = sogl.Plan(p.n, p.row, p.col, p.val, p.b, symmetric=p.symmetric)
    n_u = 0 if p.symmetric else pl.info.n_u_tiles
    bp = pl.permute_rhs(p.b)
    runs = [(0, sogl.EXEC_LEVELS, 0)] + [(links, sogl.EXEC_PERSISTENT, wg)
                                         for links in (sogl.LINK_PANEL, sogl.LINK_AFTER, 3) for wg in (1, 0)]
    out = []
    try:
        for links, mode, wg in runs:
            dev.set_task_links(links)
            fid = dev.load(pl, precision)
            for _ in range(2):
                dev.factor_config(fid, mode, 1, wg)
                st = dev.stats(fid)
                assert st.tasks_run == st.n_tasks or mode == sogl.EXEC_LEVELS, (links, wg, st.tasks_run, st.n_tasks)
            out.append((dev.factor_tiles(fid, 0, pl.info.n_l_tiles), dev.factor_tiles(fid, 1, n_u) if n_u else None,
                        dev.solve(fid, bp)))
            dev.release(fid)
    finally:
        dev.set_task_links(sogl.LINK_PANEL)
    for k in range(1, len(runs)):
        assert np.array_equal(out[0][0], out[k][0]), runs[k]
        if n_u:
            assert np.array_equal(out[0][1], out[k][1]), runs[k]
        assert np.array_equal(out[0][2], out[k][2]), runs[k]


def test_resident_factor_repeat_solves(dev):
    p = W.convdiff2d(128)
    pl = sogl.Plan(p.n, p.row, p.col, p.val, p.b)
    fid = dev.load(pl, sogl.FP64)
    dev.factor(fid)
    x1 = dev.solve(fid, pl.permute_rhs(p.b))
    b2 = W.uniform_pm1(99, p.n)
    x2 = dev.solve(fid, pl.permute_rhs(b2))
    dev.factor(fid)  # refactor from the resident inputs: same bits
    x3 = dev.solve(fid, pl.permute_rhs(p.b))
    dev.release(fid)
    assert np.array_equal(x1, x3)
    assert sogl.oracle_residual(p.n, p.row, p.col, p.val, pl.unpermute(x2), b2) < 1e-15


def block_factor(nbr, missing):
    """A block factor given directly as tiles: L block lower triangular with lower
    triangular diagonal tiles, U block upper triangular with upper triangular ones, every
    off-diagonal tile present except the two of row pair `missing` (L_2m+1,2m and
    U_2m,2m+1)."""
    lt, ut, tiles = [], [], []
    Lf = np.zeros((64 * nbr, 64 * nbr))
    Uf = np.zeros((64 * nbr, 64 * nbr))
    for i in range(nbr):
        for j in range(nbr):
            for which, full in ((0, Lf), (1, Uf)):
                if (which == 0 and j > i) or (which == 1 and j < i):
                    continue
                if missing is not None and i != j and min(i, j) == 2 * missing and max(i, j) == 2 * missing + 1:
                    continue
                t = rng.standard_normal((64, 64)) / 64
                if i == j:
                    t = np.tril(t) if which == 0 else np.triu(t)
                    t[np.diag_indices(64)] = 1.0 + rng.random(64)
                tiles.append(t)
                (lt if which == 0 else ut).append([i, j, len(tiles)])
                full[64 * i:64 * i + 64, 64 * j:64 * j + 64] = t
    return np.array(lt, np.int32), np.array(ut, np.int32), np.stack(tiles), Lf, Uf


@pytest.mark.parametrize("precision,tol", [(sogl.FP64, 1e-12), (sogl.FP32, 1e-5)])
@pytest.mark.parametrize("nbr,missing", [(3, None), (5, 1), (4, 0), (1, None)])
def test_solve_row_pairs(dev, precision, tol, nbr, missing):
    """The triangular solve takes block rows two at a time, with the off-diagonal block of
    each pair's 2 x 2 diagonal block inverse made by the factor.  On hand-built factors
    (no ops: the L and U tiles are inputs, their diagonal inverses made by inverse
    tasks): odd block-row counts (the last row alone), a pair without sub-diagonal tiles
    (no coupling block), one and 9 right-hand sides; x against dense triangular solves of
    the same (rounded) tiles."""
    import scipy.linalg as sl
    lt, ut, tiles, Lf, Uf = block_factor(nbr, missing)
    ids = np.arange(1, len(tiles) + 1, dtype=np.int32)
    fid = dev.load_ops(len(tiles) + 1, np.zeros((0, 8), np.int32), ids, tiles, lt, ut, nbr, precision)
    dev.factor(fid, sogl.EXEC_PERSISTENT)
    if precision == sogl.FP32:
        Lf, Uf = (m.astype(np.float32).astype(np.float64) for m in (Lf, Uf))
    for nrhs in (1, 9):
        b = rng.standard_normal((64 * nbr, nrhs))
        x = dev.solve(fid, b)
        want = sl.solve_triangular(Uf, sl.solve_triangular(Lf, b, lower=True), lower=False)
        assert np.abs(x - want).max() / np.abs(want).max() < tol, (nrhs, np.abs(x - want).max())
    dev.release(fid)


@pytest.mark.parametrize("nrhs", [3, 20, 70, 300])
def test_multi_rhs(dev, nrhs):
    """Few right-hand sides: the per-row dot-product kernel; from 16 on: block-row
    sums on MFMA in panels of 64 columns (70 = one full and one partial panel; 300 =
    a launch of 256 columns and one of 44)."""
    p = W.diffusion_varcoef(20, 3, 5)
    B = W.uniform_pm1(5, p.n * nrhs).reshape(nrhs, p.n).T.copy()
    pl = sogl.Plan(p.n, p.row, p.col, p.val, B[:, 0])
    fid = dev.load(pl, sogl.FP64)
    dev.factor(fid)
    Xp = dev.solve(fid, pl.permute_rhs(B))
    x0 = dev.solve(fid, pl.permute_rhs(B[:, 0]))
    if nrhs < 16:
        assert np.array_equal(Xp[:, 0], x0)  # columns are independent
    else:
        X17 = dev.solve(fid, pl.permute_rhs(B[:, :17]))
        assert np.array_equal(Xp[:, :17], X17)  # columns are independent on MFMA too
        assert np.linalg.norm(Xp[:, 0] - x0) / np.linalg.norm(x0) < 1e-14
    dev.release(fid)
    X = pl.unpermute(Xp)
    for j in range(nrhs):
        assert sogl.oracle_residual(p.n, p.row, p.col, p.val, X[:, j], B[:, j]) < 1e-15


def test_fp32_refined_many_rhs(dev):
    """FP32 factor, 40 right-hand sides refined in FP64 to 1e-12 each (MFMA block solves)."""
    p = W.convdiff2d(128)
    nrhs = 40
    B = W.uniform_pm1(21, p.n * nrhs).reshape(nrhs, p.n).T.copy()
    pl = sogl.Plan(p.n, p.row, p.col, p.val, B[:, 0])
    fid = dev.load(pl, sogl.FP32)
    dev.factor(fid)
    dev.attach_refine(fid, pl)
    X, it, res, rc = dev.solve_refined(fid, B, tol=1e-12, max_iter=20)
    dev.release(fid)
    assert rc == 0 and res <= 1e-12
    for j in range(nrhs):
        assert sogl.oracle_residual(p.n, p.row, p.col, p.val, X[:, j], B[:, j]) <= 1e-12


@pytest.mark.parametrize("precision", [sogl.FP32, sogl.BF16])
@pytest.mark.parametrize("name", ["c1", "convdiff2d_128", "diffusion3d_20"])
def test_low_precision_refined(dev, precision, name, golden):
    p = CASES[name]()
    pl = sogl.Plan(p.n, p.row, p.col, p.val, p.b)
    fid = dev.load(pl, precision)
    dev.factor(fid)
    dev.attach_refine(fid, pl)
    x, it, res, rc = dev.solve_refined(fid, p.b, tol=1e-12, max_iter=60)
    dev.release(fid)
    assert rc == 0, (it, res)
    assert sogl.oracle_residual(p.n, p.row, p.col, p.val, x, p.b) <= 1e-12
    # forward error is bounded by cond(A) x residual: cond(A) ~ 2e3 for these grids
    ref = golden(name)["xorig"]
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 1e-7


def test_repeated_entries_refined(dev, golden):
    """Where the COO repeats a position, the tiles keep one value (the one the
    reference's tile builder stores last); the FP64 refinement residual uses the same
    matrix, so FP32 + refinement converges to the reference's FP64 solution."""
    p = W.edge_repeated_entries(1200)
    ref = golden("repeated_1200")["xorig"]
    pl = sogl.Plan(p.n, p.row, p.col, p.val, p.b)
    fid = dev.load(pl, sogl.FP64)
    dev.factor(fid)
    x64 = pl.unpermute(dev.solve(fid, pl.permute_rhs(p.b)))
    dev.release(fid)
    assert np.linalg.norm(x64 - ref) / np.linalg.norm(ref) < 1e-13
    fid = dev.load(pl, sogl.FP32)
    dev.factor(fid)
    dev.attach_refine(fid, pl)
    x, it, res, rc = dev.solve_refined(fid, p.b, tol=1e-13, max_iter=40)
    dev.release(fid)
    assert rc == 0, (it, res)
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 1e-12


@pytest.mark.slow
def test_c2_full_size(dev, golden):
    """3D 48^3 (n=110,592, 1.08M unit ops) at its real size: FP64 drop-in vs the
    reference x, and FP32 factor + FP64 refinement to 1e-12."""
    g = golden("c2")
    p = W.config("c2")
    pl = sogl.Plan(p.n, p.row, p.col, p.val, p.b)
    fid = dev.load(pl, sogl.FP64)
    dev.factor(fid)
    x = pl.unpermute(dev.solve(fid, pl.permute_rhs(p.b)))
    dev.release(fid)
    idx = g["xorig_idx"]
    ref = g["xorig_sample"]
    assert np.linalg.norm(x[idx] - ref) / np.linalg.norm(ref) < 1e-12
    assert abs(np.linalg.norm(x) - float(g["xorig_norm"])) / float(g["xorig_norm"]) < 1e-12
    assert sogl.oracle_residual(p.n, p.row, p.col, p.val, x, p.b) < 1e-15
    fid = dev.load(pl, sogl.FP32)
    dev.factor(fid)
    dev.attach_refine(fid, pl)
    x32, it, res, rc = dev.solve_refined(fid, p.b, tol=1e-12, max_iter=30)
    dev.release(fid)
    assert rc == 0 and res <= 1e-12
    assert np.linalg.norm(x32[idx] - ref) / np.linalg.norm(ref) < 1e-9


def test_c5_full_size(dev, golden):
    """2D convection-diffusion 1000^2 (n=1,000,000, 3.44M unit ops) at its real size:
    FP64 vs the reference x (sample), FP32 factor + FP64 refinement to a backward error
    of 1e-10 in both forms (infinity norms and ||A||_F).  The FP32 factor's own solve
    has a backward error of about 4e-8 (x error 1e-6); one correction brings x to about
    1e-12 of the reference's (measured 1.2e-12), so the bar is 1e-11."""
    g = golden("c5")
    p = W.config("c5")
    pl = sogl.Plan(p.n, p.row, p.col, p.val, p.b)
    idx, ref = g["xorig_idx"], g["xorig_sample"]
    fid = dev.load(pl, sogl.FP64)
    dev.factor(fid)
    x = pl.unpermute(dev.solve(fid, pl.permute_rhs(p.b)))
    dev.release(fid)
    assert np.linalg.norm(x[idx] - ref) / np.linalg.norm(ref) < 1e-12
    assert abs(np.linalg.norm(x) - float(g["xorig_norm"])) / float(g["xorig_norm"]) < 1e-12
    assert sogl.oracle_residual(p.n, p.row, p.col, p.val, x, p.b) < 1e-16
    fid = dev.load(pl, sogl.FP32)
    dev.factor(fid)
    dev.attach_refine(fid, pl)
    x32, it, res, rc = dev.solve_refined(fid, p.b, tol=1e-10, max_iter=30)
    dev.release(fid)
    assert rc == 0 and res <= 1e-10 and it >= 1
    assert sogl.oracle_residual(p.n, p.row, p.col, p.val, x32, p.b) <= 1e-10
    assert sogl.backward_error_inf(p.n, p.row, p.col, p.val, x32, p.b) <= 1e-10
    assert np.linalg.norm(x32[idx] - ref) / np.linalg.norm(ref) < 1e-11


@pytest.mark.parametrize("name", ["c1", "sym_poisson2d_64"])
def test_drop_in_solveLU(dev, name, golden):
    import ctypes as C
    p = W.config("c1") if name == "c1" else SYM[name]()
    lib = sogl.load()
    x = np.empty(p.n)
    row, col = p.row.astype(np.int32), p.col.astype(np.int32)
    rc = lib.sogl_solveLU(p.n, p.nnz, 1 if p.symmetric else 0, row.ctypes.data_as(C.POINTER(C.c_int32)),
                          col.ctypes.data_as(C.POINTER(C.c_int32)),
                          p.val.ctypes.data_as(C.POINTER(C.c_double)),
                          p.b.ctypes.data_as(C.POINTER(C.c_double)), x.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc == 0
    r, c, v = W.full_matrix(p)
    assert sogl.oracle_residual(p.n, r, c, v, x, p.b) < 1e-15
    ref = golden(name)["xorig"]
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 1e-12


def test_refined_device_resident_matches_host(dev):
    p = W.convdiff2d(128)
    pl = sogl.Plan(p.n, p.row, p.col, p.val, p.b)
    fid = dev.load(pl, sogl.FP32)
    dev.factor(fid)
    dev.attach_refine(fid, pl)
    xh, it_h, res_h, rc = dev.solve_refined(fid, p.b, tol=1e-12, max_iter=30)
    d_b, d_x = dev.alloc(p.b.nbytes), dev.alloc(p.b.nbytes)
    dev.upload(d_b, p.b)
    it_d, res_d, rc_d = dev.solve_refined_device(fid, d_b, d_x, 1, 1e-12, 30)
    xd = dev.download(d_x, np.empty(p.n))
    dev.free(d_b)
    dev.free(d_x)
    dev.release(fid)
    assert rc == 0 and rc_d == 0 and it_h == it_d
    assert np.array_equal(xh, xd)


def test_serving_mix_matches_alone():
    """Requests served concurrently (3 slots, resident factors, batched RHS) give the
    same x as each request served alone; every x meets the FP64 residual bar."""
    import serve
    reqs = serve.make_stream(18, 2000.0, seed=5, max_side_2d=72, max_side_3d=20)
    m, mixed = serve.serve(reqs, slots=3)
    assert m["requests"] == 18 and m["resident_hits"] >= 1
    for r in mixed:
        alone = serve.make_stream(18, 2000.0, seed=5, max_side_2d=72, max_side_3d=20)[r.rid]
        _, (a,) = serve.serve([alone], slots=1)
        assert np.linalg.norm(a.x - r.x) <= 1e-13 * np.linalg.norm(a.x)
        prob = serve.matrix_problem(r.matrix)
        assert sogl.oracle_residual(prob.n, prob.row, prob.col, prob.val, r.x, r.b) < 1e-14


def test_c4_request_vs_reference(dev, golden):
    """One request of the serving stream (2D variable-coefficient diffusion, side 128,
    log-normal coefficients; its plan is pinned by hash in test_plan_golden) against the
    reference's own FP64 solution: the FP64 path to 1e-12, and the FP32 factor with FP64
    refinement to a 1e-12 backward error in both forms."""
    import serve
    p = serve.matrix_problem((2, 128, 117))
    ref = golden("c4_2d_128")["xorig"]
    pl = sogl.Plan(p.n, p.row, p.col, p.val, p.b)
    fid = dev.load(pl, sogl.FP64)
    dev.factor(fid)
    x = pl.unpermute(dev.solve(fid, pl.permute_rhs(p.b)))
    dev.release(fid)
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 1e-12
    fid = dev.load(pl, sogl.FP32)
    dev.factor(fid)
    dev.attach_refine(fid, pl)
    x, it, res, rc = dev.solve_refined(fid, p.b, tol=1e-12, max_iter=30)
    dev.release(fid)
    assert rc == 0 and res <= 1e-12, (rc, it, res)
    assert sogl.backward_error_inf(p.n, p.row, p.col, p.val, x, p.b) <= 1e-12
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 1e-9


@pytest.mark.slow
def test_serving_at_full_request_sizes():
    """The largest requests of the serving stream -- a 2D 512^2 and a 3D 64^3 matrix
    (262,144 unknowns each), and the 512^2 matrix again as a resident hit -- served
    concurrently with smaller ones (3 slots, pool slot reuse on, multi-RHS batching),
    give the x each gives served alone, and every x meets the FP64 residual bar."""
    import serve
    keys = [(2, 512, 201), (3, 64, 202), (2, 96, 203), (3, 24, 204), (2, 512, 201)]
    def stream():
        return [serve.Request(i, 0.02 * i, k, 1 + 3 * i % 8, 60_000 + i) for i, k in enumerate(keys)]
    print("serving: mixed stream", flush=True)
    m, mixed = serve.serve(stream(), slots=3)
    assert m["requests"] == len(keys) and m["resident_hits"] >= 1, m
    for r in mixed:
        print(f"serving: request {r.rid} {r.matrix} alone", flush=True)
        _, (a,) = serve.serve([stream()[r.rid]], slots=1)
        assert np.linalg.norm(a.x - r.x) <= 1e-13 * np.linalg.norm(a.x), r.rid
        prob = serve.matrix_problem(r.matrix)
        assert sogl.oracle_residual(prob.n, prob.row, prob.col, prob.val, r.x, r.b) < 1e-14, r.rid


@pytest.mark.parametrize("precision", [sogl.FP64, sogl.FP32])
def test_c3_shape_many_rhs(dev, precision):
    """The C3 shape (random band, half-width 32 + 4 random entries per row, N(0,1)
    values, 128 N(0,1) right-hand sides: the MFMA block-row solve) at n = 2048 against
    the oracle's factor + solve of the same plan (FP64, 1e-12 relative per column), and,
    for the FP32 factor, FP64 refinement of all 128 columns to a backward error of 1e-12
    in both forms."""
    p = W.random_band(2048, nrhs=128)
    pl = sogl.Plan(p.n, p.row, p.col, p.val, p.b[:, 0])
    bp = pl.permute_rhs(p.b)
    xo, _ = oracle.factor_solve(pl, bp, threads=8)
    fid = dev.load(pl, precision)
    dev.factor(fid, sogl.EXEC_AUTO)
    if precision == sogl.FP64:
        x = dev.solve(fid, bp)
        err = np.linalg.norm(x[: p.n] - xo[: p.n], axis=0) / np.linalg.norm(xo[: p.n], axis=0)
        assert err.max() < 1e-12, err.max()
    else:
        dev.attach_refine(fid, pl)
        x, it, res, rc = dev.solve_refined(fid, p.b, tol=1e-12, max_iter=30)
        assert rc == 0 and res <= 1e-12
        assert sogl.backward_error_inf(p.n, p.row, p.col, p.val, x, p.b) <= 1e-12
        xr = pl.unpermute(xo)
        err = np.linalg.norm(x - xr, axis=0) / np.linalg.norm(xr, axis=0)
        assert err.max() < 1e-9, err.max()
    dev.release(fid)
