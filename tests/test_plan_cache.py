"""Plan reuse by sparsity pattern (sogl_plan_create_like) and the pattern key.

A plan made from an earlier plan of the same pattern must equal, array for array,
the plan made from scratch (which the golden fixtures pin to the reference): the
ordering, both op lists, tile stages, input tile map and values, factor tile maps,
the permuted right-hand side and the permuted COO. Where the pattern differs, or
the value-dependent sortInBlock step reorders, the call plans from scratch.
"""
import numpy as np
import pytest

import sogl
import workloads as W

LIB = "libsogl_host.so"


def arrays(pl, b):
    no, ro = pl.perm()
    ids, bi, bj, tiles = pl.inputs()
    prow, pcol, pval = pl.perm_coo()
    return dict(perm=np.concatenate([no, ro]), ops=pl.ops(0), coarse=pl.ops(1),
                stages=np.concatenate(pl.tile_stages()), inputs=np.stack([ids, bi, bj], 1), tiles=tiles,
                l=np.stack(pl.factor_tiles(0), 1), u=np.stack(pl.factor_tiles(1), 1),
                bperm=pl.permute_rhs(b),
                prow=prow, pcol=pcol, pval=pval)


def assert_same(a, b):
    for k in a:
        assert a[k].shape == b[k].shape and np.array_equal(a[k], b[k]), k


def fresh_and_like(p_like, p_new):
    like = sogl.Plan(p_like.n, p_like.row, p_like.col, p_like.val, p_like.b, lib=LIB)
    got = sogl.Plan(p_new.n, p_new.row, p_new.col, p_new.val, p_new.b, lib=LIB, like=like)
    ref = sogl.Plan(p_new.n, p_new.row, p_new.col, p_new.val, p_new.b, lib=LIB)
    return like, got, ref


@pytest.mark.parametrize("dim,m", [(2, 40), (3, 20), (2, 100)])
def test_same_pattern_reuses_and_matches(dim, m):
    p1 = W.diffusion_varcoef(m, dim, seed=5)
    p2 = W.diffusion_varcoef(m, dim, seed=6)
    assert np.array_equal(p1.row, p2.row) and not np.array_equal(p1.val, p2.val)
    like, got, ref = fresh_and_like(p1, p2)
    assert got.reused and not ref.reused
    assert_same(arrays(got, p2.b), arrays(ref, p2.b))
    # the values are the new matrix's, not the earlier one's
    assert not np.array_equal(got.inputs()[3], like.inputs()[3])


def test_reuse_without_rhs_and_repeated_entries():
    p = W.edge_repeated_entries(1200)
    q = W.edge_repeated_entries(1200)
    q.val = q.val * 1.5
    like = sogl.Plan(p.n, p.row, p.col, p.val, None, lib=LIB)
    got = sogl.Plan(q.n, q.row, q.col, q.val, None, lib=LIB, like=like)
    ref = sogl.Plan(q.n, q.row, q.col, q.val, None, lib=LIB)
    assert got.reused
    assert_same(arrays(got, q.b), arrays(ref, q.b))


def test_other_pattern_plans_from_scratch():
    p1 = W.diffusion_varcoef(30, 2, seed=5)
    p2 = W.diffusion_varcoef(31, 2, seed=5)  # other n
    like, got, ref = fresh_and_like(p1, p2)
    assert not got.reused
    assert_same(arrays(got, p2.b), arrays(ref, p2.b))
    # same n, one entry moved
    p3 = W.diffusion_varcoef(30, 2, seed=6)
    k = int(np.nonzero(p3.row != p3.col)[0][7])
    p3.col = p3.col.copy()
    p3.col[k] = (p3.col[k] + 3) % p3.n
    like, got, ref = fresh_and_like(p1, p3)
    assert not got.reused
    assert_same(arrays(got, p3.b), arrays(ref, p3.b))


def test_tiny_diagonal_plans_from_scratch():
    """sortInBlock reorders blocks whose leading diagonal holds |d| <= 1e-8: the new
    matrix needs another ordering, so the earlier plan cannot be reused."""
    p1 = W.diffusion_varcoef(30, 2, seed=5)
    p2 = W.diffusion_varcoef(30, 2, seed=6)
    like = sogl.Plan(p1.n, p1.row, p1.col, p1.val, p1.b, lib=LIB)
    no, ro = like.perm()
    # an unknown that the ordering puts at position 5 of the first block
    old = int(no[5])
    p2.val = p2.val.copy()
    p2.val[(p2.row == old) & (p2.col == old)] = 1e-9
    got = sogl.Plan(p2.n, p2.row, p2.col, p2.val, p2.b, lib=LIB, like=like)
    ref = sogl.Plan(p2.n, p2.row, p2.col, p2.val, p2.b, lib=LIB)
    assert not got.reused
    assert not np.array_equal(ref.perm()[0], no)  # the block was reordered
    assert_same(arrays(got, p2.b), arrays(ref, p2.b))
    # and the other way round: a block-sorted plan is not reused either
    got2 = sogl.Plan(p1.n, p1.row, p1.col, p1.val, p1.b, lib=LIB, like=ref)
    assert not got2.reused
    assert_same(arrays(got2, p1.b), arrays(like, p1.b))


def test_zero_diagonal_workload_not_reused():
    p = W.edge_zero_diag(48)
    like = sogl.Plan(p.n, p.row, p.col, p.val, p.b, lib=LIB)
    got = sogl.Plan(p.n, p.row, p.col, p.val, p.b, lib=LIB, like=like)
    assert not got.reused


def test_pattern_hash():
    p1 = W.diffusion_varcoef(30, 2, seed=5)
    p2 = W.diffusion_varcoef(30, 2, seed=6)
    h1 = sogl.pattern_hash(p1.n, p1.row, p1.col, lib=LIB)
    assert h1 == sogl.pattern_hash(p2.n, p2.row, p2.col, lib=LIB)
    col = p1.col.copy()
    col[10] += 1
    assert h1 != sogl.pattern_hash(p1.n, p1.row, col, lib=LIB)
    assert h1 != sogl.pattern_hash(p1.n + 1, p1.row, p1.col, lib=LIB)
    # the order of the entries is part of the pattern (the ordering's ties follow it)
    assert h1 != sogl.pattern_hash(p1.n, p1.row[::-1], p1.col[::-1], lib=LIB)


def test_c2_pattern_reuse_matches():
    p1 = W.config("c2")
    p2 = W.laplace3d_dominant(48, seed=7, rhs_seed=1007)
    assert np.array_equal(p1.row, p2.row) and np.array_equal(p1.col, p2.col)
    like, got, ref = fresh_and_like(p1, p2)
    assert got.reused
    assert got.info.t_plan < 0.5 * (ref.info.t_plan + ref.info.t_reorder)
    assert_same(arrays(got, p2.b), arrays(ref, p2.b))


def test_general_matrix_never_reuses_a_symmetric_plan():
    """A general matrix whose entry sequence is exactly the mirrored COO of an L L^T
    plan (each off-diagonal entry followed by its mirror, solver.cpp:136-149) has the
    same pattern as that plan's internal matrix, but the L L^T factor ignores the
    upper triangle: create_like must plan it from scratch as a general (LU) plan."""
    import ctypes as C
    lower = W.symmetric_lower(W.diffusion_varcoef(20, 2, seed=31))
    like = sogl.Plan(lower.n, lower.row, lower.col, lower.val, lower.b, lib=LIB, symmetric=True)
    assert like.info.symmetric == 1
    rows, cols, vals = [], [], []
    for r, c, v in zip(lower.row, lower.col, lower.val):
        rows.append(r), cols.append(c), vals.append(v)
        if r != c:  # the mirror, with a different value: the matrix is not symmetric
            rows.append(c), cols.append(r), vals.append(0.5 * v)
    row = np.array(rows, np.int32)
    col = np.array(cols, np.int32)
    val = np.array(vals)
    lib = sogl.load(LIB)
    h = C.c_void_p()
    reused = C.c_int32(-1)
    P = C.POINTER
    rc = lib.sogl_plan_create_like(like.h, lower.n, val.size, row.ctypes.data_as(P(C.c_int32)),
                                   col.ctypes.data_as(P(C.c_int32)), val.ctypes.data_as(P(C.c_double)),
                                   lower.b.ctypes.data_as(P(C.c_double)), C.byref(h), C.byref(reused))
    assert rc == 0
    info = sogl.PlanInfo()
    lib.sogl_plan_get_info(h, C.byref(info))
    lib.sogl_plan_destroy(h)
    assert reused.value == 0
    assert info.symmetric == 0 and info.n_llt == 0 and info.n_lu > 0
