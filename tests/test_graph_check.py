"""Host-side checks of the executor's task graph (no GPU needed).

Every plan's graph must pass sogl_graph_check (slot indices inside the pool, one
writer per tile, every operand made at an earlier dependency level and waited for),
and the graph transforms must keep the op counts they are built from: one GEMM pair
per mul/mulneg op, one folded op per copy/neg/sub that became part of another task.
Hand-built op lists with a planted mistake must be rejected with an error.
"""
import numpy as np
import pytest

import sogl
import workloads as W

CASES = {
    "c1": lambda: W.config("c1"),
    "disconnected_40": lambda: W.edge_disconnected(40),
    "zerodiag_48": lambda: W.edge_zero_diag(48),
    "unsym_3000": lambda: W.edge_unsym_pattern(3000),
    "convdiff2d_128": lambda: W.convdiff2d(128),
    "diffusion3d_20": lambda: W.diffusion_varcoef(20, 3, 5),
    "random_band_2048": lambda: W.random_band(2048),
    "sym_poisson2d_64": lambda: W.symmetric_lower(W.poisson2d(64)),
    "sym_diffusion3d_16": lambda: W.symmetric_lower(W.diffusion_varcoef(16, 3, 21)),
}


def pair_blocks(pl):
    """Row pairs (2p, 2p+1) whose sub-diagonal L tile L_2p+1,2p is present, and those
    whose U tile U_2p,2p+1 is (an L L^T factor: U = L^T, so the L tile's)."""
    bi, bj, _ = pl.factor_tiles(0)
    n_l = int(np.sum((bi % 2 == 1) & (bj == bi - 1)))
    if pl.info.symmetric:
        return n_l, n_l
    ui, uj, _ = pl.factor_tiles(1)
    return n_l, int(np.sum((ui % 2 == 0) & (uj == ui + 1)))


@pytest.mark.parametrize("name", list(CASES))
def test_plan_graphs_pass(name):
    p = CASES[name]()
    pl = sogl.Plan(p.n, p.row, p.col, p.val, p.b, symmetric=p.symmetric)
    info = pl.info
    n_xl, n_yu = pair_blocks(pl)
    for prec in (sogl.FP64, sogl.FP32):
        g = sogl.graph_check(pl, prec)
        # the solve's two-row blocks: two GEMM tasks (one pair each) per present
        # sub-diagonal tile of a row pair, for L and for U
        assert g["pair_tasks"] == 2 * (n_xl + n_yu)
        assert g["pairs"] == info.n_mul + info.n_mulneg + info.n_mult + g["pair_tasks"]
        # folded ops are sub ops (copy / neg / sub), never more than the plan has
        assert 0 <= g["folded"] <= info.n_sub
        assert g["inv_aliased"] <= info.n_lower_inv + info.n_upper_inv
        # L L^T: one transpose task per L tile and per diagonal inverse (U = L^T)
        extra = info.n_l_tiles + info.block_rows if info.symmetric else 0
        # every unit op is a task, one more pair of a GEMM task, folded or aliased
        assert g["tasks"] <= info.n_ops - g["folded"] - g["inv_aliased"] + extra + g["pair_tasks"]
        assert g["levels"] > 0
        assert g["value_tiles"] <= info.storage_count + 2 * info.block_rows + 1 + extra + g["pair_tasks"]
        # slot reuse (checked inside graph_check): never more tiles than values, and at
        # least the factor tiles, their diagonal inverses, the two-row blocks and the
        # zero tile
        kept = (2 * info.n_l_tiles if info.symmetric else info.n_l_tiles + info.n_u_tiles) + 2 * info.block_rows + 1
        assert kept - extra + n_xl + n_yu <= g["pool_tiles"] <= g["value_tiles"]


def test_c2_graph_passes():
    p = W.config("c2")
    pl = sogl.Plan(p.n, p.row, p.col, p.val, p.b)
    g = sogl.graph_check(pl, sogl.FP32)
    # 864 of the 1024 row pairs have both sub-diagonal tiles: 4 GEMM tasks each
    assert g["pair_tasks"] == 3456
    assert g["pairs"] == 1_000_386 + g["pair_tasks"]
    assert g["levels"] < pl.info.n_stages  # fused tasks: fewer dependency levels than stages
    # most lu tasks follow the Schur-complement GEMM of their tile in the same workgroup
    assert 0 < g["chained"] <= pl.info.n_lu
    # task links (built by the check, their head order validated): panel GEMMs run by
    # their lu task, and near-critical one-pair GEMMs run by the workgroup that made
    # their last operand
    assert g["absorbed_panels"] == 892 and g["after_links"] == 442
    # slot reuse: 224,341 values in 90,316 tiles (the factor tiles, their inverses, the
    # two-row blocks and the zero tile are 80,689; a slot freed fewer than 1024 levels
    # back is not reused)
    assert g["value_tiles"] == 224_341 and 80_689 < g["pool_tiles"] < 92_000


def test_c5_pool_fits_in_15_gb():
    """The 1M-unknown convection-diffusion factor in FP64: 1,109,386 values of 32 KiB
    (36.4 GB without slot reuse) in a pool of at most 15 GB; the reference peaks at
    13.4 GB of host memory on the same matrix."""
    p = W.config("c5")
    pl = sogl.Plan(p.n, p.row, p.col, p.val, p.b)
    g = sogl.graph_check(pl, sogl.FP64)
    assert g["value_tiles"] == 1_109_386
    assert g["pool_tiles"] * 4096 * 8 <= 15e9, g["pool_tiles"]


# ---- planted mistakes in hand-built op lists (ids 1, 2 are inputs)
MUL, SUB, LU = 8, 4, 1


def _check(ops, n_ids=8, block_rows=1, lt=None, ut=None):
    tiles = np.zeros((2, 64, 64))
    ids = np.array([1, 2], np.int32)
    lt = np.array(lt if lt is not None else [[0, 0, 3]], np.int32)
    ut = np.array(ut if ut is not None else [[0, 0, 4]], np.int32)
    return sogl.graph_check_ops(n_ids, np.array(ops, np.int32), ids, tiles, lt, ut, block_rows)


def test_valid_hand_built_graph():
    g = _check([[1, 0, LU, 3, 4, 2, 0, 0]])
    assert g["tasks"] >= 1


def test_rejects_read_before_write():
    # op 0 reads tile 5, which only op 1 writes
    with pytest.raises(sogl.SoglError, match="neither an input"):
        _check([[5, 2, MUL, 6, 0, 2, 0, 0], [1, 2, MUL, 5, 0, 3, 1, 0], [6, 0, LU, 3, 4, 4, 2, 0]])


def test_rejects_write_after_read():
    # tile 5 is read by op 1 and then accumulated into again by op 2
    with pytest.raises(sogl.SoglError, match="after it was read"):
        _check([[1, 2, MUL, 5, 0, 2, 0, 0], [5, 1, MUL, 6, 0, 3, 1, 0], [1, 1, MUL, 5, 0, 4, 2, 0],
                [6, 0, LU, 3, 4, 5, 3, 0]])


def test_rejects_two_writers():
    with pytest.raises(sogl.SoglError, match="more than one non-accumulating"):
        _check([[1, 0, LU, 3, 4, 2, 0, 0], [2, 0, LU, 3, 7, 3, 1, 0]])


def test_rejects_mixed_factor_kinds():
    LLT, MULT = 10, 11
    with pytest.raises(sogl.SoglError, match="both lu and llt"):
        _check([[1, 0, LU, 3, 4, 2, 0, 0], [2, 0, LLT, 5, 0, 2, 1, 0]])
    with pytest.raises(sogl.SoglError, match="mult ops belong"):
        _check([[1, 2, MULT, 5, 0, 2, 0, 0], [1, 0, LU, 3, 4, 2, 1, 0]])
    with pytest.raises(sogl.SoglError, match="no U tiles"):
        _check([[1, 0, LLT, 3, 0, 2, 0, 0]])  # _check passes a U tile list


def test_llt_hand_built_graph():
    LLT = 10
    g = sogl.graph_check_ops(8, np.array([[1, 0, LLT, 3, 0, 2, 0, 0]], np.int32), np.array([1, 2], np.int32),
                             np.zeros((2, 64, 64)), np.array([[0, 0, 3]], np.int32), np.zeros((0, 3), np.int32), 1)
    # the llt task (with its inverse) and transposes of L and of L^-1 for the solve
    assert g["tasks"] == 3
