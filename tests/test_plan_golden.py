"""Host plan vs the reference's own output (fixtures made by tests/golden/make_golden.py
from the compiled reference): permutation, coarse and fine op lists, tile stages,
input tile map and factor tile maps must be bit-identical."""
import hashlib

import numpy as np
import pytest

import sogl
import workloads as W

SMALL = {
    "c1": lambda: W.config("c1"),
    "random_band_2048": lambda: W.random_band(2048, nrhs=1),
    "disconnected_40": lambda: W.edge_disconnected(40),
    "zerodiag_48": lambda: W.edge_zero_diag(48),
    "unsym_3000": lambda: W.edge_unsym_pattern(3000),
    "convdiff2d_128": lambda: W.convdiff2d(128),
    "diffusion3d_20": lambda: W.diffusion_varcoef(20, 3, 5),
    "repeated_1200": lambda: W.edge_repeated_entries(1200),
    "sym_poisson2d_64": lambda: W.symmetric_lower(W.poisson2d(64)),
    "sym_diffusion3d_16": lambda: W.symmetric_lower(W.diffusion_varcoef(16, 3, 21)),
    "sym_diffusion2d_100": lambda: W.symmetric_lower(W.diffusion_varcoef(100, 2, 22)),
}


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).astype("<i4").tobytes()).hexdigest()


def sha_f64(a):
    return hashlib.sha256(np.ascontiguousarray(a).astype("<f8").tobytes()).hexdigest()


def plan_arrays(p):
    b = p.b if p.b.ndim == 1 else p.b[:, 0]
    pl = sogl.Plan(p.n, p.row, p.col, p.val, b, lib="libsogl_host.so", symmetric=p.symmetric)
    no, ro = pl.perm()
    ls, st = pl.tile_stages()
    ids, bi, bj, tiles = pl.inputs()
    out = dict(perm=np.concatenate([no, ro]), ops=sogl.canonical_ops(pl.ops(0)),
               coarse_ops=sogl.canonical_ops(pl.ops(1)),
               tiles=np.concatenate([[pl.info.storage_count], ls, st]).astype(np.int32),
               inputs=np.stack([ids, bi, bj], 1), ltiles=np.stack(pl.factor_tiles(0), 1),
               utiles=np.stack(pl.factor_tiles(1), 1), bperm=pl.permute_rhs(b), input_vals=tiles)
    return pl, out


@pytest.mark.parametrize("name", list(SMALL))
def test_small_bit_exact(name, golden):
    g = golden(name)
    pl, a = plan_arrays(SMALL[name]())
    for k in ("perm", "ops", "coarse_ops", "tiles", "inputs", "ltiles", "utiles"):
        assert np.array_equal(a[k], g[k]), f"{name}: {k} differs from the reference"
    assert np.array_equal(a["bperm"], g["bperm"])
    # input tile values, bit for bit (with repeated COO positions: the same value kept)
    assert sha_f64(a["input_vals"]) == str(g["sha_input_vals"]), f"{name}: input tile values differ"


LARGE = {"c2": lambda: W.config("c2"), "c5": lambda: W.config("c5"),
         "sym_diffusion3d_48": lambda: W.symmetric_lower(W.diffusion_varcoef(48, 3, 23)),
         "c4_2d_128": lambda: W.diffusion_varcoef(128, 2, 117)}


@pytest.mark.parametrize("name", list(LARGE))
def test_large_hashes(name, golden):
    g = golden(name)
    pl, a = plan_arrays(LARGE[name]())
    assert pl.info.n_ops == int(g["n_ops"])
    assert pl.info.storage_count == int(g["storage"])
    for k in ("perm", "ops", "coarse_ops", "tiles", "inputs", "ltiles", "utiles"):
        assert sha(a[k]) == str(g["sha_" + k]), f"{name}: {k} hash differs from the reference"
    assert sha_f64(a["input_vals"]) == str(g["sha_input_vals"]), f"{name}: input tile values differ"


def test_log_lines_match_reference(golden):
    g = golden("c1")
    p = W.config("c1")
    pl = sogl.Plan(p.n, p.row, p.col, p.val, p.b, lib="libsogl_host.so")
    ref = [l for l in str(g["ref_stdout"]).splitlines() if l.startswith(("GGPS", "reduced", "structure"))]
    mine = [l for l in pl.log.splitlines() if l.startswith(("GGPS", "reduced", "structure"))]
    assert mine == ref


def test_size_config_closed_form():
    # blockRows = 2^(floor(log2(n/64)) + 1), L2 split 2 << (log2(blockRows) / 2)
    for n, br, br2 in [(64, 2, 2), (100, 2, 2), (4096, 128, 16), (110592, 2048, 64),
                       (262144, 8192, 128), (1000000, 16384, 256)]:
        k = n // 64
        assert br == 2 ** (int(np.floor(np.log2(k))) + 1)
        assert br2 == 2 << (int(np.log2(br)) // 2)


def test_plan_invariants():
    """Every op reads tiles made at an earlier stage; writers of a tile share a stage."""
    p = W.convdiff2d(64)
    pl = sogl.Plan(p.n, p.row, p.col, p.val, p.b, lib="libsogl_host.so")
    ops = pl.ops(0)
    src, src2, res, res2, stage = ops[:, 0], ops[:, 1], ops[:, 3], ops[:, 4], ops[:, 5]
    ids = pl.inputs(False)[0]
    made = np.zeros(pl.info.storage_count, np.int64)
    made[ids] = 1
    wstage = np.zeros(pl.info.storage_count, np.int64)
    for r, r2, s in zip(res, res2, stage):
        wstage[r] = s if wstage[r] == 0 else wstage[r]
        assert wstage[r] == s
        if r2 > 0:
            wstage[r2] = s
    for s1, s2, s in zip(src, src2, stage):
        for t in (s1, s2):
            if t > 0 and not made[t]:
                assert wstage[t] < s
    assert np.all(np.diff(stage) >= 0)


def test_symmetric_input_plans_llt():
    """A symmetric matrix (one triangle in the COO) is mirrored and planned as L L^T:
    llt / mult ops instead of lu / mul, no U factor tiles (U = L^T)."""
    p = W.symmetric_lower(W.poisson2d(16 * 4))
    pl = sogl.Plan(p.n, p.row, p.col, p.val, p.b, lib="libsogl_host.so", symmetric=True)
    ops = pl.ops(0)
    kinds = set(np.unique(ops[:, 2]).tolist())
    assert 10 in kinds and 11 in kinds and 1 not in kinds  # llt, mult; no lu
    assert pl.info.n_u_tiles == 0 and pl.info.n_l_tiles > 0
    assert "symmetric" in pl.log.splitlines()
    # the mirrored matrix enters the plan: its permuted COO has every entry twice
    r, c, v = pl.perm_coo()
    n_off = int((p.row != p.col).sum())
    assert r.size == p.nnz + n_off + (pl.N - p.n)
