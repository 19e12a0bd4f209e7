"""Synthetic workloads of the sparse block LU path, generated reproducibly.

Every random number comes from splitmix64 (Steele, Lea & Flood, "Fast
splittable pseudorandom number generators", OOPSLA 2014), written out here in
numpy so that any host regenerates byte-identical matrices:

    state_0  = mix(seed + 0x632BE59BD9B4E019)
    state_k  = state_0 + k * 0x9E3779B97F4A7C15        (k = 1, 2, ...)
    draw_k   = mix(state_k)
    mix(z)   = z ^= z>>30; z *= 0xBF58476D1CE4E5B9; z ^= z>>27;
               z *= 0x94D049BB133111EB; z ^= z>>31      (all mod 2**64)
    U[0,1)   = (draw >> 11) * 2**-53;  U[-1,1) = 2*U[0,1) - 1
    N(0,1)   = Box-Muller on consecutive pairs: sqrt(-2 ln(1-u1)) cos(2 pi u2)

Matrices are returned as COO sorted by (row, col) with 0-based indices and
unknowns numbered lexicographically (r = i*m + j in 2D, r = (i*m + j)*m + k in
3D), which is the entry order the planner's tile ids follow.  ``write_mtx``
writes them as Matrix Market ``coordinate real general`` with ``%.17g``.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_SEED_OFF = np.uint64(0x632BE59BD9B4E019)


def _mix(z: np.ndarray) -> np.ndarray:
    z = z.copy()
    z ^= z >> np.uint64(30)
    z *= np.uint64(0xBF58476D1CE4E5B9)
    z ^= z >> np.uint64(27)
    z *= np.uint64(0x94D049BB133111EB)
    z ^= z >> np.uint64(31)
    return z


def draws(seed: int, count: int, start: int = 0) -> np.ndarray:
    """splitmix64 outputs number start+1 .. start+count of stream `seed`."""
    with np.errstate(over="ignore"):
        base = _mix(np.array([np.uint64(seed) + _SEED_OFF], dtype=np.uint64))[0]
        k = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        return _mix(base + k * _GOLDEN)


def uniform01(seed: int, count: int, start: int = 0) -> np.ndarray:
    return (draws(seed, count, start) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def uniform_pm1(seed: int, count: int, start: int = 0) -> np.ndarray:
    return 2.0 * uniform01(seed, count, start) - 1.0


def normal(seed: int, count: int) -> np.ndarray:
    u = uniform01(seed, 2 * count)
    u1, u2 = u[0::2], u[1::2]
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * math.pi * u2)


@dataclass
class Problem:
    name: str
    n: int
    row: np.ndarray  # int32, sorted (row, col)
    col: np.ndarray
    val: np.ndarray  # float64
    b: np.ndarray    # float64, shape (n,) or (n, nrhs) column-major when written
    symmetric: bool = False  # the COO is the lower triangle of a symmetric matrix

    @property
    def nnz(self) -> int:
        return int(self.val.size)


def _sorted(rows, cols, vals):
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    key = rows * (int(cols.max()) + 1 if cols.size else 1) + cols
    order = np.argsort(key, kind="stable")
    return rows[order].astype(np.int32), cols[order].astype(np.int32), np.asarray(vals)[order]


def _grid_offsets(m: int, dim: int):
    """(offset, axis, sign) of the 2*dim nearest neighbours, by increasing offset."""
    strides = [m ** (dim - 1 - a) for a in range(dim)]
    nb = []
    for a, s in enumerate(strides):
        nb.append((-s, a, -1))
        nb.append((s, a, +1))
    nb.sort()
    return strides, nb


def _stencil_pattern(m: int, dim: int):
    """Sorted COO pattern of the 2*dim+1 point stencil; returns rows, cols, kind
    where kind = -1 for the diagonal or the index into the neighbour list."""
    n = m ** dim
    strides, nb = _grid_offsets(m, dim)
    idx = np.arange(n, dtype=np.int64)
    coord = [(idx // s) % m for s in strides]
    rows, cols, kinds = [], [], []
    entries = [(o, a, sg, k) for k, (o, a, sg) in enumerate(nb)]
    entries.append((0, -1, 0, -1))
    entries.sort()
    for o, a, sg, k in entries:
        if k < 0:
            ok = np.ones(n, dtype=bool)
        else:
            ok = coord[a] + sg >= 0 if sg < 0 else coord[a] + sg < m
        r = idx[ok]
        rows.append(r)
        cols.append(r + o)
        kinds.append(np.full(r.size, k, dtype=np.int64))
    rows = np.concatenate(rows)
    cols = np.concatenate(cols)
    kinds = np.concatenate(kinds)
    order = np.lexsort((cols, rows))
    return rows[order], cols[order], kinds[order], nb, coord


def poisson2d(m: int = 64, seed: int = 0) -> Problem:
    """5-point Poisson on an m x m grid: 4 on the diagonal, -1 to each neighbour;
    b ~ U[-1,1) from `seed`."""
    rows, cols, kinds, _, _ = _stencil_pattern(m, 2)
    vals = np.where(kinds < 0, 4.0, -1.0)
    n = m * m
    return Problem(f"poisson2d_{m}", n, rows.astype(np.int32), cols.astype(np.int32), vals,
                   uniform_pm1(seed, n))


def laplace3d_dominant(m: int = 48, seed: int = 1, rhs_seed: int = 1001) -> Problem:
    """3D 7-point Laplacian made nonsymmetric: off-diagonal -(1 + 0.2 u),
    u ~ U[-1,1) from `seed` drawn in (row, col) entry order; diagonal
    1.05 * sum |off| of its row (strictly diagonally dominant); b ~ U[-1,1)."""
    rows, cols, kinds, _, _ = _stencil_pattern(m, 3)
    off = kinds >= 0
    u = uniform_pm1(seed, int(off.sum()))
    vals = np.zeros(rows.size)
    vals[off] = -(1.0 + 0.2 * u)
    n = m ** 3
    rowsum = np.bincount(rows[off], weights=np.abs(vals[off]), minlength=n)
    vals[~off] = 1.05 * rowsum[rows[~off]]
    return Problem(f"laplace3d_{m}", n, rows.astype(np.int32), cols.astype(np.int32), vals,
                   uniform_pm1(rhs_seed, n))


def convdiff2d(m: int = 1000, pe: float = 10.0, vel=(1.0, 0.5), seed: int = 4) -> Problem:
    """-D lap(u) + v.grad(u) on an m x m grid, first-order upwind, scaled by h^2/D.
    Cell Peclet number Pe = max|v| h / D; unknown r = i*m + j with j along x
    (velocity vel[0]) and i along y (vel[1]).  Row: diagonal 4 + Pe_x + Pe_y,
    upwind neighbours -1 - Pe_x (west) / -1 - Pe_y (south), downwind -1."""
    vx, vy = vel
    vmax = max(abs(vx), abs(vy))
    pex, pey = pe * abs(vx) / vmax, pe * abs(vy) / vmax
    rows, cols, kinds, nb, _ = _stencil_pattern(m, 2)
    # nb sorted by offset: (-m: south, axis 0), (-1: west, axis 1), (+1: east), (+m: north)
    coef = {}
    for k, (o, a, sg) in enumerate(nb):
        upw = pey if a == 0 else pex
        vsign = (vy if a == 0 else vx)
        upwind = (sg < 0) == (vsign > 0)
        coef[k] = -1.0 - (upw if upwind else 0.0)
    vals = np.empty(rows.size)
    for k, c in coef.items():
        vals[kinds == k] = c
    vals[kinds < 0] = 4.0 + pex + pey
    n = m * m
    return Problem(f"convdiff2d_{m}", n, rows.astype(np.int32), cols.astype(np.int32), vals,
                   uniform_pm1(seed, n))


def random_band(n: int = 16384, half_width: int = 32, extra: int = 4, seed: int = 2,
                nrhs: int = 128, rhs_seed: int = 1002) -> Problem:
    """Band |i-j| <= half_width plus `extra` uniform random columns per row
    (stream `seed`), values N(0,1) (stream seed+100), duplicates merged (first
    kept), diagonal sum_{j!=i} |a_ij| + 1 so the pivot-free LU is stable;
    nrhs right-hand sides N(0,1) (column-major)."""
    i = np.repeat(np.arange(n, dtype=np.int64), 2 * half_width + 1)
    j = i + np.tile(np.arange(-half_width, half_width + 1, dtype=np.int64), n)
    keep = (j >= 0) & (j < n)
    i, j = i[keep], j[keep]
    ri = np.repeat(np.arange(n, dtype=np.int64), extra)
    rj = (draws(seed, n * extra) % np.uint64(n)).astype(np.int64)
    rows = np.concatenate([i, ri])
    cols = np.concatenate([j, rj])
    key = rows * n + cols
    _, first = np.unique(key, return_index=True)
    rows, cols = rows[first], cols[first]
    vals = normal(seed + 100, rows.size)
    diag = rows == cols
    vals[diag] = 0.0
    rowsum = np.bincount(rows, weights=np.abs(vals), minlength=n)
    vals[diag] = rowsum[rows[diag]] + 1.0
    b = normal(rhs_seed, n * nrhs).reshape(nrhs, n).T.copy()
    return Problem(f"random_band_{n}", n, rows.astype(np.int32), cols.astype(np.int32), vals, b)


def diffusion_varcoef(m: int, dim: int, seed: int, nrhs: int = 1, rhs_seed: int | None = None,
                      sigma: float = 1.0) -> Problem:
    """Variable-coefficient diffusion on an m^dim grid: node coefficients
    kappa ~ LogNormal(0, sigma) (stream `seed`), face weight = harmonic mean of
    its two nodes, Dirichlet boundary faces weigh kappa of the node; row =
    sum of face weights on the diagonal, -weight to each neighbour."""
    n = m ** dim
    kappa = np.exp(sigma * normal(seed, n))
    rows, cols, kinds, _, _ = _stencil_pattern(m, dim)
    off = kinds >= 0
    ka, kb = kappa[rows[off]], kappa[cols[off]]
    w = 2.0 * ka * kb / (ka + kb)
    vals = np.zeros(rows.size)
    vals[off] = -w
    inner = np.bincount(rows[off], weights=w, minlength=n)
    nnb = np.bincount(rows[off], minlength=n)
    diag = inner + (2 * dim - nnb) * kappa
    vals[~off] = diag[rows[~off]]
    rs = seed + 1000 if rhs_seed is None else rhs_seed
    b = uniform_pm1(rs, n * nrhs).reshape(nrhs, n).T.copy()
    if nrhs == 1:
        b = b[:, 0].copy()
    return Problem(f"diffusion{dim}d_{m}_s{seed}", n, rows.astype(np.int32),
                   cols.astype(np.int32), vals, b)


# Named configurations: the five workloads of the path (sizes as specified).
def config(name: str) -> Problem:
    if name == "c1":
        return poisson2d(64, seed=0)
    if name == "c2":
        return laplace3d_dominant(48, seed=1, rhs_seed=1001)
    if name == "c3":
        # the full n=262,144 is infeasible (near-dense fill after any level-set ordering);
        # the largest size run is n=16,384
        return random_band(16384, seed=2)
    if name == "c5":
        return convdiff2d(1000, pe=10.0, vel=(1.0, 0.5), seed=4)
    raise KeyError(name)


def write_mtx(prob: Problem, path: str, rhs: bool = True) -> None:
    """Matrix Market coordinate file (1-based, %.17g) and, with rhs, the
    `<base>_b.mtx` array file the reference CLI reads next to it."""
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real " + ("symmetric" if prob.symmetric else "general") + "\n")
        f.write(f"{prob.n} {prob.n} {prob.nnz}\n")
        body = np.char.add(
            np.char.add(np.char.add((prob.row + 1).astype(str), " "),
                        np.char.add((prob.col + 1).astype(str), " ")),
            np.array([f"{v:.17g}" for v in prob.val.tolist()]))
        f.write("\n".join(body.tolist()))
        f.write("\n")
    if rhs:
        base = path[: path.index(".mtx")]
        b = prob.b if prob.b.ndim == 1 else prob.b[:, 0]
        with open(base + "_b.mtx", "w") as f:
            f.write("%%MatrixMarket matrix array real general\n")
            f.write(f"{prob.n} 1\n")
            f.write("\n".join(f"{v:.17g}" for v in b.tolist()))
            f.write("\n")


if __name__ == "__main__":  # python workloads.py c1 out_dir
    import sys
    p = config(sys.argv[1])
    os.makedirs(sys.argv[2], exist_ok=True)
    write_mtx(p, os.path.join(sys.argv[2], f"{sys.argv[1]}.mtx"))
    print(p.name, p.n, p.nnz)


def edge_disconnected(m: int = 40, seed: int = 7) -> Problem:
    """Two independent 2D Poisson blocks plus isolated diagonal-only unknowns,
    numbered interleaved so that the level structure spans several components."""
    p = poisson2d(m, seed)
    n1 = p.n
    iso = 37
    n = 2 * n1 + iso
    perm = (np.arange(n, dtype=np.int64) * 7919) % n  # a fixed scatter of the unknowns
    rows = np.concatenate([p.row, p.row + n1, np.arange(2 * n1, n)])
    cols = np.concatenate([p.col, p.col + n1, np.arange(2 * n1, n)])
    vals = np.concatenate([p.val, p.val, np.full(iso, 3.0)])
    r, c, v = _sorted(perm[rows], perm[cols], vals)
    return Problem(f"disconnected_{m}", n, r, c, v, uniform_pm1(seed, n))


def edge_zero_diag(m: int = 48, seed: int = 9) -> Problem:
    """Nonsymmetric 2D stencil with some tiny/zero diagonal entries: exercises
    sortInBlock's swap network (explicit zeros are dropped by the reader, so a
    zero diagonal is an absent one)."""
    rows, cols, kinds, _, _ = _stencil_pattern(m, 2)
    u = uniform_pm1(seed, rows.size)
    vals = np.where(kinds < 0, 6.0 + u, -1.0 + 0.3 * u)
    d = kinds < 0
    sel = np.nonzero(d)[0]
    vals[sel[5::97]] = 1e-12
    keep = np.ones(rows.size, bool)
    keep[sel[11::131]] = False
    n = m * m
    return Problem(f"zerodiag_{m}", n, rows[keep].astype(np.int32), cols[keep].astype(np.int32),
                   vals[keep], uniform_pm1(seed + 1, n))


def edge_unsym_pattern(n: int = 3000, seed: int = 11) -> Problem:
    """Structurally nonsymmetric random pattern with a dominant diagonal."""
    k = 6
    ri = np.repeat(np.arange(n, dtype=np.int64), k)
    off = (draws(seed, n * k) % np.uint64(120)).astype(np.int64) - 40
    rj = np.clip(ri + off, 0, n - 1)
    rows = np.concatenate([np.arange(n), ri])
    cols = np.concatenate([np.arange(n), rj])
    key = rows * n + cols
    _, first = np.unique(key, return_index=True)
    rows, cols = rows[first], cols[first]
    vals = uniform_pm1(seed + 1, rows.size)
    diag = rows == cols
    rowsum = np.bincount(rows, weights=np.abs(vals), minlength=n)
    vals[diag] = rowsum[rows[diag]] + 1.0
    r, c, v = _sorted(rows, cols, vals)
    return Problem(f"unsym_{n}", n, r, c, v, uniform_pm1(seed + 2, n))


def edge_repeated_entries(n: int = 1200, seed: int = 13) -> Problem:
    """Band matrix whose COO repeats about one position in eight with another value,
    in a scattered (unsorted) order: the reader keeps repeats, and the tile builder's
    sort decides which value a tile keeps. Every value keeps the diagonal dominant."""
    k = 5
    ri = np.repeat(np.arange(n, dtype=np.int64), k)
    off = (draws(seed, n * k) % np.uint64(81)).astype(np.int64) - 40
    rj = np.clip(ri + off, 0, n - 1)
    rows = np.concatenate([np.arange(n), ri])
    cols = np.concatenate([np.arange(n), rj])
    key = rows * n + cols
    _, first = np.unique(key, return_index=True)
    rows, cols = rows[first], cols[first]
    rep = np.nonzero(draws(seed + 1, rows.size) % np.uint64(8) == 0)[0]
    rows = np.concatenate([rows, rows[rep], rows[rep[::3]]])
    cols = np.concatenate([cols, cols[rep], cols[rep[::3]]])
    vals = 0.5 * uniform_pm1(seed + 2, rows.size)
    diag = rows == cols
    vals[diag] = 4.0 * k + 2.0 + uniform01(seed + 3, int(diag.sum()))
    order = np.argsort(draws(seed + 4, rows.size), kind="stable")
    return Problem(f"repeated_{n}", n, rows[order].astype(np.int32), cols[order].astype(np.int32),
                   vals[order], uniform_pm1(seed + 5, n))


def symmetric_lower(p: Problem, name: str | None = None) -> Problem:
    """The lower triangle (row >= col) of a symmetric matrix, in the COO's order, as a
    Matrix Market `symmetric` file stores it."""
    keep = p.row >= p.col
    return Problem(name or f"{p.name}_sym", p.n, p.row[keep].copy(), p.col[keep].copy(), p.val[keep].copy(),
                   p.b, symmetric=True)


def full_matrix(p: Problem) -> tuple:
    """(row, col, val) of the whole matrix a symmetric Problem stands for (mirrored)."""
    if not p.symmetric:
        return p.row, p.col, p.val
    off = p.row != p.col
    return (np.concatenate([p.row, p.col[off]]), np.concatenate([p.col, p.row[off]]),
            np.concatenate([p.val, p.val[off]]))
