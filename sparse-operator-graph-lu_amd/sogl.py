"""ctypes binding of include/sogl.h for tests, bench and scripting.

The product is the C ABI in build/libsogl.so; this module only marshals numpy
arrays across it.  Loading never falls back to anything else: if the library
is missing the import of a device function raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "build")

OK = 0
FP64, FP32, BF16 = 0, 1, 2
EXEC_PERSISTENT, EXEC_LEVELS, EXEC_PER_OP, EXEC_GRAPH, EXEC_AUTO = 0, 1, 2, 3, 4
REUSE_OFF, REUSE_ON, REUSE_AUTO = 0, 1, 2
LINK_AFTER, LINK_PANEL = 1, 2  # sogl_ctx_set_task_links bits
MASKS_OFF, MASKS_ON, MASKS_AUTO = 0, 1, 2
TUNE_CANDIDATES = ("persistent chunk 1", "persistent chunk 2", "persistent chunk 4",
                   "persistent chunk 1, 1 workgroup/CU", "levels", "level graph")
OP_NAMES = {0: "inv", 1: "lu", 2: "lowerInv", 3: "upperInv", 4: "sub", 5: "add", 6: "neg",
            7: "copy", 8: "mul", 9: "mulneg", 10: "llt", 11: "mult", 12: "noop"}


class SoglError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"sogl error {code}: {msg}")
        self.code = code


class _Op(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("src", "src2", "op", "result", "result2", "stage", "seq",
                                         "group")]


class PlanInfo(C.Structure):
    _fields_ = [("m_size", C.c_int32), ("block_rows", C.c_int32), ("block_rows_l2", C.c_int32),
                ("block_size_l2", C.c_int32), ("storage_count", C.c_int32),
                ("n_stages", C.c_int32), ("gps_levels", C.c_int32), ("gps_bandwidth", C.c_int32),
                ("n_ops", C.c_int64), ("n_coarse_ops", C.c_int64), ("n_inputs", C.c_int64),
                ("n_l_tiles", C.c_int64), ("n_u_tiles", C.c_int64), ("n_mul", C.c_int64),
                ("n_mulneg", C.c_int64), ("n_sub", C.c_int64), ("n_lu", C.c_int64),
                ("n_lower_inv", C.c_int64), ("n_upper_inv", C.c_int64),
                ("t_reorder", C.c_double), ("t_plan", C.c_double),
                ("n_llt", C.c_int64), ("n_mult", C.c_int64), ("symmetric", C.c_int32), ("reserved", C.c_int32)]


class FactorDesc(C.Structure):
    _fields_ = [("n_ids", C.c_int32), ("block_rows", C.c_int32), ("ops", C.c_void_p),
                ("n_ops", C.c_int64), ("input_ids", C.c_void_p), ("input_tiles", C.c_void_p),
                ("n_inputs", C.c_int64), ("l_bi", C.c_void_p), ("l_bj", C.c_void_p),
                ("l_ids", C.c_void_p), ("n_l", C.c_int64), ("u_bi", C.c_void_p),
                ("u_bj", C.c_void_p), ("u_ids", C.c_void_p), ("n_u", C.c_int64),
                ("precision", C.c_int32)]


class FactorStats(C.Structure):
    _fields_ = [("factor_ms", C.c_double), ("exec_ms", C.c_double), ("solve_ms", C.c_double),
                ("pool_bytes", C.c_int64),
                ("n_tasks", C.c_int64), ("n_levels", C.c_int32), ("growth_tiles", C.c_int32),
                ("max_growth", C.c_double), ("zero_blocks", C.c_double), ("block_masks", C.c_int32),
                ("claimed_tasks", C.c_int32), ("dropped_tasks", C.c_int32), ("tasks_run", C.c_int32),
                ("pool_tiles", C.c_int64), ("value_tiles", C.c_int64), ("uinv_out_of_tol", C.c_int32)]


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


_libs: dict[str, C.CDLL] = {}


def load(name: str = "libsogl.so") -> C.CDLL:
    """Load build/<name>; raises if it has not been built."""
    if name in _libs:
        return _libs[name]
    path = os.path.join(BUILD, name)
    if not os.path.exists(path):
        raise OSError(f"{path} is not built (run __graft_entry__.build())")
    lib = C.CDLL(path)
    i32, i64, dbl, vp = C.c_int32, C.c_int64, C.c_double, C.c_void_p
    P = C.POINTER
    sig = {
        "sogl_plan_create": (C.c_int, [i32, i64, i32, P(i32), P(i32), P(dbl), P(dbl), i32, P(vp)]),
        "sogl_plan_create_like": (C.c_int, [vp, i32, i64, P(i32), P(i32), P(dbl), P(dbl), P(vp), P(i32)]),
        "sogl_pattern_hash": (C.c_int, [i32, i64, P(i32), P(i32), P(C.c_uint64)]),
        "sogl_plan_destroy": (None, [vp]),
        "sogl_plan_log": (C.c_char_p, [vp]),
        "sogl_plan_get_info": (C.c_int, [vp, P(PlanInfo)]),
        "sogl_plan_get_perm": (C.c_int, [vp, P(i32), P(i32)]),
        "sogl_plan_get_ops": (C.c_int, [vp, i32, P(_Op)]),
        "sogl_plan_get_tile_stages": (C.c_int, [vp, P(i32), P(i32)]),
        "sogl_plan_get_inputs": (C.c_int, [vp, P(i32), P(i32), P(i32), P(dbl)]),
        "sogl_plan_get_factor_tiles": (C.c_int, [vp, i32, P(i32), P(i32), P(i32)]),
        "sogl_plan_permute_rhs": (C.c_int, [vp, P(dbl), i32, i64, P(dbl)]),
        "sogl_plan_unpermute": (C.c_int, [vp, P(dbl), i32, P(dbl), i64]),
        "sogl_plan_perm_nnz": (i64, [vp]),
        "sogl_plan_get_perm_coo": (C.c_int, [vp, P(i32), P(i32), P(dbl)]),
        "sogl_mtx_read": (C.c_int, [C.c_char_p, P(vp)]),
        "sogl_mtx_read_rhs": (C.c_int, [vp, C.c_char_p]),
        "sogl_mtx_info": (C.c_int, [vp, P(i32), P(i64), P(i32)]),
        "sogl_mtx_arrays": (C.c_int, [vp, P(P(i32)), P(P(i32)), P(P(dbl)), P(P(dbl))]),
        "sogl_mtx_free": (None, [vp]),
        "sogl_last_error": (C.c_char_p, []),
    }
    if name == "libsogl.so":
        sig.update({
            "sogl_ctx_create": (C.c_int, [i32, P(vp)]),
            "sogl_ctx_destroy": (None, [vp]),
            "sogl_ctx_error": (C.c_char_p, [vp]),
            "sogl_ctx_mem_info": (C.c_int, [vp, P(C.c_size_t), P(C.c_size_t)]),
            "sogl_ctx_sync": (C.c_int, [vp]),
            "sogl_factor_load_plan": (C.c_int, [vp, vp, i32, P(i32)]),
            "sogl_factor_load": (C.c_int, [vp, P(FactorDesc), P(i32)]),
            "sogl_factor_run": (C.c_int, [vp, i32, i32]),
            "sogl_factor_run_config": (C.c_int, [vp, i32, i32, i32, i32]),
            "sogl_factor_autotune": (C.c_int, [vp, i32, i32, P(dbl), P(i32)]),
            "sogl_solve": (C.c_int, [vp, i32, P(dbl), P(dbl), P(dbl), i32]),
            "sogl_solve_device": (C.c_int, [vp, i32, vp, vp, i32]),
            "sogl_solve_device_noalloc": (C.c_int, [vp, i32, vp, vp, i32]),
            "sogl_factor_release": (C.c_int, [vp, i32]),
            "sogl_factor_get_stats": (C.c_int, [vp, i32, P(FactorStats)]),
            "sogl_factor_get_tiles": (C.c_int, [vp, i32, i32, P(dbl)]),
        "sogl_factor_set_masks": (C.c_int, [vp, i32, i32]),
            "sogl_factor_set_chaining": (C.c_int, [vp, i32, i32]),
            "sogl_ctx_set_pool_reuse": (C.c_int, [vp, i32]),
            "sogl_ctx_set_task_links": (C.c_int, [vp, i32]),
            "sogl_factor_get_uinv_check": (C.c_int, [vp, i32, P(i32), P(dbl)]),
            "sogl_refine_attach": (C.c_int, [vp, i32, vp]),
            "sogl_solve_refined": (C.c_int, [vp, i32, P(dbl), P(dbl), i32, dbl, i32, P(i32),
                                             P(dbl)]),
            "sogl_solve_refined_device": (C.c_int, [vp, i32, vp, vp, i32, dbl, i32, P(i32), P(dbl)]),
            "sogl_device_alloc": (C.c_int, [vp, C.c_size_t, P(vp)]),
            "sogl_device_free": (C.c_int, [vp, vp]),
            "sogl_device_copy": (C.c_int, [vp, vp, vp, C.c_size_t, i32]),
            "sogl_factor_set_trace": (C.c_int, [vp, i32, i32]),
            "sogl_factor_get_trace": (C.c_int, [vp, i32, P(i64), P(i64)]),
            "sogl_factor_get_graph": (C.c_int, [vp, i32, P(i64), P(i64), P(i32), P(i32), P(i32)]),
            "sogl_graph_check": (C.c_int, [P(FactorDesc), P(i64)]),
            "sogl_graph_check_plan": (C.c_int, [vp, i32, P(i64)]),
            "sogl_solveLU": (C.c_int, [i32, i64, i32, P(i32), P(i32), P(dbl), P(dbl), P(dbl)]),
        })
    for fn, (res, args) in sig.items():
        f = getattr(lib, fn)
        f.restype = res
        f.argtypes = args
    _libs[name] = lib
    return lib


def exported_symbols(header: str | None = None) -> list[str]:
    """Function names declared in include/sogl.h."""
    import re
    header = header or os.path.join(HERE, "..", "include", "sogl.h")
    text = open(header).read()
    return sorted(set(re.findall(r"\b(sogl_[a-z0-9_]+)\s*\(", text)))


def ops_array(ops) -> np.ndarray:
    """(n, 8) int32 in field order src src2 op result result2 stage seq group."""
    return np.ctypeslib.as_array(ops).view(np.int32).reshape(-1, 8) if len(ops) else \
        np.zeros((0, 8), np.int32)


def pattern_hash(n: int, row, col, lib: str = "libsogl.so") -> int:
    """64-bit key of a sparsity pattern (n and the row/column sequence)."""
    row = np.ascontiguousarray(row, dtype=np.int32)
    col = np.ascontiguousarray(col, dtype=np.int32)
    h = C.c_uint64()
    rc = load(lib).sogl_pattern_hash(n, row.size, _ptr(row, C.c_int32), _ptr(col, C.c_int32), C.byref(h))
    if rc != OK:
        raise SoglError(rc, load(lib).sogl_last_error().decode())
    return h.value


class Plan:
    """Host plan of one matrix: GPS ordering + staged 64x64 unit-op list."""

    def __init__(self, n: int, row, col, val, b=None, verbose=False, lib: str = "libsogl.so",
                 like: "Plan | None" = None, symmetric: bool = False):
        """like: an earlier plan; where the matrix has its sparsity pattern, its ordering
        and op list are reused (same result, see sogl_plan_create_like); self.reused
        says whether they were.  symmetric: the COO holds one triangle of a symmetric
        matrix (a Matrix Market `symmetric` file); it is mirrored and factored as L L^T."""
        self.lib = load(lib)
        row = np.ascontiguousarray(row, dtype=np.int32)
        col = np.ascontiguousarray(col, dtype=np.int32)
        val = np.ascontiguousarray(val, dtype=np.float64)
        bb = None if b is None else np.ascontiguousarray(b, dtype=np.float64)
        h = C.c_void_p()
        bptr = None if bb is None else _ptr(bb, C.c_double)
        reused = C.c_int32(0)
        if like is not None and not symmetric:
            rc = self.lib.sogl_plan_create_like(like.h, n, val.size, _ptr(row, C.c_int32), _ptr(col, C.c_int32),
                                                _ptr(val, C.c_double), bptr, C.byref(h), C.byref(reused))
        else:
            rc = self.lib.sogl_plan_create(n, val.size, 1 if symmetric else 0, _ptr(row, C.c_int32),
                                           _ptr(col, C.c_int32),
                                           _ptr(val, C.c_double), bptr, 1 if verbose else 0, C.byref(h))
        if rc != OK:
            raise SoglError(rc, self.lib.sogl_last_error().decode())
        self.reused = bool(reused.value)
        self.h = h
        self.n = n
        self.info = PlanInfo()
        self.lib.sogl_plan_get_info(self.h, C.byref(self.info))

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self.lib.sogl_plan_destroy(h)
            self.h = None

    @property
    def N(self) -> int:
        return self.info.block_rows * 64

    @property
    def log(self) -> str:
        return self.lib.sogl_plan_log(self.h).decode()

    def perm(self):
        no = np.empty(self.n, np.int32)
        ro = np.empty(self.n, np.int32)
        self.lib.sogl_plan_get_perm(self.h, _ptr(no, C.c_int32), _ptr(ro, C.c_int32))
        return no, ro

    def ops(self, level: int = 0) -> np.ndarray:
        cnt = self.info.n_ops if level == 0 else self.info.n_coarse_ops
        buf = (_Op * max(cnt, 1))()
        self.lib.sogl_plan_get_ops(self.h, level, buf)
        return np.frombuffer(buf, dtype=np.int32).reshape(-1, 8)[:cnt].copy()

    def tile_stages(self):
        S = self.info.storage_count
        ls = np.empty(S, np.int32)
        st = np.empty(S, np.int32)
        self.lib.sogl_plan_get_tile_stages(self.h, _ptr(ls, C.c_int32), _ptr(st, C.c_int32))
        return ls, st

    def inputs(self, with_tiles=True):
        T = self.info.n_inputs
        ids, bi, bj = (np.empty(T, np.int32) for _ in range(3))
        tiles = np.empty((T, 64, 64)) if with_tiles else None
        self.lib.sogl_plan_get_inputs(self.h, _ptr(ids, C.c_int32), _ptr(bi, C.c_int32),
                                      _ptr(bj, C.c_int32),
                                      None if tiles is None else _ptr(tiles, C.c_double))
        return ids, bi, bj, tiles

    def factor_tiles(self, which: int):
        cnt = self.info.n_l_tiles if which == 0 else self.info.n_u_tiles
        bi, bj, ids = (np.empty(cnt, np.int32) for _ in range(3))
        self.lib.sogl_plan_get_factor_tiles(self.h, which, _ptr(bi, C.c_int32),
                                            _ptr(bj, C.c_int32), _ptr(ids, C.c_int32))
        return bi, bj, ids

    def permute_rhs(self, b: np.ndarray) -> np.ndarray:
        b2 = np.asfortranarray(b.reshape(self.n, -1), dtype=np.float64)
        nrhs = b2.shape[1]
        out = np.empty((nrhs, self.N), np.float64)
        self.lib.sogl_plan_permute_rhs(self.h, _ptr(b2, C.c_double), nrhs, self.n,
                                       _ptr(out, C.c_double))
        return out.T if b.ndim == 2 else out[0]

    def unpermute(self, xp: np.ndarray) -> np.ndarray:
        x2 = np.asfortranarray(xp.reshape(self.N, -1), dtype=np.float64)
        nrhs = x2.shape[1]
        out = np.empty((nrhs, self.n), np.float64)
        self.lib.sogl_plan_unpermute(self.h, _ptr(x2, C.c_double), nrhs, _ptr(out, C.c_double),
                                     self.n)
        return out.T if xp.ndim == 2 else out[0]

    def perm_coo(self):
        m = self.lib.sogl_plan_perm_nnz(self.h)
        r, c = np.empty(m, np.int32), np.empty(m, np.int32)
        v = np.empty(m, np.float64)
        self.lib.sogl_plan_get_perm_coo(self.h, _ptr(r, C.c_int32), _ptr(c, C.c_int32),
                                        _ptr(v, C.c_double))
        return r, c, v


def make_desc(n_ids: int, ops, input_ids, input_tiles, ltiles, utiles, block_rows: int,
              precision: int = FP64):
    """A sogl_factor_desc over numpy copies of a raw op list (rows: src src2 op result
    result2 stage seq group); ltiles / utiles: (k, 3) block row, block col, id.
    Returns (desc, arrays that must stay alive while desc is used)."""
    ops = np.ascontiguousarray(ops, np.int32).reshape(-1, 8)
    ids = np.ascontiguousarray(input_ids, np.int32)
    tiles = np.ascontiguousarray(input_tiles, np.float64)
    lt = np.ascontiguousarray(ltiles, np.int32).reshape(-1, 3)
    ut = np.ascontiguousarray(utiles, np.int32).reshape(-1, 3)
    cols = [np.ascontiguousarray(lt[:, k]) for k in range(3)] + \
           [np.ascontiguousarray(ut[:, k]) for k in range(3)]
    keep = [ops, ids, tiles, lt, ut] + cols
    d = FactorDesc(n_ids, block_rows, ops.ctypes.data, ops.shape[0], ids.ctypes.data,
                   tiles.ctypes.data, ids.size, cols[0].ctypes.data, cols[1].ctypes.data,
                   cols[2].ctypes.data, lt.shape[0], cols[3].ctypes.data, cols[4].ctypes.data,
                   cols[5].ctypes.data, ut.shape[0], precision)
    return d, keep


GRAPH_INFO = ("tasks", "value_tiles", "levels", "pairs", "deps", "folded", "inv_aliased", "chained", "pool_tiles",
              "pair_tasks", "after_links", "absorbed_panels")


def graph_check_ops(n_ids: int, ops, input_ids, input_tiles, ltiles, utiles, block_rows: int,
                    precision: int = FP64) -> dict:
    """graph_check for a raw op list (see make_desc); raises SoglError on a bad graph."""
    lib = load("libsogl.so")
    d, keep = make_desc(n_ids, ops, input_ids, input_tiles, ltiles, utiles, block_rows, precision)
    info = np.zeros(len(GRAPH_INFO), np.int64)
    rc = lib.sogl_graph_check(C.byref(d), _ptr(info, C.c_int64))
    del keep
    if rc != OK:
        raise SoglError(rc, lib.sogl_last_error().decode())
    return dict(zip(GRAPH_INFO, (int(v) for v in info)))


def graph_check(plan: "Plan", precision: int = FP64) -> dict:
    """Build and check the executor's task graph on the host (no GPU); raises on error."""
    lib = load("libsogl.so")
    info = np.zeros(len(GRAPH_INFO), np.int64)
    rc = lib.sogl_graph_check_plan(plan.h, precision, _ptr(info, C.c_int64))
    if rc != OK:
        raise SoglError(rc, lib.sogl_last_error().decode())
    return dict(zip(GRAPH_INFO, (int(v) for v in info)))


def canonical_ops(ops: np.ndarray) -> np.ndarray:
    """Reorder sogl_op fields into the fixture layout
    stage group seq src src2 op result result2."""
    return ops[:, [5, 7, 6, 0, 1, 2, 3, 4]].copy()


class Device:
    """A device context (one GPU) holding resident factors."""

    def __init__(self, device: int = 0):
        self.lib = load("libsogl.so")
        h = C.c_void_p()
        rc = self.lib.sogl_ctx_create(device, C.byref(h))
        if rc != OK:
            raise SoglError(rc, self.lib.sogl_last_error().decode())
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.sogl_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc != OK:
            raise SoglError(rc, self.lib.sogl_ctx_error(self.h).decode())

    def mem_info(self):
        f, t = C.c_size_t(), C.c_size_t()
        self._check(self.lib.sogl_ctx_mem_info(self.h, C.byref(f), C.byref(t)))
        return f.value, t.value

    def load(self, plan: Plan, precision: int = FP64) -> int:
        fid = C.c_int32()
        self._check(self.lib.sogl_factor_load_plan(self.h, plan.h, precision, C.byref(fid)))
        return fid.value

    def load_ops(self, n_ids: int, ops: np.ndarray, input_ids, input_tiles, ltiles, utiles,
                 block_rows: int, precision: int = FP64) -> int:
        """Load a raw op list (rows: src src2 op result result2 stage seq group).
        ltiles / utiles: (k, 3) arrays of block row, block col, id."""
        d, keep = make_desc(n_ids, ops, input_ids, input_tiles, ltiles, utiles, block_rows, precision)
        fid = C.c_int32()
        self._check(self.lib.sogl_factor_load(self.h, C.byref(d), C.byref(fid)))
        return fid.value

    def factor_config(self, fid: int, mode: int = EXEC_PERSISTENT, chunk: int = 1, wg_per_cu: int = 0):
        self._check(self.lib.sogl_factor_run_config(self.h, fid, mode, chunk, wg_per_cu))

    def autotune(self, fid: int, trials: int = 1):
        """(index of the fastest configuration in TUNE_CANDIDATES, executor ms per candidate)"""
        t = np.zeros(len(TUNE_CANDIDATES))
        best = C.c_int32()
        self._check(self.lib.sogl_factor_autotune(self.h, fid, trials, _ptr(t, C.c_double), C.byref(best)))
        return best.value, t

    def factor(self, fid: int, mode: int = EXEC_PERSISTENT):
        self._check(self.lib.sogl_factor_run(self.h, fid, mode))

    def set_chaining(self, fid: int, claims):
        """Dynamic chaining of the persistent executor: at most `claims` tasks claimed per
        finished task (0 / False: off; True: the default, 2)."""
        n = 2 if claims is True else int(claims)
        self._check(self.lib.sogl_factor_set_chaining(self.h, fid, n))

    def set_task_links(self, links: int):
        """Task links (LINK_* mask) for the factors loaded from now on (sogl_ctx_set_task_links)."""
        self._check(self.lib.sogl_ctx_set_task_links(self.h, int(links)))

    def set_pool_reuse(self, mode):
        """Pool slot reuse for the factors loaded from now on: True / REUSE_ON, False /
        REUSE_OFF (the per-op executor needs it off), or REUSE_AUTO (the default)."""
        m = REUSE_ON if mode is True else REUSE_OFF if mode is False else int(mode)
        self._check(self.lib.sogl_ctx_set_pool_reuse(self.h, m))

    def uinv_check(self, fid: int):
        """The last run's upper-inverse check: (number of block rows out of tolerance,
        [(block row, 'diag' | 'middle', min, max), ...] for up to 16 of them)."""
        n = C.c_int32()
        d = np.zeros(64)
        self._check(self.lib.sogl_factor_get_uinv_check(self.h, fid, C.byref(n), _ptr(d, C.c_double)))
        rows = [(int(d[4 * q]), "diag" if d[4 * q + 1] == 0 else "middle", float(d[4 * q + 2]), float(d[4 * q + 3]))
                for q in range(min(n.value, 16))]
        return n.value, rows

    def set_masks(self, fid: int, mode: int):
        """16x16 block-mask policy: MASKS_OFF, MASKS_ON or MASKS_AUTO (the default)."""
        self._check(self.lib.sogl_factor_set_masks(self.h, fid, mode))

    def solve(self, fid: int, bp: np.ndarray, want_y=False):
        b2 = np.asfortranarray(bp.reshape(bp.shape[0], -1), dtype=np.float64)
        nrhs = b2.shape[1]
        x = np.empty_like(b2, order="F")
        y = np.empty_like(b2, order="F") if want_y else None
        self._check(self.lib.sogl_solve(self.h, fid, _ptr(b2, C.c_double), _ptr(x, C.c_double),
                                        None if y is None else _ptr(y, C.c_double), nrhs))
        if bp.ndim == 1:
            x = x[:, 0]
            y = None if y is None else y[:, 0]
        return (x, y) if want_y else x

    def solve_device(self, fid: int, d_b_ptr: int, d_x_ptr: int, nrhs: int):
        self._check(self.lib.sogl_solve_device(self.h, fid, C.c_void_p(d_b_ptr),
                                               C.c_void_p(d_x_ptr), nrhs))

    def stats(self, fid: int) -> FactorStats:
        st = FactorStats()
        self._check(self.lib.sogl_factor_get_stats(self.h, fid, C.byref(st)))
        return st

    def factor_tiles(self, fid: int, which: int, count: int) -> np.ndarray:
        out = np.empty((count, 64, 64))
        self._check(self.lib.sogl_factor_get_tiles(self.h, fid, which, _ptr(out, C.c_double)))
        return out

    def attach_refine(self, fid: int, plan: Plan):
        self._check(self.lib.sogl_refine_attach(self.h, fid, plan.h))

    def solve_refined(self, fid: int, b: np.ndarray, tol=1e-12, max_iter=20):
        b2 = np.asfortranarray(b.reshape(b.shape[0], -1), dtype=np.float64)
        x = np.empty_like(b2, order="F")
        it = C.c_int32()
        res = C.c_double()
        rc = self.lib.sogl_solve_refined(self.h, fid, _ptr(b2, C.c_double), _ptr(x, C.c_double),
                                         b2.shape[1], tol, max_iter, C.byref(it), C.byref(res))
        if rc not in (OK, 7):
            self._check(rc)
        return (x[:, 0] if b.ndim == 1 else x), it.value, res.value, rc

    def trace(self, fid: int, enable: bool = True):
        self._check(self.lib.sogl_factor_set_trace(self.h, fid, 1 if enable else 0))

    def get_trace(self, fid: int, n_tasks: int, times: bool = True):
        """(tasks (n,4): kind, pairs, level, deps; times (n,4): taken, work done, published, last wait end)"""
        tk = np.empty((n_tasks, 4), np.int64)
        tm = np.empty((n_tasks, 4), np.int64) if times else None
        self._check(self.lib.sogl_factor_get_trace(self.h, fid, _ptr(tk, C.c_int64),
                                                   None if tm is None else _ptr(tm, C.c_int64)))
        return tk, tm

    def get_graph(self, fid: int, n_tasks: int):
        """(dep_begin (n_tasks+1,), deps (n_deps,) pool slots, slot_task (n_slots,) producer)"""
        nd, ns = C.c_int64(), C.c_int64()
        self._check(self.lib.sogl_factor_get_graph(self.h, fid, C.byref(nd), C.byref(ns), None, None, None))
        db = np.empty(n_tasks + 1, np.int32)
        dp = np.empty(max(nd.value, 1), np.int32)
        st = np.empty(max(ns.value, 1), np.int32)
        self._check(self.lib.sogl_factor_get_graph(self.h, fid, None, None, _ptr(db, C.c_int32),
                                                   _ptr(dp, C.c_int32), _ptr(st, C.c_int32)))
        return db, dp[:nd.value], st[:ns.value]

    def alloc(self, nbytes: int) -> int:
        p = C.c_void_p()
        self._check(self.lib.sogl_device_alloc(self.h, nbytes, C.byref(p)))
        return p.value

    def free(self, ptr: int):
        self._check(self.lib.sogl_device_free(self.h, C.c_void_p(ptr)))

    def upload(self, ptr: int, a: np.ndarray):
        a = np.ascontiguousarray(a)
        self._check(self.lib.sogl_device_copy(self.h, C.c_void_p(ptr), a.ctypes.data_as(C.c_void_p),
                                              a.nbytes, 1))

    def download(self, ptr: int, out: np.ndarray):
        self._check(self.lib.sogl_device_copy(self.h, out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr),
                                              out.nbytes, 0))
        return out

    def solve_refined_device(self, fid: int, d_b: int, d_x: int, nrhs: int, tol=1e-12, max_iter=20):
        it = C.c_int32()
        res = C.c_double()
        rc = self.lib.sogl_solve_refined_device(self.h, fid, C.c_void_p(d_b), C.c_void_p(d_x), nrhs, tol,
                                                max_iter, C.byref(it), C.byref(res))
        if rc not in (OK, 7):
            self._check(rc)
        return it.value, res.value, rc

    def release(self, fid: int):
        self._check(self.lib.sogl_factor_release(self.h, fid))

    def sync(self):
        self._check(self.lib.sogl_ctx_sync(self.h))


def oracle_residual(n, row, col, val, x, b) -> float:
    """||b - A x||_2 / (||A||_F ||x||_2 + ||b||_2), FP64; the largest over columns
    when x and b are (n, k)."""
    if np.ndim(x) == 2:
        return max(oracle_residual(n, row, col, val, x[:, j], b[:, j]) for j in range(x.shape[1]))
    r = b - np.bincount(row, weights=val * x[col], minlength=n)
    return float(np.linalg.norm(r) / (np.linalg.norm(val) * np.linalg.norm(x) + np.linalg.norm(b)))


def backward_error_inf(n, row, col, val, x, b) -> float:
    """||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf), FP64 (||A||_inf: the largest
    row sum of |a_ij|); the largest over columns when x and b are (n, k)."""
    if np.ndim(x) == 2:
        return max(backward_error_inf(n, row, col, val, x[:, j], b[:, j]) for j in range(x.shape[1]))
    r = b - np.bincount(row, weights=val * x[col], minlength=n)
    a_inf = np.bincount(row, weights=np.abs(val), minlength=n).max()
    return float(np.abs(r).max() / (a_inf * np.abs(x).max() + np.abs(b).max()))
