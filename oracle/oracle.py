"""ctypes access to the plain-C oracle (build/liboracle.so).

Test infrastructure only: the tests, the smoke check and the benchmark's CPU
baseline leg use it as the checker; the product library never loads it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "build", "liboracle.so")
        if not os.path.exists(path):
            raise OSError(f"{path} is not built (make -C oracle oracle)")
        L = C.CDLL(path)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int32)
        L.oracle_gemm.argtypes = [dp, dp, dp, C.c_int]
        L.oracle_sub.argtypes = [dp, dp, dp]
        L.oracle_lu.argtypes = [dp, dp, dp]
        L.oracle_inv_lower.argtypes = [dp, dp]
        L.oracle_inv_upper.argtypes = [dp, dp]
        L.oracle_llt.argtypes = [dp, dp]
        L.oracle_gemm_t.argtypes = [dp, dp, dp]
        L.oracle_factor_solve.restype = C.c_int
        L.oracle_factor_solve.argtypes = [C.c_int32, ip, C.c_int64, ip, dp, C.c_int64, ip, C.c_int32,
                                          ip, ip, ip, C.c_int64, ip, ip, ip, C.c_int64, dp, dp, C.c_int,
                                          C.c_int]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def gemm(a, b, c=None, sign=1):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    c = np.zeros((64, 64)) if c is None else np.array(c, np.float64, copy=True)
    lib().oracle_gemm(_d(a), _d(b), _d(c), sign)
    return c


def gemm_t(a, b, c=None):
    """c + a b^T"""
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    c = np.zeros((64, 64)) if c is None else np.array(c, np.float64, copy=True)
    lib().oracle_gemm_t(_d(a), _d(b), _d(c))
    return c


def llt(a):
    a = np.ascontiguousarray(a, np.float64)
    l = np.empty((64, 64))
    lib().oracle_llt(_d(a), _d(l))
    return l


def sub(a, b):
    r = np.empty((64, 64))
    a = None if a is None else np.ascontiguousarray(a, np.float64)
    b = None if b is None else np.ascontiguousarray(b, np.float64)
    lib().oracle_sub(None if a is None else _d(a), None if b is None else _d(b), _d(r))
    return r


def lu(a):
    a = np.ascontiguousarray(a, np.float64)
    l, u = np.empty((64, 64)), np.empty((64, 64))
    lib().oracle_lu(_d(a), _d(l), _d(u))
    return l, u


def inv_lower(l):
    l = np.ascontiguousarray(l, np.float64)
    y = np.empty((64, 64))
    lib().oracle_inv_lower(_d(l), _d(y))
    return y


def inv_upper(u):
    u = np.ascontiguousarray(u, np.float64)
    y = np.empty((64, 64))
    lib().oracle_inv_upper(_d(u), _d(y))
    return y


def factor_solve(plan, bp: np.ndarray, threads: int = 1):
    """Factor a sogl.Plan's op list on the CPU and solve for the permuted padded
    right-hand side(s) bp (N or N x nrhs).  Returns (x, y) in the permuted order."""
    ops = np.ascontiguousarray(plan.ops(0))
    ids, _, _, tiles = plan.inputs()
    laststage, _ = plan.tile_stages()
    lbi, lbj, lid = plan.factor_tiles(0)
    ubi, ubj, uid = plan.factor_tiles(1)
    b2 = np.asfortranarray(bp.reshape(plan.N, -1), dtype=np.float64).copy(order="F")
    nrhs = b2.shape[1]
    x = np.empty_like(b2, order="F")
    tiles = np.ascontiguousarray(tiles)
    rc = lib().oracle_factor_solve(plan.info.storage_count, _i(ops), ops.shape[0], _i(ids), _d(tiles),
                                   ids.size, _i(laststage), plan.info.block_rows, _i(lbi), _i(lbj),
                                   _i(lid), lid.size, _i(ubi), _i(ubj), _i(uid), uid.size, _d(b2),
                                   _d(x), nrhs, threads)
    if rc != 0:
        raise RuntimeError("oracle_factor_solve failed")
    if bp.ndim == 1:
        return x[:, 0], b2[:, 0]
    return x, b2
