"""Matrix Market reader rules (mtx.cpp:44-182), pinned by hand-written files."""
import os

import numpy as np
import pytest

import sogl


def _read(tmp_path, text, rhs=None):
    p = tmp_path / "a.mtx"
    p.write_text(text)
    if rhs is not None:
        (tmp_path / "a_b.mtx").write_text(rhs)
    lib = sogl.load("libsogl_host.so")
    import ctypes as C
    h = C.c_void_p()
    rc = lib.sogl_mtx_read(str(p).encode(), C.byref(h))
    if rc != 0:
        return rc, None
    lib.sogl_mtx_read_rhs(h, str(tmp_path / "a_b.mtx").encode())
    n, nnz, sym = C.c_int32(), C.c_int64(), C.c_int32()
    lib.sogl_mtx_info(h, C.byref(n), C.byref(nnz), C.byref(sym))
    r, c, v, b = (C.POINTER(C.c_int32)(), C.POINTER(C.c_int32)(), C.POINTER(C.c_double)(),
                  C.POINTER(C.c_double)())
    lib.sogl_mtx_arrays(h, C.byref(r), C.byref(c), C.byref(v), C.byref(b))
    k = nnz.value
    out = dict(n=n.value, sym=sym.value, row=np.ctypeslib.as_array(r, (k,)).copy() if k else np.zeros(0),
               col=np.ctypeslib.as_array(c, (k,)).copy() if k else np.zeros(0),
               val=np.ctypeslib.as_array(v, (k,)).copy() if k else np.zeros(0),
               b=np.ctypeslib.as_array(b, (n.value,)).copy() if b else None)
    lib.sogl_mtx_free(h)
    return 0, out


def test_rules(tmp_path):
    text = ("%%MatrixMarket matrix coordinate real general\n"
            "% comment line\n"
            "\n"
            "4 4 7\n"
            "1 1 2.5\n"
            "2 2 0\n"          # explicit zero: dropped
            "2 1 -1e-3\n"
            "5 1 7\n"          # row beyond the size: dropped
            "3 3 1.0 % trailing comment -> whole line skipped\n"
            "4 4 3\n"
            "3 3 4\n")
    rc, m = _read(tmp_path, text, rhs="%%MatrixMarket matrix array real general\n4 1\n1.5\n-2\n")
    assert rc == 0
    assert m["n"] == 4 and m["sym"] == 0
    assert m["row"].tolist() == [0, 1, 3, 2]
    assert m["col"].tolist() == [0, 0, 3, 2]
    assert m["val"].tolist() == [2.5, -1e-3, 3.0, 4.0]
    assert m["b"].tolist() == [1.5, -2.0, 1.0, 1.0]  # missing entries are 1


def test_nonsquare_and_symmetric_flag(tmp_path):
    rc, m = _read(tmp_path, "%%MatrixMarket matrix coordinate real symmetric\n3 5 3\n1 1 1\n3 3 2\n2 5 9\n")
    assert rc == 0
    assert m["n"] == 3 and m["sym"] == 1
    assert m["col"].tolist() == [0, 2]  # column 5 > truncated size 3 dropped


def test_skew_symmetric_counts_as_symmetric(tmp_path):
    rc, m = _read(tmp_path, "%%MatrixMarket matrix coordinate real skew-symmetric\n2 2 1\n1 1 1\n")
    assert rc == 0 and m["sym"] == 1


def test_missing_file(tmp_path):
    rc, _ = _read(tmp_path / "nope", "") if False else (None, None)
    lib = sogl.load("libsogl_host.so")
    import ctypes as C
    h = C.c_void_p()
    assert lib.sogl_mtx_read(str(tmp_path / "absent.mtx").encode(), C.byref(h)) == 2  # SOGL_ERR_IO


def test_missing_rhs_file_gives_ones(tmp_path):
    rc, m = _read(tmp_path, "%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1\n2 2 1\n")
    assert m["b"].tolist() == [1.0, 1.0]


def test_roundtrip_generated(tmp_path):
    import workloads as W
    p = W.poisson2d(8)
    path = str(tmp_path / "p.mtx")
    W.write_mtx(p, path)
    lib = sogl.load("libsogl_host.so")
    import ctypes as C
    h = C.c_void_p()
    assert lib.sogl_mtx_read(path.encode(), C.byref(h)) == 0
    lib.sogl_mtx_read_rhs(h, (path[:-4] + "_b.mtx").encode())
    n, nnz, sym = C.c_int32(), C.c_int64(), C.c_int32()
    lib.sogl_mtx_info(h, C.byref(n), C.byref(nnz), C.byref(sym))
    assert n.value == p.n and nnz.value == p.nnz
    r, c, v, b = (C.POINTER(C.c_int32)(), C.POINTER(C.c_int32)(), C.POINTER(C.c_double)(),
                  C.POINTER(C.c_double)())
    lib.sogl_mtx_arrays(h, C.byref(r), C.byref(c), C.byref(v), C.byref(b))
    assert np.array_equal(np.ctypeslib.as_array(v, (p.nnz,)), p.val)
    assert np.array_equal(np.ctypeslib.as_array(b, (p.n,)), p.b)  # %.17g round-trips exactly
    lib.sogl_mtx_free(h)
