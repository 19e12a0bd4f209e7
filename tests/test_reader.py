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


def _restated(text):
    """The per-line rules of mtx.cpp:44-117 in Python, for files whose numbers are in
    plain forms (where float() and strtod agree)."""
    import re
    lines = text.split("\n")
    if text.endswith("\n"):
        lines = lines[:-1]
    sym = "symmetric" in lines[0]
    num_i = re.compile(r"\s*([+-]?\d+)")

    def strtol(s):
        m = num_i.match(s)
        return (int(m.group(1)), s[m.end():]) if m else (0, s)
    it = iter(lines)
    n = None
    for ln in it:
        if len(ln) <= 3 or "%" in ln:
            continue
        r, rest = strtol(ln)
        c, rest = strtol(rest)
        n = min(r, c)
        break
    rows, cols, vals = [], [], []
    num_f = re.compile(r"\s*([+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?)")
    for ln in it:
        if len(ln) <= 3 or "%" in ln or len(ln) >= 1000:
            continue
        r, rest = strtol(ln)
        c, rest = strtol(rest)
        m = num_f.match(rest)
        v = float(m.group(1)) if m else 0.0
        if v == 0 or r > n or c > n:
            continue
        rows.append(r - 1)
        cols.append(c - 1)
        vals.append(v)
    return n, sym, rows, cols, vals


def test_chunked_parse_matches_line_rules(tmp_path):
    """A file of several parallel chunks with every skip rule spread through it."""
    rng = np.random.default_rng(5)
    n = 3000
    out = ["%%MatrixMarket matrix coordinate real general", "% c", f"{n} {n} 60000"]
    for k in range(60000):
        r, c = int(rng.integers(1, n + 1)), int(rng.integers(1, n + 1))
        v = float(rng.normal())
        kind = k % 97
        if kind == 0:
            out.append(f"{r} {c} 0")                     # explicit zero
        elif kind == 1:
            out.append(f"{r} {c} {v!r} % note")          # comment: skipped
        elif kind == 2:
            out.append("")                               # short
        elif kind == 3:
            out.append(f"{r} {c}")                       # no value: 0, dropped
        elif kind == 4:
            out.append(f"{r} {c} {v!r}\r")               # CRLF
        elif kind == 5:
            out.append(f"{r} {c} {v!r}" + " " * 1000)    # long line: skipped
        elif kind == 6:
            out.append(f"{n + 3} {c} {v!r}")             # beyond the size
        elif kind == 7:
            out.append(f"  {r}\t{c}   {v:.6e}")          # other spacing and form
        else:
            out.append(f"{r} {c} {v!r}")
    text = "\n".join(out)  # no final line end
    assert len(text) > 4 * 65536
    rc, m = _read(tmp_path, text)
    assert rc == 0
    n2, sym, rows, cols, vals = _restated(text)
    assert m["n"] == n2 and m["sym"] == int(sym)
    assert m["row"].tolist() == rows and m["col"].tolist() == cols
    assert np.array_equal(m["val"], np.array(vals))
