"""The reference-side integration of INTEGRATION.md section 2: gpu_calculate and
gpu_solve, written against the reference's own globals, in place of
BlockPlanner::calculate() (solver.cpp:106) and BlockPlanner::solve() (solver.cpp:115).

`make -C oracle integration` extracts that code block verbatim and builds
oracle/_ref/ref_dump_gpu: the fixture harness (oracle/ref_dump.cpp, which drives the
reference's objects through solver.cpp:50-184) with the two functions in place of
the two calls.
  * CPU: the block compiles against the reference headers and include/sogl.h (skipped
    where the reference sources are absent);
  * GPU: the harness's x (its reference planner, our device factor and solve through
    the block) equals the reference's own FP64 x, and the forward result it leaves
    in data::b equals the reference's.
"""
import os
import subprocess

import numpy as np
import pytest

import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
HARNESS = os.path.join(ORACLE, "_ref", "ref_dump_gpu")
GOLDEN = os.path.join(ROOT, "tests", "golden")
CASES = {"c1": lambda: W.config("c1"), "sym_poisson2d_64": lambda: W.symmetric_lower(W.poisson2d(64))}


@pytest.mark.skipif(not os.path.isdir("/root/reference"), reason="reference sources not present")
def test_integration_block_compiles():
    lib = os.path.join(ROOT, "sparse-operator-graph-lu_amd", "build", "libsogl.so")
    if not os.path.exists(lib):
        pytest.fail("build the package first (libsogl.so)")
    r = subprocess.run(["make", "-C", ORACLE, "integration"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    block = open(os.path.join(ORACLE, "_ref", "integration_sites.inc")).read()
    assert "gpu_calculate" in block and "gpu_solve" in block
    assert os.path.exists(HARNESS)


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(CASES))
def test_integration_block_solves_like_the_reference(case, tmp_path):
    if not os.path.exists(HARNESS):
        pytest.fail(f"{HARNESS} is not built (make -C oracle integration, where the reference is present)")
    prob = CASES[case]()
    path = tmp_path / "m.mtx"
    W.write_mtx(prob, str(path))
    out = tmp_path / "dump"
    out.mkdir()
    print(f"integration: {case} n={prob.n}", flush=True)
    r = subprocess.run([HARNESS, str(path), str(out)], capture_output=True, text=True, timeout=120,
                       env=dict(os.environ, OMP_NUM_THREADS="4"))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    g = np.load(os.path.join(GOLDEN, f"{case}.npz"), allow_pickle=False)
    xo = np.fromfile(out / "xorig.f64", np.float64)
    assert np.linalg.norm(xo - g["xorig"]) / np.linalg.norm(g["xorig"]) < 1e-12
    y = np.fromfile(out / "y.f64", np.float64)
    assert np.linalg.norm(y - g["y"]) / np.linalg.norm(g["y"]) < 1e-12
    res = float(next(ln for ln in r.stdout.splitlines() if ln.startswith("max rhs error:")).split(":", 1)[1])
    assert res < 1e-13, res
