"""The drop-in CLI (build/solve, csrc/solve_main.cpp) against the reference's own
./solve (main.cpp:34-66) on the same files: the reference's stdout and _x.mtx were
recorded by tests/golden/make_golden.py (cli_<case>.stdout, cli_<case>_x.mtx.golden).

Checked, per case:
  * the same log lines in the same order (mtx.cpp, GPSOrder.cpp, solver.cpp:50-119,
    main.cpp): lines without timings verbatim, timer lines with their numbers
    masked;
  * `max rhs error` (mtx::checkResult, mtx.cpp:183-215) below 1e-13;
  * the sample of x (mtx::sampleArray, 6 significant digits) equal to the
    reference's to its printed precision;
  * _x.mtx (mtx::writeArray of data::b, mtx.cpp:130-143, main.cpp:61) byte for byte.

The CLI runs as a child process of the test (never exec'd in its place)."""
import os
import re
import subprocess

import numpy as np
import pytest

import workloads as W

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "sparse-operator-graph-lu_amd", "build", "solve")
GOLDEN = os.path.join(ROOT, "tests", "golden")
CASES = {"c1": lambda: W.config("c1"), "sym_poisson2d_64": lambda: W.symmetric_lower(W.poisson2d(64))}
TIMERS = ("re Order time", "plan time", "kernel time", "solve triangled", "totaltime")
NUM = re.compile(r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?")


def classify(lines):
    """(kind, text) per line: timer lines with numbers masked, the residual and the
    x sample by kind, everything else verbatim."""
    out = []
    for ln in lines:
        if ln.startswith(TIMERS):
            out.append(("timer", NUM.sub("#", ln)))
        elif ln.startswith("max rhs error:"):
            out.append(("residual", "max rhs error:#"))
        elif " ... " in ln:
            out.append(("sample", "..."))
        else:
            out.append(("text", ln))
    return out


@pytest.mark.parametrize("case", sorted(CASES))
def test_cli_matches_reference_output(case, tmp_path):
    if not os.path.exists(CLI):
        pytest.fail(f"{CLI} is not built")
    prob = CASES[case]()
    path = tmp_path / f"{case}.mtx"
    W.write_mtx(prob, str(path))
    print(f"cli: {case} n={prob.n}", flush=True)
    r = subprocess.run([CLI, str(path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    ours = r.stdout.splitlines()
    ref = open(os.path.join(GOLDEN, f"cli_{case}.stdout")).read().splitlines()
    assert classify(ours) == classify(ref), "\n".join(ours)
    res = float(next(ln for ln in ours if ln.startswith("max rhs error:")).split(":", 1)[1])
    assert res < 1e-13, res
    s_ours = [float(v) for v in NUM.findall(next(ln for ln in ours if " ... " in ln))]
    s_ref = [float(v) for v in NUM.findall(next(ln for ln in ref if " ... " in ln))]
    assert len(s_ours) == len(s_ref) == 6
    assert np.allclose(s_ours, s_ref, rtol=1e-5, atol=0), (s_ours, s_ref)
    got = (tmp_path / f"{case}_x.mtx").read_bytes()
    want = open(os.path.join(GOLDEN, f"cli_{case}_x.mtx.golden"), "rb").read()
    assert got == want
    # the CLI's own full-precision solution file holds x in the original order
    xs = np.loadtxt(str(tmp_path / f"{case}_xsol.mtx"), comments="%", skiprows=2)
    ref_x = np.load(os.path.join(GOLDEN, f"{case}.npz"), allow_pickle=False)["xorig"]
    assert np.linalg.norm(xs - ref_x) / np.linalg.norm(ref_x) < 1e-12
