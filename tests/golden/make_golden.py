"""Regenerate the golden fixtures from the reference itself.

Runs only where the reference sources are present: `make -C oracle ref`
compiles them (in place, read-only) into oracle/_ref, and its harness
oracle/_ref/ref_dump replays SOGLU::solveLU step by step on each workload,
dumping the permutation, the coarse and fine op lists, tile stages, input
tile map and values, factor tile map and the FP64 solution.  This script packs those
dumps into tests/golden/*.npz (allow_pickle=False) -- data only:

  small cases   full arrays
  large cases   SHA-256 of the canonical int32 little-endian dumps, counts,
                and a strided sample of the reference x

For the drop-in CLI it also runs the reference's own ./solve (oracle/_ref/solve) on
the CLI_CASES and keeps its stdout and the _x.mtx it writes as text files
(cli_<case>.stdout, cli_<case>_x.mtx.golden).

Usage: python tests/golden/make_golden.py [case ...]
"""
from __future__ import annotations

import hashlib
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "sparse-operator-graph-lu_amd"))
import workloads as W  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden")
REF_DUMP = os.path.join(ROOT, "oracle", "_ref", "ref_dump")

# name -> (generator, full arrays? ("x": the solution in full, the rest as hashes), plan only?)
CASES = {
    "c1": (lambda: W.config("c1"), True, False),
    "random_band_2048": (lambda: W.random_band(2048, nrhs=1), True, False),
    "disconnected_40": (lambda: W.edge_disconnected(40), True, False),
    "zerodiag_48": (lambda: W.edge_zero_diag(48), True, False),
    "unsym_3000": (lambda: W.edge_unsym_pattern(3000), True, False),
    "convdiff2d_128": (lambda: W.convdiff2d(128), True, False),
    "diffusion3d_20": (lambda: W.diffusion_varcoef(20, 3, 5), True, False),
    "repeated_1200": (lambda: W.edge_repeated_entries(1200), True, False),
    "sym_poisson2d_64": (lambda: W.symmetric_lower(W.poisson2d(64)), True, False),
    "sym_diffusion3d_16": (lambda: W.symmetric_lower(W.diffusion_varcoef(16, 3, 21)), True, False),
    "sym_diffusion2d_100": (lambda: W.symmetric_lower(W.diffusion_varcoef(100, 2, 22)), True, False),
    "sym_diffusion3d_48": (lambda: W.symmetric_lower(W.diffusion_varcoef(48, 3, 23)), False, False),
    # a serving-stream request (serve.matrix_problem((2, 128, 117))): 2D variable-coefficient
    # diffusion, log-normal coefficients
    "c4_2d_128": (lambda: W.diffusion_varcoef(128, 2, 117), "x", False),
    "c2": (lambda: W.config("c2"), False, False),
    "c5": (lambda: W.config("c5"), False, False),
}


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).astype("<i4").tobytes()).hexdigest()


def sha_f64(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).astype("<f8").tobytes()).hexdigest()


def run_case(name: str) -> None:
    gen, full, plan_only = CASES[name]
    prob = gen()
    with tempfile.TemporaryDirectory() as d:
        W.write_mtx(prob, os.path.join(d, "m.mtx"))
        cmd = [REF_DUMP, os.path.join(d, "m.mtx"), d] + (["--plan-only"] if plan_only else [])
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"{name}: ref_dump failed: {r.stderr}")
        rd = lambda f, t: np.fromfile(os.path.join(d, f), t)  # noqa: E731
        n = prob.n
        perm = rd("perm.i32", np.int32)
        ops = rd("ops.i32", np.int32).reshape(-1, 8)
        cops = rd("coarse_ops.i32", np.int32).reshape(-1, 8)
        tiles = rd("tiles.i32", np.int32)
        S = int(tiles[0])
        inputs = rd("inputs.i32", np.int32).reshape(-1, 3)
        lt = rd("ltiles.i32", np.int32).reshape(-1, 3)
        ut = rd("utiles.i32", np.int32).reshape(-1, 3)
        meta = dict(line.split() for line in open(os.path.join(d, "meta.txt")))
        out = dict(n=np.int64(n), storage=np.int64(S), n_ops=np.int64(ops.shape[0]),
                   n_coarse_ops=np.int64(cops.shape[0]),
                   n_stages=np.int64(ops[-1, 0] if ops.size else 0),
                   n_inputs=np.int64(inputs.shape[0]), n_l=np.int64(lt.shape[0]),
                   n_u=np.int64(ut.shape[0]),
                   sha_perm=np.str_(sha(perm)), sha_ops=np.str_(sha(ops)),
                   sha_coarse_ops=np.str_(sha(cops)), sha_tiles=np.str_(sha(tiles)),
                   sha_inputs=np.str_(sha(inputs)), sha_ltiles=np.str_(sha(lt)),
                   sha_utiles=np.str_(sha(ut)), sha_input_vals=np.str_(sha_f64(rd("inputs.f64", np.float64))),
                   ref_t_factor=np.float64(float(meta["t_factor"])),
                   ref_t_solve=np.float64(float(meta["t_solve"])),
                   ref_stdout=np.str_(r.stdout))
        if full is True:
            out.update(perm=perm, ops=ops, coarse_ops=cops, tiles=tiles, inputs=inputs, ltiles=lt,
                       utiles=ut, bperm=rd("bperm.f64", np.float64))
        if not plan_only:
            x = rd("x.f64", np.float64)
            y = rd("y.f64", np.float64)
            xo = rd("xorig.f64", np.float64)
            if full is True:
                out.update(x=x, y=y, xorig=xo)
            elif full == "x":
                out.update(xorig=xo)
            else:
                idx = np.unique(np.linspace(0, n - 1, 4096).astype(np.int64))
                out.update(xorig_idx=idx, xorig_sample=xo[idx], xorig_norm=np.float64(np.linalg.norm(xo)),
                           xorig_sum=np.float64(xo.sum()))
    path = os.path.join(OUT, f"{name}.npz")
    np.savez_compressed(path, **out)
    print(f"{name}: n={n} ops={ops.shape[0]} storage={S} -> {os.path.relpath(path, ROOT)} "
          f"({os.path.getsize(path) // 1024} KiB)")


REF_CLI = os.path.join(ROOT, "oracle", "_ref", "solve")
CLI_CASES = ("c1", "sym_poisson2d_64")


def run_cli(name: str) -> None:
    """The reference CLI on one case: its stdout and its _x.mtx, verbatim."""
    prob = CASES[name][0]()
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, f"{name}.mtx")
        W.write_mtx(prob, path)
        r = subprocess.run([REF_CLI, path], capture_output=True, text=True, env=dict(os.environ, OMP_NUM_THREADS="4"))
        if r.returncode != 0:
            raise RuntimeError(f"{name}: reference CLI failed: {r.stderr}")
        with open(os.path.join(OUT, f"cli_{name}.stdout"), "w") as fh:
            fh.write(r.stdout)
        with open(os.path.join(d, f"{name}_x.mtx"), "rb") as src, open(os.path.join(OUT, f"cli_{name}_x.mtx.golden"), "wb") as dst:
            dst.write(src.read())
    print(f"cli {name}: stdout {len(r.stdout.splitlines())} lines")


def main(argv):
    if not os.path.exists(REF_DUMP):
        raise SystemExit("oracle/_ref/ref_dump is not built: make -C oracle ref")
    for name in (argv or list(CASES)):
        run_case(name)
    for name in CLI_CASES:
        if not argv or name in argv:
            run_cli(name)


if __name__ == "__main__":
    main(sys.argv[1:])
