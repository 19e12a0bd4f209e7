"""The solve entry point that allocates nothing (sogl_solve_device_noalloc), called
through the C ABI and checked bit for bit against sogl_solve_device, which runs the
same kernels with buffers the factor allocates: two calls in a row, the second on a
smaller factor with fewer right-hand sides, each in both precisions; plus the
rejection of a solve larger than the scratch array."""
import ctypes as C

import numpy as np
import pytest

import sogl
import workloads as W

pytestmark = pytest.mark.gpu


def _solve_both(dev, lib, fid, plan, nrhs, seed):
    """(x from sogl_solve_device, x from sogl_solve_device_noalloc) on one device-resident
    right-hand side block, permuted padded order, column-major N x nrhs."""
    N = plan.N
    rng = np.random.default_rng(seed)
    b = rng.uniform(-1.0, 1.0, (nrhs, N))
    nbytes = b.nbytes
    d_b, d_x1, d_x2 = dev.alloc(nbytes), dev.alloc(nbytes), dev.alloc(nbytes)
    try:
        dev.upload(d_b, b)
        dev.solve_device(fid, d_b, d_x1, nrhs)
        rc = lib.sogl_solve_device_noalloc(dev.h, fid, C.c_void_p(d_b), C.c_void_p(d_x2), nrhs)
        assert rc == sogl.OK, lib.sogl_ctx_error(dev.h).decode()
        x1 = dev.download(d_x1, np.empty((nrhs, N)))
        x2 = dev.download(d_x2, np.empty((nrhs, N)))
    finally:
        for p in (d_b, d_x1, d_x2):
            dev.free(p)
    return b, x1, x2


@pytest.mark.parametrize("precision", [sogl.FP64, sogl.FP32])
def test_noalloc_solve_matches_allocating_solve(precision):
    lib = sogl.load("libsogl.so")
    dev = sogl.Device(0)
    try:
        # first call: 2D convection-diffusion 128x128 (n=16,384), 20 right-hand sides
        # (the MFMA form, whose per-row flags live in the scratch array); second call: the
        # smaller 2D Poisson 64x64 (n=4,096) with 3, then 1 right-hand side
        big = W.convdiff2d(128, pe=10.0, vel=(1.0, 0.5), seed=4)
        small = W.config("c1")
        cases = [(big, 20, 11), (small, 3, 12), (small, 1, 13)]
        for prob, nrhs, seed in cases:
            plan = sogl.Plan(prob.n, prob.row, prob.col, prob.val, prob.b)
            fid = dev.load(plan, precision)
            dev.factor(fid, sogl.EXEC_PERSISTENT)
            print(f"noalloc: {prob.name} n={prob.n} nrhs={nrhs} precision={precision}", flush=True)
            b, x1, x2 = _solve_both(dev, lib, fid, plan, nrhs, seed)
            dev.release(fid)
            assert np.all(np.isfinite(x1))
            assert np.array_equal(x1.view(np.uint64), x2.view(np.uint64)), \
                f"{prob.name} nrhs={nrhs}: max diff {np.abs(x1 - x2).max()}"
            # and the solution solves the permuted system to the factor's accuracy
            r, c, v = plan.perm_coo()
            resid = max(np.linalg.norm(b[k] - np.bincount(r, weights=v * x1[k][c], minlength=plan.N)) /
                        (np.linalg.norm(v) * np.linalg.norm(x1[k]) + np.linalg.norm(b[k])) for k in range(nrhs))
            assert resid < (1e-14 if precision == sogl.FP64 else 1e-6), resid
    finally:
        dev.close()


def test_noalloc_solve_rejects_oversize():
    lib = sogl.load("libsogl.so")
    dev = sogl.Device(0)
    try:
        prob = W.config("c1")
        plan = sogl.Plan(prob.n, prob.row, prob.col, prob.val, prob.b)
        fid = dev.load(plan, sogl.FP64)
        dev.factor(fid, sogl.EXEC_PERSISTENT)
        # 128 block rows x 1100 right-hand sides need 72 MB of y, more than the 64 MiB array;
        # the size check comes before any device work, so the pointers are never read
        nrhs = 1100
        need = 256 + 512 + 512 * plan.info.block_rows * nrhs
        assert need > 64 << 20
        d = dev.alloc(8)
        try:
            rc = lib.sogl_solve_device_noalloc(dev.h, fid, C.c_void_p(d), C.c_void_p(d), nrhs)
        finally:
            dev.free(d)
        assert rc == 1, rc  # SOGL_ERR_ARG
        assert "scratch" in lib.sogl_ctx_error(dev.h).decode()
        dev.release(fid)
    finally:
        dev.close()
