"""Shared test helpers: workloads from the generators, device runs, checks."""
import numpy as np

import oracle as O
import taichi_amd as T

# FP32 tolerance against the FP64 oracle, componentwise:
# max_i |y_i - y64_i| / max(|y64_i|, sum_j |x_j| over row i)
REL_TOL = 1e-6


def cfg1():
    """n=16,384, ones on offsets -2..2 plus 16,384 uniform draws (seed 0)."""
    n = 16384
    rp, col = T.gen_diagonals_random(n, [-2, -1, 0, 1, 2], n, 0)
    return n, n, rp, col


def rmat(scale, ef=16, seed=1, n=None):
    n = n or (1 << scale)
    rp, col = T.gen_rmat(n, scale, ef * (1 << scale), 0.499, 0.001, 0.001, seed)
    return n, n, rp, col


def host_format(M, N, rp, col):
    """Product host partition from the oracle's thumbnail (CPU only)."""
    img = O.thumbnail(M, N, rp, col)
    fmt, _ = T.partition_from_image(M, N, rp, col, img=img)
    return fmt


def check_float(y, rp, col, x64, tol=REL_TOL):
    y64 = O.csr_spmv_f64(rp, col, x64)
    absrow = O.csr_absrow_f64(rp, col, x64)
    err = O.rel_err(y, y64, absrow)
    assert err <= tol, f"relative error {err:.3e} > {tol:.0e}"
    return err


def run_spmv(fmt, x, dev, k=None, dtype=None):
    """y = A x on the GPU; x numpy float32 [N] or [N, k]."""
    import torch
    A = T.Matrix(fmt)
    xd = torch.from_numpy(np.ascontiguousarray(x)).to(dev)
    if dtype is not None:
        xd = xd.to(dtype)
    shape = (fmt.M,) if x.ndim == 1 else (fmt.M, x.shape[1])
    yd = torch.full(shape, float("nan"), dtype=torch.float32, device=dev)
    A.spmv(xd, yd)
    torch.cuda.synchronize()
    return yd.cpu().numpy(), A
