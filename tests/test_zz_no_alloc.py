"""The no-alloc H X entry point (xtpgpu_bse_matmul_device_noalloc) against the allocating one
(xtpgpu_bse_matmul_device) through the C-ABI: same operator, same device inputs, two calls in a
row with the second one smaller (fewer columns, then a smaller operator), every precision the
operator holds; a request whose fixed scratch exceeds the 256 MiB region is refused."""
import ctypes as C

import numpy as np
import pytest

from votca_amd import pyxtp as px

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _op(ctx, seed, vt, ct, P, c=(1, 2, 1, 0)):
    rng = np.random.default_rng(seed)
    n = vt + ct
    M = rng.normal(scale=0.1, size=(n, n, P))
    M = 0.5 * (M + M.transpose(1, 0, 2))
    e = np.concatenate([np.sort(rng.uniform(-1, -0.2, vt)), np.sort(rng.uniform(0.1, 2.0, ct))])
    t = px.TCMatrix_gwbse(ctx).Initialize(P, 0, n - 1, 0, n - 1).set_tensor(M)
    op = px.BSE_OPERATOR(*c, rng.uniform(0.2, 1.0, P), t, np.diag(e), precisions=("fp64", "bf16", "bf16x3")).configure(
        px.BSEOperator_Options(vt - 1, 0, 0, 0, n - 1))
    return t, op


def _both(ctx, op, Xv, prec):
    size, k = Xv.shape
    dx = px.DeviceBuffer(ctx, Xv.nbytes)
    dx.upload(np.ascontiguousarray(Xv.T))  # column-major X
    ya, yb = px.DeviceBuffer(ctx, Xv.nbytes), px.DeviceBuffer(ctx, Xv.nbytes)
    px._check(px.lib().xtpgpu_bse_matmul_device(op.h, px._PREC[prec], dx.ptr, size, k, ya.ptr, size))
    px._check(px.lib().xtpgpu_bse_matmul_device_noalloc(op.h, px._PREC[prec], dx.ptr, size, k, yb.ptr, size))
    return ya.download((size, k)), yb.download((size, k))


@pytest.mark.parametrize("c", [(1, 2, 1, 0), (0, 2, 0, 1)])
def test_noalloc_matches_allocating_entry(c):
    ctx = px.default_context()
    t1, op1 = _op(ctx, 1, 14, 120, 200, c)
    t2, op2 = _op(ctx, 2, 5, 37, 60, c)
    rng = np.random.default_rng(3)
    for op, k in ((op1, 9), (op1, 3), (op2, 2)):  # second and third calls smaller
        Xv = rng.normal(size=(op.size, k))
        for prec, tol in (("fp64", 1e-13), ("bf16x3", 1e-6), ("bf16", 1e-4)):
            ya, yb = _both(ctx, op, Xv, prec)
            assert relerr(yb, ya) < tol, (c, op.size, k, prec, relerr(yb, ya))


def test_noalloc_forced_aux_chunks_and_refusal():
    """An operator whose stage-1 block for all aux functions exceeds the region: the no-alloc call
    splits the aux index into chunks the region holds and still matches the one-chunk result; a
    block width whose fixed scratch alone exceeds the region is refused."""
    ctx = px.default_context()
    t, op = _op(ctx, 4, 40, 600, 220)
    Xv = np.random.default_rng(5).normal(size=(op.size, 16))
    # FP64 stage-1 bytes per aux function: 600 x 16 x 40 x 8 = 3.1 MB -> 220 of them exceed 256 MiB
    ya, yb = _both(ctx, op, Xv, "fp64")
    assert relerr(yb, ya) < 1e-13
    ya, yb = _both(ctx, op, Xv, "bf16x3")
    assert relerr(yb, ya) < 1e-6
    k = 1400  # X copy alone: 40 x 600 x 1400 x 8 = 269 MB
    dx = px.DeviceBuffer(ctx, 8)
    with pytest.raises(px.XtpGpuError, match="no-alloc|scratch"):
        px._check(px.lib().xtpgpu_bse_matmul_device_noalloc(op.h, 0, dx.ptr, op.size, k, dx.ptr, op.size))
