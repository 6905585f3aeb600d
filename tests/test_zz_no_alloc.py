"""The allocation-free Heff apply (tn_heff_apply_device_noalloc) against the
allocating device-pointer apply it is a variant of (tn_heff_apply_device),
through the C ABI: two calls in a row, the second on a smaller site, then a
site beyond the fixed work space (rejected), then the operator checked
against the oracle."""
import numpy as np
import pytest

import oracle as O
import tensornetworks_amd as tn
from devmem import DeviceArray

pytestmark = pytest.mark.gpu


def rand(shape, seed):
    return np.random.default_rng(seed).uniform(-1.0, 1.0, shape)


def sym_W(Dw1, Dw2, d, seed):
    W = rand((Dw1, Dw2, d, d), seed)
    W = 0.5 * (W + W.transpose(0, 1, 3, 2))
    W[np.abs(W) < 0.3] = 0.0
    return np.ascontiguousarray(W)


def with_identity_channel(F, Dw, D, w):
    Fr = F.reshape(D, D, Dw).copy()
    Fr[:, :, w] = np.eye(D)
    return Fr.ravel()


def apply_both(ctx, W, L, R, d, D1, D2, prec, seed):
    ctx.site_bind(W, L, R, d, D1, D2, prec)
    n = d * D1 * D2
    x = rand(n, seed)
    xd, ya, yb = DeviceArray(n), DeviceArray(n), DeviceArray(n)
    xd.upload(x)
    ya.upload(np.full(n, np.nan))
    yb.upload(np.full(n, np.nan))
    ctx.heff_apply_device(xd.address, ya.address, n)
    ctx.heff_apply_device_noalloc(xd.address, yb.address, n)
    out = x, ya.download(), yb.download()
    for b in (xd, ya, yb):
        b.free()
    return out


@pytest.mark.parametrize("prec,tol", [(tn.TN_FP64, 1e-13), (tn.TN_FP32, 5e-6), (tn.TN_BF16, 2e-2),
                                      (tn.TN_BF16X3, 5e-6)])
def test_noalloc_apply_matches_allocating_apply(ctx, prec, tol):
    # call 1: d=3, D1=150, D2=131, identity channels on both sides (the
    # copied-channel path); call 2: a smaller site (d=2, 37 x 45, no identity)
    d, D1, D2 = 3, 150, 131
    W = sym_W(5, 5, d, 1)
    L = with_identity_channel(rand(5 * D1 * D1, 2), 5, D1, 0)
    R = with_identity_channel(rand(5 * D2 * D2, 3), 5, D2, 4)
    x, ya, yb = apply_both(ctx, W, L, R, d, D1, D2, prec, 4)
    y_ref = O.heff_apply_ref(W, L, R, x, d, D1, D2)
    assert np.all(np.isfinite(yb))
    assert np.linalg.norm(yb - ya) / np.linalg.norm(ya) < (1e-14 if prec == tn.TN_FP64 else tol / 10)
    assert np.linalg.norm(yb - y_ref) / np.linalg.norm(y_ref) < tol
    d, D1, D2 = 2, 37, 45
    W = sym_W(5, 5, d, 5)
    L, R = rand(5 * D1 * D1, 6), rand(5 * D2 * D2, 7)
    x, ya, yb = apply_both(ctx, W, L, R, d, D1, D2, prec, 8)
    y_ref = O.heff_apply_ref(W, L, R, x, d, D1, D2)
    assert np.all(np.isfinite(yb))
    assert np.linalg.norm(yb - ya) / np.linalg.norm(ya) < (1e-14 if prec == tn.TN_FP64 else tol / 10)
    assert np.linalg.norm(yb - y_ref) / np.linalg.norm(y_ref) < tol


def test_noalloc_apply_rejects_site_beyond_work_space(ctx):
    # FP64, d=3, D=512, Dw=5: x, y and two stage buffers of 5 channels,
    # 3*512^2*8*(2 + 5 + 5) bytes = 75.5 MB > 64 MiB
    d, D = 3, 512
    W = sym_W(5, 5, d, 9)
    L, R = rand(5 * D * D, 10), rand(5 * D * D, 11)
    ctx.site_bind(W, L, R, d, D, D, tn.TN_FP64)
    n = d * D * D
    xd, yd = DeviceArray(n), DeviceArray(n)
    xd.upload(rand(n, 12))
    with pytest.raises(tn.TnError, match="work space"):
        ctx.heff_apply_device_noalloc(xd.address, yd.address, n)
    # the allocating variant still serves it
    ctx.heff_apply_device(xd.address, yd.address, n)
    assert np.all(np.isfinite(yd.download()))
    xd.free()
    yd.free()
