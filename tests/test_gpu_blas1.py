"""BLAS-1 shims of cons/sparse_gpu.py:17-69 (gpu_dot, gpu_norm_sq, gpu_axpy, gpu_xpay,
gpu_copy, gpu_fill_zero, gpu_scale, gpu_negate_and_add) through the C ABI, against numpy.

Bars: copy / fill / scale / negate_and_add are single FP32 operations, bit-exact against numpy
FP32; axpy / xpay are one fused multiply-add each (fmaf), compared with the FP64 evaluation
of the same expression rounded to FP32 (the product of two FP32 values is exact in FP64, so
only a rare double rounding can differ: at most 1 ulp, and almost every element exact); dot
and norm_sq accumulate FP32 products in FP64 and must equal math.fsum of the same products
within 1e-13 relative (reordering FP64 sums of a million terms)."""

import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

SIZES = [0, 1, 63, 257, 100_003, 1_000_003]   # empty, one wavefront, ragged, > 2048 blocks


def _vecs(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(n).astype(np.float32) * np.float32(3.0)
    b = rng.standard_normal(n).astype(np.float32)
    return a, b


def _dev(x):
    return torch.from_numpy(x.copy()).to("cuda:0")


def _fma32(alpha, x, y):
    return (np.float64(np.float32(alpha)) * x.astype(np.float64) + y.astype(np.float64)).astype(np.float32)


def _close_fma(got, ref):
    if got.size == 0:
        return
    np.testing.assert_array_max_ulp(got, ref, maxulp=1)
    assert np.count_nonzero(got != ref) <= max(2, got.size // 10000)


@pytest.mark.parametrize("n", SIZES)
def test_dot_and_norm_sq(n):
    from controllable_pbd_3d_amd.sparse_gpu import gpu_dot, gpu_norm_sq
    a, b = _vecs(n, seed=n)
    print(f"[blas1] dot n={n}", flush=True)
    got = gpu_dot(_dev(a), _dev(b), n)
    ref = math.fsum((a.astype(np.float64) * b.astype(np.float64)).tolist())
    assert abs(got - ref) <= 1e-13 * max(1.0, math.fsum(np.abs(a.astype(np.float64) * b).tolist()))
    nrm = gpu_norm_sq(_dev(a), n)
    ref2 = math.fsum((a.astype(np.float64) ** 2).tolist())
    assert abs(nrm - ref2) <= 1e-13 * max(1.0, ref2)
    if n == 0:
        assert got == 0.0 and nrm == 0.0


@pytest.mark.parametrize("n", SIZES)
def test_elementwise(n):
    from controllable_pbd_3d_amd.sparse_gpu import (gpu_axpy, gpu_copy, gpu_fill_zero,
                                                    gpu_negate_and_add, gpu_scale, gpu_xpay)
    a, b = _vecs(n, seed=7 + n)
    alpha = -0.37
    print(f"[blas1] elementwise n={n}", flush=True)
    y = _dev(b)
    gpu_axpy(alpha, _dev(a), y, n)               # y += alpha x
    _close_fma(y.cpu().numpy(), _fma32(alpha, a, b))
    y = _dev(b)
    gpu_xpay(alpha, _dev(a), y, n)               # y = x + alpha y
    _close_fma(y.cpu().numpy(), _fma32(alpha, b, a))
    y = _dev(b)
    gpu_copy(_dev(a), y, n)
    assert np.array_equal(y.cpu().numpy(), a)
    y = _dev(b)
    gpu_fill_zero(y, n)
    assert not np.any(y.cpu().numpy())
    y = _dev(b)
    gpu_scale(y, alpha, n)
    assert np.array_equal(y.cpu().numpy(), b * np.float32(alpha))
    r = _dev(b)
    gpu_negate_and_add(_dev(a), r, n)            # r = b - r
    assert np.array_equal(r.cpu().numpy(), a - b)


def test_partial_length_leaves_the_tail():
    """n below the buffer length touches only the first n entries (the reference passes n)."""
    from controllable_pbd_3d_amd.sparse_gpu import gpu_axpy, gpu_fill_zero
    a, b = _vecs(1000, seed=3)
    y = _dev(b)
    gpu_fill_zero(y, 600)
    out = y.cpu().numpy()
    assert not np.any(out[:600]) and np.array_equal(out[600:], b[600:])
    y = _dev(b)
    gpu_axpy(2.0, _dev(a), y, 10)
    out = y.cpu().numpy()
    assert np.array_equal(out[10:], b[10:])
