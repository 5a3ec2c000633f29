"""GPU parity of the ramp-filter chain (fft, filter, ifft, fused) against numpy and the oracle.

Tolerances: float32 storage <= 1e-5 relative L2 (the parity bar); float16
storage <= 1e-3 and bfloat16 <= 8e-3 relative L2 (storage rounding only:
2^-11 and 2^-8 relative per sample).
"""
import numpy as np
import pytest

import oracle
import ufomi
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _dev(a, torch_dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(torch_dev)


def _empty(shape, torch_dev, dtype=None):
    import torch
    return torch.empty(shape, dtype=dtype or torch.float32, device=torch_dev)


@pytest.mark.parametrize("W,P,N", [(512, 512, 1024), (2048, 33, 4096), (33, 17, 64), (9, 5, 16), (8, 3, 8),
                                   (1000, 7, 8192), (3, 4, 4), (100, 6, 128), (700, 9, 2048)])
def test_fft_rows_vs_numpy(ctx, torch_dev, W, P, N):
    # input of tests/test_fft.py:47 (linspace(-1, 2)), ground truth numpy (:15-23)
    x = np.linspace(-1, 2, num=W * P, dtype=np.float32).reshape(P, W)
    d_in = _dev(x, torch_dev)
    d_out = _empty((P, 2 * N), torch_dev)
    ctx.fft_rows(d_in, W, P, N, d_out)
    got = d_out.cpu().numpy()
    got = got[:, 0::2] + 1j * got[:, 1::2]
    ref = np.fft.fft(x.astype(np.float64), n=N, axis=1)
    assert rel_l2(got, ref) < 2e-6
    np.testing.assert_almost_equal(got, ref.astype(np.complex64), decimal=2)  # the reference's own bar


@pytest.mark.parametrize("N,P,crop", [(1024, 512, 512), (4096, 11, 2048), (16, 3, 0), (8, 4, 5), (2048, 2, 1024)])
def test_ifft_rows_vs_numpy(ctx, torch_dev, N, P, crop):
    rng = np.random.default_rng(N + P)
    z = (rng.standard_normal((P, N)) + 1j * rng.standard_normal((P, N))).astype(np.complex64)
    inter = np.stack([z.real, z.imag], axis=-1).reshape(P, 2 * N).astype(np.float32)
    c = crop or N
    d_in = _dev(inter, torch_dev)
    d_out = _empty((P, c), torch_dev)
    ctx.ifft_rows(d_in, N, P, crop, d_out)
    ref = np.fft.ifft(z.astype(np.complex128), axis=1).real[:, :c]
    assert rel_l2(d_out.cpu().numpy(), ref) < 2e-6
    # input left unchanged
    assert np.array_equal(d_in.cpu().numpy(), inter)


def test_fft_ifft_round_trip(ctx, torch_dev):
    # tests/test_fft.py:114-129: fft (zero padded) -> ifft (cropped) returns the input, decimal=5
    P, W, N = 17, 33, 64
    x = np.linspace(-1, 2, num=W * P, dtype=np.float32).reshape(P, W)
    d_in = _dev(x, torch_dev)
    spec = _empty((P, 2 * N), torch_dev)
    back = _empty((P, W), torch_dev)
    ctx.fft_rows(d_in, W, P, N, spec)
    ctx.ifft_rows(spec, N, P, W, back)
    np.testing.assert_almost_equal(back.cpu().numpy(), x, decimal=5)


def test_ones_round_trip_max_abs(ctx, torch_dev):
    # 1-D analogue of tests/test-177.sh: ones through fft (zero padded) and ifft crop, max |err| <= 1e-6
    x = np.ones((9, 9), dtype=np.float32)
    spec = _empty((9, 32), torch_dev)
    back = _empty((9, 9), torch_dev)
    ctx.fft_rows(_dev(x, torch_dev), 9, 9, 16, spec)
    ctx.ifft_rows(spec, 16, 9, 9, back)
    assert np.abs(back.cpu().numpy() - x).max() <= 1e-6


def test_filter_apply(ctx, torch_dev):
    rng = np.random.default_rng(3)
    P, width = 7, 256
    x = rng.standard_normal((P, width)).astype(np.float32)
    c = oracle.filter_coefficients(width)
    d_x, d_c = _dev(x, torch_dev), _dev(c, torch_dev)
    out = _empty((P, width), torch_dev)
    ctx.filter_apply(d_x, out, d_c, width, P)
    np.testing.assert_array_equal(out.cpu().numpy(), x * c[None, :])
    ctx.filter_apply(d_x, d_x, d_c, width, P)  # in place
    np.testing.assert_array_equal(d_x.cpu().numpy(), x * c[None, :])


FILTER_CASES = ["ramp", "ramp-fromreal", "butterworth", "faris-byer", "hamming", "bh3"]


@pytest.mark.parametrize("filt", FILTER_CASES)
@pytest.mark.parametrize("W,P,N", [(512, 64, 1024), (300, 5, 512), (2048, 9, 4096)])
def test_fused_chain_vs_oracle(ctx, torch_dev, filt, W, P, N):
    rng = np.random.default_rng(W + P)
    x = rng.standard_normal((P, W)).astype(np.float32)
    coef = ufomi.filter_coefficients(2 * N, filt)
    ref = oracle.ramp_chain(x, N, oracle.filter_coefficients(2 * N, filt), W)
    d_c = _dev(coef, torch_dev)
    out = _empty((P, W), torch_dev)
    ctx.ramp_filter(_dev(x, torch_dev), W, P, N, d_c, W, out)
    assert rel_l2(out.cpu().numpy(), ref) < 1e-5


@pytest.mark.parametrize("filt", ["ramp-fromreal", "faris-byer"])
def test_separate_tasks_vs_oracle(ctx, torch_dev, filt):
    P, W, N = 31, 500, 1024
    rng = np.random.default_rng(7)
    x = rng.standard_normal((P, W)).astype(np.float32)
    coef = ufomi.filter_coefficients(2 * N, filt)
    spec = _empty((P, 2 * N), torch_dev)
    out = _empty((P, W), torch_dev)
    ctx.fft_rows(_dev(x, torch_dev), W, P, N, spec)
    ctx.filter_apply(spec, spec, _dev(coef, torch_dev), 2 * N, P)
    ctx.ifft_rows(spec, N, P, W, out)
    ref = oracle.ramp_chain(x, N, oracle.filter_coefficients(2 * N, filt), W)
    assert rel_l2(out.cpu().numpy(), ref) < 1e-5


@pytest.mark.parametrize("dtype,tol", [(ufomi.F16, 1e-3), (ufomi.BF16, 8e-3)])
def test_fused_chain_low_precision_storage(ctx, torch_dev, dtype, tol):
    import torch
    P, W, N = 40, 512, 1024
    x = np.random.default_rng(11).standard_normal((P, W)).astype(np.float32)
    coef = ufomi.filter_coefficients(2 * N)
    ref = oracle.ramp_chain(x, N, oracle.filter_coefficients(2 * N), W)
    out = torch.empty((P, W), dtype=torch.float16 if dtype == ufomi.F16 else torch.bfloat16, device=torch_dev)
    ctx.ramp_filter(_dev(x, torch_dev), W, P, N, _dev(coef, torch_dev), W, out, dtype)
    assert rel_l2(out.float().cpu().numpy(), ref) < tol


def test_fused_chain_odd_rows_and_default_crop(ctx, torch_dev):
    P, W, N = 3, 100, 256
    x = np.random.default_rng(5).standard_normal((P, W)).astype(np.float32)
    coef = ufomi.filter_coefficients(2 * N)
    out = _empty((P, N), torch_dev)
    ctx.ramp_filter(_dev(x, torch_dev), W, P, N, _dev(coef, torch_dev), 0, out)
    ref = oracle.ramp_chain(x, N, oracle.filter_coefficients(2 * N), N)
    assert rel_l2(out.cpu().numpy(), ref) < 1e-5


def test_errors(ctx, torch_dev):
    x = _empty((4, 100), torch_dev)
    out = _empty((4, 512), torch_dev)
    coef = _dev(ufomi.filter_coefficients(256), torch_dev)
    with pytest.raises(ufomi.UfomiError) as e:
        ctx.ramp_filter(x, 100, 4, 64, coef, 0, out)  # size-x smaller than the input
    assert e.value.code == -6
    with pytest.raises(ufomi.UfomiError) as e:
        ctx.ifft_rows(out, 64, 4, 65, x)  # crop larger than the input
    assert e.value.code == -6
    with pytest.raises(ufomi.UfomiError) as e:
        ctx.fft_rows(x, 100, 4, 96, out)  # not a power of two
    assert e.value.code == -4
