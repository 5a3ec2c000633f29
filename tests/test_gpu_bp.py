"""GPU parity of the parallel-beam backprojector against the oracle.

Bar: float32 result within 1e-5 relative L2 of the oracle's float32
reference-semantics restatement (backproject.cl:50-94) and of the FP64 oracle.
Low-precision storage is bounded by the storage rounding (f16 1e-3, bf16 8e-3).
"""
import math

import numpy as np
import pytest

import oracle
import phantoms
import ufomi
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _dev(a, torch_dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(torch_dev)


def _bp(ctx, torch_dev, sino, n_slices=1, dtype=ufomi.F32, **props):
    import torch
    sino = np.asarray(sino, dtype=np.float32)
    if n_slices == 1 and sino.ndim == 2:
        sino = sino[None]
    _, P, W = sino.shape
    rw = props.get("roi_width") or W
    rh = props.get("roi_height") or W
    d = _dev(sino, torch_dev)
    if dtype == ufomi.F16:
        d = d.half()
    elif dtype == ufomi.BF16:
        d = d.bfloat16()
    out = torch.full((n_slices, rh, rw), float("nan"), device=torch_dev)
    ctx.bp_parallel(d, W, P, out, n_slices=n_slices, dtype=dtype, **props)
    return out.cpu().numpy()


@pytest.mark.parametrize("W,P", [(64, 37), (16, 8), (100, 180), (33, 5), (256, 256)])
def test_texture_vs_oracle(ctx, torch_dev, W, P):
    rng = np.random.default_rng(W * 1000 + P)
    sino = rng.standard_normal((P, W)).astype(np.float32)
    got = _bp(ctx, torch_dev, sino)[0]
    assert rel_l2(got, oracle.bp_texture_ref32(sino)) < 1e-5
    assert rel_l2(got, oracle.bp_texture_f64(sino)) < 1e-5


@pytest.mark.parametrize("props", [
    dict(axis_pos=20.25), dict(axis_pos=40.0, angle_step=0.013), dict(angle_offset=0.7),
    dict(roi_x=5, roi_y=9, roi_width=41, roi_height=23), dict(num_projections=60, offset=13),
    dict(roi_width=100, roi_height=70, axis_pos=31.5), dict(num_projections=45),
])
def test_texture_properties(ctx, torch_dev, props):
    rng = np.random.default_rng(42)
    sino = rng.standard_normal((45, 64)).astype(np.float32)
    got = _bp(ctx, torch_dev, sino, **props)[0]
    assert np.isfinite(got).all()
    assert rel_l2(got, oracle.bp_texture_ref32(sino, **props)) < 1e-5
    assert rel_l2(got, oracle.bp_texture_f64(sino, **props)) < 1e-5


@pytest.mark.parametrize("W,P,roi", [(1024, 128, 0), (64, 512, 0), (1024, 128, 96)])
def test_angle_split_vs_oracle(ctx, torch_dev, W, P, roi):
    """Launches of fewer than 1024 workgroups split the angles into partial volumes summed in a fixed
    order (64 to 8 x 64 angles each); 1024 tiles of one slice run unsplit.  Rows are smoothed noise
    (a Gaussian of sigma 6 px), like filtered projections: on white noise 1024 px wide the reference's
    own float32 detector coordinates (6e-5 px at u ~ 1000) put it 3e-5 from FP64."""
    raw = np.random.default_rng(W + P).standard_normal((P, W + 32))
    k = np.exp(-0.5 * (np.arange(-16, 17) / 6.0) ** 2)
    sino = np.stack([np.convolve(r, k, mode="valid")[:W] for r in raw]).astype(np.float32)
    props = dict(roi_width=roi, roi_height=roi, roi_x=W // 3, roi_y=W // 5) if roi else {}
    got = _bp(ctx, torch_dev, sino, **props)[0]
    assert rel_l2(got, oracle.bp_texture_ref32(sino, **props)) < 1e-5
    assert rel_l2(got, oracle.bp_texture_f64(sino, **props)) < 1e-5


def test_multislice_bitwise(ctx, torch_dev):
    # 1, 2 and 4 slices per workgroup run the same per-slice arithmetic
    rng = np.random.default_rng(9)
    sinos = rng.standard_normal((7, 50, 48)).astype(np.float32)
    single = np.stack([_bp(ctx, torch_dev, s, axis_pos=23.5)[0] for s in sinos])
    for n in (2, 3, 4, 5, 7):
        got = _bp(ctx, torch_dev, sinos[:n], n_slices=n, axis_pos=23.5)
        np.testing.assert_array_equal(got, single[:n])


def test_constant_sinogram_closed_form(ctx, torch_dev):
    # A constant sinogram c backprojects to c * pi inside the inscribed circle.
    W, P = 128, 90
    got = _bp(ctx, torch_dev, np.full((P, W), 2.0, np.float32))[0]
    y, x = np.mgrid[0:W, 0:W] + 0.5 - W / 2
    inside = np.hypot(x, y) < W / 2 - 2
    np.testing.assert_allclose(got[inside], 2.0 * math.pi, rtol=2e-6)


def _lerp0(row, u):
    """Textbook linear interpolation of row at positions u with a zero border."""
    i0 = np.floor(u).astype(int)
    f = u - i0
    pad = np.concatenate([[0.0], row.astype(np.float64), [0.0]])
    t0 = pad[np.clip(i0 + 1, 0, len(pad) - 1)] * ((i0 >= 0) & (i0 < len(row)))
    t1 = pad[np.clip(i0 + 2, 0, len(pad) - 1)] * ((i0 + 1 >= 0) & (i0 + 1 < len(row)))
    return (1 - f) * t0 + f * t1


def test_theta_zero_reads_pixels(ctx, torch_dev):
    # One projection at angle 0: h = x + 0.5, i.e. image column x reads detector pixel x.
    W = 40
    row = np.random.default_rng(1).standard_normal(W).astype(np.float32)
    got = _bp(ctx, torch_dev, row[None, :], axis_pos=17.3)[0]
    np.testing.assert_allclose(got, np.broadcast_to(row * math.pi, got.shape), rtol=1e-5, atol=1e-5)


def test_theta_pi_is_linear_interpolation(ctx, torch_dev):
    # One projection at angle pi: h = 2a - x - 0.5, a fractional 1-D linear
    # interpolation of the detector row at u = 2a - x - 1 with a zero border.
    W, axis = 40, 17.3
    row = np.random.default_rng(4).standard_normal(W).astype(np.float32)
    got = _bp(ctx, torch_dev, row[None, :], axis_pos=axis, angle_offset=math.pi, num_projections=1)[0]
    u = 2 * axis - np.arange(W) - 1.0
    ref = math.pi * _lerp0(row, u)
    np.testing.assert_allclose(got, np.broadcast_to(ref, got.shape), rtol=1e-4, atol=2e-5)


def test_nearest_vs_oracle(ctx, torch_dev):
    rng = np.random.default_rng(5)
    sino = rng.standard_normal((30, 64)).astype(np.float32)
    got = _bp(ctx, torch_dev, sino, mode="nearest")[0]
    ref = oracle.bp_nearest_ref32(sino)
    assert rel_l2(got, ref) < 1e-6


@pytest.mark.parametrize("slices,chunk", [(1, 8), (2, 8), (4, 8), (8, 8), (8, 16)])
def test_tile_configurations_bitwise(ctx, torch_dev, slices, chunk):
    """Slices per tile and angles per LDS chunk change the schedule, not the arithmetic."""
    rng = np.random.default_rng(12)
    sinos = rng.standard_normal((11, 70, 64)).astype(np.float32)
    ref = _bp(ctx, torch_dev, sinos, n_slices=11, axis_pos=31.75)
    try:
        ctx.bp_configure(0, 0, -1)  # bf16 reference with float staging (16-bit staging rounds more)
        ref16 = _bp(ctx, torch_dev, sinos, n_slices=11, axis_pos=31.75, dtype=ufomi.BF16)
        ctx.bp_configure(slices, chunk, -1)
        got = _bp(ctx, torch_dev, sinos, n_slices=11, axis_pos=31.75)
        got16 = _bp(ctx, torch_dev, sinos, n_slices=11, axis_pos=31.75, dtype=ufomi.BF16)
    finally:
        ctx.bp_configure(0, 0, 0)
    np.testing.assert_array_equal(got, ref)
    np.testing.assert_array_equal(got16, ref16)


@pytest.mark.parametrize("dtype,tol", [(ufomi.F16, 1e-3)])
@pytest.mark.parametrize("slices", [4, 8])
def test_half_staging(ctx, torch_dev, slices, dtype, tol):
    """f16 storage with half (v, d) staging and v_dot2_f32_f16 interpolation: one more f16
    rounding of the sample difference and the weight, still within the f16 bar, and close to
    the float-staged result."""
    rng = np.random.default_rng(13)
    sinos = rng.standard_normal((9, 64, 80)).astype(np.float32)
    ref = np.stack([oracle.bp_texture_f64(s) for s in sinos])
    try:
        ctx.bp_configure(slices, 8, 1)
        got = _bp(ctx, torch_dev, sinos, n_slices=9, dtype=dtype)
        ctx.bp_configure(slices, 8, -1)
        plain = _bp(ctx, torch_dev, sinos, n_slices=9, dtype=dtype)
    finally:
        ctx.bp_configure(0, 0, 0)
    assert rel_l2(got, ref) < tol
    assert rel_l2(got, plain) < tol / 2
    with pytest.raises(ufomi.UfomiError):
        ctx.bp_configure(3, 0, 0)
    with pytest.raises(ufomi.UfomiError):
        ctx.bp_configure(0, 32, 0)


@pytest.mark.parametrize("dtype,tol", [(ufomi.F16, 1e-3), (ufomi.BF16, 8e-3)])
def test_low_precision_storage(ctx, torch_dev, dtype, tol):
    rng = np.random.default_rng(2)
    sinos = rng.standard_normal((4, 64, 80)).astype(np.float32)
    got = _bp(ctx, torch_dev, sinos, n_slices=4, dtype=dtype)
    ref = np.stack([oracle.bp_texture_f64(s) for s in sinos])
    assert rel_l2(got, ref) < tol


def test_config0_fbp_pipeline(ctx, torch_dev):
    """512^2 phantom of 10 ellipses, 512 angles, 1% noise, Ram-Lak (ramp-fromreal)
    zero padded to 1024, axis 256 (edge-based; detector-centre index 255.5)."""
    import torch
    W = P = 512
    N = 1024
    axis = 256.0
    rng = np.random.default_rng(1000)
    ell = phantoms.random_ellipses(rng, 10, W / 2)
    sino = phantoms.sinogram(ell, W, P, axis, math.pi / P)
    sino += rng.standard_normal(sino.shape) * 0.01 * np.abs(sino).max()
    sino = sino.astype(np.float32)

    coef = ufomi.filter_coefficients(2 * N)
    filt = torch.empty((P, W), device=torch_dev)
    ctx.ramp_filter(_dev(sino, torch_dev), W, P, N, _dev(coef, torch_dev), W, filt)
    out = torch.empty((1, W, W), device=torch_dev)
    ctx.bp_parallel(filt, W, P, out, axis_pos=axis)
    got = out.cpu().numpy()[0]

    f64 = oracle.ramp_chain(sino, N, oracle.filter_coefficients(2 * N), W)
    ref32 = oracle.bp_texture_ref32(f64.astype(np.float32), axis_pos=axis)
    ref64 = oracle.bp_texture_f64(f64, axis_pos=axis)
    assert rel_l2(got, ref32) < 1e-5
    assert rel_l2(got, ref64) < 1e-5
    # FBP approaches the phantom (an RMSE bound, not a parity check)
    img = phantoms.image(ell, W, axis)
    inside = np.hypot(*np.mgrid[0:W, 0:W] - W / 2 + 0.5) < W / 2 - 8
    rmse = np.sqrt(np.mean((got - img)[inside] ** 2))
    assert rmse < 0.1 * np.abs(img).max()


def test_errors(ctx, torch_dev):
    import torch
    d = torch.zeros((10, 16), device=torch_dev)
    out = torch.zeros((1, 16, 16), device=torch_dev)
    with pytest.raises(ufomi.UfomiError) as e:
        ctx.bp_parallel(d, 16, 10, out, num_projections=5)
    assert e.value.code == -5
    with pytest.raises(ufomi.UfomiError) as e:
        ctx.bp_parallel(d, 16, 10, out, num_projections=12, offset=4)
    assert e.value.code == -1
    with pytest.raises(ufomi.UfomiError) as e:
        ctx.bp_parallel(d.half(), 16, 10, out, dtype=ufomi.F16, mode="nearest")
    assert e.value.code == -4
