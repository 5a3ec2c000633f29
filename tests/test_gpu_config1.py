"""GPU parity on the batched parallel-beam workload's own shapes.

Eight sinograms of 1800 angles x 2048 px (the middle slices of the 2048-slice
ellipsoid phantom the benchmark reconstructs), filtered at N = 4096 with the
Shepp-Logan window and with ramp-fromreal, backprojected by the default
8-slice tile of bp_texture_kernel (spans inside the detector take the
unmasked branch; 4096 tiles run unsplit).  A second case runs 8 slices of
512 px, which splits the angles into partial volumes.

Bar: filtered sinogram within 1e-6 of the FP64 ramp chain relative to the
input rows' norm (5e-5 relative to the filtered rows' own); f32 volume within
1e-5 relative L2 of the oracle's float32 reference semantics
(backproject.cl:50-94 on the float32-rounded FP64 filter output) and of the
FP64 oracle; f16 storage within 1e-3 and bf16 within 8e-3 of FP64.
"""
import math

import numpy as np
import pytest

import oracle
import ufomi
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _sinograms(ctx, torch_dev, S, W, P, first):
    """Sinograms of slices first .. first + S - 1 of a W-slice phantom of 20 seeded ellipsoids."""
    import torch
    ellipsoids = ufomi.phantom_ellipsoids(1001, 20, W / 2.0)
    ells = np.stack([ufomi.slice_ellipsoids(ellipsoids, (first + k) + 0.5 - W / 2.0) for k in range(S)])
    sino = torch.empty((S, P, W), device=torch_dev)
    ctx.sinogram_ellipses(ells, W, P, W / 2.0, math.pi / P, sino)
    return sino


def _fbp(ctx, torch_dev, sino, N, filt_name, storage):
    import torch
    S, P, W = sino.shape
    dtype = {"f32": ufomi.F32, "f16": ufomi.F16, "bf16": ufomi.BF16}[storage]
    tdtype = {"f32": torch.float32, "f16": torch.float16, "bf16": torch.bfloat16}[storage]
    coef = torch.from_numpy(ufomi.filter_coefficients(2 * N, filt_name)).to(torch_dev)
    filt = torch.empty((S, P, W), dtype=tdtype, device=torch_dev)
    ctx.ramp_filter(sino, W, S * P, N, coef, W, filt, dtype)
    vol = torch.full((S, W, W), float("nan"), device=torch_dev)
    ctx.bp_parallel(filt, W, P, vol, n_slices=S, dtype=dtype, axis_pos=W / 2.0)
    torch.cuda.synchronize()
    return filt.float().cpu().numpy(), vol.cpu().numpy()


@pytest.mark.parametrize("filt_name", ["shepp-logan", "ramp-fromreal"])
def test_config1_shapes_f32(ctx, torch_dev, filt_name):
    S, P, W, N = 8, 1800, 2048, 4096
    sino = _sinograms(ctx, torch_dev, S, W, P, W // 2 - S // 2)
    filt, vol = _fbp(ctx, torch_dev, sino, N, filt_name, "f32")
    assert np.isfinite(vol).all()
    k = 3  # one slice against both oracles (about 10 s of CPU each)
    s = sino[k].cpu().numpy()
    f64 = oracle.ramp_chain(s, N, oracle.filter_coefficients(2 * N, filt_name), W)
    # float32 FFT rounding scales with the input; the ramp removes most of the smooth sinogram's
    # energy, so against the small filtered rows the same rounding reads about 10x larger
    assert rel_l2(filt[k], f64) < 5e-5
    assert np.linalg.norm(filt[k] - f64) / np.linalg.norm(s) < 1e-6
    ref32 = oracle.bp_texture_ref32(f64.astype(np.float32), axis_pos=W / 2.0)
    ref64 = oracle.bp_texture_f64(f64, axis_pos=W / 2.0)
    err32, err64 = rel_l2(vol[k], ref32), rel_l2(vol[k], ref64)
    print("config-1 shapes %s f32: rel-L2 %.3g (ref32), %.3g (f64)" % (filt_name, err32, err64))
    assert err32 < 1e-5
    assert err64 < 1e-5


@pytest.mark.parametrize("storage,tol", [("f16", 1e-3), ("bf16", 8e-3)])
def test_config1_shapes_half_storage(ctx, torch_dev, storage, tol):
    S, P, W, N = 8, 1800, 2048, 4096
    sino = _sinograms(ctx, torch_dev, S, W, P, W // 2 - S // 2)
    _, vol = _fbp(ctx, torch_dev, sino, N, "shepp-logan", storage)
    _, vol32 = _fbp(ctx, torch_dev, sino, N, "shepp-logan", "f32")
    k = 5
    s = sino[k].cpu().numpy()
    f64 = oracle.ramp_chain(s, N, oracle.filter_coefficients(2 * N, "shepp-logan"), W)
    ref64 = oracle.bp_texture_f64(f64, axis_pos=W / 2.0)
    err = rel_l2(vol[k], ref64)
    print("config-1 shapes shepp-logan %s: rel-L2 %.3g (f64)" % (storage, err))
    assert err < tol
    # every slice of the batch against the f32-storage run of the same slices
    assert rel_l2(vol, vol32) < tol


@pytest.mark.parametrize("filt_name", ["shepp-logan", "ramp-fromreal"])
def test_eight_slices_split_angles(ctx, torch_dev, filt_name):
    # 8 slices of 512 px are 256 tiles, so the angles split into partial volumes
    S, P, W, N = 8, 900, 512, 1024
    sino = _sinograms(ctx, torch_dev, S, W, P, W // 2 - S // 2)
    filt, vol = _fbp(ctx, torch_dev, sino, N, filt_name, "f32")
    s_all = sino.cpu().numpy()
    coef = oracle.filter_coefficients(2 * N, filt_name)
    for k in range(S):
        f64 = oracle.ramp_chain(s_all[k], N, coef, W)
        assert rel_l2(filt[k], f64) < 5e-5
        assert np.linalg.norm(filt[k] - f64) / np.linalg.norm(s_all[k]) < 1e-6
        assert rel_l2(vol[k], oracle.bp_texture_ref32(f64.astype(np.float32), axis_pos=W / 2.0)) < 1e-5
        assert rel_l2(vol[k], oracle.bp_texture_f64(f64, axis_pos=W / 2.0)) < 1e-5
