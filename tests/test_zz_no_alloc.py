"""GPU test of ufomi_bp_parallel_noalloc, the backprojection entry point that
allocates nothing, against ufomi_bp_parallel through the C ABI.

Both compute the same kernels on the same inputs; only the (sin, cos) table
(computed on the device instead of uploaded from the host) and the place of the
angle-split partial volumes differ, so the results agree to float rounding of
the table (bar: 1e-6 relative L2, and 1e-5 against the float32 oracle).  Two
calls in a row, the second on a smaller input, show that the shared scratch
keeps nothing from one call to the next.
"""
import numpy as np
import pytest

import oracle
import ufomi
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _sino(shape, seed):
    return np.random.default_rng(seed).standard_normal(shape).astype(np.float32)


def _run(ctx, torch_dev, fn, sino, dtype=ufomi.F32, **props):
    import torch
    n, P, W = sino.shape
    d = torch.from_numpy(sino).to(torch_dev)
    if dtype == ufomi.F16:
        d = d.half()
    rw = props.get("roi_width") or W
    rh = props.get("roi_height") or W
    out = torch.full((n, rh, rw), float("nan"), device=torch_dev)
    fn(d, W, P, out, n_slices=n, dtype=dtype, **props)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_noalloc_matches_bp_parallel_two_calls(ctx, torch_dev):
    # call 1: one 512^2 slice of 512 angles (256 tiles: split into 4 angle parts); a positive
    # sinogram, so the sums do not cancel and float32 rounding stays far below the 1e-5 bar
    big = _sino((1, 512, 512), 11) + np.float32(2.0)
    got1 = _run(ctx, torch_dev, ctx.bp_parallel_noalloc, big)
    # call 2, smaller: 3 slices of 200 px, 180 angles, f16 storage, an ROI and an axis
    small = _sino((3, 180, 200), 12)
    props = dict(axis_pos=97.25, roi_x=3, roi_y=5, roi_width=150, roi_height=170, angle_offset=0.3)
    got2 = _run(ctx, torch_dev, ctx.bp_parallel_noalloc, small, dtype=ufomi.F16, **props)

    ref1 = _run(ctx, torch_dev, ctx.bp_parallel, big)
    ref2 = _run(ctx, torch_dev, ctx.bp_parallel, small, dtype=ufomi.F16, **props)
    assert np.isfinite(got1).all() and np.isfinite(got2).all()
    assert rel_l2(got1, ref1) < 1e-6
    assert rel_l2(got2, ref2) < 1e-6
    assert rel_l2(got1[0], oracle.bp_texture_ref32(big[0])) < 1e-5


def test_noalloc_unsplit_and_nearest(ctx, torch_dev):
    # 8 slices of 256 px: 64 tiles in one slice group, split in 4; then nearest mode
    s8 = _sino((8, 300, 256), 13)
    got = _run(ctx, torch_dev, ctx.bp_parallel_noalloc, s8)
    ref = _run(ctx, torch_dev, ctx.bp_parallel, s8)
    assert rel_l2(got, ref) < 1e-6
    # 1024^2 from 64 angles: 1024 tiles, no split (only the table is used)
    s1 = _sino((1, 64, 1024), 14)
    got = _run(ctx, torch_dev, ctx.bp_parallel_noalloc, s1)
    ref = _run(ctx, torch_dev, ctx.bp_parallel, s1)
    assert rel_l2(got, ref) < 1e-6
    sn = _sino((1, 90, 128), 15)
    got = _run(ctx, torch_dev, ctx.bp_parallel_noalloc, sn, mode="nearest")
    ref = _run(ctx, torch_dev, ctx.bp_parallel, sn, mode="nearest")
    assert rel_l2(got, ref) < 1e-6


def test_noalloc_half_staging_split(ctx, torch_dev):
    # 8 f16 slices of 256 px: half-precision (v, d) staging, 64 tiles split into 4 angle parts
    s8 = _sino((8, 300, 256), 16)
    got = _run(ctx, torch_dev, ctx.bp_parallel_noalloc, s8, dtype=ufomi.F16)
    ref = _run(ctx, torch_dev, ctx.bp_parallel, s8, dtype=ufomi.F16)
    assert rel_l2(got, ref) < 1e-6


def test_noalloc_rejects_too_many_angles(ctx, torch_dev):
    import torch
    # refused on the host before any launch: the table holds 32768 angles
    d = torch.zeros((32769, 4), device=torch_dev)
    out = torch.zeros((1, 4, 4), device=torch_dev)
    with pytest.raises(ufomi.UfomiError) as e:
        ctx.bp_parallel_noalloc(d, 4, 32769, out)
    assert e.value.code == -6  # UFOMI_ERR_SIZE
    assert "32768" in str(e.value)
