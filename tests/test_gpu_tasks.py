"""The task layer on the GPU, driven the way the reference's own tests drive
ufo-core (tests/test_fft.py:57-112): memory-in -> tasks -> memory-out, plus
the batch pipeline runner with the fused ramp filter.  Parity bar: 1e-5
relative L2 against the oracle; the FFT checks keep the reference's own
decimal=2 (forward) and decimal=5 (round trip) assertions.
"""
import math

import numpy as np
import pytest

import oracle
import ufo_tasks
import ufomi
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def run_chain(ctx, array, *tasks):
    pm = ufo_tasks.PluginManager()
    src, sink = ufo_tasks.MemoryIn(array), pm.get_task("memory-out")
    g = ufo_tasks.TaskGraph()
    nodes = [src, *tasks, sink]
    for a, b in zip(nodes, nodes[1:]):
        g.connect_nodes(a, b)
    ufo_tasks.Scheduler(ctx).run(g)
    return sink.array


def make(name, **props):
    t = ufo_tasks.PluginManager().get_task(name)
    for k, v in props.items():
        setattr(t.props, k, v)
    return t


# shapes of tests/test_fft.py:136-180 restricted to 1-D transforms
@pytest.mark.parametrize("shape,size_x", [((32, 32, 32), 0), ((32, 16, 8), 0), ((8, 16, 32), 0), ((9, 17, 33), 64),
                                          ((9, 17, 31), 32)])
def test_fft_forward_like_reference(ctx, shape, size_x):
    depth, height, width = shape
    x = np.linspace(-1, 2, num=width * height * depth, dtype=np.float32).reshape(depth, height, width)
    out = run_chain(ctx, x, make("fft", dimensions=1, size_x=size_x))
    n = size_x or width
    got = out[..., 0::2] + 1j * out[..., 1::2]
    ref = np.fft.fft(x, n=n, axis=-1).astype(np.complex64)
    np.testing.assert_almost_equal(got, ref, decimal=2)
    assert rel_l2(got, np.fft.fft(x.astype(np.float64), n=n, axis=-1)) < 2e-6


def test_auto_zeropadding(ctx):
    x = np.linspace(-1, 2, num=9 * 17 * 31, dtype=np.float32).reshape(9, 17, 31)
    out = run_chain(ctx, x, make("fft", auto_zeropadding=True))
    assert out.shape == (9, 17, 64)  # 2 * closest_smaller_pow2(31 - 1) = 32 complex values
    back = run_chain(ctx, x, make("fft", auto_zeropadding=True), make("ifft", crop_width=31))
    np.testing.assert_almost_equal(back, x, decimal=5)


def test_round_trip_crop_like_reference(ctx):
    x = np.linspace(-1, 2, num=9 * 17 * 33, dtype=np.float32).reshape(9, 17, 33)
    back = run_chain(ctx, x, make("fft", size_x=64), make("ifft", crop_width=45))
    padded = np.pad(x, ((0, 0), (0, 0), (0, 31)))[:, :, :45]
    np.testing.assert_almost_equal(back, padded, decimal=5)


def test_sizes_up_to_the_property_range(ctx):
    """size-x/y up to 32768 (the properties' range, ufo-fft-task.c:513-531) on 2-D transforms, chirp-z
    and powers of two, against numpy's FP64 transform."""
    x = np.linspace(-1, 2, 36, dtype=np.float32).reshape(1, 4, 9)
    for sx, sy in ((4097, 4), (16, 16384), (32768, 4), (20000, 5)):
        got = run_chain(ctx, x, make("fft", dimensions=2, size_x=sx, size_y=sy))
        pad = np.zeros((1, sy, sx))
        pad[:, :4, :9] = x
        ref = np.fft.fftn(pad, axes=(1, 2))
        assert rel_l2(got.view(np.complex64), ref) < 1e-5


def test_requisition_errors(ctx):
    x = np.ones((1, 4, 100), np.float32)
    with pytest.raises(ufo_tasks.TaskError) as e:
        run_chain(ctx, x, make("fft", size_x=64))
    assert (e.value.code, e.value.status) == (1, -6)
    assert "greater than or equal to input size" in str(e.value)
    with pytest.raises(ufo_tasks.TaskError) as e:           # 2N = 200 floats: not a power of two
        run_chain(ctx, x, make("fft", size_x=100), make("filter"))
    assert e.value.status in (-4, -3)
    with pytest.raises(ufo_tasks.TaskError) as e:
        run_chain(ctx, x, make("ifft"))                      # real layout into ifft
    assert (e.value.code, e.value.status) == (1, -11)
    assert "ifft input must be complex" in str(e.value)
    with pytest.raises(ufo_tasks.TaskError) as e:
        run_chain(ctx, np.ones((1, 10, 16), np.float32), make("backproject", num_projections=5))
    assert (e.value.code, e.value.status) == (1, -5)
    with pytest.raises(ufo_tasks.TaskError) as e:
        run_chain(ctx, np.ones((1, 4, 8), np.float32), make("general-backproject"))
    assert (e.value.code, e.value.status) == (0, -10)
    with pytest.raises(ufo_tasks.TaskError) as e:
        run_chain(ctx, np.ones((1, 4, 8), np.float32),
                  make("general-backproject", num_projections=4, store_type="uchar", gray_map_min=1.0, gray_map_max=1.0))
    assert (e.value.code, e.value.status) == (0, -8)


def _fbp_tasks(n, w, axis, filt="ramp-fromreal"):
    return [make("fft", dimensions=1, size_x=n), make("filter", filter=filt), make("ifft", dimensions=1, crop_width=w),
            make("backproject", axis_pos=axis)]


@pytest.mark.parametrize("filt", ["ramp-fromreal", "hamming", "faris-byer"])
def test_fbp_chain_scheduler_pipeline_oracle(ctx, torch_dev, filt):
    import torch
    rng = np.random.default_rng(3)
    n_items, p, w, n, axis = 3, 60, 64, 128, 31.5
    sinos = rng.standard_normal((n_items, p, w)).astype(np.float32)
    sched = run_chain(ctx, sinos, *_fbp_tasks(n, w, axis, filt))
    pl = ufo_tasks.Pipeline("fft dimensions=1 size-x=%d ! filter filter=%s ! ifft dimensions=1 crop-width=%d ! "
                            "backproject axis-pos=%g" % (n, filt, w, axis))
    raw, n_out, shape = pl.run(ctx, torch.from_numpy(sinos).to(torch_dev))
    assert pl.used_fused_filter and n_out == n_items and shape == (w, w)
    fused = raw[:n_out * w * w * 4].view(torch.float32).reshape(n_out, w, w).cpu().numpy()
    coef = oracle.filter_coefficients(2 * n, filt)
    for i in range(n_items):
        filt_ref = oracle.ramp_chain(sinos[i], n, coef, w)
        ref = oracle.bp_texture_ref32(filt_ref.astype(np.float32), axis_pos=axis)
        assert rel_l2(sched[i], ref) < 1e-5
        assert rel_l2(fused[i], ref) < 1e-5
        assert rel_l2(oracle.bp_texture_f64(filt_ref, axis_pos=axis), ref) < 1e-5
    assert rel_l2(fused, sched) < 1e-5


def test_general_backproject_reductor_lifecycle(ctx, torch_dev):
    rng = np.random.default_rng(4)
    p, h, w = 12, 8, 16
    projs = rng.standard_normal((p, h, w)).astype(np.float32)
    geom = dict(num_projections=p, region=[-2.0, 2.0, 1.0], center_position_x=[8.0], center_position_z=[4.0],
                source_position_y=[-40.0], detector_position_y=[30.0])
    vol = run_chain(ctx, projs, make("general-backproject", **geom))
    assert vol.shape == (4, 16, 16)
    ref = oracle.gbp(projs, region=(-2.0, 2.0, 1.0), center_position_x=[8.0], center_position_z=[4.0],
                     source_position_y=[-40.0], detector_position_y=[30.0])
    assert rel_l2(vol, ref) < 1e-5
    # fewer projections than declared: generate emits nothing (ufo-general-backproject-task.c:1733-1738)
    t = make("general-backproject", **dict(geom, num_projections=p + 1))
    assert run_chain(ctx, projs, t).size == 0
    assert "received only 12 projections out of 13" in t.last_error()


def test_fdk_pipeline_vs_oracle(ctx, torch_dev):
    """cone-beam-projection-weight ! fft ! filter ! ifft ! general-backproject: the FDK chain of the
    reference's cone-beam examples."""
    import torch
    rng = np.random.default_rng(5)
    p, h, w, n = 24, 12, 24, 64
    projs = rng.standard_normal((p, h, w)).astype(np.float32)
    sod, odd, cx, cz = 80.0, 60.0, 12.0, 6.0
    desc = ("cone-beam-projection-weight center-position-x=%g center-position-z=%g source-distance=%g "
            "detector-distance=%g ! fft dimensions=1 size-x=%d ! filter ! ifft dimensions=1 crop-width=%d ! "
            "general-backproject num-projections=%d center-position-x=%g center-position-z=%g "
            "source-position-y=%g detector-position-y=%g region=-3,3,1"
            % (cx, cz, sod, odd, n, w, p, cx, cz, -sod, odd))
    pl = ufo_tasks.Pipeline(desc)
    raw, n_out, shape = pl.run(ctx, torch.from_numpy(projs).to(torch_dev))
    assert n_out == 6 and shape == (w, w)
    vol = raw[:n_out * w * w * 4].view(torch.float32).reshape(n_out, w, w).cpu().numpy()
    coef = oracle.filter_coefficients(2 * n)
    weighted = np.stack([oracle.cone_weight_ref32(q, cx, cz, sod, odd) for q in projs])
    filt = np.stack([oracle.ramp_chain(q, n, coef, w) for q in weighted]).astype(np.float32)
    ref = oracle.gbp(filt, region=(-3.0, 3.0, 1.0), center_position_x=[cx], center_position_z=[cz],
                     source_position_y=[-sod], detector_position_y=[odd])
    assert rel_l2(vol, ref) < 1e-5


def test_read_fbp_write_pipeline(ctx, torch_dev, tmp_path):
    """read ! fft ! filter ! ifft ! backproject ! write (docs/examples.rst FBP from files):
    the written slices equal the in-memory pipeline's bit for bit and the oracle to 1e-5."""
    import torch
    from PIL import Image
    rng = np.random.default_rng(8)
    n_items, p, w, n, axis = 4, 48, 64, 128, 31.5
    sinos = rng.standard_normal((n_items, p, w)).astype(np.float32)
    ims = [Image.fromarray(s) for s in sinos]
    ims[0].save(tmp_path / "sino-a.tif", save_all=True, append_images=ims[1:3])
    ims[3].save(tmp_path / "sino-b.tif")
    chain = ("fft dimensions=1 size-x=%d ! filter ! ifft dimensions=1 crop-width=%d ! backproject axis-pos=%g"
             % (n, w, axis))
    pl = ufo_tasks.Pipeline("read path=%s/sino-*.tif ! %s ! write filename=%s/slice-%%03i.tif bytes-per-file=0 "
                            "tiff-bigtiff=false" % (tmp_path, chain, tmp_path))
    raw, n_out, shape = pl.run(ctx)
    assert raw is None and n_out == n_items and shape == (w, w) and pl.used_fused_filter
    mem = ufo_tasks.Pipeline(chain)
    raw2, n2, _ = mem.run(ctx, torch.from_numpy(sinos).to(torch_dev))
    ref_mem = raw2[:n2 * w * w * 4].view(torch.float32).reshape(n2, w, w).cpu().numpy()
    coef = oracle.filter_coefficients(2 * n)
    for i in range(n_items):
        got = np.asarray(Image.open(tmp_path / ("slice-%03d.tif" % i)))
        np.testing.assert_array_equal(got, ref_mem[i])
        ref = oracle.bp_texture_ref32(oracle.ramp_chain(sinos[i], n, coef, w).astype(np.float32), axis_pos=axis)
        assert rel_l2(got, ref) < 1e-5


def test_scheduler_read_write_tasks(ctx, torch_dev, tmp_path):
    """The item-by-item scheduler with a read generator and a write sink."""
    from PIL import Image
    rng = np.random.default_rng(9)
    p, w, axis = 40, 48, 23.5
    sinos = rng.standard_normal((3, p, w)).astype(np.float32)
    for i, s in enumerate(sinos):
        s.tofile(tmp_path / ("s%d.raw" % i))
    src = make("read", path=str(tmp_path / "s*.raw"), raw_width=w, raw_height=p, raw_bitdepth=32)
    sink = make("write", filename=str(tmp_path / "out.tif"), tiff_bigtiff=False)
    g = ufo_tasks.TaskGraph()
    bp = make("backproject", axis_pos=axis)
    g.connect_nodes(src, bp)
    g.connect_nodes(bp, sink)
    ufo_tasks.Scheduler(ctx).run(g)
    del sink
    im = Image.open(tmp_path / "out.tif")
    assert im.n_frames == 3
    for i in range(3):
        im.seek(i)
        ref = oracle.bp_texture_ref32(sinos[i], axis_pos=axis)
        assert rel_l2(np.asarray(im), ref) < 1e-5


def test_wide_detector_fbp_pipeline(ctx, torch_dev):
    """A detector wider than 4096 px: `fft size-x=16384 ! filter ! ifft ! backproject` as one
    ufo-launch description runs the unfused chain (four-step transforms) and then backprojects
    a 64 x 64 region of interest; equal to the FP64 oracle's ramp chain and backprojection."""
    import torch
    w, p, n = 8200, 16, 16384
    rng = np.random.default_rng(17)
    raw = rng.standard_normal((p, w + 32))
    k = np.exp(-0.5 * (np.arange(-16, 17) / 6.0) ** 2)
    sino = np.stack([np.convolve(r, k, mode="valid")[:w] for r in raw]).astype(np.float32)
    desc = ("fft dimensions=1 size-x=%d ! filter ! ifft dimensions=1 crop-width=%d ! "
            "backproject axis-pos=4100.5 roi-x=4068 roi-y=4068 roi-width=64 roi-height=64" % (n, w))
    raw_out, n_out, shape = ufo_tasks.Pipeline(desc).run(ctx, torch.from_numpy(sino[None]).to(torch_dev))
    assert n_out == 1 and shape == (64, 64)
    got = raw_out[:64 * 64 * 4].view(torch.float32).reshape(64, 64).cpu().numpy()
    filt = oracle.ramp_chain(sino, n, oracle.filter_coefficients(2 * n, "ramp-fromreal"), w)
    ref = oracle.bp_texture_f64(filt, axis_pos=4100.5, roi_x=4068, roi_y=4068, roi_width=64, roi_height=64)
    assert rel_l2(got, ref) < 1e-5


def test_2d_fft_filter_ifft_pipeline(ctx, torch_dev):
    """fft dimensions=2 ! filter ! ifft dimensions=2 as one description (unfused): the filter
    multiplies every row of the 2-D spectrum by its coefficients; against numpy in float64."""
    import torch
    h, w, n = 12, 30, 32
    x = np.random.default_rng(23).standard_normal((h, w)).astype(np.float32)
    desc = ("fft dimensions=2 size-x=%d size-y=16 ! filter filter=butterworth ! ifft dimensions=2 crop-width=%d "
            "crop-height=%d" % (n, w, h))
    raw, n_out, shape = ufo_tasks.Pipeline(desc).run(ctx, torch.from_numpy(x[None]).to(torch_dev))
    assert n_out == 1 and shape == (h, w)
    got = raw[:w * h * 4].view(torch.float32).reshape(h, w).cpu().numpy()
    coef = oracle.filter_coefficients(2 * n, "butterworth").astype(np.float64)
    spec = np.fft.fft2(np.pad(x.astype(np.float64), ((0, 16 - h), (0, n - w))))
    spec = spec.real * coef[0::2][None, :] + 1j * spec.imag * coef[1::2][None, :]
    ref = np.fft.ifft2(spec).real[:h, :w]
    assert rel_l2(got, ref) < 1e-5
