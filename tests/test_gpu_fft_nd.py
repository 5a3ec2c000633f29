"""fft / ifft tasks with the reference's full semantics on the GPU: every case
of tests/test_fft.py:136-180 (1-D, 2-D and 3-D, power-of-two and chirp-z
sizes, user sizes, auto zero-padding, crops), driven like the reference's
main(): a stacked 3-D item (or 2-D items) through fft (and ifft) of the task
layer.  Bar: the reference's own assertions (decimal=2 forward, decimal=5
round trip) and 1e-5 relative L2 against the FP64 oracle."""
import numpy as np
import pytest

import oracle
import ufo_tasks
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def next_pow2(n):
    return 1 << int(np.ceil(np.log2(n)))


def _run(ctx, items, *tasks):
    src, sink = ufo_tasks.MemoryIn(items), ufo_tasks.MemoryOut()
    g = ufo_tasks.TaskGraph()
    nodes = [src, *tasks, sink]
    for a, b in zip(nodes, nodes[1:]):
        g.connect_nodes(a, b)
    ufo_tasks.Scheduler(ctx).run(g)
    return sink.array


def _task(name, **props):
    t = ufo_tasks.Task(name)
    for k, v in props.items():
        setattr(t.props, k, v)
    return t


def main(ctx, input_shape, output_shape=None, crop_width=None, crop_height=None, dimensions=1, do_inverse=False,
         auto_zeropadding=False, stack=True):
    """tests/test_fft.py main(): the same inputs, task properties and checks."""
    if output_shape is None:
        if auto_zeropadding:
            output_shape = (next_pow2(input_shape[0]) if dimensions == 3 else input_shape[0],
                            next_pow2(input_shape[1]) if dimensions >= 2 else input_shape[1],
                            next_pow2(input_shape[2]))
        else:
            output_shape = input_shape
    depth, height, width = input_shape
    x = np.linspace(-1, 2, num=width * height * depth, dtype=np.float32).reshape(depth, height, width)
    fft = _task("fft", dimensions=dimensions, auto_zeropadding=auto_zeropadding, size_x=output_shape[2],
                size_y=output_shape[1], size_z=output_shape[0])
    tasks = [fft]
    if do_inverse:
        props = {"dimensions": dimensions}
        if crop_width:
            props["crop_width"] = crop_width
        if crop_height:
            props["crop_height"] = crop_height
        tasks.append(_task("ifft", **props))
    items = x[None] if stack else x
    out = _run(ctx, items, *tasks)
    out = out[0] if stack else out
    m = (output_shape[2], output_shape[1], output_shape[0])
    if do_inverse:
        pad = np.zeros(output_shape, np.float32)
        pad[:depth, :height, :width] = x
        pad = pad[:, :crop_height, :crop_width]
        np.testing.assert_almost_equal(out, pad, decimal=5)
        assert rel_l2(out, pad) < 1e-5
    else:
        got = out.view(np.complex64) if stack else out.view(np.complex64)
        if stack:
            ref = oracle.fft_nd(x, dimensions, m)
        else:
            ref = np.stack([oracle.fft_nd(item, dimensions, (m[0], m[1], 1))[0] for item in x])
        np.testing.assert_almost_equal(got, ref.astype(np.complex64), decimal=2)
        assert rel_l2(got, ref) < 1e-5, rel_l2(got, ref)


INPUT_SHAPES = [(32, 32, 32), (32, 16, 8), (8, 16, 32), (33, 17, 9), (9, 17, 33), (9, 17, 31), (33, 33, 33)]


@pytest.mark.parametrize("shape", INPUT_SHAPES)
def test_automatic_size_1d(ctx, torch_dev, shape):
    main(ctx, shape, dimensions=1)


@pytest.mark.parametrize("do_inverse", [False, True])
@pytest.mark.parametrize("out_shape,dims", [((11, 25, 50), 3), ((9, 25, 50), 2), ((9, 17, 50), 1)])
def test_user_size(ctx, torch_dev, out_shape, dims, do_inverse):
    main(ctx, (9, 17, 33), output_shape=out_shape, dimensions=dims, do_inverse=do_inverse)


def test_larger_than_chirp(ctx, torch_dev):
    main(ctx, (9, 17, 33), output_shape=(32, 128, 128), dimensions=3)


@pytest.mark.parametrize("do_inverse", [False, True])
@pytest.mark.parametrize("dims", [1, 2, 3])
def test_auto_zeropadding(ctx, torch_dev, dims, do_inverse):
    main(ctx, (9, 17, 31), auto_zeropadding=True, dimensions=dims, do_inverse=do_inverse)


@pytest.mark.parametrize("kw", [
    dict(output_shape=(11, 25, 50), dimensions=3, crop_width=45, crop_height=10),
    dict(dimensions=3, crop_width=25, crop_height=10),
    dict(output_shape=(9, 25, 50), dimensions=2, crop_width=45, crop_height=10),
    dict(dimensions=2, crop_width=25, crop_height=10),
    dict(output_shape=(9, 17, 50), dimensions=1, crop_width=45),
    dict(dimensions=1, crop_width=25),
])
def test_crops(ctx, torch_dev, kw):
    main(ctx, (9, 17, 33), do_inverse=True, **kw)


def test_unstacked_2d(ctx, torch_dev):
    main(ctx, (32, 32, 10), dimensions=2, do_inverse=True, auto_zeropadding=False, stack=False)


@pytest.mark.parametrize("shape,dims,m", [((2, 40, 3000), 1, (3000, 40, 2)), ((1, 300, 200), 2, (512, 300, 1)),
                                          ((3, 64, 4096), 1, (4096, 64, 3)), ((1, 2, 8192), 1, (8192, 2, 1)),
                                          ((5, 7, 1), 2, (1, 7, 5)),
                                          # axes longer than one LDS pass in 2-D and 3-D transforms
                                          ((1, 4, 9), 2, (4097, 4, 1)), ((1, 4, 9), 2, (16, 16384, 1)),
                                          ((1, 4, 9), 2, (16384, 4, 1)), ((3, 5, 7), 3, (16, 8, 9000)),
                                          ((2, 3, 20), 2, (32, 20000, 3)), ((2, 3, 10), 1, (16384, 5, 2)),
                                          ((2, 3, 5), 3, (10000, 4, 3)), ((1, 4, 9), 2, (32768, 4, 1))])
def test_large_and_edge_sizes(ctx, torch_dev, shape, dims, m):
    """The longest LDS chirp-z (4096 needs an 8192-point convolution), the longest LDS power of two,
    size 1, and axes past one LDS pass (gathered into rows, four-step split or Bluestein over it,
    scattered back) on every axis, up to the task's 32768."""
    x = np.random.default_rng(sum(shape)).standard_normal(shape).astype(np.float32)
    fft = _task("fft", dimensions=dims, size_x=m[0], size_y=m[1], size_z=m[2])
    got = _run(ctx, x[None], fft)[0].view(np.complex64)
    pad = np.zeros((m[2], m[1], m[0]), np.float32)
    pad[:shape[0], :shape[1], :shape[2]] = x
    # numpy's FP64 FFT (the oracle's DFT agrees with it, tests/test_oracle.py) at these sizes
    ref = np.fft.fftn(pad.astype(np.float64), axes=tuple(range(2, 2 - dims, -1)))
    assert rel_l2(got, ref) < 1e-5, rel_l2(got, ref)
    back = _run(ctx, x[None], _task("fft", dimensions=dims, size_x=m[0], size_y=m[1], size_z=m[2]),
                _task("ifft", dimensions=dims))[0]
    assert rel_l2(back, pad) < 1e-5
    if dims >= 2:  # the inverse's cropped last pass
        cw, ch = max(1, m[0] // 3), max(1, m[1] // 2)
        crop = _run(ctx, x[None], _task("fft", dimensions=dims, size_x=m[0], size_y=m[1], size_z=m[2]),
                    _task("ifft", dimensions=dims, crop_width=cw, crop_height=ch))[0]
        assert rel_l2(crop, pad[:, :ch, :cw]) < 1e-5


def test_complex_input_and_long_axes(ctx, torch_dev):
    """A complex item goes through the forward transform unchanged in size, also along an axis
    longer than one LDS pass."""
    x = np.random.default_rng(2).standard_normal((1, 6, 10)).astype(np.float32)
    spec = _run(ctx, x[None], _task("fft", dimensions=1, size_x=12))[0]
    assert spec.shape == (1, 6, 24)
    buf_in = spec.view(np.complex64)
    import torch
    dev = torch.device("cuda:0")
    xin = torch.from_numpy(spec).to(dev)

    def run_complex(t):
        # the spectrum again through fft: a complex-layout buffer, size taken from the interleaved width
        b = ufo_tasks.Buffer()
        b.data = xin.data_ptr()
        b.req.n_dims = 3
        b.req.dims[0], b.req.dims[1], b.req.dims[2] = 24, 6, 1
        b.layout = ufo_tasks.LAYOUT_COMPLEX
        b.dtype = 0
        t.setup(ctx)
        req = t.get_requisition(b)
        out = torch.empty((req.dims[2], req.dims[1], req.dims[0]), device=dev)
        o = ufo_tasks.Buffer()
        o.data = out.data_ptr()
        o.req = req
        t.process(b, o, req)
        torch.cuda.synchronize()
        return (req.dims[0], req.dims[1], req.dims[2]), out.cpu().numpy().view(np.complex64)

    dims, got = run_complex(_task("fft", dimensions=1))
    assert dims == (24, 6, 1)
    assert rel_l2(got, oracle.fft_nd(buf_in, 1, (12, 6, 1), in_complex=True)) < 1e-5
    dims, got = run_complex(_task("fft", dimensions=2, size_x=5000))
    assert dims == (10000, 6, 1)
    assert rel_l2(got, oracle.fft_nd(buf_in, 2, (5000, 6, 1), in_complex=True)) < 1e-5
    # the same items with the real layout: 24 real samples per row
    got_r = _run(ctx, spec[None], _task("fft", dimensions=2, size_x=5000))[0]
    assert rel_l2(got_r.view(np.complex64), oracle.fft_nd(spec, 2, (5000, 6, 1))) < 1e-5


@pytest.mark.parametrize("width,size_x", [(10000, 16384), (16384, 16384), (20000, 32768), (9, 32768)])
def test_long_rows_four_step(ctx, width, size_x):
    """Power-of-two rows above 8192 points (two passes of the four-step split): forward against
    numpy.fft in float64, round trip with a crop back to the input."""
    rows = 3
    x = np.random.default_rng(width).standard_normal((rows, width)).astype(np.float32)
    spec = _run(ctx, [x], _task("fft", dimensions=1, size_x=size_x))[0]
    got = spec.reshape(rows, size_x, 2)
    got = got[..., 0] + 1j * got[..., 1]
    ref = np.fft.fft(x.astype(np.float64), n=size_x, axis=-1)
    assert rel_l2(got, ref) < 1e-5
    back = _run(ctx, [x], _task("fft", dimensions=1, size_x=size_x), _task("ifft", dimensions=1,
                                                                           crop_width=width))[0]
    assert back.shape == (rows, width)
    assert rel_l2(back, x) < 1e-5


def test_long_rows_complex_input_and_filter_chain(ctx, torch_dev):
    """Complex rows of 16384 points through the C ABI, and fft ! filter ! ifft at N = 16384 (not fused)
    against the oracle's FP64 ramp chain."""
    import torch
    rows, w, n = 4, 9000, 16384
    z = np.random.default_rng(5).standard_normal((rows, 2 * w)).astype(np.float32)
    out = torch.empty((rows, 2 * n), device=torch_dev)
    ctx.fft_nd(torch.from_numpy(z).to(torch_dev), (w, rows, 1), 1, (n, rows, 1), out, in_complex=True)
    torch.cuda.synchronize()
    got = out.cpu().numpy().reshape(rows, n, 2)
    zc = z[:, 0::2].astype(np.float64) + 1j * z[:, 1::2]
    assert rel_l2(got[..., 0] + 1j * got[..., 1], np.fft.fft(zc, n=n, axis=-1)) < 1e-5
    # 65536 points (beyond the task's size-x range, through the C ABI), forward and cropped inverse
    big = 65536
    xr = np.random.default_rng(7).standard_normal((2, 40000)).astype(np.float32)
    spec = torch.empty((2, 2 * big), device=torch_dev)
    ctx.fft_nd(torch.from_numpy(xr).to(torch_dev), (40000, 2, 1), 1, (big, 2, 1), spec)
    back = torch.empty((2, 40000), device=torch_dev)
    ctx.ifft_nd(spec, (big, 2, 1), 1, back, crop_width=40000)
    torch.cuda.synchronize()
    sp = spec.cpu().numpy().reshape(2, big, 2)
    assert rel_l2(sp[..., 0] + 1j * sp[..., 1], np.fft.fft(xr.astype(np.float64), n=big, axis=-1)) < 1e-5
    assert rel_l2(back.cpu().numpy(), xr) < 1e-5
    sino = np.random.default_rng(6).standard_normal((6, w)).astype(np.float32)
    out = _run(ctx, [sino], _task("fft", dimensions=1, size_x=n), _task("filter"),
               _task("ifft", dimensions=1, crop_width=w))[0]
    coef = oracle.filter_coefficients(2 * n, "ramp-fromreal")
    ref = oracle.ramp_chain(sino, n, coef, w)
    assert rel_l2(out, ref) < 1e-5


@pytest.mark.parametrize("width,size_x", [(4500, 5000), (7000, 12345), (3, 20001)])
def test_long_rows_chirp(ctx, width, size_x):
    """Rows of other sizes above 4096 points: Bluestein over the four-step split (forward vs numpy in
    float64; round trip with a crop back to the input)."""
    rows = 3
    x = np.random.default_rng(size_x).standard_normal((rows, width)).astype(np.float32)
    spec = _run(ctx, [x], _task("fft", dimensions=1, size_x=size_x))[0]
    got = spec.reshape(rows, size_x, 2)
    ref = np.fft.fft(x.astype(np.float64), n=size_x, axis=-1)
    assert rel_l2(got[..., 0] + 1j * got[..., 1], ref) < 1e-5
    back = _run(ctx, [x], _task("fft", dimensions=1, size_x=size_x), _task("ifft", dimensions=1,
                                                                           crop_width=width))[0]
    assert rel_l2(back, x) < 1e-5


def test_reference_shell_test_177(ctx):
    """tests/test-177.sh: a 9x9 array of ones through fft dimensions=2 auto-zeropadding=True, then
    ifft dimensions=2 crop-width=9 crop-height=9; the maximum absolute error must be <= 1e-6."""
    x = np.ones((9, 9), np.float32)
    out = _run(ctx, [x], _task("fft", dimensions=2, auto_zeropadding=True),
               _task("ifft", dimensions=2, crop_width=9, crop_height=9))[0]
    assert out.shape == (9, 9)
    assert np.abs(out - 1.0).max() <= 1e-6
