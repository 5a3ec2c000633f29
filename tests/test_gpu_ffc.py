"""flat-field-correct and average on the GPU (src/kernels/ffc.cl,
src/ufo-flat-field-correct-task.c, src/ufo-average-task.c), and the
reference's file workflow (docs/examples.rst:20-48, 73-84) end to end:
[read projections, read darks ! average, read flat] ! flat-field-correct !
transpose-projections ! fft ! filter ! ifft ! backproject ! write."""
import numpy as np
import pytest

import oracle
import ufo_tasks
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _dev(x, torch_dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).to(torch_dev)


@pytest.mark.parametrize("absorb,fix,sino,scales", [(False, False, False, (1.0, 1.0)), (True, False, False, (1.0, 1.0)),
                                                    (True, True, False, (0.5, 1.5)), (False, True, True, (1.0, 2.0)),
                                                    (True, True, True, (1.0, 1.0))])
@pytest.mark.parametrize("per_item", [False, True])
def test_kernel_vs_oracle(ctx, torch_dev, absorb, fix, sino, scales, per_item):
    import torch
    rng = np.random.default_rng(7)
    n, h, w = 5, 12, 37
    data = rng.uniform(0.2, 1.0, (n, h, w)).astype(np.float32)
    dshape = (w,) if sino else (h, w)
    dark = rng.uniform(0.0, 0.1, ((n,) if per_item else ()) + dshape).astype(np.float32)
    flat = rng.uniform(1.0, 1.2, ((n,) if per_item else ()) + dshape).astype(np.float32)
    if fix:  # plant divisions by zero and logs of negatives
        d0 = dark.reshape(-1)
        data.reshape(n, -1)[:, 3] = d0[3 % d0.size]
        data[0, 0, 0] = -1.0
    out = torch.empty((n, h, w), device=torch_dev)
    nd = n if per_item else 1
    ctx.flat_field_correct(_dev(data, torch_dev), w, h, n, _dev(dark, torch_dev), nd, _dev(flat, torch_dev), nd, out,
                           sinogram_input=sino, absorption_correct=absorb, fix_nan_and_inf=fix, dark_scale=scales[0],
                           flat_scale=scales[1])
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    dk, fl = (dark[:, None] if sino else dark, flat[:, None] if sino else flat) if per_item else (dark, flat)
    ref = oracle.flat_field_correct_ref32(data, dk, fl, absorb, fix, sino and not per_item, *scales)
    fin = np.isfinite(ref)
    assert (np.isfinite(got) == fin).all()
    assert rel_l2(got[fin], ref[fin]) < 1e-6
    if fix:
        assert np.isfinite(got).all()


def test_average_bit_exact(ctx, torch_dev):
    import torch
    items = np.random.default_rng(3).standard_normal((9, 20, 30)).astype(np.float32) * 1e3
    out = torch.empty((2, 20, 30), device=torch_dev)
    ctx.average(_dev(items, torch_dev), 600, 9, out, copies=2)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    ref = oracle.average_ref32(items)
    np.testing.assert_array_equal(got[0], ref)
    np.testing.assert_array_equal(got[1], ref)


def test_task_three_inputs(ctx, torch_dev):
    """The task layer with three input buffers, and the size check of get_requisition."""
    import torch
    rng = np.random.default_rng(4)
    data, dark, flat = (rng.uniform(lo, hi, (8, 16)).astype(np.float32) for lo, hi in ((0.3, 1), (0, .1), (1, 1.1)))
    t = ufo_tasks.Task("flat-field-correct")
    t.props.absorption_correct = True
    assert t.L.ufomi_task_get_num_inputs(t.h) == 3
    t.setup(ctx)
    keep = [_dev(a, torch_dev) for a in (data, dark, flat)]
    bufs = []
    for x in keep:
        b = ufo_tasks.Buffer()
        b.data = x.data_ptr()
        b.req.n_dims = 2
        b.req.dims[0], b.req.dims[1] = 16, 8
        bufs.append(b)
    req = t.get_requisition(bufs)
    out = torch.empty((8, 16), device=torch_dev)
    o = ufo_tasks.Buffer()
    o.data = out.data_ptr()
    o.req = req
    assert t.process(bufs, o, req)
    torch.cuda.synchronize()
    ref = oracle.flat_field_correct_ref32(data, dark, flat, True)
    assert rel_l2(out.cpu().numpy(), ref) < 1e-6
    bufs[2].req.dims[0] = 15
    with pytest.raises(ufo_tasks.TaskError, match="same size"):
        t.get_requisition(bufs)


def test_file_workflow_end_to_end(ctx, torch_dev, tmp_path):
    """Projections, 3 darks and a flat as TIFF files -> corrected, transposed, reconstructed slices as
    TIFF files; equal to the oracle's correction + transposition fed to the in-memory FBP pipeline."""
    import torch
    from PIL import Image
    rng = np.random.default_rng(12)
    p, h, w, n_fft, axis = 48, 4, 64, 128, 31.5
    flat = rng.uniform(900, 1000, (h, w)).astype(np.float32)
    darks = rng.uniform(90, 110, (3, h, w)).astype(np.float32)
    att = rng.uniform(0.1, 0.9, (p, h, w)).astype(np.float32)
    projs = (darks.mean(0) + (flat - darks.mean(0)) * np.exp(-att)).astype(np.float32)
    ims = [Image.fromarray(x) for x in projs]
    ims[0].save(tmp_path / "proj.tif", save_all=True, append_images=ims[1:])
    for i, d in enumerate(darks):
        Image.fromarray(d).save(tmp_path / ("dark-%d.tif" % i))
    Image.fromarray(flat).save(tmp_path / "flat.tif")
    fbp = "fft dimensions=1 size-x=%d ! filter ! ifft dimensions=1 crop-width=%d ! backproject axis-pos=%g" % (
        n_fft, w, axis)
    desc = ("[read path=%s/proj.tif, read path=%s/dark-*.tif ! average, read path=%s/flat.tif] ! "
            "flat-field-correct absorption-correct=true ! transpose-projections number=%d ! %s ! "
            "write filename=%s/slice-%%02i.tif bytes-per-file=0 tiff-bigtiff=false"
            % (tmp_path, tmp_path, tmp_path, p, fbp, tmp_path))
    pl = ufo_tasks.Pipeline(desc)
    raw, n_out, shape = pl.run(ctx)
    assert raw is None and n_out == h and shape == (w, w)
    got = np.stack([np.asarray(Image.open(tmp_path / ("slice-%02d.tif" % i))) for i in range(h)])
    dark_avg = oracle.average_ref32(darks)
    corrected = oracle.flat_field_correct_ref32(projs, dark_avg, flat, absorption_correct=True)
    assert rel_l2(corrected, att) < 1e-4  # the correction recovers the attenuation
    sinos = oracle.transpose_projections(corrected, p)
    mem = ufo_tasks.Pipeline(fbp)
    raw2, n2, _ = mem.run(ctx, _dev(sinos, torch_dev))
    ref = raw2[:n2 * w * w * 4].view(torch.float32).reshape(n2, w, w).cpu().numpy()
    assert rel_l2(got, ref) < 1e-5
