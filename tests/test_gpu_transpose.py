"""transpose-projections on the GPU (src/ufo-transpose-projections-task.c):
the kernel, the reductor task and the batched pipeline stage, bit-exact
against the oracle (a permutation moves values unchanged)."""
import numpy as np
import pytest

import oracle
import ufo_tasks
from conftest import rel_l2

pytestmark = pytest.mark.gpu


# widths that take the 16-byte, 8-byte and 4-byte paths; fewer and more projections than `number`
@pytest.mark.parametrize("n,h,w,number", [(16, 12, 64, 16), (7, 5, 30, 10), (9, 4, 33, 9), (12, 3, 8, 5),
                                          (1, 1, 1, 1), (40, 17, 517, 40)])
def test_kernel_vs_oracle(ctx, torch_dev, n, h, w, number):
    import torch
    projs = np.random.default_rng(n * 7 + w).standard_normal((n, h, w)).astype(np.float32)
    d_in = torch.from_numpy(projs).to(torch_dev)
    out = torch.full((h, number, w), float("nan"), device=torch_dev)
    used = min(n, number)
    ctx.transpose_projections(d_in, w, h, used, out, first=0, number=number)
    if used < number:
        ctx.transpose_projections(None, w, h, number - used, out, first=used, number=number)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy(), oracle.transpose_projections(projs, number))


def test_kernel_unaligned_views(ctx, torch_dev):
    """Pointers off 16-byte alignment fall back to narrower vectors."""
    import torch
    n, h, w = 6, 5, 24
    projs = np.random.default_rng(1).standard_normal((n, h, w)).astype(np.float32)
    base_in = torch.zeros(n * h * w + 1, device=torch_dev)
    base_in[1:] = torch.from_numpy(projs.ravel()).to(torch_dev)
    base_out = torch.zeros(h * n * w + 3, device=torch_dev)
    ctx.transpose_projections(base_in[1:], w, h, n, base_out[3:], number=n)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(base_out[3:].cpu().numpy().reshape(h, n, w), oracle.transpose_projections(projs, n))
    assert base_out[:3].abs().sum().item() == 0


def test_kernel_rejects_bad_ranges(ctx, torch_dev):
    import torch
    import ufomi
    x = torch.zeros(64, device=torch_dev)
    with pytest.raises(ufomi.UfomiError):
        ctx.transpose_projections(x, 4, 2, 3, x, first=2, number=4)


def test_large_property(ctx, torch_dev):
    """A 2048-wide stack: equals torch's permutation exactly."""
    import torch
    n, h, w = 180, 96, 2048
    g = torch.Generator(device=torch_dev).manual_seed(5)
    d_in = torch.randn((n, h, w), device=torch_dev, generator=g)
    out = torch.empty((h, n, w), device=torch_dev)
    ctx.transpose_projections(d_in, w, h, n, out)
    torch.cuda.synchronize()
    assert torch.equal(out, d_in.permute(1, 0, 2).contiguous())


def _run_chain(ctx, array, *tasks):
    src, sink = ufo_tasks.MemoryIn(array), ufo_tasks.MemoryOut()
    g = ufo_tasks.TaskGraph()
    nodes = [src, *tasks, sink]
    for a, b in zip(nodes, nodes[1:]):
        g.connect_nodes(a, b)
    ufo_tasks.Scheduler(ctx).run(g)
    return sink.array


@pytest.mark.parametrize("n,number", [(10, 10), (6, 9), (12, 8)])
def test_task_reductor(ctx, torch_dev, n, number):
    h, w = 7, 21
    projs = np.random.default_rng(n + number).standard_normal((n, h, w)).astype(np.float32)
    t = ufo_tasks.Task("transpose-projections")
    t.props.number = number
    assert t.props.number == number and t.mode & ufo_tasks.MODE_REDUCTOR
    got = _run_chain(ctx, projs, t)
    np.testing.assert_array_equal(got, oracle.transpose_projections(projs, number))


def test_projections_to_fbp_pipeline(ctx, torch_dev):
    """transpose-projections ! fft ! filter ! ifft ! backproject equals the FBP of the oracle's sinograms."""
    import torch
    p, h, w, n_fft, axis = 60, 6, 64, 128, 31.5
    projs = np.random.default_rng(3).standard_normal((p, h, w)).astype(np.float32)
    chain = "fft dimensions=1 size-x=%d ! filter ! ifft dimensions=1 crop-width=%d ! backproject axis-pos=%g" % (
        n_fft, w, axis)
    pl = ufo_tasks.Pipeline("transpose-projections number=%d ! %s" % (p, chain))
    raw, n_out, shape = pl.run(ctx, torch.from_numpy(projs).to(torch_dev))
    assert n_out == h and shape == (w, w)
    got = raw[:n_out * w * w * 4].view(torch.float32).reshape(n_out, w, w).cpu().numpy()
    sinos = oracle.transpose_projections(projs, p)
    ref_pl = ufo_tasks.Pipeline(chain)
    raw2, n2, _ = ref_pl.run(ctx, torch.from_numpy(sinos).to(torch_dev))
    np.testing.assert_array_equal(got, raw2[:n2 * w * w * 4].view(torch.float32).reshape(n2, w, w).cpu().numpy())
    coef = oracle.filter_coefficients(2 * n_fft)
    for i in range(h):
        ref = oracle.bp_texture_ref32(oracle.ramp_chain(sinos[i], n_fft, coef, w).astype(np.float32), axis_pos=axis)
        assert rel_l2(got[i], ref) < 1e-5
