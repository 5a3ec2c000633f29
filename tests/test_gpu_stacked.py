"""stacked-backproject on the GPU (src/ufo-stacked-backproject-task.c,
src/kernels/stacked-backproject.cl) against the float32-semantics oracle:
single and half within 1e-5 relative L2, int8 bit for bit (the oracle keeps
the kernel's float operation order; sums are integers)."""
import numpy as np
import pytest

import oracle
import ufo_tasks
from conftest import rel_l2

pytestmark = pytest.mark.gpu

PREC = {"int8": 0, "half": 1, "single": 2}


def _run(ctx, torch_dev, sinos, precision, **kw):
    import torch
    import ufomi
    n, p, w = sinos.shape
    prm = ufomi.bp_params(**kw)
    rw, rh = kw.get("roi_width") or w, kw.get("roi_height") or w
    out = torch.full((n, rh, rw), float("nan"), device=torch_dev)
    ctx.bp_stacked(torch.from_numpy(sinos).to(torch_dev), w, p, n, prm, PREC[precision], out)
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("precision", ["single", "half", "int8"])
@pytest.mark.parametrize("n,p,w,kw", [
    (4, 90, 64, {}),
    (5, 121, 48, dict(axis_pos=23.75, roi_x=4, roi_y=3, roi_width=40, roi_height=37)),
    (7, 60, 33, dict(angle_offset=0.3, angle_step=0.05, num_projections=80, offset=11)),
    (1, 45, 20, dict(roi_width=52, roi_height=17)),
])
def test_vs_oracle(ctx, torch_dev, precision, n, p, w, kw):
    sinos = np.random.default_rng(n * 31 + p).standard_normal((n, p, w)).astype(np.float32)
    got = _run(ctx, torch_dev, sinos, precision, **kw)
    ref = oracle.stacked_bp_ref32(sinos, precision, **kw)
    if precision == "int8":
        np.testing.assert_array_equal(got, ref)
    else:
        assert rel_l2(got, ref) < 1e-5, rel_l2(got, ref)


def test_int8_constant_and_two_level_stacks(ctx, torch_dev):
    """Closed forms: a constant stack quantizes to 0 everywhere (max == min) and gives min pi / P;
    a stack of only min and max values holds exact 0 / 255 texels."""
    c = np.full((4, 30, 16), 2.5, np.float32)
    got = _run(ctx, torch_dev, c, "int8")
    np.testing.assert_allclose(got, np.float32(2.5) * np.float32(np.pi) / 30, rtol=1e-6)
    two = np.where(np.random.default_rng(3).random((4, 30, 16)) > 0.5, 1.0, -1.0).astype(np.float32)
    np.testing.assert_array_equal(_run(ctx, torch_dev, two, "int8"), oracle.stacked_bp_ref32(two, "int8"))


def test_task_stack_item(ctx, torch_dev):
    """The task on one 3-D item [depth][P][W]: output [depth][roi_h][roi_w], num-projections sticky."""
    sinos = np.random.default_rng(9).standard_normal((6, 64, 40)).astype(np.float32)
    t = ufo_tasks.Task("stacked-backproject")
    t.props.precision_mode = "half"
    t.props.roi_width = 32
    src, sink = ufo_tasks.MemoryIn(sinos[None]), ufo_tasks.MemoryOut()
    g = ufo_tasks.TaskGraph()
    g.connect_nodes(src, t)
    g.connect_nodes(t, sink)
    ufo_tasks.Scheduler(ctx).run(g)
    got = sink.array[0]
    assert got.shape == (6, 40, 32)
    ref = oracle.stacked_bp_ref32(sinos, "half", roi_width=32)
    assert rel_l2(got, ref) < 1e-5
    assert t.props.precision_mode == "half"
    bad = ufo_tasks.Task("stacked-backproject")
    bad.props.num_projections = 10
    with pytest.raises(ufo_tasks.TaskError) as e:
        g2 = ufo_tasks.TaskGraph()
        s2, k2 = ufo_tasks.MemoryIn(sinos[None]), ufo_tasks.MemoryOut()
        g2.connect_nodes(s2, bad)
        g2.connect_nodes(bad, k2)
        ufo_tasks.Scheduler(ctx).run(g2)
    assert e.value.status == -5
