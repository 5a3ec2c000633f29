"""lamino-backproject on the GPU (src/ufo-lamino-backproject-task.c, the kernel
templates src/kernels/templates/*.in) against the FP64 C oracle: every swept
parameter, rolls, offsets, integer region steps and both clamp modes within
1e-5 relative L2, plus the task layer."""
import math

import numpy as np
import pytest

import oracle
import ufo_tasks
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _run(ctx, torch_dev, projs, **kw):
    import torch
    import ufomi
    p = projs.shape[0]
    prm = ufomi.lamino_params(kw.pop("num_projections", p), **kw)
    shape = ufomi.lamino_shape(prm)
    out = torch.full(shape, float("nan"), device=torch_dev)
    d = torch.from_numpy(np.ascontiguousarray(projs, dtype=np.float32)).to(torch_dev)
    ctx.lamino_bp(d, projs.shape[2], projs.shape[1], p, prm, out)
    torch.cuda.synchronize()
    return out.cpu().numpy()


CASES = [
    dict(lamino_angle=0.5, center=(31.5, 23.0), x_region=(-24, 24, 1), y_region=(-20, 20, 1), region=(-8.0, 8.0, 1.0)),
    dict(lamino_angle=1.2, roll_angle=0.07, center=(30.25, 25.5), x_region=(-30, 31, 2), y_region=(-12, 13, 3),
         region=(-4.0, 5.0, 0.75), projection_offset=(2, -3)),
    dict(lamino_angle=math.pi / 2, center=(32.0, 0.5), x_region=(-32, 32, 1), y_region=(-32, 32, 1),
         region=(0.0, 4.0, 1.0), addressing_mode="clamp_to_edge"),
    dict(lamino_angle=0.3, overall_angle=-2 * math.pi, center=(31.0, 24.0), x_region=(-16, 16, 1),
         y_region=(-16, 16, 1), region=(-2.0, 2.0, 1.0), num_projections=80),
    dict(lamino_angle=0.8, tomo_angle=0.4, center=(31.0, 24.0), x_region=(-8, 8, 1), y_region=(-8, 8, 1),
         region=(0.0, 2.0, 1.0)),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_z_vs_oracle(ctx, torch_dev, case):
    kw = CASES[case]
    projs = np.random.default_rng(case).standard_normal((60, 48, 64)).astype(np.float32)
    got = _run(ctx, torch_dev, projs, **dict(kw))
    ref = oracle.lamino(projs, **kw)
    assert got.shape == ref.shape
    assert rel_l2(got, ref) < 1e-5, rel_l2(got, ref)


@pytest.mark.parametrize("parameter,region", [("x-center", (28.0, 32.0, 1.5)), ("lamino-angle", (0.2, 1.0, 0.25)),
                                              ("roll-angle", (-0.1, 0.1, 0.05)), ("z", (-3.0, 3.0, 2.0))])
def test_sweeps_vs_oracle(ctx, torch_dev, parameter, region):
    projs = np.random.default_rng(7).standard_normal((45, 40, 56)).astype(np.float32)
    kw = dict(lamino_angle=0.6, roll_angle=0.04, center=(28.5, 20.0), x_region=(-20, 20, 1), y_region=(-18, 18, 2),
              z=1.5, parameter=parameter, region=region)
    got = _run(ctx, torch_dev, projs, **dict(kw))
    ref = oracle.lamino(projs, **kw)
    assert got.shape == ref.shape
    assert rel_l2(got, ref) < 1e-5, rel_l2(got, ref)


def test_float32_semantics_oracle(ctx, torch_dev):
    """The reference's own float32 arithmetic (bursts of 16, mad chains) is within 1e-5 as well."""
    projs = np.random.default_rng(11).standard_normal((48, 32, 40)).astype(np.float32)
    kw = dict(lamino_angle=0.9, roll_angle=-0.05, center=(20.0, 16.0), x_region=(-16, 16, 1), y_region=(-16, 16, 1),
              region=(-2.0, 2.0, 1.0))
    got = _run(ctx, torch_dev, projs, **dict(kw))
    assert rel_l2(got, oracle.lamino(projs, mode="ref32", **kw)) < 1e-5


def _tuples(kw):
    return {k: tuple(v) if isinstance(v, list) else v for k, v in kw.items()}


def _task(**props):
    t = ufo_tasks.Task("lamino-backproject")
    for k, v in props.items():
        setattr(t.props, k, v)
    return t


def test_task_one_volume(ctx, torch_dev):
    """The task takes projections one by one and emits one 3-D volume [nz][ny][nx]."""
    projs = np.random.default_rng(3).standard_normal((32, 24, 36)).astype(np.float32)
    kw = dict(lamino_angle=0.7, center=[18.0, 12.0], x_region=[-10, 10, 1], y_region=[-8, 8, 1],
              region=[-1.0, 2.0, 1.0])
    t = _task(num_projections=32, **kw)
    src, sink = ufo_tasks.MemoryIn(projs), ufo_tasks.MemoryOut()
    g = ufo_tasks.TaskGraph()
    g.connect_nodes(src, t)
    g.connect_nodes(t, sink)
    ufo_tasks.Scheduler(ctx).run(g)
    assert len(sink.array) == 1 and sink.array[0].shape == (3, 16, 20)
    ref = oracle.lamino(projs, **_tuples(kw))
    assert rel_l2(sink.array[0], ref) < 1e-5
    assert tuple(t.props.x_region) == (-10, 10, 1)


def test_task_partial_burst_left_out(ctx, torch_dev):
    """20 of 40 projections: the reference's burst kernel runs on the first 16 only."""
    projs = np.random.default_rng(4).standard_normal((20, 16, 24)).astype(np.float32)
    kw = dict(lamino_angle=0.5, center=[12.0, 8.0], x_region=[-6, 6, 1], y_region=[-6, 6, 1], region=[0.0, 1.0, 1.0])
    t = _task(num_projections=40, **kw)
    src, sink = ufo_tasks.MemoryIn(projs), ufo_tasks.MemoryOut()
    g = ufo_tasks.TaskGraph()
    g.connect_nodes(src, t)
    g.connect_nodes(t, sink)
    ufo_tasks.Scheduler(ctx).run(g)
    ref = oracle.lamino(projs[:16], num_projections=40, **_tuples(kw))
    assert rel_l2(sink.array[0], ref) < 1e-5


def test_task_errors(ctx, torch_dev):
    with pytest.raises(ufo_tasks.TaskError, match="Number of projections has not been set"):
        _task(region=[0.0, 1.0, 1.0]).setup(ctx)
    with pytest.raises(ufo_tasks.TaskError, match="Step in region is 0"):
        _task(num_projections=4).setup(ctx)
    with pytest.raises(ufo_tasks.TaskError, match="normalized"):
        _task(num_projections=4, region=[0.0, 1.0, 1.0], addressing_mode="repeat").setup(ctx)


def test_pipeline_files(ctx, torch_dev, tmp_path):
    """read (multi-page TIFF) ! lamino-backproject ! write: one volume item, one page per slice."""
    from PIL import Image
    projs = np.random.default_rng(5).standard_normal((32, 20, 28)).astype(np.float32)
    ims = [Image.fromarray(x) for x in projs]
    ims[0].save(tmp_path / "proj.tif", save_all=True, append_images=ims[1:])
    desc = ("read path=%s/proj.tif ! lamino-backproject num-projections=32 lamino-angle=0.6 center=14,10 "
            "x-region=-8,8,1 y-region=-6,6,1 region=-1,2,1 ! write filename=%s/vol.tif tiff-bigtiff=false"
            % (tmp_path, tmp_path))
    raw, n_out, shape = ufo_tasks.Pipeline(desc).run(ctx)
    assert raw is None and n_out == 1
    im = Image.open(tmp_path / "vol.tif")
    pages = []
    for k in range(im.n_frames):
        im.seek(k)
        pages.append(np.asarray(im, dtype=np.float32))
    got = np.stack(pages)
    ref = oracle.lamino(projs, lamino_angle=0.6, center=(14.0, 10.0), x_region=(-8, 8, 1), y_region=(-6, 6, 1),
                        region=(-1.0, 2.0, 1.0))
    assert got.shape == ref.shape == (3, 12, 16)
    assert rel_l2(got, ref) < 1e-5
