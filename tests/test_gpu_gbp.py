"""GPU parity of general-backproject (parallel, cone, laminography and the
tilted / shifted variants), cone-beam weighting and the phantom generator.

Bar: float32 volume within 1e-5 relative L2 of the FP64 oracle and of the
float32 reference-semantics oracle; integer stores within one gray level.
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


TORCH_STORE = {"half": "float16", "float": "float32", "double": "float64", "uchar": "uint8",
               "ushort": "uint16", "uint": "uint32"}


def _gbp(ctx, torch_dev, projs, store_type="float", addressing_mode="clamp", overall_angle=2 * math.pi,
         region=(0, 0, 0), x_region=(0, 0, 0), y_region=(0, 0, 0), gray_map_min=0.0, gray_map_max=0.0,
         num_projections=None, compute_type="float", result_type="float", burst=0, **geometry):
    import torch
    p, h, w = projs.shape
    n = num_projections if num_projections is not None else p
    params = ufomi.gbp_params(n, overall_angle=overall_angle, region=region, x_region=x_region,
                              y_region=y_region, store_type=store_type, addressing_mode=addressing_mode,
                              gray_map_min=gray_map_min, gray_map_max=gray_map_max, compute_type=compute_type,
                              result_type=result_type, burst=burst)
    geom = ufomi.gbp_geometry(**geometry)
    shape = ufomi.gbp_volume_shape(params, w)
    vol = torch.zeros(shape, dtype=getattr(torch, TORCH_STORE[store_type]), device=torch_dev)
    ctx.gbp(_dev(projs, torch_dev), w, h, p, geom, params, vol)
    return vol.cpu().numpy()


def _check(ctx, torch_dev, projs, tol=1e-5, **kw):
    got = _gbp(ctx, torch_dev, projs, **kw)
    ref64 = oracle.gbp(projs, **kw)
    ref32 = oracle.gbp(projs, mode="ref32", **kw)
    assert got.shape == ref64.shape
    assert rel_l2(got, ref64) < tol
    assert rel_l2(got, ref32) < tol
    return got, ref64


def _projs(seed, p, h, w):
    return np.random.default_rng(seed).standard_normal((p, h, w)).astype(np.float32)


def _smooth_projs(seed, p, h, w, width=9):
    """Noise box-filtered along the detector rows. With voxels several pixels apart, white
    noise turns the float32 rounding of detector coordinates (~1e-5 px at u ~ 300) into
    errors of the same relative size in each float32 evaluation, so two float32 results
    differ by ~sqrt(2) times that; band-limited data keeps the comparison about geometry."""
    a = np.random.default_rng(seed).standard_normal((p, h, w + width - 1))
    k = np.ones(width) / width
    return np.apply_along_axis(lambda r: np.convolve(r, k, mode="valid"), 2, a).astype(np.float32)


def test_parallel_tomography(ctx, torch_dev):
    projs = _projs(1, 24, 8, 32)
    _check(ctx, torch_dev, projs, center_position_x=[16.0], center_position_z=[4.0], region=(-3.0, 3.0, 1.0))


def test_parallel_matches_backproject(ctx, torch_dev):
    # the untilted parallel general backprojection equals `backproject` (y flipped, half-pixel grid)
    import torch
    rng = np.random.default_rng(21)
    p, w, a = 40, 48, 23.5
    sino = rng.standard_normal((p, w)).astype(np.float32)
    vol = _gbp(ctx, torch_dev, sino[:, None, :], overall_angle=math.pi, x_region=(0.5 - a, 0.5 - a + w, 1.0),
               y_region=(a - 0.5, a - 0.5 - w, -1.0), center_position_x=[a], center_position_z=[0.5])
    out = torch.empty((1, w, w), device=torch_dev)
    ctx.bp_parallel(_dev(sino, torch_dev), w, p, out, axis_pos=a)
    assert rel_l2(vol[0], out.cpu().numpy()[0]) < 1e-5


@pytest.mark.parametrize("tilt", [0.0, math.radians(30), 0.9])
def test_laminography(ctx, torch_dev, tilt):
    projs = _projs(2, 30, 24, 24)
    _check(ctx, torch_dev, projs, axis_angle_x=[tilt], center_position_x=[12.0], center_position_z=[12.0],
           region=(-4.0, 4.0, 1.0))


def test_cone_beam(ctx, torch_dev):
    projs = _projs(3, 36, 20, 24)
    _check(ctx, torch_dev, projs, source_position_y=[-60.0], detector_position_y=[60.0],
           center_position_x=[12.0], center_position_z=[10.0], region=(-5.0, 5.0, 1.0))


@pytest.mark.parametrize("geo", [
    dict(axis_angle_x=[math.radians(30)]),                                   # laminography
    dict(),                                                                  # untilted parallel beam
    dict(source_position_y=[-200.0], detector_position_y=[100.0]),           # cone beam, magnification 1.5
])
def test_forty_slices(ctx, torch_dev, geo):
    # 40 slices: several voxel tiles along z, the last one partial (8 per tile for parallel
    # beam, 16 for cone beam), on the z-separable maps
    projs = _smooth_projs(7, 24, 40, 48)
    _check(ctx, torch_dev, projs, center_position_x=[24.0], center_position_z=[20.0], region=(-20.0, 20.0, 1.0),
           **geo)


@pytest.mark.parametrize("variant", [
    dict(detector_angle_x=[0.1], detector_angle_z=[0.05]),
    dict(detector_position_x=[1.5], detector_position_z=[-0.7]),
    dict(source_position_x=[2.0], source_position_z=[1.0]),
    dict(volume_angle_z=[0.3], volume_angle_x=[0.1]),
    dict(axis_angle_y=[0.05], axis_angle_x=[0.2]),
])
def test_cone_beam_variants(ctx, torch_dev, variant):
    projs = _projs(4, 20, 16, 20)
    _check(ctx, torch_dev, projs, source_position_y=[-50.0], detector_position_y=[40.0],
           center_position_x=[10.0], center_position_z=[8.0], region=(-3.0, 3.0, 1.0), **variant)


def test_parallel_tilted_detector(ctx, torch_dev):
    projs = _projs(5, 20, 16, 20)
    _check(ctx, torch_dev, projs, detector_angle_x=[0.1], detector_angle_y=[0.05], center_position_x=[10.0],
           center_position_z=[8.0], region=(-3.0, 3.0, 1.0))


def test_per_projection_geometry(ctx, torch_dev):
    p = 16
    projs = _projs(6, p, 16, 20)
    rng = np.random.default_rng(7)
    _check(ctx, torch_dev, projs, source_position_y=[-50.0], detector_position_y=[40.0],
           center_position_x=10.0 + rng.uniform(-1, 1, p), center_position_z=[8.0],
           axis_angle_z=np.sort(rng.uniform(0, 2 * math.pi, p)), region=(-3.0, 3.0, 1.0))


def test_regions_and_clamp_to_edge(ctx, torch_dev):
    projs = _projs(8, 12, 16, 16)
    _check(ctx, torch_dev, projs, addressing_mode="clamp_to_edge", x_region=(-10.0, 10.0, 0.5),
           y_region=(-6.0, 9.0, 1.5), region=(-2.0, 2.0, 0.5), center_position_x=[8.0], center_position_z=[8.0])


@pytest.mark.parametrize("cone", [False, True])
def test_wide_footprint_global_path(ctx, torch_dev, cone):
    """Voxels spaced 6 px apart: a 32-voxel tile spans ~190 detector columns, more than the
    LDS box holds, so every projection takes the global-memory path of the kernel."""
    projs = _smooth_projs(12, 10, 24, 300)
    kw = dict(x_region=(-96.0, 96.0, 6.0), y_region=(-96.0, 96.0, 6.0), region=(-4.0, 4.0, 2.0),
              center_position_x=[150.0], center_position_z=[12.0])
    if cone:
        kw.update(source_position_y=[-900.0], detector_position_y=[300.0])
    _check(ctx, torch_dev, projs, **kw)


def test_mixed_box_and_global_paths(ctx, torch_dev):
    """A volume whose outer tiles sit where the cone magnifies them beyond the LDS box."""
    projs = _smooth_projs(13, 16, 40, 160)
    _check(ctx, torch_dev, projs, x_region=(-64.0, 64.0, 1.0), y_region=(-64.0, 64.0, 1.0), region=(-4.0, 4.0, 1.0),
           center_position_x=[80.0], center_position_z=[20.0], source_position_y=[-70.0], detector_position_y=[90.0])


@pytest.mark.parametrize("store", ["half", "double", "uchar", "ushort", "uint"])
def test_store_types(ctx, torch_dev, store):
    projs = np.abs(_projs(9, 12, 12, 16))
    kw = dict(center_position_x=[8.0], center_position_z=[6.0], region=(-2.0, 2.0, 1.0))
    got = _gbp(ctx, torch_dev, projs, store_type=store, gray_map_min=0.0, gray_map_max=30.0, **kw)
    ref = oracle.gbp(projs, **kw)
    if store in ("half", "double"):
        assert rel_l2(got.astype(np.float64), ref) < (1e-3 if store == "half" else 1e-5)
    else:
        maxval = {"uchar": 255.0, "ushort": 65535.0, "uint": 4294967295.0}[store]
        exp = np.clip(maxval / 30.0 * ref, 0, maxval)
        assert np.abs(got.astype(np.float64) - exp).max() <= 1.0 + 1e-6 * maxval


def test_errors(ctx, torch_dev):
    projs = _projs(10, 4, 8, 8)
    with pytest.raises(ufomi.UfomiError) as e:
        _gbp(ctx, torch_dev, projs, num_projections=0)
    assert e.value.code == -10
    with pytest.raises(ufomi.UfomiError) as e:
        _gbp(ctx, torch_dev, projs, num_projections=6)
    assert e.value.code == -12
    with pytest.raises(ufomi.UfomiError) as e:
        _gbp(ctx, torch_dev, projs, store_type="uchar", gray_map_min=1.0, gray_map_max=1.0)
    assert e.value.code == -8
    with pytest.raises(ufomi.UfomiError) as e:
        _gbp(ctx, torch_dev, projs, burst=129)
    assert e.value.code == -1


def test_cone_weight(ctx, torch_dev):
    import torch
    p, h, w = 3, 20, 24
    x = _projs(11, p, h, w)
    out = torch.empty((p, h, w), device=torch_dev)
    cx, cz, sod = [12.0, 11.5, 12.5], 10.0, 500.0
    for tilt, odd in ((0.0, 700.0), (0.3, 500.0)):
        ctx.cone_weight(_dev(x, torch_dev), out, w, h, p, cx, cz, sod, odd, tilt)
        got = out.cpu().numpy()
        for i in range(p):
            ref = oracle.cone_weight_ref32(x[i], cx[i], cz, sod, odd, tilt)
            np.testing.assert_allclose(got[i], ref, rtol=2e-6, atol=1e-7)


def test_phantom_sinogram_vs_numpy(ctx, torch_dev):
    import torch
    W, P, axis = 96, 40, 47.5
    ell = ufomi.phantom_ellipses(1234, 10, W / 2)
    out = torch.empty((P, W), device=torch_dev)
    ctx.sinogram_ellipses(ell, W, P, axis, math.pi / P, out)
    ref = phantoms.sinogram(ell, W, P, axis, math.pi / P)
    np.testing.assert_allclose(out.cpu().numpy(), ref, rtol=1e-5, atol=1e-4)
    # ellipsoid plane sections: a slice through the centre of a single ellipsoid
    e3 = np.array([[3.0, -2.0, 1.0, 10.0, 6.0, 4.0, 0.3, 0.7]])
    sec = ufomi.slice_ellipsoids(e3, 1.0)
    np.testing.assert_allclose(sec, [[3.0, -2.0, 10.0, 6.0, 0.3, 0.7]])
    sec2 = ufomi.slice_ellipsoids(e3, 3.0)  # (z - z0) / c = 0.5 -> scale sqrt(0.75)
    np.testing.assert_allclose(sec2[0, 2:4], np.array([10.0, 6.0]) * math.sqrt(0.75))


def test_projector_sphere_closed_form(ctx, torch_dev):
    """Chord lengths through a centred sphere: 2 rho sqrt(R^2 - r^2) for parallel rays;
    the central cone ray passes through the centre (2 rho R)."""
    import torch
    W = H = 33
    R, rho = 10.0, 0.7
    sphere = np.array([[0.0, 0.0, 0.0, R, R, R, 0.0, rho]])
    params = ufomi.gbp_params(3)
    g = ufomi.gbp_geometry(center_position_x=[16.5], center_position_z=[16.5])
    rows, par = ufomi.gbp_projection_rows(g, params)
    assert par
    out = torch.empty((3, H, W), device=torch_dev)
    ctx.project_ellipsoids(sphere, W, H, rows, out)
    u = np.arange(W) + 0.5 - 16.5
    r2 = u[None, :] ** 2 + u[:, None] ** 2
    ref = 2 * rho * np.sqrt(np.maximum(R * R - r2, 0))
    for i in range(3):
        np.testing.assert_allclose(out[i].cpu().numpy(), ref, rtol=1e-5, atol=1e-4)
    gc = ufomi.gbp_geometry(center_position_x=[16.5], center_position_z=[16.5], source_position_y=[-60.0],
                            detector_position_y=[60.0])
    rows_c, par_c = ufomi.gbp_projection_rows(gc, params)
    assert not par_c
    ctx.project_ellipsoids(sphere, W, H, rows_c, out)
    assert abs(out[0, 16, 16].item() - 2 * rho * R) < 1e-4


def test_projection_rows_match_kernel_geometry(ctx, torch_dev):
    """Backprojecting a one-hot projection lights the voxels whose (u, v) the rows put on that texel."""
    rows, _ = ufomi.gbp_projection_rows(ufomi.gbp_geometry(source_position_y=[-40.0], detector_position_y=[30.0],
                                                           center_position_x=[8.0], center_position_z=[4.0]),
                                        ufomi.gbp_params(5))
    ref = oracle.gbp(np.zeros((5, 8, 16), np.float32) + 1.0, source_position_y=[-40.0], detector_position_y=[30.0],
                     center_position_x=[8.0], center_position_z=[4.0], region=(-1.0, 1.0, 1.0))
    x, y, z = -8 + 3, -8 + 5, 0.0
    X = np.array([x, y, z, 1.0])
    for p in range(5):
        U, V, D, C = rows[p] @ X
        assert 0 < D and 0 < C
    assert np.isfinite(ref).all()


def test_noise_is_seeded(ctx, torch_dev):
    import torch
    a = torch.ones(100000, device=torch_dev)
    b = torch.ones(100000, device=torch_dev)
    m = ctx.add_noise(a, 0.01, 99)
    ctx.add_noise(b, 0.01, 99)
    assert m == 1.0
    assert torch.equal(a, b)
    d = (a - 1).cpu().numpy()
    assert abs(d.std() - 0.01) < 5e-4 and abs(d.mean()) < 2e-4


SWEEP_BASE = dict(source_position_y=[-50.0], detector_position_y=[40.0], center_position_x=[10.0],
                  center_position_z=[8.0])


@pytest.mark.parametrize("parameter,region,extra", [
    ("center-position-x", (8.5, 11.5, 0.5), {}),
    ("axis-angle-x", (0.0, 0.6, 0.2), {}),
    ("detector-angle-z", (-0.1, 0.1, 0.05), {}),
    ("source-position-y", (-70.0, -40.0, 10.0), {}),
    ("center-position-z", (6.0, 10.0, 1.0), {"z": 1.5}),
    ("volume-angle-z", (0.0, 1.0, 0.25), {}),
])
def test_parameter_sweep(ctx, torch_dev, parameter, region, extra):
    """Slice k of a sweep equals the plain reconstruction at z with the parameter set to
    region[0] + k region[2] for every projection (make_parameter_assignment)."""
    import torch
    projs = _smooth_projs(21, 24, 16, 20)
    p, h, w = projs.shape
    z = extra.get("z", 0.0)
    params = ufomi.gbp_params(p, region=region)
    params.z = z
    geom = ufomi.gbp_geometry(**SWEEP_BASE)
    shape = ufomi.gbp_volume_shape(params, w)
    vol = torch.zeros(shape, device=torch_dev)
    ctx.gbp_sweep(_dev(projs, torch_dev), w, h, p, geom, params, ufomi.GBP_PARAMETERS.index(parameter), vol)
    got = vol.cpu().numpy()
    key = parameter.replace("-", "_")
    n = shape[0]
    assert n == int(math.ceil((region[1] - region[0]) / region[2]))
    for k in range(n):
        geo = dict(SWEEP_BASE)
        geo[key] = [region[0] + k * region[2]]
        ref = oracle.gbp(projs, region=(z, z + 1.0, 1.0), **geo)
        assert rel_l2(got[k], ref[0]) < 1e-5, (k, rel_l2(got[k], ref[0]))


def test_sweep_task(ctx, torch_dev):
    import ufo_tasks
    projs = _smooth_projs(22, 16, 12, 16)
    t = ufo_tasks.Task("general-backproject")
    for k, v in dict(num_projections=16, parameter="center-position-x", region=[7.0, 9.0, 0.5], z=0.5,
                     source_position_y=[-40.0], detector_position_y=[30.0], center_position_z=[6.0]).items():
        setattr(t.props, k, v)
    src, sink = ufo_tasks.MemoryIn(projs), ufo_tasks.MemoryOut()
    g = ufo_tasks.TaskGraph()
    g.connect_nodes(src, t)
    g.connect_nodes(t, sink)
    ufo_tasks.Scheduler(ctx).run(g)
    got = sink.array
    assert got.shape == (4, 16, 16)
    for k in range(4):
        ref = oracle.gbp(projs, region=(0.5, 1.5, 1.0), source_position_y=[-40.0], detector_position_y=[30.0],
                         center_position_x=[7.0 + 0.5 * k], center_position_z=[6.0])
        assert rel_l2(got[k], ref[0]) < 1e-5

def _run_task(ctx, projs, **props):
    import ufo_tasks
    t = ufo_tasks.Task("general-backproject")
    for k, v in props.items():
        setattr(t.props, k, v)
    src, sink = ufo_tasks.MemoryIn(projs), ufo_tasks.MemoryOut()
    g = ufo_tasks.TaskGraph()
    g.connect_nodes(src, t)
    g.connect_nodes(t, sink)
    ufo_tasks.Scheduler(ctx).run(g)
    return sink.array


def test_axis_angle_z_parameter(ctx, torch_dev):
    # The reference's generated kernel never reads the swept axis-angle-z: without per-projection
    # geometry it does not build; with it every slice is the reconstruction at z.
    import ufo_tasks
    projs = _smooth_projs(23, 16, 12, 16)
    base = dict(num_projections=16, parameter="axis-angle-z", region=[0.0, 1.5, 0.5], z=0.5,
                center_position_z=[6.0])
    with pytest.raises(ufo_tasks.TaskError) as e:
        _run_task(ctx, projs, center_position_x=[8.0], **base)
    assert e.value.status == -4 and "Error creating backprojection kernels" in str(e.value)
    cx = list(8.0 + 0.1 * np.arange(16))
    got = _run_task(ctx, projs, center_position_x=cx, **base)
    assert got.shape == (3, 16, 16)
    ref = oracle.gbp(projs, region=(0.5, 1.5, 1.0), center_position_x=cx, center_position_z=[6.0])
    for k in range(3):
        assert rel_l2(got[k], ref[0]) < 1e-5


def test_double_compute_and_result_types(ctx, torch_dev):
    """compute-type=double and result-type=double through the task: FP64 maps, FP64 bilinear
    weights and FP64 running sums meet the FP64 oracle to FP64 rounding with a double volume."""
    projs = _smooth_projs(5, 18, 16, 24)
    geo = dict(region=[-2.0, 2.0, 1.0], axis_angle_x=[0.4], center_position_x=[12.0], center_position_z=[8.0])
    ref = oracle.gbp(projs, region=(-2.0, 2.0, 1.0), axis_angle_x=[0.4], center_position_x=[12.0],
                     center_position_z=[8.0])
    got = _run_task(ctx, projs, num_projections=18, compute_type="double", result_type="double", **geo)
    assert rel_l2(got, ref) < 1e-6  # float volume (the task layer hands out float32 items)
    got64 = _gbp(ctx, torch_dev, projs, compute_type="double", result_type="double", store_type="double",
                 region=(-2.0, 2.0, 1.0), axis_angle_x=[0.4], center_position_x=[12.0], center_position_z=[8.0])
    assert got64.dtype == np.float64
    assert rel_l2(got64, ref) < 1e-11
    got32 = _gbp(ctx, torch_dev, projs, store_type="double", region=(-2.0, 2.0, 1.0), axis_angle_x=[0.4],
                 center_position_x=[12.0], center_position_z=[8.0])
    assert 1e-10 < rel_l2(got32, ref) < 1e-5  # the float kernel stays at float rounding


@pytest.mark.parametrize("k", [0, 3, 5, 7])
def test_double_compute_on_variants(ctx, torch_dev, k):
    # FP64 compute on the tilted / shifted / rotated / cone variants against the FP64 oracle
    import test_oracle_gbp_geometry as G
    rng = np.random.default_rng(400 + k)
    p = 6
    projs = _smooth_projs(400 + k, p, 64, 64) + np.float32(1.0)
    geo = G._random_geometry(rng, p, *G.VARIANTS[k])
    got = _gbp(ctx, torch_dev, projs, compute_type="double", result_type="double", store_type="double",
               **G.MODEL_GRID, **geo)
    ref = oracle.gbp(projs, **G.MODEL_GRID, **geo)
    assert rel_l2(got, ref) < 1e-11


@pytest.mark.parametrize("burst", [0, 3, 16])
@pytest.mark.parametrize("cone", [False, True])
def test_result_type_half(ctx, torch_dev, burst, cone):
    """result-type half: the running sum of every burst rounded to binary16 after each projection.
    Against the oracle's float32 semantics with the same rounding (a rounding flip from a float32
    difference moves one voxel by a half ulp), and visibly away from the FP64 result."""
    projs = _smooth_projs(31, 24, 20, 24) + np.float32(1.0)
    geo = dict(center_position_x=[12.0], center_position_z=[10.0], region=(-3.0, 3.0, 1.0))
    if cone:
        geo.update(source_position_y=[-60.0], detector_position_y=[60.0])
    got = _gbp(ctx, torch_dev, projs, result_type="half", burst=burst, **geo)
    emu = oracle.gbp(projs, mode="ref32", result_type="half", burst=burst or 8, **geo)
    f64 = oracle.gbp(projs, **geo)
    err_emu, err64 = rel_l2(got, emu), rel_l2(got, f64)
    print("result-type half burst %d cone %s: vs emulation %.3g, vs FP64 %.3g" % (burst, cone, err_emu, err64))
    assert err_emu < 2e-4
    assert err64 > 3 * err_emu


@pytest.mark.parametrize("addressing_mode", ["repeat", "mirrored_repeat"])
@pytest.mark.parametrize("cone", [False, True])
def test_repeat_addressing(ctx, torch_dev, addressing_mode, cone):
    # a volume wider than the detector, so samples wrap (repeat) or mirror
    projs = _smooth_projs(41, 12, 16, 20) + np.float32(1.0)
    geo = dict(center_position_x=[10.0], center_position_z=[8.0], region=(-12.0, 12.0, 4.0),
               x_region=(-24.0, 24.0, 1.5), y_region=(-24.0, 24.0, 1.5), addressing_mode=addressing_mode)
    if cone:
        geo.update(source_position_y=[-80.0], detector_position_y=[40.0])
    _check(ctx, torch_dev, projs, **geo)


@pytest.mark.parametrize("k", range(9))
def test_variants_vs_ray_model(ctx, torch_dev, k):
    # The product against the independent ray-plane model (tests/test_oracle_gbp_geometry.py) on
    # random per-projection geometries: tilted and shifted detectors, shifted sources, volume
    # rotations and axis tilts, parallel and cone beam.
    import test_oracle_gbp_geometry as G
    rng = np.random.default_rng(300 + k)
    p = 6
    projs = _smooth_projs(300 + k, p, 64, 64) + np.float32(1.0)
    geo = G._random_geometry(rng, p, *G.VARIANTS[k])
    got = _gbp(ctx, torch_dev, projs, overall_angle=2 * math.pi, **G.MODEL_GRID, **geo)
    ref = G.ray_model_volume(projs, geo, 2 * math.pi, **G.MODEL_GRID)
    assert got.shape == ref.shape
    assert rel_l2(got, ref) < 1e-5
