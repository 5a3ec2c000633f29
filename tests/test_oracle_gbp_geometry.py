"""Pins of the general-backprojection oracle's geometry branches.

The oracle (orc_gbp in oracle/ufo_oracle.c) transcribes the statements that
ufo-general-backproject-task.c generates (make_static_transformations :807-849,
make_projection_computation :869-887, make_transformations :897-1048, rotations
of general_bp_definitions.in:1-3).  Two checks here do not share that
transcription:

1. A ray-plane model of the same geometry, written from the physical setup:
   - rotations are Rodrigues matrices about the unit axes;
   - the detector is a plane through the detector position, with the frame
     B = Rx(a_x) Ry(a_y) Rz(a_z) whose columns are the detector's u axis, its
     inward normal (negated) and its v axis;
   - a voxel is projected by intersecting its ray (from the source in cone
     beam, along +y in parallel beam) with that plane and taking coordinates in
     the detector frame.
   The order of the rotations is the reference's statement order
   (rotate_z first, then rotate_y, then rotate_x, :786-793; tomographic
   rotation about z, then axis_y, then axis_x, :960-972).
2. Closed forms of the projected coordinate for one parameter at a time, derived
   by hand (each test says how).  They are read off the oracle with "coordinate
   images": a projection whose texel (i, j) holds i + 0.5 (or j + 0.5) makes the
   bilinear sampler return the sampled coordinate itself, and a projection of
   ones returns the cone weight coeff^2.

orc_gbp_plant switches on mistakes inside the oracle; each one must fail here.
"""
import math

import numpy as np
import pytest

import oracle

W = H = 256
C0 = 128.0
GRID = dict(x_region=(-20.0, 20.0, 10.0), y_region=(-15.0, 15.0, 10.0), region=(-12.0, 12.0, 8.0))


# ------------------------------------------------------------------ ray-plane model

def _rot(axis, t):
    k = np.zeros(3)
    k["xyz".index(axis)] = 1.0
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + math.sin(t) * K + (1 - math.cos(t)) * (K @ K)


def _grid(n_start_stop_step):
    start, stop, step = n_start_stop_step
    n = int(math.ceil((stop - start) / step))
    return start + step * np.arange(n)


def _bilinear_zero_border(img, x, y):
    """OpenCL linear sampler, unnormalised coordinates, CLK_ADDRESS_CLAMP (border 0)."""
    h, w = img.shape
    u, v = x - 0.5, y - 0.5
    i0, j0 = np.floor(u).astype(np.int64), np.floor(v).astype(np.int64)
    a, b = u - i0, v - j0

    def tex(i, j):
        ok = (i >= 0) & (i < w) & (j >= 0) & (j < h)
        return np.where(ok, img[np.clip(j, 0, h - 1), np.clip(i, 0, w - 1)].astype(np.float64), 0.0)

    return ((1 - a) * (1 - b) * tex(i0, j0) + a * (1 - b) * tex(i0 + 1, j0)
            + (1 - a) * b * tex(i0, j0 + 1) + a * b * tex(i0 + 1, j0 + 1))


def _geom_at(geometry, name, i, default):
    vals = geometry.get(name)
    if vals is None:
        return default
    vals = np.atleast_1d(vals)
    return float(vals[0] if len(vals) == 1 else vals[i])


def ray_model_projection(geometry, i, X0):
    """Detector coordinates (u + cx, v + cz) and cone weight of points X0 [3][n] for projection i."""
    g = lambda name, d=0.0: _geom_at(geometry, name, i, d)
    sy, dy = g("source_position_y", -math.inf), g("detector_position_y")
    cone = math.isfinite(sy)
    X = X0.astype(np.float64).copy()
    if cone:
        X *= -sy / (dy - sy)  # voxel grid is given at the detector's scale
    X = _rot("x", g("volume_angle_x")) @ _rot("y", g("volume_angle_y")) @ _rot("z", g("volume_angle_z")) @ X
    X = _rot("x", g("axis_angle_x")) @ _rot("y", g("axis_angle_y")) @ _rot("z", g("axis_angle_z")) @ X
    B = _rot("x", g("detector_angle_x")) @ _rot("y", g("detector_angle_y")) @ _rot("z", g("detector_angle_z"))
    e_u, e_v, n = B[:, 0], B[:, 2], -B[:, 1]
    D = np.array([g("detector_position_x"), dy, g("detector_position_z")])[:, None]
    if cone:
        S = np.array([g("source_position_x"), sy, g("source_position_z")])[:, None]
        origin, direction = np.broadcast_to(S, X.shape), X - S
        weight = ((sy - dy) / (sy - X[1])) ** 2
    else:
        origin, direction = X, np.broadcast_to(np.array([[0.0], [1.0], [0.0]]), X.shape)
        weight = np.ones(X.shape[1])
    t = (n @ (D - origin)) / (n @ direction)
    hit = origin + t * direction - D
    return e_u @ hit + g("center_position_x"), e_v @ hit + g("center_position_z"), weight


def ray_model_volume(projs, geometry, overall_angle, x_region, y_region, region):
    xs, ys, zs = _grid(x_region), _grid(y_region), _grid(region)
    Z, Y, Xg = np.meshgrid(zs, ys, xs, indexing="ij")
    X0 = np.stack([Xg.ravel(), Y.ravel(), Z.ravel()])
    p = len(projs)
    out = np.zeros(X0.shape[1])
    for i in range(p):
        u, v, w = ray_model_projection(geometry, i, X0)
        out += w * _bilinear_zero_border(projs[i], u, v)
    return (out * abs(overall_angle) / p).reshape(Z.shape)


# --------------------------------------------------------------- oracle probes

def _coordinate_images():
    iu = np.broadcast_to(np.arange(W, dtype=np.float32) + 0.5, (H, W))
    iv = np.broadcast_to((np.arange(H, dtype=np.float32) + 0.5)[:, None], (H, W))
    return iu, iv, np.ones((H, W), np.float32)


def oracle_probe(theta=0.0, **geometry):
    """(u + cx, v + cz, coeff^2) of every GRID voxel seen by one projection at angle theta."""
    geometry.setdefault("center_position_x", [C0])
    geometry.setdefault("center_position_z", [C0])
    vols = [oracle.gbp(img[None], num_projections=1, overall_angle=1.0, axis_angle_z=[theta],
                       **GRID, **geometry) for img in _coordinate_images()]
    return vols[0] / vols[2], vols[1] / vols[2], vols[2]


def model_probe(theta=0.0, **geometry):
    geometry.setdefault("center_position_x", [C0])
    geometry.setdefault("center_position_z", [C0])
    geometry["axis_angle_z"] = [theta]
    xs, ys, zs = (_grid(GRID[k]) for k in ("x_region", "y_region", "region"))
    Z, Y, Xg = np.meshgrid(zs, ys, xs, indexing="ij")
    u, v, w = ray_model_projection(geometry, 0, np.stack([Xg.ravel(), Y.ravel(), Z.ravel()]))
    return u.reshape(Z.shape), v.reshape(Z.shape), w.reshape(Z.shape)


def _xyz():
    xs, ys, zs = (_grid(GRID[k]) for k in ("x_region", "y_region", "region"))
    return np.meshgrid(zs, ys, xs, indexing="ij")[::-1]  # x, y, z arrays of shape [nz][ny][nx]


def _check(got, closed_u, closed_v, closed_w=None, tol=1e-9):
    u, v, w = got
    np.testing.assert_allclose(u, closed_u + C0, atol=tol, rtol=0)
    np.testing.assert_allclose(v, closed_v + C0, atol=tol, rtol=0)
    if closed_w is not None:
        np.testing.assert_allclose(w, closed_w, rtol=1e-12, atol=0)


# Each closed form: (geometry, function of (x, y, z) -> (u, v, w or None)).
ALPHA = 0.37
SY, DY = -300.0, 500.0
M = -SY / (DY - SY)  # magnification of the voxel grid in cone beam


def _cf_detector_x(x, y, z):
    # Parallel beam, detector tilted about x by a: normal (0, -cos a, -sin a), v axis (0, -sin a, cos a).
    # The ray (x, y + s, z) meets the plane at y + s - D_y = -tan(a) z, so v = sin^2 z / cos + cos z.
    return x, z / math.cos(ALPHA), None


def _cf_detector_y(x, y, z):
    # Rotation about y leaves the normal (0, -1, 0); the frame turns in the x-z plane:
    # u axis (cos a, 0, -sin a), v axis (sin a, 0, cos a).
    c, s = math.cos(ALPHA), math.sin(ALPHA)
    return c * x - s * z, s * x + c * z, None


def _cf_detector_z(x, y, z):
    # Normal (sin a, -cos a, 0), u axis (cos a, sin a, 0): the ray meets the plane at y + s - D_y = tan(a) x.
    return x / math.cos(ALPHA), z, None


def _cf_detector_shift(x, y, z):
    return x - 7.25, z + 3.5, None


def _cf_detector_shift_tilted_z(x, y, z):
    # Plane through (d, D_y, 0) tilted about z: u = (x - d) / cos a.
    return (x - 7.25) / math.cos(ALPHA), z, None


def _cf_volume_x(x, y, z):
    c, s = math.cos(ALPHA), math.sin(ALPHA)
    return x, s * y + c * z, None


def _cf_volume_y(x, y, z):
    c, s = math.cos(ALPHA), math.sin(ALPHA)
    return c * x + s * z, -s * x + c * z, None


def _cf_volume_xy(x, y, z):
    # rotate_y by b first, then rotate_x by a (:786-793): X = Rx(a) Ry(b) X0.
    a, b = ALPHA, -0.61
    xb, zb = math.cos(b) * x + math.sin(b) * z, -math.sin(b) * x + math.cos(b) * z
    return xb, math.sin(a) * y + math.cos(a) * zb, None


def _cf_axis_xy(x, y, z):
    # Same composition for the axis tilt after the tomographic rotation at theta = 0.
    return _cf_volume_xy(x, y, z)


def _cf_cone_plain(x, y, z):
    # Similar triangles from S = (0, S_y, 0) to the plane y = D_y, with the grid scaled by M.
    X, Y, Z = M * x, M * y, M * z
    t = (DY - SY) / (Y - SY)
    return t * X, t * Z, ((SY - DY) / (SY - Y)) ** 2


def _cf_cone_source_shift(x, y, z):
    sx, sz = 11.0, -6.5
    X, Y, Z = M * x, M * y, M * z
    t = (DY - SY) / (Y - SY)
    return sx + t * (X - sx), sz + t * (Z - sz), ((SY - DY) / (SY - Y)) ** 2


def _cf_cone_detector_x(x, y, z):
    # Cone beam onto a plane through (0, D_y, 0) tilted about x by a:
    # t = cos a (D_y - S_y) / (cos a (Y - S_y) + sin a Z), u = t X, v = -sin a (S_y - D_y + t (Y - S_y)) + cos a t Z.
    c, s = math.cos(ALPHA), math.sin(ALPHA)
    X, Y, Z = M * x, M * y, M * z
    t = c * (DY - SY) / (c * (Y - SY) + s * Z)
    return t * X, -s * (SY - DY + t * (Y - SY)) + c * t * Z, ((SY - DY) / (SY - Y)) ** 2


CLOSED_FORMS = {
    "detector-angle-x": (dict(detector_angle_x=[ALPHA]), _cf_detector_x),
    "detector-angle-y": (dict(detector_angle_y=[ALPHA]), _cf_detector_y),
    "detector-angle-z": (dict(detector_angle_z=[ALPHA]), _cf_detector_z),
    "detector-position-x/z": (dict(detector_position_x=[7.25], detector_position_z=[-3.5]), _cf_detector_shift),
    "detector-position-x tilted": (dict(detector_position_x=[7.25], detector_angle_z=[ALPHA]),
                                   _cf_detector_shift_tilted_z),
    "volume-angle-x": (dict(volume_angle_x=[ALPHA]), _cf_volume_x),
    "volume-angle-y": (dict(volume_angle_y=[ALPHA]), _cf_volume_y),
    "volume-angle-x,y": (dict(volume_angle_x=[ALPHA], volume_angle_y=[-0.61]), _cf_volume_xy),
    "axis-angle-x,y": (dict(axis_angle_x=[ALPHA], axis_angle_y=[-0.61]), _cf_axis_xy),
    "cone": (dict(source_position_y=[SY], detector_position_y=[DY]), _cf_cone_plain),
    "cone source-position-x/z": (dict(source_position_y=[SY], detector_position_y=[DY],
                                      source_position_x=[11.0], source_position_z=[-6.5]), _cf_cone_source_shift),
    "cone detector-angle-x": (dict(source_position_y=[SY], detector_position_y=[DY], detector_angle_x=[ALPHA]),
                              _cf_cone_detector_x),
}


@pytest.mark.parametrize("name", sorted(CLOSED_FORMS))
def test_oracle_closed_forms(name):
    geometry, cf = CLOSED_FORMS[name]
    x, y, z = _xyz()
    cu, cv, cw = cf(x, y, z)
    _check(oracle_probe(**{k: list(v) for k, v in geometry.items()}), cu, cv, cw)


@pytest.mark.parametrize("name", sorted(CLOSED_FORMS))
def test_ray_model_closed_forms(name):
    # The model itself meets the same hand-derived values, so model and closed forms pin each other.
    geometry, cf = CLOSED_FORMS[name]
    x, y, z = _xyz()
    cu, cv, cw = cf(x, y, z)
    _check(model_probe(**{k: list(v) for k, v in geometry.items()}), cu, cv, cw)


@pytest.mark.parametrize("cone", [False, True])
def test_volume_angle_z_is_a_tomographic_offset(cone):
    # Rz(theta) Rz(phi) = Rz(theta + phi): a volume rotation about z by phi equals shifting every
    # tomographic angle by phi (any projection data, any number of projections).
    rng = np.random.default_rng(5)
    p = 7
    projs = rng.standard_normal((p, 64, 64)).astype(np.float32)
    phi = 0.83
    theta = np.arange(p) * 2 * math.pi / p
    geo = dict(center_position_x=[32.0], center_position_z=[31.0])
    if cone:
        geo.update(source_position_y=[-150.0], detector_position_y=[90.0])
    grid = dict(x_region=(-12.0, 12.0, 3.0), y_region=(-9.0, 9.0, 3.0), region=(-6.0, 6.0, 4.0))
    a = oracle.gbp(projs, axis_angle_z=list(theta), volume_angle_z=[phi], **grid, **geo)
    b = oracle.gbp(projs, axis_angle_z=list(theta + phi), **grid, **geo)
    np.testing.assert_allclose(a, b, atol=1e-10)


def _random_geometry(rng, p, cone, tilted, shifted, volume, axis):
    g = dict(center_position_x=list(32.0 + rng.uniform(-2, 2, p)), center_position_z=list(30.0 + rng.uniform(-2, 2, p)),
             axis_angle_z=list(np.sort(rng.uniform(0, 2 * math.pi, p))))
    if cone:
        g["source_position_y"] = list(rng.uniform(-220, -160, p))
        g["detector_position_y"] = list(rng.uniform(60, 120, p))
        if shifted:
            g["source_position_x"] = list(rng.uniform(-5, 5, p))
            g["source_position_z"] = list(rng.uniform(-5, 5, p))
    if tilted:
        for k in ("x", "y", "z"):
            g["detector_angle_" + k] = list(rng.uniform(-0.35, 0.35, p))
    if shifted:
        g["detector_position_x"] = list(rng.uniform(-4, 4, p))
        g["detector_position_z"] = list(rng.uniform(-4, 4, p))
    if volume:
        for k in ("x", "y", "z"):
            g["volume_angle_" + k] = list(rng.uniform(-0.5, 0.5, p))
    if axis:
        g["axis_angle_x"] = list(rng.uniform(-0.4, 0.4, p))
        g["axis_angle_y"] = list(rng.uniform(-0.4, 0.4, p))
    return g


VARIANTS = [
    # cone, tilted detector, shifted source/detector, volume rotation, axis tilt
    (False, True, False, False, False),
    (False, True, True, False, False),
    (False, False, True, True, False),
    (False, True, True, True, True),
    (True, False, False, False, False),
    (True, False, True, False, False),
    (True, True, False, False, False),
    (True, True, True, True, True),
    (True, False, True, True, True),
]
MODEL_GRID = dict(x_region=(-14.0, 14.0, 4.0), y_region=(-12.0, 12.0, 4.0), region=(-9.0, 9.0, 6.0))


def _model_vs_oracle(seed, variant):
    rng = np.random.default_rng(seed)
    p = 6
    projs = rng.standard_normal((p, 64, 64)).astype(np.float32)
    geo = _random_geometry(rng, p, *variant)
    got = oracle.gbp(projs, overall_angle=2 * math.pi, **MODEL_GRID, **geo)
    ref = ray_model_volume(projs, geo, 2 * math.pi, **MODEL_GRID)
    return got, ref


@pytest.mark.parametrize("k", range(len(VARIANTS)))
def test_oracle_matches_ray_model(k):
    got, ref = _model_vs_oracle(100 + k, VARIANTS[k])
    np.testing.assert_allclose(got, ref, atol=1e-9, rtol=0)


@pytest.mark.parametrize("k", range(len(VARIANTS)))
def test_ref32_within_1e5_of_f64_on_variants(k):
    rng = np.random.default_rng(200 + k)
    p = 6
    projs = rng.uniform(0.5, 1.5, (p, 64, 64)).astype(np.float32)
    geo = _random_geometry(rng, p, *VARIANTS[k])
    f64 = oracle.gbp(projs, **MODEL_GRID, **geo)
    f32 = oracle.gbp(projs, mode="ref32", **MODEL_GRID, **geo)
    assert oracle.rel_l2(f32, f64) < 1e-5


# ------------------------------------------------------------------ planted mistakes

PLANTS = {
    1: ("inverse detector rotation about x keeps +sin", ["detector-angle-x", "cone detector-angle-x"]),
    2: ("shifted-source sign flipped (perpendicular cone)", ["cone source-position-x/z"]),
    4: ("volume rotation order reversed", ["volume-angle-x,y"]),
    8: ("detector normal rotation order reversed", []),
    16: ("cone weight without the voxel's depth", ["cone", "cone source-position-x/z"]),
}


@pytest.mark.parametrize("bit", sorted(PLANTS))
def test_planted_geometry_mistakes_are_caught(bit):
    why, forms = PLANTS[bit]
    x, y, z = _xyz()
    with oracle.gbp_planted(bit):
        for name in forms:
            geometry, cf = CLOSED_FORMS[name]
            cu, cv, cw = cf(x, y, z)
            with pytest.raises(AssertionError):
                _check(oracle_probe(**{k: list(v) for k, v in geometry.items()}), cu, cv, cw)
        # the random-geometry comparison with the ray model catches every plant on some variant
        caught = 0
        for k, variant in enumerate(VARIANTS):
            got, ref = _model_vs_oracle(100 + k, variant)
            caught += np.abs(got - ref).max() > 1e-6
        assert caught > 0, why
    # and with the plant cleared, the oracle is back to the faithful restatement
    got, ref = _model_vs_oracle(100 + 3, VARIANTS[3])
    np.testing.assert_allclose(got, ref, atol=1e-9, rtol=0)


# ------------------------------------------------------------ result type and addressing

def test_round_half_vs_numpy_float16():
    # numpy converts float64 to binary16 with its own round-to-nearest-even code
    rng = np.random.default_rng(3)
    x = np.concatenate([
        rng.standard_normal(4000) * np.exp(rng.uniform(-20, 11, 4000)),   # normals and subnormals
        (np.arange(-2048, 2048) + 0.5) * 2.0 ** -10,                       # ties at the 2^-10 quantum of [1, 2)
        [65504.0, 65519.99, 65520.0, -65520.0, 2.0 ** -24, 2.0 ** -25, 3 * 2.0 ** -26, 0.0, -0.0],
    ])
    got = np.array([oracle.round_half(v) for v in x])
    with np.errstate(over="ignore"):  # 65520 and beyond become infinity, as they should
        ref = x.astype(np.float16).astype(np.float64)
    np.testing.assert_array_equal(got, ref)


def _sample_line(values, xs, addressing_mode, mode="f64"):
    """Parallel beam, one projection at theta = 0, one detector row: voxel x samples u = x + cx."""
    w = len(values)
    projs = np.asarray(values, np.float32)[None, None, :]
    cx = 0.0
    out = [oracle.gbp(projs, num_projections=1, overall_angle=1.0, axis_angle_z=[0.0], center_position_x=[cx],
                      center_position_z=[0.5], x_region=(x, x + 1.0, 1.0), y_region=(0.0, 1.0, 1.0),
                      addressing_mode=addressing_mode, mode=mode)[0, 0, 0] for x in xs]
    return np.array(out)


@pytest.mark.parametrize("mode", ["f64", "ref32"])
def test_repeat_addressing_invariants(mode):
    rng = np.random.default_rng(8)
    w = 12
    vals = rng.uniform(-1, 1, w)
    inner = np.linspace(0.6, w - 0.6, 23)
    # inside the image every mode is the plain linear interpolation
    base = _sample_line(vals, inner, "clamp_to_edge", mode)
    for m in ("repeat", "mirrored_repeat"):
        np.testing.assert_allclose(_sample_line(vals, inner, m, mode), base, atol=1e-6)
    # repeat: period w, and at the seam u = 0 half of the last and half of the first texel
    xs = np.linspace(-30.3, 30.1, 41)
    np.testing.assert_allclose(_sample_line(vals, xs + w, "repeat", mode), _sample_line(vals, xs, "repeat", mode),
                               atol=2e-5)
    np.testing.assert_allclose(_sample_line(vals, [0.0, float(w)], "repeat", mode),
                               [(vals[-1] + vals[0]) / 2] * 2, atol=1e-6)
    # mirrored repeat: even about 0 and about w, period 2 w, edge texels held at the mirror lines
    np.testing.assert_allclose(_sample_line(vals, -xs, "mirrored_repeat", mode),
                               _sample_line(vals, xs, "mirrored_repeat", mode), atol=2e-5)
    np.testing.assert_allclose(_sample_line(vals, 2 * w - xs, "mirrored_repeat", mode),
                               _sample_line(vals, xs, "mirrored_repeat", mode), atol=2e-5)
    np.testing.assert_allclose(_sample_line(vals, [0.0, float(w)], "mirrored_repeat", mode), [vals[0], vals[-1]],
                               atol=1e-6)


@pytest.mark.parametrize("rt", ["half", "double"])
def test_result_types_ref32(rt):
    # half: the running sum is rounded to binary16 after every projection (a burst sum of
    # values near 1 is then a multiple of 2^-10 .. 2^-7), so it differs from float by about
    # the half rounding; double differs from float only by float rounding
    rng = np.random.default_rng(12)
    p = 20
    projs = rng.uniform(0.5, 1.5, (p, 24, 24)).astype(np.float32)
    geo = dict(center_position_x=[12.0], center_position_z=[12.0], region=(-2.0, 2.0, 2.0),
               x_region=(-6.0, 6.0, 1.0), y_region=(-6.0, 6.0, 1.0))
    f64 = oracle.gbp(projs, **geo)
    f32 = oracle.gbp(projs, mode="ref32", **geo)
    got = oracle.gbp(projs, mode="ref32", result_type=rt, **geo)
    if rt == "double":
        assert oracle.rel_l2(got, f64) < 1e-6
    else:
        err = oracle.rel_l2(got, f64)
        assert 1e-5 < err < 3e-3, err
        assert oracle.rel_l2(f32, f64) < 1e-6
        # with bursts of one projection every burst sum is a single rounded sample
        one = oracle.gbp(projs, mode="ref32", result_type=rt, burst=1, **geo)
        assert oracle.rel_l2(one, f64) < err
