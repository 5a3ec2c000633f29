"""CPU checks of the drop-in boundary: libufomi.so loads, exports every symbol
the headers declare, and its host-only parts (filter coefficients, phantom
parameters, volume shapes, task property tables, pipeline parsing) follow the
reference.  No GPU compute is called here.
"""
import ctypes
import math
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
import ufo_tasks
import ufomi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ufomi_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ufomi.load()
    out = subprocess.run(["nm", "-D", "--defined-only", ufomi.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (\w+)", out))
    for header in ("ufomi.h", "ufo_tasks.h"):
        names = _declared(header)
        assert names, header
        missing = [n for n in names if n not in exported]
        assert not missing, (header, missing)
        for n in names:
            getattr(lib, n)
    assert set(ufomi.EXPORTS) <= exported
    assert set(ufo_tasks.TASK_EXPORTS) <= exported


def test_version_and_no_context_without_gpu():
    assert ufomi.load().ufomi_version().decode().startswith("ufomi")


@pytest.mark.parametrize("filt", ["ramp", "ramp-fromreal", "butterworth", "faris-byer", "hamming", "bh3",
                                  "shepp-logan"])
@pytest.mark.parametrize("width", [8, 64, 1024, 8192])
def test_filter_coefficients_match_oracle(filt, width):
    a = ufomi.filter_coefficients(width, filt)
    b = oracle.filter_coefficients(width, filt)
    # both evaluate in FP64 and round once; FFT orders differ, so allow one float ulp
    np.testing.assert_allclose(a, b, rtol=2.5e-7, atol=1e-9 * np.abs(b).max())


def test_filter_coefficient_errors():
    with pytest.raises(ufomi.UfomiError) as e:
        ufomi.filter_coefficients(96)
    assert e.value.code == -3
    assert "must be a power of two" in str(e.value)


def test_shepp_logan_is_windowed_ramp():
    n = 256
    ram = ufomi.filter_coefficients(2 * n, "ramp-fromreal")[0::2]
    sl = ufomi.filter_coefficients(2 * n, "shepp-logan")[0::2]
    k = np.minimum(np.arange(n), n - np.arange(n))
    x = np.pi * k / n
    win = np.where(k == 0, 1.0, np.sin(x) / np.where(k == 0, 1, x))
    np.testing.assert_allclose(sl, (ram * win).astype(np.float32), rtol=1e-6, atol=1e-9)


def test_phantom_parameters_seeded_and_in_range():
    a = ufomi.phantom_ellipses(7, 1000, 100.0)
    b = ufomi.phantom_ellipses(7, 1000, 100.0)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, ufomi.phantom_ellipses(8, 1000, 100.0))
    assert np.all(np.hypot(a[:, 0], a[:, 1]) <= 70.0 + 1e-9)
    assert np.all((a[:, 2:4] >= 5.0) & (a[:, 2:4] <= 40.0))
    assert np.all((a[:, 4] >= 0) & (a[:, 4] < math.pi))
    assert np.all((a[:, 5] >= -0.5) & (a[:, 5] <= 1.0))
    e3 = ufomi.phantom_ellipsoids(3, 500, 100.0)
    assert np.all(np.linalg.norm(e3[:, :3], axis=1) <= 70.0 + 1e-9)


def test_gbp_volume_shape():
    p = ufomi.gbp_params(10)
    assert ufomi.gbp_volume_shape(p, 64) == (1, 64, 64)
    p = ufomi.gbp_params(10, region=(-2, 2, 0.5), x_region=(-10, 10, 0.5), y_region=(0, 7, 2))
    assert ufomi.gbp_volume_shape(p, 64) == (8, 4, 40)


# --------------------------------------------------------------- task layer

REFERENCE_PROPERTIES = {
    # name: {property: default} as the reference's tasks report them after init
    "fft": {"auto-zeropadding": False, "dimensions": 1, "size-x": 0, "size-y": 0, "size-z": 0},
    "filter": {"filter": "ramp-fromreal", "cutoff": 0.5, "order": 4, "tau": float(np.float32(0.1)), "theta": 1,
               "scale": 1},
    "ifft": {"dimensions": 1, "crop-width": 0, "crop-height": 0},
    "backproject": {"num-projections": 0, "offset": 0, "axis-pos": -1, "angle-step": -1, "angle-offset": 0,
                    "mode": "texture", "roi-x": 0, "roi-y": 0, "roi-width": 0, "roi-height": 0},
    "cone-beam-projection-weight": {"center-position-x": [], "center-position-z": [], "source-distance": [],
                                    "detector-distance": [], "axis-angle-x": 0.0},
    "general-backproject": {"burst": 0, "parameter": "z", "z": 0, "region": [0.0, 0.0, 0.0],
                            "x-region": [0.0, 0.0, 0.0], "y-region": [0.0, 0.0, 0.0], "center-position-x": 0.0,
                            "source-position-y": -math.inf, "num-projections": 0, "compute-type": "float",
                            "result-type": "float", "store-type": "float", "overall-angle": 2 * math.pi,
                            "addressing-mode": "clamp", "gray-map-min": 0, "gray-map-max": 0},
    "lamino-backproject": {"x-region": [0.0, 0.0, 0.0], "y-region": [0.0, 0.0, 0.0], "z": 0.0,
                           "region": [0.0, 0.0, 0.0], "projection-offset": [0.0, 0.0], "center": [0.0, 0.0],
                           "num-projections": 0, "overall-angle": float(np.float32(math.pi)),
                           "tomo-angle": -float(np.finfo(np.float32).max), "lamino-angle": 0.0, "roll-angle": 0.0,
                           "parameter": "z", "addressing-mode": "clamp"},
}


@pytest.mark.parametrize("name", sorted(REFERENCE_PROPERTIES))
def test_task_property_defaults(name):
    t = ufo_tasks.Task(name)
    props = t.list_properties()
    for prop, default in REFERENCE_PROPERTIES[name].items():
        assert prop in props, (name, prop)
        got = t.get_property(prop)
        if isinstance(default, float) and math.isinf(default):
            assert math.isinf(got) and got < 0
        elif isinstance(default, float):
            assert got == pytest.approx(default)
        else:
            assert got == default, (name, prop, got)


def test_task_property_ranges_and_enums():
    fft = ufo_tasks.Task("fft")
    fft.props.size_x = 1024
    assert fft.props.size_x == 1024
    with pytest.raises(ufomi.UfomiError):
        fft.props.size_x = 40000          # 0..32768
    assert fft.props.size_x == 1024       # unchanged, like g_object_set with an invalid value
    with pytest.raises(ufomi.UfomiError):
        fft.props.dimensions = 4
    filt = ufo_tasks.Task("filter")
    for nick in ("ramp", "ramp-fromreal", "butterworth", "faris-byer", "hamming", "bh3"):
        filt.props.filter = nick
        assert filt.props.filter == nick
    with pytest.raises(ufomi.UfomiError):
        filt.props.filter = "shepp"
    with pytest.raises(ufomi.UfomiError):
        filt.props.cutoff = 1.5
    bp = ufo_tasks.Task("backproject")
    with pytest.raises(ufomi.UfomiError):
        bp.props.axis_pos = -2.0          # -1..32768
    bp.props.mode = "nearest"
    assert bp.props.mode == "nearest"
    with pytest.raises(ufomi.UfomiError):
        bp.props.angle_offset = -0.1      # 0..G_MAXDOUBLE
    gbp = ufo_tasks.Task("general-backproject")
    gbp.props.source_position_y = [-2500.0]
    gbp.props.region = [-512.0, 512.0, 1.0]
    assert gbp.props.region == [-512.0, 512.0, 1.0]
    with pytest.raises(ufomi.UfomiError):
        gbp.props.burst = 129
    with pytest.raises(ufomi.UfomiError):
        gbp.set_property("no-such-property", 1)
    assert gbp.mode & ufo_tasks.MODE_REDUCTOR
    assert bp.mode & ufo_tasks.MODE_PROCESSOR


def test_lamino_shape_and_region_property():
    """REGION_SIZE on integer regions, ceil in float32 for the swept region; set_region
    (ufo-lamino-backproject-task.c:124-136) ignores a region whose "from" exceeds "to"."""
    p = ufomi.lamino_params(10, x_region=(-10, 10, 3), y_region=(0, 7, 2), region=(0.0, 1.0, 0.3))
    assert ufomi.lamino_shape(p) == (4, 4, 7)
    assert oracle.lamino_shape((-10, 10, 3), (0, 7, 2), (0.0, 1.0, 0.3)) == (7, 4, 4)
    with pytest.raises(ufomi.UfomiError):
        ufomi.lamino_shape(ufomi.lamino_params(10))
    t = ufo_tasks.Task("lamino-backproject")
    t.props.x_region = [-4, 4, 1]
    t.props.x_region = [5, 1, 1]
    assert t.props.x_region == [-4.0, 4.0, 1.0]
    t.props.parameter = "roll-angle"
    assert t.props.parameter == "roll-angle" and t.mode & ufo_tasks.MODE_REDUCTOR


def test_unknown_task_and_pipeline_errors():
    with pytest.raises(ValueError):
        ufo_tasks.Task("no-such-task")
    with pytest.raises(ufo_tasks.TaskError):
        ufo_tasks.Pipeline("fft ! frobnicate")
    with pytest.raises(ufo_tasks.TaskError):
        ufo_tasks.Pipeline("fft size-x")
    with pytest.raises(ufo_tasks.TaskError):
        ufo_tasks.Pipeline("fft size-x=abc")
    pl = ufo_tasks.Pipeline("fft dimensions=1 size-x=1024 ! filter ! ifft crop-width=512 ! backproject axis-pos=256")
    assert pl.task(0).get_property("size-x") == 1024
    assert pl.task(2).get_property("crop-width") == 512
    assert pl.task(3).get_property("axis-pos") == 256
