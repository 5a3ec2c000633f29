"""ctypes front end of the CPU oracle (oracle/ufo_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the test oracle.  The product library never loads
it.  Each wrapper names the reference code its C function restates.
"""
import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

FILTERS = {"ramp": 0, "ramp-fromreal": 1, "butterworth": 2, "faris-byer": 3, "hamming": 4, "bh3": 5,
           "shepp-logan": 6}
ADDR = {"clamp": 0, "clamp_to_edge": 1, "repeat": 2, "mirrored_repeat": 3}
RESULT_TYPES = {"half": 0, "float": 1, "double": 2}

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_lib = None


class OrcGbpGeom(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "tomo", "axis_x", "axis_y", "vol_x", "vol_y", "vol_z", "det_ax", "det_ay", "det_az",
        "cx", "cz", "sx", "sy", "sz", "dx", "dy", "dz")]


class OrcGbpFlags(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "with_axis", "with_volume", "perpendicular_detector", "shifted_detector",
        "shifted_source", "parallel_beam")]


def build():
    """Compile liboracle.so with the Makefile beside this file."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE, "-B", "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        c_int, c_uint, c_float, c_double = ctypes.c_int, ctypes.c_uint, ctypes.c_float, ctypes.c_double
        L.orc_fft_c2c.argtypes = [_f64p, _f64p, c_int, c_int]
        L.orc_dft_naive.argtypes = [_f64p, _f64p, _f64p, _f64p, c_int, c_int]
        L.orc_filter_coefficients.argtypes = [c_int, c_uint, c_float, c_float, c_float, c_float, c_float, _f32p]
        L.orc_filter_coefficients.restype = c_int
        L.orc_fft_task.argtypes = [_f32p, c_int, c_int, c_int, _f64p]
        L.orc_filter_task.argtypes = [_f64p, c_int, c_int, _f32p]
        L.orc_ifft_task.argtypes = [_f64p, c_int, c_int, c_int, _f64p]
        L.orc_ramp_chain.argtypes = [_f32p, c_int, c_int, c_int, _f32p, c_int, _f64p]
        L.orc_ramp_chain_f32.argtypes = [_f32p, c_int, c_int, c_int, _f32p, c_int, _f32p]
        L.orc_bp_lut.argtypes = [c_double, c_double, c_int, _f32p, _f32p]
        L.orc_bp_texture_ref32.argtypes = [_f32p, c_int, c_int, _f32p, _f32p, c_uint, c_float,
                                           c_uint, c_uint, c_int, c_int, _f32p]
        L.orc_bp_texture_f64.argtypes = [_f64p, c_int, c_int, c_double, c_double, c_uint, c_double,
                                         c_uint, c_uint, c_int, c_int, _f64p]
        L.orc_bp_nearest_ref32.argtypes = [_f32p, c_int, c_int, _f32p, _f32p, c_uint, c_float,
                                           c_uint, c_uint, c_int, c_int, _f32p]
        L.orc_cone_weight_ref32.argtypes = [_f32p, _f32p, c_int, c_int, c_double, c_double,
                                            c_double, c_double, c_double]
        L.orc_gbp.argtypes = [_f32p, c_int, c_int, c_int, ctypes.POINTER(OrcGbpGeom),
                              ctypes.POINTER(OrcGbpFlags), c_int, c_int, c_int, c_int,
                              c_double, c_double, c_double, c_double, c_double, c_double,
                              c_double, c_int, c_int, c_int, _f64p]
        L.orc_round_half.argtypes = [c_double]
        L.orc_round_half.restype = c_double
        L.orc_lamino.argtypes = [_f32p, c_int, c_int, c_int, c_uint, c_float, c_float, c_float, c_float,
                                 c_float, c_float, c_int, c_int, _i32p, _i32p, _f32p, c_float, c_int, c_int,
                                 c_int, c_int, c_int, c_int, c_int, _f64p]
        _lib = L
    return _lib


# --------------------------------------------------------------------------- FFT

def fft_c2c(x, direction=-1):
    """Unnormalised FP64 FFT, exp(direction * 2 pi i k n / N) (oclFFT.h:67-72)."""
    x = np.asarray(x, dtype=np.complex128)
    re = np.ascontiguousarray(x.real, dtype=np.float64)
    im = np.ascontiguousarray(x.imag, dtype=np.float64)
    lib().orc_fft_c2c(re, im, len(x), int(direction))
    return re + 1j * im


def dft_naive(x, direction=-1):
    x = np.asarray(x, dtype=np.complex128)
    re = np.ascontiguousarray(x.real, dtype=np.float64)
    im = np.ascontiguousarray(x.imag, dtype=np.float64)
    o_re = np.empty_like(re)
    o_im = np.empty_like(im)
    lib().orc_dft_naive(re, im, o_re, o_im, len(x), int(direction))
    return o_re + 1j * o_im


def _dft_matrix(m, direction):
    k = np.arange(m)
    return np.exp(direction * 2j * np.pi * (np.outer(k, k) % m) / m)


def fft_nd(x, dimensions, m, in_complex=False):
    """The fft task (src/ufo-fft-task.c:122-346, src/common/ufo-fft.c:161-354) by
    definition: x [n2][n1][n0] (real, or complex when in_complex) zero-padded to
    m = (m0, m1, m2) (x first), DFT exp(-2 pi i k n / M) over the first
    `dimensions` axes (x first), complex128 [m2][m1][m0]."""
    x = np.asarray(x, dtype=np.complex128 if in_complex else np.float64)
    while x.ndim < 3:
        x = x[None]
    out = np.zeros((m[2], m[1], m[0]), dtype=np.complex128)
    out[:x.shape[0], :x.shape[1], :x.shape[2]] = x
    for a in range(dimensions):
        ax = 2 - a
        out = np.moveaxis(np.tensordot(out, _dft_matrix(m[a], -1), axes=([ax], [0])), -1, ax)
    return out


def ifft_nd(spec, dimensions, crop_width=0, crop_height=0):
    """The ifft task (src/ufo-ifft-task.c:111-306): inverse DFT of complex
    [m2][m1][m0] over the first `dimensions` axes, divided by their sizes,
    real part cropped to [m2][crop_height][crop_width] (0 = no crop)."""
    out = np.asarray(spec, dtype=np.complex128)
    while out.ndim < 3:
        out = out[None]
    for a in range(dimensions):
        ax = 2 - a
        n = out.shape[ax]
        out = np.moveaxis(np.tensordot(out, _dft_matrix(n, +1), axes=([ax], [0])), -1, ax) / n
    ch = crop_height or out.shape[1]
    cw = crop_width or out.shape[2]
    return out.real[:, :ch, :cw]


# ------------------------------------------------------------------- filter chain

def filter_coefficients(width, filter="ramp-fromreal", scale=1.0, cutoff=0.5, order=4.0, tau=0.1, theta=1.0):
    """2N float32 coefficients of ufo-filter-task.c:290-335 (width = 2N)."""
    out = np.empty(width, dtype=np.float32)
    rc = lib().orc_filter_coefficients(FILTERS[filter], width, scale, cutoff, order, tau, theta, out)
    if rc != 0:
        raise ValueError("Invalid filter input width %u, must be a power of two" % width)
    return out


def fft_task(sino, n):
    """fft dimensions=1 size-x=n: [P][W] float32 -> [P][2n] interleaved complex (FP64)."""
    sino = np.ascontiguousarray(sino, dtype=np.float32)
    p, w = sino.shape
    out = np.empty((p, 2 * n), dtype=np.float64)
    lib().orc_fft_task(sino, w, p, n, out)
    return out


def filter_task(spec, coef):
    spec = np.ascontiguousarray(spec, dtype=np.float64).copy()
    lib().orc_filter_task(spec, spec.shape[1], spec.shape[0], np.ascontiguousarray(coef, dtype=np.float32))
    return spec


def ifft_task(spec, crop=0):
    spec = np.ascontiguousarray(spec, dtype=np.float64).copy()
    p, w2 = spec.shape
    n = w2 // 2
    crop = crop or n
    out = np.empty((p, crop), dtype=np.float64)
    lib().orc_ifft_task(spec, n, p, crop, out)
    return out


def ramp_chain(sino, n, coef, crop=0):
    """`fft size-x=n ! filter ! ifft crop-width=crop` in FP64 -> [P][crop] float64."""
    sino = np.ascontiguousarray(sino, dtype=np.float32)
    p, w = sino.shape
    crop = crop or n
    out = np.empty((p, crop), dtype=np.float64)
    lib().orc_ramp_chain(sino, w, p, n, np.ascontiguousarray(coef, dtype=np.float32), crop, out)
    return out


def ramp_chain_f32(sino, n, coef, crop=0):
    """The same chain in float32 throughout (radix-2 transforms, rows in parallel) -> [P][crop] float32."""
    sino = np.ascontiguousarray(sino, dtype=np.float32)
    p, w = sino.shape
    crop = crop or n
    out = np.empty((p, crop), dtype=np.float32)
    lib().orc_ramp_chain_f32(sino, w, p, n, np.ascontiguousarray(coef, dtype=np.float32), crop, out)
    return out


# -------------------------------------------------------------- backprojection

def bp_lut(n, angle_step=-1.0, angle_offset=0.0):
    """sin/cos LUTs of ufo-backproject-task.c:174-196, 246-251."""
    step = math.pi / n if angle_step <= 0 else angle_step
    s = np.empty(n, dtype=np.float32)
    c = np.empty(n, dtype=np.float32)
    lib().orc_bp_lut(float(angle_offset), float(step), n, s, c)
    return s, c, step


def _bp_defaults(w, p_in, num_projections, axis_pos, roi_width, roi_height):
    n_proj = num_projections or p_in
    if p_in > n_proj:
        raise ValueError("Total number of projections (%u) must be greater than or equal to sinogram "
                         "height (%u)" % (n_proj, p_in))
    axis = float(np.float32(np.float32(w) / np.float32(2.0))) if axis_pos <= 0 else float(np.float32(axis_pos))
    return n_proj, axis, roi_width or w, roi_height or w


def bp_texture_ref32(sino, axis_pos=-1.0, angle_step=-1.0, angle_offset=0.0, num_projections=0,
                     offset=0, roi_x=0, roi_y=0, roi_width=0, roi_height=0):
    """backproject mode=texture with float32 semantics (backproject.cl:50-94)."""
    sino = np.ascontiguousarray(sino, dtype=np.float32)
    p_in, w = sino.shape
    n_proj, axis, rw, rh = _bp_defaults(w, p_in, num_projections, axis_pos, roi_width, roi_height)
    s, c, _ = bp_lut(n_proj, angle_step, angle_offset)
    if offset + p_in > n_proj:
        raise ValueError("offset + sinogram height exceeds num-projections")
    out = np.empty((rh, rw), dtype=np.float32)
    lib().orc_bp_texture_ref32(sino, w, p_in, s, c, offset, axis, roi_x, roi_y, rw, rh, out)
    return out


def bp_texture_f64(sino, axis_pos=-1.0, angle_step=-1.0, angle_offset=0.0, num_projections=0,
                   offset=0, roi_x=0, roi_y=0, roi_width=0, roi_height=0):
    """The same backprojection in FP64 (exact angles, double arithmetic)."""
    sino = np.ascontiguousarray(sino, dtype=np.float64)
    p_in, w = sino.shape
    n_proj, axis, rw, rh = _bp_defaults(w, p_in, num_projections, axis_pos, roi_width, roi_height)
    step = math.pi / n_proj if angle_step <= 0 else angle_step
    out = np.empty((rh, rw), dtype=np.float64)
    lib().orc_bp_texture_f64(sino, w, p_in, float(angle_offset), float(step), offset, axis,
                             roi_x, roi_y, rw, rh, out)
    return out


def bp_nearest_ref32(sino, axis_pos=-1.0, angle_step=-1.0, angle_offset=0.0, num_projections=0,
                     offset=0, roi_x=0, roi_y=0, roi_width=0, roi_height=0):
    """backproject mode=nearest with float32 semantics (backproject.cl:24-48)."""
    sino = np.ascontiguousarray(sino, dtype=np.float32)
    p_in, w = sino.shape
    n_proj, axis, rw, rh = _bp_defaults(w, p_in, num_projections, axis_pos, roi_width, roi_height)
    s, c, _ = bp_lut(n_proj, angle_step, angle_offset)
    out = np.empty((rh, rw), dtype=np.float32)
    lib().orc_bp_nearest_ref32(sino, w, p_in, s, c, offset, axis, roi_x, roi_y, rw, rh, out)
    return out


def stacked_bp_ref32(sinos, precision="single", axis_pos=-1.0, angle_step=-1.0, angle_offset=0.0,
                     num_projections=0, offset=0, roi_x=0, roi_y=0, roi_width=0, roi_height=0):
    """stacked-backproject (src/ufo-stacked-backproject-task.c:300-430, src/kernels/stacked-backproject.cl)
    with float32 semantics: sinograms [S][P][W] -> slices [S][roi_h][roi_w].  h = bx cos - by sin + axis
    (bx = ((x - axis) + roi_x) + 0.5; a work-item's rows y0 + 4q use h(y0) - 4q sin) sampled
      single/half: linearly at h - 0.5 with a zero border (half: texels rounded to float16 first),
                   sums over the 4 projection classes p mod 4 added as (c0 + c2) + (c1 + c3), times pi / P;
      int8: texel (v - min) * (255 / (max - min)) truncated and saturated to 8 bits (min, max of the
            whole stack), nearest texel floor(h) clamped to the edge, integer sum s, then
            ((s * ((max - min) / 255)) + min) * pi / P."""
    f32 = np.float32
    sinos = np.ascontiguousarray(sinos, dtype=np.float32)
    n_sl, p_in, w = sinos.shape
    n_proj, axis, rw, rh = _bp_defaults(w, p_in, num_projections, axis_pos, roi_width, roi_height)
    s, c, _ = bp_lut(n_proj, angle_step, angle_offset)
    s, c = s[offset:offset + p_in].astype(f32), c[offset:offset + p_in].astype(f32)
    axis = f32(axis)
    x = np.arange(rw, dtype=np.int64)
    y = np.arange(rh, dtype=np.int64)
    y0 = (y // 16) * 16 + (y % 16) % 4
    q4 = ((y % 16) // 4).astype(f32) * f32(4)
    bx = (x.astype(f32) - axis + f32(roi_x)) + f32(0.5)
    by = (y0.astype(f32) - axis + f32(roi_y)) + f32(0.5)
    out = np.empty((n_sl, rh, rw), dtype=np.float32)
    if precision == "int8":
        lo, hi = sinos.min(), sinos.max()
        scale = f32(255.0) / (hi - lo)
        with np.errstate(invalid="ignore"):
            qv = (sinos - lo) * scale
        qv = np.where(qv > 0, qv, f32(0))
        qv = np.minimum(np.floor(qv), f32(255)).astype(np.uint32)
        acc = np.zeros((n_sl, rh, rw), dtype=np.uint64)
        for p in range(p_in):
            h = ((bx[None, :] * c[p] - by[:, None] * s[p]) + axis) - q4[:, None] * s[p]
            i = np.clip(np.floor(h).astype(np.int64), 0, w - 1)
            acc += qv[:, p, :][:, i]
        rs = (hi - lo) / f32(255.0)
        r = ((acc.astype(np.float32) * rs) + lo) * f32(np.pi)
        out[:] = r / f32(p_in)
        return out
    data = sinos.astype(np.float16).astype(np.float32) if precision == "half" else sinos
    classes = np.zeros((4, n_sl, rh, rw), dtype=np.float32)
    for p in range(p_in):
        h = ((bx[None, :] * c[p] - by[:, None] * s[p]) + axis) - q4[:, None] * s[p]
        u = h - f32(0.5)
        i0 = np.floor(u)
        fr = (u - i0).astype(np.float32)
        i0 = i0.astype(np.int64)
        row = data[:, p, :]
        a = np.where((i0 >= 0) & (i0 < w), row[:, np.clip(i0, 0, w - 1)], f32(0))
        b = np.where((i0 + 1 >= 0) & (i0 + 1 < w), row[:, np.clip(i0 + 1, 0, w - 1)], f32(0))
        classes[p % 4] += a * (f32(1) - fr) + b * fr
    total = (classes[0] + classes[2]) + (classes[1] + classes[3])
    out[:] = total * f32(np.pi) / f32(p_in)
    return out


def cone_weight_ref32(proj, center_x, center_z, source_distance, detector_distance, axis_angle_x=0.0):
    """weight_projection (conebeam.cl:20-36) with the task's argument setup."""
    proj = np.ascontiguousarray(proj, dtype=np.float32)
    h, w = proj.shape
    out = np.empty_like(proj)
    lib().orc_cone_weight_ref32(proj, out, w, h, center_x, center_z, source_distance,
                                detector_distance, axis_angle_x)
    return out


# ------------------------------------------------------- flat field, average

def flat_field_correct_ref32(data, dark, flat, absorption_correct=False, fix_nan_and_inf=False,
                             sinogram_input=False, dark_scale=1.0, flat_scale=1.0):
    """flat_correct of src/kernels/ffc.cl:20-48 in float32 for items [n][H][W]: d = dark * dark_scale,
    f = flat * flat_scale, (data - d) / (f - d) or log((f - d) / (data - d)); NaN / Inf -> 0 on request.
    dark / flat: one item [H][W] (or a row [W] with sinogram_input) or one per data item."""
    f32 = np.float32
    data = np.asarray(data, dtype=f32)
    dark = np.asarray(dark, dtype=f32) * f32(dark_scale)
    flat = np.asarray(flat, dtype=f32) * f32(flat_scale)
    if sinogram_input:
        dark = dark.reshape(dark.shape[:-2] + (1, dark.shape[-1])) if dark.ndim >= 2 else dark[None, :]
        flat = flat.reshape(flat.shape[:-2] + (1, flat.shape[-1])) if flat.ndim >= 2 else flat[None, :]
    with np.errstate(divide="ignore", invalid="ignore"):
        if absorption_correct:
            r = np.log((flat - dark) / (data - dark)).astype(f32)
        else:
            r = ((data - dark) / (flat - dark)).astype(f32)
    if fix_nan_and_inf:
        r = np.where(np.isfinite(r), r, f32(0))
    return r


def average_ref32(items):
    """average (src/ufo-average-task.c:102-151): float32 sum in arrival order / count."""
    items = np.asarray(items, dtype=np.float32)
    s = np.zeros(items.shape[1:], dtype=np.float32)
    for it in items:
        s = s + it
    return s / np.float32(len(items))


# ------------------------------------------------------- transpose-projections

def transpose_projections(projs, number):
    """Sinograms [H][number][W] of projections [n][H][W]
    (src/ufo-transpose-projections-task.c:64-150): row h of projection
    p is row p of sinogram h; rows of projections that never arrive stay zero
    (the store is zero-initialised) and projections past `number` are not
    taken (process returns FALSE)."""
    projs = np.asarray(projs, dtype=np.float32)
    n, h, w = projs.shape
    out = np.zeros((h, number, w), dtype=np.float32)
    m = min(n, number)
    out[:, :m] = projs[:m].transpose(1, 0, 2)
    return out


# ---------------------------------------------------------- general backproject

_GEOM_KEYS = {
    "axis-angle-z": "tomo", "axis-angle-x": "axis_x", "axis-angle-y": "axis_y",
    "volume-angle-x": "vol_x", "volume-angle-y": "vol_y", "volume-angle-z": "vol_z",
    "detector-angle-x": "det_ax", "detector-angle-y": "det_ay", "detector-angle-z": "det_az",
    "center-position-x": "cx", "center-position-z": "cz",
    "source-position-x": "sx", "source-position-y": "sy", "source-position-z": "sz",
    "detector-position-x": "dx", "detector-position-y": "dy", "detector-position-z": "dz",
}


def _almost_zero(vals):
    return all(abs(v) < 1e-7 for v in vals)


def round_half(x):
    """Nearest IEEE binary16 value of x (ties to even), as a Python float."""
    return lib().orc_round_half(float(x))


def gbp(projs, num_projections=None, overall_angle=2 * math.pi, region=(0.0, 0.0, 0.0),
        x_region=(0.0, 0.0, 0.0), y_region=(0.0, 0.0, 0.0), addressing_mode="clamp", mode="f64",
        burst=8, result_type="float", **geometry):
    """general-backproject, parameter z, in FP64 (mode "f64") or float32 semantics ("ref32",
    with the per-burst running sum in result_type: half, float or double).

    projs: [P][H][W] float32.  Geometry keyword arguments use the property names with
    '-' replaced by '_' (e.g. source_position_y=[-2500.0]); one value broadcasts.
    Defaults follow ufo_ctgeometry_new (ufo-ctgeometry.c:82-109) and the task's
    setup (ufo-general-backproject-task.c:1411-1462, 1464-1587).
    """
    projs = np.ascontiguousarray(projs, dtype=np.float32)
    p, h, w = projs.shape
    num_projections = num_projections or p
    vals = {}
    for prop, key in _GEOM_KEYS.items():
        arr = geometry.get(prop.replace("-", "_"))
        if arr is None:
            arr = [-math.inf] if prop == "source-position-y" else [0.0]
        arr = [float(v) for v in np.atleast_1d(arr)]
        vals[prop] = arr
    if len(vals["axis-angle-z"]) != num_projections:
        vals["axis-angle-z"] = [i / num_projections * overall_angle for i in range(num_projections)]

    def at(prop, i):
        a = vals[prop]
        return a[0] if len(a) == 1 else a[i]

    geoms = (OrcGbpGeom * p)()
    for i in range(p):
        for prop, key in _GEOM_KEYS.items():
            setattr(geoms[i], key, at(prop, i))
    fl = OrcGbpFlags()
    fl.with_axis = int(not (_almost_zero(vals["axis-angle-x"]) and _almost_zero(vals["axis-angle-y"])))
    fl.with_volume = int(not (_almost_zero(vals["volume-angle-x"]) and _almost_zero(vals["volume-angle-y"])
                              and _almost_zero(vals["volume-angle-z"])))
    fl.shifted_detector = int(not (_almost_zero(vals["detector-position-x"])
                                   and _almost_zero(vals["detector-position-z"])))
    fl.shifted_source = int(not (_almost_zero(vals["source-position-x"])
                                 and _almost_zero(vals["source-position-z"])))
    fl.perpendicular_detector = int(_almost_zero(vals["detector-angle-x"]) and _almost_zero(vals["detector-angle-y"])
                                    and _almost_zero(vals["detector-angle-z"]))
    fl.parallel_beam = int(all(math.isinf(at("source-position-y", i)) for i in range(num_projections)))

    def region_params(reg, length):
        if reg[2]:
            n = int(math.ceil((reg[1] - reg[0]) / reg[2]))
            return n, float(reg[0]), float(reg[2])
        return length, float(-(np.float32(length) / np.float32(2))), 1.0

    nx, x0, xs = region_params(x_region, w)
    ny, y0, ys = region_params(y_region, w)
    if abs(region[2]) < 1e-7:
        zstart, zstop, zstep = 0.0, 1.0, 1.0
    else:
        zstart, zstop, zstep = region
    nz = int(math.ceil((zstop - zstart) / zstep))
    out = np.empty((nz, ny, nx), dtype=np.float64)
    lib().orc_gbp(projs, w, h, p, geoms, ctypes.byref(fl), ADDR[addressing_mode], nx, ny, nz,
                  x0, xs, y0, ys, float(zstart), float(zstep), float(overall_angle),
                  0 if mode == "f64" else 1, int(burst), RESULT_TYPES[result_type], out)
    return out


class gbp_planted:
    """Context manager that switches on planted mistakes of orc_gbp's FP64 mode (bit mask of
    orc_gbp_plant in ufo_oracle.c), for tests that show the geometry checks catch them."""

    def __init__(self, mask):
        self.mask = int(mask)

    def __enter__(self):
        self._var = ctypes.c_int.in_dll(lib(), "orc_gbp_plant")
        self._var.value = self.mask
        return self

    def __exit__(self, *exc):
        self._var.value = 0
        return False


LAMINO_PARAMETERS = ["z", "x-center", "lamino-angle", "roll-angle"]


def lamino_shape(x_region, y_region, region):
    """REGION_SIZE for the integer x/y regions and ceil((stop - start) / step) in float32 for the
    swept region (src/ufo-lamino-backproject-task.c:41-43, 247-262)."""
    size = lambda r: 0 if r[2] == 0 else int((r[1] - r[0] - 1) / r[2]) + 1  # C division truncates
    r = np.asarray(region, np.float32)
    return size(x_region), size(y_region), int(math.ceil(float((r[1] - r[0]) / r[2])))


def lamino(projs, num_projections=None, overall_angle=math.pi, tomo_angle=-3.4028234663852886e38,
           lamino_angle=0.0, roll_angle=0.0, center=(0.0, 0.0), projection_offset=(0, 0),
           x_region=(0, 0, 0), y_region=(0, 0, 0), region=(0.0, 0.0, 0.0), z=0.0, parameter="z",
           addressing_mode="clamp", mode="f64", burst=16):
    """lamino-backproject (src/ufo-lamino-backproject-task.c:285-480 and the kernel templates),
    FP64 (mode "f64") or float32 with bursts of 16 projections ("ref32").  projs: [P][H][W].
    Returns [nz][ny][nx] float64."""
    projs = np.ascontiguousarray(projs, dtype=np.float32)
    p, h, w = projs.shape
    num_projections = num_projections or p
    nx, ny, nz = lamino_shape(x_region, y_region, region)
    out = np.empty((nz, ny, nx), dtype=np.float64)
    lib().orc_lamino(projs, w, h, p, num_projections, overall_angle, tomo_angle, lamino_angle, roll_angle,
                     center[0], center[1], int(projection_offset[0]), int(projection_offset[1]),
                     np.asarray(x_region, np.int32), np.asarray(y_region, np.int32),
                     np.asarray(region, np.float32), z, LAMINO_PARAMETERS.index(parameter),
                     ADDR[addressing_mode], nx, ny, nz, 0 if mode == "f64" else 1, burst, out)
    return out


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
