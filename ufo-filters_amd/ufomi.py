"""Python binding of libufomi.so (include/ufomi.h) over ctypes.

The library is the product: every call below goes to a HIP kernel for gfx950
through the C ABI.  There is no CPU fallback; a missing or unloadable library
raises.  Device buffers are torch tensors on a ROCm device (torch is used for
device memory and streams only), passed by data_ptr().
"""
import ctypes
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libufomi.so")

OK = 0
STATUS = {
    -1: "ERR_ARG", -2: "ERR_HIP", -3: "ERR_NOT_POW2", -4: "ERR_UNSUPPORTED", -5: "ERR_PROJECTIONS",
    -6: "ERR_SIZE", -7: "ERR_NO_DEVICE", -8: "ERR_GRAY_MAP", -9: "ERR_MEMORY", -10: "ERR_NUM_PROJECTIONS",
    -11: "ERR_LAYOUT", -12: "ERR_INCOMPLETE", -13: "ERR_IO",
}
F32, F16, BF16 = 0, 1, 2
FILTERS = {"ramp": 0, "ramp-fromreal": 1, "butterworth": 2, "faris-byer": 3, "hamming": 4, "bh3": 5,
           "shepp-logan": 6}
BP_NEAREST, BP_TEXTURE = 0, 1
STORE_TYPES = {"half": 0, "float": 1, "double": 2, "uchar": 3, "ushort": 4, "uint": 5}
ADDRESSING = {"none": 0, "clamp_to_edge": 1, "clamp": 2, "repeat": 3, "mirrored_repeat": 4}

# Every symbol include/ufomi.h declares.
EXPORTS = [
    "ufomi_create", "ufomi_destroy", "ufomi_set_stream", "ufomi_synchronize", "ufomi_last_error",
    "ufomi_version", "ufomi_filter_coefficients", "ufomi_fft_rows", "ufomi_filter_apply", "ufomi_ifft_rows",
    "ufomi_fft_nd", "ufomi_ifft_nd",
    "ufomi_ramp_filter", "ufomi_bp_params_init", "ufomi_bp_parallel",
    "ufomi_cone_weight", "ufomi_transpose_projections", "ufomi_bp_stacked", "ufomi_flat_field_correct", "ufomi_average", "ufomi_gbp_sweep", "ufomi_lamino_params_init", "ufomi_lamino_volume_shape", "ufomi_lamino_bp", "ufomi_gbp_params_init", "ufomi_gbp_geometry_init", "ufomi_gbp_volume_shape",
    "ufomi_gbp", "ufomi_phantom_ellipses", "ufomi_phantom_ellipsoids", "ufomi_slice_ellipsoids",
    "ufomi_sinogram_ellipses", "ufomi_add_noise", "ufomi_bench_lds_read", "ufomi_bench_copy",
    "ufomi_gbp_projection_rows", "ufomi_project_ellipsoids", "ufomi_device_alloc", "ufomi_device_free",
    "ufomi_memcpy", "ufomi_device_memory", "ufomi_fbp_batch_create", "ufomi_fbp_batch_configure",
    "ufomi_fbp_batch_autotune", "ufomi_fbp_batch_run", "ufomi_fbp_batch_launch", "ufomi_fbp_batch_get_info",
    "ufomi_fbp_batch_destroy", "ufomi_phantom_texture", "ufomi_sinogram_phantom", "ufomi_phantom_image",
    "ufomi_unet_create", "ufomi_unet_destroy", "ufomi_unet_num_params", "ufomi_unet_num_tensors", "ufomi_unet_tensor",
    "ufomi_unet_get_params", "ufomi_unet_set_params", "ufomi_unet_forward", "ufomi_unet_gradients",
    "ufomi_unet_train_step", "ufomi_nn_conv3x3", "ufomi_bench_mfma", "ufomi_bp_configure",
    "ufomi_bp_parallel_noalloc",
]


class UfomiError(RuntimeError):
    def __init__(self, code, message):
        super().__init__("%s (%d): %s" % (STATUS.get(code, "ERR"), code, message))
        self.code = code


class BpParams(ctypes.Structure):
    _fields_ = [("num_projections", ctypes.c_uint), ("offset", ctypes.c_uint), ("axis_pos", ctypes.c_double),
                ("angle_step", ctypes.c_double), ("angle_offset", ctypes.c_double), ("mode", ctypes.c_int),
                ("roi_x", ctypes.c_uint), ("roi_y", ctypes.c_uint), ("roi_width", ctypes.c_uint),
                ("roi_height", ctypes.c_uint)]


class FbpRequest(ctypes.Structure):
    _fields_ = [("sinogram", ctypes.c_void_p), ("slice", ctypes.c_void_p), ("width", ctypes.c_int),
                ("num_angles", ctypes.c_int), ("fft_size", ctypes.c_int), ("params", BpParams)]


class FbpBatchInfo(ctypes.Structure):
    _fields_ = [("voxel_updates", ctypes.c_double), ("tile", ctypes.c_int), ("angle_chunk", ctypes.c_int),
                ("workgroups", ctypes.c_int), ("items", ctypes.c_int), ("filter_launches", ctypes.c_int),
                ("filtered_bytes", ctypes.c_size_t)]


_GEOM_NAMES = ["axis_angle_x", "axis_angle_y", "axis_angle_z", "volume_angle_x", "volume_angle_y",
               "volume_angle_z", "detector_angle_x", "detector_angle_y", "detector_angle_z",
               "center_position_x", "center_position_z", "source_position_x", "source_position_y",
               "source_position_z", "detector_position_x", "detector_position_y", "detector_position_z"]


class GbpGeometry(ctypes.Structure):
    _fields_ = [(n, ctypes.POINTER(ctypes.c_double)) for n in _GEOM_NAMES] + \
               [("n_" + n, ctypes.c_int) for n in _GEOM_NAMES]


class GbpParams(ctypes.Structure):
    _fields_ = [("num_projections", ctypes.c_uint), ("overall_angle", ctypes.c_double), ("z", ctypes.c_double),
                ("region", ctypes.c_double * 3), ("x_region", ctypes.c_double * 3),
                ("y_region", ctypes.c_double * 3), ("store_type", ctypes.c_int), ("addressing_mode", ctypes.c_int),
                ("gray_map_min", ctypes.c_double), ("gray_map_max", ctypes.c_double),
                ("compute_type", ctypes.c_int), ("result_type", ctypes.c_int), ("burst", ctypes.c_uint)]


COMPUTE_TYPES = {"float": 0, "double": 1}
RESULT_TYPES = {"half": 0, "float": 1, "double": 2}


class LaminoParams(ctypes.Structure):
    _fields_ = [("num_projections", ctypes.c_uint), ("overall_angle", ctypes.c_double),
                ("tomo_angle", ctypes.c_double), ("lamino_angle", ctypes.c_double), ("roll_angle", ctypes.c_double),
                ("center", ctypes.c_double * 2), ("projection_offset", ctypes.c_int * 2),
                ("x_region", ctypes.c_int * 3), ("y_region", ctypes.c_int * 3), ("region", ctypes.c_double * 3),
                ("z", ctypes.c_double), ("parameter", ctypes.c_int), ("addressing_mode", ctypes.c_int)]


LAMINO_PARAMETERS = ["z", "x-center", "lamino-angle", "roll-angle"]

_lib = None


def load():
    """Load libufomi.so; raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libufomi.so is not built (%s); run make -C ufo-filters_amd libufomi.so" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    vp, ci, cu, cf, cd, csz = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint, ctypes.c_float, ctypes.c_double, ctypes.c_size_t
    dp = ctypes.POINTER(ctypes.c_double)
    sig = {
        "ufomi_create": [ci, ctypes.POINTER(vp)], "ufomi_destroy": [vp], "ufomi_set_stream": [vp, vp],
        "ufomi_synchronize": [vp], "ufomi_last_error": [], "ufomi_version": [],
        "ufomi_filter_coefficients": [ci, cu, cf, cf, cf, cf, cf, vp],
        "ufomi_fft_rows": [vp, vp, ci, ci, ci, vp], "ufomi_filter_apply": [vp, vp, vp, vp, ci, ci],
        "ufomi_fft_nd": [vp, vp, ci, vp, ci, vp, vp], "ufomi_ifft_nd": [vp, vp, vp, ci, ci, ci, vp],
        "ufomi_ifft_rows": [vp, vp, ci, ci, ci, vp], "ufomi_ramp_filter": [vp, vp, ci, ci, ci, vp, ci, vp, ci],
        "ufomi_bp_params_init": [ctypes.POINTER(BpParams)],
        "ufomi_bp_parallel": [vp, vp, ci, ci, ci, ci, csz, ctypes.POINTER(BpParams), vp, csz],
        "ufomi_bp_parallel_noalloc": [vp, vp, ci, ci, ci, ci, csz, ctypes.POINTER(BpParams), vp, csz],
        "ufomi_bp_stacked": [vp, vp, ci, ci, ci, ctypes.POINTER(BpParams), ci, vp],
        "ufomi_flat_field_correct": [vp, vp, ci, ci, ci, vp, ci, vp, ci, ci, ci, ci, ctypes.c_float, ctypes.c_float,
                                     vp],
        "ufomi_average": [vp, vp, csz, ci, ci, vp],
        "ufomi_cone_weight": [vp, vp, vp, ci, ci, ci, dp, dp, dp, dp, dp, ci],
        "ufomi_transpose_projections": [vp, vp, ci, ci, ci, ci, ci, vp],
        "ufomi_gbp_params_init": [ctypes.POINTER(GbpParams)], "ufomi_gbp_geometry_init": [ctypes.POINTER(GbpGeometry)],
        "ufomi_gbp_volume_shape": [ctypes.POINTER(GbpParams), ci, ctypes.POINTER(ci), ctypes.POINTER(ci),
                                   ctypes.POINTER(ci)],
        "ufomi_gbp": [vp, vp, ci, ci, ci, ctypes.POINTER(GbpGeometry), ctypes.POINTER(GbpParams), vp],
        "ufomi_lamino_params_init": [ctypes.POINTER(LaminoParams)],
        "ufomi_lamino_volume_shape": [ctypes.POINTER(LaminoParams), ctypes.POINTER(ci), ctypes.POINTER(ci),
                                      ctypes.POINTER(ci)],
        "ufomi_lamino_bp": [vp, vp, ci, ci, ci, ctypes.POINTER(LaminoParams), vp],
        "ufomi_gbp_sweep": [vp, vp, ci, ci, ci, ctypes.POINTER(GbpGeometry), ctypes.POINTER(GbpParams), ci, vp],
        "ufomi_phantom_ellipses": [ctypes.c_uint64, ci, cd, dp],
        "ufomi_phantom_ellipsoids": [ctypes.c_uint64, ci, cd, dp],
        "ufomi_slice_ellipsoids": [dp, ci, cd, dp],
        "ufomi_sinogram_ellipses": [vp, dp, ci, ci, ci, ci, cd, cd, vp],
        "ufomi_add_noise": [vp, vp, csz, cd, ctypes.c_uint64, ctypes.POINTER(cf)],
        "ufomi_gbp_projection_rows": [ctypes.POINTER(GbpGeometry), ctypes.POINTER(GbpParams), dp,
                                      ctypes.POINTER(ci)],
        "ufomi_project_ellipsoids": [vp, dp, ci, ci, ci, ci, dp, vp],
        "ufomi_bench_lds_read": [vp, ci, ci, vp, ctypes.POINTER(cd)],
        "ufomi_bench_copy": [vp, vp, vp, csz],
        "ufomi_bench_mfma": [vp, ci, ci, vp, ctypes.POINTER(cd)],
        "ufomi_bp_configure": [vp, ci, ci, ci],
        "ufomi_fbp_batch_create": [vp, ctypes.POINTER(FbpRequest), ci, ci, cf, ci, ctypes.POINTER(vp)],
        "ufomi_fbp_batch_configure": [vp, vp, ci, ci],
        "ufomi_fbp_batch_autotune": [vp, vp, ci, ci, ctypes.POINTER(ci), ctypes.POINTER(ci), ctypes.POINTER(cf)],
        "ufomi_fbp_batch_run": [vp, vp], "ufomi_fbp_batch_launch": [vp, vp, ci],
        "ufomi_fbp_batch_get_info": [vp, ctypes.POINTER(FbpBatchInfo)], "ufomi_fbp_batch_destroy": [vp],
        "ufomi_phantom_texture": [ctypes.c_uint64, ci, cd, cd, cd, dp],
        "ufomi_sinogram_phantom": [vp, dp, ci, dp, ci, cd, ci, ci, ci, cd, cd, vp],
        "ufomi_phantom_image": [vp, dp, ci, dp, ci, cd, ci, ci, cd, vp],
        "ufomi_unet_create": [vp, ci, ci, ci, ci, ctypes.c_uint64, ctypes.POINTER(vp)], "ufomi_unet_destroy": [vp],
        "ufomi_unet_num_params": [vp, ctypes.POINTER(csz), ctypes.POINTER(csz)], "ufomi_unet_num_tensors": [vp],
        "ufomi_unet_tensor": [vp, ci, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(csz), ctypes.POINTER(ci)],
        "ufomi_unet_get_params": [vp, vp, vp], "ufomi_unet_set_params": [vp, vp, vp],
        "ufomi_unet_forward": [vp, vp, vp, ci, vp],
        "ufomi_unet_gradients": [vp, vp, vp, vp, ci, ctypes.POINTER(cf), vp],
        "ufomi_unet_train_step": [vp, vp, vp, vp, ci, cf, ctypes.POINTER(cf)],
        "ufomi_nn_conv3x3": [vp, vp, ci, ci, ci, ci, vp, vp, ci, ci, vp],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ci
    L.ufomi_last_error.restype = ctypes.c_char_p
    L.ufomi_version.restype = ctypes.c_char_p
    L.ufomi_bp_params_init.restype = None
    L.ufomi_gbp_params_init.restype = None
    L.ufomi_lamino_params_init.restype = None
    L.ufomi_gbp_geometry_init.restype = None
    _lib = L
    return L


def check(rc):
    if rc != OK:
        raise UfomiError(rc, load().ufomi_last_error().decode())
    return rc


class _dptr:
    """A float64 array argument: the converted copy lives as long as this object, which ctypes
    keeps for the duration of the call it is passed to."""

    def __init__(self, a):
        self.a = np.ascontiguousarray(a, dtype=np.float64)
        self._as_parameter_ = self.a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ptr(t):
    """Device pointer of a torch tensor; refuses host tensors and non-contiguous views, whose
    pointer would make the kernels read host memory or the wrong elements."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("expected a device tensor, got one on %s" % t.device)
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor (got strides %s for shape %s)" % (t.stride(), tuple(t.shape)))
    return ctypes.c_void_p(t.data_ptr())


# ------------------------------------------------------------------ host-only

def filter_coefficients(width, filter="ramp-fromreal", scale=1.0, cutoff=0.5, order=4.0, tau=0.1, theta=1.0):
    """float32[width] coefficients of the filter task (width = 2N)."""
    out = np.empty(width, dtype=np.float32)
    check(load().ufomi_filter_coefficients(FILTERS[filter], width, scale, cutoff, order, tau, theta,
                                           out.ctypes.data_as(ctypes.c_void_p)))
    return out


def phantom_ellipses(seed, n, radius_px):
    out = np.empty((n, 6), dtype=np.float64)
    check(load().ufomi_phantom_ellipses(seed, n, radius_px, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return out


def phantom_texture(seed, n, period_min_px, period_max_px, rms):
    """[n][4] plane waves (kx, ky, phase, amplitude) of a band-limited texture."""
    out = np.empty((n, 4), dtype=np.float64)
    check(load().ufomi_phantom_texture(seed, n, period_min_px, period_max_px, rms,
                                       out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return out


def phantom_ellipsoids(seed, n, radius_px):
    out = np.empty((n, 8), dtype=np.float64)
    check(load().ufomi_phantom_ellipsoids(seed, n, radius_px, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return out


def slice_ellipsoids(ellipsoids, z):
    e = np.ascontiguousarray(ellipsoids, dtype=np.float64)
    out = np.empty((len(e), 6), dtype=np.float64)
    check(load().ufomi_slice_ellipsoids(_dptr(e), len(e), z, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return out


def gbp_volume_shape(params, width):
    nx, ny, nz = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    check(load().ufomi_gbp_volume_shape(ctypes.byref(params), width, ctypes.byref(nx), ctypes.byref(ny),
                                        ctypes.byref(nz)))
    return nz.value, ny.value, nx.value


def gbp_projection_rows(geometry, params):
    """[P][4][4] FP64 rows (U, V, D, C) of the general backprojector's maps, and the parallel flag."""
    rows = np.empty((params.num_projections, 4, 4), dtype=np.float64)
    par = ctypes.c_int()
    check(load().ufomi_gbp_projection_rows(ctypes.byref(geometry), ctypes.byref(params),
                                           rows.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(par)))
    return rows, bool(par.value)


def gbp_params(num_projections, overall_angle=2 * math.pi, region=(0, 0, 0), x_region=(0, 0, 0),
               y_region=(0, 0, 0), store_type="float", addressing_mode="clamp", gray_map_min=0.0,
               gray_map_max=0.0, compute_type="float", result_type="float", burst=0):
    p = GbpParams()
    load().ufomi_gbp_params_init(ctypes.byref(p))
    p.num_projections = num_projections
    p.overall_angle = overall_angle
    p.region[:] = [float(v) for v in region]
    p.x_region[:] = [float(v) for v in x_region]
    p.y_region[:] = [float(v) for v in y_region]
    p.store_type = STORE_TYPES[store_type]
    p.addressing_mode = ADDRESSING[addressing_mode]
    p.gray_map_min = gray_map_min
    p.gray_map_max = gray_map_max
    p.compute_type = COMPUTE_TYPES[compute_type]
    p.result_type = RESULT_TYPES[result_type]
    p.burst = burst
    return p


def lamino_params(num_projections, **kw):
    """LaminoParams from keyword arguments named like the task's properties ('-' -> '_')."""
    p = LaminoParams()
    load().ufomi_lamino_params_init(ctypes.byref(p))
    p.num_projections = num_projections
    for k, v in kw.items():
        if k == "parameter":
            v = LAMINO_PARAMETERS.index(v) if isinstance(v, str) else v
        elif k == "addressing_mode":
            v = ADDRESSING[v] if isinstance(v, str) else v
        if k in ("center", "projection_offset", "x_region", "y_region", "region"):
            getattr(p, k)[:] = list(v)
        else:
            setattr(p, k, v)
    return p


def lamino_shape(params):
    nx, ny, nz = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    check(load().ufomi_lamino_volume_shape(ctypes.byref(params), ctypes.byref(nx), ctypes.byref(ny),
                                           ctypes.byref(nz)))
    return nz.value, ny.value, nx.value


def gbp_geometry(**arrays):
    """Geometry struct from keyword arrays (names with '_'); keeps the numpy arrays alive."""
    g = GbpGeometry()
    load().ufomi_gbp_geometry_init(ctypes.byref(g))
    keep = []
    for name, val in arrays.items():
        if name not in _GEOM_NAMES:
            raise KeyError(name)
        a = np.ascontiguousarray(np.atleast_1d(val), dtype=np.float64)
        keep.append(a)
        setattr(g, name, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        setattr(g, "n_" + name, len(a))
    g._keep = keep
    return g


# ---------------------------------------------------------------- device API

class Context:
    """A ufomi context on one device, launching on torch's current stream."""

    def __init__(self, device=0, stream=None):
        self.L = load()
        self.h = ctypes.c_void_p()
        check(self.L.ufomi_create(device, ctypes.byref(self.h)))
        self.device = device
        if stream is not None:
            self.set_stream(stream)

    def set_stream(self, stream):
        """stream: a torch.cuda.Stream (or raw hipStream_t int)."""
        raw = getattr(stream, "cuda_stream", stream)
        check(self.L.ufomi_set_stream(self.h, ctypes.c_void_p(raw)))

    def use_current_stream(self):
        import torch
        self.set_stream(torch.cuda.current_stream(self.device))

    def synchronize(self):
        check(self.L.ufomi_synchronize(self.h))

    def close(self):
        if self.h:
            self.L.ufomi_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- ramp filter chain
    def fft_rows(self, d_in, W, P, N, d_out):
        check(self.L.ufomi_fft_rows(self.h, _ptr(d_in), W, P, N, _ptr(d_out)))

    def fft_nd(self, d_in, n, dimensions, m, d_out, in_complex=False):
        """Forward DFT over the first `dimensions` axes of [n2][n1][n0] zero-padded to m (x first)."""
        ni, mi = (ctypes.c_int * 3)(*n), (ctypes.c_int * 3)(*m)
        check(self.L.ufomi_fft_nd(self.h, _ptr(d_in), int(in_complex), ni, dimensions, mi, _ptr(d_out)))

    def ifft_nd(self, d_in, m, dimensions, d_out, crop_width=0, crop_height=0):
        mi = (ctypes.c_int * 3)(*m)
        check(self.L.ufomi_ifft_nd(self.h, _ptr(d_in), mi, dimensions, crop_width, crop_height, _ptr(d_out)))

    def filter_apply(self, d_in, d_out, d_coef, width, P):
        check(self.L.ufomi_filter_apply(self.h, _ptr(d_in), _ptr(d_out), _ptr(d_coef), width, P))

    def ifft_rows(self, d_in, N, P, crop, d_out):
        check(self.L.ufomi_ifft_rows(self.h, _ptr(d_in), N, P, crop, _ptr(d_out)))

    def ramp_filter(self, d_in, W, P, N, d_coef, crop, d_out, dtype=F32):
        check(self.L.ufomi_ramp_filter(self.h, _ptr(d_in), W, P, N, _ptr(d_coef), crop, _ptr(d_out), dtype))

    # -- backprojection
    def bp_configure(self, slices_per_tile=0, angle_chunk=0, half_staging=0):
        """Tile configuration of bp_parallel (0 = automatic)."""
        check(self.L.ufomi_bp_configure(self.h, slices_per_tile, angle_chunk, int(half_staging)))

    def bp_parallel(self, d_sino, W, P_in, d_out, n_slices=1, dtype=F32, sino_pitch=0, out_pitch=0, **props):
        p = bp_params(**props)
        check(self.L.ufomi_bp_parallel(self.h, _ptr(d_sino), dtype, W, P_in, n_slices, sino_pitch,
                                       ctypes.byref(p), _ptr(d_out), out_pitch))

    def bp_parallel_noalloc(self, d_sino, W, P_in, d_out, n_slices=1, dtype=F32, sino_pitch=0, out_pitch=0,
                            **props):
        """bp_parallel on the library's preallocated device scratch (no allocation)."""
        p = bp_params(**props)
        check(self.L.ufomi_bp_parallel_noalloc(self.h, _ptr(d_sino), dtype, W, P_in, n_slices, sino_pitch,
                                               ctypes.byref(p), _ptr(d_out), out_pitch))

    def flat_field_correct(self, d_data, W, H, n, d_dark, n_dark, d_flat, n_flat, d_out, sinogram_input=False,
                           absorption_correct=False, fix_nan_and_inf=False, dark_scale=1.0, flat_scale=1.0):
        check(self.L.ufomi_flat_field_correct(self.h, _ptr(d_data), W, H, n, _ptr(d_dark), n_dark, _ptr(d_flat),
                                              n_flat, int(sinogram_input), int(absorption_correct),
                                              int(fix_nan_and_inf), dark_scale, flat_scale, _ptr(d_out)))

    def average(self, d_items, elems, n, d_out, copies=1):
        check(self.L.ufomi_average(self.h, _ptr(d_items), elems, n, copies, _ptr(d_out)))

    def bp_stacked(self, d_sino, W, P_in, n_slices, params, precision, d_out):
        """stacked-backproject of n_slices sinograms; precision 0 int8, 1 half, 2 single."""
        check(self.L.ufomi_bp_stacked(self.h, _ptr(d_sino), W, P_in, n_slices, ctypes.byref(params), precision,
                                      _ptr(d_out)))

    def cone_weight(self, d_in, d_out, W, H, n_proj, center_x, center_z, source_distance, detector_distance,
                    axis_angle_x=0.0):
        arrs = [np.ascontiguousarray(np.atleast_1d(a), dtype=np.float64)
                for a in (center_x, center_z, source_distance, detector_distance, axis_angle_x)]
        n = max(len(a) for a in arrs)
        arrs = [np.full(n, a[0]) if len(a) == 1 else a for a in arrs]
        check(self.L.ufomi_cone_weight(self.h, _ptr(d_in), _ptr(d_out), W, H, n_proj,
                                       *[a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) for a in arrs], n))

    def transpose_projections(self, d_proj, W, H, n, d_sino, first=0, number=None):
        """Projections [n][H][W] (None: zero rows) -> rows first.. of sinograms [H][number][W]."""
        check(self.L.ufomi_transpose_projections(self.h, _ptr(d_proj) if d_proj is not None else None, W, H, n,
                                                 first, number if number is not None else n, _ptr(d_sino)))

    def gbp(self, d_proj, W, H, P, geometry, params, d_vol):
        check(self.L.ufomi_gbp(self.h, _ptr(d_proj), W, H, P, ctypes.byref(geometry), ctypes.byref(params),
                               _ptr(d_vol)))

    def lamino_bp(self, d_proj, W, H, P, params, d_vol):
        """lamino-backproject: projections [P][H][W] -> volume [nz][ny][nx]."""
        check(self.L.ufomi_lamino_bp(self.h, _ptr(d_proj), W, H, P, ctypes.byref(params), _ptr(d_vol)))

    def gbp_sweep(self, d_proj, W, H, P, geometry, params, parameter, d_vol):
        """general-backproject with `parameter` (index of GBP_PARAMETERS) swept over params.region."""
        check(self.L.ufomi_gbp_sweep(self.h, _ptr(d_proj), W, H, P, ctypes.byref(geometry), ctypes.byref(params),
                                     parameter, _ptr(d_vol)))

    # -- synthetic data
    def sinogram_phantom(self, ellipses, waves, radius, W, P, axis_pos, angle_step, d_out):
        """ellipses [S][n][6] (or [n][6]), waves [S][m][4] (or [m][4], or None)."""
        e = np.ascontiguousarray(ellipses, dtype=np.float64)
        n_sl = 1 if e.ndim == 2 else e.shape[0]
        n_ell = e.shape[-2]
        w = None if waves is None else np.ascontiguousarray(waves, dtype=np.float64)
        n_w = 0 if w is None else w.shape[-2]
        check(self.L.ufomi_sinogram_phantom(self.h, _dptr(e), n_ell, None if w is None else _dptr(w), n_w, radius,
                                            n_sl, W, P, axis_pos, angle_step, _ptr(d_out)))

    def phantom_image(self, ellipses, waves, radius, N, axis_pos, d_out):
        e = np.ascontiguousarray(ellipses, dtype=np.float64)
        n_sl = 1 if e.ndim == 2 else e.shape[0]
        n_ell = e.shape[-2]
        w = None if waves is None else np.ascontiguousarray(waves, dtype=np.float64)
        n_w = 0 if w is None else w.shape[-2]
        check(self.L.ufomi_phantom_image(self.h, _dptr(e), n_ell, None if w is None else _dptr(w), n_w, radius,
                                         n_sl, N, axis_pos, _ptr(d_out)))

    def nn_conv3x3(self, d_in, batch, height, width, cin, d_w, d_bias, relu, cout, d_out):
        check(self.L.ufomi_nn_conv3x3(self.h, _ptr(d_in), batch, height, width, cin, _ptr(d_w), _ptr(d_bias),
                                      int(relu), cout, _ptr(d_out)))

    def sinogram_ellipses(self, ellipses, W, P, axis_pos, angle_step, d_out):
        e = np.ascontiguousarray(ellipses, dtype=np.float64)
        if e.ndim == 2:
            e = e[None]
        check(self.L.ufomi_sinogram_ellipses(self.h, _dptr(e), e.shape[1], e.shape[0], W, P, axis_pos,
                                             angle_step, _ptr(d_out)))

    def project_ellipsoids(self, ellipsoids, W, H, rows, d_out):
        e = np.ascontiguousarray(ellipsoids, dtype=np.float64)
        r = np.ascontiguousarray(rows, dtype=np.float64)
        check(self.L.ufomi_project_ellipsoids(self.h, _dptr(e), len(e), W, H, r.shape[0], _dptr(r), _ptr(d_out)))

    def add_noise(self, d, rel_sigma, seed):
        m = ctypes.c_float()
        check(self.L.ufomi_add_noise(self.h, _ptr(d), d.numel(), rel_sigma, seed, ctypes.byref(m)))
        return m.value

    # -- roofline microbenchmarks
    def bench_lds_read(self, blocks, iters, d_sink):
        b = ctypes.c_double()
        check(self.L.ufomi_bench_lds_read(self.h, blocks, iters, _ptr(d_sink), ctypes.byref(b)))
        return b.value

    def bench_copy(self, d_in, d_out):
        check(self.L.ufomi_bench_copy(self.h, _ptr(d_in), _ptr(d_out), d_in.numel()))

    def bench_mfma(self, blocks, iters, d_sink):
        f = ctypes.c_double()
        check(self.L.ufomi_bench_mfma(self.h, blocks, iters, _ptr(d_sink), ctypes.byref(f)))
        return f.value


class FbpBatch:
    """Uneven single-slice FBP requests as one HIP graph (ufomi_fbp_batch_*).

    requests: dicts with torch device tensors `sinogram` ([P][W] float32) and
    `slice` ([roi_height][roi_width] float32), optional `fft_size` and any
    `backproject` property of bp_params().  The tensors must outlive the batch.
    """

    def __init__(self, ctx, requests, filter="ramp-fromreal", scale=1.0, storage=F32):
        self.ctx = ctx
        self.L = ctx.L
        self.h = ctypes.c_void_p()
        arr = (FbpRequest * len(requests))()
        self._keep = []
        for i, r in enumerate(requests):
            sino, out = r["sinogram"], r["slice"]
            p, w = sino.shape
            props = {k: v for k, v in r.items() if k not in ("sinogram", "slice", "fft_size")}
            arr[i].sinogram = _ptr(sino).value
            arr[i].slice = _ptr(out).value
            arr[i].width = w
            arr[i].num_angles = p
            arr[i].fft_size = r.get("fft_size", 0)
            arr[i].params = bp_params(**props)
            self._keep += [sino, out]
        check(self.L.ufomi_fbp_batch_create(ctx.h, arr, len(requests), FILTERS[filter], scale, storage,
                                            ctypes.byref(self.h)))

    def configure(self, tile, angle_chunk):
        check(self.L.ufomi_fbp_batch_configure(self.ctx.h, self.h, tile, angle_chunk))

    def autotune(self, reps=2, force=False):
        t, a, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_float()
        check(self.L.ufomi_fbp_batch_autotune(self.ctx.h, self.h, reps, int(force), ctypes.byref(t),
                                              ctypes.byref(a), ctypes.byref(ms)))
        return t.value, a.value, ms.value

    def run(self):
        check(self.L.ufomi_fbp_batch_run(self.ctx.h, self.h))

    def launch(self, stages=3):
        check(self.L.ufomi_fbp_batch_launch(self.ctx.h, self.h, stages))

    def info(self):
        i = FbpBatchInfo()
        check(self.L.ufomi_fbp_batch_get_info(self.h, ctypes.byref(i)))
        return {k: getattr(i, k) for k, _ in FbpBatchInfo._fields_}

    def close(self):
        if self.h:
            self.L.ufomi_fbp_batch_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Unet:
    """The learned sparse-view stage (ufomi_unet_*): images are float32 torch tensors [B][H][W]."""

    def __init__(self, ctx, max_batch, height, width, training=False, seed=0):
        self.ctx = ctx
        self.L = ctx.L
        self.h = ctypes.c_void_p()
        check(self.L.ufomi_unet_create(ctx.h, max_batch, height, width, int(training), seed, ctypes.byref(self.h)))
        c, p = ctypes.c_size_t(), ctypes.c_size_t()
        check(self.L.ufomi_unet_num_params(self.h, ctypes.byref(c), ctypes.byref(p)))
        self.count, self.padded = c.value, p.value
        self.tensors = []
        for i in range(self.L.ufomi_unet_num_tensors(self.h)):
            name, off, dims = ctypes.c_char_p(), ctypes.c_size_t(), (ctypes.c_int * 4)()
            check(self.L.ufomi_unet_tensor(self.h, i, ctypes.byref(name), ctypes.byref(off), dims))
            self.tensors.append((name.value.decode(), off.value, tuple(dims)))

    def get_params(self):
        out = np.empty(self.padded, dtype=np.float32)
        check(self.L.ufomi_unet_get_params(self.ctx.h, self.h, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == self.padded
        check(self.L.ufomi_unet_set_params(self.ctx.h, self.h, p.ctypes.data_as(ctypes.c_void_p)))

    def unpack(self, flat):
        """{name: array of the tensor's dims} views of a flat parameter array."""
        out = {}
        for name, off, dims in self.tensors:
            n = int(np.prod(dims))
            out[name] = flat[off:off + n].reshape(dims)
        return out

    def forward(self, d_in, d_out, batch=None):
        batch = d_in.shape[0] if batch is None else batch
        check(self.L.ufomi_unet_forward(self.ctx.h, self.h, _ptr(d_in), batch, _ptr(d_out)))

    def gradients(self, d_in, d_target, batch=None):
        batch = d_in.shape[0] if batch is None else batch
        loss = ctypes.c_float()
        g = np.empty(self.padded, dtype=np.float32)
        check(self.L.ufomi_unet_gradients(self.ctx.h, self.h, _ptr(d_in), _ptr(d_target), batch, ctypes.byref(loss),
                                          g.ctypes.data_as(ctypes.c_void_p)))
        return loss.value, g

    def train_step(self, d_in, d_target, lr, batch=None, want_loss=True):
        batch = d_in.shape[0] if batch is None else batch
        loss = ctypes.c_float()
        check(self.L.ufomi_unet_train_step(self.ctx.h, self.h, _ptr(d_in), _ptr(d_target), batch, lr,
                                           ctypes.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def close(self):
        if self.h:
            self.L.ufomi_unet_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# the `parameter` enum of general-backproject, in order
GBP_PARAMETERS = ["axis-angle-x", "axis-angle-y", "axis-angle-z", "volume-angle-x", "volume-angle-y", "volume-angle-z",
                  "detector-angle-x", "detector-angle-y", "detector-angle-z", "detector-position-x",
                  "detector-position-y", "detector-position-z", "source-position-x", "source-position-y",
                  "source-position-z", "center-position-x", "center-position-z", "z"]


def bp_params(num_projections=0, offset=0, axis_pos=-1.0, angle_step=-1.0, angle_offset=0.0, mode="texture",
              roi_x=0, roi_y=0, roi_width=0, roi_height=0):
    p = BpParams()
    load().ufomi_bp_params_init(ctypes.byref(p))
    p.num_projections = num_projections
    p.offset = offset
    p.axis_pos = axis_pos
    p.angle_step = angle_step
    p.angle_offset = angle_offset
    p.mode = BP_TEXTURE if mode == "texture" else BP_NEAREST
    p.roi_x, p.roi_y, p.roi_width, p.roi_height = roi_x, roi_y, roi_width, roi_height
    return p
