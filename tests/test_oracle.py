"""Pins the CPU oracle (oracle/ufo_oracle.c) before the GPU parity tests trust it.

Each check uses a source independent of the oracle's own formula: numpy.fft
golden vectors, closed forms, brute force on tiny inputs, invariants
(linearity, symmetry), or a second geometric formulation (matrix algebra).
A planted mistake in the oracle (dropped term, wrong sign, swapped operand)
fails at least one of them; see test_planted_mistakes_are_caught.
"""
import math
import os

import numpy as np
import pytest

import oracle
from conftest import rel_l2

GOLDEN = np.load(os.path.join(os.path.dirname(__file__), "golden", "golden.npz"))


# ------------------------------------------------------------------------- FFT

@pytest.mark.parametrize("n", [4, 8, 16, 64, 128])
def test_fft_vs_naive_dft(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for d in (-1, 1):
        assert np.abs(oracle.fft_c2c(x, d) - oracle.dft_naive(x, d)).max() < 1e-12 * n


def test_naive_dft_vs_numpy():
    x = np.random.default_rng(3).standard_normal(48)
    np.testing.assert_allclose(oracle.dft_naive(x), np.fft.fft(x), atol=1e-11)


def test_fft_task_vs_golden():
    for key in GOLDEN.files:
        if key.startswith("fft_in_"):
            p, w, n = map(int, key.split("_")[2:])
            got = oracle.fft_task(GOLDEN[key], n)
            ref = GOLDEN["fft_out_%d_%d_%d" % (p, w, n)]
            assert np.abs(got[:, 0::2] + 1j * got[:, 1::2] - ref).max() < 1e-9


def test_fft_round_trip_and_parseval():
    x = np.random.default_rng(1).standard_normal(256) + 0j
    X = oracle.fft_c2c(x, -1)
    np.testing.assert_allclose(oracle.fft_c2c(X, +1) / 256, x, atol=1e-13)
    assert abs(np.sum(np.abs(X) ** 2) / 256 - np.sum(np.abs(x) ** 2)) < 1e-9


def test_ifft_task_crop_and_scale():
    p, w, n = 5, 30, 64
    x = np.linspace(-1, 2, num=p * w, dtype=np.float32).reshape(p, w)
    back = oracle.ifft_task(oracle.fft_task(x, n), w)
    np.testing.assert_allclose(back, x, atol=1e-12)


# --------------------------------------------------------- filter coefficients

def _mirror_ok(c):
    w = len(c)
    for k in range(w // 2 + 2, w, 2):
        assert c[k] == c[w - k] and c[k + 1] == c[w - k + 1]


@pytest.mark.parametrize("n", [8, 64, 1024, 4096])
def test_ramp_fromreal_closed_form(n):
    c = oracle.filter_coefficients(2 * n, "ramp-fromreal")
    h = GOLDEN["ramp_fromreal_%d" % n]
    np.testing.assert_allclose(c[0::2], h, atol=2e-8)
    np.testing.assert_allclose(c[1::2], h, atol=2e-8)
    assert np.array_equal(c[0::2], c[1::2])           # real spectrum (re == im of (1+i) H)
    assert np.array_equal(c[2::2], c[2::2][::-1])     # even: H_k == H_{N-k}
    assert 0 < c[0] < 0.25 / n * 4                   # small positive DC


@pytest.mark.parametrize("n", [64, 1024, 4096])
def test_shepp_logan_closed_form(n):
    """The Shepp-Logan-windowed ramp |f| sin(pi f)/(pi f) = sin(pi |f|)/pi, f in cycles per
    sample: the oracle's spectrum is within the truncation error of the real-space Ram-Lak
    kernel (~0.2/N, the same as ramp-fromreal's distance to |f|) of that closed form, and
    exactly 2/pi times ramp-fromreal at Nyquist."""
    sl = oracle.filter_coefficients(2 * n, "shepp-logan")
    rf = oracle.filter_coefficients(2 * n, "ramp-fromreal")[0::2].astype(np.float64)
    k = np.arange(n)
    f = np.minimum(k, n - k) / n
    assert np.abs(sl[0::2] - np.sin(np.pi * f) / np.pi).max() < 0.21 / n
    assert np.abs(rf - f).max() < 0.21 / n
    assert sl[n] == pytest.approx(rf[n // 2] * 2 / np.pi, rel=1e-6)
    assert sl[0] == np.float32(rf[0])
    assert np.array_equal(sl[0::2], sl[1::2])
    _mirror_ok(sl)


@pytest.mark.parametrize("scale", [1.0, 0.5])
def test_ramp_closed_form(scale):
    n = 64
    c = oracle.filter_coefficients(2 * n, "ramp", scale=scale)
    k = np.arange(1, n // 2 + 1)
    np.testing.assert_allclose(c[2 * k], (k / n * scale).astype(np.float32), rtol=1e-7)
    assert c[0] == np.float32(0.5 / (2 * n))
    _mirror_ok(c)


def test_window_filters_closed_form():
    n = 128
    f = np.arange(n // 2 + 1) / n
    bw = f / (1 + (f / np.float32(0.5)) ** (2 * 4.0))
    np.testing.assert_allclose(oracle.filter_coefficients(2 * n, "butterworth")[0:n + 2:2], bw, rtol=1e-6, atol=1e-9)
    hm = np.where(f < 0.5, f * (0.54 + 0.46 * np.cos(np.pi * f / 0.5)), 0.0)
    np.testing.assert_allclose(oracle.filter_coefficients(2 * n, "hamming")[0:n + 2:2], hm, rtol=1e-6, atol=1e-9)
    bh = f * (0.42 + 0.5 * np.cos(f * np.pi) + 0.08 * np.cos(2 * f * np.pi))
    got = oracle.filter_coefficients(2 * n, "bh3")
    np.testing.assert_allclose(got[2:n + 2:2], bh[1:], rtol=1e-6, atol=1e-9)
    assert got[0] == np.float32(0.5 / (2 * n))  # bh3 keeps the DC of :303-304
    for name in ("butterworth", "hamming", "bh3"):
        _mirror_ok(oracle.filter_coefficients(2 * n, name))


def test_faris_byer_index_pattern():
    # Value of every index as derived by hand from ufo-filter-task.c:254-276:
    # odd j <= W/2-3 -> -sin(theta)/2 / ((j+2)^2 pi^2 tau); j = W/2-1 -> 1/(pi^2 tau (W/2-1));
    # other indices below W/2 -> 0, except index 1 when W/2-1 == 1; then the mirror.
    w, tau, theta = 32, np.float32(0.1), np.float32(1.0)
    c = oracle.filter_coefficients(w, "faris-byer", tau=tau, theta=theta)
    pst = np.pi ** 2 * float(tau)
    exp = np.zeros(w)
    for j in range(1, w // 2 - 2, 2):
        exp[j] = -math.sin(theta) / 2 / ((j + 2) ** 2 * pst)
    exp[w // 2 - 1] = 1 / (pst * (w // 2 - 1))
    exp[w // 2 + 1] = 1 / (pst * (w // 2 + 1))  # x = W/2+1 is written by the first loop (x <= W/2 only when odd)... see below
    # The first loop stops at x <= W/2, so index W/2+1 is never written before the mirror:
    exp[w // 2 + 1] = 0.0
    for k in range(w // 2 + 2, w, 2):
        exp[k] = exp[w - k]
        exp[k + 1] = exp[w - k + 1]
    np.testing.assert_allclose(c, exp.astype(np.float32), rtol=1e-6, atol=1e-12)


def test_non_power_of_two_rejected():
    with pytest.raises(ValueError):
        oracle.filter_coefficients(96)


def test_ramp_chain_vs_numpy():
    p, w, n = 6, 50, 128
    x = np.random.default_rng(8).standard_normal((p, w)).astype(np.float32)
    c = oracle.filter_coefficients(2 * n, "hamming")
    X = np.fft.fft(x.astype(np.float64), n=n, axis=1)
    Y = X.real * c[0::2] + 1j * X.imag * c[1::2]
    ref = np.fft.ifft(Y, axis=1).real[:, :w]
    np.testing.assert_allclose(oracle.ramp_chain(x, n, c, w), ref, atol=1e-12)


# ------------------------------------------------------------ backprojection

@pytest.mark.parametrize("filt", ["ramp-fromreal", "shepp-logan", "faris-byer"])
def test_ramp_chain_f32_vs_f64(filt):
    # the CPU baseline's float32 chain against the FP64 chain (both restate fft ! filter ! ifft)
    rng = np.random.default_rng(4)
    sino = rng.standard_normal((9, 300)).astype(np.float32)
    coef = oracle.filter_coefficients(1024, filt)
    f64 = oracle.ramp_chain(sino, 512, coef, 300)
    f32 = oracle.ramp_chain_f32(sino, 512, coef, 300)
    assert f32.dtype == np.float32 and f32.shape == (9, 300)
    assert oracle.rel_l2(f32, f64) < 1e-6


def test_bp_vs_brute_force_golden():
    sino = GOLDEN["bp_brute_sino"]
    axis = float(GOLDEN["bp_brute_axis"])
    ref = GOLDEN["bp_brute_out"]
    np.testing.assert_allclose(oracle.bp_texture_f64(sino, axis_pos=axis), ref, atol=1e-12)
    assert rel_l2(oracle.bp_texture_ref32(sino, axis_pos=axis), ref) < 1e-6


def test_bp_ref32_close_to_f64():
    s = np.random.default_rng(4).standard_normal((90, 64)).astype(np.float32)
    assert rel_l2(oracle.bp_texture_ref32(s), oracle.bp_texture_f64(s)) < 1e-5


def test_bp_constant_sinogram():
    w = 64
    img = oracle.bp_texture_f64(np.full((50, w), 3.0))
    y, x = np.mgrid[0:w, 0:w] + 0.5 - w / 2
    inside = np.hypot(x, y) < w / 2 - 1.5
    np.testing.assert_allclose(img[inside], 3.0 * math.pi, rtol=1e-12)
    assert np.all(img[np.hypot(x, y) > w / 2 * math.sqrt(2)] == 0)


def test_bp_linearity_and_pi_flip():
    rng = np.random.default_rng(6)
    a, b = rng.standard_normal((2, 20, 32))
    np.testing.assert_allclose(oracle.bp_texture_f64(2 * a - 3 * b),
                               2 * oracle.bp_texture_f64(a) - 3 * oracle.bp_texture_f64(b), atol=1e-11)
    # shifting every angle by pi rotates the image by pi about the axis (axis = W/2)
    r0 = oracle.bp_texture_f64(a)
    r1 = oracle.bp_texture_f64(a, angle_offset=math.pi)
    np.testing.assert_allclose(r1, r0[::-1, ::-1], atol=1e-10)


def test_bp_nearest_vs_numpy_float32():
    rng = np.random.default_rng(12)
    p, w = 9, 16
    s = rng.standard_normal((p, w)).astype(np.float32)
    got = oracle.bp_nearest_ref32(s)
    sl, cl, _ = oracle.bp_lut(p)
    a = np.float32(w / 2)
    f32 = np.float32
    out = np.zeros((w, w), np.float32)
    flat = s.ravel()
    for iy in range(w):
        for ix in range(w):
            bx = f32(f32(f32(ix) - a) + f32(0)) + f32(0.5)
            by = f32(f32(f32(iy) - a) + f32(0)) + f32(0.5)
            acc = f32(0)
            for q in range(p):
                h = f32(f32(a + f32(bx * cl[q])) + f32(by * sl[q]))
                i = int(f32(f32(q * w) + h))
                if 0 <= i < flat.size:
                    acc = f32(acc + flat[i])
            out[iy, ix] = f32(f32(acc * f32(math.pi)) / f32(p))
    np.testing.assert_array_equal(got, out)


def test_fbp_of_disc_approaches_phantom():
    # analytic sinogram of a centred disc of radius R: 2 sqrt(R^2 - t^2)
    w, p, n, r = 128, 256, 256, 40.0
    t = np.arange(w) + 0.5 - w / 2
    sino = np.tile(2 * np.sqrt(np.maximum(r * r - t * t, 0)), (p, 1)).astype(np.float32)
    filt = oracle.ramp_chain(sino, n, oracle.filter_coefficients(2 * n), w)
    img = oracle.bp_texture_f64(filt)
    y, x = np.mgrid[0:w, 0:w] + 0.5 - w / 2
    rr = np.hypot(x, y)
    assert abs(np.mean(img[rr < r - 4]) - 1.0) < 0.02
    assert abs(np.mean(img[(rr > r + 4) & (rr < w / 2 - 4)])) < 0.02


# --------------------------------------------------------------- cone weight

def test_cone_weight_closed_form():
    h, w = 5, 7
    x = np.random.default_rng(2).standard_normal((h, w)).astype(np.float32)
    cx, cz, sod, odd, tilt = 3.2, 2.1, 500.0, 700.0, 0.3
    got = oracle.cone_weight_ref32(x, cx, cz, sod, odd, tilt)
    m = sod / (sod + odd)
    i = (np.arange(w)[None, :] - cx + 0.5) * m
    j = (np.arange(h)[:, None] - cz + 0.5) * m
    L = sod + odd
    ref = x * sod * math.cos(tilt) / np.sqrt(L * L + i * i + j * j)
    np.testing.assert_allclose(got, ref, rtol=2e-6)
    got0 = oracle.cone_weight_ref32(x, cx, cz, sod, odd, 0.0)  # cos > 0.9999999 -> L = SOD
    ref0 = x * sod / np.sqrt(sod * sod + i * i + j * j)
    np.testing.assert_allclose(got0, ref0, rtol=2e-6)


# -------------------------------------------------------- general backproject

def test_gbp_parallel_equals_backproject():
    # untilted parallel general-backproject == backproject after the y flip and
    # half-pixel grid shift: x = bx, y = -by, center-x = axis, one detector row.
    rng = np.random.default_rng(21)
    p, w = 24, 32
    sino = rng.standard_normal((p, w)).astype(np.float32)
    a = 15.5
    projs = sino[:, None, :]
    vol = oracle.gbp(projs, overall_angle=math.pi, x_region=(0.5 - a, 0.5 - a + w, 1.0),
                     y_region=(a - 0.5, a - 0.5 - w, -1.0), center_position_x=[a], center_position_z=[0.5])
    ref = oracle.bp_texture_f64(sino, axis_pos=a)
    np.testing.assert_allclose(vol[0], ref, atol=1e-10)
    vol32 = oracle.gbp(projs, overall_angle=math.pi, x_region=(0.5 - a, 0.5 - a + w, 1.0),
                       y_region=(a - 0.5, a - 0.5 - w, -1.0), center_position_x=[a], center_position_z=[0.5],
                       mode="ref32")
    assert rel_l2(vol32[0], ref) < 1e-5


def _matrix_chain(x, y, z, th, tilt, cx, cz):
    # Second formulation of parallel laminography: R = Rx(tilt) Rz(theta) as matrices.
    rz = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]])
    rx = np.array([[1, 0, 0], [0, math.cos(tilt), -math.sin(tilt)], [0, math.sin(tilt), math.cos(tilt)]])
    v = rx @ rz @ np.array([x, y, z])
    return v[0] + cx, v[2] + cz


def _bilinear0(img, u, v):
    h, w = img.shape
    u, v = u - 0.5, v - 0.5
    i0, j0 = math.floor(u), math.floor(v)
    a, b = u - i0, v - j0
    def t(i, j):
        return float(img[j, i]) if 0 <= i < w and 0 <= j < h else 0.0
    return ((1 - a) * (1 - b) * t(i0, j0) + a * (1 - b) * t(i0 + 1, j0)
            + (1 - a) * b * t(i0, j0 + 1) + a * b * t(i0 + 1, j0 + 1))


def test_gbp_laminography_vs_matrix_brute_force():
    rng = np.random.default_rng(31)
    p, h, w = 6, 9, 10
    projs = rng.standard_normal((p, h, w)).astype(np.float32)
    tilt, cx, cz = 0.4, 5.0, 4.5
    vol = oracle.gbp(projs, axis_angle_x=[tilt], center_position_x=[cx], center_position_z=[cz],
                     region=(-2.0, 2.0, 1.0))
    nz, ny, nx = vol.shape
    ref = np.zeros_like(vol)
    for iz in range(nz):
        for iy in range(ny):
            for ix in range(nx):
                x, y, z = -nx / 2 + ix, -ny / 2 + iy, -2.0 + iz
                s = 0.0
                for i in range(p):
                    u, v = _matrix_chain(x, y, z, i / p * 2 * math.pi, tilt, cx, cz)
                    s += _bilinear0(projs[i], u, v)
                ref[iz, iy, ix] = s * 2 * math.pi / p
    np.testing.assert_allclose(vol, ref, atol=1e-9)


def test_gbp_cone_isocentre_closed_form():
    # Constant projections: the isocentre voxel sees coeff = (S_y - D_y) / S_y at every angle,
    # so the volume there is norm * P * coeff^2 = |overall| * coeff^2.
    p, h, w = 12, 16, 16
    projs = np.ones((p, h, w), np.float32)
    sy, dy = -40.0, 40.0
    vol = oracle.gbp(projs, source_position_y=[sy], detector_position_y=[dy], center_position_x=[8.0],
                     center_position_z=[8.0], region=(0.0, 1.0, 1.0))
    coeff = (sy - dy) / sy
    assert abs(vol[0, 8, 8] - 2 * math.pi * coeff ** 2) < 1e-9


def test_gbp_cone_magnification_closed_form():
    # A voxel (x, 0, 0) at theta = 0 lands on u = x + cx (magnification S/(D-S) then (D-S)/S).
    p, h, w = 1, 8, 32
    projs = np.zeros((p, h, w), np.float32)
    projs[0, 4, 20] = 1.0  # texel centre (20.5, 4.5)
    sy, dy, cx, cz = -100.0, 100.0, 16.0, 4.5
    vol = oracle.gbp(projs, source_position_y=[sy], detector_position_y=[dy], center_position_x=[cx],
                     center_position_z=[cz], x_region=(4.5, 4.5 + 1, 1.0), y_region=(0.0, 1.0, 1.0),
                     region=(0.0, 1.0, 1.0))
    # x = 4.5 -> u = 20.5 exactly: the texel's full weight; coeff = (S_y - D_y) / S_y = 2 at y = 0
    assert abs(vol[0, 0, 0] - 2 * math.pi * 4) < 1e-9


def test_planted_mistakes_are_caught():
    """A wrong sign or a swapped operand in the BP formula is caught by the brute-force golden."""
    sino = GOLDEN["bp_brute_sino"]
    axis = float(GOLDEN["bp_brute_axis"])
    ref = GOLDEN["bp_brute_out"]
    good = oracle.bp_texture_f64(sino, axis_pos=axis)
    assert np.abs(good - ref).max() < 1e-12
    # planted on the CPU: sin and cos swapped (equivalently, transposed image axes),
    # the +0.5 pixel centre dropped (axis shifted by half a pixel), the angle
    # order reversed
    assert np.abs(good.T - ref).max() > 1e-3
    assert np.abs(oracle.bp_texture_f64(sino, axis_pos=axis + 0.5) - ref).max() > 1e-3
    assert np.abs(oracle.bp_texture_f64(sino[::-1], axis_pos=axis) - ref).max() > 1e-3


def _transpose_flat_brute(projs, number):
    """The reference's own loop over flat memory (ufo-transpose-projections-task.c:73-90):
    projection p (1-based) lands at (p - 1) * width + i * width * number for row i."""
    n, h, w = projs.shape
    flat = np.zeros(number * w * h, dtype=np.float32)
    for proj in range(1, n + 1):
        if proj > number:
            break
        sino_index = (proj - 1) * w
        src = projs[proj - 1].ravel()
        for i in range(h):
            flat[sino_index + i * w * number: sino_index + i * w * number + w] = src[i * w:(i + 1) * w]
    return flat.reshape(h, number, w)


@pytest.mark.parametrize("n,h,w,number", [(5, 3, 7, 5), (4, 6, 2, 9), (9, 2, 5, 6), (1, 1, 1, 1)])
def test_transpose_projections_vs_reference_loop(n, h, w, number):
    projs = np.random.default_rng(n * 100 + number).standard_normal((n, h, w)).astype(np.float32)
    got = oracle.transpose_projections(projs, number)
    np.testing.assert_array_equal(got, _transpose_flat_brute(projs, number))
    # involution when every row is filled: transposing the sinograms back gives the projections
    if n == number:
        np.testing.assert_array_equal(oracle.transpose_projections(got, h), projs)


@pytest.mark.parametrize("shape,dims,m", [((3, 5, 7), 1, (7, 5, 3)), ((3, 5, 7), 2, (9, 6, 3)),
                                          ((2, 3, 4), 3, (5, 3, 4)), ((1, 4, 6), 1, (16, 4, 1))])
def test_fft_nd_vs_numpy_and_brute_force(shape, dims, m):
    """fft_nd is the DFT of the zero-padded item: numpy's FFT (an independent routine) and the C
    brute-force DFT agree with it; ifft_nd inverts it."""
    x = np.random.default_rng(sum(shape) + dims).standard_normal(shape)
    got = oracle.fft_nd(x, dims, m)
    pad = np.zeros((m[2], m[1], m[0]))
    pad[:shape[0], :shape[1], :shape[2]] = x
    ref = np.fft.fftn(pad, axes=tuple(range(2, 2 - dims, -1)))
    np.testing.assert_allclose(got, ref, atol=1e-9)
    for line in pad.reshape(-1, m[0])[:3]:
        if dims >= 1:
            np.testing.assert_allclose(np.fft.fft(line), oracle.dft_naive(line), atol=1e-9)
    np.testing.assert_allclose(oracle.ifft_nd(got, dims), pad, atol=1e-9)
    # a planted mistake (conjugated kernel) is caught
    assert np.abs(np.conj(got) - ref).max() > 1e-3


def _stacked_int8_brute(sinos, axis, roi=(0, 0, 0, 0)):
    """stacked-backproject.cl texture_uint / uninterleave_uint, one pixel at a time in float32."""
    f32 = np.float32
    n_sl, p_in, w = sinos.shape
    rx, ry, rw, rh = roi
    rw, rh = rw or w, rh or w
    s, c, _ = oracle.bp_lut(p_in)
    lo, hi = sinos.min(), sinos.max()
    scale = f32(255.0) / (hi - lo)
    out = np.zeros((n_sl, rh, rw), np.float32)
    for sl in range(n_sl):
        for y in range(rh):
            q = (y % 16) // 4
            y0 = y - 4 * q
            for x in range(rw):
                bx = f32(f32(f32(x) - f32(axis)) + f32(rx)) + f32(0.5)
                by = f32(f32(f32(y0) - f32(axis)) + f32(ry)) + f32(0.5)
                acc = 0
                for p in range(p_in):
                    h = f32(f32(f32(bx * c[p]) - f32(by * s[p])) + f32(axis))
                    h = f32(h - f32(f32(4 * q) * s[p]))
                    i = min(max(int(math.floor(h)), 0), w - 1)
                    v = f32(f32(sinos[sl, p, i] - lo) * scale)
                    acc += 0 if not v > 0 else min(int(math.floor(v)), 255)
                r = f32(f32(f32(f32(acc) * f32((hi - lo) / f32(255.0))) + lo) * f32(math.pi))
                out[sl, y, x] = f32(r / f32(p_in))
    return out


def test_stacked_oracle_int8_vs_brute_force():
    sinos = np.random.default_rng(4).standard_normal((3, 7, 11)).astype(np.float32)
    got = oracle.stacked_bp_ref32(sinos, "int8", roi_x=1, roi_y=2, roi_width=9, roi_height=18)
    ref = _stacked_int8_brute(sinos, np.float32(5.5), roi=(1, 2, 9, 18))
    np.testing.assert_array_equal(got, ref)


def test_stacked_oracle_single_is_backproject_of_negated_angles():
    """h = axis + bx cos - by sin is backproject's h with every angle negated: the C oracle
    with a (-sin, cos) table gives the same slices up to float rounding; a plain table does not."""
    sinos = np.random.default_rng(5).standard_normal((2, 40, 32)).astype(np.float32)
    got = oracle.stacked_bp_ref32(sinos, "single", axis_pos=15.25, roi_x=3, roi_width=24, roi_height=20)
    s, c, _ = oracle.bp_lut(40)
    for k in range(2):
        out = np.empty((20, 24), np.float32)
        oracle.lib().orc_bp_texture_ref32(np.ascontiguousarray(sinos[k]), 32, 40, -s, c, 0, 15.25, 3, 0, 24, 20, out)
        assert oracle.rel_l2(got[k], out) < 5e-6
        plain = oracle.bp_texture_ref32(sinos[k], axis_pos=15.25, roi_x=3, roi_width=24, roi_height=20)
        assert oracle.rel_l2(got[k], plain) > 1e-2
    half = oracle.stacked_bp_ref32(sinos, "half", axis_pos=15.25, roi_x=3, roi_width=24, roi_height=20)
    rounded = oracle.stacked_bp_ref32(sinos.astype(np.float16).astype(np.float32), "single", axis_pos=15.25,
                                      roi_x=3, roi_width=24, roi_height=20)
    np.testing.assert_array_equal(half, rounded)


# ------------------------------------------------------------ lamino-backproject

def test_lamino_oracle_at_right_angle_is_backproject():
    """lamino-angle pi/2, roll 0: slice z reads projection row z + y-centre, u = x cos t + y sin t + x-centre,
    which is backproject's h = axis + bx cos + by sin with bx = x (axis 15.5, x-region from -15)."""
    rng = np.random.default_rng(41)
    p, h, w = 24, 5, 32
    projs = rng.standard_normal((p, h, w)).astype(np.float32)
    vol = oracle.lamino(projs, lamino_angle=math.pi / 2, center=(15.5, 0.5), x_region=(-15, 17, 1),
                        y_region=(-15, 17, 1), region=(0.0, float(h), 1.0))
    for r in range(h):
        ref = oracle.bp_texture_f64(projs[:, r, :], axis_pos=15.5)
        # float32 pi/2 leaves cos(l) = -4.4e-8: rows blend by at most 1e-6
        assert rel_l2(vol[r], ref) < 2e-6


def test_lamino_oracle_is_general_backproject_laminography():
    """lamino-backproject(l) equals general-backproject with axis-angle-x = l - pi/2 and axis-angle-z = -t
    (the latter pinned by the matrix brute force above)."""
    rng = np.random.default_rng(42)
    p, h, w = 12, 20, 24
    projs = rng.standard_normal((p, h, w)).astype(np.float32)
    lam, xc, yc = 0.6, 11.25, 9.5
    vol = oracle.lamino(projs, lamino_angle=lam, center=(xc, yc), x_region=(-8, 9, 2), y_region=(-6, 6, 1),
                        region=(-3.0, 3.0, 1.5))
    lam32 = float(np.float32(lam))
    pi32 = float(np.float32(math.pi))  # the float property
    t = [-pi32 * i / p for i in range(p)]
    ref = oracle.gbp(projs, overall_angle=pi32, axis_angle_x=[lam32 - math.pi / 2], axis_angle_z=t,
                     center_position_x=[xc], center_position_z=[yc], x_region=(-8.0, 10.0, 2.0),
                     y_region=(-6.0, 6.0, 1.0), region=(-3.0, 3.0, 1.5))
    assert vol.shape == ref.shape == (4, 12, 9)
    np.testing.assert_allclose(vol, ref, atol=1e-9)


def test_lamino_oracle_roll_closed_forms():
    """definitions.in:1-6 rotate() updates x before computing y from it.  Roll pi is a point reflection
    about the centre either way; roll pi/2 maps (a, b) to (-b, -b) (a true rotation would give (-b, a)),
    so projections linear in u - v (which bilinear sampling reproduces) reconstruct to a constant."""
    rng = np.random.default_rng(43)
    p, h, w = 8, 16, 16
    projs = rng.standard_normal((p, h, w)).astype(np.float32)
    kw = dict(lamino_angle=0.7, center=(8.0, 8.0), x_region=(-4, 4, 1), y_region=(-4, 4, 1), region=(-1.0, 1.0, 1.0))
    flipped = oracle.lamino(projs[:, ::-1, ::-1], **kw)
    # the float32 roll property leaves sin(pi) = -8.7e-8: sub-1e-6 pixel shifts
    np.testing.assert_allclose(oracle.lamino(projs, roll_angle=math.pi, **kw), flipped, atol=1e-5)
    u, v = np.meshgrid(np.arange(w), np.arange(h))
    diag = np.broadcast_to(0.3 * (u - v) + 1.7, (p, h, w)).astype(np.float32)
    kw2 = dict(kw, center=(8.5, 8.5), x_region=(-2, 2, 1), y_region=(-2, 2, 1))
    rolled = oracle.lamino(diag, roll_angle=math.pi / 2, **kw2)
    # every sample at u - v = 0 inside the image: 1.7 per projection, times pi / p
    np.testing.assert_allclose(rolled, float(np.float32(math.pi)) * 1.7, atol=1e-5)
    assert np.ptp(oracle.lamino(diag, roll_angle=0.0, **kw2)) > 1e-2


def test_lamino_oracle_sweeps_are_slices():
    """Each swept slice equals a z run with the parameter at that value; the x-center sweep's rotate()
    pivots on the first centre, which at roll pi equals a z run with centre 2 c0 - c_k.  (Swept values are
    float32 region values stepped in double, z runs take float32 properties: 1e-7 apart.)"""
    rng = np.random.default_rng(44)
    p, h, w = 10, 14, 18
    projs = rng.standard_normal((p, h, w)).astype(np.float32)
    base = dict(lamino_angle=0.5, roll_angle=0.1, center=(9.0, 7.0), x_region=(-5, 5, 1), y_region=(-4, 4, 1),
                z=0.75)
    zrun = lambda **kw: oracle.lamino(projs, **dict(base, region=(0.75, 1.75, 1.0), **kw))[0]
    sw = oracle.lamino(projs, **dict(base, parameter="lamino-angle", region=(0.3, 0.9, 0.2)))
    for k in range(3):
        np.testing.assert_allclose(sw[k], zrun(lamino_angle=0.3 + 0.2 * k), atol=1e-6)
    sw = oracle.lamino(projs, **dict(base, parameter="roll-angle", region=(-0.2, 0.2, 0.2)))
    for k in range(2):
        np.testing.assert_allclose(sw[k], zrun(roll_angle=-0.2 + 0.2 * k), atol=1e-6)
    sw = oracle.lamino(projs, **dict(base, roll_angle=math.pi, parameter="x-center", region=(8.0, 10.0, 0.5)))
    for k in range(4):
        c = 8.0 + 0.5 * k
        np.testing.assert_allclose(sw[k], zrun(roll_angle=math.pi, center=(2 * 8.0 - c, 7.0)), atol=1e-5)
    ref32 = oracle.lamino(projs, **dict(base, parameter="x-center", region=(8.0, 10.0, 0.5)), mode="ref32")
    f64 = oracle.lamino(projs, **dict(base, parameter="x-center", region=(8.0, 10.0, 0.5)))
    assert rel_l2(ref32, f64) < 1e-5


def test_lamino_oracle_ref32_bursts():
    """float32 semantics with bursts of 16: close to FP64 with all projections; with fewer projections
    than num-projections the reference launches its 16-wide kernel only on full bursts, so a trailing
    partial burst is left out."""
    rng = np.random.default_rng(45)
    projs = rng.standard_normal((40, 12, 16)).astype(np.float32)
    kw = dict(lamino_angle=0.4, center=(8.0, 6.0), x_region=(-6, 6, 1), y_region=(-6, 6, 1), region=(-1.0, 1.0, 1.0))
    assert rel_l2(oracle.lamino(projs, mode="ref32", **kw), oracle.lamino(projs, **kw)) < 1e-5
    part = oracle.lamino(projs[:20], num_projections=40, mode="ref32", **kw)
    first16 = oracle.lamino(projs[:16], num_projections=40, **kw)
    assert rel_l2(part, first16) < 1e-5


# ------------------------------------------------------ flat-field-correct, average

def test_ffc_oracle_closed_forms():
    """Data built as dark + k (flat - dark) corrects to k, and as dark + (flat - dark) exp(-a) to a
    with absorption; scales act on dark and flat; a sinogram-input row broadcasts over the rows;
    zero denominators give NaN/Inf, or 0 with the fix."""
    rng = np.random.default_rng(51)
    dark = rng.uniform(90, 110, (6, 9))
    flat = rng.uniform(900, 1000, (6, 9))
    k = rng.uniform(0.1, 0.9, (3, 6, 9))
    data = (dark + k * (flat - dark)).astype(np.float32)
    np.testing.assert_allclose(oracle.flat_field_correct_ref32(data, dark, flat), k, rtol=1e-4)
    att = -np.log(k)
    np.testing.assert_allclose(oracle.flat_field_correct_ref32(data, dark, flat, absorption_correct=True), att,
                               rtol=1e-4, atol=1e-5)
    np.testing.assert_array_equal(oracle.flat_field_correct_ref32(data, dark, flat, dark_scale=0.5, flat_scale=2.0),
                                  oracle.flat_field_correct_ref32(data, np.float32(dark) * np.float32(0.5),
                                                                  np.float32(flat) * np.float32(2.0)))
    row_d, row_f = dark[0], flat[0]
    np.testing.assert_array_equal(oracle.flat_field_correct_ref32(data, row_d, row_f, sinogram_input=True),
                                  oracle.flat_field_correct_ref32(data, np.tile(row_d, (6, 1)),
                                                                  np.tile(row_f, (6, 1))))
    z = data.copy()
    z[0, 0, 0] = np.float32(dark[0, 0])
    bad = oracle.flat_field_correct_ref32(z, dark, flat, absorption_correct=True)
    assert np.isinf(bad[0, 0, 0])
    assert oracle.flat_field_correct_ref32(z, dark, flat, absorption_correct=True, fix_nan_and_inf=True)[0, 0, 0] == 0


def test_average_oracle_closed_forms():
    """Integer-valued items average exactly to the mean; identical items average to themselves; the sum is
    float32 in arrival order (1e8 + 1 - 1e8 loses the 1, as the reference's += does)."""
    rng = np.random.default_rng(52)
    items = rng.integers(-1000, 1000, (7, 5, 4)).astype(np.float32)
    np.testing.assert_array_equal(oracle.average_ref32(items), (items.astype(np.float64).sum(0) / 7).astype(np.float32))
    one = rng.standard_normal((5, 4)).astype(np.float32)
    np.testing.assert_allclose(oracle.average_ref32(np.stack([one] * 4)), one, rtol=1e-7)
    seq = np.array([[1e8], [1.0], [-1e8]], np.float32)
    assert oracle.average_ref32(seq)[0] == 0.0
