"""The band-limited texture of the learned-stage phantoms: the closed-form line
integral (tests/phantoms.py) against brute-force quadrature along the ray, and
the seeded generator's statistics."""
import numpy as np

import phantoms
import ufomi


def test_wave_line_integral_closed_form_vs_quadrature():
    rng = np.random.default_rng(0)
    waves = np.stack([rng.uniform(-0.5, 0.5, 6), rng.uniform(-0.5, 0.5, 6), rng.uniform(0, 6.3, 6),
                      rng.uniform(0.1, 1.0, 6)], axis=1)
    R, W, P, axis = 20.0, 48, 7, 24.0
    step = np.pi / P
    closed = phantoms.texture_sinogram(waves, R, W, P, axis, step)
    s = np.linspace(-R, R, 400001)
    ds = s[1] - s[0]
    for p in range(P):
        th = p * step
        n = np.array([np.cos(th), np.sin(th)])
        d = np.array([-np.sin(th), np.cos(th)])
        for j in (0, 5, 17, 24, 40, 47):
            t = j + 0.5 - axis
            pts = t * n[None, :] + s[:, None] * d[None, :]
            inside = (pts ** 2).sum(axis=1) < R ** 2
            val = sum(a * np.cos(kx * pts[:, 0] + ky * pts[:, 1] + ph) for kx, ky, ph, a in waves)
            quad = np.sum(np.where(inside, val, 0.0)) * ds
            assert abs(quad - closed[p, j]) < 2e-3, (p, j, quad, closed[p, j])


def test_texture_generator():
    w = ufomi.phantom_texture(5, 32, 8.0, 64.0, 0.1)
    assert w.shape == (32, 4)
    k = np.hypot(w[:, 0], w[:, 1])
    assert np.all((k >= 2 * np.pi / 64 - 1e-12) & (k <= 2 * np.pi / 8 + 1e-12))
    assert np.allclose(w[:, 3], 0.1 * np.sqrt(2 / 32))
    assert np.array_equal(w, ufomi.phantom_texture(5, 32, 8.0, 64.0, 0.1))
    img = phantoms.texture_image(w, 1e9, 256, 128.0)
    assert 0.07 < img.std() < 0.13  # RMS 0.1 for incoherent waves
