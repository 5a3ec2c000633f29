"""numpy restatement of the synthetic phantoms used by the tests.

Independent of the GPU generator in ufo-filters_amd/csrc/phantom.hip, against
which it is compared (tests/test_gpu_phantom.py).  Ellipse: (x0, y0, a, b,
phi, density) in pixels relative to the rotation axis; detector pixel j
samples x cos(theta) + y sin(theta) = j + 0.5 - axis.
"""
import numpy as np


def sinogram(ellipses, W, P, axis, step):
    th = np.arange(P)[:, None] * step
    c, s = np.cos(th), np.sin(th)
    t = np.arange(W)[None, :] + 0.5 - axis
    out = np.zeros((P, W))
    for x0, y0, a, b, phi, rho in np.atleast_2d(ellipses):
        tp = t - (x0 * c + y0 * s)
        ang = th - phi
        a2 = (a * np.cos(ang)) ** 2 + (b * np.sin(ang)) ** 2
        d = a2 - tp ** 2
        out += np.where(d > 0, 2 * rho * a * b * np.sqrt(np.maximum(d, 0)) / a2, 0.0)
    return out


def image(ellipses, n, axis, roi_x=0, roi_y=0):
    """Phantom sampled at the backprojector's pixel centres (x - axis + 0.5)."""
    x = np.arange(n)[None, :] - axis + roi_x + 0.5
    y = np.arange(n)[:, None] - axis + roi_y + 0.5
    out = np.zeros((n, n))
    for x0, y0, a, b, phi, rho in np.atleast_2d(ellipses):
        dx, dy = x - x0, y - y0
        u = dx * np.cos(phi) + dy * np.sin(phi)
        v = -dx * np.sin(phi) + dy * np.cos(phi)
        out += np.where((u / a) ** 2 + (v / b) ** 2 <= 1.0, rho, 0.0)
    return out


def random_ellipses(rng, n, radius):
    r = np.sqrt(rng.uniform(size=n)) * 0.7
    psi = rng.uniform(0, 2 * np.pi, size=n)
    return np.stack([r * np.cos(psi) * radius, r * np.sin(psi) * radius,
                     rng.uniform(0.05, 0.4, n) * radius, rng.uniform(0.05, 0.4, n) * radius,
                     rng.uniform(0, np.pi, n), rng.uniform(-0.5, 1.0, n)], axis=1)


def texture_sinogram(waves, radius, W, P, axis, step):
    """Line integrals of plane waves (kx, ky, phase, amp) restricted to the disk |x| < radius."""
    th = np.arange(P)[:, None] * step
    c, s = np.cos(th), np.sin(th)
    t = np.arange(W)[None, :] + 0.5 - axis
    L2 = radius ** 2 - t ** 2
    L = np.sqrt(np.maximum(L2, 0.0))
    out = np.zeros((P, W))
    for kx, ky, ph, amp in np.atleast_2d(waves):
        alpha = (kx * c + ky * s) * t + ph
        beta = -kx * s + ky * c
        with np.errstate(divide="ignore", invalid="ignore"):
            f = np.where(np.abs(beta * L) < 1e-8, 2 * L, 2 * np.sin(beta * L) / beta)
        out += np.where(L2 > 0, amp * np.cos(alpha) * f, 0.0)
    return out


def texture_image(waves, radius, n, axis):
    x = np.arange(n)[None, :] + 0.5 - axis
    y = np.arange(n)[:, None] + 0.5 - axis
    out = np.zeros((n, n))
    for kx, ky, ph, amp in np.atleast_2d(waves):
        out += amp * np.cos(kx * x + ky * y + ph)
    return np.where(x ** 2 + y ** 2 < radius ** 2, out, 0.0)
