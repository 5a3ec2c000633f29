"""Writes the golden vectors the oracle is pinned against (tests/golden/*.npz).

Run from the repository root: python tests/golden/make_golden.py
Sources of truth, none of which is the oracle itself:
  fft_*        numpy.fft (the ground truth of the reference's tests/test_fft.py:15-23)
               on the reference test's own input, np.linspace(-1, 2) (:47);
  ramp_*       the closed form of the DFT of the real-space Ram-Lak kernel,
               H_k = 1/4 - (2/pi^2) sum_{odd n < N/2} cos(2 pi k n / N) / n^2;
  bp_brute_*   a pure-Python loop over the OpenCL 1.2 linear-sampler definition
               (u = h - 0.5, i0 = floor(u), a = u - i0, zero border) on a tiny
               sinogram, in float64.
"""
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def fft_vectors():
    out = {}
    for (p, w, n) in [(17, 33, 64), (16, 32, 32), (8, 8, 16), (3, 9, 16), (2, 100, 128), (1, 512, 1024)]:
        x = np.linspace(-1, 2, num=w * p, dtype=np.float32).reshape(p, w)
        out["fft_in_%d_%d_%d" % (p, w, n)] = x
        out["fft_out_%d_%d_%d" % (p, w, n)] = np.fft.fft(x.astype(np.float64), n=n, axis=1)
    return out


def ramp_closed_form():
    out = {}
    for n in (8, 64, 1024, 4096):
        k = np.arange(n)[:, None]
        odd = np.arange(1, n // 2, 2)[None, :]
        h = 0.25 - (2 / np.pi ** 2) * np.sum(np.cos(2 * np.pi * k * odd / n) / odd ** 2, axis=1)
        out["ramp_fromreal_%d" % n] = h
    return out


def bp_brute():
    rng = np.random.default_rng(20261202)
    w, p = 6, 4
    sino = rng.standard_normal((p, w)).astype(np.float32)
    axis = 3.25
    step = math.pi / p
    img = np.zeros((w, w))
    for iy in range(w):
        for ix in range(w):
            bx = ix - axis + 0.5
            by = iy - axis + 0.5
            acc = 0.0
            for q in range(p):
                th = q * step
                h = by * math.sin(th) + bx * math.cos(th) + axis
                u = h - 0.5
                i0 = math.floor(u)
                a = u - i0
                t0 = float(sino[q, i0]) if 0 <= i0 < w else 0.0
                t1 = float(sino[q, i0 + 1]) if 0 <= i0 + 1 < w else 0.0
                acc += (1 - a) * t0 + a * t1
            img[iy, ix] = acc * math.pi / p
    return {"bp_brute_sino": sino, "bp_brute_axis": np.array(axis), "bp_brute_out": img}


if __name__ == "__main__":
    data = {}
    data.update(fft_vectors())
    data.update(ramp_closed_form())
    data.update(bp_brute())
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **data)
    print("wrote", len(data), "arrays")
