"""CPU tests of host-side tasks: center-of-rotation
(src/ufo-center-of-rotation-task.c:111-166) through ufomi_task_process_host."""
import numpy as np
import pytest

import ufo_tasks as U


def _center(sino):
    t = U.Task("center-of-rotation")
    t.setup()
    t.process_host(np.ascontiguousarray(sino, dtype=np.float32))
    return t.props.center


def _center_brute(sino):
    """The reference's loops one index at a time (float32 sums in order); reads past the last
    row (mirror index width - x + 1) are 0."""
    sino = np.asarray(sino, dtype=np.float32)
    w = sino.shape[1]
    p0, p180 = sino[0], sino[-1]
    at = lambda i: p180[i] if i < w else np.float32(0)
    maxd = w // 2
    scores = np.zeros(2 * maxd - 1, np.float32)
    for d in range(-maxd + 1, 0):
        idx, max_x = d + maxd - 1, w + d
        for x in range(max_x):
            diff = np.float32(p0[x] - at(max_x - x + 1))
            scores[idx] = np.float32(scores[idx] + diff * diff)
    for d in range(maxd):
        idx = d + maxd - 1
        for x in range(w - d):
            diff = np.float32(p0[x + d] - at(w - x + 1))
            scores[idx] = np.float32(scores[idx] + diff * diff)
    best = int(np.argmin(scores))  # first minimum, like minimum_index
    return (w + best - maxd + 1) / 2.0


@pytest.mark.parametrize("w,seed", [(16, 0), (33, 1), (64, 2), (101, 3)])
def test_center_vs_reference_loops(w, seed):
    sino = np.random.default_rng(seed).standard_normal((7, w)).astype(np.float32)
    assert _center(sino) == _center_brute(sino)


@pytest.mark.parametrize("axis", [30.0, 34.5, 40.0])
def test_center_of_mirrored_object(axis):
    """The last row is the first mirrored about the axis (projection at 180 degrees): the estimate
    lands within two pixels of it (the reference's mirror index is offset by two elements)."""
    w = 96
    x = np.arange(w) + 0.5
    obj = lambda u: np.exp(-((u - 20.0) / 4.0) ** 2) + 0.5 * np.exp(-((u - 55.0) / 6.0) ** 2)
    sino = np.stack([obj(x), obj(2 * axis - x)]).astype(np.float32)
    assert abs(_center(sino) - axis) <= 2.0


def test_center_property_is_read_only():
    t = U.Task("center-of-rotation")
    assert t.props.center == 0.0 and abs(t.props.angle_step - np.pi / 180) < 1e-12
    with pytest.raises(Exception):
        t.props.center = 3.0
    assert t.mode & U.MODE_CPU
