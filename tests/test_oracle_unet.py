"""The U-Net oracle (oracle/unet_oracle.py) pinned on the CPU: bf16 rounding
against torch's bfloat16 conversion, each layer against brute-force loops, and
every parameter gradient of the float64 network against central finite
differences."""
import numpy as np
import pytest

import unet_oracle as U


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(10000) * 10.0 ** rng.integers(-6, 6, 10000),
                        [0.0, -0.0, 1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 65504.0]]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(U.bf16(x), ref)
    assert U.bf16(np.float32(1.0 + 2.0 ** -8)) == 1.0        # tie to even
    assert U.bf16(np.float32(1.0 + 3 * 2.0 ** -8)) == 1.0 + 2.0 ** -6


def _brute_conv(x, w, b):
    B, H, W, _ = x.shape
    out = np.zeros((B, H, W, w.shape[0]))
    for n in range(B):
        for y in range(H):
            for xx in range(W):
                for ky in range(3):
                    for kx in range(3):
                        yy, xs = y + ky - 1, xx + kx - 1
                        if 0 <= yy < H and 0 <= xs < W:
                            out[n, y, xx] += w[:, ky, kx, :] @ x[n, yy, xs]
    return out + b


def test_conv3x3_brute_force():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 5, 6, 3))
    w = rng.standard_normal((4, 3, 3, 3))
    b = rng.standard_normal(4)
    np.testing.assert_allclose(U.conv3x3(x, w, b), _brute_conv(x, w, b), rtol=1e-12, atol=1e-12)


def test_conv3x3_gradients_are_adjoints():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((2, 6, 5, 3))
    w = rng.standard_normal((4, 3, 3, 3))
    g = rng.standard_normal((2, 6, 5, 4))
    # <conv(x, w), g> = <x, dgrad(g, w)> = <w, wgrad(g, x)>
    lhs = np.sum(U.conv3x3(x, w) * g)
    assert np.isclose(lhs, np.sum(x * U.conv3x3_dgrad(g, w)), rtol=1e-12)
    assert np.isclose(lhs, np.sum(w * U.conv3x3_wgrad(g, x)), rtol=1e-12)


def test_up2x2_brute_force_and_adjoint():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((2, 3, 4, 5))
    w = rng.standard_normal((4, 2, 1, 5))
    b = rng.standard_normal(2)
    out = U.up2x2(x, w, b)
    ref = np.zeros((2, 6, 8, 2))
    for n in range(2):
        for y in range(3):
            for xx in range(4):
                for a in range(4):
                    ref[n, 2 * y + a // 2, 2 * xx + a % 2] = w[a, :, 0, :] @ x[n, y, xx] + b
    np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-12)
    g = rng.standard_normal(out.shape)
    dx, dw, db = U.up2x2_grads(g, x, w)
    lin = np.sum((out - b) * g)
    assert np.isclose(lin, np.sum(x * dx), rtol=1e-12)
    assert np.isclose(lin, np.sum(w * dw), rtol=1e-12)
    np.testing.assert_allclose(db, g.sum(axis=(0, 1, 2)))


def test_maxpool_routes_to_first_maximum():
    x = np.array([[1.0, 3.0], [3.0, 2.0]]).reshape(1, 2, 2, 1)
    g = U.maxpool_grad(x, np.array([[[[5.0]]]]))
    assert g.reshape(-1).tolist() == [0.0, 5.0, 0.0, 0.0]
    assert U.maxpool(x).item() == 3.0


def test_network_gradients_finite_differences():
    """float64 network (no bf16 rounding) on a 16x16 image: sampled entries of
    every parameter tensor against central differences."""
    rng = np.random.default_rng(4)
    params = U.random_params(rng)
    img = rng.standard_normal((1, 16, 16))
    tgt = img + 0.3 * rng.standard_normal((1, 16, 16))
    net = U.UNet(params, bf16_storage=False)
    _, grads = net.gradients(img, tgt)
    eps = 1e-6
    for name, dims in U.layout():
        flat = params[name].reshape(-1)
        for idx in rng.choice(flat.size, size=min(2, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + eps
            lp = np.mean((U.UNet(params, False).forward(img)[0] - tgt) ** 2)
            flat[idx] = old - eps
            lm = np.mean((U.UNet(params, False).forward(img)[0] - tgt) ** 2)
            flat[idx] = old
            fd = (lp - lm) / (2 * eps)
            an = grads[name].reshape(-1)[idx]
            assert abs(fd - an) <= 1e-6 + 1e-4 * abs(fd), (name, idx, fd, an)


def test_layout_parameter_count():
    n = sum(int(np.prod(d)) for _, d in U.layout())
    assert 1_900_000 < n < 2_000_000
