"""Learned sparse-view FBP on the GPU: the textured phantom kernels against the
numpy restatements (tests/phantoms.py), and what a briefly trained U-Net gains
over sparse-view FBP alone (the method without it)."""
import math

import numpy as np
import pytest

import learned
import phantoms
import ufomi

pytestmark = pytest.mark.gpu


def test_textured_sinogram_and_image_vs_numpy(ctx, torch_dev):
    import torch
    W, P, axis = 96, 40, 48.0
    ell, waves, R = learned.phantom_params(77, W)
    sino = torch.empty((2, P, W), device=torch_dev)
    ctx.sinogram_phantom(np.stack([ell, ell]), np.stack([waves, waves * [1, 1, 1, 2]]), R, W, P, axis,
                         math.pi / P, sino)
    ref = phantoms.sinogram(ell, W, P, axis, math.pi / P) + phantoms.texture_sinogram(waves, R, W, P, axis,
                                                                                      math.pi / P)
    np.testing.assert_allclose(sino[0].cpu().numpy(), ref, rtol=1e-5, atol=1e-4)
    ref2 = ref + phantoms.texture_sinogram(waves, R, W, P, axis, math.pi / P)
    np.testing.assert_allclose(sino[1].cpu().numpy(), ref2, rtol=1e-5, atol=1e-4)
    img = torch.empty((1, W, W), device=torch_dev)
    ctx.phantom_image(ell, waves, R, W, axis, img)
    ref_img = phantoms.image(ell, W, axis) + phantoms.texture_image(waves, R, W, axis)
    np.testing.assert_allclose(img[0].cpu().numpy(), ref_img, rtol=1e-5, atol=1e-5)


def test_trained_unet_beats_sparse_fbp(ctx, torch_dev):
    """30-view FBP of 128^2 textured phantoms vs 360-view FBP: 400 Adam steps on
    512 pairs, evaluated on 64 unseen phantoms inside the reconstruction disk."""
    import torch
    size, sparse, full = 128, 30, 360
    pm = learned.PairMaker(ctx, size, sparse, full, chunk=128)
    X = torch.empty((512, size, size), device=torch_dev)
    Y = torch.empty_like(X)
    pm.make(np.arange(512) + 10000, X, Y)
    Xt = torch.empty((64, size, size), device=torch_dev)
    Yt, Pt = torch.empty_like(Xt), torch.empty_like(Xt)
    pm.make(np.arange(64) + 90000, Xt, Yt, Pt)
    net = ufomi.Unet(ctx, 32, size, size, training=True, seed=1)
    losses = learned.train(net, X, Y, steps=400, batch=32, lr=1e-3, log_every=100)
    assert losses[-1][1] < losses[0][1]
    inf = ufomi.Unet(ctx, 64, size, size)
    inf.set_params(net.get_params())
    out = torch.empty_like(Xt)
    learned.infer(inf, Xt, out, 64)
    mask = learned.disk_mask(size)
    r_sparse, r_learned = learned.rmse(Xt, Yt, mask), learned.rmse(out, Yt, mask)
    print("RMSE to the 360-view FBP: sparse FBP %.4f, learned %.4f" % (r_sparse, r_learned))
    assert r_learned < 0.7 * r_sparse
    p_sparse, p_learned = learned.rmse(Xt, Pt, mask), learned.rmse(out, Pt, mask)
    print("RMSE to the phantom: sparse FBP %.4f, learned %.4f, 360-view FBP %.4f"
          % (p_sparse, p_learned, learned.rmse(Yt, Pt, mask)))
    assert p_learned < p_sparse
    net.close()
    inf.close()
