"""GPU parity of the learned sparse-view stage (ufomi_unet_*, ufomi_nn_conv3x3)
against the CPU oracle (oracle/unet_oracle.py, float64 with the device's bf16
storage points emulated).

Bars: a single conv within one bf16 rounding step (rel-L2 < 4e-3, element-wise
<= 2 bf16 ulps of the result).  Through the whole network, bf16 rounding flips
(an f32 sum landing on the other side of a bf16 rounding boundary than the f64
sum) compound layer after layer, so the oracle is evaluated twice, with f64 and
with f32 accumulation; the distance between those two correct evaluations is
the floor, and the device must stay within 3x that floor (+2e-3) of the f64
oracle: for the output's learned part (out - img) and for every parameter
gradient tensor; the loss within 5x its floor + 1e-4 relative.
"""
import numpy as np
import pytest

import unet_oracle as U
import ufomi
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _bf16_tensor(a, torch_dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(torch.bfloat16).to(torch_dev)


@pytest.mark.parametrize("cin,cout,H,W", [(32, 32, 16, 32), (64, 64, 16, 32), (32, 128, 16, 32), (128, 64, 16, 32),
                                          (32, 32, 8, 128), (64, 64, 4, 64), (32, 128, 16, 64), (256, 64, 8, 64),
                                          (64, 48, 16, 64), (32, 8, 8, 64), (64, 192, 8, 64), (32, 64, 64, 64),
                                          (32, 128, 64, 64)])
def test_conv3x3_mfma_vs_oracle(ctx, torch_dev, cin, cout, H, W):
    """W >= 64 runs the LDS-staged kernels (8-row tiles from 8 rows on, 4-row below), narrower
    images the generic GEMM.  cout 48 and 8 leave channel tiles partly past cout; cout 192 is three
    channel tiles of one pixel tile; 64 x 64 images of two are 16 pixel tiles, which the grid
    spreads over the XCDs in contiguous ranges (with 128 output channels, two channel tiles each)."""
    import torch
    rng = np.random.default_rng(cin + cout)
    B = 2
    x = U.bf16(rng.standard_normal((B, H, W, cin)))
    w = U.bf16(rng.standard_normal((cout, 3, 3, cin)) / np.sqrt(9 * cin))
    b = rng.standard_normal(cout).astype(np.float32)
    out = torch.empty((B, H, W, cout), dtype=torch.bfloat16, device=torch_dev)
    ctx.nn_conv3x3(_bf16_tensor(x, torch_dev), B, H, W, cin, _bf16_tensor(w, torch_dev),
                   torch.from_numpy(b).to(torch_dev), True, cout, out)
    got = out.float().cpu().numpy()
    ref = np.maximum(U.conv3x3(x, w, b.astype(np.float64)), 0)
    assert rel_l2(got, ref) < 4e-3
    assert np.all(np.abs(got - ref) <= 2 * 2.0 ** -7 * np.abs(ref) + 1e-6)


def test_layout_matches_oracle(ctx):
    net = ufomi.Unet(ctx, 1, 64, 64)
    assert [(n, d) for n, _, d in net.tensors] == U.layout()
    assert net.count == sum(int(np.prod(d)) for _, d in U.layout())
    flat = net.get_params()
    p = net.unpack(flat)
    assert np.all(p["final.w"] == 0) and np.all(p["enc1.conv1.b"] == 0)
    assert 0.5 < np.std(p["enc2.conv2.w"]) * np.sqrt(9 * 128 / 2.0) < 1.5   # He-normal
    net.close()


def _random_net(ctx, seed, B, H, W, training):
    net = ufomi.Unet(ctx, B, H, W, training=training)
    params = U.random_params(np.random.default_rng(seed))
    flat = np.zeros(net.padded, dtype=np.float32)
    for name, off, dims in net.tensors:
        flat[off:off + int(np.prod(dims))] = params[name].reshape(-1)
    net.set_params(flat)
    return net, params


# (1, 32, 256): levels 0-2 are at least 64 pixels wide and 8 rows high, so the 8-row LDS conv
# kernels run there, with the fused 2x2 pool on 32 and 64 channels
SHAPES = [(2, 64, 64), (1, 32, 256)]


@pytest.mark.parametrize("B,H,W", SHAPES)
def test_forward_vs_oracle(ctx, torch_dev, B, H, W):
    import torch
    net, params = _random_net(ctx, 5, B, H, W, False)
    img = np.random.default_rng(6).standard_normal((B, H, W)).astype(np.float32)
    d_in = torch.from_numpy(img).to(torch_dev)
    d_out = torch.empty_like(d_in)
    net.forward(d_in, d_out)
    got = d_out.cpu().numpy()
    ref, _ = U.UNet(params).forward(img.astype(np.float64))
    r32, _ = U.UNet(params, acc=np.float32).forward(img.astype(np.float64))
    floor = rel_l2(r32 - img, ref - img)
    assert floor < 2e-2
    assert rel_l2(got - img, ref - img) < 3 * floor + 2e-3
    net.close()


@pytest.mark.parametrize("B,H,W", SHAPES)
def test_gradients_vs_oracle(ctx, torch_dev, B, H, W):
    import torch
    net, params = _random_net(ctx, 7, B, H, W, True)
    rng = np.random.default_rng(8)
    img = rng.standard_normal((B, H, W)).astype(np.float32)
    tgt = (img + 0.5 * rng.standard_normal((B, H, W))).astype(np.float32)
    loss, g = net.gradients(torch.from_numpy(img).to(torch_dev), torch.from_numpy(tgt).to(torch_dev))
    ref_loss, ref = U.UNet(params).gradients(img.astype(np.float64), tgt.astype(np.float64))
    loss32, r32 = U.UNet(params, acc=np.float32).gradients(img.astype(np.float64), tgt.astype(np.float64))
    assert abs(loss - ref_loss) <= 5 * abs(loss32 - ref_loss) + 1e-4 * ref_loss
    got = net.unpack(g)
    # The float32-accumulation oracle measures how far float32 sums and the bf16 rounding flips
    # they cause carry the gradients from the FP64 ones through the 20 layers. Its float32 sums are
    # blocked BLAS products, the device's are long sequential MFMA chains, so its distance can
    # understate the device's: where it is small (0.3% at 1 x 32 x 256, against 2% at 2 x 64 x 64)
    # the bar keeps 3% of room, still below the 6% the 64 x 64 case allows. (The 4-row and 8-row
    # LDS conv kernels give the same gradients to three digits here.)
    for name, _ in U.layout():
        floor = rel_l2(r32[name], ref[name])
        print("%s: device %.3g, float32-accumulation floor %.3g" % (name, rel_l2(got[name], ref[name]), floor))
    for name, _ in U.layout():
        floor = rel_l2(r32[name], ref[name])
        assert floor < 5e-2, name
        assert rel_l2(got[name], ref[name]) < max(3 * floor + 2e-3, 3e-2), (name, floor)
    net.close()


def test_set_get_roundtrip_and_errors(ctx, torch_dev):
    import torch
    net = ufomi.Unet(ctx, 2, 64, 64)
    flat = np.random.default_rng(9).standard_normal(net.padded).astype(np.float32)
    net.set_params(flat)
    assert np.array_equal(net.get_params(), flat)
    x = torch.zeros((3, 64, 64), device=torch_dev)
    with pytest.raises(ufomi.UfomiError) as e:
        net.forward(x, x)  # batch 3 > max_batch 2
    assert e.value.code == -6
    with pytest.raises(ufomi.UfomiError) as e:
        net.gradients(x[:2], x[:2])
    assert e.value.code == -4  # created without training state
    with pytest.raises(ufomi.UfomiError):
        ufomi.Unet(ctx, 1, 48, 64)
    with pytest.raises(ufomi.UfomiError):
        ufomi.Unet(ctx, 1, 32, 32, training=True)
    net.close()


def test_adam_steps_reduce_loss(ctx, torch_dev):
    import torch
    B, H, W = 4, 64, 64
    net = ufomi.Unet(ctx, B, H, W, training=True, seed=3)
    rng = np.random.default_rng(10)
    img = rng.standard_normal((B, H, W)).astype(np.float32)
    tgt = (0.5 * img).astype(np.float32)
    d_in, d_t = torch.from_numpy(img).to(torch_dev), torch.from_numpy(tgt).to(torch_dev)
    losses = [net.train_step(d_in, d_t, 1e-3) for _ in range(30)]
    assert losses[0] == pytest.approx(np.mean((img - tgt) ** 2), rel=1e-4)  # identity at start
    assert losses[-1] < 0.8 * losses[0]
    net.close()
