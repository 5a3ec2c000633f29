"""GPU parity of batched uneven FBP (ufomi_fbp_batch_*): segmented ramp filter +
one persistent backprojection launch, replayed as a HIP graph.

Bar: every slice within 1e-5 relative L2 of the FP64 oracle (FP64
fft-filter-ifft + FP64 backprojection) and of the float32-semantics oracle, in
every tile / angle-chunk configuration; bit-identical to the single-request
path (ufomi_ramp_filter + ufomi_bp_parallel) with the default configuration;
f16 / bf16 storage bounded by the storage rounding (1e-3, 8e-3).
"""
import math

import numpy as np
import pytest

import oracle
import ufomi
from conftest import rel_l2

pytestmark = pytest.mark.gpu

CONFIGS = [(32, 8), (32, 16), (32, 32), (64, 8), (64, 16)]

# (W, P, extra properties): uneven widths, angle counts, ROIs, axes and angle tables
SPECS = [
    (64, 48, {}),
    (100, 73, {"roi_x": 10, "roi_y": 3, "roi_width": 70, "roi_height": 45}),
    (33, 20, {"axis_pos": 15.25}),
    (128, 90, {"angle_offset": 0.3, "angle_step": math.pi / 75}),
    (80, 31, {"num_projections": 40, "offset": 5, "roi_width": 97, "roi_height": 13}),
    (200, 17, {"axis_pos": 103.5, "roi_x": 40, "roi_width": 130}),
]


def _sino(seed, p, w):
    return np.random.default_rng(seed).standard_normal((p, w)).astype(np.float32)


def _roi(w, props):
    return props.get("roi_height", 0) or w, props.get("roi_width", 0) or w


def _make(torch_dev, specs=SPECS, seed0=100):
    import torch
    reqs, host = [], []
    for i, (w, p, props) in enumerate(specs):
        s = _sino(seed0 + i, p, w)
        rh, rw = _roi(w, props)
        reqs.append(dict(sinogram=torch.from_numpy(s).to(torch_dev),
                         slice=torch.full((rh, rw), float("nan"), device=torch_dev), **props))
        host.append(s)
    return reqs, host


def _oracle(s, props, filt="ramp-fromreal"):
    p, w = s.shape
    n = 16
    while n < 2 * w:
        n *= 2
    coef = oracle.filter_coefficients(2 * n, filt)
    f64 = oracle.ramp_chain(s, n, coef, w)
    kw = {k: v for k, v in props.items()}
    ref64 = oracle.bp_texture_f64(f64, **kw)
    ref32 = oracle.bp_texture_ref32(f64.astype(np.float32), **kw)
    return ref64, ref32


def _slices(reqs):
    import torch
    torch.cuda.synchronize()
    return [r["slice"].cpu().numpy() for r in reqs]


@pytest.mark.parametrize("tile,chunk", CONFIGS)
def test_batch_vs_oracle(ctx, torch_dev, tile, chunk):
    reqs, host = _make(torch_dev)
    b = ufomi.FbpBatch(ctx, reqs)
    b.configure(tile, chunk)
    info = b.info()
    assert (info["tile"], info["angle_chunk"]) == (tile, chunk)
    b.run()
    for got, s, (w, p, props) in zip(_slices(reqs), host, SPECS):
        ref64, ref32 = _oracle(s, props)
        assert got.shape == ref64.shape
        assert rel_l2(got, ref64) < 1e-5, (w, p, props)
        assert rel_l2(got, ref32) < 1e-5, (w, p, props)
    b.close()


def test_batch_bitwise_single_request_path(ctx, torch_dev):
    import torch
    reqs, host = _make(torch_dev)
    b = ufomi.FbpBatch(ctx, reqs)
    assert (b.info()["tile"], b.info()["angle_chunk"]) == (32, 16)
    b.run()
    got = _slices(reqs)
    for g, r in zip(got, reqs):
        s = r["sinogram"]
        p, w = s.shape
        n = 16
        while n < 2 * w:
            n *= 2
        coef = torch.from_numpy(ufomi.filter_coefficients(2 * n)).to(torch_dev)
        filt = torch.empty_like(s)
        ctx.ramp_filter(s, w, p, n, coef, w, filt)
        out = torch.empty_like(r["slice"])
        props = {k: v for k, v in r.items() if k not in ("sinogram", "slice")}
        ctx.bp_parallel(filt, w, p, out, **props)
        torch.cuda.synchronize()
        assert np.array_equal(g, out.cpu().numpy())
    b.close()


def test_graph_replays_match_eager(ctx, torch_dev):
    """Replays with eager work and synchronisations between them, each checked
    against the directly launched kernels."""
    import torch
    reqs, _ = _make(torch_dev, seed0=300)
    b = ufomi.FbpBatch(ctx, reqs)
    b.configure(64, 8)
    b.launch(3)
    eager = _slices(reqs)
    for rep in range(4):
        for r in reqs:
            r["slice"].fill_(float("nan"))
        junk = torch.randn(1 << 20, device=torch_dev)
        (junk * 2).sum().item()
        b.run()
        for g, e in zip(_slices(reqs), eager):
            assert np.array_equal(g, e), rep
    b.close()


def test_new_inputs_between_replays(ctx, torch_dev):
    """The graph reads the request buffers at replay time."""
    reqs, host = _make(torch_dev, specs=SPECS[:3], seed0=400)
    b = ufomi.FbpBatch(ctx, reqs)
    b.run()
    first = _slices(reqs)
    for r in reqs:
        r["sinogram"].mul_(-2.0)
    b.run()
    for g, f in zip(_slices(reqs), first):
        np.testing.assert_allclose(g, -2.0 * f, rtol=1e-5, atol=1e-6 * np.abs(f).max())
    b.close()


def test_autotune(ctx, torch_dev):
    reqs, host = _make(torch_dev, seed0=500)
    b = ufomi.FbpBatch(ctx, reqs)
    t, a, ms = b.autotune(reps=2, force=True)
    assert (t, a) in CONFIGS and ms > 0
    assert (b.info()["tile"], b.info()["angle_chunk"]) == (t, a)
    # A second batch of the same shapes takes the cached choice without timing.
    b2 = ufomi.FbpBatch(ctx, reqs)
    assert b2.autotune(reps=2) == (t, a, 0.0)
    b2.run()
    for got, s, (w, p, props) in zip(_slices(reqs), host, SPECS):
        ref64, _ = _oracle(s, props)
        assert rel_l2(got, ref64) < 1e-5
    b.close()
    b2.close()


@pytest.mark.parametrize("storage,tol", [(ufomi.F16, 1e-3), (ufomi.BF16, 8e-3)])
def test_low_precision_storage(ctx, torch_dev, storage, tol):
    reqs, host = _make(torch_dev, seed0=600)
    b = ufomi.FbpBatch(ctx, reqs, storage=storage)
    b.run()
    for got, s, (w, p, props) in zip(_slices(reqs), host, SPECS):
        ref64, _ = _oracle(s, props)
        assert rel_l2(got, ref64) < tol
    b.close()


def test_shepp_logan_filter(ctx, torch_dev):
    reqs, host = _make(torch_dev, specs=SPECS[:4], seed0=700)
    b = ufomi.FbpBatch(ctx, reqs, filter="shepp-logan")
    b.run()
    for got, s, (w, p, props) in zip(_slices(reqs), host, SPECS):
        ref64, _ = _oracle(s, props, "shepp-logan")
        assert rel_l2(got, ref64) < 1e-5
    b.close()


def test_many_requests_fill_the_chip(ctx, torch_dev):
    """More work items than resident workgroups, mixed FFT sizes."""
    rng = np.random.default_rng(9)
    specs = []
    for i in range(60):
        w = int(rng.integers(40, 500))
        p = int(rng.choice([12, 30, 45]))
        rw = int(rng.integers(w // 2, w + 1))
        specs.append((w, p, {"roi_width": rw, "roi_height": int(rng.integers(w // 2, w + 1)),
                             "roi_x": (w - rw) // 2}))
    reqs, host = _make(torch_dev, specs=specs, seed0=800)
    b = ufomi.FbpBatch(ctx, reqs)
    info = b.info()
    assert info["items"] > info["workgroups"] and info["filter_launches"] >= 3
    b.run()
    for got, s, (w, p, props) in zip(_slices(reqs), host, specs):
        ref64, _ = _oracle(s, props)
        assert rel_l2(got, ref64) < 1e-5
    b.close()


def test_request_errors(ctx, torch_dev):
    import torch
    s = torch.zeros((10, 16), device=torch_dev)
    out = torch.zeros((16, 16), device=torch_dev)
    ok = dict(sinogram=s, slice=out)
    with pytest.raises(ufomi.UfomiError) as e:
        ufomi.FbpBatch(ctx, [ok, dict(sinogram=s, slice=out, num_projections=5)])
    assert e.value.code == -5 and "request 1" in str(e.value)
    assert "must be greater than or equal to sinogram height" in str(e.value)
    with pytest.raises(ufomi.UfomiError) as e:
        ufomi.FbpBatch(ctx, [dict(sinogram=s, slice=out, fft_size=48)])
    assert e.value.code == -3
    with pytest.raises(ufomi.UfomiError) as e:
        ufomi.FbpBatch(ctx, [dict(sinogram=s, slice=out, fft_size=8)])
    assert e.value.code == -6
    with pytest.raises(ufomi.UfomiError) as e:
        ufomi.FbpBatch(ctx, [dict(sinogram=s, slice=out, mode="nearest")])
    assert e.value.code == -4
    b = ufomi.FbpBatch(ctx, [ok])
    with pytest.raises(ufomi.UfomiError) as e:
        b.configure(64, 32)
    assert e.value.code == -1
    with pytest.raises(ufomi.UfomiError):
        b.launch(4)
    b.close()
