"""Learned sparse-view FBP: training pairs, training loop and evaluation for the
U-Net of csrc/unet.hip.  No counterpart in the reference.

A pair is (FBP of `sparse` views, FBP of `full` views) of one seeded phantom:
10 random ellipses (the config-0 distribution) plus a band-limited texture of
plane waves inside the reconstruction disk, with exact analytic sinograms.
The sparse sinogram takes every (full / sparse)-th angle of the full one, so
both share the angle grid p * pi / full.  Everything runs on the GPU through
libufomi (phantom kernels, fused ramp filter, batched backprojection).
"""
import math

import numpy as np

import ufomi

N_ELLIPSES = 10
N_WAVES = 32
TEXTURE_RMS = 0.1
TEXTURE_PERIODS = (8.0, 64.0)  # px


def phantom_params(seed, size):
    """(ellipses [10][6], waves [32][4], radius) of phantom `seed` for a size^2 image."""
    radius = size / 2.0
    ell = ufomi.phantom_ellipses(seed, N_ELLIPSES, radius)
    waves = ufomi.phantom_texture(seed ^ 0x5EED, N_WAVES, TEXTURE_PERIODS[0], TEXTURE_PERIODS[1], TEXTURE_RMS)
    return ell, waves, radius


def fft_size(width):
    n = 1
    while n < 2 * width:
        n *= 2
    return n


class PairMaker:
    """Generates (sparse FBP, full FBP, phantom) batches of size^2 slices on the GPU."""

    def __init__(self, ctx, size, sparse, full, chunk=256):
        import torch
        if full % sparse:
            raise ValueError("full view count must be a multiple of the sparse one")
        self.ctx, self.size, self.sparse, self.full, self.chunk = ctx, size, sparse, full, chunk
        self.N = fft_size(size)
        self.coef = torch.from_numpy(ufomi.filter_coefficients(2 * self.N, "ramp-fromreal")).to("cuda")
        self.sino = torch.empty((chunk, full, size), device="cuda")
        self.filt = torch.empty_like(self.sino)
        self.sino_s = torch.empty((chunk, sparse, size), device="cuda")
        self.filt_s = torch.empty_like(self.sino_s)

    def fbp(self, sino, filt, n, views, out):
        """Ram-Lak FBP of n sinograms [views][size] into out [n][size][size]."""
        W = self.size
        self.ctx.ramp_filter(sino[:n], W, n * views, self.N, self.coef, W, filt[:n])
        self.ctx.bp_parallel(filt[:n], W, views, out[:n], n_slices=n, axis_pos=W / 2.0)

    def make(self, seeds, inputs, targets, phantoms=None):
        """Fill inputs / targets (and phantoms) [len(seeds)][size][size] float32."""
        W, P = self.size, self.full
        step = self.full // self.sparse
        for c0 in range(0, len(seeds), self.chunk):
            ss = seeds[c0:c0 + self.chunk]
            n = len(ss)
            pars = [phantom_params(int(s), W) for s in ss]
            ell = np.stack([p[0] for p in pars])
            wav = np.stack([p[1] for p in pars])
            radius = pars[0][2]
            self.ctx.sinogram_phantom(ell, wav, radius, W, P, W / 2.0, math.pi / P, self.sino[:n])
            self.fbp(self.sino, self.filt, n, P, targets[c0:c0 + n])
            self.sino_s[:n].copy_(self.sino[:n, ::step])
            self.fbp(self.sino_s, self.filt_s, n, self.sparse, inputs[c0:c0 + n])
            if phantoms is not None:
                self.ctx.phantom_image(ell, wav, radius, W, W / 2.0, phantoms[c0:c0 + n])


def disk_mask(size, device="cuda"):
    """Pixels whose centre lies inside the reconstruction disk (radius size / 2)."""
    import torch
    c = torch.arange(size, device=device, dtype=torch.float32) + 0.5 - size / 2.0
    return (c[None, :] ** 2 + c[:, None] ** 2) < (size / 2.0) ** 2


def rmse(a, b, mask):
    """Root mean squared difference inside the mask, over a batch [n][size][size]."""
    d = (a - b)[:, mask]
    return float((d.double() ** 2).mean().sqrt())


def train(net, inputs, targets, steps, batch, lr, seed=0, log=None, log_every=50):
    """Adam steps on random batches (drawn without replacement per epoch); returns the losses logged."""
    import torch
    n = inputs.shape[0]
    g = torch.Generator(device="cpu").manual_seed(seed)
    perm = torch.randperm(n, generator=g)
    pos = 0
    xb = torch.empty((batch,) + tuple(inputs.shape[1:]), device=inputs.device)
    yb = torch.empty_like(xb)
    losses = []
    for it in range(steps):
        if pos + batch > n:
            perm = torch.randperm(n, generator=g)
            pos = 0
        idx = perm[pos:pos + batch].to(inputs.device)
        pos += batch
        torch.index_select(inputs, 0, idx, out=xb)
        torch.index_select(targets, 0, idx, out=yb)
        want = log_every and (it % log_every == 0 or it == steps - 1)
        loss = net.train_step(xb, yb, lr, want_loss=bool(want))
        if want:
            losses.append((it, loss))
            if log:
                log("step %d loss %.6g" % (it, loss))
    return losses


def infer(net, inputs, out, batch):
    for b0 in range(0, inputs.shape[0], batch):
        n = min(batch, inputs.shape[0] - b0)
        net.forward(inputs[b0:b0 + n], out[b0:b0 + n], batch=n)
