"""CPU oracle of the learned sparse-view stage (the residual U-Net of
ufo-filters_amd/csrc/unet.hip).  Test infrastructure only.

There is no reference counterpart: ufo-filters has no learned stage.  The
network is restated here in numpy (float64 arithmetic) from its definition,
with an optional emulation of the device's bf16 storage points
(`bf16=True`): the input image, every stored activation and every stored
activation gradient are rounded to bf16 (round to nearest even) where the
device stores them, weights are used as their bf16 copies and biases as
float32.  Accumulation is float64 here and float32 (MFMA) on the device.

What pins it (tests/test_oracle_unet.py):
  * conv3x3 / transposed conv against brute-force loops on tiny inputs;
  * every parameter gradient against central finite differences of the
    float64 forward pass (bf16=False) on a tiny image;
  * the bf16 rounding against torch's float32 -> bfloat16 conversion.
"""
import numpy as np

LEVELS = 4
BASE = 32


def bf16(x):
    """Round to the nearest bf16 (ties to even), returned as float64."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def layout():
    """[(name, dims)] of the parameters in the device's order."""
    out = [("enc0.conv1.w", (BASE, 3, 3, 1)), ("enc0.conv1.b", (BASE, 1, 1, 1))]
    for lev in range(LEVELS):
        C = BASE << lev
        if lev > 0:
            out += [("enc%d.conv1.w" % lev, (C, 3, 3, C // 2)), ("enc%d.conv1.b" % lev, (C, 1, 1, 1))]
        out += [("enc%d.conv2.w" % lev, (C, 3, 3, C)), ("enc%d.conv2.b" % lev, (C, 1, 1, 1))]
    for lev in range(LEVELS - 2, -1, -1):
        C = BASE << lev
        out += [("up%d.w" % lev, (4, C, 1, 2 * C)), ("up%d.b" % lev, (C, 1, 1, 1)),
                ("dec%d.conv1.w" % lev, (C, 3, 3, 2 * C)), ("dec%d.conv1.b" % lev, (C, 1, 1, 1)),
                ("dec%d.conv2.w" % lev, (C, 3, 3, C)), ("dec%d.conv2.b" % lev, (C, 1, 1, 1))]
    out += [("final.w", (1, 1, 1, BASE)), ("final.b", (1, 1, 1, 1))]
    return out


def random_params(rng, bias_scale=0.05, final_scale=0.1):
    """He-normal weights, small random biases and a non-zero final layer."""
    p = {}
    for name, dims in layout():
        if name == "final.w":
            p[name] = rng.standard_normal(dims) * final_scale
        elif name.endswith(".w"):
            fan_in = dims[3] if name.startswith("up") else 9 * dims[3]
            p[name] = rng.standard_normal(dims) * np.sqrt(2.0 / fan_in)
        else:
            p[name] = rng.standard_normal(dims) * bias_scale
    return p


def _cols(x):
    """im2col of NHWC x for a 3x3 'same' window: [B][H][W][9][C], tap = 3 (dy+1) + (dx+1)."""
    b, h, w, c = x.shape
    p = np.zeros((b, h + 2, w + 2, c))
    p[:, 1:-1, 1:-1] = x
    return np.stack([p[:, dy:dy + h, dx:dx + w] for dy in range(3) for dx in range(3)], axis=3)


def conv3x3(x, w, bias=None, acc=np.float64):
    """x [B][H][W][Cin], w [Cout][3][3][Cin] -> [B][H][W][Cout] (zero padding).
    acc: accumulation type of the contraction (float64, or float32 like the device)."""
    co = w.shape[0]
    out = np.tensordot(_cols(x).astype(acc), w.reshape(co, 9, -1).astype(acc), axes=([3, 4], [1, 2]))
    out = out.astype(np.float64)
    return out if bias is None else out + bias


def conv3x3_dgrad(g, w, acc=np.float64):
    """Gradient w.r.t. x of conv3x3(x, w) for output gradient g."""
    wf = w[:, ::-1, ::-1, :].transpose(3, 1, 2, 0)  # [Cin][3][3][Cout], spatially flipped
    return conv3x3(g, wf, acc=acc)


def conv3x3_wgrad(g, x, acc=np.float64):
    """Gradient w.r.t. w [Cout][3][3][Cin]."""
    out = np.einsum("bhwo,bhwtc->otc", g.astype(acc), _cols(x).astype(acc))
    return out.astype(np.float64).reshape(g.shape[-1], 3, 3, x.shape[-1])


def up2x2(x, w, bias, acc=np.float64):
    """2x2 stride-2 transposed conv: x [B][h][w][Cin], w [4][Cout][1][Cin] -> [B][2h][2w][Cout]."""
    b, h, ww, _ = x.shape
    co = w.shape[1]
    y = np.einsum("bhwc,aoc->bhwao", x.astype(acc), w[:, :, 0, :].astype(acc)).astype(np.float64)  # a = 2 ky + kx
    out = y.reshape(b, h, ww, 2, 2, co).transpose(0, 1, 3, 2, 4, 5).reshape(b, 2 * h, 2 * ww, co)
    return out + bias


def up2x2_grads(g, x, w, acc=np.float64):
    """(dx, dw, db) of up2x2 for output gradient g."""
    b, h, ww, _ = x.shape
    co = w.shape[1]
    gr = g.reshape(b, h, 2, ww, 2, co).transpose(0, 1, 3, 2, 4, 5).reshape(b, h, ww, 4, co).astype(acc)
    dx = np.einsum("bhwao,aoc->bhwc", gr, w[:, :, 0, :].astype(acc)).astype(np.float64)
    dw = np.einsum("bhwao,bhwc->aoc", gr, x.astype(acc)).astype(np.float64)[:, :, None, :]
    return dx, dw, g.sum(axis=(0, 1, 2))


def maxpool(x):
    b, h, w, c = x.shape
    return x.reshape(b, h // 2, 2, w // 2, 2, c).max(axis=(2, 4))


def maxpool_grad(x, g):
    """Routes g to the first maximum of each 2x2 window in row-major order."""
    b, h, w, c = x.shape
    win = x.reshape(b, h // 2, 2, w // 2, 2, c).transpose(0, 1, 3, 2, 4, 5).reshape(b, h // 2, w // 2, 4, c)
    arg = np.argmax(win, axis=3)
    out = np.zeros_like(win)
    np.put_along_axis(out, arg[:, :, :, None, :], g[:, :, :, None, :], axis=3)
    return out.reshape(b, h // 2, w // 2, 2, 2, c).transpose(0, 1, 3, 2, 4, 5).reshape(b, h, w, c)


class UNet:
    """params: {name: array} with the device's tensor names and dims."""

    def __init__(self, params, bf16_storage=True, acc=np.float64):
        """acc: accumulation type of every contraction (float32 gives a second,
        equally valid evaluation; its distance to the float64 one measures how
        much bf16 rounding flips amplify through the network)."""
        self.acc = acc
        self.q = bf16 if bf16_storage else (lambda a: np.asarray(a, dtype=np.float64))
        self.p = {k: np.asarray(v, dtype=np.float64) for k, v in params.items()}
        # weights as stored on the device (bf16), biases float32
        f32 = (lambda a: np.float32(a).astype(np.float64)) if bf16_storage else (lambda a: a)
        self.w = {k: (self.q(v) if k.endswith(".w") else f32(v).reshape(-1)) for k, v in self.p.items()}

    def forward(self, img):
        q, w, acc = self.q, self.w, self.acc
        c = {"img": img}
        x0 = q(img)[..., None]
        c["x0"] = x0
        h = q(np.maximum(conv3x3(x0, w["enc0.conv1.w"], w["enc0.conv1.b"], acc), 0))
        skips = []
        for lev in range(LEVELS):
            if lev > 0:
                c["pool%d" % (lev - 1)] = h
                h = q(np.maximum(conv3x3(h, w["enc%d.conv1.w" % lev], w["enc%d.conv1.b" % lev], acc), 0))
            c["e1_%d" % lev] = h
            h = q(np.maximum(conv3x3(h, w["enc%d.conv2.w" % lev], w["enc%d.conv2.b" % lev], acc), 0))
            c["e2_%d" % lev] = h
            if lev < LEVELS - 1:
                skips.append(h)
                h = maxpool(h)
        for lev in range(LEVELS - 2, -1, -1):
            c["src%d" % lev] = h
            u = q(up2x2(h, w["up%d.w" % lev], w["up%d.b" % lev], acc))
            cat = np.concatenate([u, skips[lev]], axis=-1)
            c["cat%d" % lev] = cat
            h = q(np.maximum(conv3x3(cat, w["dec%d.conv1.w" % lev], w["dec%d.conv1.b" % lev], acc), 0))
            c["d1_%d" % lev] = h
            h = q(np.maximum(conv3x3(h, w["dec%d.conv2.w" % lev], w["dec%d.conv2.b" % lev], acc), 0))
            c["d2_%d" % lev] = h
        out = img + (h @ w["final.w"].reshape(-1) + w["final.b"].reshape(-1)[0])
        return out, c

    def gradients(self, img, target):
        """(loss, {name: gradient}) of the mean squared error."""
        q, w, acc = self.q, self.w, self.acc
        out, c = self.forward(img)
        n = out.size
        loss = float(np.mean((out - target) ** 2))
        g_out = 2.0 * (out - target) / n
        gr = {}
        d2 = c["d2_0"]
        gr["final.w"] = np.einsum("bhw,bhwc->c", g_out, d2).reshape(self.p["final.w"].shape)
        gr["final.b"] = np.array([g_out.sum()]).reshape(self.p["final.b"].shape)
        g = q(np.where(d2 > 0, g_out[..., None] * w["final.w"].reshape(-1), 0.0))
        g_skip = {}
        for lev in range(LEVELS - 1):
            # d2 = relu(conv(d1)); g is the pre-activation gradient of d2
            d1, cat = c["d1_%d" % lev], c["cat%d" % lev]
            gr["dec%d.conv2.w" % lev] = conv3x3_wgrad(g, d1, acc)
            gr["dec%d.conv2.b" % lev] = g.sum(axis=(0, 1, 2))
            g = q(np.where(d1 > 0, conv3x3_dgrad(g, w["dec%d.conv2.w" % lev], acc), 0.0))
            gr["dec%d.conv1.w" % lev] = conv3x3_wgrad(g, cat, acc)
            gr["dec%d.conv1.b" % lev] = g.sum(axis=(0, 1, 2))
            gcat = q(conv3x3_dgrad(g, w["dec%d.conv1.w" % lev], acc))
            C = cat.shape[-1] // 2
            g_skip[lev] = gcat[..., C:]
            src = c["src%d" % lev]
            dx, dw, db = up2x2_grads(gcat[..., :C], src, w["up%d.w" % lev], acc)
            gr["up%d.w" % lev] = dw
            gr["up%d.b" % lev] = db
            g = q(np.where(src > 0, dx, 0.0))  # pre-activation gradient of src
        for lev in range(LEVELS - 1, -1, -1):
            if lev < LEVELS - 1:
                e2 = c["e2_%d" % lev]
                g = q(np.where(e2 > 0, g_skip[lev] + maxpool_grad(e2, g_pool), 0.0))
            e1 = c["e1_%d" % lev]
            gr["enc%d.conv2.w" % lev] = conv3x3_wgrad(g, e1, acc)
            gr["enc%d.conv2.b" % lev] = g.sum(axis=(0, 1, 2))
            g = q(np.where(e1 > 0, conv3x3_dgrad(g, w["enc%d.conv2.w" % lev], acc), 0.0))
            if lev > 0:
                x = c["pool%d" % (lev - 1)]
                gr["enc%d.conv1.w" % lev] = conv3x3_wgrad(g, x, acc)
                gr["enc%d.conv1.b" % lev] = g.sum(axis=(0, 1, 2))
                g_pool = q(conv3x3_dgrad(g, w["enc%d.conv1.w" % lev], acc))
                g = g_pool
            else:
                gr["enc0.conv1.w"] = conv3x3_wgrad(g, c["x0"], acc)
                gr["enc0.conv1.b"] = g.sum(axis=(0, 1, 2))
        for k in gr:
            gr[k] = np.asarray(gr[k]).reshape(self.p[k].shape)
        return loss, gr
