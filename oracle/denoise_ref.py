"""CPU oracle of the learned denoiser (test infrastructure only).

The denoiser has no counterpart in the reference: gartus returns plain Monte
Carlo averages (RayCamera::render_world*_hdr_image and
render_world*_denoising_aovs, src/graphics/camera.rs:4175-4336,3064-3116),
and those are this network's inputs. This module restates the network and
its gradients in float64 numpy, for checking the GPU kernels:

  input    x0 = [log1p(max(beauty, 0)), albedo, normal, 0 x 7]   (16 channels),
           edge-replicated to a multiple of 2^(levels-1)
  encoder  level l: e_l = relu(conv(relu(conv(in_l)))), in_0 = x0, in_l = maxpool2(e_{l-1})
  decoder  level l (from levels-2 down to 0):
           d_l = relu(conv(relu(conv(concat(up2(d_{l+1} or e_{L-1}), e_l)))))
  output   z = conv(d_0) (16 channels, 3 used); radiance = expm1(max(z, 0))
  loss     mean over valid pixels and 3 channels of |T(z) - t / (1 + t)|,
           T(z) = 1 - exp(-z) for z >= 0, z below
  conv     3x3, zero padding 1, weights [cout][9][cin] (tap 3 (dy + 1) + (dx + 1)), bias [cout]
  maxpool  2x2; the gradient goes to the window's first maximum in scan order

Parameter order matches gt_denoiser_get_params: encoder convolutions
(level 0..L-1, two each), decoder convolutions (level L-2..0, two each), the
output convolution; each as weights then bias.

With bf16=True, every tensor the GPU stores in bf16 is rounded to bf16
(round to nearest even) at the same point: input, weights, each layer's
output, pooled and concatenated tensors, the loss gradient and each
activation gradient. Sums stay float64.

Pins (tests/test_denoise_oracle.py): conv against a brute-force loop, the
whole gradient against central finite differences, pooling routing, and the
loss against its closed form.
"""
from __future__ import annotations

import numpy as np

IN_CH, OUT_CH = 16, 16


def bf16(a: np.ndarray) -> np.ndarray:
    """Round float values to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    nan = np.isnan(f)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    out = (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32).astype(np.float64)
    out[nan] = np.nan
    return out


class Net:
    def __init__(self, levels: int, channels, params: np.ndarray | None = None):
        self.L = levels
        self.ch = list(channels)[:levels]
        self.convs = []  # (cin, cout, relu)
        for l in range(levels):
            self.convs.append((IN_CH if l == 0 else self.ch[l - 1], self.ch[l], True))
            self.convs.append((self.ch[l], self.ch[l], True))
        for l in range(levels - 2, -1, -1):
            self.convs.append((self.ch[l + 1] + self.ch[l], self.ch[l], True))
            self.convs.append((self.ch[l], self.ch[l], True))
        self.convs.append((self.ch[0], OUT_CH, False))
        self.offsets = []
        off = 0
        for cin, cout, _ in self.convs:
            self.offsets.append((off, off + cout * 9 * cin))
            off += cout * 9 * cin + cout
        self.nparams = off
        self.params = np.zeros(off) if params is None else np.asarray(params, np.float64).copy()

    def wb(self, i, params=None):
        p = self.params if params is None else params
        cin, cout, _ = self.convs[i]
        w0, b0 = self.offsets[i]
        return p[w0:b0].reshape(cout, 9, cin), p[b0:b0 + cout]

    def enc(self, l, k):
        return 2 * l + k

    def dec(self, l, k):
        return 2 * self.L + 2 * (self.L - 2 - l) + k


# ---- layers -------------------------------------------------------------------

def _shift(x, dy, dx):
    """x[n, h + dy, w + dx, :] with zeros outside."""
    out = np.zeros_like(x)
    H, W = x.shape[1:3]
    hs, he = max(0, -dy), min(H, H - dy)
    ws, we = max(0, -dx), min(W, W - dx)
    out[:, hs:he, ws:we] = x[:, hs + dy:he + dy, ws + dx:we + dx]
    return out


def conv(x, w, b):
    """y[p, co] = b[co] + sum_t sum_ci x[p + off_t, ci] w[co, t, ci]."""
    y = np.zeros(x.shape[:3] + (w.shape[0],))
    for t in range(9):
        dy, dx = t // 3 - 1, t % 3 - 1
        y += _shift(x, dy, dx) @ w[:, t, :].T
    return y + b


def conv_backward(x, w, g):
    """Given dL/dy = g: (dL/dx, dL/dw, dL/db)."""
    dw = np.zeros_like(w)
    dx = np.zeros_like(x)
    for t in range(9):
        dy, dxo = t // 3 - 1, t % 3 - 1
        xs = _shift(x, dy, dxo)
        dw[:, t, :] = np.einsum("nhwo,nhwi->oi", g, xs)
        dx += _shift(g @ w[:, t, :], -dy, -dxo)
    return dx, dw, g.sum(axis=(0, 1, 2))


def pool(x):
    n, H, W, c = x.shape
    v = x.reshape(n, H // 2, 2, W // 2, 2, c)
    return v.max(axis=(2, 4))


def pool_backward(x, g):
    """Routes g to the first maximum of each 2x2 window in scan order."""
    n, H, W, c = x.shape
    v = x.reshape(n, H // 2, 2, W // 2, 2, c).transpose(0, 1, 3, 2, 4, 5).reshape(n, H // 2, W // 2, 4, c)
    arg = v.argmax(axis=3)  # first maximum
    out = np.zeros_like(v)
    np.put_along_axis(out, arg[:, :, :, None, :], g[:, :, :, None, :], axis=3)
    return out.reshape(n, H // 2, W // 2, 2, 2, c).transpose(0, 1, 3, 2, 4, 5).reshape(n, H, W, c)


def up2(x):
    return x.repeat(2, axis=1).repeat(2, axis=2)


def up2_backward(g):
    n, H, W, c = g.shape
    return g.reshape(n, H // 2, 2, W // 2, 2, c).sum(axis=(2, 4))


HALVES_CLIP = 2.0


def pack_input(beauty, albedo, normal, multiple, beauty_b=None):
    """With beauty_b (two half-sample renders a = beauty, b): log1p of their
    mean in channels 0-2 and log1p(a) - log1p(b), clipped to +-2, in
    channels 9-11."""
    n, h, w, _ = beauty.shape
    hp, wp = -(-h // multiple) * multiple, -(-w // multiple) * multiple
    x = np.zeros((n, hp, wp, IN_CH))
    yy = np.minimum(np.arange(hp), h - 1)
    xx = np.minimum(np.arange(wp), w - 1)
    sel = (slice(None), yy[:, None], xx[None, :])
    if beauty_b is None:
        x[..., 0:3] = np.log1p(np.maximum(beauty, 0.0))[sel]
    else:
        x[..., 0:3] = np.log1p(np.maximum(0.5 * (beauty + beauty_b), 0.0))[sel]
        d = np.log1p(np.maximum(beauty, 0.0)) - np.log1p(np.maximum(beauty_b, 0.0))
        x[..., 9:12] = np.clip(d, -HALVES_CLIP, HALVES_CLIP)[sel]
    x[..., 3:6] = albedo[sel]
    x[..., 6:9] = normal[sel]
    return x


def tone(z):
    return np.where(z >= 0.0, 1.0 - np.exp(-np.maximum(z, 0.0)), z)


def dtone(z):
    return np.where(z >= 0.0, np.exp(-np.maximum(z, 0.0)), 1.0)


# ---- network ------------------------------------------------------------------

def forward(net: Net, beauty, albedo, normal, params=None, bf=False, beauty_b=None):
    """Returns (z, cache)."""
    r = bf16 if bf else (lambda a: a)
    x0 = r(pack_input(beauty, albedo, normal, 1 << (net.L - 1), beauty_b))
    W = {}
    for i in range(len(net.convs)):
        w, b = net.wb(i, params)
        W[i] = (r(w), b)

    def c(i, x):
        w, b = W[i]
        y = conv(x, w, b)
        return r(np.maximum(y, 0.0)) if net.convs[i][2] else y

    cache = {"x0": x0, "W": W, "in": {0: x0}, "ea": {}, "e": {}, "cat": {}, "da": {}, "d": {}}
    for l in range(net.L):
        inp = cache["in"][l]
        cache["ea"][l] = c(net.enc(l, 0), inp)
        cache["e"][l] = c(net.enc(l, 1), cache["ea"][l])
        if l + 1 < net.L:
            cache["in"][l + 1] = r(pool(cache["e"][l]))
    for l in range(net.L - 2, -1, -1):
        upsrc = cache["e"][net.L - 1] if l == net.L - 2 else cache["d"][l + 1]
        cache["cat"][l] = np.concatenate([up2(upsrc), cache["e"][l]], axis=-1)
        cache["da"][l] = c(net.dec(l, 0), cache["cat"][l])
        cache["d"][l] = c(net.dec(l, 1), cache["da"][l])
    z = c(len(net.convs) - 1, cache["d"][0])
    return z, cache


LOSS_REL_L2, LOSS_L1_TONEMAP = 0, 1
REL_L2_EPS = 0.01


def rel_l2_elem(zc, t):
    """Relative L2 per element, (y - t)^2 / (max(y, 0)^2 + eps) with
    y = expm1(z) (z below 0), and its derivative in z with the denominator
    held constant."""
    pos = zc >= 0.0
    y = np.where(pos, np.expm1(np.maximum(zc, 0.0)), zc)
    dy = np.where(pos, np.exp(np.maximum(zc, 0.0)), 1.0)
    inv = 1.0 / (np.maximum(y, 0.0) ** 2 + REL_L2_EPS)
    d = y - t
    return d * d * inv, 2.0 * d * inv * dy


def loss_and_dz(z, target, bf=False, kind=LOSS_REL_L2):
    n, h, w, _ = target.shape
    zc = z[:, :h, :w, :3]
    t = np.maximum(target, 0.0)
    count = n * h * w * 3
    dz = np.zeros_like(z)
    if kind == LOSS_REL_L2:
        e, g = rel_l2_elem(zc, t)
        loss = e.sum() / count
        dz[:, :h, :w, :3] = g / count
    else:
        d = tone(zc) - t / (1.0 + t)
        loss = np.abs(d).sum() / count
        dz[:, :h, :w, :3] = np.sign(d) * dtone(zc) / count
    return loss, (bf16(dz) if bf else dz)


def backward(net: Net, cache, dz, bf=False):
    """Gradient of every parameter (parameter order) given dL/dz."""
    r = bf16 if bf else (lambda a: a)
    grads = np.zeros(net.nparams)
    W = cache["W"]

    def put(i, dw, db):
        w0, b0 = net.offsets[i]
        grads[w0:b0] = dw.ravel()
        grads[b0:b0 + len(db)] = db

    def cb(i, x, y, g):
        """Backward through conv i (input x, output y after ReLU)."""
        if net.convs[i][2]:
            g = g * (y > 0.0)
        dx, dw, db = conv_backward(x, W[i][0], g)
        put(i, dw, db)
        return r(dx)

    last = len(net.convs) - 1
    g_d = {0: cb(last, cache["d"][0], None, dz)}
    g_cat = {}
    for l in range(0, net.L - 1):
        g_da = cb(net.dec(l, 1), cache["da"][l], cache["d"][l], g_d[l])
        g_cat[l] = cb(net.dec(l, 0), cache["cat"][l], cache["da"][l], g_da)
        ca = net.ch[l + 1]
        g_up = r(up2_backward(g_cat[l][..., :ca]))
        if l == net.L - 2:
            g_e_last = g_up
        else:
            g_d[l + 1] = g_up
    g_in = {}
    for l in range(net.L - 1, -1, -1):
        if l == net.L - 1:
            g_e = g_e_last
        else:
            ca = net.ch[l + 1]
            g_e = r(pool_backward(cache["e"][l], g_in[l + 1]) + g_cat[l][..., ca:])
        g_ea = cb(net.enc(l, 1), cache["ea"][l], cache["e"][l], g_e)
        g_in[l] = cb(net.enc(l, 0), cache["in"][l], cache["ea"][l], g_ea)
    return grads


def loss_and_gradients(net: Net, beauty, albedo, normal, target, params=None, bf=False, kind=LOSS_REL_L2,
                       beauty_b=None):
    z, cache = forward(net, beauty, albedo, normal, params, bf, beauty_b)
    loss, dz = loss_and_dz(z, target, bf, kind)
    return loss, backward(net, cache, dz, bf), z


def radiance(z, h, w):
    return np.expm1(np.maximum(z[:, :h, :w, :3], 0.0))
