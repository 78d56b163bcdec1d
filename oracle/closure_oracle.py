"""CPU oracle of the learned subgrid closure. TEST INFRASTRUCTURE ONLY: the tests use it as
the checker of the HIP kernels in fip_amd/csrc/closure.hip; the product never imports it.

The closure has no counterpart in the reference (which solves one grid at one resolution,
src/kernels.cu:264-438), so this restatement is pinned by mathematics instead of by
reference vectors: the restriction by closed forms and conservation invariants, the
network's backward pass by central finite differences in binary64
(tests/test_closure_oracle.py).

Network: 4 layers of 3x3 convolutions with zero padding, 32 channels, an activation after
the first three (act="gelu": GELU, tanh form; act="hsilu": hard SiLU,
x * clamp(x / 4 + 1 / 2, 0, 1)); inputs 6 features, outputs 2 (corrections of qx and qy). Parameters
in the padded layout of include/fip_closure.h: per layer W[tap][ci][co] (32 x 32) then
bias[co]; tap t = (dy + 1) * 3 + (dx + 1) reads the input at (y + dy, x + dx).

With emulate=True the forward/backward follow the GPU's rounding: inputs, weights,
pre-activations Z, activations A and the backward tensors dZ are rounded to bf16, sums
are formed in float32 (the order of a sum differs from the GPU's, so agreement is to
float32 rounding of the sums, not bitwise).
"""
from __future__ import annotations

import numpy as np

C = 32
TAPS = 9
FEAT = 6
OUT = 2
LAYERS = 4
LP = TAPS * C * C + C
NPARAMS = LAYERS * LP
S = 0.7978845608028654  # sqrt(2/pi)


def bf16(x):
    """Round float32 to bf16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    nan = np.isnan(a)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32).reshape(a.shape)
    return np.where(nan, np.float32(np.nan), out).astype(np.float32)


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(S * (x + 0.044715 * x ** 3)))


def dgelu(x):
    u = S * (x + 0.044715 * x ** 3)
    t = np.tanh(u)
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * S * (1.0 + 3.0 * 0.044715 * x * x)


def hsilu(x):
    """Hard SiLU, x * clamp(x / 4 + 1 / 2, 0, 1), in the dtype of x (x / 4 is exact, so in
    float32 this is the GPU's fused multiply-add with clamp and multiply, bit for bit)."""
    x = np.asarray(x)
    t = np.clip(x * x.dtype.type(0.25) + x.dtype.type(0.5), 0, 1).astype(x.dtype)
    return x * t


def dhsilu(x):
    """Derivative of hsilu: 0 for x <= -2, x / 2 + 1 / 2 between, 1 for x >= 2."""
    x = np.asarray(x)
    mid = x * x.dtype.type(0.5) + x.dtype.type(0.5)
    return np.where(x >= 2, x.dtype.type(1), np.where(x <= -2, x.dtype.type(0), mid)).astype(x.dtype)


ACTS = {"gelu": (gelu, dgelu), "hsilu": (hsilu, dhsilu)}


def split(params):
    """Padded parameter vector -> list of (W[9, 32, 32], b[32])."""
    p = np.asarray(params).reshape(LAYERS, LP)
    return [(p[l, :TAPS * C * C].reshape(TAPS, C, C), p[l, TAPS * C * C:]) for l in range(LAYERS)]


def join(layers, dtype=np.float64):
    out = np.zeros((LAYERS, LP), dtype=dtype)
    for l, (w, b) in enumerate(layers):
        out[l, :TAPS * C * C] = w.reshape(-1)
        out[l, TAPS * C * C:] = b
    return out.reshape(-1)


def shifted(x, dy, dx):
    """x[n, y + dy, x + dx, :] with zeros outside the image."""
    N, H, W, K = x.shape
    out = np.zeros_like(x)
    ys, yd = (slice(dy, H), slice(0, H - dy)) if dy >= 0 else (slice(0, H + dy), slice(-dy, H))
    xs, xd = (slice(dx, W), slice(0, W - dx)) if dx >= 0 else (slice(0, W + dx), slice(-dx, W))
    out[:, yd, xd, :] = x[:, ys, xs, :]
    return out


def conv(x, w, b=None, acc=np.float64):
    """3x3 convolution, zero padding: out[p] = sum_t x[p + off(t)] @ w[t] (+ b)."""
    N, H, W, _ = x.shape
    out = np.zeros((N, H, W, w.shape[2]), dtype=acc)
    for t in range(TAPS):
        dy, dx = t // 3 - 1, t % 3 - 1
        out += shifted(x, dy, dx).astype(acc) @ w[t].astype(acc)
    if b is not None:
        out += b.astype(acc)
    return out


def conv_input_grad(dz, w, acc=np.float64):
    """dX[q] = sum_t dZ[q - off(t)] @ w[t].T"""
    N, H, W, _ = dz.shape
    out = np.zeros((N, H, W, w.shape[1]), dtype=acc)
    for t in range(TAPS):
        dy, dx = t // 3 - 1, t % 3 - 1
        out += shifted(dz, -dy, -dx).astype(acc) @ w[t].T.astype(acc)
    return out


def conv_weight_grad(x, dz, acc=np.float64):
    """dW[t][ci][co] = sum_p x[p + off(t)][ci] dz[p][co]; db[co] = sum_p dz[p][co]."""
    K, Co = x.shape[3], dz.shape[3]
    dw = np.zeros((TAPS, K, Co), dtype=acc)
    dzf = dz.reshape(-1, Co).astype(acc)
    for t in range(TAPS):
        dy, dx = t // 3 - 1, t % 3 - 1
        dw[t] = shifted(x, dy, dx).reshape(-1, K).astype(acc).T @ dzf
    return dw, dzf.sum(axis=0)


def pad_features(feat):
    f = np.zeros(feat.shape[:3] + (C,), dtype=feat.dtype)
    f[..., :FEAT] = feat
    return f


def forward(feat, params, emulate=True, act="gelu"):
    """feat [N, H, W, 6] -> raw outputs o [N, H, W, 2] and the cache for backward."""
    f_act = ACTS[act][0]
    layers = split(params)
    acc = np.float32 if emulate else np.float64
    rnd = bf16 if emulate else (lambda a: a)
    x = rnd(pad_features(np.asarray(feat, np.float32 if emulate else np.float64)))
    cache = {"x": [x], "z": []}
    for l in range(LAYERS - 1):
        w, b = layers[l]
        z = rnd(conv(x, rnd(w.astype(acc)), b.astype(acc), acc=acc).astype(acc))
        a = rnd(f_act(z).astype(acc))
        cache["z"].append(z)
        cache["x"].append(a)
        x = a
    w, b = layers[-1]
    o = conv(x, rnd(w.astype(acc)), b.astype(acc), acc=acc)[..., :OUT]
    return o, cache


def loss_terms(o, qp, qt, mask, bound):
    th = np.tanh(o)
    sc = bound * np.abs(qp)
    r = qp + th * sc - qt
    m = mask.astype(o.dtype)
    return th, sc, r, m


def loss_and_grad(feat, params, qp, qt, mask, bound=0.1, emulate=True, act="gelu"):
    """Loss sum m r^2 / (2 N H W) and its gradient in the padded parameter layout."""
    o, cache = forward(feat, params, emulate, act)
    d_act = ACTS[act][1]
    acc = np.float32 if emulate else np.float64
    rnd = bf16 if emulate else (lambda a: a)
    N, H, W, _ = o.shape
    npix = N * H * W
    th, sc, r, m = loss_terms(o.astype(np.float64), np.asarray(qp, np.float64), np.asarray(qt, np.float64),
                              np.asarray(mask), bound)
    loss = float((m * r * r).sum() / (2.0 * npix))
    do = m * r * sc * (1.0 - th * th) / npix
    dz = np.zeros((N, H, W, C), dtype=acc)
    dz[..., :OUT] = do
    dz = rnd(dz.astype(acc))
    layers = split(params)
    grads = []
    for l in range(LAYERS - 1, -1, -1):
        w, _ = layers[l]
        dw, db = conv_weight_grad(cache["x"][l], dz, acc=acc)
        grads.append((dw, db))
        if l > 0:
            dx = conv_input_grad(dz, rnd(w.astype(acc)), acc=acc)
            dz = rnd((dx * d_act(cache["z"][l - 1].astype(acc))).astype(acc))
    grads = grads[::-1]
    g = join(grads, dtype=np.float64)
    return loss, g


def live_mask():
    """Parameters that exist (not padding): layer 0 reads 6 channels, layer 3 writes 2."""
    m = np.ones((LAYERS, LP), dtype=bool)
    w0 = m[0, :TAPS * C * C].reshape(TAPS, C, C)
    w0[:, FEAT:, :] = False
    w3 = m[LAYERS - 1, :TAPS * C * C].reshape(TAPS, C, C)
    w3[:, :, OUT:] = False
    m[LAYERS - 1, TAPS * C * C + OUT:] = False
    return m.reshape(-1)


def adam(w, m, v, g, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step in float32 (as the GPU kernel: 1 - beta is formed in binary32 from
    the binary32 beta); padded entries stay 0."""
    f = np.float32
    live = live_mask()
    m = (f(b1) * m + (f(1) - f(b1)) * g).astype(f)
    v = (f(b2) * v + (f(1) - f(b2)) * g * g).astype(f)
    c1 = f(1) - f(b1) ** f(t)
    c2 = f(1) - f(b2) ** f(t)
    w = (w - f(lr) * (m / c1) / (np.sqrt(v / c2) + f(eps))).astype(f)
    w = np.where(live, w, f(0))
    return w, m, v


# ---------------------------------------------------------------------------------------
# restriction


def restrict(h, qx, qy, F):
    """Fine [H, W] fields -> coarse [H/F, W/F]: h = block mean; qx on the coarse east face
    = mean of the F fine east faces at column F I + F - 1; qy on the north face likewise.
    Each sum is a sequential binary64 sum, rows of the block in order and columns in order
    within a row, scaled by 1/F^2 (1/F) and rounded once to binary32. (A binary64 sum of
    16 floats is exact only while their exponents lie within 25 bits of each other; a block
    with both a nearly dry cell, say 1e-35 m, and a deep one rounds, so the order is part
    of the definition.)"""
    H, W = h.shape
    Hc, Wc = H // F, W // F
    h = np.asarray(h, np.float32)
    qx = np.asarray(qx, np.float32)
    qy = np.asarray(qy, np.float32)
    sh = np.zeros((Hc, Wc))
    for r in range(F):
        for q in range(F):
            sh += h[r::F, q::F][:Hc, :Wc].astype(np.float64)
    sx = np.zeros((Hc, Wc))
    sy = np.zeros((Hc, Wc))
    for r in range(F):
        sx += qx[r::F, F - 1::F][:Hc, :Wc].astype(np.float64)
    for q in range(F):
        sy += qy[F - 1::F, q::F][:Hc, :Wc].astype(np.float64)
    inv = 1.0 / F
    return (sh * inv * inv).astype(np.float32), (sx * inv).astype(np.float32), (sy * inv).astype(np.float32)


def coarse_bed(bed, F):
    H, W = bed.shape
    b = np.asarray(bed, np.float64).reshape(H // F, F, W // F, F)
    return (b.sum(axis=(1, 3)) / (F * F)).astype(np.float32)


def bed_std(bed, F):
    H, W = bed.shape
    b = np.asarray(bed, np.float64).reshape(H // F, F, W // F, F)
    mu = b.mean(axis=(1, 3), keepdims=True)
    return np.sqrt(((b - mu) ** 2).mean(axis=(1, 3))).astype(np.float32)


def features(h, qx, qy, bed, bstd, kst, zref):
    """The 6 input channels of cl_features (float32 before bf16 rounding)."""
    f = np.float32
    z = (h.astype(f) + bed.astype(f) - f(zref)) * f(0.25)
    kk = np.broadcast_to(np.asarray(kst, f), h.shape) * f(1.0 / 30.0)
    return np.stack([z, qx.astype(f), qy.astype(f), (bed.astype(f) - f(zref)) * f(0.25),
                     bstd.astype(f) * f(2.0), kk], axis=-1).astype(f)


def interior_face_mask(Hc, Wc):
    m = np.zeros((Hc, Wc, 2), dtype=np.uint8)
    m[2:Hc - 2, 1:Wc - 2, 0] = 1
    m[1:Hc - 2, 2:Wc - 2, 1] = 1
    return m


# ---------------------------------------------------------------------------------------
# metrics of a coarse run against the restricted fine run


def depth_rmse(h, h_ref, valid=None):
    d = (np.asarray(h, np.float64) - np.asarray(h_ref, np.float64))
    if valid is not None:
        d = d[valid]
    return float(np.sqrt((d * d).mean()))


def max_depth_error(hmax, hmax_ref, valid=None):
    d = np.abs(np.asarray(hmax, np.float64) - np.asarray(hmax_ref, np.float64))
    if valid is not None:
        d = d[valid]
    return float(d.max())


def wet_iou(h, h_ref, thresh=0.01):
    a = np.asarray(h) > thresh
    b = np.asarray(h_ref) > thresh
    union = np.logical_or(a, b).sum()
    return float(np.logical_and(a, b).sum() / union) if union else 1.0
