"""CPU oracle of the radiance-cache MLP (test infrastructure only).

The network of include/lisa.h (`lisa_cache_*`): 12 inputs -> (sin, cos)
(pi 2^b x), b = 0..5 -> 144 features -> 4 x (64, ReLU) -> 3 linear outputs,
no biases; weights packed per layer as W_l^T[out_pad][in_pad].

`Net(precision="bf16")` mirrors the device arithmetic: BF16 (round to
nearest even) weights, encoded features, hidden activations and
backpropagated dZ; products accumulated in wide precision and rounded to
FP32.  `precision="f64"` runs everything in float64 -- the mode the
finite-difference checks use to pin the backward pass itself.
"""
from __future__ import annotations

import numpy as np

IN, BANDS, ENC, ENC_PAD, HID, OUT, OUT_PAD = 12, 6, 144, 160, 64, 3, 16
LAYERS = 5
OUT_DIM = (HID, HID, HID, HID, OUT_PAD)
IN_DIM = (ENC_PAD, HID, HID, HID, HID)
OFFSET = (0, 10240, 14336, 18432, 22528, 23552)
PARAMS = 23552
REAL_IN = (ENC, HID, HID, HID, HID)
REAL_OUT = (HID, HID, HID, HID, OUT)


def bf16(x):
    """Round float32 values to the nearest bfloat16 (ties to even), as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, ((u >> 16) | 0x40) << 16, r)
    return (r.astype(np.uint32)).view(np.float32).reshape(a.shape)


def layer(w, l):
    """W_l^T (out_pad x in_pad) view of the packed weights."""
    return w[OFFSET[l]:OFFSET[l + 1]].reshape(OUT_DIM[l], IN_DIM[l])


def mask_padding(w):
    w = np.array(w, dtype=np.float64 if w.dtype == np.float64 else np.float32)
    for l in range(LAYERS):
        W = layer(w, l)
        W[REAL_OUT[l]:, :] = 0
        W[:, REAL_IN[l]:] = 0
    return w


def encode(feat, dtype=np.float64):
    """(sin, cos)(pi 2^b p) for every input p, feature 12 i + 2 b + {0, 1}; padded to 160."""
    p = np.asarray(feat, np.float64).reshape(-1, IN)
    out = np.zeros((p.shape[0], ENC_PAD), dtype)
    for i in range(IN):
        for b in range(BANDS):
            x = np.pi * p[:, i] * (1 << b)
            out[:, 2 * BANDS * i + 2 * b] = np.sin(x)
            out[:, 2 * BANDS * i + 2 * b + 1] = np.cos(x)
    return out


class Net:
    def __init__(self, w, precision="bf16"):
        self.p = precision
        self.w = np.asarray(w, np.float64 if precision == "f64" else np.float32)

    def _r(self, x):
        if self.p == "f64":
            return np.asarray(x, np.float64)
        return bf16(np.asarray(x, np.float32))

    def _mm(self, a, b):
        prod = np.asarray(a, np.float64) @ np.asarray(b, np.float64)
        return prod if self.p == "f64" else prod.astype(np.float32)

    def forward(self, feat):
        X = self._r(encode(feat, np.float64 if self.p == "f64" else np.float32))
        acts = [X]
        h = X
        for l in range(LAYERS - 1):
            z = self._mm(h, self._r(layer(self.w, l)).T)
            h = self._r(np.maximum(z, 0))
            acts.append(h)
        y = self._mm(h, self._r(layer(self.w, LAYERS - 1)).T)[:, :OUT]
        return y, acts

    def loss_and_grads(self, feat, target):
        """Mean relative-L2 loss (denominator y^2 + 0.01 held constant) and dL/dW (packed)."""
        y, acts = self.forward(feat)
        t = np.asarray(target, y.dtype).reshape(-1, OUT)
        n = y.shape[0]
        den = y * y + 0.01
        loss = float(((y - t) ** 2 / den).sum() / (3 * n))
        g = 2.0 * (y - t) / den / (3.0 * n)
        dz = np.zeros((n, OUT_PAD), y.dtype)
        dz[:, :OUT] = g
        dz = self._r(dz)
        grad = np.zeros(PARAMS, np.float64)
        for l in range(LAYERS - 1, -1, -1):
            gw = np.asarray(dz, np.float64).T @ np.asarray(acts[l], np.float64)  # out_pad x in_pad
            grad[OFFSET[l]:OFFSET[l + 1]] = gw.reshape(-1)
            if l > 0:
                da = self._mm(dz, self._r(layer(self.w, l)))  # n x in
                dz = self._r(np.where(acts[l] > 0, da, 0))
        return loss, mask_padding(grad)


def relative_l2(w, feat, target):
    """Loss of the float64 network, for finite differences."""
    return Net(w, "f64").loss_and_grads(feat, target)[0]


def adam(w, m, v, g, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step on the packed FP32 master weights (padding untouched)."""
    w, m, v = (np.array(a, np.float32) for a in (w, m, v))
    g = np.asarray(g, np.float32)
    real = np.zeros(PARAMS, bool)
    for l in range(LAYERS):
        r = real[OFFSET[l]:OFFSET[l + 1]].reshape(OUT_DIM[l], IN_DIM[l])
        r[:REAL_OUT[l], :REAL_IN[l]] = True
    m = np.where(real, b1 * m + (1 - b1) * g, m).astype(np.float32)
    v = np.where(real, b2 * v + (1 - b2) * g * g, v).astype(np.float32)
    c1 = np.float32(1 - b1 ** step)
    c2 = np.float32(1 - b2 ** step)
    w = np.where(real, w - lr * (m / c1) / (np.sqrt(v / c2) + eps), w).astype(np.float32)
    return w, m, v
