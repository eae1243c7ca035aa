"""FP64 numpy restatement of the warm-start network (3D convolutions and the
U-Net built from them), its gradients and its Adam step.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP
convolution kernels and of the device-trained U-Net; never a product path.

The network has no counterpart in the reference: it is project-defined, so
this file is pinned by mathematics instead of reference vectors (tests in
tests/test_surrogate_cpu.py): the convolution against scipy.ndimage.correlate,
its two adjoints against the inner-product identity, the network gradient
against central finite differences, bf16 rounding against torch's CPU
conversion.

Conventions (shared with the device code): channels-last volumes
x[n][z][y][x][c]; weights W[co][tap][ci] with tap = (dz*k + dy)*k + dx for
offsets -(k-1)/2..(k-1)/2; zero padding, stride 1.  With ``emulate=True`` every
tensor the device stores in bf16 is rounded to bf16 at the same place.
"""
from __future__ import annotations

import numpy as np

C0 = 16  # padded input / output channels


def bf16_round(a) -> np.ndarray:
    """Round to the nearest bf16 value (ties to even) via float32, as the
    device converts its FP32 results; returned as float64."""
    f = np.asarray(a, np.float64).astype(np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    nan = np.isnan(f)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, (u >> 16) | 0x40, r).astype(np.uint32)
    return (r << 16).view(np.float32).astype(np.float64)


def to_bf16_bits(a) -> np.ndarray:
    return (bf16_round(a).astype(np.float32).view(np.uint32) >> 16).astype(np.uint16)


def from_bf16_bits(b) -> np.ndarray:
    return (np.asarray(b, np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _taps(ks):
    r = (ks - 1) // 2
    return [(dz, dy, dx) for dz in range(-r, r + 1) for dy in range(-r, r + 1) for dx in range(-r, r + 1)]


def im2col(x: np.ndarray, ks: int) -> np.ndarray:
    """[n,d,h,w,c] -> [n,d,h,w,T,c]: the input at voxel + offset(tap), 0 outside."""
    n, d, h, w, c = x.shape
    r = (ks - 1) // 2
    xp = np.zeros((n, d + 2 * r, h + 2 * r, w + 2 * r, c), x.dtype)
    xp[:, r:r + d, r:r + h, r:r + w] = x
    cols = [xp[:, r + dz:r + dz + d, r + dy:r + dy + h, r + dx:r + dx + w] for dz, dy, dx in _taps(ks)]
    return np.stack(cols, axis=4)


def col2im(col: np.ndarray, ks: int) -> np.ndarray:
    """Adjoint of im2col: [n,d,h,w,T,c] -> [n,d,h,w,c] (scatter-add)."""
    n, d, h, w, T, c = col.shape
    r = (ks - 1) // 2
    xp = np.zeros((n, d + 2 * r, h + 2 * r, w + 2 * r, c), col.dtype)
    for t, (dz, dy, dx) in enumerate(_taps(ks)):
        xp[:, r + dz:r + dz + d, r + dy:r + dy + h, r + dx:r + dx + w] += col[:, :, :, :, t]
    return xp[:, r:r + d, r:r + h, r:r + w]


def conv_fwd(x, W, b=None, ks=3):
    """y[v, co] = sum_{tap, ci} x[v + off(tap), ci] W[co, tap, ci] + b[co]."""
    co = W.shape[0]
    col = im2col(np.asarray(x, np.float64), ks)
    y = col.reshape(-1, W.shape[1] * W.shape[2]) @ np.asarray(W, np.float64).reshape(co, -1).T
    if b is not None:
        y = y + np.asarray(b, np.float64)[None, :]
    return y.reshape(*x.shape[:4], co)


def conv_bwd_data(gy, W, ks=3):
    """Gradient of sum(conv_fwd(x) * gy) with respect to x."""
    co, T, ci = W.shape
    g = np.asarray(gy, np.float64).reshape(-1, co) @ np.asarray(W, np.float64).reshape(co, T * ci)
    return col2im(g.reshape(*gy.shape[:4], T, ci), ks)


def conv_bwd_weight(x, gy, ks=3):
    """(dW[co, tap, ci], db[co]) of sum(conv_fwd(x) * gy)."""
    co = gy.shape[-1]
    col = im2col(np.asarray(x, np.float64), ks)
    T, ci = col.shape[4], col.shape[5]
    g = np.asarray(gy, np.float64).reshape(-1, co)
    dW = g.T @ col.reshape(-1, T * ci)
    return dW.reshape(co, T, ci), g.sum(axis=0)


def pool(x):
    n, d, h, w, c = x.shape
    return x.reshape(n, d // 2, 2, h // 2, 2, w // 2, 2, c).mean(axis=(2, 4, 6))


def pool_bwd(g):
    return upsample(g) / 8.0


def upsample(x):
    return x.repeat(2, axis=1).repeat(2, axis=2).repeat(2, axis=3)


def upsample_bwd(g):
    n, d, h, w, c = g.shape
    return g.reshape(n, d // 2, 2, h // 2, 2, w // 2, 2, c).sum(axis=(2, 4, 6))


class UNet:
    """The device network's architecture and parameter layout (see the
    nx_unet_* declarations): L levels, widths wd, cubes of side S."""

    def __init__(self, levels, widths, in_ch, out_ch):
        self.L = int(levels)
        self.wd = [int(v) for v in widths[:self.L]]
        self.in_ch, self.out_ch = int(in_ch), int(out_ch)
        self.convs = []   # (cin, cout, ks, w_off, b_off)
        off = 0

        def add(ci, co, ks):
            nonlocal off
            w_off = off
            off += co * ks ** 3 * ci
            self.convs.append((ci, co, ks, w_off, off))
            off += co

        L, wd = self.L, self.wd
        for l in range(L):
            add(C0 if l == 0 else wd[l - 1], wd[l], 3)
            add(wd[l], wd[l], 3)
        for l in range(L - 1):
            add(wd[l + 1] + wd[l], wd[l], 3)
            add(wd[l], wd[l], 3)
        add(wd[0], C0, 1)
        self.n_params = off

    def enc(self, l, k):
        return 2 * l + k

    def dec(self, l, k):
        return 2 * self.L + 2 * l + k

    def out(self):
        return 4 * self.L - 2

    def wb(self, p, i):
        ci, co, ks, w_off, b_off = self.convs[i]
        return p[w_off:b_off].reshape(co, ks ** 3, ci), p[b_off:b_off + co]

    def forward(self, p, x, emulate=True):
        """x: [n, S, S, S, in_ch] -> y [n, S, S, S, out_ch]; keeps activations."""
        R = bf16_round if emulate else (lambda a: a)
        pw = np.asarray(p, np.float64)
        W = [(R(w), b) for w, b in (self.wb(pw, i) for i in range(len(self.convs)))]
        self.W = W
        n = x.shape[0]
        inp = np.zeros((*x.shape[:4], C0))
        inp[..., :self.in_ch] = x
        A = {"in0": R(inp)}
        relu = lambda a: np.maximum(a, 0.0)

        def cv(i, t, act=True):
            w, b = W[i]
            y = conv_fwd(t, w, b, self.convs[i][2])
            return R(relu(y)) if act else y

        L = self.L
        for l in range(L):
            A[f"a{l}"] = cv(self.enc(l, 0), A[f"in{l}"])
            A[f"e{l}"] = cv(self.enc(l, 1), A[f"a{l}"])
            if l < L - 1:
                A[f"in{l + 1}"] = R(pool(A[f"e{l}"]))
        low = A[f"e{L - 1}"]
        for l in range(L - 2, -1, -1):
            A[f"cat{l}"] = np.concatenate([upsample(low), A[f"e{l}"]], axis=-1)
            A[f"b{l}"] = cv(self.dec(l, 0), A[f"cat{l}"])
            A[f"d{l}"] = cv(self.dec(l, 1), A[f"b{l}"])
            low = A[f"d{l}"]
        y = cv(self.out(), low, act=False)
        A["y"] = y
        self.A, self.emulate = A, emulate
        return y[..., :self.out_ch]

    def backward(self, gy):
        """Parameter gradient (float64 vector) of sum(y * gy) for the last forward."""
        R = bf16_round if self.emulate else (lambda a: a)
        A, W, L, wd = self.A, self.W, self.L, self.wd
        grads = np.zeros(self.n_params)

        def put(i, dW, db):
            ci, co, ks, w_off, b_off = self.convs[i]
            grads[w_off:b_off] = dW.reshape(-1)
            grads[b_off:b_off + co] = db

        def cb(i, x, g, need_dx=True):
            put(i, *conv_bwd_weight(x, g, self.convs[i][2]))
            return R(conv_bwd_data(g, W[i][0], self.convs[i][2])) if need_dx else None

        relu_b = lambda y, g: np.where(y > 0, g, 0.0)
        g0 = np.zeros((*gy.shape[:4], C0))
        g0[..., :self.out_ch] = gy
        g = cb(self.out(), A["d0"], R(g0))
        gcat = {}
        for l in range(L - 1):
            g = relu_b(A[f"d{l}"], g)
            g = cb(self.dec(l, 1), A[f"b{l}"], g)
            g = relu_b(A[f"b{l}"], g)
            gcat[l] = cb(self.dec(l, 0), A[f"cat{l}"], g)
            g = R(upsample_bwd(gcat[l][..., :wd[l + 1]]))
        gin_next = None
        for l in range(L - 1, -1, -1):
            if l < L - 1:
                g = R(gcat[l][..., wd[l + 1]:] + pool_bwd(gin_next))
            g = relu_b(A[f"e{l}"], g)
            g = cb(self.enc(l, 1), A[f"a{l}"], g)
            g = relu_b(A[f"a{l}"], g)
            gin_next = cb(self.enc(l, 0), A[f"in{l}"], g, need_dx=l > 0)
            g = None
        return grads

    def loss_grad(self, y, target, weight):
        """loss = sum_v weight[v] sum_c (y - t)^2 and its gradient w.r.t. y."""
        d = y - target
        return float((weight[..., None] * d * d).sum()), 2.0 * weight[..., None] * d


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015) with bias correction; returns new (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
