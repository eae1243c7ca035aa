"""Learned warm start for batched cavities: a 3D U-Net (nx_unet_*, BF16
matrix-core convolutions trained on the device) predicts the steady (rho, u)
of a cavity from [solid mask, lid mask, Re/1000]; the LBM then iterates from
that start to the same residual as from rest, and any member whose
predicted start has the worse first-interval residual is restarted from rest.

No counterpart in the reference (project-defined).  Data flow:
  generate(): cavities of side S (wall layer included) stepped to a residual
      by one CavityBatch, their steady moments -> features / targets / weights
  train(): minibatch Adam steps on the device (nx_unet_train_step)
  predict() / warm_start(): network forward, equilibrium init, fallback test
"""
from __future__ import annotations

import ctypes as C
import math
import time

import numpy as np

from . import abi
from .abi import check, lib
from .lattice import CavityBatch, DeviceBuffer, cavity_classes, init

IN_CH, OUT_CH = 3, 4


# ---- BF16 convolution entry points (thin wrappers over device pointers) ----

def _shape(shape):
    return (C.c_int32 * 4)(*[int(v) for v in shape])


def conv3d_fwd(x, w, bias, y, shape, cin, cout, ksize=3, relu=False, out_f32=False, stream=None):
    flags = (abi.NX_CONV_RELU if relu else 0) | (abi.NX_CONV_OUT_F32 if out_f32 else 0)
    check(lib.nx_conv3d_fwd_bf16(x.ptr, w.ptr, bias.ptr if bias is not None else None, y.ptr,
                                 _shape(shape), int(cin), int(cout), int(ksize), flags, stream))


def conv3d_bwd_data(dy, w, dx, shape, cin, cout, ksize=3, out_f32=False, stream=None):
    check(lib.nx_conv3d_bwd_data_bf16(dy.ptr, w.ptr, dx.ptr, _shape(shape), int(cin), int(cout),
                                      int(ksize), abi.NX_CONV_OUT_F32 if out_f32 else 0, stream))


def conv3d_bwd_weight(x, dy, dw, db, shape, cin, cout, ksize=3, stream=None):
    check(lib.nx_conv3d_bwd_weight_bf16(x.ptr, dy.ptr, dw.ptr, db.ptr if db is not None else None,
                                        _shape(shape), int(cin), int(cout), int(ksize), stream))


class UNet:
    """Device U-Net (nx_unet_*): levels, widths, cubes of side `size`."""

    def __init__(self, size: int, levels: int = 4, widths=(16, 32, 64, 64), max_batch: int = 16,
                 seed: int = 1, in_ch: int = IN_CH, out_ch: int = OUT_CH):
        init()
        d = abi.UNetDesc()
        d.levels, d.size, d.max_batch = int(levels), int(size), int(max_batch)
        for k in range(4):
            d.width[k] = int(widths[k]) if k < len(widths) else 0
        d.in_ch, d.out_ch, d.seed = int(in_ch), int(out_ch), int(seed)
        h = C.c_void_p()
        check(lib.nx_unet_create(C.byref(d), C.byref(h)))
        self.h = h.value
        self.size, self.levels, self.widths = int(size), int(levels), tuple(widths[:levels])
        self.max_batch, self.in_ch, self.out_ch = int(max_batch), int(in_ch), int(out_ch)
        n, c = C.c_uint64(), C.c_uint32()
        check(lib.nx_unet_info(self.h, C.byref(n), C.byref(c)))
        self.n_params, self.n_convs = int(n.value), int(c.value)
        self.voxels = self.size ** 3

    def params(self) -> np.ndarray:
        out = np.empty(self.n_params, np.float32)
        check(lib.nx_unet_params(self.h, out.ctypes.data, 0))
        return out

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n_params
        check(lib.nx_unet_params(self.h, p.ctypes.data, 1))

    def grads(self) -> np.ndarray:
        out = np.empty(self.n_params, np.float32)
        check(lib.nx_unet_grads(self.h, out.ctypes.data))
        return out

    def forward(self, x: DeviceBuffer, batch: int, y: DeviceBuffer, stream=None) -> None:
        check(lib.nx_unet_forward(self.h, x.ptr, int(batch), y.ptr, stream))

    def backward(self, dy: DeviceBuffer, batch: int, stream=None) -> None:
        check(lib.nx_unet_backward(self.h, dy.ptr, int(batch), stream))

    def train_step(self, x, target, weight, batch: int, lr: float, want_loss=False, stream=None):
        loss = C.c_double(float("nan"))
        check(lib.nx_unet_train_step(self.h, x.ptr, target.ptr, weight.ptr, int(batch), float(lr),
                                     C.byref(loss) if want_loss else None, stream))
        return float(loss.value) if want_loss else None

    def __del__(self):
        try:
            lib.nx_unet_destroy(self.h)
        except Exception:
            pass


# ---- cavities, features, targets ------------------------------------------

def omega_of(re: float, size: int, ulb: float = 0.05) -> float:
    """nu = ulb * N / Re (N = cells along the lid, wall layer included)."""
    nu = ulb * size / re
    return 1.0 / (3.0 * nu + 0.5)


def make_cavities(n: int, size: int, seed: int, re_range=(20.0, 200.0), boxes=(0, 4),
                  edges=(4, 16), ulb: float = 0.05):
    """n cubes of side `size`: Re uniform in re_range, 0..4 seeded boxes with
    edges uniform in {edges} (clipped to the interior), lid speed ulb along x."""
    rng = np.random.default_rng(seed)
    ms = []
    for k in range(n):
        re = float(rng.uniform(*re_range))
        nb = int(rng.integers(boxes[0], boxes[1] + 1))
        cls = cavity_classes((size, size, size), nb, edges[0], edges[1],
                             seed=int(rng.integers(0, 2 ** 63)))
        ms.append({"dim": (size, size, size), "classes": cls, "omega": omega_of(re, size, ulb),
                   "ulb": ulb, "ulid": (1.0, 0.0, 0.0), "re": re})
    return ms


def features(members) -> np.ndarray:
    """[n, S^3, 3] float32: solid (bounce-back) mask, lid mask, Re/1000."""
    out = []
    for m in members:
        c = np.asarray(m["classes"]).reshape(-1)
        f = np.empty((c.size, IN_CH), np.float32)
        f[:, 0] = c == abi.NX_CELL_BOUNCEBACK
        f[:, 1] = c == abi.NX_CELL_MOVINGWALL
        f[:, 2] = m["re"] / 1000.0
        out.append(f)
    return np.stack(out)


def targets(batch: CavityBatch, members):
    """Steady moments of every member -> [n, S^3, 4] = ((rho-1)/ulb^2, u/ulb)
    and loss weights [n, S^3] (bulk cells, each member's mean)."""
    ys, ws = [], []
    for k, m in enumerate(members):
        rho, u = batch.moments(k)
        ulb = m["ulb"]
        y = np.empty((rho.size, OUT_CH), np.float32)
        y[:, 0] = (rho.reshape(-1) - 1.0) / ulb ** 2
        y[:, 1:] = u.reshape(3, -1).T / ulb
        bulk = np.asarray(m["classes"]).reshape(-1) == abi.NX_CELL_BULK
        y[~bulk] = 0.0
        ys.append(y)
        ws.append(bulk.astype(np.float32) / (bulk.sum() * OUT_CH))
    return np.stack(ys), np.stack(ws)


def fields_from_prediction(pred: np.ndarray, members):
    """Network output [n, S^3, 4] -> (rho, u) in the batch's arena order
    (rho per cell; u as 3 planes per member), float32."""
    rho, u = [], []
    for k, m in enumerate(members):
        ulb = m["ulb"]
        rho.append(1.0 + pred[k, :, 0] * ulb ** 2)
        u.append((pred[k, :, 1:] * ulb).T.reshape(-1))
    return (np.concatenate(rho).astype(np.float32), np.concatenate(u).astype(np.float32))


def generate(members, tol=1e-6, max_steps=200000, check_every=50, storage="f32dev"):
    """Step every member to the residual with one batch; returns
    (batch, steps, residual, unconverged, seconds)."""
    b = CavityBatch(members, storage)
    b.init()
    t0 = time.perf_counter()
    steps, res, left = b.run(max_steps=max_steps, check_every=check_every, tol=tol)
    return b, steps, res, left, time.perf_counter() - t0


# ---- training -----------------------------------------------------------------

class Dataset:
    """Features, targets and weights resident on the device, [n, S^3, ...]."""

    def __init__(self, x: np.ndarray, y: np.ndarray, w: np.ndarray):
        self.n, self.V = x.shape[0], x.shape[1]
        self.x = DeviceBuffer.from_array(np.ascontiguousarray(x, np.float32))
        self.y = DeviceBuffer.from_array(np.ascontiguousarray(y, np.float32))
        self.w = DeviceBuffer.from_array(np.ascontiguousarray(w, np.float32))


def train(net: UNet, data: Dataset, steps: int, batch: int, lr: float = 2e-3, lr_min: float = 5e-5,
          seed: int = 0, log_every: int = 100, log=None):
    """Minibatch Adam with a cosine learning rate; samples are gathered into
    contiguous device buffers by device-to-device copies.  Returns the loss
    history [(step, per-member loss)]."""
    assert batch <= net.max_batch and batch <= data.n
    V = data.V
    bx = DeviceBuffer(batch * V * IN_CH * 4)
    by = DeviceBuffer(batch * V * OUT_CH * 4)
    bw = DeviceBuffer(batch * V * 4)
    rng = np.random.default_rng(seed)
    perm, pos = rng.permutation(data.n), 0
    hist = []
    for s in range(steps):
        if pos + batch > data.n:
            perm, pos = rng.permutation(data.n), 0
        idx = perm[pos:pos + batch]
        pos += batch
        for j, i in enumerate(idx):
            i = int(i)
            check(lib.nx_d2d(bx.ptr + j * V * IN_CH * 4, data.x.ptr + i * V * IN_CH * 4, V * IN_CH * 4, None))
            check(lib.nx_d2d(by.ptr + j * V * OUT_CH * 4, data.y.ptr + i * V * OUT_CH * 4, V * OUT_CH * 4, None))
            check(lib.nx_d2d(bw.ptr + j * V * 4, data.w.ptr + i * V * 4, V * 4, None))
        cur = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * s / max(1, steps)))
        want = (s % log_every == 0) or s == steps - 1
        loss = net.train_step(bx, by, bw, batch, cur, want_loss=want)
        if want:
            loss /= batch   # per-member mean
            hist.append((s, loss))
            if log:
                log(f"step {s} loss {loss:.4e} lr {cur:.2e}")
    check(lib.nx_sync())
    return hist


def predict(net: UNet, x: np.ndarray) -> np.ndarray:
    """Network output for features [n, S^3, 3] -> [n, S^3, 4] (float32)."""
    n, V = x.shape[0], x.shape[1]
    B = net.max_batch
    dx = DeviceBuffer(B * V * IN_CH * 4)
    dy = DeviceBuffer(B * V * OUT_CH * 4)
    out = np.empty((n, V, OUT_CH), np.float32)
    for k in range(0, n, B):
        m = min(B, n - k)
        dx.upload(np.ascontiguousarray(x[k:k + m], np.float32))
        net.forward(dx, m, dy)
        out[k:k + m] = dy.download(np.float32, m * V * OUT_CH).reshape(m, V, OUT_CH)
    return out


def loss_of(pred: np.ndarray, y: np.ndarray, w: np.ndarray) -> float:
    """Mean over members of the weighted squared error (the training loss)."""
    return float((w[..., None] * (pred - y) ** 2).sum() / pred.shape[0])


def warm_start(batch: CavityBatch, rho: np.ndarray, u: np.ndarray, check_every: int):
    """Initialise `batch` from the predicted (rho, u), falling back to the
    rest start for every member whose residual after one check interval is
    not below the rest start's.  Leaves the batch initialised (0 steps taken)
    and returns the boolean fallback mask."""
    dr, du = DeviceBuffer.from_array(rho), DeviceBuffer.from_array(u)
    batch.init(dr, du)
    _, res_w, _ = batch.run(max_steps=check_every, check_every=check_every, tol=0.0)
    batch.init()
    _, res_c, _ = batch.run(max_steps=check_every, check_every=check_every, tol=0.0)
    fallback = ~(res_w < res_c)
    if fallback.any():
        rho = rho.copy()
        u = u.copy()
        off = 0
        for k, d in enumerate(batch.dims):
            n = d[0] * d[1] * d[2]
            if fallback[k]:
                rho[off:off + n] = 1.0
                u[3 * off:3 * off + 3 * n] = 0.0
            off += n
        dr.upload(rho)
        du.upload(u)
    batch.init(dr, du)
    return fallback
