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
    """Steady moments of every member -> [n, S^3, 4] = ((rho-1)/ulb, u/ulb)
    and loss weights [n, S^3] (bulk cells, each member's mean).  The density
    target is scaled by ulb, not ulb^2: the pressure peaks at the lid corners
    would otherwise dominate the loss (the constrained start re-forms the
    density field anyway)."""
    ys, ws = [], []
    for k, m in enumerate(members):
        rho, u = batch.moments(k)
        ulb = m["ulb"]
        y = np.empty((rho.size, OUT_CH), np.float32)
        y[:, 0] = (rho.reshape(-1) - 1.0) / ulb
        y[:, 1:] = u.reshape(3, -1).T / ulb
        bulk = np.asarray(m["classes"]).reshape(-1) == abi.NX_CELL_BULK
        y[~bulk] = 0.0
        ys.append(y)
        ws.append(bulk.astype(np.float32) / (bulk.sum() * OUT_CH))
    return np.stack(ys), np.stack(ws)


def mirror_z(x: np.ndarray, y: np.ndarray, w: np.ndarray, size: int):
    """The z-mirrored copies of a data set: the lid-driven cube is symmetric
    under z -> N-1-z with u_z -> -u_z (lid along x on the y = N-1 face)."""
    n = x.shape[0]
    f = lambda a: a.reshape(n, size, size, size, -1)[:, ::-1].reshape(n, size ** 3, -1)
    ym = f(y).copy()
    ym[..., 3] *= -1.0
    return f(x).copy(), ym, f(w[..., None])[..., 0].copy()


def velocity_error(pred: np.ndarray, y: np.ndarray, w: np.ndarray) -> float:
    """Relative L2 error of the predicted velocity over bulk cells."""
    bulk = w > 0
    du = pred[..., 1:] - y[..., 1:]
    return float(np.sqrt((du[bulk] ** 2).sum() / (y[..., 1:][bulk] ** 2).sum()))


def fields_from_prediction(pred: np.ndarray, members):
    """Network output [n, S^3, 4] -> (rho, u) in the batch's arena order
    (rho per cell; u as 3 planes per member), float32."""
    rho, u = [], []
    for k, m in enumerate(members):
        ulb = m["ulb"]
        rho.append(1.0 + pred[k, :, 0] * ulb)
        u.append((pred[k, :, 1:] * ulb).T.reshape(-1))
    return (np.concatenate(rho).astype(np.float32), np.concatenate(u).astype(np.float32))


def converge(b: CavityBatch, tol, max_steps, check_every, chunk=10000, log=None):
    """Run an initialised batch until every member's residual is below tol
    (or max_steps), in chunks so that progress can be reported; returns
    (steps, residual, unconverged, seconds)."""
    t0 = time.perf_counter()
    done = 0
    while True:
        n = min(chunk, max_steps - done)
        steps, res, left = b.run(max_steps=n, check_every=check_every, tol=tol)
        done += n
        if log:
            log(f"  {done} steps, {left} of {b.n} unconverged")
        if left == 0 or done >= max_steps:
            return steps, res, left, time.perf_counter() - t0


def converged(res: np.ndarray, tol: float) -> np.ndarray:
    """Members that reached the residual (a diverged member reports NaN)."""
    return np.isfinite(res) & (res < tol)


def generate(members, tol=1e-6, max_steps=200000, check_every=50, storage="f32dev", log=None, chunk=10000):
    """Step every member to the residual with one batch (FP32 deviation
    storage by default: raw FP32 populations stall near a residual of 1e-5
    on small cavities); returns (batch, steps, residual, unconverged, seconds)."""
    b = CavityBatch(members, storage)
    b.init()
    steps, res, left, secs = converge(b, tol, max_steps, check_every, chunk=chunk, log=log)
    return b, steps, res, left, secs


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


CONSTRAIN_ITERATIONS = 10
PROBE_INTERVALS = 2      # check intervals per start in the fallback test


def warm_start(batch: CavityBatch, rho: np.ndarray, u: np.ndarray, check_every: int,
               iterations: int = CONSTRAIN_ITERATIONS):
    """Initialise `batch` from the predicted (rho, u): equilibrium populations,
    then `iterations` constrained double steps (nx_batch_constrain) so that
    density and non-equilibrium parts match the field.

    Fallback test: the predicted start and the rest start each run
    PROBE_INTERVALS check intervals, and a member falls back to rest when the
    predicted start's residual over the last interval,
    r = ||u(t) - u(t - K)|| / ||u(t)||, is not below the rest start's over the
    same interval.  The first interval cannot rank the two: the rest start's
    previous velocity is 0, which makes its first residual exactly 1.  Leaves
    the batch initialised (0 steps counted); returns the boolean fallback
    mask.  Cost beyond the returned state: warm_start_overhead steps of every
    member."""
    dr, du = DeviceBuffer.from_array(rho), DeviceBuffer.from_array(u)
    never = 1e-150            # no member retires during the probe intervals
    probe = PROBE_INTERVALS * check_every
    batch.init(dr, du)
    batch.constrain(du, iterations)
    _, res_w, _ = batch.run(max_steps=probe, check_every=check_every, tol=never)
    batch.init()
    _, res_c, _ = batch.run(max_steps=probe, check_every=check_every, tol=never)
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
    batch.constrain(du, iterations, members=~fallback)
    return fallback


def warm_start_overhead(check_every: int, iterations: int = CONSTRAIN_ITERATIONS) -> int:
    """Steps of every member spent by warm_start before the free iteration:
    the constrained start twice (probe and final) and both probes."""
    return 4 * iterations + 2 * PROBE_INTERVALS * check_every


def forward_flops(net: UNet) -> float:
    """Multiply-add FLOPs of one forward pass per sample: sum over the
    convolutions of 2 k^3 cin cout voxels."""
    S, L, wd = net.size, net.levels, net.widths
    f = 0.0
    for l in range(L):
        v = (S >> l) ** 3
        cin = 16 if l == 0 else wd[l - 1]
        f += 2 * 27 * v * (cin * wd[l] + wd[l] * wd[l])
        if l < L - 1:
            f += 2 * 27 * v * ((wd[l + 1] + wd[l]) * wd[l] + wd[l] * wd[l])
    f += 2 * S ** 3 * wd[0] * 16
    return f


def conv_layers(levels: int, widths, out_ch_padded: int = 16):
    """(cin, cout, ksize, level) of every convolution of the U-Net, in the
    parameter order of nx_unet_*."""
    L, wd = int(levels), [int(v) for v in widths]
    convs = []
    for l in range(L):
        convs.append((16 if l == 0 else wd[l - 1], wd[l], 3, l))
        convs.append((wd[l], wd[l], 3, l))
    for l in range(L - 1):
        convs.append((wd[l + 1] + wd[l], wd[l], 3, l))
        convs.append((wd[l], wd[l], 3, l))
    convs.append((wd[0], out_ch_padded, 1, 0))
    return convs


MFMA_BF16_PEAK_TFLOPS = 2500.0   # MI355X dense BF16 matrix-core peak


def conv_bench(size: int = 32, batch: int = 16, levels: int = 3, widths=(48, 96, 96), reps: int = 10):
    """Times nx_conv3d_{fwd,bwd_data,bwd_weight}_bf16 on every layer shape of
    the network (HIP events on one stream); FLOPs = 2 k^3 cin cout voxels per
    pass.  Returns a list of per-layer records."""
    from .lattice import Event, Stream
    init()
    st = Stream()
    rng = np.random.default_rng(0)
    out = []
    for ci, co, ks, lvl in conv_layers(levels, widths):
        side = size >> lvl
        shape = (batch, side, side, side)
        V = batch * side ** 3
        bits = lambda n: (rng.integers(0x3C00, 0x3F80, n, dtype=np.uint16))
        x = DeviceBuffer.from_array(bits(V * ci))
        dyb = DeviceBuffer.from_array(bits(V * co))
        w = DeviceBuffer.from_array(bits(co * ks ** 3 * ci))
        y = DeviceBuffer(V * co * 2)
        dx = DeviceBuffer(V * ci * 2)
        dw = DeviceBuffer(co * ks ** 3 * ci * 4)
        db = DeviceBuffer(co * 4)
        flops = 2.0 * ks ** 3 * ci * co * V
        rec = {"cin": ci, "cout": co, "k": ks, "shape": list(shape), "gflop": round(flops / 1e9, 3)}
        for name, fn in (("fwd", lambda: conv3d_fwd(x, w, None, y, shape, ci, co, ks, relu=True, stream=st.handle)),
                         ("bwd_data", lambda: conv3d_bwd_data(dyb, w, dx, shape, ci, co, ks, stream=st.handle)),
                         ("bwd_weight", lambda: conv3d_bwd_weight(x, dyb, dw, db, shape, ci, co, ks, stream=st.handle))):
            for _ in range(2):
                fn()
            e0, e1 = Event(), Event()
            e0.record(st)
            for _ in range(reps):
                fn()
            e1.record(st)
            e1.sync()
            ms = e0.elapsed_ms(e1) / reps
            tf = flops / (ms * 1e-3) / 1e12
            rec[name] = {"us": round(ms * 1e3, 1), "TFLOPs": round(tf, 1),
                         "frac": round(tf / MFMA_BF16_PEAK_TFLOPS, 4)}
        out.append(rec)
    return out


def main(argv=None):
    """Warm-start experiment: generate steady cavities with the batched
    solver, train the U-Net on the device, then iterate held-out cavities to
    the residual from rest and from the predicted start.  One JSON line."""
    import argparse
    import json
    ap = argparse.ArgumentParser(prog="python -m neon_amd.surrogate")
    ap.add_argument("--size", type=int, default=32)
    ap.add_argument("--train", type=int, default=512)
    ap.add_argument("--test", type=int, default=64)
    ap.add_argument("--re", type=float, nargs=2, default=(20.0, 200.0))
    ap.add_argument("--boxes", type=int, nargs=2, default=(0, 4))
    ap.add_argument("--edges", type=int, nargs=2, default=(4, 16))
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--check-every", type=int, default=50)
    ap.add_argument("--max-steps", type=int, default=200000)
    ap.add_argument("--chunk", type=int, default=10000, help="steps between progress lines of the solver runs")
    ap.add_argument("--storage", default="f32dev")
    ap.add_argument("--levels", type=int, default=4)
    ap.add_argument("--widths", default="16,32,64,64")
    ap.add_argument("--train-steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--save-params", default=None)
    ap.add_argument("--mirror", action="store_true", help="add the z-mirrored copies to the training set")
    ap.add_argument("--conv-bench", action="store_true",
                    help="only time the BF16 convolutions on every layer shape of the network")
    a = ap.parse_args(argv)
    if a.conv_bench:
        widths = tuple(int(v) for v in a.widths.split(","))
        recs = conv_bench(a.size, a.batch, a.levels, widths)
        tot = {k: sum(r[k]["us"] for r in recs) for k in ("fwd", "bwd_data", "bwd_weight")}
        fl = sum(r["gflop"] for r in recs) * 1e9
        for r in recs:
            print(json.dumps(r), flush=True)
        summary = {"conv_bench": True, "size": a.size, "batch": a.batch, "widths": list(widths),
                   "total_us": tot,
                   "TFLOPs": {k: round(fl / (v * 1e-6) / 1e12, 1) for k, v in tot.items()},
                   "peak_TFLOPs": MFMA_BF16_PEAK_TFLOPS}
        print(json.dumps(summary), flush=True)
        return summary
    log = lambda m: print(m, flush=True)
    widths = tuple(int(v) for v in a.widths.split(","))
    ms = make_cavities(a.train + a.test, a.size, a.seed, tuple(a.re), tuple(a.boxes), tuple(a.edges))
    train_ms, test_ms = ms[:a.train], ms[a.train:]
    log(f"generate {a.train} cavities {a.size}^3 to residual {a.tol:g}")
    b, steps_g, res_g, left_g, t_gen = generate(train_ms, a.tol, a.max_steps, a.check_every, a.storage,
                                                log=log, chunk=a.chunk)
    ok = converged(res_g, a.tol)
    x = features(train_ms)[ok]
    y, w = targets(b, train_ms)
    y, w = y[ok], w[ok]
    cells_g = int(sum(np.prod(m["dim"]) for m in train_ms))
    del b
    log(f"training set: {int(ok.sum())} converged, {int((~np.isfinite(res_g)).sum())} diverged, "
        f"{int(left_g)} unconverged")
    if a.mirror:
        xm, ym, wm = mirror_z(x, y, w, a.size)
        x, y, w = np.concatenate([x, xm]), np.concatenate([y, ym]), np.concatenate([w, wm])
    net = UNet(a.size, a.levels, widths, max_batch=a.batch, seed=a.seed + 1)
    data = Dataset(x, y, w)
    log(f"train {net.n_params} parameters, {a.train_steps} steps of {a.batch}")
    check(lib.nx_sync())
    t0 = time.perf_counter()
    hist = train(net, data, a.train_steps, a.batch, a.lr, seed=a.seed, log_every=200, log=log)
    t_train = time.perf_counter() - t0
    flops = 3.0 * forward_flops(net) * a.batch * a.train_steps
    if a.save_params:
        np.save(a.save_params, net.params())
    xt = features(test_ms)

    log(f"held out: {a.test} cavities from rest")
    cold = CavityBatch(test_ms, a.storage)
    # (members that diverge from rest are dropped from both comparisons)
    cold.init()
    check(lib.nx_sync())
    steps_c, res_c, left_c, t_cold = converge(cold, a.tol, a.max_steps, a.check_every, chunk=a.chunk, log=log)
    yt, wt = targets(cold, test_ms)
    del cold
    keep = converged(res_c, a.tol)
    if not keep.all():
        test_ms = [m for m, k in zip(test_ms, keep) if k]
        xt, yt, wt = xt[keep], yt[keep], wt[keep]
        steps_c = steps_c[keep]
        log(f"  {int((~keep).sum())} held-out cavities did not converge from rest; dropped, cold run repeated")
        cold = CavityBatch(test_ms, a.storage)
        cold.init()
        check(lib.nx_sync())
        steps_c, res_c, left_c, t_cold = converge(cold, a.tol, a.max_steps, a.check_every, chunk=a.chunk, log=log)
        del cold
    n_test = len(test_ms)

    log("held out: from the predicted start")
    warm = CavityBatch(test_ms, a.storage)
    check(lib.nx_sync())
    t0 = time.perf_counter()
    pred = predict(net, xt)
    t_inf = time.perf_counter() - t0
    rho, u = fields_from_prediction(pred, test_ms)
    fb = warm_start(warm, rho, u, a.check_every)
    t_probe = time.perf_counter() - t0 - t_inf
    steps_w, res_w, left_w, t_w = converge(warm, a.tol, a.max_steps, a.check_every, chunk=a.chunk, log=log)
    t_warm = time.perf_counter() - t0
    probe = warm_start_overhead(a.check_every)
    out = {
        "experiment": "learned warm start vs rest start, held-out cavities",
        "size": a.size, "train_fields": int(len(x)), "test": n_test, "re": list(a.re), "boxes": list(a.boxes),
        "edges": list(a.edges), "tol": a.tol, "check_every": a.check_every, "storage": a.storage,
        "generate": {"seconds": round(t_gen, 2), "unconverged": int(left_g),
                     "mean_steps": float(steps_g.mean()), "max_steps": int(steps_g.max()),
                     "MLUPS": round(cells_g * float(steps_g.mean()) / t_gen / 1e6, 1)},
        "network": {"levels": a.levels, "widths": list(widths), "params": net.n_params},
        "train": {"steps": a.train_steps, "batch": a.batch, "seconds": round(t_train, 2),
                  "loss_first": hist[0][1], "loss_last": hist[-1][1],
                  "TFLOPs": round(flops / t_train / 1e12, 2)},
        "held_out_loss": loss_of(pred, yt, wt),
        "held_out_velocity_rel_l2": velocity_error(pred, yt, wt),
        "cold": {"steps_mean": float(steps_c.mean()), "steps_total": int(steps_c.sum()),
                 "seconds": round(t_cold, 3), "unconverged": int(left_c)},
        "warm": {"steps_mean": float(steps_w.mean() + probe), "steps_total": int(steps_w.sum() + probe * n_test),
                 "seconds": round(t_warm, 3), "inference_s": round(t_inf, 3),
                 "probe_s": round(t_probe, 3), "unconverged": int(left_w),
                 "fallback": int(fb.sum())},
    }
    out["speedup_steps"] = round(out["cold"]["steps_total"] / max(1, out["warm"]["steps_total"]), 3)
    out["speedup_time"] = round(t_cold / t_warm, 3)
    print(json.dumps(out), flush=True)
    return out


if __name__ == "__main__":
    main()
