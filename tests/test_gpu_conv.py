"""BF16 matrix-core convolutions and the device U-Net against the FP64
oracle (oracle/unet_oracle.py) on the same bf16-valued inputs.

Tolerances: a single convolution accumulates up to 27*256 products in FP32,
so |gpu - exact| <= 1e-5 * (|x| conv |w|) elementwise (float outputs) or one
bf16 rounding step of that (bf16 outputs).  The network compares against the
oracle with bf16 rounding emulated at the same places; a rounding tie that
falls differently moves a value by one bf16 step (2^-8 relative) and
propagates, hence relative L2 bounds of 1e-2 (output) and 3e-2 (gradients)."""
import numpy as np
import pytest

from oracle import unet_oracle as U

pytestmark = pytest.mark.gpu


def _dev(nx, a):
    return nx.DeviceBuffer.from_array(np.ascontiguousarray(a))


def _bf(rng, shape, scale=1.0):
    bits = U.to_bf16_bits(rng.standard_normal(shape) * scale)
    return bits, U.from_bf16_bits(bits)


CASES = [  # (shape n,d,h,w), cin, cout, ks
    ((2, 4, 8, 16), 16, 16, 3),
    ((1, 8, 8, 8), 48, 32, 3),
    ((1, 4, 4, 4), 128, 64, 3),
    ((2, 2, 4, 8), 32, 96, 3),
    ((1, 1, 1, 1), 64, 16, 3),
    ((2, 8, 8, 8), 16, 16, 1),
    ((3, 4, 2, 4), 96, 48, 3),
]

# rows of >= 32 voxels: the all-taps weight-gradient kernel (k_bwd_w27),
# including 1, 3 and 4 output-channel tiles and several input-channel blocks
WIDE = [((1, 4, 4, 32), 16, 16, 3), ((2, 2, 4, 64), 48, 16, 3), ((1, 4, 2, 32), 32, 64, 3),
        ((1, 2, 2, 32), 16, 48, 3), ((2, 2, 2, 64), 96, 32, 3)]
# blocks of 32 x 4 x 4 voxels: the LDS-tiled forward, 16 input channels at a
# time (k_fwd_tile; also the backward-data pass of these layers)
TILED = [((1, 4, 4, 32), 16, 48, 3), ((2, 4, 4, 32), 32, 64, 3), ((1, 8, 8, 64), 16, 16, 3),
         ((1, 4, 4, 64), 32, 32, 3), ((1, 4, 4, 32), 48, 32, 3), ((2, 4, 4, 64), 48, 16, 3),
         ((1, 4, 8, 32), 96, 32, 3)]


@pytest.mark.parametrize("shape,cin,cout,ks", CASES + WIDE + TILED)
def test_conv_fwd_f32_vs_oracle(nx, shape, cin, cout, ks):
    from neon_amd import surrogate as S
    rng = np.random.default_rng(cin * 7 + cout)
    xb, x = _bf(rng, (*shape, cin))
    wb, w = _bf(rng, (cout, ks ** 3, cin), 0.2)
    b = rng.standard_normal(cout).astype(np.float32)
    y = nx.DeviceBuffer(int(np.prod(shape)) * cout * 4)
    S.conv3d_fwd(_dev(nx, xb), _dev(nx, wb), _dev(nx, b), y, shape, cin, cout, ks, out_f32=True)
    got = y.download(np.float32).reshape(*shape, cout)
    ref = U.conv_fwd(x, w, b.astype(np.float64), ks)
    bound = U.conv_fwd(np.abs(x), np.abs(w), np.abs(b.astype(np.float64)), ks)
    assert np.all(np.abs(got - ref) <= 1e-5 * bound + 1e-30)


@pytest.mark.parametrize("shape,cin,cout,ks", CASES[:4] + TILED[:2])
def test_conv_fwd_relu_bf16(nx, shape, cin, cout, ks):
    from neon_amd import surrogate as S
    rng = np.random.default_rng(11)
    xb, x = _bf(rng, (*shape, cin))
    wb, w = _bf(rng, (cout, ks ** 3, cin), 0.2)
    y = nx.DeviceBuffer(int(np.prod(shape)) * cout * 2)
    S.conv3d_fwd(_dev(nx, xb), _dev(nx, wb), None, y, shape, cin, cout, ks, relu=True)
    got = U.from_bf16_bits(y.download(np.uint16)).reshape(*shape, cout)
    exact = np.maximum(U.conv_fwd(x, w, None, ks), 0.0)
    bound = U.conv_fwd(np.abs(x), np.abs(w), None, ks)
    # one bf16 rounding (<= 2^-8 relative) of a value within the FP32 bound
    assert np.all(np.abs(got - exact) <= 2.0 ** -8 * np.abs(exact) + 1e-5 * bound + 1e-30)
    assert np.all(got >= 0)


@pytest.mark.parametrize("shape,cin,cout,ks", CASES + WIDE + TILED)
def test_conv_bwd_data_vs_oracle(nx, shape, cin, cout, ks):
    from neon_amd import surrogate as S
    rng = np.random.default_rng(cin + 3 * cout)
    gb, g = _bf(rng, (*shape, cout))
    wb, w = _bf(rng, (cout, ks ** 3, cin), 0.2)
    dx = nx.DeviceBuffer(int(np.prod(shape)) * cin * 4)
    S.conv3d_bwd_data(_dev(nx, gb), _dev(nx, wb), dx, shape, cin, cout, ks, out_f32=True)
    got = dx.download(np.float32).reshape(*shape, cin)
    ref = U.conv_bwd_data(g, w, ks)
    bound = U.conv_bwd_data(np.abs(g), np.abs(w), ks)
    assert np.all(np.abs(got - ref) <= 1e-5 * bound + 1e-30)


@pytest.mark.parametrize("shape,cin,cout,ks", CASES + WIDE)
def test_conv_bwd_weight_vs_oracle(nx, shape, cin, cout, ks):
    from neon_amd import surrogate as S
    rng = np.random.default_rng(5 * cin + cout)
    xb, x = _bf(rng, (*shape, cin))
    gb, g = _bf(rng, (*shape, cout))
    dw = nx.DeviceBuffer(cout * ks ** 3 * cin * 4)
    db = nx.DeviceBuffer(cout * 4)
    S.conv3d_bwd_weight(_dev(nx, xb), _dev(nx, gb), dw, db, shape, cin, cout, ks)
    got_w = dw.download(np.float32).reshape(cout, ks ** 3, cin)
    got_b = db.download(np.float32)
    ref_w, ref_b = U.conv_bwd_weight(x, g, ks)
    bw, bb = U.conv_bwd_weight(np.abs(x), np.abs(g), ks)
    assert np.all(np.abs(got_w - ref_w) <= 1e-5 * bw + 1e-30)
    assert np.all(np.abs(got_b - ref_b) <= 1e-5 * bb + 1e-30)


def test_conv_rejects_bad_shapes(nx):
    from neon_amd import surrogate as S
    buf = nx.DeviceBuffer(1 << 16)
    with pytest.raises(nx.abi.NeonError):
        S.conv3d_fwd(buf, buf, None, buf, (1, 3, 4, 4), 16, 16, 3)      # depth not a power of two
    with pytest.raises(nx.abi.NeonError):
        S.conv3d_fwd(buf, buf, None, buf, (1, 4, 4, 4), 24, 16, 3)     # cin not a multiple of 16
    with pytest.raises(nx.abi.NeonError):
        S.conv3d_fwd(buf, buf, None, buf, (1, 4, 4, 4), 16, 16, 5)     # kernel size


def _net_pair(nx, levels, widths, size, batch, seed):
    from neon_amd import surrogate as S
    net = S.UNet(size, levels, widths, max_batch=batch, seed=seed)
    orc = U.UNet(levels, widths, S.IN_CH, S.OUT_CH)
    assert orc.n_params == net.n_params
    return net, orc


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("levels,widths,size,batch", [(4, (16, 32, 64, 64), 8, 2),
                                                      (3, (16, 32, 48), 16, 1),
                                                      (2, (16, 32), 32, 1),
                                                      (2, (32, 16), 4, 3)])
def test_unet_forward_backward_vs_oracle(nx, levels, widths, size, batch):
    from neon_amd import surrogate as S
    net, orc = _net_pair(nx, levels, widths, size, batch, seed=7)
    p = net.params().astype(np.float64)
    rng = np.random.default_rng(3)
    V = size ** 3
    x = rng.random((batch, V, S.IN_CH)).astype(np.float32)
    x[..., :2] = x[..., :2] > 0.7
    dx = _dev(nx, x)
    y = nx.DeviceBuffer(batch * V * S.OUT_CH * 4)
    net.forward(dx, batch, y)
    got = y.download(np.float32).reshape(batch, V, S.OUT_CH)
    ref = orc.forward(p, x.reshape(batch, size, size, size, S.IN_CH).astype(np.float64))
    ref = ref.reshape(batch, V, S.OUT_CH)
    assert _rel(got, ref) <= 1e-2, _rel(got, ref)
    gy = rng.standard_normal((batch, V, S.OUT_CH)).astype(np.float32)
    net.backward(_dev(nx, gy), batch)
    g_gpu = net.grads().astype(np.float64)
    g_ref = orc.backward(gy.reshape(batch, size, size, size, S.OUT_CH).astype(np.float64))
    for i, (ci, co, ks, w_off, b_off) in enumerate(orc.convs):
        e_w = _rel(g_gpu[w_off:b_off], g_ref[w_off:b_off])
        e_b = _rel(g_gpu[b_off:b_off + co], g_ref[b_off:b_off + co])
        assert e_w <= 3e-2 and e_b <= 3e-2, (i, e_w, e_b)


def test_unet_train_step_loss_and_adam(nx):
    """First step: loss equals the oracle's; the parameter update equals the
    oracle Adam step applied to the device gradient; repeated steps on one
    batch drive the loss down."""
    from neon_amd import surrogate as S
    size, batch = 8, 2
    net, orc = _net_pair(nx, 3, (16, 32, 32), size, batch, seed=3)
    rng = np.random.default_rng(5)
    V = size ** 3
    x = rng.random((batch, V, S.IN_CH)).astype(np.float32)
    t = rng.standard_normal((batch, V, S.OUT_CH)).astype(np.float32)
    w = (rng.random((batch, V)) > 0.3).astype(np.float32)
    w /= w.sum(axis=1, keepdims=True) * S.OUT_CH
    dx, dt, dw = _dev(nx, x), _dev(nx, t), _dev(nx, w)
    p0 = net.params().astype(np.float64)
    y0 = orc.forward(p0, x.reshape(batch, size, size, size, S.IN_CH).astype(np.float64))
    l_ref, _ = orc.loss_grad(y0.reshape(batch, V, S.OUT_CH), t.astype(np.float64), w.astype(np.float64))
    lr = 1e-3
    l0 = net.train_step(dx, dt, dw, batch, lr, want_loss=True)
    assert abs(l0 - l_ref) <= 2e-2 * abs(l_ref)
    g = net.grads().astype(np.float64)
    p1 = net.params().astype(np.float64)
    q, _, _ = U.adam(p0, g, np.zeros_like(g), np.zeros_like(g), 1, lr)
    assert np.allclose(p1, q, rtol=1e-5, atol=1e-7)
    losses = [net.train_step(dx, dt, dw, batch, lr, want_loss=True) for _ in range(60)]
    assert losses[-1] < 0.5 * l0, (l0, losses[-1])
