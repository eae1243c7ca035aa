"""Pins of the FP64 network oracle (oracle/unet_oracle.py) by mathematics:
the convolution against scipy's correlate, its adjoints against the
inner-product identity, the network gradient against central differences,
bf16 rounding against torch's CPU conversion."""
import numpy as np
import pytest
import scipy.ndimage as ndi

from oracle import unet_oracle as U


def test_bf16_round_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    a = np.concatenate([rng.standard_normal(20000) * 10.0 ** rng.integers(-30, 30, 20000),
                        [0.0, -0.0, 1.0, 1.00390625, 1.005859375, 65504.0, 3.0e38]])
    ref = torch.tensor(a.astype(np.float32)).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(U.bf16_round(a), ref)
    bits = U.to_bf16_bits(a)
    assert np.array_equal(U.from_bf16_bits(bits), ref)


@pytest.mark.parametrize("ks", [1, 3])
def test_conv_fwd_matches_scipy_correlate(ks):
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 4, 5, 6, 3))
    W = rng.standard_normal((4, ks ** 3, 3))
    b = rng.standard_normal(4)
    y = U.conv_fwd(x, W, b, ks)
    for n in range(2):
        for co in range(4):
            ref = np.full(x.shape[1:4], b[co])
            for ci in range(3):
                k = W[co, :, ci].reshape(ks, ks, ks)       # [dz, dy, dx]
                ref += ndi.correlate(x[n, :, :, :, ci], k, mode="constant", cval=0.0)
            assert np.allclose(y[n, ..., co], ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("ks", [1, 3])
def test_conv_adjoints(ks):
    """<conv(x, W), g> = <x, bwd_data(g, W)> = <W, bwd_weight(x, g)>; db = sum g."""
    rng = np.random.default_rng(2)
    x = rng.standard_normal((2, 4, 4, 8, 5))
    W = rng.standard_normal((6, ks ** 3, 5))
    g = rng.standard_normal((2, 4, 4, 8, 6))
    lhs = (U.conv_fwd(x, W, None, ks) * g).sum()
    assert np.isclose(lhs, (x * U.conv_bwd_data(g, W, ks)).sum(), rtol=1e-12)
    dW, db = U.conv_bwd_weight(x, g, ks)
    assert np.isclose(lhs, (W * dW).sum(), rtol=1e-12)
    assert np.allclose(db, g.reshape(-1, 6).sum(0))
    # a single tap/channel by brute force: dW[co, t, ci] = sum_v g[v, co] x[v + off(t), ci]
    t = ks ** 3 - 1
    off = [o for o in U._taps(ks)][t]
    xs = np.zeros_like(x)
    d, h, w = x.shape[1:4]
    zs, ys, xs_ = [slice(max(0, -o), min(n, n - o)) for o, n in zip(off, (d, h, w))]
    zd, yd, xd = [slice(max(0, o), min(n, n + o)) for o, n in zip(off, (d, h, w))]
    xs[:, zs, ys, xs_] = x[:, zd, yd, xd]
    assert np.isclose(dW[3, t, 2], (g[..., 3] * xs[..., 2]).sum(), rtol=1e-12)


def test_pool_upsample_adjoints():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((2, 4, 4, 4, 3))
    g = rng.standard_normal((2, 2, 2, 2, 3))
    assert np.isclose((U.pool(x) * g).sum(), (x * U.pool_bwd(g)).sum())
    y = rng.standard_normal((2, 4, 4, 4, 3))
    assert np.isclose((U.upsample(g) * y).sum(), (g * U.upsample_bwd(y)).sum())


def _random_params(net, rng, scale=0.3):
    p = rng.standard_normal(net.n_params) * scale
    for ci, co, ks, w_off, b_off in net.convs:
        p[b_off:b_off + co] = rng.standard_normal(co) * 0.1
    return p


@pytest.mark.parametrize("levels,widths", [(2, [16, 16]), (3, [16, 32, 16])])
def test_network_gradient_finite_differences(levels, widths):
    rng = np.random.default_rng(4)
    net = U.UNet(levels, widths, in_ch=3, out_ch=4)
    p = _random_params(net, rng)
    x = rng.standard_normal((1, 4, 4, 4, 3))
    gy = rng.standard_normal((1, 4, 4, 4, 4))
    net.forward(p, x, emulate=False)
    g = net.backward(gy)
    f = lambda q: (net.forward(q, x, emulate=False) * gy).sum()
    # a few coordinates in every convolution (weights and biases)
    idx = []
    for ci, co, ks, w_off, b_off in net.convs:
        idx += list(rng.integers(w_off, b_off, 3)) + [b_off + int(rng.integers(0, 4))]
    eps = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = eps
        fd = (f(p + e) - f(p - e)) / (2 * eps)
        assert abs(fd - g[i]) <= 2e-5 * max(1.0, abs(fd)), (i, fd, g[i])


def test_network_layout_counts():
    net = U.UNet(4, [16, 32, 64, 64], in_ch=3, out_ch=4)
    # 15 convolutions; parameter count of the warm-start network
    assert len(net.convs) == 15
    expect = 0
    for ci, co, ks in [(16, 16, 3), (16, 16, 3), (16, 32, 3), (32, 32, 3), (32, 64, 3), (64, 64, 3),
                       (64, 64, 3), (64, 64, 3), (48, 16, 3), (16, 16, 3), (96, 32, 3), (32, 32, 3),
                       (128, 64, 3), (64, 64, 3), (16, 16, 1)]:
        expect += co * ci * ks ** 3 + co
    assert net.n_params == expect


def test_adam_first_step_is_signed_lr():
    p = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -1e-3, 0.0])
    q, m, v = U.adam(p, g, np.zeros(3), np.zeros(3), 1, lr=0.01)
    # first bias-corrected step: lr * g / (|g| + eps)
    assert np.allclose(q, p - 0.01 * g / (np.abs(g) + 1e-8))
