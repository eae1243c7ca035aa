"""The learned warm start end to end at a small scale: steady fields from the
batched solver, a U-Net trained on the device, held-out cavities iterated to
the same residual from the predicted start and from rest, and the fallback
to rest for a start that is worse than rest."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_warm_start_beats_rest_on_held_out_cavities(nx):
    from neon_amd import surrogate as S
    size, tol, every = 16, 1e-6, 50
    ms = S.make_cavities(1024 + 32, size, seed=1, re_range=(10.0, 40.0), boxes=(0, 2), edges=(2, 5))
    train_ms, test_ms = ms[:1024], ms[1024:]
    b, steps, res, left, _ = S.generate(train_ms, tol=tol, max_steps=30000, check_every=every)
    ok = S.converged(res, tol)
    assert ok.sum() >= 1000
    x = S.features(train_ms)[ok]
    y, w = S.targets(b, train_ms)
    y, w = y[ok], w[ok]
    del b
    xm, ym, wm = S.mirror_z(x, y, w, size)
    data = S.Dataset(np.concatenate([x, xm]), np.concatenate([y, ym]), np.concatenate([w, wm]))
    net = S.UNet(size, 3, (32, 64, 64), max_batch=16, seed=1)
    hist = S.train(net, data, steps=6000, batch=16, lr=5e-4, log_every=1000)
    print("loss", hist)
    assert hist[-1][1] < 0.01 * hist[0][1]

    cold = nx.CavityBatch(test_ms, "f32dev")
    cold.init()
    steps_c, res_c, left_c = cold.run(max_steps=30000, check_every=every, tol=tol)
    yt, wt = S.targets(cold, test_ms)
    del cold
    pred = S.predict(net, S.features(test_ms))
    err = S.velocity_error(pred, yt, wt)
    rho, u = S.fields_from_prediction(pred, test_ms)
    warm = nx.CavityBatch(test_ms, "f32dev")
    fb = S.warm_start(warm, rho, u, every)
    steps_w, res_w, left_w = warm.run(max_steps=30000, check_every=every, tol=tol)
    total_w = int(steps_w.sum()) + S.warm_start_overhead(every) * len(test_ms)
    print("velocity error", err, "cold", int(steps_c.sum()), "warm", total_w, "fallback", int(fb.sum()))
    assert left_c == 0 and left_w == 0
    assert err < 0.06
    assert not fb.any()
    # the warm start's own steps (constrained start, two probe intervals) count against it
    assert total_w < 0.95 * steps_c.sum()

    # a start with the flow reversed is worse than rest: every member falls back
    fb2 = S.warm_start(warm, rho, -u, every)
    assert fb2.all()
    steps_f, _, left_f = warm.run(max_steps=30000, check_every=every, tol=tol)
    assert left_f == 0 and np.array_equal(steps_f, steps_c)

    # the vortex mirrored top to bottom (the right flow in the wrong place:
    # the return flow under the lid) must fall back as well, while the right
    # field at 30% of its speed is closer to steady than rest and is kept
    u_mirror = np.concatenate([
        uk.reshape(3, d[2], d[1], d[0])[:, :, ::-1, :].reshape(-1)
        for uk, d in zip(np.split(u, np.cumsum([3 * d[0] * d[1] * d[2] for d in warm.dims])[:-1]),
                         warm.dims)])
    fb3 = S.warm_start(warm, rho, u_mirror, every)
    fb4 = S.warm_start(warm, rho, 0.3 * u, every)
    print("fallback: mirrored", int(fb3.sum()), "scaled 0.3", int(fb4.sum()), "of", len(test_ms))
    assert fb3.mean() > 0.9
    assert fb4.mean() < 0.1
