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
    ms = S.make_cavities(80, size, seed=1, re_range=(10.0, 40.0), boxes=(0, 2), edges=(2, 5))
    train_ms, test_ms = ms[:64], ms[64:]
    b, steps, res, left, _ = S.generate(train_ms, tol=tol, max_steps=60000, check_every=every)
    assert left == 0
    x = S.features(train_ms)
    y, w = S.targets(b, train_ms)
    del b
    net = S.UNet(size, 3, (16, 32, 32), max_batch=8, seed=1)
    hist = S.train(net, S.Dataset(x, y, w), steps=600, batch=8, lr=3e-3, log_every=100)
    print("loss", hist)
    assert hist[-1][1] < 0.1 * hist[0][1]

    xt = S.features(test_ms)
    pred = S.predict(net, xt)
    rho, u = S.fields_from_prediction(pred, test_ms)
    cold = nx.CavityBatch(test_ms, "f32dev")
    cold.init()
    steps_c, res_c, left_c = cold.run(max_steps=60000, check_every=every, tol=tol)
    warm = nx.CavityBatch(test_ms, "f32dev")
    fb = S.warm_start(warm, rho, u, every)
    steps_w, res_w, left_w = warm.run(max_steps=60000, check_every=every, tol=tol)
    print("cold", steps_c.tolist(), "warm", steps_w.tolist(), "fallback", fb.tolist())
    assert left_c == 0 and left_w == 0
    assert not fb.any()
    # the warm start's own steps (constrained start, two probe intervals) count against it
    assert (steps_w + S.warm_start_overhead(every)).sum() < 0.8 * steps_c.sum()

    # a start with the flow reversed is worse than rest: every member falls back
    fb2 = S.warm_start(warm, rho, -u, every)
    assert fb2.all()
    steps_f, _, left_f = warm.run(max_steps=60000, check_every=every, tol=tol)
    assert left_f == 0 and np.array_equal(steps_f, steps_c)
