"""Pins of the two oracle collisions that the textbook stepper alone leaves open.

collide_ref (oracle/lbm_oracle.c): the reference's float instantiation of
collideBgkUnrolled (LbmTools.h:171-195 moments, :209-281 collision, :311-313
usqr).  Its double literals promote parts of each expression to binary64 and
the declarations (`const LbmComputeType v = ...`, LbmComputeType = float)
round them back.  The oracle writes these promotions as explicit casts.  The
check here does not: `ref_float_collision` evaluates the reference's
expressions with binary32 arrays and binary64 literals and lets numpy's type
rules (float32 op float64 -> float64, float32 op float32 -> float32, the same
usual arithmetic conversions C++ applies) place every promotion; only the
declared types round.  A misplaced or dropped cast in the oracle then shows
as a bit difference.  A second, statistical check shows the promotions are
observable at all: the float path differs from the all-binary32 collision
on about half the populations and lies closer to the binary64 collision.

forced_dev (deviation storage with Guo forcing, project-defined): compared
with the textbook forced step of tests/textbook.py (Guo, Zheng & Shi 2002,
written from the formulas in float64) over 10 steps in a periodic box, and
with the closed-form momentum gain rho*g per step, in both binary32 and
binary16 deviation storage.
"""
import numpy as np
import pytest

import textbook as tb

L = np.float64  # a double literal of the C++ source


def _decl(x):
    """`const LbmComputeType v = expr;` with LbmComputeType = float."""
    return np.asarray(x).astype(np.float32)


def _sum(f, idx):
    acc = f[idx[0]]
    for i in idx[1:]:
        acc = acc + f[i]  # left to right, binary32
    return acc


def ref_float_collision(f: np.ndarray, omega: np.float32) -> np.ndarray:
    """Reference float path, f: binary32 pulled populations [19, n]."""
    assert f.dtype == np.float32 and isinstance(omega, np.float32)
    xm1, xp1 = _sum(f, (0, 3, 4, 5, 6)), _sum(f, (10, 13, 14, 15, 16))
    x0 = _sum(f, (9, 1, 2, 7, 8, 11, 12, 17, 18))
    ym1, yp1 = _sum(f, (1, 3, 7, 8, 14)), _sum(f, (4, 11, 13, 17, 18))
    zm1, zp1 = _sum(f, (2, 5, 7, 16, 18)), _sum(f, (6, 8, 12, 15, 17))
    rho = xm1 + xp1 + x0
    u0, u1, u2 = (xp1 - xm1) / rho, (yp1 - ym1) / rho, (zp1 - zm1) / rho
    usqr = _decl(L(1.5) * (u0 * u0 + u1 * u1 + u2 * u2))
    ck = [u0, u1, u2, u0 + u1, u0 - u1, u0 + u2, u0 - u2, u1 + u2, u1 - u2]
    out = np.empty_like(f)
    for i in range(9):
        w = L(1.0) / L(18.0) if i < 3 else L(1.0) / L(36.0)
        c = ck[i]
        eq = _decl(rho * w * (L(1.0) - L(3.0) * c + L(4.5) * c * c - usqr))
        eqopp = _decl(eq + rho * w * L(6.0) * c)
        out[i] = _decl((L(1.0) - omega) * f[i] + omega * eq)
        out[i + 10] = _decl((L(1.0) - omega) * f[i + 10] + omega * eqopp)
    eq9 = _decl(rho * (L(1.0) / L(3.0)) * (L(1.0) - usqr))
    out[9] = _decl((L(1.0) - omega) * f[9] + omega * eq9)
    return out


def _random_state(N, seed, du=0.05, noneq=0.01):
    rng = np.random.default_rng(seed)
    n = N ** 3
    rho = 1 + 0.02 * rng.standard_normal(n)
    u = du * rng.standard_normal((3, n))
    feq = tb.equilibrium(rho, u)
    return feq * (1 + noneq * rng.standard_normal(feq.shape))


def _one_step(orc, N, f0, storage, compute, omega):
    o = orc.DenseOracle((N, N, N), storage=storage, compute=compute, periodic=7)
    o.set_classes(np.full(N ** 3, orc.BULK, np.uint32))
    o.pop[0][:] = f0.reshape(-1)
    o.pop[1][:] = f0.reshape(-1)
    o.run(1, omega)
    return o.fin.reshape(19, -1)


def _pulled(f0, N):
    """Pull streaming in a fully periodic all-bulk box: f_q(x) = f0_q(x - c_q)."""
    g = f0.reshape(19, N, N, N)
    return np.stack([np.roll(g[q], shift=tuple(tb.C[q][::-1]), axis=(0, 1, 2)) for q in range(19)])


@pytest.mark.parametrize("omega", [1.85, 1.3, 0.7, 0.35])  # 1 - 0.35f is inexact in binary32
def test_ref_float_path_matches_type_rules(orc, omega):
    N = 16
    f0 = _random_state(N, seed=7).astype(np.float32)
    om = np.float32(omega)
    got = _one_step(orc, N, f0, orc.F32, orc.COMPUTE_REF, float(om))
    want = ref_float_collision(_pulled(f0, N).reshape(19, -1), om)
    assert got.dtype == np.float32
    diff = np.count_nonzero(got.view(np.uint32) != want.view(np.uint32))
    assert diff == 0, f"{diff} of {got.size} populations differ"


def test_ref_float_path_promotions_are_observable(orc):
    """The float path is neither the all-binary32 nor the binary64 collision,
    and its mean error against binary64 arithmetic on the same binary32
    inputs is at most 0.9x the all-binary32 one (measured 0.82x; a collision
    with every promotion dropped gives 1.0x)."""
    N, om = 24, np.float32(1.85)
    f0 = _random_state(N, seed=7).astype(np.float32)
    ref = _one_step(orc, N, f0, orc.F32, orc.COMPUTE_REF, float(om)).astype(np.float64)
    f32 = _one_step(orc, N, f0, orc.F32, orc.COMPUTE_F32, float(om)).astype(np.float64)
    exact = _one_step(orc, N, f0.astype(np.float64), orc.F64, orc.COMPUTE_F64, float(om))
    assert np.mean(ref != f32) > 0.4
    assert np.mean(ref != exact.astype(np.float32)) > 0.4
    e_ref, e_f32 = np.abs(ref - exact).mean(), np.abs(f32 - exact).mean()
    assert e_ref < 0.9 * e_f32, (e_ref, e_f32)


@pytest.mark.parametrize("storage,tol,mom_rtol", [("F32DEV", 1e-7, 1e-6), ("F16DEV", 2e-5, 2e-4)])
def test_forced_deviation_matches_textbook_and_momentum(orc, storage, tol, mom_rtol):
    """10 steps of Guo-forced BGK in deviation storage against the float64
    textbook step from the same decoded start.  Tolerances: binary32
    deviations round each d' to 2^-24 relative (|d| < 1e-2, so < 1e-9 per
    step); binary16 deviations to 2^-11 relative (< 5e-6 per step)."""
    N, om, g = 8, 1.3, (2e-3, -1e-3, 5e-4)
    st = getattr(orc, storage)
    d0 = _random_state(N, seed=11, du=0.04) - tb.W[:, None]
    if st == orc.F16DEV:
        d = d0.astype(np.float16)
        raw = d.view(np.uint16)
    else:
        d = d0.astype(np.float32)
        raw = d
    # decoded start f = w_q + d (binary32 addition, the storage's decode rule)
    f0 = (tb.W[:, None].astype(np.float32) + d.astype(np.float32)).astype(np.float64)
    o = orc.DenseOracle((N, N, N), storage=st, compute=orc.COMPUTE_F32, periodic=7)
    o.set_classes(np.full(N ** 3, orc.BULK, np.uint32))
    o.pop[0][:] = raw.reshape(-1)
    o.pop[1][:] = raw.reshape(-1)
    fluid = np.ones((N, N, N), bool)
    ref = f0.reshape(19, N, N, N).copy()
    for _ in range(10):
        ref = tb.step(ref, fluid, np.zeros((3, N, N, N)), om, g=g)
    o.run(10, om, force=g)
    got = o.decoded().reshape(19, N, N, N)
    err = np.abs(got - ref).max()
    assert err < tol, err
    # closed form: momentum of the post-collision populations gains rho*g per step
    mom = lambda f: tb.C.T.astype(float) @ f.reshape(19, -1).sum(1)
    gain = mom(got) - mom(f0)
    want = 10 * f0.sum() * np.array(g)
    assert np.abs(gain - want).max() < mom_rtol * np.abs(want).max(), (gain, want)
    assert abs(got.sum() - f0.sum()) < 50 * tol * N ** 3
