"""Pin the CPU oracle (oracle/lbm_oracle.c) before trusting it as the checker.

Each test is chosen so that a plausible mistake in the oracle (a wrong
velocity, weight or opposite index, a dropped moving-wall term, a sign error
in the equilibrium, a wrong pull direction) fails it:
  * closed forms of the D3Q19 lattice,
  * binary16 rounding vs numpy's IEEE float16 conversion,
  * a numpy textbook collide-and-stream implementation (tests/textbook.py)
    written from the textbook formulas, compared at round-off tolerance,
  * invariants of the cavity: mass conservation, z-mirror symmetry,
  * Guo forcing: exact momentum gain in a fully periodic box.
"""
import numpy as np
import pytest

import textbook as tb


def test_lattice_closed_forms(orc):
    c, w, opp = orc.lattice()
    assert np.array_equal(c, tb.C) and np.allclose(w, tb.W, rtol=0, atol=0)
    assert abs(w.sum() - 1.0) < 1e-15
    assert np.abs((w[:, None] * c).sum(0)).max() < 1e-16
    second = np.einsum("q,qa,qb->ab", w, c, c)
    assert np.allclose(second, np.eye(3) / 3, atol=1e-16)
    # fourth-order isotropy of D3Q19: sum w c_a^2 c_b^2 = 1/9 (a != b), 1/3 (a == b)
    four = np.einsum("q,qa,qb->ab", w, c.astype(float) ** 2, c.astype(float) ** 2)
    assert np.allclose(four, np.full((3, 3), 1 / 9) + np.eye(3) * (1 / 3 - 1 / 9), atol=1e-16)
    for q in range(19):  # D3Q19.h:84-94 self-check
        assert np.array_equal(c[opp[q]], -c[q])
    assert opp[9] == 9 and list(opp[:9]) == list(range(10, 19))


def test_f16_rounding_matches_ieee(orc):
    rng = np.random.default_rng(0)
    vals = np.concatenate([
        rng.standard_normal(20000).astype(np.float32) * 1e-3,
        rng.standard_normal(20000).astype(np.float32) * 10,
        (rng.random(5000).astype(np.float32) * 2 - 1) * 6.2e-5,      # subnormal range
        np.array([65504, 65519.99, 65520, 1e9, -1e9, 0.0, -0.0, 2 ** -25, 2 ** -24,
                  3 * 2 ** -26, 1 / 18, 1 / 36, 1 / 3], np.float32),
    ]).astype(np.float32)
    ref = vals.astype(np.float16).view(np.uint16)
    got = np.array([orc.lib.or_f32_to_f16(float(v)) for v in vals], np.uint16)
    assert np.array_equal(got, ref)
    back = np.array([orc.lib.or_f16_to_f32(int(h)) for h in ref], np.float32)
    assert np.array_equal(back, ref.view(np.float16).astype(np.float32))


def _cavity_textbook(N, ulb, omega, steps):
    fluid = np.zeros((N, N, N), bool)
    fluid[1:-1, 1:-1, 1:-1] = True
    wall_u = np.zeros((3, N, N, N))
    wall_u[0, :, N - 1, :] = ulb          # lid y = N-1 moves along +x
    post = np.where(fluid[None], tb.W.reshape(19, 1, 1, 1), 0.0)
    for _ in range(steps):
        post = tb.step(post, fluid, wall_u, omega)
    return post, fluid


@pytest.mark.parametrize("N,steps", [(6, 7), (9, 23)])
def test_oracle_matches_textbook_cavity(orc, N, steps):
    ulb, omega = 0.05, 1.7
    o = orc.DenseOracle((N, N, N))
    o.setup_cavity(ulb)
    o.run(steps, omega)
    got = o.fin.reshape(19, N, N, N)
    ref, fluid = _cavity_textbook(N, ulb, omega, steps)
    err = np.abs(got - ref)[:, fluid].max()
    assert err < 5e-15, err


def test_oracle_matches_textbook_obstacle_and_open_links(orc):
    """Random solid cells inside a periodic box: exercises every wall-link
    combination of the mask, and the periodic wrap."""
    rng = np.random.default_rng(3)
    N = 10
    solid = rng.random((N, N, N)) < 0.2
    fluid = ~solid
    cls = np.where(solid, orc.BOUNCEBACK, orc.BULK).astype(np.uint32)
    o = orc.DenseOracle((N, N, N), periodic=7)
    o.set_classes(cls.reshape(-1))
    post0 = o.decoded().reshape(19, N, N, N).copy()
    rng2 = np.random.default_rng(4)
    pert = post0 * (1 + 0.05 * rng2.standard_normal(post0.shape)) * fluid[None]
    o.pop[0][:] = pert.reshape(-1)
    o.pop[1][:] = pert.reshape(-1)
    o.run(9, 1.3)
    ref = pert
    for _ in range(9):
        ref = tb.step(ref, fluid, np.zeros((3, N, N, N)), 1.3)
    err = np.abs(o.fin.reshape(19, N, N, N) - ref)[:, fluid].max()
    assert err < 5e-15, err


def test_oracle_guo_force_matches_textbook_and_momentum(orc):
    N = 6
    rng = np.random.default_rng(5)
    g = (1e-3, -2e-4, 5e-4)
    o = orc.DenseOracle((N, N, N), periodic=7)
    o.set_classes(np.full(N ** 3, orc.BULK, np.uint32))
    f0 = (tb.W.reshape(19, 1, 1, 1) * (1 + 0.01 * rng.standard_normal((19, N, N, N))))
    o.pop[0][:] = f0.reshape(-1)
    o.pop[1][:] = f0.reshape(-1)
    fluid = np.ones((N, N, N), bool)
    ref = f0.copy()
    for _ in range(5):
        ref = tb.step(ref, fluid, np.zeros((3, N, N, N)), 1.1, g=g)
    o.run(5, 1.1, force=g)
    got = o.fin.reshape(19, N, N, N)
    assert np.abs(got - ref).max() < 5e-15
    # Guo: momentum of post-collision populations grows by rho*g per step
    mom = lambda f: tb.C.T.astype(float) @ f.reshape(19, -1).sum(1)
    rho_tot = f0.sum()
    o2 = orc.DenseOracle((N, N, N), periodic=7)
    o2.set_classes(np.full(N ** 3, orc.BULK, np.uint32))
    o2.pop[0][:] = f0.reshape(-1)
    o2.pop[1][:] = f0.reshape(-1)
    m0 = mom(f0)
    o2.run(3, 1.1, force=g)
    m3 = mom(o2.fin.reshape(19, N, N, N))
    assert np.allclose(m3 - m0, 3 * rho_tot * np.array(g), rtol=1e-10, atol=1e-15)


def test_collision_fixed_point_and_conservation(orc):
    """Rest equilibrium is a fixed point; a random state conserves mass and
    momentum across one step in a periodic box (streaming conserves, the
    collision conserves locally)."""
    N = 5
    o = orc.DenseOracle((N, N, N), periodic=7)
    o.set_classes(np.full(N ** 3, orc.BULK, np.uint32))
    o.run(3, 1.9)
    f = o.fin.reshape(19, -1)
    assert np.abs(f - tb.W[:, None]).max() < 1e-15  # a few ulp
    rng = np.random.default_rng(1)
    f0 = tb.W[:, None] * (1 + 0.1 * rng.standard_normal((19, N ** 3)))
    o.pop[0][:] = f0.reshape(-1)
    o.pop[1][:] = f0.reshape(-1)
    o.parity = 0
    o.run(1, 1.6)
    f1 = o.fin.reshape(19, -1)
    assert abs(f1.sum() - f0.sum()) < 1e-12
    mom = lambda f: tb.C.T.astype(float) @ f.sum(1)
    assert np.abs(mom(f1) - mom(f0)).max() < 1e-13


def test_cavity_mass_conservation_and_mirror_symmetry(orc):
    N = 16
    o = orc.DenseOracle((N, N, N))
    o.setup_cavity(0.05)
    m0 = o.mass()
    o.run(200, 1.68)
    assert abs(o.mass() - m0) / m0 < 1e-13
    rho, u = o.moments()
    inner = (slice(1, -1),) * 3
    # z-mirror: u_x, u_y even, u_z odd
    assert np.abs(u[0] - u[0][::-1]).max() < 1e-14
    assert np.abs(u[1] - u[1][::-1]).max() < 1e-14
    assert np.abs(u[2] + u[2][::-1]).max() < 1e-14
    assert u[0][inner].max() > 1e-3   # flow is driven


def test_wall_mask_rule(orc):
    """bit q <=> the upstream cell x - c_q is not bulk (LbmTools.h:327-377)."""
    N = 5
    o = orc.DenseOracle((N, N, N))
    o.setup_cavity(0.05)
    fl = o.flags.reshape(N, N, N)
    c, _, _ = orc.lattice()
    for z, y, x in [(1, 1, 1), (2, 2, 2), (1, 2, 3), (3, 3, 3)]:
        m = int(fl[z, y, x]) & orc.MASK_BITS
        for q in range(19):
            up = (x - c[q][0], y - c[q][1], z - c[q][2])
            wall = q != 9 and not all(1 <= v <= N - 2 for v in up)
            assert bool(m >> q & 1) == wall, (x, y, z, q)
    assert (fl[2, 2, 2] & orc.MASK_BITS) == 0


def test_deviation_storage_close_to_f64(orc):
    """FP32/FP16 deviation storage and the REF float path stay close to FP64."""
    N, steps = 12, 60
    ref = orc.DenseOracle((N, N, N))
    ref.setup_cavity(0.05)
    ref.run(steps, 1.7)
    _, u64 = ref.moments()
    for st, cm, tol in [(orc.F32, orc.COMPUTE_F32, 1e-5), (orc.F32, orc.COMPUTE_REF, 1e-5),
                        (orc.F32, orc.COMPUTE_F64, 1e-5), (orc.F32DEV, orc.COMPUTE_F32, 1e-5),
                        (orc.F16DEV, orc.COMPUTE_F32, 3e-3)]:
        o = orc.DenseOracle((N, N, N), storage=st, compute=cm)
        o.setup_cavity(0.05)
        o.run(steps, 1.7)
        _, u = o.moments()
        rel = np.abs(u.astype(np.float64) - u64).max() / np.abs(u64).max()
        assert rel < tol, (st, cm, rel)
