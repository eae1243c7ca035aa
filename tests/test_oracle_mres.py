"""The two-level grid-refinement oracle (oracle/mres_oracle.c, restating
apps/lbmMultiRes for depth 2 with BGK D3Q19) pinned without the reference:

* grid: the library's builder (nx_mgrid_create, host code) and the oracle
  agree voxel for voxel, and the structure has the properties the mGrid
  constructor guarantees -- every level-0 position is covered by exactly one
  leaf, a level-1 voxel is refined iff a child is active, and a refined
  voxel is kept iff one of its in-domain neighbours is not refined;
* step: with every voxel refined to level 0 the scheme is the uniform
  collide-then-stream LBM, so its velocities equal the uniform dense oracle's
  to round-off; a fluid at rest stays at rest across the interface (the
  coalescence weights 1 / (2 n) average the fine contributions exactly); and
  the two-level cavity stays within a few percent of the uniform fine one
  (a misplaced explosion / coalescence direction or relaxation time moves
  it by tens of percent).
"""
import numpy as np
import pytest

import neon_amd.mgrid as mg


def _uniform_velocity(orc, f, dims):
    c, _, _ = orc.lattice()
    f = f.reshape(19, -1).astype(np.float64)
    rho = f.sum(0)
    u = (c.T.astype(np.float64) @ f) / np.where(rho == 0, 1.0, rho)
    return u.reshape(3, dims[2], dims[1], dims[0])


GEOMS = [((48, 48, 48), 1 / 3), ((37, 30, 41), 0.3), ((64, 64, 64), 0.25), ((20, 26, 18), 0.45)]


@pytest.mark.parametrize("dim,a", GEOMS)
def test_builder_matches_oracle(orc, dim, a):
    lam0, lam1 = mg.cavity_lambdas(dim, (0.0, -a, -1.0))
    g = mg.MGrid(dim, lam0, lam1, "f64")
    s0, s1 = g.state()
    m = orc.MresOracle(dim, lam0, lam1, orc.F64)
    assert np.array_equal(s0 & 3, m.st0 & 3)
    assert np.array_equal(s1 & 3, m.st1 & 3)
    info = g.info()
    assert info["active"] == [int(m.active(0).sum()), int(m.active(1).sum())]


@pytest.mark.parametrize("dim,a", GEOMS)
def test_grid_structure(orc, dim, a):
    lam0, lam1 = mg.cavity_lambdas(dim, (0.0, -a, -1.0))
    m = orc.MresOracle(dim, lam0, lam1, orc.F64)
    X, Y, Z = dim
    X1, Y1, Z1 = m.dim1
    a0 = m.active(0).reshape(Z, Y, X)
    a1 = m.active(1).reshape(Z1, Y1, X1)
    r1 = m.refined().reshape(Z1, Y1, X1)
    assert a0.any() and (a1 & ~r1).any() and r1.any()
    # children of every level-1 position (padded to even sizes)
    pad = np.zeros((2 * Z1, 2 * Y1, 2 * X1), bool)
    pad[:Z, :Y, :X] = a0
    child_any = pad.reshape(Z1, 2, Y1, 2, X1, 2).any(axis=(1, 3, 5))
    child_all = pad.reshape(Z1, 2, Y1, 2, X1, 2).all(axis=(1, 3, 5))
    inside = np.zeros_like(pad)
    inside[:Z, :Y, :X] = True
    # level-0 activity comes in whole 2^3 groups (within the domain)
    grp_inside = inside.reshape(Z1, 2, Y1, 2, X1, 2)
    assert np.array_equal(child_any, (pad | ~inside).reshape(Z1, 2, Y1, 2, X1, 2).all(axis=(1, 3, 5)) & child_any)
    assert (~child_any | child_all | ~grp_inside.all(axis=(1, 3, 5))).all()
    # refined <=> active with an active child
    assert np.array_equal(r1, a1 & child_any)
    # every level-0 position: exactly one leaf (a level-0 voxel or the
    # footprint of an unrefined level-1 voxel)
    coarse_leaf = np.repeat(np.repeat(np.repeat(a1 & ~r1, 2, 0), 2, 1), 2, 2)[:Z, :Y, :X]
    assert np.array_equal(a0.astype(int) + coarse_leaf.astype(int), np.ones((Z, Y, X), int))
    # overlap culling: refined positions kept iff an in-domain neighbour is unrefined
    ref = np.pad(child_any, 1, constant_values=True)
    unrefined_nbr = np.zeros_like(child_any)
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                if dx or dy or dz:
                    unrefined_nbr |= ~ref[1 + dz:1 + dz + Z1, 1 + dy:1 + dy + Y1, 1 + dx:1 + dx + X1]
    assert np.array_equal(r1, child_any & unrefined_nbr)


@pytest.mark.parametrize("storage,tol", [("F64", 2e-8), ("F32", 1e-6)])
def test_rest_state_is_kept(orc, storage, tol):
    """No lid (ulb = 0): every bulk leaf population stays w_q through 20
    coarse steps, across explosion and coalescence.  The bound is the
    rounding of the coalescence weights, which the reference keeps in
    binary32 for binary64 fields too (1/12 is 1 + 3e-8 of itself; measured
    deviation 4.4e-9 of max w_q in binary64)."""
    dim = (32, 28, 36)
    lam0, lam1 = mg.cavity_lambdas(dim, (0.0, -0.3, -1.0))
    m = orc.MresOracle(dim, lam0, lam1, getattr(orc, storage))
    m.setup_cavity(0.0)
    m.run(20, 1.6)
    _, w, _ = orc.lattice()
    for lvl in (0, 1):
        act = m.active(lvl) & (m.classes(lvl) == orc.BULK)
        if lvl:
            act &= ~m.refined()
        f = m.fin[lvl].reshape(19, -1)[:, act].astype(np.float64)
        assert act.sum() > 100
        assert np.abs(f - w[:, None]).max() <= tol * w.max(), lvl


def test_all_fine_equals_uniform(orc):
    """Level 0 everywhere (level 1 empty after culling): T coarse steps are
    2T uniform collide-then-stream steps; the velocities (pre-collision here,
    post-collision in the uniform oracle: BGK conserves momentum) agree to
    round-off."""
    N, T, re, ulb = 24, 60, 20.0, 0.04
    lam0, lam1 = mg.cavity_lambdas((N, N, N), (0.0, -5.0, -6.0))
    m = orc.MresOracle((N, N, N), lam0, lam1, orc.F64)
    assert m.active(0).all() and not m.active(1).any()
    m.setup_cavity(ulb)
    om0 = 1.0 / (3.0 * ulb * m.dim1[0] / re + 0.5)
    m.run(T, om0)
    u = orc.DenseOracle((N, N, N), orc.F64, orc.COMPUTE_F64)
    u.setup_cavity(ulb)
    u.run(2 * T, orc.MresOracle.omega(om0, 0))
    bulk = ((u.flags >> 24) & 3).reshape(N, N, N) == orc.BULK
    um = _uniform_velocity(orc, m.fin[0], (N, N, N))
    uu = _uniform_velocity(orc, u.fin, (N, N, N))
    d = np.abs(um - uu)[:, bulk].max()
    assert d < 1e-14, d
    assert np.abs(uu[:, bulk]).max() > 0.01


def test_two_levels_close_to_uniform(orc):
    """32^3 cavity, a level-0 shell of about a third of the half width, Re 20,
    200 coarse steps against 400 uniform steps at level-0 resolution: largest
    velocity difference 2.7% of the lid speed on level-0 leaves (next to the
    interface), 1.6% on level-1 leaves, RMS 0.4% / 0.7%; mass changes by 2e-5
    relative (measured).  A weight 1 / n instead of 1 / (2 n), an explosion
    from the voxel's own parent or a level-0 relaxation equal to level 1's
    each multiply these differences several times."""
    N, T, re, ulb = 32, 200, 20.0, 0.04
    lam0, lam1 = mg.cavity_lambdas((N, N, N), (0.0, -0.34, -1.0))
    m = orc.MresOracle((N, N, N), lam0, lam1, orc.F64)
    m.setup_cavity(ulb)
    Nc = m.dim1[0]
    om0 = 1.0 / (3.0 * ulb * Nc / re + 0.5)
    m.run(T, om0)
    u = orc.DenseOracle((N, N, N), orc.F64, orc.COMPUTE_F64)
    u.setup_cavity(ulb)
    u.run(2 * T, orc.MresOracle.omega(om0, 0))
    uu = _uniform_velocity(orc, u.fin, (N, N, N))
    u0 = _uniform_velocity(orc, m.fin[0], (N, N, N))
    u1 = _uniform_velocity(orc, m.fin[1], m.dim1)
    leaf0 = (m.active(0) & (m.classes(0) == orc.BULK)).reshape(N, N, N)
    leaf1 = (m.active(1) & ~m.refined()).reshape(Nc, Nc, Nc)
    uc = uu.reshape(3, Nc, 2, Nc, 2, Nc, 2).mean(axis=(2, 4, 6))
    d0 = np.sqrt(((u0 - uu) ** 2).sum(0))[leaf0] / ulb
    d1 = np.sqrt(((u1 - uc) ** 2).sum(0))[leaf1] / ulb
    print(f"two levels vs uniform: max {d0.max():.4f} / {d1.max():.4f}, "
          f"rms {np.sqrt((d0 ** 2).mean()):.4f} / {np.sqrt((d1 ** 2).mean()):.4f}")
    assert d0.max() < 0.045 and d1.max() < 0.03
    assert np.sqrt((d0 ** 2).mean()) < 0.01 and np.sqrt((d1 ** 2).mean()) < 0.015
    bulk_u = ((u.flags >> 24) & 3) == orc.BULK
    mass_u = u.fin.reshape(19, -1)[:, bulk_u].sum()
    mass_m = (m.fin[0].reshape(19, -1)[:, leaf0.reshape(-1)].sum()
              + 8 * m.fin[1].reshape(19, -1)[:, leaf1.reshape(-1)].sum())
    assert abs(mass_m - mass_u) / mass_u < 1e-4
