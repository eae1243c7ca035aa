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


# ---- more than two levels ------------------------------------------------------

DEEP = [((48, 48, 48), mg.LID_SCALES[0][1]),              # the app's --scale 0: 48^3, depth 3
        ((64, 64, 64), mg.LID_SCALES[6][1]),              # --scale 6's level bands at 64^3, depth 4
        ((37, 30, 41), (0.0, -0.25, -0.5, -1.0)),
        ((50, 44, 40), (0.0, -0.2, -0.4, -0.7, -1.0))]


def _cover_levels(m):
    """Per level l, for every level-l position: the lowest level >= l whose
    voxel covering it is active (depth when none is)."""
    out = [None] * m.depth
    for l in range(m.depth - 1, -1, -1):
        d = m.dims[l]
        a = m.active(l).reshape(d[2], d[1], d[0])
        up = np.full(a.shape, m.depth) if l == m.depth - 1 else \
            np.repeat(np.repeat(np.repeat(out[l + 1], 2, 0), 2, 1), 2, 2)[:d[2], :d[1], :d[0]]
        out[l] = np.where(a, l, up)
    return out


def _balance_violations(m):
    """Active voxels of level l with an in-domain neighbour position whose
    finest active cover lies two or more levels up."""
    cover = _cover_levels(m)
    bad = 0
    for l in range(m.depth):
        d = m.dims[l]
        a = m.active(l).reshape(d[2], d[1], d[0])
        c = np.pad(cover[l], 1, constant_values=-1)  # outside the domain: never a violation
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    if dx or dy or dz:
                        n = c[1 + dz:1 + dz + d[2], 1 + dy:1 + dy + d[1], 1 + dx:1 + dx + d[0]]
                        bad += int((a & (n > l + 1) & (n < m.depth)).sum())
    return bad


@pytest.mark.parametrize("dim,sdf", DEEP)
def test_deep_builder_matches_oracle(orc, dim, sdf):
    lams = mg.cavity_lambdas(dim, sdf)
    m = orc.MresOracle(dim, list(lams), orc.F64)
    g = mg.MGrid(dim, list(lams), "f64")
    st = g.state()
    for l in range(m.depth):
        assert np.array_equal(st[l] & 3, m.st[l] & 3), l
    assert g.info()["active"] == [int(m.active(l).sum()) for l in range(m.depth)]


@pytest.mark.parametrize("dim,sdf", DEEP)
def test_deep_grid_structure(orc, dim, sdf):
    """Every level-0 position is covered by exactly one leaf, refined <=> an
    active child, kept refined voxels have an unrefined in-domain neighbour,
    and every neighbour position of an active voxel of level l is covered on
    level l or l + 1 (the strong-balance condition)."""
    m = orc.MresOracle(dim, list(mg.cavity_lambdas(dim, sdf)), orc.F64)
    X, Y, Z = dim
    cover = np.zeros((Z, Y, X), int)
    for l in range(m.depth):
        d = m.dims[l]
        a = m.active(l).reshape(d[2], d[1], d[0])
        assert a.any(), l
        leaf = a & ~m.refined(l).reshape(a.shape) if l else a
        s = 1 << l
        cover += np.repeat(np.repeat(np.repeat(leaf, s, 0), s, 1), s, 2)[:Z, :Y, :X]
        if l:
            f = m.dims[l - 1]
            pad = np.zeros((2 * d[2], 2 * d[1], 2 * d[0]), bool)
            pad[:f[2], :f[1], :f[0]] = m.active(l - 1).reshape(f[2], f[1], f[0])
            child = pad.reshape(d[2], 2, d[1], 2, d[0], 2).any(axis=(1, 3, 5))
            assert np.array_equal(m.refined(l).reshape(a.shape), a & child), l
    assert (cover == 1).all()
    assert _balance_violations(m) == 0 and m.balanced == 0


def test_strong_balance_pass(orc):
    """Two isolated fine groups inside a coarse level-2 region with no level-1
    band between them: the pass adds level-1 voxels (38 here) until every
    neighbour of an active voxel is covered one level up at most; the
    library's builder agrees voxel for voxel."""
    dim = (24, 24, 24)
    dims = mg.level_dims(dim, 3)
    lam = [np.zeros(int(np.prod(d)), np.uint8) for d in dims]
    for x, y, z in ((6, 6, 6), (13, 9, 11)):
        lam[0][x + 24 * (y + 24 * z)] = 1
    lam[2][:] = 1
    m = orc.MresOracle(dim, lam, orc.F64)
    assert m.balanced == 38
    assert _balance_violations(m) == 0
    g = mg.MGrid(dim, lam, "f64")
    st = g.state()
    for l in range(3):
        assert np.array_equal(st[l] & 3, m.st[l] & 3), l


@pytest.mark.parametrize("depth", [3, 4])
def test_deep_rest_state_is_kept(orc, depth):
    dim = (40, 36, 44)
    sdf = (0.0, -0.2, -0.45, -1.0) if depth == 3 else (0.0, -0.2, -0.35, -0.6, -1.0)
    m = orc.MresOracle(dim, list(mg.cavity_lambdas(dim, sdf)), orc.F64)
    m.setup_cavity(0.0)
    m.run(6, 1.6)
    _, w, _ = orc.lattice()
    for lvl in range(depth):
        act = m.active(lvl) & (m.classes(lvl) == orc.BULK)
        if lvl:
            act &= ~m.refined(lvl)
        f = m.fin[lvl].reshape(19, -1)[:, act].astype(np.float64)
        assert act.sum() > 0
        assert np.abs(f - w[:, None]).max() <= 2e-8 * w.max(), lvl


def test_deep_all_fine_equals_uniform(orc):
    """Depth 3 with every voxel on level 0: one coarse step is four uniform
    steps with computeOmega(omega0, 0, 3)."""
    N, T, re, ulb = 24, 30, 20.0, 0.04
    m = orc.MresOracle((N, N, N), list(mg.cavity_lambdas((N, N, N), (0.0, -5.0, -6.0, -7.0))), orc.F64)
    assert m.active(0).all() and not m.active(1).any() and not m.active(2).any()
    m.setup_cavity(ulb)
    om0 = 1.0 / (3.0 * ulb * m.dims[2][0] / re + 0.5)
    m.run(T, om0)
    u = orc.DenseOracle((N, N, N), orc.F64, orc.COMPUTE_F64)
    u.setup_cavity(ulb)
    u.run(4 * T, orc.MresOracle.omega(om0, 0, 3))
    bulk = ((u.flags >> 24) & 3).reshape(N, N, N) == orc.BULK
    d = np.abs(_uniform_velocity(orc, m.fin[0], (N, N, N)) - _uniform_velocity(orc, u.fin, (N, N, N)))[:, bulk]
    assert d.max() < 1e-14, d.max()


def test_three_levels_close_to_uniform(orc):
    """The app's --scale 0 grid (48^3, three levels), Re 20, 100 coarse steps
    against 400 uniform steps at level-0 resolution: largest velocity
    difference on the leaves of levels 0 / 1 / 2 measured 1.9% / 1.8% / 2.0%
    of the lid speed, RMS 0.23% / 0.63% / 1.1%.  Two level-0 sub-steps per
    level-1 step instead of four, or level 1 relaxing with level 0's omega,
    move these by tens of percent."""
    N, T, re, ulb = 48, 100, 20.0, 0.04
    m = orc.MresOracle((N, N, N), list(mg.cavity_lambdas((N, N, N), mg.LID_SCALES[0][1])), orc.F64)
    m.setup_cavity(ulb)
    om0 = 1.0 / (3.0 * ulb * m.dims[2][0] / re + 0.5)
    m.run(T, om0)
    u = orc.DenseOracle((N, N, N), orc.F64, orc.COMPUTE_F64)
    u.setup_cavity(ulb)
    u.run(4 * T, orc.MresOracle.omega(om0, 0, 3))
    uu = _uniform_velocity(orc, u.fin, (N, N, N))
    bounds = [(0.03, 0.005), (0.03, 0.01), (0.03, 0.02)]
    for l in range(3):
        d, s = m.dims[l], 1 << l
        ul = _uniform_velocity(orc, m.fin[l], d)
        leaf = (m.active(l) & ((m.st[l] & 2) == 0) & (m.classes(l) == orc.BULK)).reshape(d[2], d[1], d[0])
        pad = np.zeros((3, d[2] * s, d[1] * s, d[0] * s))
        pad[:, :N, :N, :N] = uu
        uc = pad.reshape(3, d[2], s, d[1], s, d[0], s).mean(axis=(2, 4, 6))
        dd = np.sqrt(((ul - uc) ** 2).sum(0))[leaf] / ulb
        print(f"level {l}: {int(leaf.sum())} leaves, max {dd.max():.4f}, rms {np.sqrt((dd ** 2).mean()):.4f}")
        assert dd.max() < bounds[l][0] and np.sqrt((dd ** 2).mean()) < bounds[l][1], l


# ---- KBC D3Q27 (the multi-resolution app's default collision) -----------------

def _d3q27():
    from oracle.oracle import D3Q27_C
    c = D3Q27_C.astype(np.float64)
    n = np.abs(D3Q27_C).sum(1)
    w = np.array([8 / 27, 2 / 27, 1 / 54, 1 / 216])[n]
    return c, w


def test_d3q27_lattice_closed_forms():
    """The D3Q27 order of lattice.h:14-42 (z fastest, then y, then x, each over
    0, -1, +1) with the tensor-product weights: isotropy up to fourth order."""
    c, w = _d3q27()
    assert c.shape == (27, 3) and len({tuple(v) for v in c}) == 27
    assert abs(w.sum() - 1) < 1e-15
    np.testing.assert_allclose(w @ c, 0, atol=1e-16)
    np.testing.assert_allclose(np.einsum("q,qa,qb->ab", w, c, c), np.eye(3) / 3, atol=1e-16)
    m4 = np.einsum("q,qa,qb,qc,qd->abcd", w, c, c, c, c)
    d = np.eye(3)
    want = (np.einsum("ab,cd->abcd", d, d) + np.einsum("ac,bd->abcd", d, d) + np.einsum("ad,bc->abcd", d, d)) / 9
    np.testing.assert_allclose(m4, want, atol=1e-16)


def _equilibrium(c, w, rho, u):
    cu = 3.0 * (u @ c.T)
    return rho[:, None] * w * (1 + cu + 0.5 * cu * cu - 1.5 * (u * u).sum(1)[:, None])


def _shear_part(c, rho, fneq):
    """The KBC shear part from the mathematics: the deviatoric stress of fneq,
    carried by the six face directions (half of dev_aa each) and the twelve
    edge directions (a quarter of Pi_ab each), scaled by rho as the reference
    does (collide.h:196-226, 252-253)."""
    pi = np.einsum("kq,qa,qb->kab", fneq, c, c)
    dev = pi - np.trace(pi, axis1=1, axis2=2)[:, None, None] / 3 * np.eye(3)
    n = np.abs(c).sum(1)
    s = np.zeros_like(fneq)
    for q in range(27):
        if n[q] == 1:
            a = int(np.nonzero(c[q])[0][0])
            s[:, q] = dev[:, a, a] / 2
        elif n[q] == 2:
            a, b = np.nonzero(c[q])[0]
            s[:, q] = c[q, a] * c[q, b] * pi[:, a, b] / 4
    return rho[:, None] * s


def _random_states(k, seed):
    rng = np.random.default_rng(seed)
    c, w = _d3q27()
    rho = 1 + 0.05 * rng.standard_normal(k)
    u = 0.06 * rng.standard_normal((k, 3))
    f = _equilibrium(c, w, rho, u) * (1 + 0.08 * rng.standard_normal((k, 27)))
    return c, w, f


@pytest.mark.parametrize("omega", [0.7, 1.3, 1.95])
def test_kbc27_closed_forms(orc, omega):
    """On random non-equilibrium states (binary64): mass and momentum are
    conserved; out = f - beta (2 dS + gamma dH) with beta = omega / 2, the
    shear part dS of the deviatoric decomposition and dH = fneq - dS, for one
    gamma per voxel; and gamma is the entropic one: the post-collision
    non-equilibrium part is orthogonal to dH in the 1 / feq weighted product,
    sum dH (out - feq) / feq = 0, up to the regularisation tiny = 1e-7 of
    gamma's denominator, which leaves (1 - 2 beta) e0 tiny / (tiny + e1)
    (e0 = sum dS dH / feq, e1 = sum dH^2 / feq)."""
    c, w, f = _random_states(400, 3)
    out = orc.kbc27_collide(f, omega)
    rho = f.sum(1)
    np.testing.assert_allclose(out.sum(1), rho, rtol=1e-14)
    np.testing.assert_allclose(out @ c, f @ c, atol=2e-15)
    u = (f @ c) / rho[:, None]
    feq = _equilibrium(c, w, rho, u)
    fneq = f - feq
    ds = _shear_part(c, rho, fneq)
    dh = fneq - ds
    beta = omega / 2
    # gamma from the output, per voxel, by least squares on the one-parameter form
    r = (f - out) / beta - 2 * ds
    gamma = (r * dh).sum(1) / (dh * dh).sum(1)
    np.testing.assert_allclose(out, f - beta * (2 * ds + gamma[:, None] * dh), atol=1e-13)
    orth = (dh * (out - feq) / feq).sum(1)
    e0, e1 = (ds * dh / feq).sum(1), (dh * dh / feq).sum(1)
    want = (1 - 2 * beta) * e0 * 1e-7 / (1e-7 + e1)
    scale = (np.abs(dh) * np.abs(out - feq) / feq).sum(1)
    assert np.abs((orth - want) / scale).max() < 1e-9
    assert np.abs(want / scale).max() > 1e-7  # the regularisation is visible at this scale
    assert (gamma > 0).all()


def test_kbc27_fixed_point_and_omega_one(orc):
    """An equilibrium state is kept; at omega = 1 (beta = 1/2, gamma = 2
    whatever e0) every state goes to its equilibrium."""
    c, w, f = _random_states(200, 5)
    rho = f.sum(1)
    feq = _equilibrium(c, w, rho, (f @ c) / rho[:, None])
    np.testing.assert_allclose(orc.kbc27_collide(feq, 1.6), feq, atol=2e-16 * 30)
    np.testing.assert_allclose(orc.kbc27_collide(f, 1.0), feq, atol=1e-15)


def test_kbc27_binary32_follows_binary64(orc):
    c, w, f = _random_states(200, 7)
    o64 = orc.kbc27_collide(f, 1.7, orc.F64)
    o32 = orc.kbc27_collide(f.astype(np.float32), 1.7, orc.F32).astype(np.float64)
    assert np.abs(o32 - o64).max() < 2e-6 * np.abs(o64).max()


def test_kbc27_rest_state_is_kept(orc):
    """No lid: every bulk leaf population of the three-level KBC grid stays
    w_q through 6 coarse steps (equilibrium is the collision's fixed point)."""
    dim = (40, 36, 44)
    m = orc.MresOracle(dim, list(mg.cavity_lambdas(dim, (0.0, -0.2, -0.45, -1.0))), orc.F64, lattice=orc.KBC27)
    m.setup_cavity(0.0)
    m.run(6, 1.6)
    _, w = _d3q27()
    for lvl in range(3):
        act = m.active(lvl) & (m.classes(lvl) == orc.BULK)
        if lvl:
            act &= ~m.refined(lvl)
        f = m.fin[lvl].reshape(27, -1)[:, act]
        assert act.sum() > 0
        assert np.abs(f - w[:, None]).max() <= 2e-8 * w.max(), lvl


def test_kbc27_three_levels_close_to_bgk(orc):
    """The app's --scale 0 grid at Re 20, 100 coarse steps: the KBC D3Q27 and
    BGK D3Q19 runs give the same flow to discretisation accuracy -- level-0
    velocities differ by at most 3.0% of the lid speed (at the lid's
    corners), RMS 0.16% (measured) -- and in both the mass of the level-0
    leaves moves by less than 2e-5 relative (measured 5.3e-6 and 6.0e-6:
    the interface exchange is not exactly conservative)."""
    N, T, re, ulb = 48, 100, 20.0, 0.04
    lams = list(mg.cavity_lambdas((N, N, N), mg.LID_SCALES[0][1]))
    u = {}
    for lat in (orc.BGK19, orc.KBC27):
        m = orc.MresOracle((N, N, N), lams, orc.F64, lattice=lat)
        m.setup_cavity(ulb)
        leaf = m.active(0) & (m.classes(0) == orc.BULK)
        m0 = m.fin[0].reshape(m.Q, -1)[:, leaf].sum()
        m.run(T, 1.0 / (3.0 * ulb * m.dims[2][0] / re + 0.5))
        c = _d3q27()[0] if m.Q == 27 else orc.lattice()[0].astype(np.float64)
        f = m.fin[0].reshape(m.Q, -1)
        assert abs(f[:, leaf].sum() - m0) < 2e-5 * m0
        u[lat] = ((c.T @ f) / f.sum(0))[:, leaf]
    d = np.sqrt(((u[orc.BGK19] - u[orc.KBC27]) ** 2).sum(0)) / ulb
    assert d.max() < 0.045 and np.sqrt((d ** 2).mean()) < 0.003


def _random_levels(seed):
    rng = np.random.default_rng(seed)
    depth = int(rng.integers(2, 6))
    dim = tuple(int(v) for v in rng.integers(2, 40, size=3))
    dims = mg.level_dims(dim, depth)
    dens = rng.uniform(0, 0.05, size=depth)
    lams = [(rng.random(int(np.prod(d))) < dens[l]).astype(np.uint8) for l, d in enumerate(dims)]
    lams[-1][:] = 1
    return dim, lams


@pytest.mark.parametrize("seed", range(24))
def test_random_grids_builder_matches_oracle(orc, seed):
    """Scattered activation at depth 2..5 on odd and even boxes (most of them
    make the strong-balance pass add voxels, some cull the top level away):
    the library's builder equals the oracle's voxel for voxel, and the
    result keeps every neighbour of an active voxel covered at most one
    level up."""
    dim, lams = _random_levels(seed)
    m = orc.MresOracle(dim, lams, orc.F64)
    g = mg.MGrid(dim, lams, "f64")
    for l, st in enumerate(g.state()):
        assert np.array_equal(st & 3, m.st[l] & 3), l
    assert _balance_violations(m) == 0


# ---- the flow over a sphere (the app's second problem) -------------------------

def test_sphere_problem_geometry():
    """flowOverShape.h:115-190 at --scale 1: the domain and level boxes, the
    inlet at x = 0, the sphere as the integer points strictly inside radius 8
    around (52, 52, 68) on level 0, and the bounce-back y / z faces on every
    level at 0 and at the level's last voxel."""
    from neon_amd.abi import NX_CELL_BULK, NX_CELL_BOUNCEBACK, NX_CELL_MOVINGWALL
    dim, lams, cls, sph = mg.sphere_problem(1)
    assert dim == (136, 96, 136) and sph == (52, 52, 68, 8)
    d0 = mg.level_dims(dim, 3)
    l0 = lams[0].reshape(d0[0][2], d0[0][1], d0[0][0])
    assert l0.sum() == 56 * 24 * 56 and l0[40:96, 40:64, 40:96].all()
    assert lams[2].all()
    c0 = cls[0].reshape(d0[0][2], d0[0][1], d0[0][0])
    z, y, x = np.mgrid[0:136, 0:96, 0:136]
    ball = (52 - x) ** 2 + (52 - y) ** 2 + (68 - z) ** 2 < 64
    assert ((c0 == NX_CELL_BOUNCEBACK) & ball).sum() == ball.sum() == 2103  # lattice points with r^2 < 64
    assert (c0[1:-1, 1:-1, 0] == NX_CELL_MOVINGWALL).all()
    for l, (c, d) in enumerate(zip(cls, d0)):
        c = c.reshape(d[2], d[1], d[0])
        assert (c[:, 0, :] == NX_CELL_BOUNCEBACK).all() and (c[:, -1, :] == NX_CELL_BOUNCEBACK).all()
        assert (c[0] == NX_CELL_BOUNCEBACK).all() and (c[-1] == NX_CELL_BOUNCEBACK).all()
        assert (c[1:-1, 1:-1, -1] != NX_CELL_BOUNCEBACK).any()  # the outlet face stays open


def test_sphere_flow_kbc_oracle_is_driven_by_the_inlet(orc):
    """KBC D3Q27 on the sphere grid (--scale 1, Re 100, binary64), 60 coarse
    steps: finite everywhere, mass enters (the level-0 leaves gain mass), and
    the flow reaches the sphere's box: the largest level-0 u_x exceeds a tenth
    of the inlet speed."""
    dim, lams, cls, _ = mg.sphere_problem(1)
    uin, om0 = mg.sphere_parameters(1, 100.0)
    m = orc.MresOracle(dim, lams, orc.F64, lattice=orc.KBC27)
    m.setup_cells(cls, (uin, 0.0, 0.0))
    m.run(60, om0)
    c = _d3q27()[0]
    for l in range(3):
        assert np.isfinite(m.fin[l]).all(), l
    leaf = m.active(0) & (m.classes(0) == orc.BULK)
    f = m.fin[0].reshape(27, -1)[:, leaf]
    ux = (c[:, 0] @ f) / f.sum(0)
    assert ux.max() > 0.1 * uin
