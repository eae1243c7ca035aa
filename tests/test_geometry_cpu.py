"""nx_geometry_spheres against an independent numpy restatement of its
definition (SplitMix64 stream per Steele, Lea & Flood 2014; sphere test
|v - c| < r; periodic axes across the wrap)."""
import numpy as np

M64 = (1 << 64) - 1


def splitmix(seed):
    s = seed
    while True:
        s = (s + 0x9E3779B97F4A7C15) & M64
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        yield (z ^ (z >> 31)) & M64


def reference(dim, n, rmin, rmax, seed, periodic):
    X, Y, Z = dim
    g = splitmix(seed)
    u = lambda: (next(g) >> 11) * (1.0 / 9007199254740992.0)
    zz, yy, xx = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    solid = np.zeros((Z, Y, X), bool)
    for _ in range(n):
        cx, cy, cz = u() * X, u() * Y, u() * Z
        r = rmin + (rmax - rmin) * u()
        dx, dy, dz = xx - cx, yy - cy, zz - cz
        if periodic & 1:
            dx = dx - X * np.round(dx / X)
        if periodic & 2:
            dy = dy - Y * np.round(dy / Y)
        if periodic & 4:
            dz = dz - Z * np.round(dz / Z)
        solid |= dx * dx + dy * dy + dz * dz < r * r
    return solid.reshape(-1).astype(np.uint8)


def test_splitmix_known_answer():
    # first outputs of SplitMix64 seeded with 0 (published test vector)
    g = splitmix(0)
    assert next(g) == 0xE220A8397B1DCDAF
    assert next(g) == 0x6E789E6AA1B965F4


def test_spheres_match_reference():
    import neon_amd as nx
    for dim, per in [((40, 24, 32), 1), ((17, 33, 21), 0), ((24, 24, 24), 7)]:
        got = nx.spheres(dim, 25, 2.0, 6.0, seed=123, periodic=per)
        ref = reference(dim, 25, 2.0, 6.0, 123, per)
        assert np.array_equal(got, ref), (dim, per, int((got != ref).sum()))
        assert 0 < got.sum() < got.size


def cavity_reference(dim, n_boxes, emin, emax, seed):
    X, Y, Z = dim
    c = np.full((Z, Y, X), 2, np.uint8)
    c[0, :, :] = c[-1, :, :] = 0
    c[:, 0, :] = 0
    c[:, :, 0] = c[:, :, -1] = 0
    c[:, -1, :] = 1
    g = splitmix(seed)
    u = lambda: (next(g) >> 11) * (1.0 / 9007199254740992.0)
    pick = lambda lo, hi: lo + int(u() * (hi - lo + 1))
    for _ in range(n_boxes):
        e = [pick(emin, emax) for _ in range(3)]
        o = [pick(1, max(1, d - 1 - ea)) for d, ea in zip(dim, e)]
        c[o[2]:min(o[2] + e[2], Z - 1), o[1]:min(o[1] + e[1], Y - 1),
          o[0]:min(o[0] + e[0], X - 1)] = 0
    return c.reshape(-1)


def test_cavity_geometry_matches_reference():
    import neon_amd as nx
    for dim, nb, seed in [((32, 40, 24), 4, 7), ((20, 20, 20), 0, 1), ((17, 9, 33), 3, 99)]:
        got = nx.cavity_classes(dim, nb, 4, 16, seed)
        ref = cavity_reference(dim, nb, 4, 16, seed)
        assert np.array_equal(got, ref), (dim, int((got != ref).sum()))
