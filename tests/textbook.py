"""Independent textbook D3Q19 BGK lattice-Boltzmann step in numpy.

Used only to pin the oracle.  It is written from the textbook formulation
(Krueger et al., "The Lattice Boltzmann Method", 2017, ch. 3 and 5), not from
the oracle's expression tree:

  * collide-then-stream with whole-array numpy operations,
  * equilibrium f_eq = w rho (1 + 3 c.u + 9/2 (c.u)^2 - 3/2 u.u),
  * streaming as an explicit shift of every population plane,
  * half-way bounce-back on every link that crosses into a solid cell, with
    the moving-wall term  -6 w_q rho_w (c_q . u_w),  rho_w = 1, u_w taken from
    the wall cell's class.

Arrays are [19, Z, Y, X] (x fastest, the dGrid SoA order).
"""
import numpy as np

C = np.array([[-1, 0, 0], [0, -1, 0], [0, 0, -1], [-1, -1, 0], [-1, 1, 0], [-1, 0, -1],
              [-1, 0, 1], [0, -1, -1], [0, -1, 1], [0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1],
              [1, 1, 0], [1, -1, 0], [1, 0, 1], [1, 0, -1], [0, 1, 1], [0, 1, -1]])
W = np.array([1 / 18] * 3 + [1 / 36] * 6 + [1 / 3] + [1 / 18] * 3 + [1 / 36] * 6)
OPP = np.array([int(np.where((C == -c).all(1))[0][0]) for c in C])


def equilibrium(rho, u):
    cu = np.einsum("qa,a...->q...", C.astype(float), u)
    usq = (u * u).sum(0)
    return W.reshape(19, *([1] * rho.ndim)) * rho * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * usq)


def moments(f):
    rho = f.sum(0)
    with np.errstate(divide="ignore", invalid="ignore"):
        u = np.einsum("qa,q...->a...", C.astype(float), f) / rho
    return rho, u


def step(post, fluid, wall_u, omega, g=None):
    """One step.  `post` holds post-collision populations of the previous step
    (what the pull scheme reads); returns the new post-collision field.

    fluid:  bool [Z,Y,X]  cells that are updated
    wall_u: float [3,Z,Y,X] wall velocity of solid cells (0 for resting walls)
    """
    Z, Y, X = fluid.shape
    f = np.zeros_like(post)
    for q in range(19):
        cx, cy, cz = C[q]
        # streaming: f_q(x) <- post_q(x - c_q); the source is the upstream cell
        src = np.roll(post[q], shift=(cz, cy, cx), axis=(0, 1, 2))
        up_fluid = np.roll(fluid, shift=(cz, cy, cx), axis=(0, 1, 2))
        up_wu = np.roll(wall_u, shift=(cz, cy, cx), axis=(1, 2, 3))
        # link crosses into a solid upstream cell: reflect own opposite population
        o = OPP[q]
        cu_w = (C[o][:, None, None, None] * up_wu).sum(0)
        bounce = post[o] - 6.0 * W[o] * 1.0 * cu_w
        f[q] = np.where(up_fluid, src, bounce)
    rho, u = moments(f)
    u = np.where(fluid[None], u, 0.0)
    if g is None:
        feq = equilibrium(rho, u)
        new = f - omega * (f - feq)
    else:
        g = np.asarray(g, float).reshape(3, 1, 1, 1)
        u = u + 0.5 * g
        feq = equilibrium(rho, u)
        cu = np.einsum("qa,a...->q...", C.astype(float), u)
        cg = (C.astype(float) @ g.reshape(3)).reshape(19, 1, 1, 1)
        ug = (u * g).sum(0)
        src = (1 - 0.5 * omega) * W.reshape(19, 1, 1, 1) * rho * (3 * (cg - ug) + 9 * cu * cg)
        new = f - omega * (f - feq) + src
    return np.where(fluid[None], new, post)
