"""Zou-He boundary cells in the CPU oracle, pinned by what the mathematics
fixes (CPU only).

* One step from a perturbed state: the post-collision populations of a
  density cell sum to the prescribed density, and those of a velocity cell
  carry the prescribed normal velocity (sum c_n f / sum f).  The completion
  f_i = f_opp + feq_i - feq_opp restores density and normal momentum
  exactly and BGK conserves both, so a wrong unknown set, a wrong sign or a
  wrong equilibrium breaks these identities.
* Pressure-driven slab (walls in y, periodic z, density faces in x): the
  steady profile is the Poiseuille parabola for the imposed gradient, and
  the mass flux is the same through every x plane.
* Velocity inlet + density outlet: the steady flux equals the inlet's.
The flow-over-sphere benchmark's own Zou-He path cannot run (its problem
set-up header is absent from the reference tree), so these identities are
the pin; the GPU kernel is then checked against this oracle bit for bit.
"""
import numpy as np
import pytest

orc = pytest.importorskip("oracle.oracle")

CX = np.array([-1, 0, 0, -1, -1, -1, -1, 0, 0, 0, 1, 0, 0, 1, 1, 1, 1, 0, 0])
CY = np.array([0, -1, 0, -1, 1, 0, 0, -1, -1, 0, 0, 1, 0, 1, -1, 0, 0, 1, 1])
CZ = np.array([0, 0, -1, 0, 0, -1, 1, -1, 1, 0, 0, 0, 1, 0, 0, 1, -1, 1, -1])
C3 = np.stack([CX, CY, CZ])


def slab(X, Y, Z, inlet, outlet, st=orc.F64, cm=orc.COMPUTE_F64):
    """Walls (bounce-back) at y = 0 and Y-1, periodic z, Zou-He cells on the
    x = 0 and x = X-1 planes with the given codes."""
    cls = np.full((Z, Y, X), 2, np.uint32)
    cls[:, 0, :] = 0
    cls[:, Y - 1, :] = 0
    zh = np.zeros((Z, Y, X), np.uint32)
    zh[:, 1:Y - 1, 0] = inlet
    zh[:, 1:Y - 1, X - 1] = outlet
    o = orc.DenseOracle((X, Y, Z), st, cm, periodic=4)
    o.set_classes(cls.reshape(-1), zouhe=zh.reshape(-1))
    return o, cls, zh


def perturb(o, seed):
    rng = np.random.default_rng(seed)
    for p in o.pop:
        v = p.reshape(19, -1)
        v *= 1.0 + 0.05 * rng.standard_normal(v.shape)


@pytest.mark.parametrize("face,axis,sign", [(1, 0, 1), (2, 0, -1)])
def test_density_and_velocity_faces_impose_their_value(face, axis, sign):
    X, Y, Z = 9, 7, 3
    rho_p, u_p = 1.013, 0.021
    # density face on this side, velocity face (code + 8) on the other
    other = 2 if face == 1 else 1
    o, cls, zh = slab(X, Y, Z, face if face == 1 else other + 8, face if face == 2 else other + 8)
    perturb(o, 5)
    vals = [0.0] * 6
    vals[face - 1] = rho_p
    vals[other - 1] = u_p
    o.run_zouhe(1, 1.3, vals)
    f = o.fin.reshape(19, -1)
    flags = o.flags
    code = (flags >> orc.ZH_SHIFT) & 15
    dens = code == face
    vel = code == other + 8
    assert dens.sum() == (Y - 2) * Z and vel.sum() == (Y - 2) * Z
    rho = f.sum(axis=0)
    np.testing.assert_allclose(rho[dens], rho_p, rtol=1e-13, atol=0)
    # inward normal of the velocity face: +x for face 1, -x for face 2
    n_in = 1 if other == 1 else -1
    un = (CX[:, None] * f[:, vel]).sum(axis=0) / rho[vel] * n_in
    np.testing.assert_allclose(un, u_p, rtol=1e-11, atol=0)


def test_identities_hold_in_binary32():
    X, Y, Z = 8, 6, 3
    o, _, _ = slab(X, Y, Z, 1, 2 + 8, st=orc.F32, cm=orc.COMPUTE_F32)
    perturb(o, 9)
    o.run_zouhe(1, 1.1, [1.004, -0.01, 0, 0, 0, 0])
    f = o.fin.reshape(19, -1).astype(np.float64)
    code = (o.flags >> orc.ZH_SHIFT) & 15
    rho = f.sum(axis=0)
    np.testing.assert_allclose(rho[code == 1], 1.004, rtol=2e-6)
    vel = code == 10
    un = -(CX[:, None] * f[:, vel]).sum(axis=0) / rho[vel]
    np.testing.assert_allclose(un, -0.01, rtol=1e-4)


def test_pressure_driven_slab_is_poiseuille():
    X, Y, Z = 40, 11, 3
    d_rho = 1e-3
    o, cls, _ = slab(X, Y, Z, 1, 2)
    omega = 1.0
    nu = (1.0 / omega - 0.5) / 3.0
    o.run_zouhe(5000, omega, [1.0 + d_rho, 1.0 - d_rho, 0, 0, 0, 0])
    rho, u = o.moments()
    ux = u[0]
    assert np.isfinite(ux).all()
    # flux through every x plane of the fluid (y = 1..Y-2)
    # (moments are post-stream, so the Zou-He planes x = 0, X-1 are left out)
    flux = (rho * ux)[:, 1:Y - 1, 1:X - 1].sum(axis=(0, 1))
    assert flux.min() > 0
    assert np.abs(flux / flux.mean() - 1).max() < 2e-3
    # mid-channel profile against the parabola for the local gradient; the
    # bounce-back walls sit half way between nodes (y = 0.5 and Y - 1.5)
    xm = X // 2
    grad = -(rho[1, Y // 2, xm + 1] - rho[1, Y // 2, xm - 1]) / 2.0 / 3.0   # -dp/dx, p = rho / 3
    y = np.arange(1, Y - 1)
    ya, yb = 0.5, Y - 1.5
    analytic = grad / (2.0 * nu * rho[1, Y // 2, xm]) * (y - ya) * (yb - y)
    prof = ux[1, 1:Y - 1, xm]
    assert np.abs(prof - analytic).max() < 0.02 * analytic.max()


def test_velocity_inlet_flux_reaches_the_outlet():
    X, Y, Z = 32, 9, 3
    u_in = 2e-3
    o, cls, _ = slab(X, Y, Z, 1 + 8, 2)
    o.run_zouhe(4000, 1.0, [u_in, 1.0, 0, 0, 0, 0])
    rho, u = o.moments()
    flux = (rho * u[0])[:, 1:Y - 1, 1:X - 1].sum(axis=(0, 1))
    assert np.abs(flux / flux.mean() - 1).max() < 5e-3
    # the section-mean velocity downstream is the inlet's (mass conservation)
    mid = u[0][:, 1:Y - 1, X // 2].mean()
    assert abs(mid / u_in - 1) < 0.05
