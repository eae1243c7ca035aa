"""Batched cavities: every member of a batch evolves exactly like the same
cavity stepped on its own (bit for bit), converged members stop stepping at
the step count the batch reports, and the captured hipGraph replays agree
with eager steps."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def make_members(nx, seed=0, sizes=(12, 16, 20, 14), re_range=(50, 400)):
    rng = np.random.default_rng(seed)
    ms = []
    for k, n in enumerate(sizes):
        dim = (n, n + (k % 2) * 4, n - (k % 3))
        re = rng.uniform(*re_range)
        omega = 1.0 / (3.0 * 0.05 * dim[0] / re + 0.5)
        ms.append({"dim": dim, "classes": nx.cavity_classes(dim, k % 3, 2, 5, seed=100 + k),
                   "omega": omega, "ulb": 0.05, "ulid": (1.0, 0.0, 0.0)})
    return ms


def standalone(nx, m, st, steps):
    g = nx.DenseLattice(m["dim"], st)
    g.set_classes(m["classes"], ulb=m["ulb"], ulid=m["ulid"])
    g.run_plain(steps, m["omega"])
    return g


@pytest.mark.parametrize("st", ["f32", "f64", "f16dev", "f32dev"])
@pytest.mark.parametrize("launch", [(256, 4, 1, 0), (512, 0, 0, 1), (256, 16, 0, 1), (256, 2, 1, 1),
                                    (256, 0, 1, 1), (512, 0, 1, 0)])
def test_batch_members_equal_standalone(nx, st, launch):
    """Every launch (wg, workgroups per CU, XCD chunks, edge speculation):
    the members' populations equal the standalone single-kernel step's."""
    ms = make_members(nx)
    b = nx.CavityBatch(ms, st)
    b.set_launch(*launch[:2])
    b.set_variant(xcd_chunks=launch[2], edge=launch[3])
    lc = b.launch()
    assert lc["edge"] == launch[3] and lc["xcd_chunks"] == launch[2]
    assert lc["wgs_per_cu"] == launch[1]
    if launch[1] == 0:                      # one workgroup per tile
        assert lc["grid"] == b.tiles() and lc["chunked"] == (2 if launch[2] else 0)
    b.init()
    b.step(14)
    for k, m in enumerate(ms):
        ref = standalone(nx, m, st, 14)
        assert np.array_equal(b.populations(k).view(np.uint8), ref.populations().view(np.uint8)), k


def test_batch_run_retires_converged_members(nx):
    ms = make_members(nx, seed=3, sizes=(10, 12, 14), re_range=(5, 30))
    b = nx.CavityBatch(ms, "f64")
    b.init()
    steps, res, left = b.run(max_steps=20000, check_every=20, tol=1e-5)
    print("steps", steps, "residual", res)
    assert left == 0 and (res < 1e-5).all() and (steps > 0).all()
    assert len(set(steps.tolist())) > 1          # different members stop at different times
    for k, m in enumerate(ms):
        ref = standalone(nx, m, "f64", int(steps[k]))
        assert np.array_equal(b.populations(k), ref.populations()), k
    # graph replays after the eager steps above: run again from rest
    b.init()
    steps2, res2, left2 = b.run(max_steps=20000, check_every=20, tol=1e-5)
    assert np.array_equal(steps, steps2) and np.array_equal(res, res2) and left2 == 0


def test_batch_warm_start_from_converged_fields(nx):
    # FP32 deviation storage: its residual floor (~2e-7 here) is below the
    # tolerance; raw FP32 populations stall near 1e-5 on such small cavities
    ms = make_members(nx, seed=5, sizes=(12, 12), re_range=(5, 30))
    b = nx.CavityBatch(ms, "f32dev")
    b.init()
    steps, _, left0 = b.run(max_steps=20000, check_every=20, tol=1e-6)
    assert left0 == 0
    # gather converged (rho, u) in arena order and restart from them
    rho = []
    u = []
    for k in range(len(ms)):
        r, uu = b.moments(k)
        rho.append(r.reshape(-1))
        u.append(uu.reshape(-1))
    drho = nx.DeviceBuffer.from_array(np.concatenate(rho).astype(np.float32))
    du = nx.DeviceBuffer.from_array(np.concatenate(u).astype(np.float32))
    # an equilibrium start alone lacks the non-equilibrium part and relaxes
    # almost as slowly as a start from rest; 10 constrained double steps
    # restore a consistent state
    b.init(drho, du)
    b.constrain(du, 10)
    steps_w, res_w, left = b.run(max_steps=20000, check_every=20, tol=1e-6)
    print("cold", steps, "warm", steps_w)
    assert left == 0
    assert (steps_w + 20 < steps / 2).all(), (steps, steps_w)


def test_batch_variant_change_recaptures_graph(nx):
    """nx_batch_run after nx_batch_set_variant runs the new variant: results
    stay those of the standalone step and the grid reported is the new one."""
    ms = make_members(nx, seed=7, sizes=(12, 14))
    b = nx.CavityBatch(ms, "f32")
    b.set_launch(256, 4)
    b.set_variant(xcd_chunks=0, edge=0)
    b.init()
    b.run(max_steps=20, check_every=10, tol=1e-30)
    b.set_variant(xcd_chunks=0, edge=1)
    b.set_launch(256, 0)
    assert b.launch()["grid"] == b.tiles()
    b.init()
    steps, _, _ = b.run(max_steps=20, check_every=10, tol=1e-30)
    assert (steps == 20).all()
    for k, m in enumerate(ms):
        assert np.array_equal(b.populations(k), standalone(nx, m, "f32", 20).populations()), k
