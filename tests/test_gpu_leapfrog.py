"""Leapfrog driver: graph-captured KDK steps equal the same steps done by
hand through the tree API, bit for bit.

The reference integrates only through scipy (examples/galaxy/
galaxy_simulation.py:277-284), so the checker here is the step written out
with the library's own build / refit / acceleration calls and NumPy kicks:
v += a dt/2, x += v dt, refit (or rebuild on schedule), a = F(x), v += a dt/2.
The kick and drift are single IEEE operations, identical in NumPy and in the
kernels; the tree calls are the ones the parity tests pin to the oracle.
"""

import numpy as np
import pytest

from bhut_amd.workloads import plummer, plummer_velocities

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def native():
    from bhut_amd import _native
    if _native.device_count() < 1:
        pytest.skip("no HIP device")
    return _native


def initial(n, seed=3):
    x, m = plummer(n, seed=seed)
    v = plummer_velocities(n, seed=seed + 1, sigma=0.3)
    return x, v, m


def by_hand(native, x, v, m, steps, dt, rebuild_every, precision, leaf=32, theta=0.5, eps=0.01, opts=None):
    x, v = x.copy(), v.copy()

    def force(t):
        if precision == 64:
            return t.accel_self(theta, eps, 1.0)
        return t.accel_f32(None, theta, eps, 1.0, opts)

    t = native.DeviceTree.build(x, m, leaf)
    a = force(t)
    for k in range(steps):
        v = v + a * (0.5 * dt)
        x = x + v * dt
        if rebuild_every and (k + 1) % rebuild_every == 0:
            t = native.DeviceTree.build(x, m, leaf)
        else:
            t.refit(x, m)
        a = force(t)
        v = v + a * (0.5 * dt)
    return x, v, a


def run_sim(native, x, v, m, steps, dt, rebuild_every, precision, use_graph, leaf=32, caps=None, opts=None):
    import torch
    dev = torch.device("cuda:0")
    tx = torch.from_numpy(x.copy()).to(dev)
    tv = torch.from_numpy(v.copy()).to(dev)
    tm = torch.from_numpy(m.copy()).to(dev)
    torch.cuda.synchronize()
    sim = native.Simulation(tx.data_ptr(), tv.data_ptr(), tm.data_ptr(), len(x), x.shape[1], leaf_size=leaf,
                            theta=0.5, softening=0.01, G=1.0, precision=precision, opts=opts)
    if caps is not None:
        sim.set_list_capacity(*caps)
    sim.run(dt, steps, rebuild_every, use_graph)
    a = torch.empty_like(tx)
    sim.accelerations_dev(a.data_ptr())
    torch.cuda.synchronize()
    return tx.cpu().numpy(), tv.cpu().numpy(), a.cpu().numpy(), sim.info()


def bits(a, b, name):
    np.testing.assert_array_equal(a.view(np.uint64), b.view(np.uint64), err_msg=name)


@pytest.mark.parametrize("precision", [64, 32])
def test_graph_steps_equal_hand_steps(native, precision):
    x, v, m = initial(20_000)
    steps, dt, rb = 12, 1e-3, 4
    hx, hv, ha = by_hand(native, x, v, m, steps, dt, rb, precision)
    for use_graph in (False, True):
        gx, gv, ga, info = run_sim(native, x, v, m, steps, dt, rb, precision, use_graph)
        bits(gx, hx, f"positions graph={use_graph}")
        bits(gv, hv, f"velocities graph={use_graph}")
        bits(ga, ha, f"accelerations graph={use_graph}")
        assert info["steps"] == steps and info["rebuilds"] == 3 and info["refits"] == 9
        if use_graph:
            # refit steps: eager once, captured at the second, replayed after;
            # rebuild steps (k = 3, 7, 11): eager, captured, replayed
            assert info["graph_captures"] == 2
            assert info["graph_launches"] == steps - 2
        else:
            assert info["graph_captures"] == 0 and info["graph_launches"] == 0


def test_list_overflow_is_regrown(native):
    """Capacities far too small: every overflow is caught by the device-side
    check, the force redone eagerly, and the trajectory is unchanged."""
    x, v, m = initial(8000, seed=5)
    steps, dt, rb = 6, 1e-3, 3
    hx, hv, ha = by_hand(native, x, v, m, steps, dt, rb, 32)
    gx, gv, ga, info = run_sim(native, x, v, m, steps, dt, rb, 32, True, caps=(16, 16))
    bits(gx, hx, "positions")
    bits(gv, hv, "velocities")
    assert info["list_regrows"] >= 1
    assert info["list_region"] > 16 and info["item_capacity"] > 16


def test_group128_and_monopole_options(native):
    x, v, m = initial(6000, seed=8)
    opts = native.F32Opts(128, 1, 1024, 1)
    hx, hv, ha = by_hand(native, x, v, m, 5, 1e-3, 2, 32, opts=opts)
    gx, gv, ga, info = run_sim(native, x, v, m, 5, 1e-3, 2, 32, True, opts=opts)
    bits(gx, hx, "positions")
    bits(ga, ha, "accelerations")


def test_energy_drift_small(native):
    from bhut_amd.integrate import Leapfrog
    x, v, m = initial(4000, seed=11)
    for prec in ("fp64", "fp32"):
        lf = Leapfrog(x, v, m, theta=0.5, softening=0.01, precision=prec)
        e0 = lf.energy()
        lf.run(40, 1e-3, rebuild_every=10)
        e1 = lf.energy()
        assert abs(e1 - e0) / abs(e0) < 1e-3, prec


def test_two_dimensional_sim(native):
    rng = np.random.default_rng(4)
    x = rng.standard_normal((3000, 2))
    v = 0.1 * rng.standard_normal((3000, 2))
    m = np.full(3000, 1.0 / 3000)
    for prec in (64, 32):
        hx, hv, ha = by_hand(native, x, v, m, 4, 1e-3, 2, prec, leaf=16)
        gx, gv, ga, info = run_sim(native, x, v, m, 4, 1e-3, 2, prec, True, leaf=16)
        bits(gx, hx, f"positions {prec}")
        bits(gv, hv, f"velocities {prec}")
