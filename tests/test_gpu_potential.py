"""Softened potential through the plain walk on the GPU against the oracle's
(same visits, same summation order, IEEE sqrt and divide): to 1e-12."""

import numpy as np
import pytest

from oracle import oracle as O
from bhut_amd.workloads import plummer

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def native():
    from bhut_amd import _native
    if _native.device_count() < 1:
        pytest.skip("no HIP device")
    return _native


@pytest.mark.parametrize("theta", [0.0, 0.5, 0.8])
def test_potential_matches_oracle(native, theta):
    x, m = plummer(20_000, seed=7)
    rng = np.random.default_rng(2)
    tg = rng.uniform(-3, 3, (2000, 3))
    for leaf in (8, 32):
        dt = native.DeviceTree.build(x, m, leaf)
        ot = O.build(x, m, leaf)
        np.testing.assert_allclose(dt.potential(None, theta, 0.01, 1.0),
                                   O.potential(ot, x, m, softening=0.01, theta=theta), rtol=1e-12)
        np.testing.assert_allclose(dt.potential(tg, theta, 0.01, 1.0),
                                   O.potential(ot, x, m, tg, softening=0.01, theta=theta), rtol=1e-12)


def test_potential_two_dimensional_and_zero_softening(native):
    rng = np.random.default_rng(3)
    x = rng.standard_normal((5000, 2))
    m = rng.uniform(0.5, 1.5, 5000)
    x[10] = x[11]  # coincident pair: zero softened distance contributes nothing
    dt = native.DeviceTree.build(x, m, 16)
    ot = O.build(x, m, 16)
    np.testing.assert_allclose(dt.potential(None, 0.5, 0.0, 1.0), O.potential(ot, x, m, softening=0.0, theta=0.5),
                               rtol=1e-12)


def test_potential_api_and_forest(native):
    import bhut_amd
    x, m = plummer(5000, seed=4)
    phi = bhut_amd.potential(x, m, theta=0.5, softening=0.01)
    ref = O.potential(O.build(x, m, 32), x, m, softening=0.01, theta=0.5)
    np.testing.assert_allclose(phi, ref, rtol=1e-12)
    t = bhut_amd.Tree(x, m).build()
    np.testing.assert_allclose(t.potential(theta=0.5, softening=0.01), ref, rtol=1e-12)
    # a forest of two systems: each system's potential as if alone
    y, my = plummer(3000, seed=5)
    f = native.DeviceTree.build_forest(np.concatenate([x, y]), np.concatenate([m, my]), [5000, 3000], 32)
    pf = f.potential(None, 0.5, 0.01, 1.0)
    np.testing.assert_allclose(pf[:5000], ref, rtol=1e-12)
    np.testing.assert_allclose(pf[5000:], O.potential(O.build(y, my, 32), y, my, softening=0.01, theta=0.5),
                               rtol=1e-12)
