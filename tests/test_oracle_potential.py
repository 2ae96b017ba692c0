"""The oracle's softened potential, pinned by closed forms (bhut itself has no
potential; docs/getting-started.md:36 names one).

* two bodies: phi_0 = -G m_1 / sqrt(r^2 + eps^2) and phi_1 = -G m_0 / ...;
* theta = 0 opens every node, so the walk equals the direct sum (a separate
  loop in source order), for self and external targets;
* a ring of N equal masses around an external target at its centre:
  phi = -G M / sqrt(R^2 + eps^2) at theta = 0;
* a planted mistake (the softening dropped) fails the two-body check.
"""

import numpy as np

from oracle import oracle as O


def test_two_body_closed_form():
    x = np.array([[0.0, 0.0, 0.0], [3.0, 4.0, 0.0]])
    m = np.array([2.0, 5.0])
    for eps in (0.0, 0.5):
        t = O.build(x, m, 1)
        phi = O.potential(t, x, m, softening=eps, theta=0.5, G=2.0)
        r = np.sqrt(25.0 + eps * eps)
        np.testing.assert_allclose(phi, [-2.0 * 5.0 / r, -2.0 * 2.0 / r], rtol=1e-15)
        # a planted mistake: without the softening the two-body value differs
        if eps > 0:
            assert not np.allclose(phi, [-2.0 * 5.0 / 5.0, -2.0 * 2.0 / 5.0], rtol=1e-6)


def test_theta_zero_equals_direct_sum():
    rng = np.random.default_rng(5)
    for dim in (2, 3):
        x = rng.standard_normal((1500, dim))
        m = rng.uniform(0.2, 2.0, 1500)
        t = O.build(x, m, 16) if dim == 3 else O.build(x, m, 8)
        phi = O.potential(t, x, m, softening=0.01, theta=0.0)
        ref = O.direct_potential(x, m, softening=0.01)
        np.testing.assert_allclose(phi, ref, rtol=1e-12)
        tg = rng.uniform(-2, 2, (300, dim))
        np.testing.assert_allclose(O.potential(t, x, m, tg, softening=0.01, theta=0.0),
                                   O.direct_potential(x, m, tg, softening=0.01), rtol=1e-12)


def test_ring_centre():
    n, R, eps = 257, 2.5, 0.1
    a = 2 * np.pi * np.arange(n) / n
    x = np.stack([R * np.cos(a), R * np.sin(a), np.zeros(n)], axis=1)
    m = np.full(n, 0.25)
    t = O.build(x, m, 8)
    phi = O.potential(t, x, m, np.zeros((1, 3)), softening=eps, theta=0.0)
    np.testing.assert_allclose(phi, [-n * 0.25 / np.sqrt(R * R + eps * eps)], rtol=1e-13)


def test_theta_half_close_to_direct():
    rng = np.random.default_rng(9)
    x = rng.standard_normal((4000, 3))
    m = np.ones(4000)
    t = O.build(x, m, 32)
    phi = O.potential(t, x, m, softening=0.01, theta=0.5)
    ref = O.direct_potential(x, m, softening=0.01)
    err = np.sqrt(np.mean(((phi - ref) / ref) ** 2))
    assert 0 < err < 2e-3
