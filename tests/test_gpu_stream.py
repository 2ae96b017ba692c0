"""Streamed external targets (bh_accel_targets_streamed): host-resident
targets through the device in chunks on two lane streams, the counterpart of
the reference's Dask target chunks (bhut/api.py:28-78, 690-731;
bhut/backends/dask_.py:173-228).  Every result must be bit for bit the one a
single call on all targets gives, in the plain walk and in the FP32 form,
whatever the chunk size; the plain form equals the oracle on a small case."""

import time

import numpy as np
import pytest

from bhut_amd import workloads
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def native():
    from bhut_amd import _native
    if _native.device_count() < 1:
        pytest.skip("no HIP device")
    return _native


def bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


def test_streamed_equals_single_call(native):
    x, m = workloads.plummer(1_000_000, seed=1)
    dt = native.DeviceTree.build(x, m, 32)
    rng = np.random.default_rng(61)
    q = rng.uniform(-5.0, 5.0, (300_001, 3))
    q[1000:1010] = q[5]  # coincident targets (tied Morton codes)
    q[2000] = x[17]      # a target on a source particle
    one64 = dt.accel_targets(q, 0.5, 0.01, 1.0)
    one32 = dt.accel_f32(q, 0.5, 0.01, 1.0)
    # chunks of one group of 64 on a part of the targets (thousands of chunks)
    np.testing.assert_array_equal(bits(dt.accel_targets_streamed(q[:5000], 0.5, 0.01, 1.0, chunk=1)),
                                  bits(dt.accel_targets(q[:5000], 0.5, 0.01, 1.0)))
    np.testing.assert_array_equal(bits(dt.accel_targets_streamed(q[:5000], 0.5, 0.01, 1.0, precision="fp32",
                                                                 chunk=1)),
                                  bits(dt.accel_f32(q[:5000], 0.5, 0.01, 1.0)))
    for chunk in (50_000, 65_536, 1 << 20):
        t0 = time.time()
        s64 = dt.accel_targets_streamed(q, 0.5, 0.01, 1.0, precision="fp64", chunk=chunk)
        t1 = time.time()
        s32 = dt.accel_targets_streamed(q, 0.5, 0.01, 1.0, precision="fp32", chunk=chunk)
        print(f"chunk {chunk}: plain {t1 - t0:.3f} s, FP32 {time.time() - t1:.3f} s", flush=True)
        np.testing.assert_array_equal(bits(s64), bits(one64), err_msg=f"fp64 chunk {chunk}")
        np.testing.assert_array_equal(bits(s32), bits(one32), err_msg=f"fp32 chunk {chunk}")
    # unsoftened, another theta
    np.testing.assert_array_equal(bits(dt.accel_targets_streamed(q[:70_000], 0.7, 0.0, 2.0, chunk=20_000)),
                                  bits(dt.accel_targets(q[:70_000], 0.7, 0.0, 2.0)))


def test_streamed_small_against_oracle_and_api(native):
    import bhut_amd
    from bhut_amd.traverse.bh import barnes_hut_accelerations_targets
    rng = np.random.default_rng(62)
    for dim in (3, 2):
        x = rng.standard_normal((3000, dim))
        mass = rng.uniform(0.5, 1.5, 3000)
        q = rng.uniform(-3, 3, (777, dim))
        ot = O.build(x, mass, 16)
        ref = O.accelerations(ot, x, mass, q, softening=0.01, theta=0.5)
        t = bhut_amd.Tree(x, mass, leaf_size=16, dim=dim).build()
        got = t.accelerations(q, theta=0.5, softening=0.01, chunk_size=100)
        nr = np.linalg.norm(ref, axis=1)
        assert np.all(np.linalg.norm(got - ref, axis=1) <= 1e-12 * np.maximum(nr, nr.max()))
        np.testing.assert_array_equal(bits(got), bits(t.accelerations(q, theta=0.5, softening=0.01)))
        g32 = t.accelerations(q, theta=0.5, softening=0.01, precision="fp32", chunk_size=64)
        np.testing.assert_array_equal(bits(g32), bits(t.accelerations(q, theta=0.5, softening=0.01,
                                                                       precision="fp32")))
        np.testing.assert_array_equal(
            bits(barnes_hut_accelerations_targets(t._tree, t.positions, t.masses, q, 0.01, 0.5, chunk_size=5)),
            bits(got))
    with pytest.raises(ValueError):
        t.accelerations(q, theta=0.5, softening=0.01, chunk_size=0)
    # no targets, one target
    dt = t._tree.device_tree
    assert dt.accel_targets_streamed(np.zeros((0, 2)), 0.5, 0.01, 1.0).shape == (0, 2)
    np.testing.assert_array_equal(bits(dt.accel_targets_streamed(q[:1], 0.5, 0.01, 1.0, chunk=7)),
                                  bits(dt.accel_targets(q[:1], 0.5, 0.01, 1.0)))
