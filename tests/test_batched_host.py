"""Host-side validation of the batched API (no device work)."""

import numpy as np
import pytest

import bhut_amd


def test_zero_particle_system_rejected():
    with pytest.raises(ValueError, match="zero particles"):
        bhut_amd.Forest(np.zeros((3, 3)), np.ones(3), [3, 0])


def test_bad_leaf_rejected():
    with pytest.raises(ValueError, match="leaf_size"):
        bhut_amd.Forest(np.zeros((3, 3)), np.ones(3), [3], leaf_size=0)


def test_mismatched_system_rejected():
    with pytest.raises(ValueError, match="masses"):
        bhut_amd.accelerations_batched([(np.zeros((4, 3)), np.ones(3))])
    with pytest.raises(ValueError, match="positions must be"):
        bhut_amd.accelerations_batched([(np.zeros((4, 3)), np.ones(4)), (np.zeros((4, 2)), np.ones(4))])
    with pytest.raises(ValueError, match="at least one"):
        bhut_amd.accelerations_batched([])
