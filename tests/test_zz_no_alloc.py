"""bh_accel_targets_dev_noalloc (the plain external-target walk with its
workspace in a fixed device array, no allocation) against
bh_accel_targets_dev on the same inputs, through the C ABI: bit for bit on
two calls in a row, the second smaller than the first (the workspace keeps
the first call's contents), at the size limit, and the limit's error."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LIMIT = 1 << 20


@pytest.fixture(scope="module")
def native():
    from bhut_amd import _native
    if _native.device_count() < 1:
        pytest.skip("no HIP device")
    return _native


def test_noalloc_targets_equal_allocating_entry_point(native):
    import torch
    from bhut_amd import workloads
    dev = torch.device("cuda", 0)
    x, m = workloads.plummer(200_000, seed=41)
    dt = native.DeviceTree.build(x, m, 32)
    rng = np.random.default_rng(41)
    for step, (nt, theta, eps) in enumerate(((300_000, 0.5, 0.01), (1_000, 0.7, 0.0), (LIMIT, 0.6, 0.02))):
        print(f"no-alloc call {step}: {nt} targets, theta {theta}, eps {eps}", flush=True)
        tg = torch.from_numpy(rng.uniform(-5, 5, (nt, 3))).to(dev)
        a = torch.full((nt, 3), float("nan"), dtype=torch.float64, device=dev)
        b = torch.full((nt, 3), float("nan"), dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        dt.accel_targets_dev_noalloc(tg.data_ptr(), nt, a.data_ptr(), theta, eps, 1.0)
        dt.accel_targets_dev(tg.data_ptr(), nt, b.data_ptr(), theta, eps, 1.0)
        dt.context.synchronize()
        an, bn = a.cpu().numpy(), b.cpu().numpy()
        assert np.isfinite(an).all()
        np.testing.assert_array_equal(an.view(np.uint64), bn.view(np.uint64))
    # host-array form against the same walk, to tie both to the tested path
    tg = rng.uniform(-5, 5, (2_000, 3))
    out = torch.empty((2_000, 3), dtype=torch.float64, device=dev)
    tgd = torch.from_numpy(tg).to(dev)
    torch.cuda.synchronize()
    dt.accel_targets_dev_noalloc(tgd.data_ptr(), 2_000, out.data_ptr(), 0.5, 0.01, 1.0)
    dt.context.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy(), dt.accel_targets(tg, 0.5, 0.01, 1.0))
    # zero targets: nothing to do; above the limit: refused before any launch
    dt.accel_targets_dev_noalloc(0, 0, 0, 0.5, 0.01, 1.0)
    with pytest.raises(ValueError, match="at most"):
        dt.accel_targets_dev_noalloc(tgd.data_ptr(), LIMIT + 1, out.data_ptr(), 0.5, 0.01, 1.0)
