"""The plain walk's middle-range FP64 square root and division
(bhut_amd/csrc/bh_fp64.h) against sqrt() and '/' on the device, bit for bit
(bhut_amd/fp64_check: random operands over the whole middle range, exact
squares and their neighbours, every exponent of the range with boundary
mantissas)."""

import json
import pathlib
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = pathlib.Path(__file__).resolve().parents[1] / "bhut_amd" / "fp64_check"


def test_middle_range_sqrt_and_division_bitwise():
    from bhut_amd import _native
    if _native.device_count() < 1:
        pytest.skip("no HIP device")
    assert EXE.exists(), "bhut_amd/fp64_check is not built (make -C bhut_amd/csrc)"
    print("fp64_check: 2^28 random operands per function", flush=True)
    r = subprocess.run([str(EXE), str(1 << 28)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    res = json.loads(r.stdout.strip().splitlines()[-1])
    assert res["operands"] > 1 << 28
    for f in ("sqrt", "recip", "div", "recip_sqrt"):
        assert res[f]["mismatch"] == 0, (f, res[f])
    assert r.returncode == 0
