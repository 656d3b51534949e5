"""The operator on the real south-building-24 problem (24 views, 3,000 landmarks, 12,706
observations of COLMAP's SIFT matches; tests/real_scene.py) against the reference's recorded runs
and the oracle: final RMSE within 1e-9 relative of the record (the reference's WebGPU tier rounds
differently), 4 accepted steps, and the oracle's parity bars (final cost within 1e-9 relative,
coordinates within 1e-7·max(1, |x|)). The direct route (`directSolve`) solves each step exactly
where the record's PCG stops at a relative residual of 1e-6: its final RMSE is within 1e-7."""
import numpy as np
import pytest

from oracle import oracle as O
from typecv_amd import bundle as B
from problems import packed_problem
import real_scene as R

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("focal", ["fixed", "shared"])
@pytest.mark.parametrize("cache", [False, True])
def test_operator_reproduces_the_recorded_run(device, focal, cache):
    print(f"\n[south-building-24] {focal} directSolve={cache}", flush=True)
    model = R.reconstruction()
    options = {**R.OPTIONS, "refineSharedFocal": focal == "shared", "directSolve": cache}
    result = B.bundle_adjust_gpu(device, model, options)
    d = result["diagnostics"]
    record = R.RECORD[focal]
    print(f"  RMSE {d['initialRmsePx']:.16g} -> {d['finalRmsePx']:.16g} px (record "
          f"{record['finalRmsePx']}); PCG {result['cgIterations']} (record "
          f"{record['cgIterations']})", flush=True)
    assert abs(d["initialRmsePx"] - record["initialRmsePx"]) <= 1e-9 * record["initialRmsePx"]
    if cache:
        # exact steps where the record's PCG stops at a relative residual of 1e-6: measured 6.6e-9
        # (fixed) and 1.4e-9 (shared) from the record
        assert abs(d["finalRmsePx"] - record["finalRmsePx"]) <= 1e-7 * record["finalRmsePx"]
    else:
        assert abs(d["finalRmsePx"] - record["finalRmsePx"]) <= 1e-9 * record["finalRmsePx"]
    assert d["acceptedSteps"] == record["acceptedSteps"]
    if not cache:
        packed, control = packed_problem(model, options)
        expected, _ = O.solve(control, packed)
        got, _ = B.solve_packed(device, control, packed)
        base = packed.C * 12 + packed.P * 3
        assert abs(got[base + 6] - expected[base + 6]) <= 1e-9 * expected[base + 6]
        scale = np.maximum(1.0, np.abs(expected[:base]))
        assert np.max(np.abs(got[:base] - expected[:base]) / scale) <= 1e-7
