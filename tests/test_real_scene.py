"""The real-data pin: the south-building-24 problem (tests/golden/south_building_24.npz, built by
tests/golden/south_building_24.py from the reference's own sample data) against the reference's
recorded runs of it (`docs/experiments/gpu-budget/south-building-24-ba-{fixed,shared}.json`).

The reference ran on its WebGPU tier (emulated ~48-bit f64), so its RMSEs are matched to 1e-9
relative rather than bit for bit, and its PCG counts to within 1 %. Measured: initial RMSE
1.8e-11 relative, final 6.6e-12 (fixed) and 6.5e-10 (shared); 302 / 315 PCG iterations against the
record's 304 / 317."""
import numpy as np
import pytest

from oracle import oracle as O
from problems import packed_problem
import real_scene as R


def test_fixture_has_the_recorded_sizes():
    with np.load(R.FIXTURE) as f:
        z = {k: f[k] for k in f.files}
    assert len(z["view_id"]) == 24
    assert len(z["xyz"]) == 3000
    assert len(z["obs_view"]) == 12706  # cameras / points / observations of both records
    starts = z["obs_start"]
    assert starts[0] == 0 and starts[-1] == 12706 and np.all(np.diff(starts) >= 2)


@pytest.mark.parametrize("focal", ["fixed", "shared"])
def test_oracle_reproduces_the_recorded_run(focal):
    model = R.reconstruction()
    packed, control = packed_problem(model, {**R.OPTIONS, "refineSharedFocal": focal == "shared"})
    out, _ = O.solve(control, packed)
    base = packed.C * 12 + packed.P * 3
    record = R.RECORD[focal]
    initial, final = out[base + 24], out[base + 25]
    print(f"\n[{focal}] RMSE {initial:.16g} -> {final:.16g} px; record {record['initialRmsePx']} -> "
          f"{record['finalRmsePx']}; PCG {out[base + 16]:.0f} (record {record['cgIterations']})")
    assert abs(initial - record["initialRmsePx"]) <= 1e-9 * record["initialRmsePx"]
    assert abs(final - record["finalRmsePx"]) <= 1e-9 * record["finalRmsePx"]
    assert out[base + 10] == record["acceptedSteps"]
    assert abs(out[base + 16] - record["cgIterations"]) <= 0.01 * record["cgIterations"]
