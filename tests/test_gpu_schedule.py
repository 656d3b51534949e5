"""The two-sweep Schur product's schedule tuner changes speed only.

A context's first PCG chunk on a new problem shape runs iterations 1..6 with the three candidate
schedules in turn (the plain sweeps with no cap, with the camera sweep capped at three workgroups
per CU by LDS reservation, and the fused product) and keeps the fastest; no schedule changes
the task partition or any sum. So a solve that tunes (and so mixes all three) and a later solve of
the same problem that runs only the picked one return the same bits, in both arithmetic forms.
The scene is one where the one-pass product does not run (points along 200 cameras in a random
order, so that a workgroup's camera window passes 64 cameras): the two sweeps, and so the tuner,
run.
"""
import numpy as np
import pytest

from problems import shuffle_points
from typecv_amd import bundle as B
from typecv_amd import synthetic as S

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_tuned_schedule_keeps_the_bits(precision):
    packed, _ = S.trajectory_scene(cameras=200, points=600, seed=5, track=(5, 30))
    packed = shuffle_points(packed, seed=1)
    control = B.make_control(packed.C, packed.P, packed.M, shared_focal=True, huber=1.0,
                             iterations=6, max_cg=64, precision=precision)
    print(f"\n[schedule] {precision} C={packed.C} P={packed.P} M={packed.M}", flush=True)
    device = B.HipDevice(0)  # a fresh context: its first solve tunes
    try:
        first, r1 = B.solve_packed(device, control, packed)
        second, r2 = B.solve_packed(device, control, packed)
    finally:
        device.dispose()
    print(f"[schedule] cap {r1['camera_sweep_cap']} / {r2['camera_sweep_cap']}, fused "
          f"{r1['fused_product']} / {r2['fused_product']}, PCG {r1['cg_iterations']}", flush=True)
    assert r1["cg_iterations"] > 6
    assert r1["onepass_rows"] == 0 and r2["onepass_rows"] == 0  # the two sweeps ran
    assert r1["camera_sweep_cap"] in (0, 3)
    assert r2["camera_sweep_cap"] == r1["camera_sweep_cap"]
    assert r1["fused_product"] in (0, 1) and r2["fused_product"] == r1["fused_product"]

    assert np.array_equal(first, second)
