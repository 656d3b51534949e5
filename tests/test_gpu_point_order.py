"""Points given in an order the one-pass Schur product cannot take, on the GPU.

tcv_ba_solve orders the points of host arrays by the lowest camera of their tracks when their given
order would leave some one-pass workgroup's camera window wider than 64 cameras, solves, and puts
the output points back in the given order. The scene here is a 150-camera trajectory with its
points shuffled:

* the solve runs the one-pass product (report.onepass_rows > 0) and matches the oracle on the same
  shuffled input at tests/test_gpu_parity.py's bars (cost 1e-9 relative, coordinates
  1e-7·max(1, |x|), accept/reject/failure/input-error words equal);
* it equals, bit for bit, the solve of the same points given in the engine's own order (a stable
  sort by lowest camera, directional points after the others) with the output points permuted back:
  the reordering is the whole difference, and putting the points back loses nothing.
"""
import numpy as np
import pytest

from oracle import oracle as O
from problems import shuffle_points
from typecv_amd import bundle as B
from typecv_amd import synthetic as S
from typecv_amd.packing import pack_arrays

from test_gpu_parity import assert_solve_parity

pytestmark = pytest.mark.gpu

KW = dict(iterations=10, max_cg=256, shared_focal=True, huber=1.0)


def scene():
    traj, _ = S.trajectory_scene(cameras=150, points=8000, seed=2)
    return shuffle_points(traj, seed=3)


def engine_order(packed, directional=0):
    """The order tcv_ba_solve gives the points: stable by the lowest camera of each track, the
    directional points (the last `directional`) after the others."""
    ps = packed.section("pointStart").astype(np.int64)
    cams = packed.section("observationCamera").astype(np.int64)
    key = np.array([cams[ps[p]:ps[p + 1]].min() if ps[p + 1] > ps[p] else packed.C
                    for p in range(packed.P)])
    first_dir = packed.P - directional
    return np.concatenate([np.argsort(key[:first_dir], kind="stable"),
                           first_dir + np.argsort(key[first_dir:], kind="stable")])


def reordered(packed, perm):
    ps = packed.section("pointStart").astype(np.int64)
    order = np.concatenate([np.arange(ps[p], ps[p + 1]) for p in perm])
    return pack_arrays(packed.poses(), packed.models(), packed.points()[perm], np.diff(ps)[perm],
                       packed.section("observationCamera")[order], packed.pixels()[order],
                       packed.section("cameraActive"))


SHUFFLED = scene()


def test_shuffled_points_run_the_one_pass_product_and_match_the_oracle(device):
    packed = SHUFFLED
    print(f"\n[point order] C={packed.C} P={packed.P} M={packed.M}", flush=True)
    control = B.make_control(packed.C, packed.P, packed.M, **KW)
    got, report = B.solve_packed(device, control, packed)
    assert report["onepass_rows"] > 0  # the points were ordered for the one-pass product
    expected, _ = O.solve(control, packed)
    base = packed.C * 12 + packed.P * 3
    assert expected[base + 6] < expected[base + 8]
    assert_solve_parity(got, expected, base)


def test_reordering_is_the_whole_difference(device):
    packed = SHUFFLED
    control = B.make_control(packed.C, packed.P, packed.M, **KW)
    got, _ = B.solve_packed(device, control, packed)
    perm = engine_order(packed)
    ordered = reordered(packed, perm)
    direct, report = B.solve_packed(device, control, ordered)
    assert report["onepass_rows"] > 0
    base_c = packed.C * 12
    back = direct.copy()
    back[base_c + 3 * perm[:, None] + np.arange(3)] = direct[base_c:base_c + 3 * packed.P].reshape(-1, 3)
    assert np.array_equal(got, back)
