"""Points given in an order the one-pass Schur product cannot take, on the GPU.

tcv_ba_solve orders the points of host arrays by the lowest camera of their tracks, then by their
highest, when their given order would leave some one-pass workgroup's camera window wider than 64
cameras, solves, and puts the output points back in the given order. The scene here is a
150-camera trajectory with its points shuffled:

* the solve runs the one-pass product (report.onepass_rows > 0) and matches the oracle on the same
  shuffled input at tests/test_gpu_parity.py's bars (cost 1e-9 relative, coordinates
  1e-7·max(1, |x|), accept/reject/failure/input-error words equal);
* it equals, bit for bit, the solve of the same points given in the engine's own order (a stable
  sort by lowest camera, then highest camera, directional points after the others) with the output
  points permuted back: the reordering is the whole difference, and putting the points back loses
  nothing.

Both entries also take that order when the given one fits but writes at least an eighth more
one-pass partial rows (the trajectory generator's order, by the camera that created each point):
its solve then writes fewer rows.

tcv_ba_solve_device orders device arrays the same way on the device: its solves of the shuffled and
of the generator's scene equal tcv_ba_solve's bit for bit and leave the caller's arrays as they were;
an index the engine refuses is refused by both entries with the same error.
"""
import copy
import ctypes

import numpy as np
import pytest
# torch's HIP runtime loads before the engine's library does (as in bench.py): loaded after it,
# torch finds no device
import torch

from oracle import oracle as O
from problems import shuffle_points
from typecv_amd import _native
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
    """The order tcv_ba_solve gives the points: stable by the lowest camera of each track, then by
    its highest camera, the directional points (the last `directional`) after the others."""
    ps = packed.section("pointStart").astype(np.int64)
    cams = packed.section("observationCamera").astype(np.int64)
    lo = np.array([cams[ps[p]:ps[p + 1]].min() if ps[p + 1] > ps[p] else packed.C
                   for p in range(packed.P)])
    hi = np.array([cams[ps[p]:ps[p + 1]].max() if ps[p + 1] > ps[p] else packed.C
                   for p in range(packed.P)])
    d = packed.P - directional
    return np.concatenate([np.lexsort((hi[:d], lo[:d])), d + np.lexsort((hi[d:], lo[d:]))])


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


def solve_device(device, control, packed):
    """tcv_ba_solve_device on copies of the packed arrays in device memory: (out, report, whether
    the device arrays still hold the packed arrays afterwards)."""
    lib = _native.library()
    scene_h = np.ascontiguousarray(packed.scene, dtype=np.float64)
    index_h = np.ascontiguousarray(packed.index, dtype=np.uint32)
    scene = torch.from_numpy(scene_h).cuda()
    index = torch.from_numpy(index_h.view(np.int32)).cuda()
    out = torch.empty(lib.tcv_ba_output_size(ctypes.byref(control)), dtype=torch.float64,
                      device="cuda")
    torch.cuda.synchronize()
    report = _native.Report()
    status = lib.tcv_ba_solve_device(device.handle, ctypes.byref(control), scene.data_ptr(),
                                     index.data_ptr(), out.data_ptr(), ctypes.byref(report), None)
    _native.check(status, "tcv_ba_solve_device")
    torch.cuda.synchronize()
    kept = (np.array_equal(scene.cpu().numpy(), scene_h)
            and np.array_equal(index.cpu().numpy().view(np.uint32), index_h))
    return out.cpu().numpy(), report.as_dict(), kept


def test_device_entry_orders_points_on_the_device(device):
    packed = SHUFFLED
    print(f"\n[point order] device entry C={packed.C} P={packed.P} M={packed.M}", flush=True)
    control = B.make_control(packed.C, packed.P, packed.M, **KW)
    host, _ = B.solve_packed(device, control, packed)
    got, report, kept = solve_device(device, control, packed)
    assert report["onepass_rows"] > 0
    assert kept
    assert np.array_equal(got, host)


def test_generator_order_is_reordered_for_fewer_rows_by_both_entries(device):
    packed, _ = S.trajectory_scene(cameras=150, points=8000, seed=2)
    print("\n[point order] generator's order", flush=True)
    control = B.make_control(packed.C, packed.P, packed.M, **KW)
    host, host_report = B.solve_packed(device, control, packed)
    got, report, kept = solve_device(device, control, packed)
    assert report["onepass_rows"] == host_report["onepass_rows"] > 0
    assert kept
    assert np.array_equal(got, host)
    # the engine's order given: kept as it is, and the rows are the reordered solve's
    perm = engine_order(packed)
    ordered, ordered_report = B.solve_packed(device, control, reordered(packed, perm))
    assert ordered_report["onepass_rows"] == report["onepass_rows"]
    base_c = packed.C * 12
    back = ordered.copy()
    back[base_c + 3 * perm[:, None] + np.arange(3)] = ordered[base_c:base_c + 3 * packed.P].reshape(-1, 3)
    assert np.array_equal(host, back)


def test_device_entry_refuses_a_bad_index_as_the_host_entry_does(device):
    print("\n[point order] device entry, corrupted index", flush=True)
    packed = SHUFFLED
    control = B.make_control(packed.C, packed.P, packed.M, iterations=2, max_cg=10)
    swapped = copy.deepcopy(packed)
    lists = swapped.section("cameraObservations")
    lists[[3, 4]] = lists[[4, 3]]  # breaks the ascending order of one camera list
    far = copy.deepcopy(packed)
    far.index[0] = 10_000  # a camera id past C
    for bad in (swapped, far):
        with pytest.raises(_native.EngineError, match="index"):
            B.solve_packed(device, control, bad)
        with pytest.raises(_native.EngineError, match="index"):
            solve_device(device, control, bad)
    got, _, _ = solve_device(device, control, packed)  # the context still works
    assert np.isfinite(got).all()
