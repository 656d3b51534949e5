"""Host half of the operator on the CPU: the reference's validation errors, packing bit for bit
against a loop restatement of packBundleScene, and the C ABI's exported symbols (no device call)."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from oracle.pack_reference import pack_bundle_scene as pack_loop
from typecv_amd import _native
from typecv_amd import bundle as B
from typecv_amd import synthetic as S

ROOT = Path(__file__).resolve().parent.parent
FIXED = lambda C: min(96, max(24, 12 * C))  # noqa: E731


@pytest.mark.parametrize("make,options", [
    (lambda: S.bundle_fixture(), {}),
    (lambda: S.bundle_fixture(views=5), {"variableViewIds": [30, 40]}),
    (lambda: S.joint_focal_fixture()["input"], {"refineSharedFocal": True}),
    (lambda: S.ring_bundle_fixture(views=24, points=200), {}),
])
def test_packing_matches_loop_restatement_bit_for_bit(make, options):
    model = make()
    prob = B.bundle_problem(model, options, FIXED)
    packed = B.pack_bundle_scene(prob["input"], prob["byId"], prob["active"])
    scene, index = pack_loop(prob["input"], prob["byId"], prob["active"])
    assert np.array_equal(packed.scene.view(np.uint64), scene.view(np.uint64))
    assert np.array_equal(packed.index, index)
    # every observation exactly once in the camera lists, ascending within each camera
    lists = packed.section("cameraObservations")
    starts = packed.section("cameraStart")
    assert np.array_equal(np.sort(lists), np.arange(packed.M))
    cams = packed.section("observationCamera")
    for c in range(packed.C):
        seg = lists[starts[c]: starts[c + 1]]
        assert np.all(np.diff(seg.astype(np.int64)) > 0)
        assert np.all(cams[seg] == c)


def test_synthetic_scene_packing_invariants():
    packed, info = S.trajectory_scene(cameras=30, points=800, seed=9)
    ps = packed.section("pointStart")
    lengths = np.diff(ps.astype(np.int64))
    assert lengths.min() >= 5 and lengths.max() <= 20
    assert B.finite_gpu_array(packed.scene)


def expect_range_error(model, options, pattern):
    with pytest.raises(B.RangeError, match=pattern):
        B.bundle_problem(model, options, FIXED)


def test_problem_rejections():
    m = S.bundle_fixture()
    m["landmarks"][1]["observations"][0] = dict(m["landmarks"][0]["observations"][0])
    expect_range_error(m, {}, "multiply owned")
    m = S.bundle_fixture()
    m["landmarks"][0]["observations"] = m["landmarks"][0]["observations"][:1]
    expect_range_error(m, {}, "at least two observations")
    expect_range_error(S.bundle_fixture(), {"anchorViewId": 10, "scaleViewId": 10}, "gauge cameras")
    expect_range_error(S.bundle_fixture(), {"anchorViewId": 99}, "gauge cameras")
    expect_range_error(S.bundle_fixture(), {"variableViewIds": [10]}, "variable camera window")
    expect_range_error(S.bundle_fixture(), {"variableViewIds": [30, 30]}, "variable camera window")
    expect_range_error(S.bundle_fixture(views=5), {"variableViewIds": [20, 30]},
                       "scale camera must be fixed")
    expect_range_error(S.bundle_fixture(), {"iterations": 101}, "scene/options")
    expect_range_error(S.bundle_fixture(), {"maxCgIterations": 257}, "scene/options")
    expect_range_error(S.bundle_fixture(), {"cgRelativeTolerance": 0.5}, "scene/options")
    m = S.bundle_fixture(views=3)
    m["views"].append({"id": 99, "camera": dict(m["views"][0]["camera"]),
                       "pose": {"R": np.eye(3).reshape(9), "t": np.zeros(3)}})
    expect_range_error(m, {}, "connected")
    m = S.bundle_fixture()
    m["views"][2]["pose"]["R"] = np.eye(3).reshape(9) * 1.1
    expect_range_error(m, {}, "SO\\(3\\)")


def test_option_enum_errors_are_synchronous():
    with pytest.raises(B.RangeError, match="cameraProductWorkgroupSize"):
        B._check_option_enums({"cameraProductWorkgroupSize": 48})
    with pytest.raises(B.RangeError, match="Schur reduction"):
        B._check_option_enums({"schurReduction": "fast"})
    with pytest.raises(B.RangeError, match="schurCache"):
        B._check_option_enums({"schurCache": 1})
    with pytest.raises(B.RangeError, match="directSolve"):
        B._check_option_enums({"directSolve": "yes"})
    with pytest.raises(B.RangeError, match="stallTolerance"):
        B.bundle_stall_tolerance({"stallTolerance": 1.0})


def test_directional_constraints_append_unit_bearings():
    m = S.bundle_fixture(count=10)
    extra = [{"trackId": 500, "direction": np.array([0.0, 0.0, 2.0]),
              "observations": [{"viewId": 10, "featureId": 900, "pixel": [480.0, 270.0]},
                               {"viewId": 20, "featureId": 900, "pixel": [470.0, 271.0]}]}]
    prob = B.bundle_problem(m, {"directionalConstraints": extra}, FIXED)
    assert prob["P"] == 11 and prob["finitePoints"] == 10
    assert np.allclose(prob["input"]["landmarks"][-1]["xyz"], [0, 0, 1])
    extra[0]["observations"][1]["featureId"] = 3  # owned by a finite landmark of view 20
    with pytest.raises(B.RangeError, match="directional observation"):
        B.bundle_problem(m, {"directionalConstraints": extra}, FIXED)


def declared_functions():
    text = (ROOT / "include" / "typecv_ba.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.findall(r"^\s*[a-z_0-9 ]+\*?\s+\**(tcv_[a-z_]+)\s*\(", text, flags=re.M)


def test_c_abi_exports_every_declared_symbol():
    declared = declared_functions()
    assert set(declared) == set(_native.EXPORTS)
    lib = ctypes.CDLL(str(_native.LIBRARY))
    for name in declared:
        assert hasattr(lib, name), name
    # size helpers need no device
    control = B.make_control(10, 100, 1000, shared_focal=True)
    lib = _native.library()
    assert lib.tcv_ba_output_size(ctypes.byref(control)) == 120 + 300 + 27
    assert lib.tcv_ba_scene_size(ctypes.byref(control)) == 180 + 300 + 2000
    assert lib.tcv_ba_index_size(ctypes.byref(control)) == 3000 + 100 + 22
    assert lib.tcv_ba_status_string(2).decode().startswith("GPU BA index")


def test_node_addon_loads_and_exports_the_operator():
    """The Node-API addon over the C ABI loads in the installed Node (no device touched) and
    exports the operator's entry points."""
    import json
    import shutil
    import subprocess
    from pathlib import Path

    node = shutil.which("node")
    addon = Path(__file__).resolve().parent.parent / "typecv_amd" / "node" / "typecv_ba.node"
    if node is None or not addon.exists():
        pytest.skip("no Node or no built addon (build() builds it where Node-API headers exist)")
    out = subprocess.run([node, str(addon.parent / "run.js"), "--exports"], capture_output=True,
                         text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    assert json.loads(out.stdout) == ["backward", "deviceBytes", "evaluate", "jacobian", "open",
                                       "outputSize", "polish", "release", "setWorkingBytes",
                                       "solve", "solveBatch"]
