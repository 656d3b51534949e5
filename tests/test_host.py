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


# ------------------------------------------------------------------ AddressSanitizer (host code)

def asan_runtime():
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        return None
    path = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True,
                          text=True).stdout.strip()
    return path if path and Path(path).is_file() else None


ASAN_FLAGS = ["-fsanitize=address", "-fno-omit-frame-pointer", "-g", "-O1"]


@pytest.mark.parametrize("shared", [False, True])
def test_oracle_under_address_sanitizer(tmp_path, shared):
    """The oracle's solve, chart gradient and dense solve built with -fsanitize=address (leak
    checking on) run a fixture clean and return the regular build's arrays."""
    import os
    import subprocess
    from oracle import oracle as O
    from problems import packed_problem
    if asan_runtime() is None:
        pytest.skip("no AddressSanitizer runtime for gcc")
    exe = tmp_path / "oracle_asan"
    # the runtime linked statically, so that it comes first whatever the environment preloads
    build = subprocess.run(["gcc", *ASAN_FLAGS, "-static-libasan", "-std=c11", "-fopenmp",
                            "-ffp-contract=off", "-D_POSIX_C_SOURCE=199309L",
                            "-I", str(ROOT / "include"),
                            str(ROOT / "tests" / "asan" / "oracle_driver.c"),
                            str(ROOT / "oracle" / "ba_oracle.c"), "-o", str(exe), "-lm"],
                           capture_output=True, text=True)
    assert build.returncode == 0, build.stderr
    if shared:
        model = S.joint_focal_fixture()["input"]
        options = {"iterations": 12, "refineSharedFocal": True, "maxCgIterations": 64}
    else:
        model, options = S.bundle_fixture(outlier=True), {"iterations": 12, "huberPx": 1}
    packed, control = packed_problem(model, options)
    (tmp_path / "control.bin").write_bytes(bytes(control))
    packed.scene.astype(np.float64).tofile(tmp_path / "scene.f64")
    packed.index.astype(np.uint32).tofile(tmp_path / "index.u32")
    run = subprocess.run([str(exe), str(tmp_path / "control.bin"), str(tmp_path / "scene.f64"),
                          str(tmp_path / "index.u32"), str(tmp_path / "out")],
                         capture_output=True, text=True, timeout=300,
                         env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1",
                                  OMP_NUM_THREADS="2"))
    assert run.returncode == 0, run.stdout + run.stderr
    assert "AddressSanitizer" not in run.stderr
    out = np.fromfile(tmp_path / "out.solve.f64")
    want, _ = O.solve(control, packed)
    assert np.array_equal(out, want)
    grad = np.fromfile(tmp_path / "out.gradient.f64")
    assert np.array_equal(grad, O.gradient(control, packed.scene, packed.index)[0])
    dense = np.fromfile(tmp_path / "out.dense.f64")
    assert np.array_equal(dense, O.dense_solve(control, packed))


def test_addon_host_code_under_address_sanitizer(tmp_path):
    """The Node-API addon built with -fsanitize=address: every export's argument checks (valid and
    invalid controls, wrong types and lengths, a non-context, opening a device) run clean in Node."""
    import json
    import os
    import shutil
    import subprocess
    from typecv_amd.node import build as addon_build
    runtime = asan_runtime()
    include = addon_build.node_include()
    if runtime is None or include is None or shutil.which("node") is None:
        pytest.skip("no AddressSanitizer runtime, Node or Node-API headers")
    addon = tmp_path / "typecv_ba.node"
    package = ROOT / "typecv_amd"
    build = subprocess.run(["gcc", *ASAN_FLAGS, "-std=c11", "-fPIC", "-shared",
                            "-DNODE_GYP_MODULE_NAME=typecv_ba", "-I", str(include),
                            "-I", str(ROOT / "include"), str(package / "node" / "typecv_ba_napi.c"),
                            "-L", str(package), "-ltypecv_ba", f"-Wl,-rpath,{package}",
                            "-o", str(addon)], capture_output=True, text=True)
    assert build.returncode == 0, build.stderr
    # the runtime goes first in the preload list (Node itself is not instrumented), ahead of
    # whatever the environment already preloads; Node's own allocations are not leak-checked
    preload = " ".join(x for x in (runtime, os.environ.get("LD_PRELOAD", "")) if x)
    env = dict(os.environ, LD_PRELOAD=preload, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1")
    run = subprocess.run(["node", str(ROOT / "tests" / "asan" / "addon_checks.js"), str(addon)],
                         capture_output=True, text=True, timeout=120, env=env)
    assert run.returncode == 0, run.stdout + run.stderr
    assert "AddressSanitizer" not in run.stderr
    result = json.loads(run.stdout.strip().splitlines()[-1])
    assert result["outputSize"] == 120 + 300 + 27
    assert all(result["threw"].values()), result["threw"]
