"""The Node-API addon end to end on the GPU: a packed solve through `typecv_amd/node/run.js`
(Node, napi_async_work, tcv_ba_solve) returns exactly the array the Python binding returns for the
same inputs, and its diagnostics; a batch through `solveBatch` (tcv_ba_solve_batch) returns each
problem's array exactly as the Python batch does; an observation evaluation through `evaluate`
(tcv_obs_evaluate) returns the Python path's flags and summary."""
import json
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from typecv_amd import bundle as B
from typecv_amd import synthetic as S
from problems import packed_problem

pytestmark = pytest.mark.gpu

NODE_DIR = Path(__file__).resolve().parent.parent / "typecv_amd" / "node"


def control_vector(c) -> np.ndarray:
    return np.array([c.cameras, c.points, c.directional_points, c.observations, c.anchor,
                     c.scale_view, c.normalize_scale, c.shared_focal, c.iterations, c.max_cg,
                     c.precision, c.flags, c.huber, c.relative_tolerance, c.stall_tolerance,
                     c.refinement_tolerance], dtype=np.float64)


@pytest.mark.parametrize("shared", [False, True])
def test_node_solve_matches_python(device, tmp_path, shared):
    node = shutil.which("node")
    if node is None or not (NODE_DIR / "typecv_ba.node").exists():
        pytest.skip("no Node or no built addon")
    print(f"\n[node] shared={shared}", flush=True)
    if shared:
        model = S.joint_focal_fixture()["input"]
        options = {"iterations": 40, "refineSharedFocal": True, "maxCgIterations": 256,
                   "cgRelativeTolerance": 1e-10}
    else:
        model, options = S.bundle_fixture(outlier=True), {"iterations": 20, "huberPx": 1}
    packed, control = packed_problem(model, options)
    control_vector(control).tofile(tmp_path / "control.f64")
    packed.scene.astype(np.float64).tofile(tmp_path / "scene.f64")
    packed.index.astype(np.uint32).tofile(tmp_path / "index.u32")
    run = subprocess.run([node, str(NODE_DIR / "run.js"), str(tmp_path / "control.f64"),
                          str(tmp_path / "scene.f64"), str(tmp_path / "index.u32"),
                          str(tmp_path / "out.f64")], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stderr
    got = np.fromfile(tmp_path / "out.f64", dtype=np.float64)
    expected, _ = B.solve_packed(device, control, packed)
    assert np.array_equal(got, expected)
    diag = json.loads(run.stdout.strip().splitlines()[-1])["diagnostics"]
    base = packed.C * 12 + packed.P * 3
    assert diag["finalCost"] == expected[base + 6] and diag["observations"] == packed.M


def test_node_batch_matches_python(device, tmp_path):
    node = shutil.which("node")
    if node is None or not (NODE_DIR / "typecv_ba.node").exists():
        pytest.skip("no Node or no built addon")
    print("\n[node] batch", flush=True)
    scenes = [S.trajectory_scene(cameras=c, points=p, seed=s, track=(3, min(c, 8)))[0]
              for c, p, s in ((6, 600, 1), (9, 900, 2))]
    controls = [B.make_control(x.C, x.P, x.M, iterations=6, max_cg=128, shared_focal=True,
                               huber=1.0, direct=True) for x in scenes]
    args = [node, str(NODE_DIR / "run.js"), "--batch", str(tmp_path / "out")]
    for b, (x, c) in enumerate(zip(scenes, controls)):
        control_vector(c).tofile(tmp_path / f"c{b}.f64")
        x.scene.astype(np.float64).tofile(tmp_path / f"s{b}.f64")
        x.index.astype(np.uint32).tofile(tmp_path / f"i{b}.u32")
        args += [str(tmp_path / f"c{b}.f64"), str(tmp_path / f"s{b}.f64"), str(tmp_path / f"i{b}.u32")]
    run = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stderr
    expected, _, _ = B.solve_batch(device, controls, scenes)
    for b, want in enumerate(expected):
        got = np.fromfile(tmp_path / f"out.{b}.f64", dtype=np.float64)
        assert np.array_equal(got, want)


@pytest.mark.parametrize("guarded", [False, True])
def test_node_observations_match_python(device, tmp_path, guarded):
    node = shutil.which("node")
    if node is None or not (NODE_DIR / "typecv_ba.node").exists():
        pytest.skip("no Node or no built addon")
    import sys
    sys.path.insert(0, str(Path(__file__).resolve().parent))
    from obs_cases import observation_model
    from typecv_amd import observations as OB
    print(f"\n[node] observations guarded={guarded}", flush=True)
    model = observation_model(points=300, views=5, baseline=0.2)
    model["landmarks"][7]["observations"][1]["pixel"][0] += 9.0
    options = {"maxRelativeDepthStdDev": 0.02} if guarded else {}
    req = OB.observation_request(model, 3.0, options, False)
    tracks = len(req["lengths"]) if guarded else 0
    np.array([req["views"], req["points"], req["count"], tracks, 3.0,
              req["depth"]["maxRelativeDepthStdDev"], req["depth"]["pixelNoiseStdDevPx"]],
             dtype=np.float64).tofile(tmp_path / "c.f64")
    req["cameras"].tofile(tmp_path / "cam.f64")
    req["points_xyz"].tofile(tmp_path / "pts.f64")
    req["pixels"].tofile(tmp_path / "px.f64")
    req["indices"].tofile(tmp_path / "ix.u32")
    offsets = "-"
    if guarded:
        req["offsets"].tofile(tmp_path / "off.u32")
        offsets = str(tmp_path / "off.u32")
    run = subprocess.run([node, str(NODE_DIR / "run.js"), "--observations",
                          str(tmp_path / "c.f64"), str(tmp_path / "cam.f64"),
                          str(tmp_path / "pts.f64"), str(tmp_path / "px.f64"),
                          str(tmp_path / "ix.u32"), offsets, str(tmp_path / "out")],
                         capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stderr
    flags, summary, _ = OB.evaluate_arrays(device, req, 3.0)
    got_flags = np.fromfile(tmp_path / "out.flags.u32", dtype=np.uint32)
    got_summary = np.fromfile(tmp_path / "out.summary.f64", dtype=np.float64)
    assert np.array_equal(got_flags, flags[:req["count"] + tracks])
    assert np.array_equal(got_summary, summary)


def test_node_concurrent_solves_on_one_context(device, tmp_path):
    """Two un-awaited `addon.solve` calls on one context (libuv workers, no FIFO in between) and a
    third after them return the Python binding's array, bit for bit: the context's lock runs them
    one after another."""
    node = shutil.which("node")
    if node is None or not (NODE_DIR / "typecv_ba.node").exists():
        pytest.skip("no Node or no built addon")
    print("\n[node] concurrent solves", flush=True)
    packed, _ = S.trajectory_scene(cameras=20, points=3000, seed=2)
    control = B.make_control(packed.C, packed.P, packed.M, iterations=10, max_cg=241,
                             shared_focal=True, huber=1.0)
    control_vector(control).tofile(tmp_path / "control.f64")
    packed.scene.astype(np.float64).tofile(tmp_path / "scene.f64")
    packed.index.astype(np.uint32).tofile(tmp_path / "index.u32")
    run = subprocess.run([node, str(NODE_DIR / "run.js"), "--concurrent",
                          str(tmp_path / "control.f64"), str(tmp_path / "scene.f64"),
                          str(tmp_path / "index.u32"), str(tmp_path / "out")],
                         capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stderr
    expected, _ = B.solve_packed(device, control, packed)
    for name in ("0", "1", "seq"):
        got = np.fromfile(tmp_path / f"out.{name}.f64", dtype=np.float64)
        assert np.array_equal(got, expected), name


@pytest.mark.parametrize("mode", ["polish", "jacobian"])
def test_node_derivative_entries_match_python(device, tmp_path, mode):
    node = shutil.which("node")
    if node is None or not (NODE_DIR / "typecv_ba.node").exists():
        pytest.skip("no Node or no built addon")
    print(f"\n[node] {mode}", flush=True)
    packed, _ = S.trajectory_scene(cameras=10, points=600, seed=6)
    control = B.make_control(packed.C, packed.P, packed.M, iterations=15, max_cg=121,
                             shared_focal=True, huber=1.0)
    out, _ = B.solve_packed(device, control, packed)
    at = B.at_solution(packed, out)
    lf = B.output_log_focal(control, out)
    control_vector(control).tofile(tmp_path / "control.f64")
    at.scene.astype(np.float64).tofile(tmp_path / "scene.f64")
    at.index.astype(np.uint32).tofile(tmp_path / "index.u32")
    run = subprocess.run([node, str(NODE_DIR / "run.js"), f"--{mode}",
                          str(tmp_path / "control.f64"), str(tmp_path / "scene.f64"),
                          str(tmp_path / "index.u32"), repr(lf), str(tmp_path / "out.f64")],
                         capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stderr
    got = np.fromfile(tmp_path / "out.f64", dtype=np.float64)
    if mode == "polish":
        expected, _ = B.polish(device, control, at, lf, iterations=3, max_cg=4000,
                               tolerance=1e-12, lm_iterations=30)
    else:
        expected = B.jacobian(device, control, at, lf).reshape(-1)
    assert np.array_equal(got, expected)


@pytest.mark.parametrize("mode", ["exact", "gaussNewton"])
def test_node_backward_matches_python(device, tmp_path, mode):
    """`backward` through the addon (tcv_ba_backward on a libuv worker) returns the Python binding's
    pixel and input-focal cotangents bit for bit, in both adjoint modes."""
    node = shutil.which("node")
    if node is None or not (NODE_DIR / "typecv_ba.node").exists():
        pytest.skip("no Node or no built addon")
    print(f"\n[node] backward {mode}", flush=True)
    packed, _ = S.trajectory_scene(cameras=10, points=600, seed=6)
    control = B.make_control(packed.C, packed.P, packed.M, iterations=15, max_cg=121,
                             shared_focal=True, huber=1.0)
    out, _ = B.solve_packed(device, control, packed)
    at = B.at_solution(packed, out)
    lf = B.output_log_focal(control, out)
    pol_out, _ = B.polish(device, control, at, lf, iterations=3, tolerance=1e-12,
                          lm_iterations=30)
    pol = B.at_solution(at, pol_out)
    lf = B.output_log_focal(control, pol_out)
    cot = np.random.default_rng(3).normal(size=12 * packed.C + 3 * packed.P + 1)
    control_vector(control).tofile(tmp_path / "control.f64")
    pol.scene.astype(np.float64).tofile(tmp_path / "scene.f64")
    pol.index.astype(np.uint32).tofile(tmp_path / "index.u32")
    cot.tofile(tmp_path / "cot.f64")
    run = subprocess.run([node, str(NODE_DIR / "run.js"), "--backward",
                          str(tmp_path / "control.f64"), str(tmp_path / "scene.f64"),
                          str(tmp_path / "index.u32"), repr(lf), str(tmp_path / "cot.f64"), mode,
                          str(tmp_path / "out")], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stderr
    pixels, focal, _, _ = B.backward(device, control, pol, lf, cot, max_cg=4000,
                                     tolerance=1e-10, mode=mode)
    got_pixels = np.fromfile(tmp_path / "out.pixels.f64", dtype=np.float64)
    got_focal = np.fromfile(tmp_path / "out.focal.f64", dtype=np.float64)
    assert np.array_equal(got_pixels, pixels.reshape(-1))
    assert np.array_equal(got_focal, focal)
