"""Local bundles on the GPU: windows extracted from one map (extractLocalBundle) and solved together
in one batched engine call equal, bit for bit, the same problems solved one by one through the
operator, and merging them back (mergeLocalBundle) lowers the map's robust cost."""
import numpy as np
import pytest

from typecv_amd import bundle as B
from typecv_amd import local_bundle as L
from typecv_amd import synthetic as S

pytestmark = pytest.mark.gpu


def windowed_map():
    """A 10-view strip: landmark i seen by the four consecutive views starting at view i % 7."""
    model, rng = S.reference_scene(count=140, views=10, noise=0.3)
    views = [{**v, "pose": S.update_pose(v["pose"], [0.002, -0.001, 0.001, 0.01, 0.005, -0.01])}
             if v["id"] in (40, 70, 90) else v for v in model["views"]]
    landmarks = []
    for i, p in enumerate(model["landmarks"]):
        first = (i % 7) * 10 + 10
        keep = range(first, first + 40, 10)
        landmarks.append({**p, "xyz": p["xyz"] + np.array([(rng() - 0.5) * 0.05 for _ in range(3)]),
                          "observations": [o for o in p["observations"] if o["viewId"] in keep]})
    return {"views": views, "landmarks": landmarks}


def test_batched_local_bundles_equal_single_solves(device):
    print("\n[local] batched windows", flush=True)
    model = windowed_map()
    options = {"iterations": 10, "huberPx": 2, "maxCgIterations": 96}
    windows = [[30, 40], [60, 70], [80, 90]]
    problems = [L.extract_local_bundle(model, w) for w in windows]
    assert all(p is not None for p in problems)
    batch = L.solve_local_bundles(device, problems, options)
    merged = model
    for problem, solved in zip(problems, batch):
        single = B.bundle_adjust_gpu(device, problem["reconstruction"], {
            **options, "variableViewIds": problem["variableViewIds"],
            "anchorViewId": problem["anchorViewId"], "scaleViewId": problem["scaleViewId"]})
        assert single["diagnostics"]["finalCost"] == solved["diagnostics"]["finalCost"]
        assert solved["diagnostics"]["finalCost"] < solved["diagnostics"]["initialCost"]
        for a, b in zip(single["reconstruction"]["views"], solved["reconstruction"]["views"]):
            np.testing.assert_array_equal(a["pose"]["R"], b["pose"]["R"])
            np.testing.assert_array_equal(a["pose"]["t"], b["pose"]["t"])
        merged = L.merge_local_bundle(merged, problem, solved["reconstruction"])
    B.validate_reconstruction(merged)
    from test_local_bundle import huber_cost
    assert huber_cost(merged) < huber_cost(model)
