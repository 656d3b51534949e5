"""Shared test helper: the packed problem and control that `bundle_adjust_gpu` builds for a
reconstruction and options (the reference's bundleProblem + packBundleScene)."""
from typecv_amd import bundle as B

FIXED = lambda C: min(96, max(24, 12 * C))  # noqa: E731
SHARED = lambda C: min(256, max(32, 12 * C + 1))  # noqa: E731


def packed_problem(model, options):
    shared = bool(options.get("refineSharedFocal"))
    problem = B.bundle_problem(model, options, SHARED if shared else FIXED)
    packed = B.pack_bundle_scene(problem["input"], problem["byId"], problem["active"])
    control = B.make_control(
        packed.C, packed.P, packed.M, directional=packed.P - problem["finitePoints"],
        anchor=problem["byId"][problem["anchorId"]], scale_view=problem["byId"][problem["scaleId"]],
        normalize_scale=problem["fixed"] == 1, shared_focal=shared,
        iterations=problem["iterations"], max_cg=problem["maxCg"], huber=problem["huber"],
        tolerance=problem["tolerance"], precision=options.get("precision", "fp64"))
    return packed, control
