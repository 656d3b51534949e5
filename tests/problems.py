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


def repeat_tracks(packed, lengths: dict, seed: int = 0, noise_px: float = 0.5):
    """The packed scene with the tracks of some points lengthened by repeating their observations:
    `lengths` maps a point to (new track length, "cycle" | "grouped"). "cycle" walks the original
    track again and again (cameras 0 1 2 .. 0 1 2 ..), "grouped" repeats each observation in place
    (0 0 0 1 1 1 ..). Every repeat keeps its camera and gets N(0, noise_px^2) more pixel noise, so a
    camera observes the point several times: tracks over 64 observations whose cameras still fit a
    window of 64."""
    from typecv_amd.packing import pack_arrays
    import numpy as np

    rng = np.random.default_rng(seed)
    ps = packed.section("pointStart").astype(np.int64)
    cams = packed.section("observationCamera")
    pixels = packed.pixels()
    out_len, out_cam, out_px = [], [], []
    for p in range(packed.P):
        s, e = ps[p], ps[p + 1]
        n = e - s
        if p not in lengths:
            out_len.append(n)
            out_cam.append(cams[s:e])
            out_px.append(pixels[s:e])
            continue
        total, order = lengths[p]
        if order == "cycle":
            pick = np.arange(total) % n
        else:
            pick = np.minimum(np.arange(total) * n // total, n - 1)
        extra = np.ones(total, dtype=bool)
        extra[np.unique(pick, return_index=True)[1]] = False
        px = pixels[s + pick] + rng.normal(0, noise_px, size=(total, 2)) * extra[:, None]
        out_len.append(total)
        out_cam.append(cams[s + pick])
        out_px.append(px)
    return pack_arrays(packed.poses(), packed.models(), packed.points(), np.array(out_len),
                       np.concatenate(out_cam), np.concatenate(out_px),
                       packed.section("cameraActive"))


def shuffle_points(packed, seed: int = 0):
    """The packed scene with its points (and their tracks) in a random order: consecutive points no
    longer share cameras, so the one-pass product's camera windows pass 64 cameras and the two
    sweeps run."""
    from typecv_amd.packing import pack_arrays
    import numpy as np

    perm = np.random.default_rng(seed).permutation(packed.P)
    ps = packed.section("pointStart").astype(np.int64)
    order = np.concatenate([np.arange(ps[p], ps[p + 1]) for p in perm])
    return pack_arrays(packed.poses(), packed.models(), packed.points()[perm],
                       np.diff(ps)[perm], packed.section("observationCamera")[order],
                       packed.pixels()[order], packed.section("cameraActive"))
