"""Tracks of more than 64 observations whose cameras still fit one product window, on the GPU.

A point observed more than 64 times gets a tile of its own that the one-pass Schur product walks in
64-slot passes: the point's terms summed pass by pass, then its camera terms added pass by pass.
The scenes here lengthen a few tracks by repeating their observations (a camera observing the point
several times, each repeat with its own pixel noise): tracks of exactly 64 (one full pass), 65 (a
second pass of one slot), 70, 96, 130 and 200, in both repeat orders, beside ordinary tiles.

Bars, as in tests/test_gpu_parity.py and tests/test_gpu_forms.py: the FP64 assembled system within
1e-10 relative and whole solves to cost 1e-9 relative, coordinates 1e-7·max(1, |x|), integer words
equal; the mixed form's Schur product within 1e-6 relative of the FP64 oracle's (FP32 Jacobian
factors, unit roundoff 6e-8, products and sums in FP64; measured 1e-8) and its solve within 1e-6
relative in cost; streamed solves equal to resident ones bit for bit.
"""
import numpy as np
import pytest

from oracle import oracle as O
from problems import repeat_tracks
from typecv_amd import bundle as B
from typecv_amd import synthetic as S

from test_gpu_budget import stored_bytes
from test_gpu_parity import assert_solve_parity

pytestmark = pytest.mark.gpu


def scenes():
    ring, _ = S.ring_scene(seed=3, cameras=10, points=400)
    ring = repeat_tracks(ring, {0: (65, "cycle"), 5: (130, "grouped"), 9: (64, "cycle"),
                                200: (200, "grouped"), 201: (70, "cycle")}, seed=1)
    yield "ring", ring, dict(iterations=20, max_cg=50)
    traj, _ = S.trajectory_scene(cameras=20, points=1500, seed=2)
    traj = repeat_tracks(traj, {0: (96, "grouped"), 1: (65, "cycle"), 700: (150, "cycle"),
                                1499: (70, "grouped")}, seed=2)
    yield "trajectory_shared", traj, dict(iterations=15, max_cg=256, shared_focal=True, huber=1.0)


SCENES = list(scenes())
IDS = [s[0] for s in SCENES]


def longest_track(packed) -> int:
    return int(np.max(np.diff(packed.section("pointStart").astype(np.int64))))


@pytest.mark.parametrize("name,packed,kw", SCENES, ids=IDS)
def test_long_track_probe_matches_oracle(device, name, packed, kw):
    print(f"\n[long tracks probe] {name} M={packed.M} longest={longest_track(packed)}", flush=True)
    assert longest_track(packed) > 128
    control = B.make_control(packed.C, packed.P, packed.M, **kw)
    m = packed.C * 6 + control.shared_focal
    p = np.random.default_rng(5).normal(size=m)
    for lam in (1e-3, 1.0):
        rhs, sp, z, status = O.schur_probe(control, packed, lam, 0.01, p)
        grhs, gsp, gz, gstatus, report = B.schur_probe(device, control, packed, lam, 0.01, p)
        assert report["onepass_rows"] > 0  # the one-pass product ran
        assert gstatus == status
        for a, b in ((grhs, rhs), (gsp, sp), (gz, z)):
            assert np.max(np.abs(a - b)) <= 1e-10 * max(1.0, np.max(np.abs(b)))


@pytest.mark.parametrize("name,packed,kw", SCENES, ids=IDS)
def test_long_track_mixed_probe_near_oracle(device, name, packed, kw):
    print(f"\n[long tracks mixed probe] {name}", flush=True)
    control = B.make_control(packed.C, packed.P, packed.M, **kw)
    mixed = B.make_control(packed.C, packed.P, packed.M, precision="mixed", **kw)
    m = packed.C * 6 + control.shared_focal
    p = np.random.default_rng(6).normal(size=m)
    _, sp, _, status = O.schur_probe(control, packed, 1e-3, 0.01, p)
    _, gsp, _, gstatus, report = B.schur_probe(device, mixed, packed, 1e-3, 0.01, p)
    assert report["onepass_rows"] > 0
    assert gstatus == status
    rel = np.max(np.abs(gsp - sp)) / np.max(np.abs(sp))
    print(f"  mixed S p against FP64: {rel:.2e} relative", flush=True)
    assert rel <= 1e-6


@pytest.mark.parametrize("name,packed,kw", SCENES, ids=IDS)
def test_long_track_solve_matches_oracle(device, name, packed, kw):
    print(f"\n[long tracks solve] {name}", flush=True)
    control = B.make_control(packed.C, packed.P, packed.M, **kw)
    expected, _ = O.solve(control, packed)
    got, report = B.solve_packed(device, control, packed)
    base = packed.C * 12 + packed.P * 3
    assert report["onepass_rows"] > 0
    assert expected[base + 6] < expected[base + 8]
    assert_solve_parity(got, expected, base)


@pytest.mark.parametrize("name,packed,kw", SCENES, ids=IDS)
def test_long_track_mixed_solve_against_fp64_oracle(device, name, packed, kw):
    print(f"\n[long tracks mixed solve] {name}", flush=True)
    control = B.make_control(packed.C, packed.P, packed.M, **kw)
    expected, _ = O.solve(control, packed)
    got, report = B.solve_packed(
        device, B.make_control(packed.C, packed.P, packed.M, precision="mixed", **kw), packed)
    base = packed.C * 12 + packed.P * 3
    cost = expected[base + 6]
    rel = abs(got[base + 6] - cost) / max(1.0, abs(cost))
    print(f"  final cost {got[base + 6]:.12g} vs {cost:.12g} (rel {rel:.2e})", flush=True)
    assert report["onepass_rows"] > 0
    assert rel <= 1e-6


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("name,packed,kw", SCENES, ids=IDS)
def test_long_track_streamed_equals_resident(name, packed, kw, precision):
    """The streamed product relinearises each pass's observations in registers: its long tiles
    against the resident ones that read the stored words."""
    print(f"\n[long tracks streamed] {name} {precision}", flush=True)
    control = B.make_control(packed.C, packed.P, packed.M, precision=precision, **kw)
    with B.HipDevice(0) as device:
        resident, report = B.solve_packed(device, control, packed)
        assert report["streamed"] == 0 and report["onepass_rows"] > 0
        held = device.statistics["liveBytes"]
    floor = held - stored_bytes(control, packed)
    with B.HipDevice(0, working_bytes=floor) as device:
        streamed, report = B.solve_packed(device, control, packed)
        assert report["streamed"] == 1 and report["onepass_rows"] > 0
        assert np.array_equal(streamed, resident)
