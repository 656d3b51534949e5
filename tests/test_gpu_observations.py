"""Observation evaluation on the device against the oracle's restatement of the reference's CPU
tier (tcv_oracle_observations), on the reference's ten cases and a 160,000-observation scene.

Bars: the flags (masks and depth rejections) equal the oracle's, bit for bit; the summary (error
sum and inlier count) equals the oracle's bit for bit — the kernels compute without fused
multiply-adds and sum in the CPU tier's order (64-observation partials in lane order, then the
partials in block order)."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from typecv_amd import observations as OB

from obs_cases import observation_cases, observation_model

pytestmark = pytest.mark.gpu

CASES = observation_cases()


def oracle_of(model, max_error, options, last_only):
    req = OB.observation_request(model, max_error, options, last_only)
    flags, _, summary = O.observations(
        req["cameras"], req["points_xyz"], req["pixels"], req["indices"][:2 * req["count"]],
        req["offsets"], max_error, req["depth"]["maxRelativeDepthStdDev"],
        req["depth"]["pixelNoiseStdDevPx"])
    return OB.unpack_observations(summary, flags, req["lengths"], req["count"], req["guarded"])


def same(got, want):
    assert len(got["masks"]) == len(want["masks"])
    for a, b in zip(got["masks"], want["masks"]):
        assert a.tolist() == b.tolist()
    assert ("depthRejected" in got) == ("depthRejected" in want)
    if "depthRejected" in want:
        assert got["depthRejected"].tolist() == want["depthRejected"].tolist()
    assert got["inlierCount"] == want["inlierCount"]
    if math.isinf(want["meanErrorPx"]):
        assert math.isinf(got["meanErrorPx"])
    else:
        assert got["meanErrorPx"] == want["meanErrorPx"], (got["meanErrorPx"], want["meanErrorPx"])


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_cases_match_the_oracle(device, case):
    options = case.get("options", {})
    print(f"\n[observations] {case['name']}", flush=True)
    got = OB.evaluate_observations_gpu(device, case["model"], 3, options)
    want = oracle_of(case["model"], 3, options, False)
    print(f"  inliers {got['inlierCount']:.0f}, mean error {got['meanErrorPx']!r}", flush=True)
    same(got, want)


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_continued_cases_match_the_oracle(device, case):
    options = case.get("options", {})
    guarded = OB.depth_sensitivity_controls(options)["maxRelativeDepthStdDev"] > 0
    got = OB.evaluate_continued_observations_gpu(device, case["model"], 3, options)
    want = oracle_of(case["model"], 3, options, not guarded)
    same(got, want)


def test_large_scene_matches_the_oracle(device):
    """20,000 landmarks in 8 views (160,000 observations), one outlier in 97, thresholds 3 px and
    guarded at 1 %."""
    model = observation_model(points=20000, views=8, baseline=0.2)
    for p, lm in enumerate(model["landmarks"]):
        if p % 97 == 0:
            lm["observations"][p % 8]["pixel"][1] += 5.0
    print("\n[observations] 160,000 observations", flush=True)
    for options in ({}, {"maxRelativeDepthStdDev": 0.01, "pixelNoiseStdDevPx": 0.7}):
        got = OB.evaluate_observations_gpu(device, model, 3, options)
        want = oracle_of(model, 3, options, False)
        print(f"  {options}: inliers {got['inlierCount']:.0f}", flush=True)
        same(got, want)
        if options:
            assert 0 < got["depthRejected"].sum() < 20000
