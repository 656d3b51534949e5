"""The allocation-free observation evaluation (tcv_obs_evaluate_noalloc: device arrays in and out,
its scratch a fixed device array of the library) against tcv_obs_evaluate, the entry that
allocates, through the C ABI: flags and summary bit for bit, on two calls in a row (the second
with fewer observations, so it runs on the first call's leftover scratch), guarded and unguarded;
out-of-range indices refused before any kernel reads through them; more observations than the
scratch holds refused."""
import ctypes

import numpy as np
import pytest

from typecv_amd import _native
from typecv_amd import observations as OB
from typecv_amd import synthetic as S

pytestmark = pytest.mark.gpu


def trajectory_request(cameras, points, guarded):
    """The trajectory scene's observations at its starting poses, as observationRequest packs
    them (one camera record per view, landmarks in order)."""
    packed, _ = S.trajectory_scene(cameras=cameras, points=points, seed=2)
    C, P, M = packed.C, packed.P, packed.M
    sc, ix = packed.scene, packed.index
    poses, pts = sc[:12 * C].reshape(C, 12), sc[12 * C:12 * C + 3 * P]
    models = sc[12 * C + 3 * P:12 * C + 3 * P + 6 * C].reshape(C, 6)
    offsets = np.ascontiguousarray(ix[2 * M:2 * M + P + 1])
    return {
        "depth": {"maxRelativeDepthStdDev": 0.02 if guarded else 0.0, "pixelNoiseStdDevPx": 1.0},
        "guarded": guarded, "lengths": np.diff(offsets).tolist() if guarded else [], "count": M,
        "views": C, "points": P,
        "cameras": np.ascontiguousarray(np.concatenate([poses, models], axis=1).reshape(-1)),
        "points_xyz": np.ascontiguousarray(pts),
        "pixels": np.ascontiguousarray(sc[12 * C + 3 * P + 6 * C:]),
        "indices": np.ascontiguousarray(np.stack([ix[:M], ix[M:2 * M]], axis=1).reshape(-1)),
        "offsets": offsets if guarded else None,
    }


class DeviceArrays:
    """Device copies of host arrays through the HIP runtime the engine library links (the test
    allocates its inputs and outputs; the entry under test allocates nothing)."""

    def __init__(self):
        _native.library()  # loads libamdhip64.so.7 with the engine
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                       ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.held = []

    def upload(self, array: np.ndarray) -> int:
        array = np.ascontiguousarray(array)
        ptr = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(ptr), max(8, array.nbytes)) == 0
        self.held.append(ptr)
        if array.nbytes:
            assert self.hip.hipMemcpy(ptr, array.ctypes.data, array.nbytes, 1) == 0  # to device
        return ptr.value

    def download(self, ptr: int, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0  # to host
        return out

    def free(self):
        for ptr in self.held:
            self.hip.hipFree(ptr)
        self.held = []


def noalloc(request, max_error):
    """tcv_obs_evaluate_noalloc on device copies of the request: (status, flags, summary)."""
    lib = _native.library()
    count, guarded = request["count"], request["guarded"]
    tracks = len(request["lengths"]) if guarded else 0
    control = _native.ObsControl(
        views=request["views"], points=request["points"], observations=count, tracks=tracks,
        max_error_px=float(max_error),
        max_relative_depth_std_dev=request["depth"]["maxRelativeDepthStdDev"],
        pixel_noise_std_dev_px=request["depth"]["pixelNoiseStdDevPx"])
    mem = DeviceArrays()
    try:
        arrays = [mem.upload(request[k]) for k in ("cameras", "points_xyz", "pixels", "indices")]
        offsets = mem.upload(request["offsets"]) if guarded else None
        flags_host = np.full(max(1, count + tracks), 7, dtype=np.uint32)
        summary_host = np.full(2, -1.0)
        flags, summary = mem.upload(flags_host), mem.upload(summary_host)
        status = lib.tcv_obs_evaluate_noalloc(ctypes.byref(control), *arrays, offsets, flags,
                                              summary, None)
        return status, mem.download(flags, flags_host), mem.download(summary, summary_host)
    finally:
        mem.free()


@pytest.mark.parametrize("guarded", [False, True])
def test_noalloc_equals_the_allocating_entry(device, guarded):
    print(f"\n[noalloc] guarded={guarded}", flush=True)
    # the 500-camera scene's 1.89 M observations, then a 40-camera scene of far fewer
    for cameras, points in ((500, 200_000), (40, 3000)):
        request = trajectory_request(cameras, points, guarded)
        want_flags, want_summary, _ = OB.evaluate_arrays(device, request, 3.0)
        status, flags, summary = noalloc(request, 3.0)
        n = request["count"] + (len(request["lengths"]) if guarded else 0)
        print(f"  M={request['count']}: status {status}, inliers {summary[1]:.0f} "
              f"(allocating entry {want_summary[1]:.0f})", flush=True)
        assert status == 0, _native.library().tcv_ba_last_error().decode()
        assert np.array_equal(flags[:n], want_flags[:n])
        assert np.array_equal(summary, want_summary)


def test_noalloc_refuses_bad_indices_and_sizes():
    print("\n[noalloc] refusals", flush=True)
    request = trajectory_request(40, 3000, False)
    request["indices"] = request["indices"].copy()
    request["indices"][2 * 17 + 1] = request["points"]  # a landmark past the records
    status, _, _ = noalloc(request, 3.0)
    assert status == 2
    assert "outside" in _native.library().tcv_ba_last_error().decode()
    guarded = trajectory_request(40, 3000, True)
    guarded["offsets"] = guarded["offsets"].copy()
    guarded["offsets"][5], guarded["offsets"][6] = guarded["offsets"][6], guarded["offsets"][5] - 1
    status, _, _ = noalloc(guarded, 3.0)
    assert status == 2
    # sizes beyond the fixed scratch: refused before any device call
    lib = _native.library()
    control = _native.ObsControl(views=1, points=1,
                                 observations=lib.TCV_OBS_NOALLOC_MAX_OBSERVATIONS + 1, tracks=0,
                                 max_error_px=3.0, max_relative_depth_std_dev=0.0,
                                 pixel_noise_std_dev_px=1.0)
    one = ctypes.c_double(0)
    word = ctypes.c_uint32(0)
    status = lib.tcv_obs_evaluate_noalloc(ctypes.byref(control), ctypes.addressof(one),
                                          ctypes.addressof(one), ctypes.addressof(one),
                                          ctypes.addressof(word), None, ctypes.addressof(word),
                                          ctypes.addressof(one), None)
    assert status == 1
    assert "at most 2097152 observations" in lib.tcv_ba_last_error().decode()
