"""The working budget, the streamed route and the failure paths of the C ABI, on the GPU.

* Streamed = resident, bit for bit (reference tests/typeshade/bundle-streamed.test.mjs:67-128): a
  budget below the resident arrays but above the floor runs the solve streamed (every Schur sweep
  relinearises the cross blocks in registers instead of reading them) with the resident run's
  output on both routes and in both arithmetic forms, and the context holds no more than the budget.
* Budget refusal (focal.test.mjs:75-80): below the floor the solve is refused with
  TCV_BA_ERR_MEMORY and the reference's "GPU BA needs N bytes ... working budget" text before
  anything is allocated; the context holds no buffers afterwards.
* An injected read failure (cpu.test.mjs:508-531): the solve fails with TCV_BA_ERR_DEVICE, the
  context drops its buffers, and the next solve returns the undisturbed result.
* One context used from two threads at once: the calls run one after another (the context's lock)
  and give the sequential results bit for bit.
"""
import ctypes
import threading

import numpy as np
import pytest

from typecv_amd import _native
from typecv_amd import bundle as B
from typecv_amd import synthetic as S

pytestmark = pytest.mark.gpu


def tile_slots(control, packed) -> int:
    """Slots of the engine's tiles of whole points (at most 64 observations and 64 points, a longer
    point alone), each tile a whole number of 64-slot passes: the count the engine sizes its
    camera-major stored words by."""
    M, P = control.observations, control.points
    ps = np.asarray(packed.index[2 * M: 2 * M + P + 1], dtype=np.int64)
    slots, first = 0, 0
    for p in range(1, P + 1):
        if p == P or ps[p + 1] - ps[first] > 64 or p - first >= 64:
            slots += (ps[p] - ps[first] + 63) // 64 * 64
            first = p
    return int(slots)


def stored_bytes(control, packed) -> int:
    """Bytes of the arrays only a resident solve holds: each observation's Jacobian factors
    (FP64: K = 8 words, 6 on the fixed route; FP32 in the mixed form: 11 or 9), point-major and
    camera-major (K words per slot of 64-slot blocks)."""
    mixed = control.precision == _native.MIXED
    K = (11 if control.shared_focal else 9) if mixed else (8 if control.shared_focal else 6)
    ws = 4 if mixed else 8
    return (K * control.observations + K * tile_slots(control, packed)) * ws


def scenes():
    ring, _ = S.ring_scene(seed=3, cameras=10, points=400)
    yield "fixed", ring, dict(iterations=10, max_cg=50, huber=2.0)
    traj, _ = S.trajectory_scene(cameras=20, points=3000, seed=2)
    yield "shared", traj, dict(iterations=10, max_cg=241, shared_focal=True, huber=1.0)


SCENES = list(scenes())


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("name,packed,kw", SCENES, ids=[s[0] for s in SCENES])
def test_streamed_equals_resident(name, packed, kw, precision):
    print(f"\n[budget] streamed vs resident: {name} {precision}", flush=True)
    control = B.make_control(packed.C, packed.P, packed.M, precision=precision, **kw)
    with B.HipDevice(0) as device:
        resident, report = B.solve_packed(device, control, packed)
        assert report["streamed"] == 0
        held = device.statistics["liveBytes"]
    floor = held - stored_bytes(control, packed)
    for budget in (held - 1, floor):
        with B.HipDevice(0, working_bytes=budget) as device:
            streamed, report = B.solve_packed(device, control, packed)
            assert report["streamed"] == 1, budget
            assert device.statistics["liveBytes"] <= budget
            assert np.array_equal(streamed, resident), budget
    with B.HipDevice(0, working_bytes=held) as device:
        again, report = B.solve_packed(device, control, packed)
        assert report["streamed"] == 0
        assert np.array_equal(again, resident)


def test_budget_refusal_before_allocation():
    print("\n[budget] refusal", flush=True)
    name, packed, kw = SCENES[1]
    control = B.make_control(packed.C, packed.P, packed.M, **kw)
    with B.HipDevice(0) as device:
        expected, _ = B.solve_packed(device, control, packed)
        floor = device.statistics["liveBytes"] - stored_bytes(control, packed)
    with B.HipDevice(0, working_bytes=floor - 1) as device:
        assert device.statistics["liveBytes"] == 0
        with pytest.raises(_native.EngineError, match="GPU BA needs") as err:
            B.solve_packed(device, control, packed)
        assert err.value.status == 4
        assert f"needs {floor} bytes" in str(err.value)
        assert f"working budget of {floor - 1} bytes" in str(err.value)
        assert device.statistics["liveBytes"] == 0
        # the budget raised to the floor: the same device completes the solve
        device.set_working_bytes(floor)
        got, report = B.solve_packed(device, control, packed)
        assert report["streamed"] == 1
        assert np.array_equal(got, expected)


def test_a_context_over_its_budget_releases_first():
    """Buffers held from a larger problem are released when the next solve's plan would pass the
    budget with them."""
    print("\n[budget] held buffers", flush=True)
    big = SCENES[1]
    small, _ = S.trajectory_scene(cameras=12, points=800, seed=4)
    kw = big[2]
    with B.HipDevice(0) as device:
        B.solve_packed(device, B.make_control(big[1].C, big[1].P, big[1].M, **kw), big[1])
        held = device.statistics["liveBytes"]
        control = B.make_control(small.C, small.P, small.M, **kw)
        alone_device = B.HipDevice(0)
        expected, _ = B.solve_packed(alone_device, control, small)
        small_held = alone_device.statistics["liveBytes"]
        alone_device.dispose()
        assert small_held < held
        device.set_working_bytes(small_held)
        got, report = B.solve_packed(device, control, small)
        assert report["streamed"] == 0
        assert device.statistics["liveBytes"] <= small_held
        assert np.array_equal(got, expected)


@pytest.mark.parametrize("read", [1, 3, 40])
def test_injected_read_failure_drops_the_buffers(device, read):
    print(f"\n[failure] read {read} fails", flush=True)
    name, packed, kw = SCENES[1]
    control = B.make_control(packed.C, packed.P, packed.M, **kw)
    expected, _ = B.solve_packed(device, control, packed)
    assert device.statistics["liveBytes"] > 0
    lib = _native.library()
    _native.check(lib.tcv_ba_fail_reads(device.handle, read), "tcv_ba_fail_reads")
    with pytest.raises(_native.EngineError, match="injected read failure") as err:
        B.solve_packed(device, control, packed)
    assert err.value.status == 3
    assert device.statistics["liveBytes"] == 0
    got, _ = B.solve_packed(device, control, packed)
    assert np.array_equal(got, expected)


def test_one_context_from_two_threads(device):
    print("\n[threads] two solves on one context at once", flush=True)
    lib = _native.library()
    work = []
    for name, packed, kw in SCENES:
        control = B.make_control(packed.C, packed.P, packed.M, **kw)
        expected, _ = B.solve_packed(device, control, packed)
        work.append((control, packed, expected))
    outs = [np.empty_like(w[2]) for w in work]
    status = [None, None]

    def run(k):
        control, packed, _ = work[k]
        scene = np.ascontiguousarray(packed.scene)
        index = np.ascontiguousarray(packed.index)
        report = _native.Report()
        # the C ABI directly, without the Python device's own job lock
        status[k] = lib.tcv_ba_solve(device.handle, ctypes.byref(control), scene.ctypes.data,
                                     index.ctypes.data, outs[k].ctypes.data, ctypes.byref(report))

    for _ in range(3):
        threads = [threading.Thread(target=run, args=(k,)) for k in range(2)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        assert status == [0, 0]
        for k in range(2):
            assert np.array_equal(outs[k], work[k][2])
