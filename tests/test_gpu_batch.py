"""Batched solves: many uneven bundles in one tcv_ba_solve_batch call.

Every bundle of a batch gets exactly the arithmetic of its own solve (the per-bundle reductions
partition rows, tiles, tasks and cost blocks relative to the bundle's start, as a lone solve does),
so the batched result must equal the separate tcv_ba_solve results bit for bit; each bundle is also
held to the FP64 oracle bars of tests/test_gpu_parity.py.
"""
import numpy as np
import pytest

from oracle import oracle as O
from typecv_amd import bundle as B
from typecv_amd import synthetic as S

from test_gpu_parity import assert_solve_parity

pytestmark = pytest.mark.gpu


def controls_for(bundles, **kw):
    out = []
    for packed, info in bundles:
        out.append(B.make_control(packed.C, packed.P, packed.M, shared_focal=True, huber=1.0,
                                  iterations=kw.get("iterations", 12), max_cg=256,
                                  precision=kw.get("precision", "fp64")))
    return out


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_equals_separate_solves(device, precision):
    print(f"\n[batch] {precision}", flush=True)
    bundles = S.local_bundles(count=7, seed=8, cameras=(5, 14), points=(200, 1500))
    controls = controls_for(bundles, precision=precision)
    outs, report, finish = B.solve_batch(device, controls, [p for p, _ in bundles])
    assert len(finish) == 7 and np.all(finish > 0)
    for (packed, _), control, got in zip(bundles, controls, outs):
        alone, _ = B.solve_packed(device, control, packed)
        assert np.array_equal(got, alone)


def test_batch_matches_oracle(device):
    print("\n[batch] oracle", flush=True)
    bundles = S.local_bundles(count=5, seed=9, cameras=(5, 12), points=(200, 800))
    controls = controls_for(bundles)
    outs, report, _ = B.solve_batch(device, controls, [p for p, _ in bundles])
    assert report["lm_iterations"] > 0
    for (packed, _), control, got in zip(bundles, controls, outs):
        expected, _ = O.solve(control, packed)
        base = packed.C * 12 + packed.P * 3
        assert_solve_parity(got, expected, base)
        assert abs(got[base + 26] - expected[base + 26]) <= 1e-9


def test_batch_refuses_mixed_options(device):
    bundles = S.local_bundles(count=2, seed=10, cameras=(5, 6), points=(100, 200))
    controls = controls_for(bundles)
    controls[1].huber = 2.0
    from typecv_amd import _native
    with pytest.raises(_native.EngineError, match="scene/options"):
        B.solve_batch(device, controls, [p for p, _ in bundles])
