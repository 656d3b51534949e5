"""Benchmark of the MI355X bundle-adjustment engine on the shared-focal trajectory scene.

Workload (one GPU): 2,000 cameras on a smooth closed trajectory, 1,000,000 points each seen by
5-20 cameras (~9.7M observations), N(0, 0.7 px) noise, 2 % uniform outliers, common focal error
U(0.95, 1.05); joint pose/point/shared-focal LM with Huber delta = 1 px, 30 LM iterations, PCG up
to 256 iterations (the route's default). A step is one whole solve from the device-resident packed
scene (tcv_ba_solve_device); the scene is generated and uploaded before timing.

Prints one JSON line: the plain FP64 form's LM iterations/s over the timed steps (`value`),
observations/s and the Schur matvec roofline (algorithmic bytes of the two sweeps / their HIP-event
time); the mixed form on the same scene in the same process (`mixed`: FP32 cross blocks and
products with FP64 accumulation, refinement of the FP64 residual to 1e-10; its speed-up and its
final-cost, pose and point differences from the FP64 run); the directSolve route; the 500-camera
scene and the 256 local bundles as sub-records; and the CPU oracle timed on bounded samples of the
same solve on the host cores.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = ("LM iters/sec + obs/sec, mixed vs FP64 at ≤1e-6 rel cost; Schur matvec HBM GB/s")
HBM_PEAK_GBS = 8000.0

# HBM bytes of one Schur product measured with `rocprofv3 --pmc FETCH_SIZE` and `--pmc WRITE_SIZE`
# over `bench.py --mode probe` (profiles/schur_trajectory2000_traffic.md): FETCH_SIZE x 2.00 and
# WRITE_SIZE x 1.00, the factors that make a 1 GiB device copy read and write 1 GiB. Keyed by
# (cameras, points, seed, precision, operator); other workloads report null.
MEASURED_TRAFFIC = {
    # the one-pass product over FP64 or FP32 factors as the PCG loop runs it (k_schur_onepass +
    # k_onepass_product): HBM bytes per product, 2 x FETCH_SIZE + WRITE_SIZE (KiB per dispatch)
    (2000, 1_000_000, 2, "fp64", "onepass"):
        2 * (391699.6 + 4384.9) * 1024 + (5536.4 + 140.6) * 1024,
    (2000, 1_000_000, 2, "mixed", "onepass"):
        2 * (288646.1 + 4384.9) * 1024 + (5536.1 + 140.6) * 1024,
}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def matvec_bytes(C: int, P: int, M: int, shared: bool, factored: bool = False) -> int:
    """Algorithmic HBM bytes of one Schur product (both sweeps): every observation's stored words
    read once per sweep (FP64 cross blocks: 21 doubles shared, 18 fixed; the mixed form's FP32
    Jacobian factors: 11 or 9 floats, with each camera's FP64 rotation read once per sweep), the
    camera id (point sweep) and point id (camera sweep) of each observation, V^-1 and the point
    offsets, y written and read back, the camera blocks, offsets and vectors."""
    if factored:
        per_obs = 2 * (11 if shared else 9) * 4
        per_cam = 2 * 9 * 8
    else:
        per_obs = 2 * (21 if shared else 18) * 8
        per_cam = 0
    total = M * (per_obs + 8) + P * (9 * 8 + 4 + 24 + 24) + C * (36 * 8 + 4 + 3 * 6 * 8 + per_cam)
    if shared:
        total += C * 6 * 8 * 2 + 24
    return total


def onepass_bytes(C: int, P: int, M: int, shared: bool, factored: bool, rows: int,
                  tiles: int) -> int:
    """Algorithmic HBM bytes of one one-pass Schur product as the PCG loop runs it (k_schur_onepass +
    k_onepass_product): every observation's stored Jacobian factors read once (FP32 in the mixed
    form, `factored`: 11 words shared, 9 fixed; FP64 in the FP64 form: 8 or 6, the camera-frame
    point recomputed from each point's coordinates, read once, and the window's poses) and its code
    word (window row, rank, tile point, place in its point's run); V^-1 of every point; each
    tile's 8-word record; per camera p, its rotation (FP64 form: its pose) and damped block A
    read, its rows of S p and its p.Sp and focal share written; the rows of camera partials
    written once and read once (8 doubles each)."""
    if factored:
        per_obs, per_point, pose = (11 if shared else 9) * 4, 9 * 8, 9 * 8
    else:
        per_obs, per_point, pose = (8 if shared else 6) * 8, 12 * 8, 12 * 8
    per_cam = 6 * 8 + pose + 36 * 8 + 8 * 8
    return (M * (per_obs + 4) + P * per_point + tiles * 8 * 4 + C * per_cam + rows * 8 * 8 * 2
            + (8 if shared else 0))


def solve_forms(args, B, _native, lib, device, packed, scene, index, out, precision, steps,
                warmup, refine=0.0, direct=False, label=""):
    """`warmup` untimed and `steps` timed device-resident solves of one form of the workload:
    (reports, elapsed seconds of the timed solves, final output array)."""
    import torch

    C, P, M = packed.C, packed.P, packed.M
    control = B.make_control(C, P, M, shared_focal=True, huber=1.0, iterations=args.iterations,
                             max_cg=min(256, max(32, 12 * C + 1)), precision=precision,
                             refinement_tolerance=refine, timing=True, direct=direct)

    def solve() -> dict:
        report = _native.Report()
        status = lib.tcv_ba_solve_device(device.handle, ctypes.byref(control), scene.data_ptr(),
                                         index.data_ptr(), out.data_ptr(), ctypes.byref(report),
                                         None)
        _native.check(status, "tcv_ba_solve_device")
        return report.as_dict()

    for k in range(warmup):
        r = solve()
        log(f"{label} warmup {k}: {r['solve_ms']:.1f} ms, {r['lm_iterations']} LM, "
            f"{r['cg_iterations']} CG, {r['refinement_steps']:.0f} refinements")
    torch.cuda.synchronize()
    reports = []
    start = time.perf_counter()
    for k in range(steps):
        reports.append(solve())
        log(f"{label} step {k}: {reports[-1]['solve_ms']:.1f} ms")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - start
    return control, reports, elapsed, out.cpu().numpy()


def form_record(packed, seed, precision, operator, reports, elapsed, values):
    """LM iterations/s, observations/s and the Schur product's roofline of timed solves."""
    C, P, M = packed.C, packed.P, packed.M
    steps = len(reports)
    lm = sum(r["lm_iterations"] for r in reports)
    cg = sum(r["cg_iterations"] for r in reports)
    timed = sum(r["timed_matvecs"] for r in reports)  # the engine times every 8th product
    rows = reports[-1].get("onepass_rows", 0)
    onepass = rows > 0
    bytes_per = (onepass_bytes(C, P, M, True, precision == "mixed", rows,
                               reports[-1]["point_tiles"]) if onepass else
                 matvec_bytes(C, P, M, True, factored=precision == "mixed"))
    mv_ms = sum(r["schur_ms"] for r in reports) / max(1, timed)
    achieved = bytes_per / (mv_ms * 1e-3) / 1e9 if mv_ms > 0 else 0.0
    base = C * 12 + P * 3
    fused = reports[-1]["fused_product"]
    return {
        "value": lm / elapsed, "unit": "LM iterations/s", "steps": steps,
        "ms_per_step": elapsed * 1e3 / steps,
        "observations_per_s": M * lm / elapsed,
        "pcg_iterations_per_s": cg / elapsed,
        "lm_iterations_per_step": lm / steps, "cg_iterations_per_step": cg / steps,
        "matvecs_per_step": sum(r["matvecs"] for r in reports) / steps,
        "refinements_per_step": sum(r["refinement_steps"] for r in reports) / steps,
        "direct_solves_per_step": sum(r["direct_solves"] for r in reports) / steps,
        "final_cost": float(values[base + 6]), "initial_cost": float(values[base + 8]),
        "final_rmse_px": float(values[base + 25]),
        "schur_matvec": {
            "ms": mv_ms, "point_sweep_ms": sum(r["schur_point_ms"] for r in reports) / max(1, timed),
            "camera_sweep_ms": sum(r["schur_camera_ms"] for r in reports) / max(1, timed),
            "bytes": bytes_per, "timed_products": timed, "onepass_rows": rows,
            "share_of_solve": mv_ms * sum(r["matvecs"] for r in reports)
            / max(1e-9, sum(r["solve_ms"] for r in reports)),
            # camera-sweep workgroups per CU picked by the context's schedule tuner (0: no cap)
            "camera_sweep_cap": reports[-1]["camera_sweep_cap"],
            # 1: the tuner picked the fused FP64 product (k_point_sweep_fused + k_camera_tsum)
            "fused_product": fused,
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": MEASURED_TRAFFIC.get((C, P, seed, precision,
                                              "onepass" if onepass else operator)),
            "algorithmic_bytes": bytes_per,
            "kernel": ("Schur product: k_schur_onepass + k_onepass_product" if onepass else
                       "Schur matvec: k_point_sweep_factored + k_camera_sweep_factored"
                       if precision == "mixed" else
                       "Schur matvec: k_point_sweep_fused + k_camera_tsum" if fused else
                       "Schur matvec: k_point_sweep + k_camera_sweep"),
        },
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cameras", type=int, default=2000)
    ap.add_argument("--points", type=int, default=1_000_000)
    ap.add_argument("--iterations", type=int, default=30)
    ap.add_argument("--refine", type=float, default=1e-10,
                    help="mixed form: FP64 residual target of the refinement")
    ap.add_argument("--sub-steps", type=int, default=0,
                    help="timed solves of the mixed, direct and sub-workload records "
                         "(0: min(steps, 5))")
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--no-subworkloads", action="store_true",
                    help="skip the 500-camera and 256-bundle records")
    ap.add_argument("--cpu", action=argparse.BooleanOptionalAction, default=True,
                    help="time the CPU oracle baselines")
    ap.add_argument("--forms", default="fp64,mixed,direct",
                    help="the forms of the main workload to run (a profile takes one at a time); "
                         "the value is the first")
    args = ap.parse_args()
    forms = args.forms.split(",")
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the matvec would shard by points with one all-reduce of 6C+1 doubles per PCG "
              "iteration); run it with --gpus 1", file=sys.stderr)
        return 2
    sub_steps = args.sub_steps or min(args.steps, 5)

    import numpy as np
    import torch

    from typecv_amd import _native
    from typecv_amd import bundle as B
    from typecv_amd import synthetic as S

    t0 = time.time()
    packed, info = S.trajectory_scene(cameras=args.cameras, points=args.points, seed=args.seed)
    C, P, M = packed.C, packed.P, packed.M
    log(f"scene C={C} P={P} M={M} generated in {time.time() - t0:.1f}s")
    lib = _native.library()
    device = B.HipDevice(0)
    scene = torch.from_numpy(packed.scene).cuda()
    index = torch.from_numpy(packed.index.view(np.int32)).cuda()
    control0 = B.make_control(C, P, M, shared_focal=True)
    out = torch.empty(lib.tcv_ba_output_size(ctypes.byref(control0)), dtype=torch.float64,
                      device="cuda")
    torch.cuda.synchronize()
    base = C * 12 + P * 3

    # the value: the plain FP64 form, exactly `steps` timed solves
    first = "mixed" if forms[0] == "mixed" else "fp64"
    control, reports, elapsed, f64 = solve_forms(args, B, _native, lib, device, packed, scene,
                                                 index, out, first, args.steps, args.warmup,
                                                 refine=args.refine if first == "mixed" else 0.0,
                                                 label=first)
    plain = form_record(packed, args.seed, first, "stored", reports, elapsed, f64)
    result = {
        "metric": METRIC,
        "value": plain["value"],
        "unit": "LM iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": plain["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if first == "fp64" else "f32-products/f64-accumulation",
        "data": "synthetic (seeded trajectory scene, generated on the host, resident in HBM)",
        "config": {
            "workload": "shared-focal LM bundle adjustment, trajectory scene",
            "cameras": C, "points": P, "observations": M, "lm_iterations_max": args.iterations,
            "max_cg": control.max_cg, "huber_px": 1.0, "precision": first,
            "outliers": info["outliers"], "focal_factor": info["focal_factor"],
            # the one-pass Schur product over each observation's stored Jacobian factors (FP64
            # in the FP64 form, FP32 in the mixed one)
            "operator": "stored Jacobian factors" if plain.get("schur_matvec", {}).get(
                "onepass_rows") else "stored cross blocks",
            "linear_solver": "matrix-free PCG",
        },
    }
    for key in ("observations_per_s", "pcg_iterations_per_s", "lm_iterations_per_step",
                "cg_iterations_per_step", "matvecs_per_step", "final_cost", "initial_cost",
                "final_rmse_px", "schur_matvec", "roofline"):
        result[key] = plain[key]

    # the mixed form on the same scene, in the same process: FP32 cross blocks and Schur products,
    # FP64 accumulation, V^-1, preconditioner, dots and LM; refinement of the FP64 residual
    if first != "fp64" or "mixed" not in forms:
        return finish(args, result, device, packed, control, plain)
    _, reports, elapsed, mx = solve_forms(args, B, _native, lib, device, packed, scene, index,
                                          out, "mixed", sub_steps, 1, refine=args.refine,
                                          label="mixed")
    mixed = form_record(packed, args.seed, "mixed", "stored", reports, elapsed, mx)
    cost = f64[base + 6]
    mixed.update({
        "dtype": "f32-products/f64-accumulation", "refinement_tolerance": args.refine,
        "speedup_vs_fp64": mixed["value"] / plain["value"],
        "final_cost_rel_err_vs_fp64": abs(mx[base + 6] - cost) / cost,
        "max_pose_abs_diff_vs_fp64": float(np.abs(mx[:12 * C] - f64[:12 * C]).max()),
        "max_point_abs_diff_vs_fp64": float(np.abs(mx[12 * C:base] - f64[12 * C:base]).max()),
        "log_focal_diff_vs_fp64": float(abs(mx[base + 26] - f64[base + 26])),
    })
    result["mixed"] = mixed
    log(f"mixed: {mixed['value']:.2f} LM it/s ({mixed['speedup_vs_fp64']:.2f}x FP64), cost "
        f"{mixed['final_cost_rel_err_vs_fp64']:.1e} from FP64")
    # the mixed form refined to the PCG tolerance only (the linear solve the FP64 form does)
    _, reports, elapsed, mt = solve_forms(args, B, _native, lib, device, packed, scene, index,
                                          out, "mixed", sub_steps, 1, refine=0.0,
                                          label="mixed at the PCG tolerance")
    at_tol = form_record(packed, args.seed, "mixed", "stored", reports, elapsed, mt)
    result["mixed_at_pcg_tolerance"] = {
        "value": at_tol["value"], "unit": "LM iterations/s", "steps": at_tol["steps"],
        "ms_per_step": at_tol["ms_per_step"], "observations_per_s": at_tol["observations_per_s"],
        "refinement_tolerance": control.relative_tolerance,
        "cg_iterations_per_step": at_tol["cg_iterations_per_step"],
        "refinements_per_step": at_tol["refinements_per_step"],
        "speedup_vs_fp64": at_tol["value"] / plain["value"],
        "final_cost_rel_err_vs_fp64": abs(mt[base + 6] - cost) / cost,
        "max_pose_abs_diff_vs_fp64": float(np.abs(mt[:12 * C] - f64[:12 * C]).max()),
        "max_point_abs_diff_vs_fp64": float(np.abs(mt[12 * C:base] - f64[12 * C:base]).max()),
    }
    log(f"mixed at the PCG tolerance: {at_tol['value']:.2f} LM it/s "
        f"({at_tol['value'] / plain['value']:.2f}x FP64)")

    # the directSolve option on the same scene: a dense Cholesky solve per LM iteration
    if "direct" not in forms:
        return finish(args, result, device, packed, control, plain)
    _, reports, elapsed, dv = solve_forms(args, B, _native, lib, device, packed, scene, index,
                                          out, "fp64", sub_steps, 1, direct=True, label="direct")
    d_lm = sum(r["lm_iterations"] for r in reports)
    result["direct_route"] = {
        "value": d_lm / elapsed, "unit": "LM iterations/s", "steps": len(reports),
        "ms_per_step": elapsed * 1e3 / len(reports),
        "direct_solves_per_step": sum(r["direct_solves"] for r in reports) / len(reports),
        "cg_iterations_per_step": sum(r["cg_iterations"] for r in reports) / len(reports),
        "final_cost": float(dv[base + 6]), "final_rmse_px": float(dv[base + 25]),
        "linear_solver": "directSolve: dense Cholesky (64-column panels), certified by the "
                         "matrix-free residual",
    }
    return finish(args, result, device, packed, control, plain)


def finish(args, result, device, packed, control, plain) -> int:
    """The sub-workload records and the CPU baselines, then the JSON line."""
    import torch

    device.dispose()
    torch.cuda.empty_cache()
    sub_steps = args.sub_steps or min(args.steps, 5)
    if not args.no_subworkloads:
        result["trajectory_500"] = trajectory_500(args, sub_steps)
        result["batched_256"] = batched_256(args, sub_steps)
    if args.cpu:
        result["cpu_baseline"] = cpu_baseline(packed, control, plain)
        result["cpu_dense_ring_10"] = cpu_dense_ring()
    print(json.dumps(result), flush=True)
    return 0


def trajectory_500(args, steps) -> dict:
    """The 500-camera / 200,000-point trajectory scene in both forms (a sub-record)."""
    import numpy as np
    import torch

    from typecv_amd import _native
    from typecv_amd import bundle as B
    from typecv_amd import synthetic as S

    packed, _ = S.trajectory_scene(cameras=500, points=200_000, seed=args.seed)
    log(f"sub-workload C={packed.C} P={packed.P} M={packed.M}")
    lib = _native.library()
    scene = torch.from_numpy(packed.scene).cuda()
    index = torch.from_numpy(packed.index.view(np.int32)).cuda()
    control0 = B.make_control(packed.C, packed.P, packed.M, shared_focal=True)
    out = torch.empty(lib.tcv_ba_output_size(ctypes.byref(control0)), dtype=torch.float64,
                      device="cuda")
    record = {"config": {"cameras": packed.C, "points": packed.P, "observations": packed.M}}
    with B.HipDevice(0) as device:
        for precision in ("fp64", "mixed"):
            _, reports, elapsed, values = solve_forms(
                args, B, _native, lib, device, packed, scene, index, out, precision, steps, 1,
                refine=args.refine if precision == "mixed" else 0.0, label=f"500 {precision}")
            record[precision] = form_record(packed, args.seed, precision, "stored", reports,
                                            elapsed, values)
    record["mixed"]["speedup_vs_fp64"] = record["mixed"]["value"] / record["fp64"]["value"]
    return record


def batched_256(args, steps) -> dict:
    """256 local bundles in one tcv_ba_solve_batch per step (a sub-record): the directSolve route
    and the mixed PCG route, bundles/s and per-bundle LM-loop latency."""
    import numpy as np

    from typecv_amd import bundle as B
    from typecv_amd import synthetic as S

    bundles = S.local_bundles(count=256, seed=4)
    packeds = [p for p, _ in bundles]
    record = {"config": {"bundles": len(packeds), "observations": sum(p.M for p in packeds),
                         "cameras": "U{5..40}", "points": "U{500..20000}"}}
    with B.HipDevice(0) as device:
        for name, precision, direct, n in (("direct_fp64", "fp64", True, steps),
                                           ("pcg_mixed", "mixed", False, 1),
                                           ("pcg_fp64", "fp64", False, 1)):
            controls = [B.make_control(p.C, p.P, p.M, shared_focal=True, huber=1.0,
                                       iterations=30, max_cg=min(256, max(32, 12 * p.C + 1)),
                                       precision=precision, direct=direct) for p in packeds]
            loops, walls, latencies, totals = run_batch(B, device, controls, packeds, n, 1)
            lat = np.array(latencies)
            record[name] = {
                "bundles_per_s": len(packeds) * n / sum(loops), "steps": n,
                "ms_per_step": 1e3 * sum(loops) / n,
                "latency_ms": {"p50": float(np.percentile(lat, 50)),
                               "p99": float(np.percentile(lat, 99))},
                "end_to_end_bundles_per_s": len(packeds) * n / sum(walls),
                "pcg_iterations_per_step": totals["cg_iterations"] / n,
                "direct_solves_per_step": totals["direct_solves"] / n,
            }
    return record


def cpu_baseline(packed, control, plain) -> dict:
    """The C oracle (the reference's CPU-tier algorithm restated, -O3, OpenMP) on the host cores,
    on bounded samples of the same solve: the first LM iteration on every host thread, and the
    first LM iteration with its PCG capped at 4 iterations on one thread."""
    from oracle import oracle as O
    from typecv_amd import bundle as B

    threads = O.set_threads(0)
    C, P, M = packed.C, packed.P, packed.M
    log(f"CPU oracle sample: 1 LM iteration on {threads} threads")
    _, st = O.solve(control, packed, max_lm=1)
    lm_s = st["lm_ms"] * 1e-3
    one = B.make_control(C, P, M, shared_focal=True, huber=1.0, iterations=1, max_cg=4)
    O.set_threads(1)
    log("CPU oracle sample: 1 LM iteration, PCG <= 4, on 1 thread")
    _, st1 = O.solve(one, packed, max_lm=1)
    O.set_threads(0)
    return {
        "value": st["lm_iterations"] / lm_s, "unit": "LM iterations/s", "cores": threads,
        "kind": "port",
        "sample": f"first {int(st['lm_iterations'])} LM iteration of the same solve "
                  f"({int(st['cg_iterations'])} PCG iterations, {lm_s:.1f} s), C oracle -O3 "
                  f"OpenMP on {threads} threads",
        "pcg_iterations_per_s": st["cg_iterations"] / lm_s,
        "gpu_vs_cpu": {"lm_iterations": plain["value"] / (st["lm_iterations"] / lm_s),
                       "pcg_iterations": plain["pcg_iterations_per_s"]
                       / (st["cg_iterations"] / lm_s)},
        "single_thread": {
            "sample": f"first LM iteration with PCG capped at 4 ({int(st1['cg_iterations'])} "
                      f"PCG iterations, {st1['lm_ms'] * 1e-3:.1f} s), 1 thread",
            "seconds": st1["lm_ms"] * 1e-3, "cores": 1,
        },
    }


def cpu_dense_ring() -> dict:
    """The reference's dense reduced-camera CPU solver (bundleAdjust, bundle-adjustment.ts:80-291,
    restated in the oracle) on the 10-camera ring (2,000 points, fixed focal, 20 LM iterations,
    PCG <= 50 on the matrix-free routes), one thread, beside the GPU engine and the matrix-free
    oracle on the same problem."""
    from oracle import oracle as O
    from typecv_amd import bundle as B
    from typecv_amd import synthetic as S

    ring, _ = S.ring_scene(seed=1)
    control = B.make_control(ring.C, ring.P, ring.M, iterations=20, max_cg=50)
    O.set_threads(1)
    t = time.perf_counter()
    dense = O.dense_solve(control, ring)
    dense_s = time.perf_counter() - t
    _, st = O.solve(control, ring)
    O.set_threads(0)
    base = ring.C * 12 + ring.P * 3
    lm = dense[base + 10] + dense[base + 11]
    with B.HipDevice(0) as device:
        B.solve_packed(device, control, ring)
        t = time.perf_counter()
        got, rep = B.solve_packed(device, control, ring)
        gpu_s = time.perf_counter() - t
    return {
        "value": lm / dense_s, "unit": "LM iterations/s", "cores": 1, "kind": "port",
        "sample": f"the whole dense solve of the 10-camera ring ({ring.M} observations, "
                  f"{lm:.0f} LM iterations, {dense_s:.2f} s)",
        "matrix_free_oracle_lm_per_s": st["lm_iterations"] / (st["lm_ms"] * 1e-3),
        "gpu_lm_per_s_host_arrays": rep["lm_iterations"] / gpu_s,
        "final_cost": {"dense_cpu": float(dense[base + 6]), "gpu": float(got[base + 6])},
    }


def run_batch(B, device, controls, packeds, steps, warmup):
    """Times `steps` batched solves after `warmup`: (LM-loop seconds per step, wall seconds per
    step, per-bundle LM-loop latencies in ms, summed report counters)."""
    for _ in range(warmup):
        B.solve_batch(device, controls, packeds)
    loops, walls, latencies = [], [], []
    totals = {"lm_iterations": 0, "cg_iterations": 0, "direct_solves": 0}
    for k in range(steps):
        t = time.perf_counter()
        outs, report, finish = B.solve_batch(device, controls, packeds)
        walls.append(time.perf_counter() - t)
        loops.append(float(finish.max()) * 1e-3)
        latencies.extend(finish.tolist())
        for key in totals:
            totals[key] += report[key]
        log(f"step {k}: wall {walls[-1] * 1e3:.1f} ms, LM loop {loops[-1] * 1e3:.1f} ms, "
            f"{report['lm_iterations']} LM, {report['cg_iterations']} PCG, "
            f"{report['direct_solves']} direct")
    return loops, walls, latencies, totals


def batched(args) -> int:
    """256 independent local bundles (cameras U{5..40}, points U{500..20000}, each point seen by
    3-12 cameras of its bundle; the trajectory scene's noise, outliers and focal error), shared
    focal, Huber 1 px, 30 LM iterations, mixed precision, solved together by one
    tcv_ba_solve_batch call per step. The value is the direct route (directSolve: each bundle's
    reduced camera system factored, certified by the matrix-free residual); the PCG route is timed
    beside it unless --no-pcg. Reports bundles/s over the batches' LM loops and the p50/p99 of the
    per-bundle LM-loop latency."""
    import numpy as np

    from typecv_amd import bundle as B
    from typecv_amd import synthetic as S

    t0 = time.time()
    bundles = S.local_bundles(count=args.bundles, seed=args.seed)
    packeds = [p for p, _ in bundles]
    M = sum(p.M for p in packeds)
    log(f"{len(packeds)} bundles, {sum(p.C for p in packeds)} cameras, "
        f"{sum(p.P for p in packeds)} points, {M} observations in {time.time() - t0:.1f}s")

    def controls_for(direct):
        cs = [B.make_control(p.C, p.P, p.M, shared_focal=True, huber=1.0,
                             iterations=args.iterations, max_cg=min(256, max(32, 12 * p.C + 1)),
                             precision=args.precision, direct=direct) for p in packeds]
        # the bundles' default budgets differ (12C+1); a batch shares one budget: the largest
        budget = max(c.max_cg for c in cs)
        for c in cs:
            c.max_cg = budget
        return cs, budget

    device = B.HipDevice(0)
    controls, budget = controls_for(True)
    log("direct route")
    loops, walls, latencies, totals = run_batch(B, device, controls, packeds, args.steps,
                                                args.warmup)
    pcg = None
    if not args.no_pcg:
        log("PCG route")
        pcg_controls, _ = controls_for(False)
        p_loops, p_walls, p_lat, p_totals = run_batch(B, device, pcg_controls, packeds,
                                                      args.pcg_steps, args.warmup)
        p_lat = np.array(p_lat)
        pcg = {"bundles_per_s": len(packeds) * args.pcg_steps / sum(p_loops),
               "ms_per_step": 1e3 * sum(p_loops) / args.pcg_steps,
               "steps": args.pcg_steps,
               "latency_ms": {"p50": float(np.percentile(p_lat, 50)),
                              "p99": float(np.percentile(p_lat, 99))},
               "pcg_iterations_per_step": p_totals["cg_iterations"] / args.pcg_steps}
    device.dispose()
    lat = np.array(latencies)
    lm = totals["lm_iterations"]
    result = {
        "metric": "batched local bundles: bundles/s, p50/p99 per-bundle LM-loop latency",
        "value": len(packeds) * args.steps / sum(loops),
        "unit": "bundles/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * sum(loops) / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        # the direct route stores no FP32 cross blocks: it computes in FP64 in either form
        "dtype": "f64",
        "data": "synthetic (seeded local bundles)",
        "config": {"workload": "batched local bundles", "bundles": len(packeds),
                   "observations": M, "lm_iterations_max": args.iterations,
                   "precision": args.precision, "max_cg": budget,
                   "linear_solver": "direct (directSolve), certified by the matrix-free residual"},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "max": float(lat.max())},
        "end_to_end_bundles_per_s": len(packeds) * args.steps / sum(walls),
        "lm_iterations_per_s": lm / sum(loops),
        "pcg_iterations_per_step": totals["cg_iterations"] / max(1, args.steps),
        "direct_solves_per_step": totals["direct_solves"] / max(1, args.steps),
        "observations_per_s": M * (lm / len(packeds)) / (sum(loops) / args.steps) / args.steps
        if args.steps else 0.0,
        "pcg_route": pcg,
    }
    print(json.dumps(result), flush=True)
    return 0


def observations(args) -> int:
    """Fixed-scene observation evaluation (`evaluateObservationsGpu`, the mapper's observation
    job) over the trajectory scene's observations at its perturbed starting poses: projection and
    threshold of every observation, the optional depth guard per landmark, the inlier summary.
    A step is one `tcv_obs_evaluate` call from host arrays (upload and readback of the flags
    included; the per-landmark masks of the Python mirror are not built); the CPU baseline is the
    oracle's restatement of the reference's CPU tier on the same arrays, one thread."""
    import numpy as np

    from oracle import oracle as O
    from typecv_amd import bundle as B
    from typecv_amd import observations as OB
    from typecv_amd import synthetic as S

    packed, _ = S.trajectory_scene(cameras=args.cameras, points=args.points, seed=2)
    C, P, M = packed.C, packed.P, packed.M
    sc, ix = packed.scene, packed.index
    poses, pts = sc[:12 * C].reshape(C, 12), sc[12 * C:12 * C + 3 * P]
    models = sc[12 * C + 3 * P:12 * C + 3 * P + 6 * C].reshape(C, 6)
    pixels = sc[12 * C + 3 * P + 6 * C:]
    guarded = args.max_relative_depth_std > 0
    request = {
        "depth": {"maxRelativeDepthStdDev": args.max_relative_depth_std, "pixelNoiseStdDevPx": 1.0},
        "guarded": guarded, "lengths": [], "count": M, "views": C, "points": P,
        "cameras": np.ascontiguousarray(np.concatenate([poses, models], axis=1).reshape(-1)),
        "points_xyz": np.ascontiguousarray(pts), "pixels": np.ascontiguousarray(pixels),
        "indices": np.ascontiguousarray(np.stack([ix[:M], ix[M:2 * M]], axis=1).reshape(-1)),
        "offsets": np.ascontiguousarray(ix[2 * M:2 * M + P + 1]) if guarded else None,
    }
    if guarded:
        request["lengths"] = np.diff(request["offsets"]).tolist()
    device = B.HipDevice(0)
    for _ in range(args.warmup):
        OB.evaluate_arrays(device, request, args.max_error)
    times = []
    for k in range(args.steps):
        g_flags, g_summary, elapsed = OB.evaluate_arrays(device, request, args.max_error)
        times.append(elapsed * 1e-3)
        log(f"step {k}: {elapsed:.2f} ms, {g_summary[1]:.0f} inliers")
    device.dispose()
    result = {"inlierCount": float(g_summary[1]),
              "meanErrorPx": float(g_summary[0] / g_summary[1]) if g_summary[1] else None}
    t = time.perf_counter()
    flags, _, summary = O.observations(request["cameras"], request["points_xyz"],
                                       request["pixels"], request["indices"], request["offsets"],
                                       args.max_error, args.max_relative_depth_std, 1.0)
    cpu = time.perf_counter() - t
    ms = float(np.median(times)) * 1e3
    print(json.dumps({
        "metric": "fixed-scene observation evaluation: observations/s", "value": M / (ms * 1e-3),
        "unit": "observations/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded trajectory scene at its starting poses)",
        "config": {"workload": "evaluateObservationsGpu", "cameras": C, "points": P,
                   "observations": M, "max_error_px": args.max_error,
                   "max_relative_depth_std_dev": args.max_relative_depth_std},
        "note": "host arrays in and flags out: the timed step includes the PCIe transfers",
        "inliers": result["inlierCount"], "mean_error_px": result["meanErrorPx"],
        "matches_oracle": bool(np.array_equal(flags, g_flags[:len(flags)])
                               and np.array_equal(summary, g_summary)),
        "cpu_baseline": {"value": M / cpu, "unit": "observations/s", "cores": 1, "kind": "port",
                         "sample": f"the whole evaluation ({M} observations), C oracle -O3"},
    }), flush=True)
    return 0


def real_scene(args) -> int:
    """The reference's real-data BA measurement (south-building-24: 24 views, 3,000 landmarks,
    12,706 observations; 4 LM iterations, Huber 1 px; tests/golden/south_building_24.npz) through
    the operator on the packed problem, fixed and shared focal, PCG and direct routes: wall time of
    the device-resident LM loop (upload excluded) and the final RMSE against the reference's record
    (docs/experiments/gpu-budget/south-building-24-ba-*.json, 6,741 / 8,938 ms on its WebGPU tier)."""
    import numpy as np

    sys.path.insert(0, str(Path(__file__).resolve().parent / "tests"))
    import real_scene as R
    from typecv_amd import bundle as B

    model = R.reconstruction()
    rows = []
    with B.HipDevice(0) as device:
        for focal in ("fixed", "shared"):
            for cache in (False, True):
                options = {**R.OPTIONS, "refineSharedFocal": focal == "shared",
                           "directSolve": cache}
                problem = B.bundle_problem(model, options, (lambda C: min(256, max(32, 12 * C + 1)))
                                           if focal == "shared" else
                                           (lambda C: min(96, max(24, 12 * C))))
                packed = B.pack_bundle_scene(problem["input"], problem["byId"], problem["active"])
                control = B.make_control(
                    packed.C, packed.P, packed.M, directional=packed.P - problem["finitePoints"],
                    anchor=problem["byId"][problem["anchorId"]],
                    scale_view=problem["byId"][problem["scaleId"]],
                    normalize_scale=problem["fixed"] == 1, shared_focal=focal == "shared",
                    iterations=problem["iterations"], max_cg=problem["maxCg"],
                    huber=problem["huber"], tolerance=problem["tolerance"], direct=cache)
                times, out, rep = [], None, None
                for _ in range(args.warmup + args.steps):
                    out, rep = B.solve_packed(device, control, packed)
                    times.append(rep["solve_ms"] - rep["setup_ms"])
                base = packed.C * 12 + packed.P * 3
                rows.append({"focal": focal, "route": "direct" if cache else "PCG",
                             "lm_loop_ms": float(np.median(times[args.warmup:])),
                             "final_rmse_px": float(out[base + 25]),
                             "record_final_rmse_px": R.RECORD[focal]["finalRmsePx"],
                             "pcg_iterations": int(rep["cg_iterations"]),
                             "direct_solves": int(rep["direct_solves"])})
                log(str(rows[-1]))
    print(json.dumps({"mode": "south-building-24", "cameras": 24, "points": 3000,
                      "observations": 12706, "lm_iterations": 4, "runs": rows}), flush=True)
    return 0


def probe(args) -> int:
    """Counter run: a 1 GiB device copy (the calibration kernel of known size), then `--repeat`
    Schur products of the workload through tcv_ba_schur_probe, for `rocprofv3 --pmc FETCH_SIZE` /
    `WRITE_SIZE` runs. Prints the HIP-event time per sweep."""
    import numpy as np
    import torch

    from typecv_amd import bundle as B
    from typecv_amd import synthetic as S

    packed, _ = S.trajectory_scene(cameras=args.cameras, points=args.points, seed=args.seed)
    src = torch.empty(1 << 27, dtype=torch.float64, device="cuda")  # 1 GiB
    src.fill_(1.0)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    control = B.make_control(packed.C, packed.P, packed.M, shared_focal=True, huber=1.0,
                             iterations=1, max_cg=256, precision=args.precision, timing=True)
    m = 6 * packed.C + 1
    vec = np.random.default_rng(1).normal(size=m)
    with B.HipDevice(0) as device:
        _, _, _, status, rep = B.schur_probe(device, control, packed, 1e-3, 0.0, vec,
                                             repeat=args.repeat)
    n = max(1, rep["matvecs"])
    print(json.dumps({"mode": "probe", "observations": packed.M, "status": status,
                      "point_sweep_us": rep["schur_point_ms"] / n * 1e3,
                      "camera_sweep_us": rep["schur_camera_ms"] / n * 1e3,
                      "copy_bytes": 2 * src.numel() * 8}), flush=True)
    return 0


def grad(args) -> int:
    """Differentiation on a shared-focal trajectory scene (FP64; the generator of the benchmark's
    main workload, with its noise, outliers and focal error): the forward solve, the Jacobian
    kernels against central differences (h = 1e-6) of the oracle's projection on a sample of
    observations (`--jacobian-only` stops there: the 500-camera scene's check), a polish to a
    stationary point (damped exact-Hessian LM, then Newton), the adjoint backward for a seeded
    N(0, 1) cotangent on every output (poses, points, log focal), and central differences
    (h = 1e-3 px) of re-polished solutions on 64 random single-observation pixel perturbations.
    The default scene is 200 cameras / 80,000 points (the workload's 400 points per camera): the
    largest size tried where the damped polish reaches a stationary point; at 500 cameras it does
    not within 200 steps, so the exact derivative is undefined there. Points the polish holds fixed
    (tracks run off to infinity, their Gauss-Newton blocks singular to rounding) get a zero
    cotangent."""
    import numpy as np

    from oracle import oracle as O
    from typecv_amd import bundle as B
    from typecv_amd import synthetic as S

    packed, info = S.trajectory_scene(cameras=args.cameras, points=args.points, seed=args.seed,
                                      outliers=args.outliers)
    C, P, M = packed.C, packed.P, packed.M
    log(f"scene C={C} P={P} M={M}")
    control = B.make_control(C, P, M, shared_focal=True, huber=args.huber,
                             iterations=args.iterations, max_cg=min(256, max(32, 12 * C + 1)))
    device = B.HipDevice(0)
    t = time.perf_counter()
    out, rep = B.solve_packed(device, control, packed)
    forward_s = time.perf_counter() - t
    at = B.at_solution(packed, out)
    lf = B.output_log_focal(control, out)
    log(f"forward {forward_s:.2f}s, {rep['lm_iterations']} LM, log focal {lf:.6f}")
    # Jacobian kernels against central differences of the oracle's projection
    J = B.jacobian(device, control, at, lf)
    rng = np.random.default_rng(args.seed)
    sample = rng.choice(M, size=args.jacobian_sample, replace=False)
    cams = packed.index[:M].astype(np.int64)
    pts = packed.index[M: 2 * M].astype(np.int64)
    models = at.scene[12 * C + 3 * P: 18 * C + 3 * P].reshape(C, 6)
    active = packed.index[3 * M + P + C + 2:].astype(bool)

    def pixel(pose, X, model, lfs):
        s = O.focal_scale(lfs)
        k = model.copy()
        k[0] *= s
        k[1] *= s
        R = pose[:9].reshape(3, 3)
        ok, px, _, _ = O.project(k, R @ X + pose[9:])
        return np.asarray(px)

    h, worst = 1e-6, 0.0
    for i in sample:
        c, p = cams[i], pts[i]
        pose = at.scene[12 * c: 12 * c + 12]
        X = at.scene[12 * C + 3 * p: 12 * C + 3 * p + 3]
        cols = []
        for k in range(6):
            if not active[c]:
                cols.append(np.zeros(2))
                continue
            d = np.zeros(6)
            d[k] = h
            cols.append((pixel(O.move_pose(pose, d), X, models[c], lf)
                         - pixel(O.move_pose(pose, -d), X, models[c], lf)) / (2 * h))
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            cols.append((pixel(pose, X + e, models[c], lf) - pixel(pose, X - e, models[c], lf))
                        / (2 * h))
        cols.append((pixel(pose, X, models[c], lf + h) - pixel(pose, X, models[c], lf - h)) / (2 * h))
        fd = np.array(cols).T  # 2 x 10
        an = np.concatenate([J[i, 2:14].reshape(2, 6), J[i, 14:20].reshape(2, 3),
                             J[i, 20:22].reshape(2, 1)], axis=1)
        worst = max(worst, float(np.max(np.abs(fd - an)) / max(1.0, np.max(np.abs(an)))))
    log(f"jacobian vs central differences on {len(sample)} observations: max rel err {worst:.2e}")
    if args.jacobian_only:
        print(json.dumps({
            "metric": "differentiation: Jacobian kernels vs central differences (h = 1e-6)",
            "value": worst, "unit": "max relative error", "higher_is_better": False, "n_gpus": 1,
            "dtype": "f64", "data": "synthetic (seeded trajectory scene)",
            "config": {"workload": "shared-focal trajectory scene, Jacobian kernels", "cameras": C,
                       "points": P, "observations": M, "sample": len(sample)},
            "forward_s": forward_s}), flush=True)
        device.dispose()
        return 0
    # Newton polish to a stationary point
    t = time.perf_counter()
    pol_out, pol_rep = B.polish(device, control, at, lf, iterations=4, tolerance=1e-12,
                                lm_iterations=args.polish_lm, max_cg=args.max_cg)
    polish_s = time.perf_counter() - t
    pol = B.at_solution(at, pol_out)
    lf = B.output_log_focal(control, pol_out)
    g_before = np.linalg.norm(O.gradient(control, at.scene, packed.index, B.output_log_focal(
        control, out))[0])
    g_after = np.linalg.norm(O.gradient(control, pol.scene, packed.index, lf)[0])
    log(f"polish {polish_s:.2f}s ({pol_rep['lm_iterations']} damped, "
        f"{pol_rep['refinement_steps']:.0f} Newton steps, {pol_rep['cg_iterations']} CG): "
        f"|g| {g_before:.3e} -> {g_after:.3e}")
    result = {
        "metric": "differentiation: Jacobian vs central differences, adjoint vs FD of re-solves",
        "n_gpus": 1, "dtype": "f64", "data": "synthetic (seeded trajectory scene)",
        "config": {"workload": "shared-focal trajectory scene, implicit differentiation",
                   "cameras": C, "points": P, "observations": M, "outliers": info["outliers"]},
        "forward_s": forward_s, "polish_s": polish_s,
        "polish": {"damped_steps": pol_rep["lm_iterations"], "newton_steps": pol_rep["refinement_steps"],
                   "cg_iterations": pol_rep["cg_iterations"]},
        "gradient_norm": {"forward": g_before, "polished": g_after},
        "jacobian": {"sample": len(sample), "h": h, "max_rel_err": worst},
    }
    if pol_rep["refinement_steps"] < 1:
        result["backward"] = {"error": "the polish reached no stationary point (no Newton step "
                                       "could be taken): the implicit derivative is undefined"}
        print(json.dumps(result), flush=True)
        device.dispose()
        return 1
    # the backward
    n_out = 12 * C + 3 * P + 1
    cot = np.random.default_rng(args.seed + 1).normal(size=n_out)
    # no cotangent on the points the exact-Hessian entries hold fixed (their Gauss-Newton blocks
    # singular to rounding)
    Jpol = B.jacobian(device, control, pol, lf)
    Jp, w = Jpol[:, 14:20].reshape(-1, 2, 3), Jpol[:, 22]
    ok = w >= 0
    V = np.zeros((P, 3, 3))
    np.add.at(V, pts[ok], w[ok, None, None] * np.einsum("mai,maj->mij", Jp[ok], Jp[ok]))
    ev = np.linalg.eigvalsh(V)
    for p in np.nonzero(ev[:, 0] < 1e-11 * ev[:, 2])[0]:
        cot[12 * C + 3 * p: 12 * C + 3 * p + 3] = 0
    B.backward(device, control, pol, lf, cot, tolerance=args.tolerance, max_cg=args.max_cg)
    t = time.perf_counter()
    pixels, focal, adjoint, back = B.backward(device, control, pol, lf, cot,
                                              tolerance=args.tolerance, max_cg=args.max_cg)
    backward_s = time.perf_counter() - t
    log(f"backward {backward_s * 1e3:.1f} ms, {back['cg_iterations']} CG")
    # central differences of re-polished solutions; a perturbation that can carry its observation's
    # residual across the Huber radius (|r| within 2h of it at the solution) meets the kink of the
    # robust cost, where the derivative is one-sided: reported apart
    radius = np.hypot(*B.jacobian(device, control, pol, lf)[:, 0:2].T)
    errs, kinks = [], []
    pix = 18 * C + 3 * P
    hp = 1e-3
    for k in range(args.perturbations):
        i = int(rng.integers(0, M))
        d = rng.normal(size=2)
        d /= np.linalg.norm(d)
        outs = []
        for sign in (1, -1):
            scene = pol.scene.copy()
            scene[pix + 2 * i: pix + 2 * i + 2] += sign * hp * d
            o, _ = B.polish(device, control, B.dataclasses.replace(pol, scene=scene), lf,
                            iterations=2, tolerance=1e-12, max_cg=args.max_cg)
            outs.append(o)
        fd = cot[:n_out - 1] @ ((outs[0][:n_out - 1] - outs[1][:n_out - 1]) / (2 * hp))
        fd += cot[-1] * (B.output_log_focal(control, outs[0])
                         - B.output_log_focal(control, outs[1])) / (2 * hp)
        an = float(pixels[i] @ d)
        err = abs(fd - an) / max(abs(fd), 1e-12)
        kink = abs(radius[i] - control.huber) < 2 * hp
        (kinks if kink else errs).append(err)
        log(f"perturbation {k}: observation {i} (|r| {radius[i]:.4f}{', at the Huber radius' if kink else ''}),"
            f" fd {fd:.6e}, adjoint {an:.6e}, rel {err:.1e}")
    device.dispose()
    result.update({
        "value": backward_s * 1e3, "unit": "ms per backward", "higher_is_better": False,
        "backward": {"ms": backward_s * 1e3, "cg_iterations": back["cg_iterations"],
                     "tolerance": args.tolerance},
        "finite_differences": {"perturbations": len(errs) + len(kinks), "h_px": hp,
                               "max_rel_err": float(max(errs)) if errs else None,
                               "median_rel_err": float(np.median(errs)) if errs else None,
                               "at_huber_radius": len(kinks),
                               "at_huber_radius_rel_err": [float(e) for e in kinks]},
    })
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    if "--mode" in sys.argv and "grad" in sys.argv:
        ap = argparse.ArgumentParser()
        ap.add_argument("--mode")
        ap.add_argument("--cameras", type=int, default=200)
        ap.add_argument("--points", type=int, default=80_000)
        ap.add_argument("--jacobian-only", action="store_true")
        ap.add_argument("--seed", type=int, default=2)
        ap.add_argument("--iterations", type=int, default=30)
        ap.add_argument("--jacobian-sample", type=int, default=2000)
        ap.add_argument("--perturbations", type=int, default=64)
        ap.add_argument("--tolerance", type=float, default=1e-10)
        ap.add_argument("--polish-lm", type=int, default=200)
        ap.add_argument("--max-cg", type=int, default=20000)
        ap.add_argument("--outliers", type=float, default=0.02)
        ap.add_argument("--huber", type=float, default=1.0)
        sys.exit(grad(ap.parse_args()))
    if "--mode" in sys.argv and "probe" in sys.argv:
        ap = argparse.ArgumentParser()
        ap.add_argument("--mode")
        ap.add_argument("--cameras", type=int, default=500)
        ap.add_argument("--points", type=int, default=200_000)
        ap.add_argument("--seed", type=int, default=2)
        ap.add_argument("--repeat", type=int, default=10)
        ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64")
        sys.exit(probe(ap.parse_args()))
    if "--workload" in sys.argv and "observations" in sys.argv:
        ap = argparse.ArgumentParser()
        ap.add_argument("--workload")
        ap.add_argument("--steps", type=int, default=5)
        ap.add_argument("--warmup", type=int, default=1)
        ap.add_argument("--cameras", type=int, default=500)
        ap.add_argument("--points", type=int, default=200_000)
        ap.add_argument("--max-error", type=float, default=3.0)
        ap.add_argument("--max-relative-depth-std", type=float, default=0.0)
        sys.exit(observations(ap.parse_args()))
    if "--workload" in sys.argv and "south-building-24" in sys.argv:
        ap = argparse.ArgumentParser()
        ap.add_argument("--workload")
        ap.add_argument("--steps", type=int, default=5)
        ap.add_argument("--warmup", type=int, default=1)
        sys.exit(real_scene(ap.parse_args()))
    if "--workload" in sys.argv and "batched" in sys.argv:
        ap = argparse.ArgumentParser()
        ap.add_argument("--workload")
        ap.add_argument("--gpus", type=int, default=1)
        ap.add_argument("--steps", type=int, default=2)
        ap.add_argument("--warmup", type=int, default=1)
        ap.add_argument("--bundles", type=int, default=256)
        ap.add_argument("--iterations", type=int, default=30)
        ap.add_argument("--precision", choices=("fp64", "mixed"), default="mixed")
        ap.add_argument("--seed", type=int, default=4)
        ap.add_argument("--no-pcg", action="store_true", help="skip timing the PCG route")
        ap.add_argument("--pcg-steps", type=int, default=1)
        sys.exit(batched(ap.parse_args()))
    sys.exit(main())
