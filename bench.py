#!/usr/bin/env python3
"""Benchmark of the joint flow+depth solve on one MI355X.

Workload: one 640x480 seeded synthetic pair (intrinsics f=500, c=(320,240); camera-1 pose
|rotation| <= 3 deg, |t| = 0.1; 8 planes at depths 2-10), ReconFlow METHODR_TVL1_MS, 5 levels
(scale divisor 2), 5 warps, 50 iterations, lambda 50, tau 0.125, alphaTv 33.3, alphaProj 60,
lambdaf 0.1, lambdams 100. A step is one full solve (gray conversion, pyramid, all levels,
warps and iterations) from images already resident on the GPU.

Prints one JSON line: frames/s, the roofline of the dominant kernel (its algorithmic bytes
per launch over its average launch time, timed with HIP events on the solver's stream), the
FP64 CPU oracle's time on the same frame (the reported CPU baseline) and the mean EPE of the
GPU flow against that FP64 result.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

W, H, LEVELS, WARPS, ITERS = 640, 480, 5, 5, 50
PEAK_HBM_GBS = 8000.0

# Algorithmic bytes per pixel of one launch (each distinct element read once and written
# once, FP32): see the kernel table of the design notes.
BYTES_PER_PX = {
    "fused_iter": 140.0,  # u,v,I(3) read; uhat(2), p(4), uproj(2), X(3), uc(2), sk(2) read + written
    "data": 68.0, "dual": 40.0, "uvproj": 40.0, "solve3d": 32.0, "median3d": 24.0, "cleanup": 36.0,
    "solve_uc": 20.0, "update_sk": 40.0,
}
KID = {"fused_iter": 0, "data": 1, "dual": 2, "uvproj": 3, "solve3d": 4, "median3d": 5, "cleanup": 6,
       "solve_uc": 7, "update_sk": 8}


def make_solver(rf, path, graph, tri, joint_tile=0):
    r = rf.ReconFlow()
    r._initializeR(W, H, 1, LEVELS, 2.0, rf.ReconFlow.METHODR_TVL1_MS, 50.0, 0.0, 0.1, 100.0, 33.3, 60.0, 0.125,
                   WARPS, ITERS, cvType=rf.INPUT_32F, path=path, use_graph=graph)
    r.params.tri_f64 = int(tri)
    r.params.joint_tile = int(joint_tile)
    r._init_params(r.params)
    return r


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--path", choices=["fused", "plain"], default="fused")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--tri-f64", type=int, default=1, help="FP64 triangulation (1) or the reference FP32 (0)")
    ap.add_argument("--joint-tile", type=int, default=0, choices=[0, 1, 2, 3, 4],
                    help="fused joint tiling: 0 autotuned per level, 1 64x8 two rows per lane, 2 32x16, 3 16x16, "
                         "4 16x16 capped at 96 VGPRs")
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain-compare", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 256 MiB device-to-device copy 5 times (scales the FETCH_SIZE / WRITE_SIZE "
                         "counters to bytes) and exit")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this solver runs one image pair per device and the "
              "multi-GPU sharding (independent pairs per GPU) is not exercised here.", file=sys.stderr)
        return 2

    import ctypes as C

    import torch

    import reconflow_amd as rf

    assert torch.cuda.is_available(), "bench needs a GPU"
    if args.copy_calibration:
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")  # 256 MiB
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": src.numel() * 4, "launches": 5}))
        return 0
    d = rf.make_recon_pair(W, H, args.seed)
    path = rf.PATH_FUSED if args.path == "fused" else rf.PATH_PLAIN
    r = make_solver(rf, path, not args.no_graph, args.tri_f64, args.joint_tile)
    r.setCameraMatrices(d["K"], d["K"])
    r.copyImagesToDevice(d["i0"], d["i1"])  # resident for every step
    for _ in range(args.warmup):
        r.solveR(d["R"], d["t"], 50.0)
    r.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r.solveR(d["R"], d["t"], 50.0)
    r.synchronize()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms = dt / args.steps * 1e3
    fps = args.steps / dt
    u, v = r.copyOpticalFlowToHost()
    X, Y, Z = r.copy3dToHost()

    # dominant kernel: time it live with HIP events on the solver's stream (eager solves)
    lib = r._lib
    cand = ["fused_iter"] if path == rf.PATH_FUSED else ["data", "dual", "uvproj", "solve3d", "median3d",
                                                        "cleanup", "solve_uc", "update_sk"]
    per = {}
    for name in cand:
        assert lib.rf_profile_arm(r._ctx, KID[name], 0) == 0
        r.solveR(d["R"], d["t"], 50.0)
        tot, n = C.c_double(), C.c_int()
        assert lib.rf_profile_read(r._ctx, C.byref(tot), C.byref(n)) == 0
        per[name] = (tot.value, n.value)
    dom = max(per, key=lambda k: per[k][0])
    tot_ms, nl = per[dom]
    per_level_us = []
    for lv in range(LEVELS):  # the dominant kernel's average launch per pyramid level
        assert lib.rf_profile_arm(r._ctx, KID[dom], lv) == 0
        r.solveR(d["R"], d["t"], 50.0)
        t_, n_ = C.c_double(), C.c_int()
        assert lib.rf_profile_read(r._ctx, C.byref(t_), C.byref(n_)) == 0
        per_level_us.append(round(t_.value * 1e3 / max(n_.value, 1), 3))
    avg_s = tot_ms / 1e3 / max(nl, 1)
    bytes_launch = BYTES_PER_PX[dom] * W * H
    achieved = bytes_launch / avg_s / 1e9
    traffic = None
    tfile = ROOT / "profiles" / "traffic.json"
    if tfile.exists():
        tj = json.loads(tfile.read_text())
        traffic = tj.get(f"{dom}_{args.path}_{W}x{H}")

    plain_fps = None
    if path == rf.PATH_FUSED and not args.no_plain_compare:
        rp = make_solver(rf, rf.PATH_PLAIN, True, args.tri_f64)
        rp.setCameraMatrices(d["K"], d["K"])
        rp.copyImagesToDevice(d["i0"], d["i1"])
        rp.solveR(d["R"], d["t"], 50.0)
        rp.synchronize()
        n_plain = 3
        t1 = time.perf_counter()
        for _ in range(n_plain):
            rp.solveR(d["R"], d["t"], 50.0)
        rp.synchronize()
        plain_fps = n_plain / (time.perf_counter() - t1)
        del rp

    cpu = None
    epe = None
    if not args.no_cpu_baseline:
        from oracle import oracle as orc
        p = orc.recon_params(W, H, LEVELS)
        tc = time.perf_counter()
        ref = orc.solve_recon(p, d["K"], d["K"], np.eye(3), np.zeros(3), d["R"], d["t"], d["i0"], d["i1"], "f64")
        tcpu = time.perf_counter() - tc
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu = {"value": round(1.0 / tcpu, 4), "unit": "frames/s", "cores": threads, "kind": "port",
               "sample": "one full 640x480 joint frame (5 levels x 5 warps x 50 iterations) of the FP64 oracle, "
                         f"OpenMP over rows, {tcpu:.2f} s"}
        epe = float(np.hypot(u[:, :W] - ref[0][:, :W], v[:, :W] - ref[1][:, :W]).mean())

    line = {
        "metric": "640x480 flow+depth frames/s at EPE<=0.05px vs FP64 oracle; % of HBM peak",
        "value": round(fps, 3),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" + ("+f64-triangulation" if args.tri_f64 else ""),
        "data": "synthetic (seeded piecewise-planar scene pair, rendered on the host)",
        "config": {"workload": "joint flow+depth 640x480, 5 levels, 5 warps, 50 iterations, one pair per step",
                   "path": args.path, "graph": not args.no_graph, "seed": args.seed, "joint_tile": args.joint_tile},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
                     "avg_launch_us": round(avg_s * 1e6, 3), "algorithmic_bytes_per_launch": bytes_launch,
                     "avg_launch_us_per_level": per_level_us},
        "cpu_baseline": cpu,
        "epe_vs_fp64_oracle": None if epe is None else round(epe, 5),
        "accuracy_ok": None if epe is None else bool(epe <= 0.05),
        "plain_fp32_path_fps": None if plain_fps is None else round(plain_fps, 3),
        "mpix_iter_per_s": round(fps * 409200 * WARPS * ITERS / 1e6, 1),
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
