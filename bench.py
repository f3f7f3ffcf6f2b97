#!/usr/bin/env python3
"""Benchmark of the joint flow+depth solve on one MI355X.

Default workload: one 640x480 seeded synthetic pair (intrinsics f=500, c=(320,240); camera-1
pose |rotation| <= 3 deg, |t| = 0.1; 8 planes at depths 2-10), ReconFlow METHODR_TVL1_MS,
5 levels (scale divisor 2), 5 warps, 50 iterations, lambda 50, tau 0.125, alphaTv 33.3,
alphaProj 60, lambdaf 0.1, lambdams 100. A step is one full solve (gray conversion, pyramid,
all levels, warps and iterations) from images already resident on the GPU, replayed from a
captured graph.

Prints one JSON line: frames/s, the roofline of the dominant kernel (its algorithmic bytes per
launch over its average launch time at level 0, timed with HIP events on the solver's stream),
the FP64 CPU oracle's time on the same frame (the reported CPU baseline), the mean EPE of the
GPU flow against that FP64 result, and the per-step (plain) GPU path's frames/s beside it.
`batched_640x480` adds the same solve over --batch-pairs seeded 640x480 pairs per batched launch
(FP32 storage, FP16 derivatives, FP16 derivatives and duals), frames/s and pair 0's EPE against the
FP64 result.

--workload tvl1-320: the 320x240 TV-L1 case (fused, fused FP16, plain; EPE and time of the FP64
oracle). --workload methods: each CudaFlow data term at 320x240 on the per-step path.
--workload stream720: consecutive pairs of a 300-frame 1280x720 synthetic camera trajectory: one
pair per launch cold and warm-started from the previous pair's flow, eight pairs per batched launch
cold, and the sequence in eight chunks whose pairs share each batched launch, each warm-started from
the previous pair of its chunk (the headline of this workload): frames/s and p50/p99 latency.
--workload batch64: 64 seeded pairs of uneven sizes (W in 160..1280, H in 120..720, multiples of
16), even seeds TV-L1 (4 levels, 3 warps, 30 iterations; FP16 images, derivatives and duals),
odd seeds joint (as above; FP16 derivatives, FP32 images); the two batches run on their own
streams.
"""
from __future__ import annotations

import argparse
import ctypes
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
# VALU issue peak: 256 CUs x 4 SIMDs x 2.4 GHz, one wave64 VALU instruction per SIMD every 4 cycles
PEAK_VALU_GINST = 256 * 4 * 2.4 / 4

# Algorithmic bytes per pixel of one launch (each distinct element read once and written
# once, FP32).
BYTES_PER_PX = {
    "fused_iter": 140.0,  # u,v,I(3) read; uhat(2), p(4), uproj(2), X(3), uc(2), sk(2) read + written
    "data": 68.0, "dual": 40.0, "uvproj": 40.0, "solve3d": 32.0, "median3d": 24.0, "cleanup": 36.0,
    "solve_uc": 20.0, "update_sk": 40.0,
}
FLOW_ITER_BYTES = 60.0  # one fused TV-L1 iteration: I(3) read; uhat(2), p(4) read + written
KID = {"fused_iter": 0, "data": 1, "dual": 2, "uvproj": 3, "solve3d": 4, "median3d": 5, "cleanup": 6,
       "solve_uc": 7, "update_sk": 8}
TILING = {0: "auto", 1: "64x8-rpt2", 2: "32x16", 3: "16x16", 4: "16x16-occ5", 5: "16x16-k2", 6: "16x16-k3"}
ITERS_PER_LAUNCH = {"16x16-k2": 2, "16x16-k3": 3, "32x16-k4": 4, "32x32-k8": 8, "32x16-k2": 2, "32x16-k6": 6}
FLOW_TILING = {0: "auto", 1: "64x16-rpt4", 2: "16x16", 3: "32x16", 4: "32x16-k4", 5: "32x32-k8", 6: "32x16-k2",
               7: "32x16-k6"}


def make_solver(rf, path, graph, tri, joint_tile=0, fix=0):
    """ReconFlow with the demo's parameters; every solve ends with the flow's colour map, as the
    reference's solveReconFlow does (Reconflow.cu:460)."""
    r = rf.ReconFlow()
    r._initializeR(W, H, 1, LEVELS, 2.0, rf.ReconFlow.METHODR_TVL1_MS, 50.0, 0.0, 0.1, 100.0, 33.3, 60.0, 0.125,
                   WARPS, ITERS, cvType=rf.INPUT_32F, path=path, use_graph=graph, withVisualization=True)
    r.params.tri_f64 = int(tri)
    r.params.joint_tile = int(joint_tile)
    r.params.fix_coarse_k = int(fix)
    r._init_params(r.params)
    return r


REF64 = {}  # the FP64 oracle's flow of the single-pair workload, kept for the batched EPE


def depth_rel_err(Z, Z64):
    zr = np.abs(Z[:, :W] - Z64[:, :W]) / np.abs(Z64[:, :W])
    zr = zr[np.isfinite(zr)]
    return {"median": float(f"{np.median(zr):.3g}"), "p99": float(f"{np.percentile(zr, 99):.3g}")}


def cpu_baseline(d, u, v):
    """The FP64 restatement on the host: median of 5 timed frames after one warm-up at the
    default OpenMP thread count, one frame on one thread, one FP32 frame; and the GPU flow's
    mean EPE against the FP64 result."""
    from oracle import oracle as orc

    p = orc.recon_params(W, H, LEVELS)
    args = (p, d["K"], d["K"], np.eye(3), np.zeros(3), d["R"], d["t"], d["i0"], d["i1"])
    threads = orc.get_threads()

    def timed(kind):
        t0 = time.perf_counter()
        res = orc.solve_recon(*args, kind)
        return time.perf_counter() - t0, res

    _, ref = timed("f64")
    REF64["joint"] = ref
    runs = [timed("f64")[0] for _ in range(5)]
    t_med = float(np.median(runs))
    t32, _ = timed("f32")
    orc.set_threads(1)
    t1, _ = timed("f64")
    orc.set_threads(threads)
    epe = float(np.hypot(u[:, :W] - ref[0][:, :W], v[:, :W] - ref[1][:, :W]).mean())
    cpu = {"value": round(1.0 / t_med, 4), "unit": "frames/s", "cores": threads, "kind": "port",
           "sample": f"one 640x480 joint frame (5 levels x 5 warps x 50 iterations) of the FP64 oracle, OpenMP "
                     f"over rows on {threads} threads: median of 5 frames after a warm-up, {t_med:.2f} s/frame; "
                     f"1 thread {t1:.2f} s/frame; FP32 restatement {t32:.2f} s/frame",
           "one_thread_frames_per_s": round(1.0 / t1, 4), "fp32_frames_per_s": round(1.0 / t32, 4)}
    return cpu, epe


def _fetch(dptr, n):
    """n floats at a raw device address -> host numpy (after a device synchronisation)."""
    hip = ctypes.CDLL("libamdhip64.so.7")  # by soname: the runtime torch and librfhip already share
    hip.hipDeviceSynchronize.restype = ctypes.c_int
    hip.hipMemcpy.restype = ctypes.c_int
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    assert hip.hipDeviceSynchronize() == 0
    out = np.empty(n, np.float32)
    assert hip.hipMemcpy(out.ctypes.data, dptr, n * 4, 2) == 0  # device to host
    return out


def _copy_async(dst, src, nbytes, stream):
    """hipMemcpyAsync device to device on `stream` (raw addresses)."""
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMemcpyAsync.restype = ctypes.c_int
    hip.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p]
    assert hip.hipMemcpyAsync(dst, src, nbytes, 3, stream) == 0


def run_joint_batched(args, rf, torch, n, storage):
    """n seeded 640x480 pairs (pair 0 is the single-pair workload's scene) in one batch: every
    launch covers the tiles of all n pairs. Returns frames/s and pair 0's flow."""
    p = rf.batch_params(14, LEVELS, WARPS, ITERS, lam=50.0, tau=0.125, alpha_tv=33.3, alpha_proj=60.0,
                        lambdaf=0.1, lambdams=100.0, tri_f64=int(args.tri_f64), storage=storage,
                        joint_tile=args.joint_tile)
    b = rf.Batch(p, [W] * n, [H] * n)
    keep = []
    for i in range(n):
        d = rf.make_recon_pair(W, H, args.seed + i, stride=b.stride(i))
        t0, t1 = torch.from_numpy(d["i0"]).cuda(), torch.from_numpy(d["i1"]).cuda()
        keep += [t0, t1]
        b.set_images_device(i, t0.data_ptr(), t1.data_ptr())
        b.set_camera(i, d["K"], d["K"], d["R"], d["t"])
    torch.cuda.synchronize()
    for _ in range(max(1, args.warmup)):  # the first solve is the tuning solve
        b.solve()
    b.synchronize()
    steps = max(2, args.steps // 4)
    t0 = time.perf_counter()
    for _ in range(steps):
        b.solve()
    b.synchronize()
    dt = time.perf_counter() - t0
    s = b.stride(0)
    u_ptr, v_ptr = b.result_ptrs(0)[:2]
    u, v = _fetch(u_ptr, H * s).reshape(H, s), _fetch(v_ptr, H * s).reshape(H, s)
    tiling = [TILING.get(c, str(c)) for c in b.tiling(LEVELS)]
    del b, keep
    return n * steps / dt, u, v, tiling


def run_drop_in_cycle(args, rf, d):
    """The reference demo's call cycle from host memory (test_reconflow/main.cpp:100-106):
    copyImagesToDevice from pageable host images, solveR, copyOpticalFlowToHost(u, v, uvrgb) and
    copy3dToHost into host buffers, per frame."""
    r = make_solver(rf, rf.PATH_FUSED if args.path == "fused" else rf.PATH_PLAIN, not args.no_graph, args.tri_f64)
    r.setCameraMatrices(d["K"], d["K"])

    def cycle():
        r.copyImagesToDevice(d["i0"], d["i1"])
        r.solveR(d["R"], d["t"], 50.0)
        r.copyOpticalFlowToHost(with_rgb=True)
        r.copy3dToHost()

    for _ in range(max(1, args.warmup)):
        cycle()
    steps = max(5, args.steps // 2)
    t0 = time.perf_counter()
    for _ in range(steps):
        cycle()
    dt = (time.perf_counter() - t0) / steps
    return {"frames_per_s": round(1.0 / dt, 2), "ms_per_frame": round(dt * 1e3, 3), "frames": steps,
            "cycle": "copyImagesToDevice (host) + solveR + copyOpticalFlowToHost(u, v, uvrgb) + copy3dToHost"}


def run_fixed_k(args, rf, d, path):
    """The same solve with K[2][2] = 1 on the coarse levels (the reference's setCameraMatrices leaves
    it 0 there, Reconflow.cu:99-110, which makes the coarse triangulation degenerate): frames/s, its
    mean EPE against the FP64 oracle of that mode, and both modes' median EPE against the generator's
    ground-truth flow."""
    r = make_solver(rf, path, not args.no_graph, args.tri_f64, fix=1)
    r.setCameraMatrices(d["K"], d["K"])
    r.copyImagesToDevice(d["i0"], d["i1"])
    dt = _time(lambda: r.solveR(d["R"], d["t"], 50.0), r.synchronize, max(5, args.steps // 2), max(1, args.warmup))
    u, v = r.copyOpticalFlowToHost()
    out = {"frames_per_s": round(1.0 / dt, 2)}
    gt = lambda a, b: float(np.median(np.hypot(a[:, :W] - d["gt_u"][:, :W], b[:, :W] - d["gt_v"][:, :W])))  # noqa
    out["median_epe_vs_ground_truth"] = round(gt(u, v), 4)
    if "joint" in REF64:
        out["reference_semantics_median_epe_vs_ground_truth"] = round(gt(REF64["joint"][0], REF64["joint"][1]), 3)
    if not args.no_cpu_baseline:
        from oracle import oracle as orc
        ref = orc.solve_recon(orc.recon_params(W, H, LEVELS, fix_coarse_k=1), d["K"], d["K"], np.eye(3), np.zeros(3),
                              d["R"], d["t"], d["i0"], d["i1"], "f64")
        REF64["joint_fixed_k"] = ref
        e = np.hypot(u[:, :W] - ref[0][:, :W], v[:, :W] - ref[1][:, :W])
        out["epe_vs_fp64_oracle"] = round(float(e.mean()), 5)
        out["median_epe_vs_fp64_oracle"] = round(float(np.median(e)), 6)
    return out


def run_warm_started(args, rf, torch, d, ref_fixed):
    """The learned warm start (reconflow_amd.warmnet, trained weights in reconflow_amd/weights):
    per frame the network's prediction (pyramid, six BF16 MFMA convolutions, upscale to level 0) and
    a solve that starts from it at level 3 (4 levels, the network's resolution) or level 2 (3 levels),
    5 warps x 50 iterations, K[2][2] = 1, FP64 mode. One pair per step (FP32 storage) and 8 pairs per
    batched launch (FP16 derivatives and duals, 4 levels); the network runs on the solver's stream inside the
    timed step. Accuracy against the FP64 oracle's full cold solve (5 levels) of the same mode, and the
    median distance from the generator's ground-truth flow (pair 0), as reported for the cold solve."""
    from reconflow_amd import warmnet as wn
    path = wn.WEIGHTS / "warmnet_640x480.npz"
    if not path.exists():
        return None
    out = {"weights": str(path.relative_to(ROOT))}
    for name, B, storage, L in (("1_pair_4_levels_f32", 1, rf.STORAGE_F32, 4),
                                ("1_pair_3_levels_f32", 1, rf.STORAGE_F32, 3),
                                ("8_pairs_4_levels_f16_derivatives_duals", 8, rf.STORAGE_F16_DERIV_DUALS, 4)):
        net = wn.WarmStartNet.load(path)
        p = rf.batch_params(14, L, WARPS, ITERS, lam=50.0, tau=0.125, alpha_tv=33.3, alpha_proj=60.0, lambdaf=0.1,
                            lambdams=100.0, tri_f64=int(args.tri_f64), storage=storage, fix_coarse_k=1)
        b = rf.Batch(p, [W] * B, [H] * B)
        keep, ws = [], []
        for i in range(B):
            di = d if i == 0 else rf.make_recon_pair(W, H, args.seed + i)
            a, c = torch.from_numpy(di["i0"]).cuda(), torch.from_numpy(di["i1"]).cuda()
            u = torch.zeros(H * W, dtype=torch.float32, device="cuda")
            v = torch.zeros_like(u)
            keep += [a, c, u, v]
            ws.append((a, c, u, v))
            b.set_images_device(i, a.data_ptr(), c.data_ptr())
            b.set_camera(i, di["K"], di["K"], di["R"], di["t"])
            b.set_warm_start_device(i, u.data_ptr(), v.data_ptr())
        stream = b.stream()

        def step():
            for a, c, u, v in ws:
                net.warm_start(a.data_ptr(), c.data_ptr(), W, H, W, u.data_ptr(), v.data_ptr(), stream)
            b.solve()

        torch.cuda.synchronize()
        dt = _time(step, b.synchronize, max(4, args.steps // 2), max(1, args.warmup))
        ent = {"frames_per_s": round(B / dt, 2), "ms_per_step": round(dt * 1e3, 3),
               "warm_start_used": bool(b.warm_start_in_use(0))}
        u_ptr, v_ptr = b.result_ptrs(0)[:2]
        uu, vv = _fetch(u_ptr, H * W).reshape(H, W), _fetch(v_ptr, H * W).reshape(H, W)
        ent["median_epe_vs_ground_truth"] = round(float(np.median(np.hypot(uu - d["gt_u"][:, :W],
                                                                            vv - d["gt_v"][:, :W]))), 4)
        if ref_fixed is not None:
            e = np.hypot(uu - ref_fixed[0][:, :W], vv - ref_fixed[1][:, :W])
            ent["epe_vs_fp64_oracle_5_levels"] = {"mean": round(float(e.mean()), 4),
                                                  "median": round(float(np.median(e)), 5)}
        out[name] = ent
        del b, net
    return out


def _sub(line, keys):
    """A workload's line as a sub-object of the default line."""
    return {k: line[k] for k in keys if k in line}


def run_joint(args, rf, torch):
    import ctypes as C

    d = rf.make_recon_pair(W, H, args.seed)
    path = rf.PATH_FUSED if args.path == "fused" else rf.PATH_PLAIN
    r = make_solver(rf, path, not args.no_graph, args.tri_f64, args.joint_tile)
    r.setCameraMatrices(d["K"], d["K"])
    r.copyImagesToDevice(d["i0"], d["i1"])  # resident for every step
    for _ in range(args.warmup):  # with an autotuned tiling the first solve is the tuning solve
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

    lib = r._lib
    tiling = None
    if path == rf.PATH_FUSED:
        tl = (C.c_int * LEVELS)()
        assert lib.rf_joint_tiling(r._ctx, tl, LEVELS) == 0
        tiling = [TILING[t] for t in tl]

    # dominant kernel: time it live with HIP events on the solver's stream (eager solves)
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
    for lv in range(LEVELS):  # the dominant kernel's average time per iteration and pyramid level
        assert lib.rf_profile_arm(r._ctx, KID[dom], lv) == 0
        r.solveR(d["R"], d["t"], 50.0)
        t_, n_ = C.c_double(), C.c_int()
        assert lib.rf_profile_read(r._ctx, C.byref(t_), C.byref(n_)) == 0
        per_level_us.append(round(t_.value * 1e3 / max(n_.value, 1), 3))
    # level 0: time per iteration; a launch runs k0 iterations and moves the state through HBM once
    k0 = ITERS_PER_LAUNCH.get(tiling[0], 1) if tiling else 1
    avg_s = tot_ms / 1e3 / max(nl, 1) * k0
    bytes_launch = BYTES_PER_PX[dom] * W * H
    achieved = bytes_launch / avg_s / 1e9
    traffic = None
    tfile = ROOT / "profiles" / "traffic.json"
    if tfile.exists():
        key = f"{dom}_{tiling[0] if tiling else args.path}_{W}x{H}"
        traffic = json.loads(tfile.read_text()).get(key)
    # the VALU work of the same launches (rocprofv3 SQ_INSTS_VALU per launch, profiles/valu.json) over
    # the same measured time: the issue rate the level-0 launch is bound by
    valu = None
    vfile = ROOT / "profiles" / "valu.json"
    if vfile.exists() and tiling:
        n_valu = json.loads(vfile.read_text()).get(f"{dom}_{tiling[0]}_{W}x{H}")
        if n_valu:
            rate = n_valu / avg_s / 1e9
            valu = {"wave_instructions_per_launch": n_valu, "achieved": round(rate, 1), "peak": PEAK_VALU_GINST,
                    "unit": "G wave-instructions/s", "frac": round(rate / PEAK_VALU_GINST, 4)}

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

    X, Y, Z = r.copy3dToHost()
    cpu, epe = (None, None) if args.no_cpu_baseline else cpu_baseline(d, u, v)
    zerr = depth_rel_err(Z, REF64["joint"][4]) if "joint" in REF64 else None
    cycle = None if args.no_sub else run_drop_in_cycle(args, rf, d)
    fixed = None if args.no_sub else run_fixed_k(args, rf, d, path)
    warm = None if args.no_sub else run_warm_started(args, rf, torch, d, REF64.get("joint_fixed_k"))
    batched = None
    if path == rf.PATH_FUSED and args.batch_pairs > 0:
        batched = {"pairs_per_launch": args.batch_pairs}
        for name, storage in (("f32", rf.STORAGE_F32), ("f16_derivatives", rf.STORAGE_F16_DERIV),
                              ("f16_derivatives_duals", rf.STORAGE_F16_DERIV_DUALS)):
            bfps, bu, bv, btl = run_joint_batched(args, rf, torch, args.batch_pairs, storage)
            ent = {"frames_per_s": round(bfps, 2), "joint_tiling_per_level": btl}
            if plain_fps:
                ent["speedup_vs_plain_path"] = round(bfps / plain_fps, 3)
            if "joint" in REF64:
                ru, rv = REF64["joint"][0], REF64["joint"][1]
                ent["epe_vs_fp64_oracle"] = round(float(np.hypot(bu[:, :W] - ru[:, :W], bv[:, :W] - rv[:, :W]).mean()), 5)
            batched[name] = ent
    px_iter = sum((W >> l) * (H >> l) for l in range(LEVELS)) * WARPS * ITERS
    return {
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
                   "path": args.path, "graph": not args.no_graph, "seed": args.seed,
                   "joint_tile": TILING[args.joint_tile], "joint_tiling_per_level": tiling},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
                     "level": 0, "avg_launch_us": round(avg_s * 1e6, 3), "iterations_per_launch": k0,
                     "algorithmic_bytes_per_launch": bytes_launch, "avg_iteration_us_per_level": per_level_us,
                     "valu_issue": valu},
        "cpu_baseline": cpu,
        "epe_vs_fp64_oracle": None if epe is None else round(epe, 5),
        "depth_rel_err_vs_fp64_oracle": zerr,
        "accuracy_ok": None if epe is None else bool(epe <= 0.05),
        "drop_in_cycle": cycle,
        "fixed_coarse_k": fixed,
        "learned_warm_start": warm,
        "plain_fp32_path_fps": None if plain_fps is None else round(plain_fps, 3),
        "speedup_vs_plain_path": None if plain_fps is None else round(fps / plain_fps, 3),
        "mpix_iter_per_s": round(fps * px_iter / 1e6, 1),
        "batched_640x480": batched,
    }


def run_batch(args, rf, torch):
    rng = np.random.default_rng(2024)
    sizes = [(160 + 16 * int(rng.integers(0, 71)), 120 + 16 * int(rng.integers(0, 38))) for _ in range(64)]
    groups = {}
    for kind, method, L, nw, ni in (("flow", 0, 4, 3, 30), ("joint", 14, 5, 5, 50)):
        idx = [k for k in range(64) if (k % 2 == 0) == (kind == "flow")]
        ws, hs = [sizes[k][0] for k in idx], [sizes[k][1] for k in idx]
        if kind == "flow":
            p = rf.batch_params(0, L, nw, ni, lam=50.0, theta=0.3, tau=0.25, storage=rf.STORAGE_F16)
        else:
            p = rf.batch_params(14, L, nw, ni, lam=50.0, tau=0.125, alpha_tv=33.3, alpha_proj=60.0, lambdaf=0.1,
                                lambdams=100.0, tri_f64=args.tri_f64, storage=rf.STORAGE_F16_DERIV)
        b = rf.Batch(p, ws, hs)
        keep = []
        for i, k in enumerate(idx):
            s = b.stride(i)
            if kind == "flow":
                i0, i1, _, _ = rf.make_flow_pair(ws[i], hs[i], k, stride=s)
            else:
                dd = rf.make_recon_pair(ws[i], hs[i], k, stride=s)
                i0, i1 = dd["i0"], dd["i1"]
                b.set_camera(i, dd["K"], dd["K"], dd["R"], dd["t"])
            t0, t1 = torch.from_numpy(i0).cuda(), torch.from_numpy(i1).cuda()
            keep += [t0, t1]
            b.set_images_device(i, t0.data_ptr(), t1.data_ptr())
        px_iter = sum(sum((w_ >> l) * (h_ >> l) for l in range(L)) * nw * ni for w_, h_ in zip(ws, hs))
        groups[kind] = dict(b=b, keep=keep, px_iter=px_iter, n=len(idx), ws=ws, hs=hs, L=L, nw=nw, ni=ni)
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        for g in groups.values():
            g["b"].solve()
        for g in groups.values():
            g["b"].synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for g in groups.values():  # the two batches run concurrently on their own streams
            g["b"].solve()
        for g in groups.values():
            g["b"].synchronize()
    dt = time.perf_counter() - t0
    pairs_s = 64 * args.steps / dt
    # algorithmic bytes: one pass of the state per launch (a launch running k iterations moves it
    # through HBM once), summed over levels with the tuned iterations per launch
    alg, tilings = 0.0, {}
    for kind, g in groups.items():
        codes = g["b"].tiling(g["L"])
        names = FLOW_TILING if kind == "flow" else TILING
        tilings[kind] = [names.get(c, str(c)) for c in codes]
        per_px = FLOW_ITER_BYTES if kind == "flow" else BYTES_PER_PX["fused_iter"]
        for l, name in enumerate(tilings[kind]):
            k = ITERS_PER_LAUNCH.get(name, 1)
            launches = g["ni"] // k + g["ni"] % k
            alg += per_px * launches * g["nw"] * sum((w_ >> l) * (h_ >> l) for w_, h_ in zip(g["ws"], g["hs"]))
    alg_gbs = alg * args.steps / dt / 1e9
    return {
        "metric": "batched pairs/s: 64 seeded pairs of uneven sizes, half TV-L1 half joint flow+depth",
        "value": round(pairs_s, 3),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 updates; f16 storage: TV-L1 images, derivatives and duals, joint derivatives"
                 + (", f64 triangulation" if args.tri_f64 else ""),
        "data": "synthetic (seeded pairs, rendered on the host)",
        "config": {"workload": "64 pairs, W 160..1280, H 120..720 (multiples of 16); TV-L1 4 levels x 3 warps x "
                               "30 iterations, joint 5 levels x 5 warps x 50 iterations",
                   "pixels": int(sum(w_ * h_ for w_, h_ in sizes))},
        "mpix_iter_per_s": round((groups["flow"]["px_iter"] + groups["joint"]["px_iter"]) * args.steps / dt / 1e6, 1),
        "algorithmic_gb_per_s_fp32_equivalent": round(alg_gbs, 1),
        "tiling_per_level": tilings,
        "roofline": {"bound": "hbm", "achieved": round(alg_gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(alg_gbs / PEAK_HBM_GBS, 4), "traffic": None},
        "cpu_baseline": None,
    }


def _time(fn, sync, steps, warmup):
    for _ in range(warmup):
        fn()
    sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    sync()
    return (time.perf_counter() - t0) / steps


def run_tvl1(args, rf, torch):
    """TV-L1 at 320x240 (4 levels, 3 warps, 30 iterations, lambda 50, theta 0.3, tau 0.25):
    fused graph replays, the plain path beside it, EPE and time of the FP64 oracle."""
    w, h, L, nw, ni = 320, 240, 4, 3, 30
    i0, i1, _, _ = rf.make_flow_pair(w, h, args.seed)

    def solver(path, graph, storage=rf.STORAGE_F32):
        f = rf.CudaFlow()
        f.initialize(w, h, 1, rf.INPUT_32F, L, 2.0, rf.CudaFlow.METHOD_TVL1, 50.0, 0.0, 0.3, 0.25, nw, ni, False,
                     path=path, use_graph=graph, storage=storage)
        f.copyImagesToDevice(i0, i1)
        return f

    out, tiling = {}, {}
    for name, path, storage in (("fused", rf.PATH_FUSED, rf.STORAGE_F32), ("fused_fp16", rf.PATH_FUSED,
                                                                          rf.STORAGE_F16),
                                ("plain", rf.PATH_PLAIN, rf.STORAGE_F32)):
        f = solver(path, not args.no_graph, storage)
        dt = _time(lambda: f.solveOpticalFlow(5.0), f.synchronize, args.steps, args.warmup)
        out[name] = (dt, f.copyOpticalFlowToHost())
        if path == rf.PATH_FUSED:
            tl = (ctypes.c_int * L)()
            assert f._lib.rf_flow_tiling(f._ctx, tl, L) == 0
            tiling[name] = [FLOW_TILING[t] for t in tl]
    from oracle import oracle as orc
    tc = time.perf_counter()
    ref = orc.solve_flow(orc.flow_params(w, h, L), i0, i1, "f64")
    tcpu = time.perf_counter() - tc
    epe = {k: float(np.hypot(v[1][0][:, :w] - ref[0][:, :w], v[1][1][:, :w] - ref[1][:, :w]).mean())
           for k, v in out.items()}
    px_iter = sum((w >> l) * (h >> l) for l in range(L)) * nw * ni
    fps = 1.0 / out["fused"][0]
    return {
        "metric": "320x240 TV-L1 frames/s (4 levels x 3 warps x 30 iterations)", "value": round(fps, 2),
        "unit": "frames/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(out["fused"][0] * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded affine-flow pair)",
        "config": {"workload": "TV-L1 320x240", "graph": not args.no_graph},
        "frames_per_s": {k: round(1.0 / v[0], 2) for k, v in out.items()},
        "mpix_iter_per_s": round(fps * px_iter / 1e6, 1),
        "epe_vs_fp64_oracle": {k: round(v, 6) for k, v in epe.items()},
        "flow_tiling_per_level": tiling,
        "cpu_baseline": {"value": round(1.0 / tcpu, 3), "unit": "frames/s", "cores": orc.get_threads(),
                         "kind": "port", "sample": f"one 320x240 TV-L1 frame of the FP64 oracle, {tcpu:.3f} s"},
    }


def run_stream(args, rf, torch):
    """Streaming 1280x720 joint flow+depth: consecutive pairs of a synthetic camera trajectory
    (per-frame rotation <= 0.5 deg, camera displacement <= 0.02, 16 planes at depths 2-20), 5 levels x
    5 warps x 50 iterations, FP64 mode, K[2][2] = 1 on the coarse levels (with the reference's
    degenerate coarse cameras the joint flow is not a usable warm start). Runs: one pair per launch
    cold; one pair per launch, each pair warm-started from a device copy of the previous pair's flow
    (rf_batch_set_warm_start_device; the rule falls back to a cold start when that flow is not usable);
    eight consecutive pairs per batched launch, cold; the 299 pairs cut into eight chunks of 37
    consecutive pairs, launch j solving pair j of every chunk, each slot warm-started from its own
    previous result (a chunk's first pair reads zero warm-start buffers, the cold solve). A frame's
    latency is its launch's time from submission to completion (waiting for a batch's later frames is
    not counted)."""
    w, h, n = 1280, 720, args.stream_frames
    t_gen = time.perf_counter()
    seq = rf.make_sequence(w, h, n, args.seed)
    t_gen = time.perf_counter() - t_gen
    frames = torch.from_numpy(seq["frames"]).cuda()
    K, R, t = seq["K"], seq["R"], seq["t"]
    runs, pair0 = {}, None
    f16 = rf.STORAGE_F16_DERIV_DUALS
    for name, B, storage, mode in (("1_pair_cold_f16_derivatives_duals", 1, f16, "cold"),
                                   ("1_pair_warm_started_f16_derivatives_duals", 1, f16, "warm"),
                                   ("8_pairs_cold_f16_derivatives_duals", 8, f16, "cold"),
                                   ("8_chunks_warm_started_f16_derivatives_duals", 8, f16, "chunks")):
        p = rf.batch_params(14, LEVELS, WARPS, ITERS, lam=50.0, tau=0.125, alpha_tv=33.3, alpha_proj=60.0,
                            lambdaf=0.1, lambdams=100.0, tri_f64=int(args.tri_f64), storage=storage, fix_coarse_k=1)
        b = rf.Batch(p, [w] * B, [h] * B)
        st = b.stride(0)
        nws = B if mode == "chunks" else 1
        ws = [(torch.zeros(h * st, dtype=torch.float32, device="cuda"),
               torch.zeros(h * st, dtype=torch.float32, device="cuda")) for _ in range(nws)]
        chunk = (n - 1) // B  # chunks: slot i runs pairs i * chunk .. i * chunk + chunk - 1

        def pair_of(step, i):
            return i * chunk + step if mode == "chunks" else step * B + i

        def submit(step):
            for i in range(B):
                k = pair_of(step, i)
                b.set_images_device(i, frames[k].data_ptr(), frames[k + 1].data_ptr())
                b.set_camera(i, K, K, R[k], t[k])
            b.solve()
            # the next warm start of each warm-started slot: this solve's flow, copied on the batch stream
            for i in range(nws if mode != "cold" else 0):
                u_ptr, v_ptr = b.result_ptrs(i)[:2]
                _copy_async(ws[i][0].data_ptr(), u_ptr, h * st * 4, b.stream())
                _copy_async(ws[i][1].data_ptr(), v_ptr, h * st * 4, b.stream())

        def restart():
            # a chunk's first pair reads zero warm-start buffers: the cold solve bit for bit
            b.synchronize()
            for wu, wv in ws:
                wu.zero_()
                wv.zero_()
            torch.cuda.synchronize()

        submit(0)  # the first solve of a batch includes the tiling tuning solves
        b.synchronize()
        if pair0 is None:
            u_ptr, v_ptr = b.result_ptrs(0)[:2]
            pair0 = (_fetch(u_ptr, h * st).reshape(h, st), _fetch(v_ptr, h * st).reshape(h, st))
        if mode == "warm":
            b.set_warm_start_device(0, ws[0][0].data_ptr(), ws[0][1].data_ptr())
        if mode == "chunks":
            for i in range(B):
                b.set_warm_start_device(i, ws[i][0].data_ptr(), ws[i][1].data_ptr())
            restart()
        nsteps = chunk if mode == "chunks" else (n - 1) // B
        npairs = nsteps * B
        lat, used, nonfinite = [], 0, 0
        t0 = time.perf_counter()
        for step in range(nsteps):
            ts = time.perf_counter()
            submit(step)
            b.synchronize()
            lat.append(time.perf_counter() - ts)
        dt = time.perf_counter() - t0
        if mode != "cold":  # one more pass, untimed, for the rule's decisions and the results' finiteness
            if mode == "chunks":
                restart()
            else:
                submit(0)
                b.synchronize()
            for step in range(0 if mode == "chunks" else 1, nsteps if mode == "chunks" else npairs):
                submit(step)
                b.synchronize()
                for i in range(nws):
                    if mode != "chunks" or step > 0:
                        used += int(b.warm_start_in_use(i))
                    if not np.isfinite(_fetch(b.result_ptrs(i)[0], h * st)).all():
                        nonfinite += 1
        runs[name] = {"frames_per_s": round(npairs / dt, 2), "pairs": npairs,
                      "latency_ms_p50": round(float(np.percentile(lat, 50)) * 1e3, 3),
                      "latency_ms_p99": round(float(np.percentile(lat, 99)) * 1e3, 3),
                      "joint_tiling_per_level": [TILING.get(c, str(c)) for c in b.tiling(LEVELS)]}
        if mode != "cold":
            runs[name].update({"warm_start_used": used, "pairs_with_nonfinite_flow": nonfinite})
        if mode == "chunks":
            runs[name].update({"chunks": B, "pairs_per_chunk": chunk,
                               "warm_start_offered": (chunk - 1) * B})
        del b
    epe = None
    if not args.no_cpu_baseline:
        from oracle import oracle as orc
        f = seq["frames"]
        ref = orc.solve_recon(orc.recon_params(w, h, LEVELS, fix_coarse_k=1), K, K, np.eye(3), np.zeros(3), R[0],
                              t[0], f[0], f[1], "f64")
        e = np.hypot(pair0[0][:, :w] - ref[0][:, :w], pair0[1][:, :w] - ref[1][:, :w])
        epe = {"mean": round(float(e.mean()), 5), "median": round(float(np.median(e)), 6)}
    best = runs["8_chunks_warm_started_f16_derivatives_duals"]
    return {
        "metric": "1280x720 streaming joint flow+depth frames/s (the sequence in 8 chunks of consecutive pairs, one "
                  "pair of each chunk per batched launch, each pair warm-started from the previous pair of its chunk)",
        "value": best["frames_per_s"], "unit": "frames/s", "n_gpus": 1, "steps": best["pairs"], "warmup": 1,
        "ms_per_step": round(1e3 / best["frames_per_s"], 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 updates, f16 derivatives and duals" + (", f64 mode" if args.tri_f64 else ""),
        "data": f"synthetic {n}-frame camera trajectory over a piecewise-planar scene (generated in {t_gen:.1f} s)",
        "config": {"workload": "stream 1280x720, 5 levels x 5 warps x 50 iterations, K[2][2] = 1 on coarse levels"},
        "runs": runs, "epe_pair0_vs_fp64_oracle": epe, "cpu_baseline": None,
    }


def run_methods(args, rf, torch):
    """Each CudaFlow data term at 320x240 (4 levels, 3 warps, 30 iterations) on the per-step
    path, graph-replayed: frames/s per method."""
    w, h, L, nw, ni = 320, 240, 4, 3, 30
    i0, i1, _, _ = rf.make_flow_pair(w, h, args.seed)
    mask = (np.random.default_rng(1).random(i0.shape) > 0.2).astype(np.float32)
    depth = np.random.default_rng(2).integers(500, 2500, i0.shape).astype(np.uint16)
    res = {}
    for name, method in (("tvl1", 0), ("charbonnier", 1), ("charbonnier_gradient", 2), ("patch_derivatives", 3),
                         ("masked_tvl1", 8), ("scene_flow", 4)):
        f = rf.CudaFlow()
        f.initialize(w, h, 1, rf.INPUT_32F, L, 2.0, method, 50.0, 20.0, 0.3, 0.25, nw, ni, False,
                     path=rf.PATH_PLAIN, use_graph=not args.no_graph)
        f.copyImagesToDevice(i0, i1)
        if method == 8:
            f.copyMaskToDevice(mask, mask)
            fn = lambda f=f: f.solveInpainting(5.0)  # noqa: E731
        elif method == 4:
            f.copyDepthToDevice(depth, depth)
            fn = lambda f=f: f.solveSceneFlowAndPointCloud(500.0, 500.0, 160.0, 120.0, 5.0, 1.0)  # noqa: E731
        else:
            fn = lambda f=f: f.solveOpticalFlow(5.0)  # noqa: E731
        dt = _time(fn, f.synchronize, args.steps, args.warmup)
        res[name] = round(1.0 / dt, 2)
    return {
        "metric": "320x240 CudaFlow frames/s per data term (per-step path)", "value": res["tvl1"],
        "unit": "frames/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 / res["tvl1"], 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded affine-flow pair, random masks/depth)",
        "config": {"workload": "CudaFlow methods 320x240, 4 levels x 3 warps x 30 iterations"},
        "frames_per_s": res, "cpu_baseline": None,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=["joint640", "batch64", "tvl1-320", "methods", "stream720"],
                    default="joint640")
    ap.add_argument("--stream-frames", type=int, default=300, help="stream720: frames of the sequence")
    ap.add_argument("--path", choices=["fused", "plain"], default="fused")
    ap.add_argument("--batch-pairs", type=int, default=32,
                    help="joint640: also time this many 640x480 pairs per batched launch (0: skip)")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--tri-f64", type=int, default=1, help="FP64 triangulation (1) or the reference FP32 (0)")
    ap.add_argument("--joint-tile", type=int, default=0, choices=sorted(TILING),
                    help="fused joint tiling of the one-pair and batched 640x480 runs: 0 autotuned per level, 1 64x8 two rows per lane, 2 32x16, 3 16x16, "
                         "4 16x16 capped at 96 VGPRs")
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain-compare", action="store_true")
    ap.add_argument("--no-sub", action="store_true",
                    help="joint640: skip the drop-in cycle, fixed-intrinsics, TV-L1 320x240 and 64-pair sub-objects")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 256 MiB device-to-device copy 5 times (scales the FETCH_SIZE / WRITE_SIZE "
                         "counters to bytes) and exit")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this solver runs one image pair per device and the "
              "multi-GPU sharding (independent pairs per GPU) is not exercised here.", file=sys.stderr)
        return 2

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
    run = {"batch64": run_batch, "tvl1-320": run_tvl1, "methods": run_methods, "stream720": run_stream}.get(
        args.workload, run_joint)
    line = run(args, rf, torch)
    if run is run_joint and not args.no_sub:
        t = run_tvl1(args, rf, torch)
        line["tvl1_320"] = _sub(t, ("metric", "value", "unit", "ms_per_step", "frames_per_s", "mpix_iter_per_s",
                                    "epe_vs_fp64_oracle", "flow_tiling_per_level", "cpu_baseline"))
        b = run_batch(args, rf, torch)
        line["batch64"] = _sub(b, ("metric", "value", "unit", "ms_per_step", "dtype", "config", "mpix_iter_per_s",
                                   "algorithmic_gb_per_s_fp32_equivalent", "tiling_per_level", "roofline"))
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
