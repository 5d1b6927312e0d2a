#!/usr/bin/env python3
"""Benchmark of the differentiable 2D-Gaussian-surfel training step on one MI355X.

Workload (default): 1M surfels, SH degree 3, heavy-tailed log-normal scales
(mu -5, sigma 1), 4 views of 1920x1080 per step drawn from 16 cameras on a
sphere of radius 3, loss 0.8 L1 + 0.2 (1 - SSIM) against targets rendered
from a jittered copy of the scene, one Adam step per training step.

--workload c4: the streaming-frame batch instead -- 16 frames of 1280x720
from a camera on an SE(3) random walk (1 cm / 0.5 deg per frame) over a scene
of 500k surfels that grows 10 % per frame (frame k renders its first
500k + 50k k surfels; 1.25M in all), every frame in one graph launch per
step, targets rendered from a jittered copy.

A "step" = project, scan, emit keys, 64-bit radix sort, tile ranges,
rasterize forward, L1 + D-SSIM, rasterize backward, geometry backward and
fused Adam over the 4 views. `value` is steps/s of the batched step captured
in a hipGraph; the plain per-view step (every kernel launched per view, no
graph) is timed beside it. Inputs are resident in HBM before timing starts.

Roofline: the dominant stage's algorithmic bytes (per-stage formulas of the
design notes) divided by its duration measured with HIP events on the step
stream (profiled batched steps after the timed region).
CPU baseline: the FP64 oracle (oracle/, OpenMP tile-binned mode) on a
bounded sample -- one view's forward + loss + backward over a centred crop,
scaled to a full step -- timed on this host.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "train steps/s, 1M surfels @1080p, batched vs plain at equal loss; % HBM peak"
METRIC_C4 = "train steps/s, 16-frame 720p streaming batch over a growing 500k..1.25M-surfel scene"
HBM_PEAK_GBS = 8000.0
# Non-packed VALU issue rate: 256 CUs x 4 SIMDs, one wave64 instruction per
# SIMD every 4 cycles, 2.4 GHz (packed FP32 is not used).
VALU_PEAK_WINST = 256 * 2.4e9
COUNTERS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "c3_step_counters.json")
STAGE_KERNEL = {"raster_bwd": "raster_bwd_kernel", "raster_fwd": "raster_fwd_kernel",
                "sort": "onesweep_kernel", "project": "project_kernel", "l1_ssim": "ssim_fwd_kernel",
                "adam": "adam_kernel", "geom_bwd": "geom_bwd_kernel"}


def profile_figures(stage):
    """HBM bytes and VALU rate of the stage's main kernel at its largest
    launch, from the committed rocprofv3 summary (None if absent)."""
    if not os.path.exists(COUNTERS) or stage not in STAGE_KERNEL:
        return None
    with open(COUNTERS) as f:
        summ = json.load(f)
    cands = [k for k in summ["kernels"] if STAGE_KERNEL[stage] in k["kernel"]]
    if not cands:
        return None
    # The largest launches of the kernel (several views per launch).
    k = max((c.get("large", c) | {"kernel": c["kernel"]} for c in cands), key=lambda k: k["mean_ms"])
    out = {"source": os.path.relpath(COUNTERS, os.path.dirname(COUNTERS) + "/.."), "kernel": k["kernel"],
           "profiled_ms": round(k["mean_ms"], 4)}
    if "hbm_read_bytes" in k and "hbm_write_bytes" in k:
        out["traffic"] = int(k["hbm_read_bytes"] + k["hbm_write_bytes"])
    cs = k.get("counters_per_dispatch", {})
    if "SQ_INSTS_VALU" in cs:
        rate = cs["SQ_INSTS_VALU"] / (k["mean_ms"] * 1e-3)
        out["valu"] = {"achieved": round(rate / 1e9, 1), "peak": round(VALU_PEAK_WINST / 1e9, 1),
                       "unit": "G wave-instr/s", "frac": round(rate / VALU_PEAK_WINST, 4)}
    return out


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--workload", choices=["c3", "c4"], default="c3",
                   help="c3: 1M surfels, 4 x 1080p views per step (default); c4: 16-frame 720p "
                        "streaming batch over a growing 500k..1.25M-surfel scene")
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--n", type=int, default=1_000_000, help="surfels")
    p.add_argument("--width", type=int, default=1920)
    p.add_argument("--height", type=int, default=1080)
    p.add_argument("--views-per-step", type=int, default=4)
    p.add_argument("--num-views", type=int, default=16)
    p.add_argument("--plain-steps", type=int, default=None)
    p.add_argument("--profile-steps", type=int, default=3)
    p.add_argument("--digit-bits", type=int, default=8)
    p.add_argument("--binning", choices=["tight", "presort", "sort64"], default="tight",
                   help="tile binning: presort with footprint-clipped rects (same images), "
                        "depth presort + 32-bit tile sort, or one 64-bit key sort")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-autotune", action="store_true",
                   help="keep --digit-bits and the default tile queue instead of autotuning")
    p.add_argument("--psnr-steps", type=int, default=500,
                   help="steps of the batched-vs-plain equal-loss comparison (0 = skip)")
    p.add_argument("--schedule", default=None,
                   help="DIGIT_BITS,BLOCKS_PER_CU[,TILE,FWD_BATCH,BWD_BATCH]: set this schedule "
                        "instead of autotuning")
    p.add_argument("--cpu-crop", default="480x270")
    p.add_argument("--copy-calibration", action="store_true",
                   help="only run 5 device copies of 1 GiB (read 1 GiB + write 1 GiB each), the "
                        "known-size reference for scaling FETCH_SIZE / WRITE_SIZE counters")
    return p.parse_args()


def copy_calibration():
    import torch
    n = (1 << 30) // 4
    src = torch.rand(n, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "copies": 5}))


def stage_bytes(n, V, E, P, T, sort_passes, S=232, nv=None):
    """Algorithmic HBM bytes per stage for one step over V views (E, P, T
    summed over the views; nv = sum over the views of the surfels each sees,
    V n by default): compulsory traffic only."""
    nv = V * n if nv is None else nv
    return {
        "project": nv * (S + 60),
        "scan": 8 * nv,
        "emit_keys": 24 * nv + 12 * E,
        "sort": (8 + 24 * sort_passes) * E,
        "tile_ranges": 8 * E + 8 * T,
        "raster_fwd": 56 * E + 20 * P,
        "l1_ssim": 36 * P,
        "grad_clear": 80 * nv,
        "raster_bwd": 56 * E + 20 * P + 52 * nv,
        "geom_bwd": nv * (2 * S + 52),
        "adam": 7 * S * n,
    }


def cpu_baseline(scene, spec, crop, views_per_step, n_view=None):
    """n_view: surfels the sampled view renders (a prefix; all by default)."""
    import oracle
    n = scene.n if n_view is None else n_view
    deg = scene.sh_degree
    ncoef = (deg + 1) ** 2
    s64 = oracle.Scene(scene.positions[:n].astype(np.float64), scene.scales[:n].astype(np.float64),
                       scene.quats[:n].astype(np.float64), scene.opacities[:n].astype(np.float64),
                       scene.sh[:n].astype(np.float64).reshape(n, ncoef, 3), deg)
    W, H = spec.width, spec.height
    cw, ch = crop
    x0, y0 = (W - cw) // 2, (H - ch) // 2

    def cam(c):
        return oracle.Camera(np.asarray(spec.center, np.float64), np.asarray(spec.quat, np.float64),
                             float(spec.focal), W, H, c)

    # Fixed per-view cost (prepare: projection + SH + depth sort; tile lists).
    t0 = time.perf_counter()
    oracle.render(s64, cam((x0, y0, 1, 1)))
    oracle.gradients(s64, cam((x0, y0, 1, 1)), np.zeros((1, 3)))
    t_fixed = time.perf_counter() - t0
    t0 = time.perf_counter()
    img = oracle.render(s64, cam((x0, y0, cw, ch)))["color"]
    tgt = np.clip(img + 0.01, 0, 1)
    _, dl, _, _ = oracle.l1_ssim(img, tgt, 0.2, 1.0)
    oracle.gradients(s64, cam((x0, y0, cw, ch)), dl.reshape(-1, 3))
    t_crop = time.perf_counter() - t0
    raw = np.zeros(scene.n * 58, np.float32)
    g = np.full(scene.n * 58, 1e-3, np.float32)
    m = np.zeros_like(raw)
    v = np.zeros_like(raw)
    t0 = time.perf_counter()
    oracle.adam_step_f32(raw, g, m, v, 1e-3, 0, 1)
    t_adam = time.perf_counter() - t0
    scale = (W * H) / float(cw * ch)
    t_view = t_fixed + max(t_crop - t_fixed, 0.0) * scale
    t_step = views_per_step * t_view + t_adam
    return {"value": 1.0 / t_step, "unit": "steps/s", "cores": oracle.num_threads(), "kind": "port",
            "sample": (f"FP64 oracle (tile-binned, OpenMP) on view 0: fixed per-view cost {t_fixed:.2f}s + "
                       f"centred {cw}x{ch} crop fwd+L1/SSIM+bwd {t_crop:.2f}s scaled x{scale:.1f} to "
                       f"{W}x{H}; x{views_per_step} views + f32 Adam over {scene.n}x58 params "
                       f"{t_adam:.2f}s = {t_step:.1f}s/step"),
            "sample_seconds": round(t_fixed + t_crop + t_adam, 2)}


def main():
    a = parse()
    if a.copy_calibration:
        copy_calibration()
        return
    if a.gpus != 1:
        print(f"bench.py runs on exactly one GPU (asked for --gpus {a.gpus}); the multi-GPU "
              "sharding of the step (by view, one gradient all-reduce) is designed, not built.",
              file=sys.stderr)
        sys.exit(2)
    import splattar_amd as sa
    from splattar_amd._lib import GS_BINNING_PRESORT, GS_BINNING_SORT64, GS_BINNING_TIGHT
    from splattar_amd import scenes

    ctx = sa.GpuContext(0)
    name, cus = ctx.device_info()
    log(f"device {name}, {cus} CUs")
    if a.workload == "c4":
        W, H, V = 1280, 720, 16
        a.num_views = 16
        counts = scenes.frame_surfel_counts(16, 500_000)
        n = counts[-1]
        scene = scenes.config_scene(4, n)
        target_scene = scenes.jittered(scene, 0.01, (scenes.CONFIG_SEED + 4) ^ 1)
        specs = scenes.walk_views(16, scenes.CONFIG_SEED + 4, (W, H))
    else:
        n, W, H, V = a.n, a.width, a.height, a.views_per_step
        scene = scenes.config_scene(3, n)
        target_scene = scenes.jittered(scene, 0.01, (scenes.CONFIG_SEED + 3) ^ 1)
        specs = scenes.sphere_views(a.num_views, scenes.CONFIG_SEED + 3, [(W, H)])
        counts = [n] * a.num_views
    deg = scene.sh_degree

    # Targets rendered on the GPU from the jittered scene (setup, untimed).
    t0 = time.perf_counter()
    cap_entries = min(64 * n, (1 << 30) // 2)
    r = sa.Rasterizer(ctx, n, W, H, cap_entries)
    r.upload_scene(target_scene.positions, target_scene.scales, target_scene.quats,
                   target_scene.opacities, target_scene.sh)
    targets, ents = [], []
    for sp, m in zip(specs, counts):
        r.forward(sa.Camera(sp.center, sp.quat, sp.focal, deg), m)
        ents.append(r.entry_count())
        targets.append(r.out_color.read().copy())
    r.close()
    log(f"targets rendered in {time.perf_counter() - t0:.1f}s; entries per view "
        f"min {min(ents)} mean {int(np.mean(ents))} max {max(ents)}")

    extent = float(np.linalg.norm(scene.positions.max(0) - scene.positions.min(0)) * 0.5)
    per_view = int(max(ents) * 1.3) + 1024
    cfg = sa.TrainConfig(lam=0.2, dssim_scale=1.0, lr_pos=1.6e-4 * extent, lr_scales=5e-3, lr_quat=1e-3,
                         lr_opacity=5e-2, lr_sh=2.5e-3, max_views_per_step=V, entries_per_view=per_view,
                         digit_bits=a.digit_bits,
                         binning={"tight": GS_BINNING_TIGHT, "presort": GS_BINNING_PRESORT,
                                  "sort64": GS_BINNING_SORT64}[a.binning])
    views = [(sa.Camera(sp.center, sp.quat, sp.focal, deg), t, m) for sp, t, m in zip(specs, targets, counts)]
    del targets

    def ids_for(step):
        return [(step * V + j) % a.num_views for j in range(V)]

    def run(trainer, mode, steps, start):
        for s in range(steps):
            trainer.step(ids_for(start + s), mode=mode)

    # ---- batched + graph (the headline)
    tr = sa.Trainer(ctx, cfg, scene.as_init(), views)
    schedule = {"digit_bits": a.digit_bits, "blocks_per_cu": None, "tile_size": 16, "fwd_batch": 256,
                "bwd_batch": 128}
    if a.schedule:
        f = [int(x) for x in a.schedule.split(",")]
        tr.set_schedule(f[0], f[1])
        schedule.update(digit_bits=f[0], blocks_per_cu=f[1])
        if len(f) == 5:
            tr.set_tiling(f[2], f[3], f[4])
            schedule.update(tile_size=f[2], fwd_batch=f[3], bwd_batch=f[4])
    elif not a.no_autotune:
        # Untimed: real training steps under each candidate schedule.
        t0 = time.perf_counter()
        schedule = tr.autotune(ids_for(0), reps=4)
        log(f"autotune picked {schedule} in {time.perf_counter() - t0:.1f}s")
    run(tr, tr.GRAPH, a.warmup, 0)
    ctx.synchronize()
    t0 = time.perf_counter()
    run(tr, tr.GRAPH, a.steps, a.warmup)
    ctx.synchronize()
    t_graph = time.perf_counter() - t0
    graph_losses = tr.last_losses(V)
    entries_step = tr.entry_count()
    log(f"graph: {a.steps} steps in {t_graph:.3f}s -> {a.steps / t_graph:.2f} steps/s; "
        f"entries/step {entries_step}")

    # ---- batched without graph, then profiled batched steps (stage times)
    run(tr, tr.BATCHED, 1, 0)
    ctx.synchronize()
    kb = max(2, a.steps // 4)
    t0 = time.perf_counter()
    run(tr, tr.BATCHED, kb, 1)
    ctx.synchronize()
    t_batched = (time.perf_counter() - t0) / kb
    tr.set_profiling(True)
    run(tr, tr.BATCHED, a.profile_steps, 3)
    stages, psteps = tr.stage_times()
    tr.set_profiling(False)
    stage_ms = {k: v / max(psteps, 1) for k, v in stages.items()}
    log("stage ms/step (batched): " + ", ".join(f"{k} {v:.3f}" for k, v in stage_ms.items()))

    # ---- plain per-view steps (same views, same math, no graph)
    kp = a.plain_steps if a.plain_steps is not None else max(2, a.steps // 4)
    run(tr, tr.PLAIN, 1, 0)
    ctx.synchronize()
    t0 = time.perf_counter()
    run(tr, tr.PLAIN, kp, 1)
    ctx.synchronize()
    t_plain = (time.perf_counter() - t0) / kp
    tr.set_profiling(True)
    run(tr, tr.PLAIN, 1, 2)
    plain_stages, _ = tr.stage_times()
    tr.set_profiling(False)
    tr.close()

    # ---- equal-loss check: the same K steps batched+graph and plain, from
    # the same initial state, then mean PSNR over every view (untimed).
    equal_loss = None
    if a.psnr_steps > 0:
        def train_and_score(mode):
            t = sa.Trainer(ctx, cfg, scene.as_init(), views)
            if schedule["blocks_per_cu"] is not None:
                t.set_schedule(schedule["digit_bits"], schedule["blocks_per_cu"])
            t.set_tiling(schedule["tile_size"], schedule["fwd_batch"], schedule["bwd_batch"])
            t0 = time.perf_counter()
            run(t, mode, a.psnr_steps, 0)
            ctx.synchronize()
            secs = time.perf_counter() - t0
            ps = []
            for v, (_, tgt, _) in enumerate(views):
                img = t.render(v)[..., :3].astype(np.float64)
                mse = float(np.mean((img - tgt[..., :3].astype(np.float64)) ** 2))
                ps.append(10.0 * np.log10(1.0 / max(mse, 1e-20)))
            t.close()
            return float(np.mean(ps)), a.psnr_steps / secs
        pb, rb = train_and_score(sa.Trainer.GRAPH)
        pp, rp = train_and_score(sa.Trainer.PLAIN)
        equal_loss = {"steps": a.psnr_steps, "psnr_batched_graph_db": round(pb, 4),
                      "psnr_plain_db": round(pp, 4), "psnr_diff_db": round(pb - pp, 4),
                      "steps_per_s_batched_graph": round(rb, 2), "steps_per_s_plain": round(rp, 2),
                      "ratio": round(rb / rp, 3)}
        log(f"equal-loss check: {equal_loss}")

    ms_graph = 1000.0 * t_graph / a.steps
    value = a.steps / t_graph
    # Roofline of the dominant kernel.
    P = V * W * H
    ts = schedule["tile_size"]
    T = V * ((W + ts - 1) // ts) * ((H + ts - 1) // ts)
    import math
    tile_bits = max(1, math.ceil(math.log2(T)))
    passes = math.ceil((32 + tile_bits) / schedule["digit_bits"])
    nv = sum(counts[j] for j in ids_for(0))
    sb = stage_bytes(n, V, entries_step, P, T, passes, nv=nv)
    dom = max(stage_ms, key=stage_ms.get)
    achieved = sb[dom] / (stage_ms[dom] * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None, "kernel": dom,
                "ms_per_launch": round(stage_ms[dom], 4), "algorithmic_bytes": int(sb[dom])}
    # The committed counter summary is of the default (C3) workload.
    prof = profile_figures(dom) if a.workload == "c3" else None
    if prof is not None:
        roofline["traffic"] = prof.pop("traffic", None)
        roofline["profile"] = prof
    total_bytes = sum(sb.values())
    result = {
        "metric": METRIC if a.workload == "c3" else METRIC_C4, "value": round(value, 4), "unit": "steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(ms_graph, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": (f"C3: {n} surfels SH3 heavy-tailed scales, {V}x{W}x{H} views/step, "
                                f"L1+SSIM, Adam") if a.workload == "c3" else
                               (f"C4: {V} frames {W}x{H}, SE(3) walk 1 cm / 0.5 deg per frame, "
                                f"{counts[0]}..{counts[-1]} surfels per frame (+10 %/frame), one graph "
                                f"launch per step, L1+SSIM, Adam"),
                   "surfels": n, "views_per_step": V, "width": W,
                   "height": H, "entries_per_step": entries_step, "digit_bits": schedule["digit_bits"],
                   "raster_blocks_per_cu": schedule["blocks_per_cu"], "tile_size": schedule["tile_size"],
                   "fwd_batch": schedule["fwd_batch"], "bwd_batch": schedule["bwd_batch"],
                   "binning": a.binning,
                   "parallelism": "single-gpu"},
        "plain_steps_per_s": round(1.0 / t_plain, 4),
        "batched_nograph_steps_per_s": round(1.0 / t_batched, 4),
        "batched_vs_plain": round(value * t_plain, 3),
        "equal_loss": equal_loss,
        "step_hbm_frac": round(total_bytes / (ms_graph * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
        "mpix_per_s": round(P * value / 1e6, 2),
        "sorted_keys_per_s": round(entries_step / (stage_ms["sort"] * 1e-3), 1),
        "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()},
        "plain_stage_ms": {k: round(v, 4) for k, v in plain_stages.items()},
        "losses_l1_dssim": graph_losses.round(6).tolist(),
        "roofline": roofline,
    }
    if not a.no_cpu_baseline:
        cw, chh = (int(x) for x in a.cpu_crop.split("x"))
        log("timing CPU oracle baseline ...")
        result["cpu_baseline"] = cpu_baseline(scene, specs[0], (cw, chh), V, counts[0])
    else:
        result["cpu_baseline"] = None
    ctx.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
