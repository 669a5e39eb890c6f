#!/usr/bin/env python3
"""Training-step benchmark: train steps/s + sorted keys/s at 1M splats, 1080p,
batched adjoint vs plain, one GPU.

Workload (default): 1,000,000 seeded Gaussians (means U[-1,1]^3, log-scales
N(-4.5,0.3), SH degree 3), 16 cameras on a ring of radius 3 at 1920x1080,
targets rendered from a copy with means jittered by N(0,0.02). One step =
render + L2/L1 loss + adjoint backward + geometry backward for all 16 views,
gradients summed, one Adam step over all 59 parameters of every Gaussian.

value      = steps/s of the batched adjoint path (all views in one pipeline)
plain      = the same step with the views run one after another
keys/s     = intersections per step / time to build the depth-ordered tile
             lists from the projections (emit + sort + ranges + fill)
roofline   = the dominant kernel's algorithmic bytes per launch / its mean
             launch time (hipEvents on the library stream), vs 8 TB/s
cpu_baseline = the CPU oracle on one of the 16 views, scaled to a step

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "skyfall-viewer_amd"))

import numpy as np  # noqa: E402

import sv  # noqa: E402

METRIC = "train steps/s + sorted keys/s at 1M splats 1080p, batched adjoint vs plain, 1 GPU"
HBM_PEAK_GBPS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def configure(ctx: sv.Context, args, keep_grads=0):
    c = ctx.config()
    c.sh_deg = 3
    c.w_l2, c.w_l1, c.w_ssim = 0.8, 0.2, 0.0
    c.lr_pos = c.lr_sh = c.lr_opacity = c.lr_scale = c.lr_rot = 1e-3
    c.beta1, c.beta2, c.eps = 0.9, 0.999, 1e-15
    c.sh_update_interval, c.pos_lr_schedule = 1, 0
    c.batch = args.batch
    c.persistent = args.persistent
    c.keep_grads = keep_grads
    c.xcd_remap = args.xcd_remap
    c.binned_build = args.binned
    c.geom_flags = geom_flags(args)
    ctx.set_config(c)


def geom_flags(args) -> int:
    """The growth run trains with the mean's y in the forward's convention
    unless asked otherwise; every other workload runs the geometry backward
    as the reference's shader writes it."""
    if args.geom_flags is not None:
        return args.geom_flags
    return sv.GEOM_MEAN_Y_FORWARD if args.workload == "growth4m" else 0


def kernel_bytes(phase: str, N: int, V: int, I: int, P: int, key_bytes: int, passes: int) -> float:
    """Algorithmic HBM bytes of one launch of a phase: the per-unit figures of
    the step's byte model (per intersection: a 40 B gather of the id and the
    36 B record in each rasterizer, a 36 B 2-D gradient contribution; per
    pixel: 12 B image + 8 B (alpha, last index) written by the forward, 12 B
    target + 8 B (alpha, last index) read by the backward) x the units the
    launch processes."""
    if phase == "forward":
        return I * 40.0 + P * 20.0
    if phase == "backward":
        return I * 76.0 + P * 20.0
    if phase == "vjp_adam":  # per Gaussian: 59 x (p, m, v) read + write; per (G, view) the 36 B 2-D grad record
        return N * 1416.0 + N * V * 36.0
    if phase == "project":  # per Gaussian: 59 parameters read; per (G, view): 64 B record, 8 B aux, 24 B counts
        return N * 236.0 + N * V * 96.0
    if phase == "sort":
        return passes * I * (key_bytes + 2 * (key_bytes + 4.0))
    return 0.0


def step_b_alg(N: int, V: int, I: int, P: int, d: int) -> float:
    """Algorithmic bytes of one step: 1488 N + 92 N V + I (136 + 24 d) + 40 P."""
    return 1488.0 * N + 92.0 * N * V + I * (136.0 + 24.0 * d) + 40.0 * P


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_1M_16x1080p.json")
SQ_FILE = os.path.join(ROOT, "profiles", "sq_counters_1M_16x1080p.json")
# MI355X: 256 CUs x 4 SIMDs; a wave64 FP32 add / mul / fma occupies its
# 16-lane SIMD's FP32 pipe for 4 cycles (unpacked FP32: 78.6 TFLOP/s of the
# 157.3 packed peak); peak engine clock 2.4 GHz. Other VALU classes (integer,
# conversions, selects, ldexp) can issue beside it.
VALU_ISSUE_PER_S = 256 * 4 * 2.4e9 / 4.0


def measured_sq(phase: str, N: int, V: int, W: int, H: int, workload: str):
    """Shader-sequencer counters of `phase` from the committed rocprofv3
    summary, when it was measured on this workload; else None."""
    if workload != "ring16" or not os.path.exists(SQ_FILE):
        return None
    with open(SQ_FILE) as fh:
        t = json.load(fh)
    if t["workload"] != {"splats": N, "views": V, "width": W, "height": H} or phase not in t["phases"]:
        return None
    return t["phases"][phase]


def measured_traffic(phase: str, N: int, V: int, W: int, H: int, workload: str):
    """HBM bytes per launch of `phase` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary, when it was measured on this workload; else None."""
    if workload != "ring16" or not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as fh:
        t = json.load(fh)
    if t["workload"] != {"splats": N, "views": V, "width": W, "height": H} or phase not in t["phases"]:
        return None
    k = t["phases"][phase]
    return k["fetch_bytes"] + k["write_bytes"]


def run_mode(mode, g, cams, targets, args, graph=0):
    """Returns (s/step of the timed loop, per-phase timings, phase steps,
    stats of one more step, s/step of eager launches or None)."""
    ctx = sv.Context(capacity=g["pos"].shape[0])
    configure(ctx, args)
    ctx.set_config(use_graph=graph)
    ctx.set_gaussians(g)
    ctx.set_views(cams)
    for v, t in enumerate(targets):
        ctx.set_target(v, t)
    views = list(range(len(cams)))
    tune = None
    if mode == sv.MODE_BATCHED and args.autotune:
        tune = ctx.autotune(views, 3)
        best = min(tune, key=lambda e: e["ms_per_step"])
        log("autotune: " + ", ".join(f"t{e['tile']}/b{e['batch']}/p{e['persistent']}/x{e['xcd_remap']}"
                                      f"/s{e['split_backward']} {e['ms_per_step']:.2f}" for e in tune) + f" -> {best}")
    for i in range(args.warmup):
        ctx.step_views(views, mode)
    ctx.sync()
    t0 = time.perf_counter()
    for i in range(args.steps):
        ctx.step_views(views, mode)
    ctx.sync()
    t1 = time.perf_counter()
    # per-phase kernel time: further steps with HIP events around each phase
    # on the library stream (these launch eagerly)
    phase_steps = max(1, min(args.steps, 5))
    ctx.reset_timings()
    ctx.enable_timing(True)
    for i in range(phase_steps):
        ctx.step_views(views, mode)
    timings = ctx.timings()
    ctx.enable_timing(False)
    eager = None
    if graph and not args.only_batched:
        ctx.set_config(use_graph=0)
        ctx.step_views(views, mode)
        ctx.sync()
        e0 = time.perf_counter()
        for i in range(args.steps):
            ctx.step_views(views, mode)
        ctx.sync()
        eager = (time.perf_counter() - e0) / args.steps
    st = ctx.step_views(views, mode, stats=True)
    cfg = ctx.config()
    chosen = {"batch": cfg.batch, "persistent": cfg.persistent, "xcd_remap": cfg.xcd_remap, "raster_px": cfg.raster_px,
              "split_backward": cfg.split_backward, "tile_order": cfg.tile_order}
    ctx.close()
    return (t1 - t0) / args.steps, timings, phase_steps, st, eager, tune, chosen


def cpu_threads():
    """Threads of the CPU baseline: OMP_NUM_THREADS when set (the host's
    share of cores for this GPU), else every core this process may run on."""
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return env or avail, avail


def cpu_baseline(g, cams, targets, args):
    """The oracle (C, OpenMP; parallel LSD radix sort) on one of the step's
    views, scaled to a step: steps/s, and keys/s of sorting that view's
    emitted (tile, depth) pairs."""
    threads, avail = cpu_threads()
    os.environ["OMP_NUM_THREADS"] = str(threads)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    orc.lib()
    w, h = int(cams[0].viewport[0]), int(cams[0].viewport[1])
    view = orc.view_params(w, h, sh_deg=3)
    t0 = time.perf_counter()
    r = orc.render_and_grad(g, cams[0], view, target=targets[0], w_l2=0.8, w_l1=0.2)
    t_view = time.perf_counter() - t0
    t0 = time.perf_counter()
    bc1, bc2 = orc.adam_bias(1, 0.9, 0.999)
    for k in ("pos", "logscale", "quat", "logit", "sh"):
        p = np.array(g[k], np.float32, copy=True)
        orc.adam(p, np.zeros_like(p), np.zeros_like(p), r["grads"][k], 1e-3, 0.9, 0.999, 1e-15, bc1, bc2)
    t_adam = time.perf_counter() - t0
    # keys/s: the emitted pairs of this view, unsorted, through the radix sort
    keys, vals = orc.emit(r["proj"], view)
    t0 = time.perf_counter()
    orc.sort_pairs_radix(keys, vals)
    t_sort = time.perf_counter() - t0
    V = len(cams)
    return {
        "value": 1.0 / (V * t_view + t_adam),
        "unit": "steps/s",
        "cores": threads,
        "kind": "port",
        "keys_per_s": keys.size / t_sort if t_sort > 0 else None,
        "host_cores_available": avail,
        "sample": f"CPU oracle (C, OpenMP, {threads} threads) render+loss+backward+geometry backward of 1 of the "
                  f"{V} views ({g['pos'].shape[0]} splats, {w}x{h}, {r['n_isect']} intersections, {t_view:.2f} s) "
                  f"+ Adam over all parameters ({t_adam:.2f} s), scaled to {V} views; keys/s = that view's "
                  f"{keys.size} (tile, depth) pairs through the parallel LSD radix sort ({t_sort:.3f} s)",
    }


def run_growth4m(args):
    """The growth run: 1M splats, 16 ring views at 1920x1080, batched steps;
    every 100 steps prune sigmoid(logit) < 0.005 and densify (split above 8
    px where a tile's L1 error is above its view's 90th percentile, else
    clone). With --growth-cap 1 (default) an event adds at most
    max(1, floor(0.1 N)) splats, the best-scored candidates first (the
    reference's cap, gaussian-renderer.ts:3297-3306), and never more than
    the 4M budget; with 0 every candidate is taken up to the budget (in
    index order). With --growth-params reference (default) the run trains
    with the reference trainer's defaults: learning rates means 1.6e-4
    (decayed by the position schedule), SH 2.5e-3 every 16 steps, opacity
    0.025, scale 0.005, rotation 1e-3 (gaussian-renderer.ts:86-96,
    284-298, 2476-2484) and loss 0.8 L1 + 0.2 D-SSIM (:506-522); with
    "unit" it keeps the parity configs' 1e-3 rates, sh interval 1 and
    loss (0.8 L2, 0.2 L1), under which added splats are never trained
    away and the image error grows. The timed region holds every step and
    every prune / densify event. An event densifies, then prunes, as the
    reference's runDensifyAndPruneCombined does; --densify-rule tile_l1
    (default) decides with the tile-error rule of the 300-step
    configuration (the scene grows to the 4M budget), gradient with the
    trainer's rule (mean screen-space gradients since the last event,
    multi-view importance, DEFAULT_DENSIFICATION_CONFIG thresholds; the
    scene grows to about 1.4M). --geom-flags 1 (the default here) gives the
    geometry backward's mean path the forward's y convention: with the
    shader's +f_y (:264-266, --geom-flags 0) every position-y update moves
    against the rendered image's dependence, so the positions drift and the
    image error grows over the run."""
    N0, V, W, H, NMAX, EVERY = 1_000_000, 16, 1920, 1080, 4_000_000, 100
    g = sv.scene_gaussians(0, N0, 1.0)
    cams = sv.ring_cameras(V, 3.0, W, H)
    tg = dict(g)
    tg["pos"] = sv.scene_jitter(1, g["pos"], 0.02)
    c = sv.Context(capacity=NMAX)
    configure(c, args)
    c.set_config(use_graph=args.graph)
    if args.growth_params == "reference":
        c.set_config(lr_pos=0.00016, lr_sh=0.0025, lr_opacity=0.025, lr_scale=0.005, lr_rot=0.001,
                     sh_update_interval=16, pos_lr_schedule=1, w_l2=0.0, w_l1=0.8, w_ssim=0.2)
    c.set_gaussians(tg)
    c.set_views(cams)
    targets = []
    for v in range(V):
        c.forward(v)
        targets.append(c.image(v))
    c.set_gaussians(g)
    for v in range(V):
        c.set_target(v, targets[v])
    views = list(range(V))
    # the trainer's densify rule: its default thresholds (scene extent 2 for
    # the [-1, 1]^3 scene, pixel scale = width / 2), importance from 10 of
    # the views (MULTI_VIEW_SAMPLE_COUNT; every other view here)
    rule = sv.densify_rule(2.0, W)
    sample = views[::2][:10]
    log(f"growth4m: {N0} splats, {V} views {W}x{H}, {args.warmup} warmup + {args.steps} steps, "
        f"densify rule {args.densify_rule}")
    for _ in range(args.warmup):
        c.step_views(views, sv.MODE_BATCHED)
    c.sync()
    events, isects, st = [], 0, None
    t0 = time.perf_counter()
    for it in range(1, args.steps + 1):
        st = c.step_views(views, sv.MODE_BATCHED, stats=True)
        isects += int(st.n_isect)
        if it % EVERY == 0:
            # runDensifyAndPruneCombined (gaussian-renderer.ts:3468-3488):
            # densify, then prune
            l1 = float(st.l1_sum) / (3.0 * V * W * H)
            budget = NMAX - c.count
            cap = min(max(1, c.count // 10), budget) if args.growth_cap else 0
            max_new = max(cap, 0) if args.growth_cap else 0
            if args.densify_rule == "gradient":
                imp, _, ns_views = c.multiview_scores(sample)
                act, nc, ns = c.densify_decide_grad(rule, imp, ns_views, max_new=max_new)
            else:
                act, nc, ns = c.densify_decide(views, 0.9, 8.0, max_new=max_new)
            if args.growth_cap and cap <= 0:
                act[:] = 0
            if nc + ns > budget:
                idx = np.flatnonzero(act)
                act[idx[max(budget, 0):]] = 0
            added = c.densify(act, seed=it) if act.any() else 0
            removed = c.prune_opacity(0.005)
            events.append({"step": it, "mean_l1": l1, "added": added, "pruned": removed, "count": c.count})
            log(f"growth4m step {it}: mean L1 {l1:.5f}, added {added}, pruned {removed}, {c.count} splats")
    c.sync()
    dt = time.perf_counter() - t0
    P = V * W * H
    # per kernel at the final size (after the timed run): three more steps
    # with the phase timers (eager launches), their time and algorithmic HBM
    # rate from the byte model, and the tile-list key rate
    c.reset_timings()
    c.enable_timing(True)
    for _ in range(3):
        st_k = c.step_views(views, sv.MODE_BATCHED, stats=True)
    tim = c.timings()
    c.enable_timing(False)
    I_k, N_k = int(st_k.n_isect), c.count
    kb = 8 if st_k.key_bits > 32 else 4
    per_kernel = {}
    for k, (ms, per_step) in phase_table(tim, 3).items():
        b = kernel_bytes(k, N_k, V, I_k, P, kb, (int(st_k.key_bits) + 7) // 8) if k != "sort" else 0.0
        per_kernel[k] = {"ms_per_launch": ms, "launches_per_step": per_step,
                         "algorithmic_GBps": b / (ms * 1e-3) / 1e9 if b and ms > 0 else None}
    list_ms, list_kps = list_build(tim, 3, I_k)
    log("growth4m final size: " + ", ".join(f"{k} {v['ms_per_launch']:.3f} ms" +
                                             (f" {v['algorithmic_GBps']:.0f} GB/s" if v["algorithmic_GBps"] else "")
                                             for k, v in per_kernel.items()))
    out = {
        "metric": "train steps/s, growth run (1M to 4M splats, 16 views at 1080p)", "value": args.steps / dt, "unit": "steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * dt / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded scene generator; targets rendered from a jittered copy)",
        "config": {"workload": f"{N0} splats growing to <= {NMAX}, {V} views {W}x{H}, prune/densify every "
                               f"{EVERY} steps", "splats_start": N0, "splats_end": c.count, "views": V,
                   "width": W, "height": H, "growth_cap": "max(1, floor(0.1 N)) per event, best score first"
                   if args.growth_cap else "none (4M budget only)",
                   "densify_rule": "trainer's gradient rule (gaussian-renderer.ts:3249-3306), default thresholds"
                   if args.densify_rule == "gradient" else "tile L1 above the 90th percentile, split above 8 px",
                   "geometry_backward": "mean y as the forward (SV_GEOM_MEAN_Y_FORWARD)"
                   if geom_flags(args) & sv.GEOM_MEAN_Y_FORWARD else "the shader as written (+f_y mean path)",
                   "training": "reference defaults: lr means 1.6e-4 scheduled, SH 2.5e-3 every 16 steps, opacity "
                               "0.025, scale 0.005, rotation 1e-3; loss 0.8 L1 + 0.2 D-SSIM"
                   if args.growth_params == "reference" else
                   "unit: lr 1e-3 every group, SH every step; loss 0.8 L2 + 0.2 L1"},
        "intersections_per_s": isects / dt, "final_mean_l1": float(st.l1_sum) / (3.0 * P) if st else None,
        "events": events,
        "final_size": {"splats": N_k, "intersections": I_k, "keys_per_s": list_kps, "tile_list_ms": list_ms,
                       "per_kernel": per_kernel,
                       "note": "three steps after the timed run with the phase timers (eager launches); "
                               "algorithmic_GBps = the byte model's bytes of the launch / its time"},
    }
    c.close()
    print(json.dumps(out), flush=True)


def run_adjoint2m(args):
    """The memory / recompute comparison at 2,000,000 splats and one 1920x1080
    view (camera at (0,0,3), 60 degree FOV): the same training step with the
    adjoint backward, which rebuilds transmittance by reverse traversal from
    the stored final alpha and last index (8 B per pixel kept from the
    forward; 36 B adjoint records per reached pair written during the step),
    against the reference's method, which stores 54 fixed-point gradient
    slots per intersection (216 B, backward-tiles-local.wgsl:44-53, allocated
    at gaussian-renderer.ts:1281-1349) and reduces them. For each: steps/s
    and the device memory the context holds (peak, sv_memory_stats). A stored
    per-pixel contributor list (every visited (pixel, splat) pair's index and
    alpha, 8 B) is not built; its size is counted from the forward's last
    indices and tile ranges."""
    N, W, H = 2_000_000, 1920, 1080
    g = sv.scene_gaussians(0, N, 1.0)
    cams = [sv.scene_camera([0.0, 0.0, 3.0], [0.0, 0.0, 0.0], W, H, 60.0)]
    targets = render_targets(g, cams, args)
    res = {}
    for name, mode, graph in (("adjoint_batched", sv.MODE_BATCHED, args.graph),
                              ("adjoint_sequential", sv.MODE_SEQUENTIAL, 0),
                              ("reference_stored_tile_grads", sv.MODE_PLAIN, 0)):
        c = sv.Context(capacity=N)
        configure(c, args)
        c.set_config(use_graph=graph)
        c.set_gaussians(g)
        c.set_views(cams)
        c.set_target(0, targets[0])
        base = c.memory_stats()[0]
        for _ in range(args.warmup):
            c.step_views([0], mode)
        c.sync()
        c.reset_memory_peak()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            st = c.step_views([0], mode, stats=True)
        c.sync()
        dt = (time.perf_counter() - t0) / args.steps
        now, peak = c.memory_stats()
        res[name] = {"ms_per_step": 1e3 * dt, "steps_per_s": 1.0 / dt, "device_bytes_peak": peak,
                     "device_bytes_before_first_step": base, "intersections": int(st.n_isect)}
        if mode == sv.MODE_SEQUENTIAL:
            last = c.last_ids(0).ravel().astype(np.int64)
            offs = c.tile_offsets(0).astype(np.int64)
            tx = (W + 15) // 16
            yy, xx = np.divmod(np.arange(W * H), W)
            start = offs[(yy // 16) * tx + xx // 16]
            ok = last != 0xFFFFFFFF
            visited = int(np.sum(last[ok] - start[ok] + 1))
            res["stored_pixel_lists"] = {"built": False, "visited_pairs": visited,
                                         "bytes": 8 * visited, "note": "index (4 B) + alpha (4 B) per visited "
                                         "(pixel, splat) pair, counted from the forward's last indices"}
        c.close()
        log(f"adjoint2m {name}: {1e3 * dt:.2f} ms/step, peak {peak / 2**30:.2f} GiB")
    I = res["adjoint_sequential"]["intersections"]
    out = {
        "metric": "train step time and device memory at 2M splats 1080p: reverse-traversal adjoint vs stored "
                  "per-intersection gradients",
        "value": res["adjoint_batched"]["steps_per_s"], "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": res["adjoint_batched"]["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded scene generator; target rendered from a jittered copy)",
        "config": {"workload": f"{N} splats, 1 view {W}x{H} from (0,0,3)", "splats": N, "views": 1, "width": W,
                   "height": H},
        "methods": res,
        "per_unit_bytes": {"reverse_traversal_saved_per_pixel": 8, "adjoint_record_per_reached_pair": 36,
                           "stored_tile_grads_per_intersection": 216, "stored_pixel_list_per_visited_pair": 8},
        "stored_tile_grads_bytes": 216 * I,
    }
    print(json.dumps(out), flush=True)


def visited_pairs(c, v: int, w: int, h: int) -> int:
    """(pixel, splat) pairs the compositing of view v visits, from the last
    forward's per-pixel last indices and the view's tile offsets: each pixel
    walks its tile's list from the first entry to its last index."""
    last = c.last_ids(v).ravel().astype(np.int64)
    offs = c.tile_offsets(v).astype(np.int64)
    tx = (w + 15) // 16
    yy, xx = np.divmod(np.arange(w * h), w)
    start = offs[(yy // 16) * tx + xx // 16]
    ok = last != 0xFFFFFFFF
    return int(np.sum(last[ok] - start[ok] + 1))


def pairs_and_single_view(g, cams, targets, args, fwd_ms: float, bwd_ms: float):
    """The secondary FLOP figure and the single-view plain point: pairs E =
    visited (pixel, splat) pairs of a step (every view rendered on its own,
    counted as visited_pairs), forward about 16 FLOP + 1 exp per pair and
    backward about 45 + 1, over the batched launches' mean times, against
    MI355X's 157.3 TFLOP/s FP32 vector peak (half of it unpacked); and the
    reference method on view 0 alone."""
    c = sv.Context(capacity=g["pos"].shape[0])
    configure(c, args)
    c.set_gaussians(g)
    c.set_views(cams)
    for v, t in enumerate(targets):
        c.set_target(v, t)
    E = 0
    for v, cam in enumerate(cams):
        c.forward(v)
        w, h = int(round(cam.viewport[0])), int(round(cam.viewport[1]))
        E += visited_pairs(c, v, w, h)
    for _ in range(2):
        c.step_views([0], sv.MODE_PLAIN)
    c.sync()
    t0 = time.perf_counter()
    n = 5
    for _ in range(n):
        c.step_views([0], sv.MODE_PLAIN)
    c.sync()
    plain1 = (time.perf_counter() - t0) / n
    c.close()
    peak = 157.3e12
    fw = 17.0 * E / (fwd_ms * 1e-3) if fwd_ms else None
    bw = 46.0 * E / (bwd_ms * 1e-3) if bwd_ms else None
    return {"visited_pairs_per_step": E,
            "forward_flop_per_s": fw, "backward_flop_per_s": bw,
            "forward_frac_of_fp32_peak": fw / peak if fw else None,
            "backward_frac_of_fp32_peak": bw / peak if bw else None,
            "model": "forward 16 FLOP + 1 exp, backward 45 FLOP + 1 exp per visited pair (the exp counted as 1); "
                     "peak 157.3 TFLOP/s FP32 vector (packed; the kernels issue unpacked FP32)"}, {
        "ms_per_step": 1e3 * plain1, "steps_per_s": 1.0 / plain1,
        "workload": "view 0 alone, the reference's method (wavefront fixed-point backward, reduce, geometry "
                    "backward, Adam)"}


def make_workload(name: str, args):
    """(splats, cameras, sizes, width, height, description) of a workload."""
    N, V, W, H = args.n, args.views, args.width, args.height
    if name == "small256":
        N, V, W, H = 16384, 1, 256, 256
    elif name == "ring8_720p":
        N, V, W, H = 500_000, 8, 1280, 720
    if name == "uneven64":
        V = 64
        g = sv.scene_gaussians(0, N, 2.0)
        cams, sizes = sv.sphere_cameras(2, V, 4.0)
        desc = (f"{N} splats in [-2,2]^3, {V} views of seeded sizes 320-1920 x 240-1080 "
                f"({sum(w * h for w, h in sizes)} pixels) per step, batched adjoint training step")
        return g, cams, sizes, None, None, desc
    g = sv.scene_gaussians(0, N, 1.0)
    if name == "small256":
        cams = [sv.scene_camera([0.0, 0.0, 3.0], [0.0, 0.0, 0.0], W, H, 60.0)]
        desc = f"{N} splats, 1 view {W}x{H} from (0,0,3), training step"
    else:
        cams = sv.ring_cameras(V, 3.0, W, H)
        desc = f"{N} splats, {V} views {W}x{H} per step, batched adjoint training step"
    return g, cams, [(W, H)] * V, W, H, desc


def render_targets(g, cams, args):
    tg = dict(g)
    tg["pos"] = sv.scene_jitter(1, g["pos"], 0.02)
    targets = []
    with sv.Context(capacity=g["pos"].shape[0]) as tc:
        configure(tc, args)
        tc.set_gaussians(tg)
        tc.set_views(cams)
        for v in range(len(cams)):
            tc.forward(v)
            targets.append(tc.image(v))
    return targets


def phase_table(tim, steps):
    """{phase: (ms per launch, launches per step)}."""
    return {k: (v[0] / max(v[1], 1), v[1] / steps) for k, v in tim.items() if v[1]}


def list_build(tim, steps, I):
    """(ms per step, keys/s) of building the tile lists from the projections
    (emit + sort + ranges [+ fill])."""
    ms = sum(tim.get(k, (0.0, 0))[0] for k in ("emit", "sort", "ranges", "fill")) / steps
    return ms, (I / (ms * 1e-3) if ms > 0 else None)


def secondary(name, args):
    """Batched and plain steps/s, keys/s and the CPU oracle of a smaller
    workload (reported beside the headline)."""
    g, cams, sizes, W, H, desc = make_workload(name, args)
    targets = render_targets(g, cams, args)
    sub = argparse.Namespace(**vars(args))
    sub.steps, sub.warmup, sub.autotune = max(3, min(args.steps, 10)), 2, 0
    sec_b, tim_b, ps_b, st_b, *_ = run_mode(sv.MODE_BATCHED, g, cams, targets, sub, graph=args.graph)
    sec_p, *_ = run_mode(sv.MODE_PLAIN, g, cams, targets, sub)
    ms_list, kps = list_build(tim_b, ps_b, int(st_b.n_isect))
    out = {"workload": desc, "batched_steps_per_s": 1.0 / sec_b, "plain_steps_per_s": 1.0 / sec_p,
           "keys_per_s": kps, "intersections_per_step": int(st_b.n_isect)}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(g, cams, targets, args)
    log(f"{name}: batched {1e3 * sec_b:.3f} ms/step, plain {1e3 * sec_p:.3f} ms/step")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=("ring16", "uneven64", "small256", "ring8_720p", "growth4m", "adjoint2m"),
                    default="ring16",
                    help="ring16: 16 ring views at --width x --height (the headline); uneven64: 64 views "
                         "with seeded sizes 320-1920 x 240-1080 on a sphere of radius 4, means in [-2,2]^3; "
                         "small256: 16,384 splats, one 256x256 view from (0,0,3); ring8_720p: 500,000 splats, 8 ring "
                         "views at 1280x720; growth4m: the growth run (1M splats growing to at most 4M, 16 ring "
                         "views at 1080p, prune / densify every 100 steps; give --steps, e.g. 2000); adjoint2m: "
                         "2M splats, one 1080p view: step time and device memory of the reverse-traversal adjoint "
                         "against the reference's stored per-intersection gradients")
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--views", type=int, default=16)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--persistent", type=int, default=0)
    ap.add_argument("--graph", type=int, default=1, help="1: the batched step is one hipGraph launch")
    ap.add_argument("--xcd-remap", type=int, default=1, help="1: contiguous tile ranges per XCD")
    ap.add_argument("--binned", type=int, default=1,
                    help="1: tile lists by supertile binning; 0: sort every (tile, splat) pair")
    ap.add_argument("--autotune", type=int, default=1,
                    help="1: time the tuner's candidates first and run the batched step with the fastest")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--only-batched", action="store_true",
                    help="run the batched step alone (no eager, sequential, plain, pair-sort, CPU or other "
                         "workload legs): the target of the profiler runs")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--growth-cap", type=int, default=1,
                    help="growth4m: 1 = at most max(1, N // 10) new splats per event, best score first")
    ap.add_argument("--growth-params", choices=("reference", "unit"), default="reference",
                    help="growth4m: the reference trainer's default learning rates, schedules and loss, or the "
                         "parity configs' 1e-3 rates")
    ap.add_argument("--densify-rule", choices=("gradient", "tile_l1"), default="tile_l1",
                    help="growth4m: the trainer's gradient rule or the tile-error rule")
    ap.add_argument("--geom-flags", type=int, default=None,
                    help="SV_GEOM_* flags of the geometry backward (default: 1 for growth4m, 0 otherwise)")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the small256 and ring8_720p measurements reported beside the headline")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy (torch) three times, to scale the "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.only_batched:
        args.no_plain = args.no_cpu_baseline = args.no_configs = True
    if args.calibrate_copy:
        import torch

        x = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_dispatch": x.numel() * 4}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported "
              "(multi-GPU sharding by views is a design only)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "growth4m":
        run_growth4m(args)
        return
    if args.workload == "adjoint2m":
        run_adjoint2m(args)
        return
    g, cams, sizes, W, H, workload = make_workload(args.workload, args)
    N, V = g["pos"].shape[0], len(cams)
    P = sum(w * h for w, h in sizes)
    log(f"scene: {workload}")
    targets = render_targets(g, cams, args)
    log("targets rendered")

    sec_b, tim_b, ps_b, st_b, eager_b, tune_b, chosen_b = run_mode(sv.MODE_BATCHED, g, cams, targets, args,
                                                                   graph=args.graph)
    log(f"batched: {1e3 * sec_b:.2f} ms/step (graph {args.graph}), I={st_b.n_isect}, "
        f"key bits {st_b.key_bits}, passes {st_b.sort_passes}, binned {st_b.binned}"
        + (f"; eager launches {1e3 * eager_b:.2f} ms/step" if eager_b else ""))
    plain = sequential = None
    if not args.no_plain:
        sec_q, tim_q, ps_q, *_ = run_mode(sv.MODE_SEQUENTIAL, g, cams, targets, args)
        sequential = {"ms_per_step": 1e3 * sec_q, "steps_per_s": 1.0 / sec_q,
                      "phase_ms_per_step": {k: v[0] / ps_q for k, v in tim_q.items() if v[1]}}
        log(f"sequential adjoint: {1e3 * sec_q:.2f} ms/step")
        sec_p, tim_p, ps_p, *_ = run_mode(sv.MODE_PLAIN, g, cams, targets, args)
        plain = {"ms_per_step": 1e3 * sec_p, "steps_per_s": 1.0 / sec_p,
                 "method": "reference method per view: wavefront fixed-point backward, per-Gaussian reduce, "
                           "geometry backward, Adam",
                 "phase_ms_per_step": {k: v[0] / ps_p for k, v in tim_p.items() if v[1]}}
        log(f"plain: {1e3 * sec_p:.2f} ms/step")
    # the same batched step with every (tile, splat) pair sorted: sorted keys/s of the pair sort
    pair_sort = None
    if args.binned and not args.only_batched:
        ps_args = argparse.Namespace(**vars(args))
        ps_args.binned, ps_args.autotune, ps_args.steps, ps_args.warmup = 0, 0, 3, 1
        _, tim_s, ps_s, st_s, *_ = run_mode(sv.MODE_BATCHED, g, cams, targets, ps_args, graph=args.graph)
        ms_s, kps_s = list_build(tim_s, ps_s, int(st_s.n_isect))
        sort_ms = tim_s.get("sort", (0.0, 0))[0] / ps_s
        pair_sort = {"keys_per_s": kps_s, "tile_list_ms_per_step": ms_s, "sort_ms_per_step": sort_ms,
                     "sort_only_keys_per_s": int(st_s.n_isect) / (sort_ms * 1e-3) if sort_ms > 0 else None,
                     "key_bits": int(st_s.key_bits), "sort_passes": int(st_s.sort_passes)}
        log(f"pair sort: {ms_s:.3f} ms/step list build, {sort_ms:.3f} ms sort, {st_s.key_bits}-bit keys")

    I = int(st_b.n_isect)
    key_bytes = 8 if st_b.key_bits > 32 else 4
    passes = int(st_b.sort_passes)
    pair_passes = (int(st_b.key_bits) + 7) // 8  # passes of a sort of every (tile, depth) key
    phase = phase_table(tim_b, ps_b)
    list_ms_step, keys_per_s = list_build(tim_b, ps_b, I)
    # dominant single-kernel phase
    cands = {k: phase[k][0] for k in ("forward", "backward", "vjp_adam") if k in phase}
    dom = max(cands, key=cands.get)
    dom_ms = cands[dom]
    dom_bytes = kernel_bytes(dom, N, V, I, P, key_bytes, pair_passes)
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    b_alg = step_b_alg(N, V, I, P, pair_passes)
    traffic = measured_traffic(dom, N, V, W, H, args.workload)
    sq = measured_sq(dom, N, V, W, H, args.workload)
    valu = None
    if sq:
        valu = {"insts_per_launch": sq["waves"] * sq["valu_per_wave"],
                "all_classes_per_cu_cycle": sq.get("valu_per_cu_cycle"), "lane_util": sq.get("valu_lane_util"),
                "active_frac": sq["valu_active_frac"], "waiting_frac": sq["waiting_frac"]}
        if "fp32_per_wave" in sq:
            f32 = sq["waves"] * sq["fp32_per_wave"]
            valu.update({"fp32_insts_per_launch": f32,
                         "issue_frac": f32 / (dom_ms * 1e-3 * VALU_ISSUE_PER_S),
                         "model": "FP32 add + mul + fma wave-instructions (SQ_INSTS_VALU_{ADD,MUL,FMA}_F32) / "
                                  "(launch time x 256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles): the share of the "
                                  "FP32 pipe's issue rate"})
    # the bound the counters show: the resource closer to its peak
    hbm_frac = traffic / (dom_ms * 1e-3) / 1e9 / HBM_PEAK_GBPS if traffic else None
    bound = "hbm"
    if valu and "issue_frac" in valu and (hbm_frac is None or valu["issue_frac"] > hbm_frac):
        bound = "valu"
    limiter = None
    if traffic or sq:
        parts = []
        if valu and "issue_frac" in valu:
            parts.append(f"FP32 pipe issue {valu['issue_frac']:.2f} of its peak rate (all VALU classes "
                         f"{valu['all_classes_per_cu_cycle']:.2f} wave-instructions per CU-cycle, {valu['lane_util']:.2f} "
                         f"of lanes active), waiting {valu['waiting_frac']:.2f} of wave-cycles")
        if traffic:
            parts.append(f"measured HBM {traffic / (dom_ms * 1e-3) / 1e9:.0f} GB/s = {hbm_frac:.2f} of peak")
        limiter = "; ".join(parts)
    out = {
        "metric": METRIC,
        "value": 1.0 / sec_b,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * sec_b,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scene generator; targets rendered from a jittered copy)",
        "config": {"workload": workload, "splats": N, "views": V, "width": W, "height": H, "pixels": P,
                   "tile": 16, "raster_tile": chosen_b["raster_px"], "lds_batch": chosen_b["batch"],
                   "persistent": chosen_b["persistent"], "graph": args.graph, "xcd_remap": chosen_b["xcd_remap"],
                   "split_backward": chosen_b["split_backward"], "tile_order": chosen_b["tile_order"],
                   "binned_build": args.binned},
        "autotune": tune_b,
        "eager_launch_ms_per_step": 1e3 * eager_b if eager_b else None,
        "keys_per_s": keys_per_s,
        "tile_list_ms_per_step": list_ms_step,
        "intersections_per_step": I,
        "tile_lists": "binned (supertile pairs sorted, lists filled)" if st_b.binned else "pair sort",
        "sort_key_bits": int(st_b.key_bits),
        "sort_passes": passes,
        "pair_sort": pair_sort,
        "plain": plain,
        "speedup_vs_plain": (plain["ms_per_step"] / (1e3 * sec_b)) if plain else None,
        "sequential_adjoint": sequential,
        "speedup_vs_sequential": (sequential["ms_per_step"] / (1e3 * sec_b)) if sequential else None,
        "phase_ms_per_launch": {k: v[0] for k, v in phase.items()},
        "step_b_alg_bytes": b_alg,
        "step_b_alg_GBps": b_alg / sec_b / 1e9,
        "step_b_alg_frac": b_alg / sec_b / 1e9 / HBM_PEAK_GBPS,
        "roofline": {"bound": bound, "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBPS, "traffic": traffic,
                     "frac_basis": "algorithmic HBM bytes per launch / mean launch time, of 8 TB/s; the bound "
                                   "is the resource the counters put nearest its peak",
                     "traffic_source": "profiles/hbm_traffic_1M_16x1080p.md (rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                                       "copy-calibrated)" if traffic else None,
                     "algorithmic_bytes_per_launch": dom_bytes, "mean_launch_ms": dom_ms,
                     "limiter": limiter, "valu": valu},
        "cpu_baseline": None,
    }
    if args.workload == "ring16" and not args.only_batched:
        pairs, single = pairs_and_single_view(g, cams, targets, args, phase.get("forward", (0,))[0],
                                              phase.get("backward", (0,))[0])
        out["pairs"] = pairs
        out["plain_single_view"] = single
        log(f"visited pairs {pairs['visited_pairs_per_step']}, forward {pairs['forward_flop_per_s'] / 1e12:.1f} "
            f"TFLOP/s, backward {pairs['backward_flop_per_s'] / 1e12:.1f} TFLOP/s; plain single view "
            f"{single['ms_per_step']:.2f} ms/step")
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle, one view)...")
        out["cpu_baseline"] = cpu_baseline(g, cams, targets, args)
    if args.workload == "ring16" and not args.no_configs:
        out["other_workloads"] = {name: secondary(name, args) for name in ("small256", "ring8_720p")}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
