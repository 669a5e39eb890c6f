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
    ctx.set_config(c)


def kernel_bytes(phase: str, N: int, V: int, I: int, P: int, key_bytes: int, passes: int) -> float:
    """Algorithmic HBM bytes of one launch of a phase (per-unit figures x units)."""
    if phase == "forward":  # per intersection: id 4 + record 36; per pixel: rgb 12 + A 4 + last 4
        return I * 40.0 + P * 20.0
    if phase == "backward":  # per isect: id 4 + record 36 + rect/offset 12 + adjoint record 36; per pixel 32
        return I * 88.0 + P * 32.0
    if phase == "vjp_adam":  # per Gaussian: 59 x (p, m, v) read + write + stats; per (G, view) 24; per isect 36
        return N * 1480.0 + N * V * 24.0 + I * 36.0
    if phase == "sort":
        return passes * I * (key_bytes + 2 * (key_bytes + 4.0))
    return 0.0


def step_b_alg(N: int, V: int, I: int, P: int, d: int) -> float:
    """Algorithmic bytes of one step: 1488 N + 92 N V + I (136 + 24 d) + 40 P."""
    return 1488.0 * N + 92.0 * N * V + I * (136.0 + 24.0 * d) + 40.0 * P


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_1M_16x1080p.json")


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
        log("autotune: " + ", ".join(f"b{e['batch']}/p{e['persistent']}/x{e['xcd_remap']} {e['ms_per_step']:.2f}"
                                      for e in tune) + f" -> {best}")
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
    if graph:
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
    chosen = {"batch": cfg.batch, "persistent": cfg.persistent, "xcd_remap": cfg.xcd_remap}
    ctx.close()
    return (t1 - t0) / args.steps, timings, phase_steps, st, eager, tune, chosen


def cpu_baseline(g, cams, targets, args):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(threads)
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
    V = len(cams)
    return {
        "value": 1.0 / (V * t_view + t_adam),
        "unit": "steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"CPU oracle (C, OpenMP) render+loss+backward+geometry backward of 1 of the {V} views "
                  f"({g['pos'].shape[0]} splats, {w}x{h}, {r['n_isect']} intersections, {t_view:.2f} s) "
                  f"+ Adam over all parameters ({t_adam:.2f} s), scaled to {V} views",
    }


def run_growth4m(args):
    """The growth run: 1M splats, 16 ring views at 1920x1080, batched steps;
    every 100 steps prune sigmoid(logit) < 0.005 and densify (split above 8
    px where a tile's L1 error is above its view's 90th percentile, else
    clone), the added splats capped so the count stays <= 4M (actions past
    the budget are dropped in index order). The timed region holds every
    step and every prune / densify event."""
    N0, V, W, H, NMAX, EVERY = 1_000_000, 16, 1920, 1080, 4_000_000, 100
    g = sv.scene_gaussians(0, N0, 1.0)
    cams = sv.ring_cameras(V, 3.0, W, H)
    tg = dict(g)
    tg["pos"] = sv.scene_jitter(1, g["pos"], 0.02)
    c = sv.Context(capacity=NMAX)
    configure(c, args)
    c.set_config(use_graph=args.graph)
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
    log(f"growth4m: {N0} splats, {V} views {W}x{H}, {args.warmup} warmup + {args.steps} steps")
    for _ in range(args.warmup):
        c.step_views(views, sv.MODE_BATCHED)
    c.sync()
    events, isects, st = [], 0, None
    t0 = time.perf_counter()
    for it in range(1, args.steps + 1):
        st = c.step_views(views, sv.MODE_BATCHED, stats=True)
        isects += int(st.n_isect)
        if it % EVERY == 0:
            removed = c.prune_opacity(0.005)
            act, nc, ns = c.densify_decide(views, 0.9, 8.0)
            budget = NMAX - c.count
            if nc + ns > budget:
                idx = np.flatnonzero(act)
                act[idx[max(budget, 0):]] = 0
            added = c.densify(act, seed=it)
            events.append({"step": it, "pruned": removed, "added": added, "count": c.count})
            log(f"growth4m step {it}: pruned {removed}, added {added}, {c.count} splats, "
                f"mean L1 {float(st.l1_sum) / (3.0 * V * W * H):.5f}")
    c.sync()
    dt = time.perf_counter() - t0
    P = V * W * H
    out = {
        "metric": "train steps/s, growth run (1M to 4M splats, 16 views at 1080p)", "value": args.steps / dt, "unit": "steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * dt / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded scene generator; targets rendered from a jittered copy)",
        "config": {"workload": f"{N0} splats growing to <= {NMAX}, {V} views {W}x{H}, prune/densify every "
                               f"{EVERY} steps", "splats_start": N0, "splats_end": c.count, "views": V,
                   "width": W, "height": H},
        "intersections_per_s": isects / dt, "final_mean_l1": float(st.l1_sum) / (3.0 * P) if st else None,
        "events": events,
    }
    c.close()
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=("ring16", "uneven64", "small256", "ring8_720p", "growth4m"), default="ring16",
                    help="ring16: 16 ring views at --width x --height (the headline); uneven64: 64 views "
                         "with seeded sizes 320-1920 x 240-1080 on a sphere of radius 4, means in [-2,2]^3; "
                         "small256: 16,384 splats, one 256x256 view from (0,0,3); ring8_720p: 500,000 splats, 8 ring "
                         "views at 1280x720; growth4m: the growth run (1M splats growing to at most 4M, 16 ring "
                         "views at 1080p, prune / densify every 100 steps; give --steps, e.g. 2000)")
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
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy (torch) three times, to scale the "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
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
    N, V, W, H = args.n, args.views, args.width, args.height
    if args.workload == "small256":
        N, V, W, H = 16384, 1, 256, 256
    elif args.workload == "ring8_720p":
        N, V, W, H = 500_000, 8, 1280, 720
    if args.workload == "uneven64":
        V = 64
        g = sv.scene_gaussians(0, N, 2.0)
        cams, sizes = sv.sphere_cameras(2, V, 4.0)
        P = sum(w * h for w, h in sizes)
        workload = (f"{N} splats in [-2,2]^3, {V} views of seeded sizes 320-1920 x 240-1080 "
                    f"({P} pixels) per step, batched adjoint training step")
    elif args.workload == "small256":
        g = sv.scene_gaussians(0, N, 1.0)
        cams = [sv.scene_camera([0.0, 0.0, 3.0], [0.0, 0.0, 0.0], W, H, 60.0)]
        P = W * H
        workload = f"{N} splats, 1 view {W}x{H} from (0,0,3), training step"
    else:
        g = sv.scene_gaussians(0, N, 1.0)
        cams = sv.ring_cameras(V, 3.0, W, H)
        P = V * W * H
        workload = f"{N} splats, {V} views {W}x{H} per step, batched adjoint training step"
    log(f"scene: {workload}")
    tg = dict(g)
    tg["pos"] = sv.scene_jitter(1, g["pos"], 0.02)
    targets = []
    with sv.Context(capacity=N) as tc:
        configure(tc, args)
        tc.set_gaussians(tg)
        tc.set_views(cams)
        for v in range(V):
            tc.forward(v)
            targets.append(tc.image(v))
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
                 "method": "reference method per view: wavefront fixed-point backward, atomic reduce, "
                           "geometry backward, Adam",
                 "phase_ms_per_step": {k: v[0] / ps_p for k, v in tim_p.items() if v[1]}}
        log(f"plain: {1e3 * sec_p:.2f} ms/step")

    I = int(st_b.n_isect)
    key_bytes = 8 if st_b.key_bits > 32 else 4
    passes = int(st_b.sort_passes)
    pair_passes = (int(st_b.key_bits) + 7) // 8  # passes of a sort of every (tile, depth) key
    phase = {k: (v[0] / max(v[1], 1), v[1] / ps_b) for k, v in tim_b.items() if v[1]}
    # sorted keys/s: intersections per second of building the depth-ordered
    # tile lists from the projections (emit + sort + ranges [+ fill])
    list_ms_step = sum(tim_b.get(k, (0.0, 0))[0] for k in ("emit", "sort", "ranges", "fill")) / ps_b
    keys_per_s = I / (list_ms_step * 1e-3) if list_ms_step > 0 else None
    # dominant single-kernel phase
    cands = {k: phase[k][0] for k in ("forward", "backward", "vjp_adam") if k in phase}
    dom = max(cands, key=cands.get)
    dom_ms = cands[dom]
    dom_bytes = kernel_bytes(dom, N, V, I, P, key_bytes, pair_passes)
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    b_alg = step_b_alg(N, V, I, P, pair_passes)
    traffic = measured_traffic(dom, N, V, W, H, args.workload)
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
        "config": {"workload": workload, "splats": N, "views": V,
                   "width": W if args.workload != "uneven64" else None,
                   "height": H if args.workload != "uneven64" else None, "pixels": P,
                   "tile": 16, "lds_batch": chosen_b["batch"],
                   "persistent": chosen_b["persistent"], "graph": args.graph, "xcd_remap": chosen_b["xcd_remap"],
                   "binned_build": args.binned},
        "autotune": tune_b,
        "eager_launch_ms_per_step": 1e3 * eager_b if eager_b else None,
        "keys_per_s": keys_per_s,
        "tile_list_ms_per_step": list_ms_step,
        "intersections_per_step": I,
        "tile_lists": "binned (supertile pairs sorted, lists filled)" if st_b.binned else "pair sort",
        "sort_key_bits": int(st_b.key_bits),
        "sort_passes": passes,
        "plain": plain,
        "speedup_vs_plain": (plain["ms_per_step"] / (1e3 * sec_b)) if plain else None,
        "sequential_adjoint": sequential,
        "speedup_vs_sequential": (sequential["ms_per_step"] / (1e3 * sec_b)) if sequential else None,
        "phase_ms_per_launch": {k: v[0] for k, v in phase.items()},
        "step_b_alg_bytes": b_alg,
        "step_b_alg_GBps": b_alg / sec_b / 1e9,
        "step_b_alg_frac": b_alg / sec_b / 1e9 / HBM_PEAK_GBPS,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBPS, "traffic": traffic,
                     "traffic_source": "profiles/hbm_traffic_1M_16x1080p.md (rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                                       "copy-calibrated)" if traffic else None,
                     "algorithmic_bytes_per_launch": dom_bytes, "mean_launch_ms": dom_ms},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle, one view)...")
        out["cpu_baseline"] = cpu_baseline(g, cams, targets, args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
