"""The raster tile of the forward (8, 16 or 32 pixels; sv_config.raster_px)
regroups the same per-pixel work: the tile lists stay the 16-px binning
tiles' (gaussian-renderer.ts:366), and every pixel composites its tile's
list in the same order with the same operations. So across raster tiles
8 / 16 / 32 and LDS batches 64 / 256, on views whose sizes are not multiples
of 32 (partial blocks and quadrants at the edges), images, alpha and last
indices are identical bit for bit, images match the oracle within 1e-5 and
alpha exactly, and a batched training step (graph launch) leaves identical
parameters: after an 8-px forward the backward culls every entry itself,
after a 16- or 32-px forward it visits only the entries and quadrants the
forward's taken masks name, and both write the same records. The
autotuner times all three raster tiles and keeps the configuration it
reports as fastest.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("pos", "logscale", "quat", "logit", "sh")
SIZES = [(200, 136), (96, 72), (330, 250)]


def _cfg(c, **kw):
    c.set_config(sh_deg=3, w_l2=0.8, w_l1=0.2, w_ssim=0.0, lr_pos=1e-3, lr_sh=1e-3, lr_opacity=1e-3,
                 lr_scale=1e-3, lr_rot=1e-3, beta1=0.9, beta2=0.999, eps=1e-15, sh_update_interval=1,
                 pos_lr_schedule=0, **kw)


@pytest.fixture(scope="module")
def scene(sv, orc):
    g = sv.scene_gaussians(0, 40000, 1.0)
    cams = []
    for k, (w, h) in enumerate(SIZES):
        az = 2.0 * np.pi * k / len(SIZES)
        cams.append(sv.scene_camera([3.0 * np.sin(az), 0.4 * k, 3.0 * np.cos(az)], [0, 0, 0], w, h, 60.0))
    tg = dict(g)
    tg["pos"] = sv.scene_jitter(1, g["pos"], 0.02)
    targets = [orc.render_and_grad(tg, c, orc.view_params(w, h, sh_deg=3))["color"] for c, (w, h) in zip(cams, SIZES)]
    return g, cams, targets


def _render(sv, g, cams, raster, batch):
    with sv.Context(capacity=g["pos"].shape[0]) as c:
        _cfg(c, raster_px=raster, batch=batch)
        c.set_gaussians(g)
        c.set_views(cams)
        c.render_batched(list(range(len(cams))))
        return [(c.image(v), c.alpha(v), c.last_ids(v)) for v in range(len(cams))]


def test_images_identical_across_raster_tiles(sv, orc, scene):
    g, cams, _ = scene
    base = _render(sv, g, cams, 16, 64)
    for v, (cam, (w, h)) in enumerate(zip(cams, SIZES)):
        ref = orc.render_and_grad(g, cam, orc.view_params(w, h, sh_deg=3))
        assert np.abs(base[v][0] - ref["color"]).max() <= 1e-5
        # (last indices of a batched render count in the batch's lists, so
        # they are compared across raster tiles only)
        assert np.array_equal(base[v][1], ref["alpha"])
    for raster in (8, 16, 32):
        for batch in (64, 256):
            got = _render(sv, g, cams, raster, batch)
            for v in range(len(cams)):
                for a, b in zip(got[v], base[v]):
                    assert np.array_equal(a, b), (raster, batch, v)
            print(f"[raster] tile {raster} batch {batch}: identical to tile 16 batch 64", flush=True)


def test_training_step_identical_across_raster_tiles(sv, scene):
    g, cams, targets = scene
    out = {}
    for raster in (8, 16, 32):
        with sv.Context(capacity=g["pos"].shape[0]) as c:
            _cfg(c, raster_px=raster, use_graph=1)
            c.set_gaussians(g)
            c.set_views(cams)
            for v, t in enumerate(targets):
                c.set_target(v, t)
            c.step_views(list(range(len(cams))), sv.MODE_BATCHED)
            c.step_views(list(range(len(cams))), sv.MODE_BATCHED)
            out[raster] = c.get_gaussians()
    for raster in (8, 32):
        for k in KEYS:
            assert np.array_equal(out[raster][k], out[16][k]), (raster, k)


def test_autotune_sweeps_raster_tiles(sv, scene):
    g, cams, targets = scene
    with sv.Context(capacity=g["pos"].shape[0]) as c:
        _cfg(c)
        c.set_gaussians(g)
        c.set_views(cams)
        for v, t in enumerate(targets):
            c.set_target(v, t)
        before = c.get_gaussians()
        tune = c.autotune(list(range(len(cams))), 2)
        assert {e["tile"] for e in tune} == {8, 16, 32}
        assert {e["batch"] for e in tune} == {64, 128, 256}
        best = min(tune, key=lambda e: e["ms_per_step"])
        cfg = c.config()
        assert (cfg.raster_px, cfg.batch, cfg.persistent, cfg.xcd_remap) == (best["tile"], best["batch"],
                                                                            best["persistent"], best["xcd_remap"])
        after = c.get_gaussians()
        for k in KEYS:
            assert np.array_equal(before[k], after[k]), k
        print("[raster] autotune: " + ", ".join(f"t{e['tile']}/b{e['batch']}/p{e['persistent']}/x{e['xcd_remap']} "
                                               f"{e['ms_per_step']:.3f}" for e in tune), flush=True)
