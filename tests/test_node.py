"""The Node N-API addon and the JS drop-in renderer (skyfall-viewer_amd/node).

CPU: the addon loads in node, exports the whole binding, throws a JS Error
(not a crash) when no device is present, and its seeded scene helpers give
the same bytes as the library through ctypes.
GPU: one training iteration through createGaussianRenderer(...).train(1) and
a RasterizeFunction forward/backward give bit-identical parameters, image
and gradients to the same calls made through ctypes.
"""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ADDON = os.path.join(ROOT, "skyfall-viewer_amd", "node", "skyfall_sv.node")
INDEX = os.path.join(ROOT, "skyfall-viewer_amd", "node", "index.js")
NODE = shutil.which("node")

needs_node = pytest.mark.skipif(NODE is None or not os.path.exists(ADDON), reason="node or the addon is missing")

EXPORTS = ["version", "create", "destroy", "getConfig", "setConfig", "setGaussians", "getGaussians", "setMoments",
           "getMoments", "getCount", "setViews", "setTarget", "setResidual", "buildTiles", "forward", "loss",
           "backward", "backwardFixed", "zeroGrad", "adam", "step", "stepAsync", "readImage", "readResidual",
           "readAlpha", "readLastIds", "readGrads", "readKeys", "pruneOpacity", "pruneMask", "densify",
           "resetOpacity", "densifyDecide", "densifyDecideGrad", "enableTiming", "timings", "sync", "autotune", "sceneGaussians",
           "sceneJitter", "sceneCamera"]


def _node(script, *args, timeout=120):
    return subprocess.run([NODE, "-e", script, *args], capture_output=True, text=True, timeout=timeout, cwd=ROOT)


@needs_node
def test_addon_loads_and_exports(tmp_path):
    js = f"""
const sv = require({json.dumps(INDEX)});
const out = {{ exports: Object.keys(sv.native), api: Object.keys(sv), version: sv.native.version() }};
try {{ sv.native.create(1000); out.create = 'ok'; }} catch (e) {{ out.create = e.message; }}
console.log(JSON.stringify(out));
"""
    r = _node(js)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert set(EXPORTS) <= set(out["exports"])
    assert {"RasterizeFunction", "createGaussianRenderer", "Context"} <= set(out["api"])
    assert "gfx950" in out["version"]


@needs_node
def test_addon_has_no_silent_device_fallback(sv):
    import torch

    if torch.cuda.is_available():
        pytest.skip("a device is present")
    r = _node(f"const sv = require({json.dumps(INDEX)}); try {{ sv.native.create(10); console.log('created'); }}"
              " catch (e) { console.log('error: ' + e.message); }")
    assert r.returncode == 0 and r.stdout.startswith("error: sv_create failed"), r.stdout + r.stderr


@needs_node
def test_scene_helpers_match_ctypes(sv, tmp_path):
    out = tmp_path / "g.bin"
    js = f"""
const fs = require('fs');
const sv = require({json.dumps(INDEX)});
const g = sv.native.sceneGaussians(7, 1000, 1.5);
const j = sv.native.sceneJitter(8, g.pos, 0.02);
const c = sv.native.sceneCamera([2, 1, 3], [0, 0, 0], 640, 360, 60);
const parts = [g.pos, g.logscale, g.quat, g.logit, g.sh, j, c.view, c.proj, c.focal];
fs.writeFileSync({json.dumps(str(out))}, Buffer.concat(parts.map((a) => Buffer.from(a.buffer, a.byteOffset, a.byteLength))));
"""
    r = _node(js)
    assert r.returncode == 0, r.stderr
    b = np.fromfile(out, np.float32)
    g = sv.scene_gaussians(7, 1000, 1.5)
    j = sv.scene_jitter(8, g["pos"], 0.02)
    cam = sv.camera_arrays(sv.scene_camera([2, 1, 3], [0, 0, 0], 640, 360, 60.0))
    want = np.concatenate([g["pos"].ravel(), g["logscale"].ravel(), g["quat"].ravel(), g["logit"], g["sh"].ravel(),
                           j.ravel(), cam["view"], cam["proj"], cam["focal"]])
    assert np.array_equal(b.view(np.uint32), want.view(np.uint32))


@needs_node
@pytest.mark.gpu
def test_js_renderer_and_function_match_ctypes(sv, orc, tmp_path):
    n, w, h = 16384, 256, 256
    g = sv.scene_gaussians(0, n, 1.0)
    cam = sv.scene_camera([0, 0, 3], [0, 0, 0], w, h, 60.0)
    tg = dict(g)
    tg["pos"] = sv.scene_jitter(1, g["pos"], 0.02)
    target = orc.render_and_grad(tg, cam, orc.view_params(w, h, sh_deg=3))["color"]
    tfile, ofile = tmp_path / "target.bin", tmp_path / "out.bin"
    target.astype(np.float32).tofile(tfile)
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "node", "train_step.js"), str(tfile), str(ofile),
                        str(n), str(w), str(h)], capture_output=True, text=True, timeout=300, cwd=ROOT)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    meta = json.loads(r.stdout.strip().splitlines()[-1])
    assert meta["iteration"] == 1 and meta["count"] == n
    blob = np.fromfile(ofile, np.float32)
    # the same through ctypes
    c = sv.Context(capacity=2 * n)
    c.set_config(sh_deg=3, w_l2=0.8, w_l1=0.2, w_ssim=0.0, lr_pos=1e-3, lr_sh=1e-3, lr_opacity=1e-3,
                 lr_scale=1e-3, lr_rot=1e-3, beta1=0.9, beta2=0.999, eps=1e-15, sh_update_interval=1,
                 pos_lr_schedule=0)
    c.set_gaussians(g)
    c.set_views([cam])
    c.set_target(0, target)
    c.step_views([0], sv.MODE_SEQUENTIAL)
    after = c.get_gaussians()
    c.close()
    c = sv.Context(capacity=2 * n)
    c.set_gaussians(g)
    c.set_views([cam])
    c.forward(0)
    img = c.image(0)
    c.forward(0)
    c.set_residual(0, (target - img).astype(np.float32))
    c.zero_grad()
    c.backward(0)
    grads = c.grads()
    c.close()
    want = [after[k] for k in ("pos", "logscale", "quat", "logit", "sh")] + [img] + \
           [grads[k] for k in ("pos", "logscale", "quat", "logit", "sh", "geom")]
    off = 0
    for name, a in zip(meta["order"], want):
        m = a.size
        got = blob[off:off + m]
        off += m
        assert np.array_equal(got.view(np.uint32), a.ravel().view(np.uint32)), name
    assert off == blob.size


@needs_node
@pytest.mark.gpu
def test_js_train_stops_mid_loop():
    """train(1000) stops at the iteration after stopTraining(); a second
    train() while the first runs warns and returns."""
    script = os.path.join(ROOT, "tests", "node", "stop_train.js")
    print("[node] stop mid-train", flush=True)
    res = subprocess.run([NODE, script], capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert res.returncode == 0, res.stderr
    st = json.loads(res.stdout.strip().splitlines()[-1])
    print(f"[node] {st}", flush=True)
    assert st["during_training"] and not st["training"]
    assert st["warned"] >= 1
    assert 2 <= st["iterations"] < 1000 and st["target"] == 1000


@pytest.mark.gpu
@needs_node
def test_js_schedule_densifies_and_prunes():
    """train() runs the reference's schedule (gaussian-renderer.ts:3796-3828):
    with it brought forward to every 10 iterations, the splat count changes
    only at iterations 10, 20, 30 and 40 (densify with the trainer's rule,
    then prune), some event grows the scene, and the parameters stay
    finite."""
    script = os.path.join(ROOT, "tests", "node", "schedule_train.js")
    print("[node] scheduled train", flush=True)
    res = subprocess.run([NODE, script], capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert res.returncode == 0, res.stderr
    st = json.loads(res.stdout.strip().splitlines()[-1])
    counts = st["counts"]
    print(f"[node] counts at events: {[counts[i] for i in (9, 19, 29, 39)]}, max opacity {st['maxOpacity']:.4f}",
          flush=True)
    assert st["iteration"] == 40 and st["finite"]
    for i in range(1, 40):
        if (i + 1) % 10:
            assert counts[i] == counts[i - 1], i
    assert max(counts) > 8192


@needs_node
def test_training_controller_schedule():
    """The JS TrainingController's predicates at the edges of the reference's
    DEFAULT_TRAINING_SCHEDULE (training-controller.ts:24-44, 86-126):
    densify and prune every 100 iterations from 500 (densify until 15000,
    prune until 25000), opacity reset every 3000 inside the densify window,
    screen-size pruning after 3000."""
    js = f"""
const sv = require({json.dumps(INDEX)});
const c = new sv.TrainingController(2.0);
const out = {{}};
for (const i of [0, 100, 499, 500, 550, 600, 3000, 3100, 15000, 15100, 25000, 25100]) {{
  c.setIteration(i);
  out[i] = [c.shouldDensify(), c.shouldPrune(), c.shouldResetOpacity(), c.getPruneParams().maxScreenRadius];
}}
const r = new sv.TrainingController(2.0, {{ maxScaleAbsolute: 0, maxScaleRatio: 0.5 }});
out.maxScale = r.getPruneParams().maxScale;
out.reset = c.getOpacityResetValue();
console.log(JSON.stringify(out));
"""
    r = _node(js)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    F, T = False, True
    expect = {0: [F, F, F, 10000], 100: [F, F, F, 10000], 499: [F, F, F, 10000], 500: [T, T, F, 10000],
              550: [F, F, F, 10000], 600: [T, T, F, 10000], 3000: [T, T, T, 10000], 3100: [T, T, F, 20],
              15000: [T, T, T, 20], 15100: [F, T, F, 20], 25000: [F, T, F, 20], 25100: [F, F, F, 20]}
    for i, e in expect.items():
        assert out[str(i)] == e, i
    assert out["maxScale"] == 1.0 and out["reset"] == 0.01
