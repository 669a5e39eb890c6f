"""The binned tile build (supertile pairs sorted, tile lists filled from
them; sv_config.binned_build = 1) against the sort of every (tile, splat)
pair (binned_build = 0) and the oracle: sort keys, values and tile offsets
bit-exact, batched training steps bit-identical (images, gradients,
parameters), and the fall back to the pair sort when the supertile key and
the rectangle bits exceed 32 bits."""
import numpy as np
import pytest

from conftest import small_scene

pytestmark = pytest.mark.gpu


def _cfg(ctx, **kw):
    c = ctx.config()
    c.sh_deg = 3
    c.w_l2, c.w_l1, c.w_ssim = 0.8, 0.2, 0.0
    c.lr_pos = c.lr_sh = c.lr_opacity = c.lr_scale = c.lr_rot = 1e-3
    c.beta1, c.beta2, c.eps = 0.9, 0.999, 1e-15
    c.sh_update_interval = 1
    c.pos_lr_schedule = 0
    for k, v in kw.items():
        setattr(c, k, v)
    ctx.set_config(c)


def test_binned_lists_match_oracle(sv, orc):
    g, _, cam = small_scene(sv)
    view = orc.view_params(256, 256, sh_deg=3)
    ref = orc.render_and_grad(g, cam, view)
    with sv.Context(capacity=32768) as c:
        _cfg(c, binned_build=1)
        c.set_gaussians(g)
        c.set_views([cam])
        c.set_target(0, ref["color"])
        st = c.step_views([0], sv.MODE_BATCHED, stats=True)  # lists of the step's build
        n = st.n_isect
        keys, vals = c.keys(0, n)
        offs = c.tile_offsets(0)
        assert st.binned == 1
    assert n == ref["n_isect"]
    assert np.array_equal(keys, ref["keys"]) and np.array_equal(vals, ref["vals"])
    assert np.array_equal(offs, ref["offsets"])


@pytest.mark.parametrize("shape", [(2, 1920, 1080, 200_000), (5, 1000, 700, 60_000)])
def test_binned_equals_pair_sort(sv, shape):
    nv, w, h, n = shape
    g = sv.scene_gaussians(21, n, 1.0)
    cams = sv.ring_cameras(nv, 3.0, w, h)
    tg = dict(g)
    tg["pos"] = sv.scene_jitter(22, g["pos"], 0.02)
    with sv.Context(capacity=n) as t:
        _cfg(t)
        t.set_gaussians(tg)
        t.set_views(cams)
        targets = []
        for v in range(nv):
            t.forward(v)
            targets.append(t.image(v))
    out = {}
    for binned in (0, 1):
        with sv.Context(capacity=n) as c:
            _cfg(c, binned_build=binned, keep_grads=1)
            c.set_gaussians(g)
            c.set_views(cams)
            for v in range(nv):
                c.set_target(v, targets[v])
            lists = []
            for v in range(nv):  # single-view builds: keys, values, offsets
                m = c.build_tiles(v)
                k, vv = c.keys(v, m)
                lists.append((k, vv, c.tile_offsets(v)))
            st = c.step_views(list(range(nv)), sv.MODE_BATCHED, stats=True)
            st2 = c.step_views(list(range(nv)), sv.MODE_BATCHED, stats=True)
            out[binned] = (lists, [c.image(v) for v in range(nv)], c.grads(), c.get_gaussians(), st, st2)
            print(f"[binned] {nv}x{w}x{h} binned={binned}: I={st.n_isect} key bits {st.key_bits} "
                  f"passes {st.sort_passes} binned {st.binned}", flush=True)
    (la, ia, ga, pa, sa, _), (lb, ib, gb, pb, sb, sb2) = out[0], out[1]
    assert sa.binned == 0 and sb.binned == 1 and sb2.binned == 1
    assert sa.n_isect == sb.n_isect and sa.key_bits == sb.key_bits
    for (k0, v0, o0), (k1, v1, o1) in zip(la, lb):
        assert np.array_equal(k0, k1) and np.array_equal(v0, v1) and np.array_equal(o0, o1)
    for v in range(nv):
        assert np.array_equal(ia[v], ib[v])
    for k in ("pos", "logscale", "quat", "logit", "sh"):
        assert np.array_equal(ga[k], gb[k]), k
        assert np.array_equal(pa[k], pb[k]), k


def test_binned_falls_back_when_key_too_wide(sv):
    # depths of thousands of units: depth_q needs 13 bits, 8 supertile bits
    # + 13 + 12 rectangle bits > 32, so the build sorts the tile pairs
    n = 20_000
    g = sv.scene_gaussians(3, n, 1.0)
    g["pos"] = (g["pos"] * np.float32(200.0)).astype(np.float16).astype(np.float32)
    g["logscale"] = g["logscale"] + np.float32(np.log(200.0))
    g["logscale"] = g["logscale"].astype(np.float16).astype(np.float32)
    cam = sv.scene_camera([0.0, 0.0, 5000.0], [0, 0, 0], 1920, 1080, 60.0)
    target = np.zeros((1080, 1920, 3), np.float32)
    res = {}
    for binned in (0, 1):
        with sv.Context(capacity=n) as c:
            _cfg(c, binned_build=binned, far_plane=1e4)
            c.set_gaussians(g)
            c.set_views([cam])
            c.set_target(0, target)
            st = c.step_views([0], sv.MODE_BATCHED, stats=True)
            k, v = c.keys(0, st.n_isect)
            res[binned] = (k, v, c.tile_offsets(0), c.image(0), st.binned, st.key_bits, st.n_isect)
            print(f"[binned] deep scene binned={binned}: I={st.n_isect} key bits {st.key_bits} "
                  f"binned {st.binned}", flush=True)
    assert res[1][6] > 0 and res[1][4] == 0 and res[0][4] == 0, "binning must have fallen back"
    assert res[1][5] >= 8 + 13
    for a, b in zip(res[0][:4], res[1][:4]):
        assert np.array_equal(a, b)
