"""The longest-first work order of the 16-px raster launches
(sv_config.tile_order): only which tile a workgroup renders, and when,
changes, so three steps leave bit-identical parameters and moments with and
without it, in every launch shape that takes the order: one workgroup per
tile with and without the XCD mapping, the persistent queues, the split
backward and the single-launch one, graph and eager launches, ring views and
views of uneven sizes. The forward's images, alpha and last indices are
compared too.
"""
import numpy as np
import pytest

from test_gpu_split import KEYS, _run, _scene

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("sizes", [[(256, 192)] * 4, [(200, 136), (96, 72), (330, 250), (64, 48), (160, 160)]],
                         ids=["ring4", "uneven5"])
@pytest.mark.parametrize("graph,persistent,xcd_remap,split", [(1, 0, 1, 1), (1, 0, 0, 1), (0, 1, 1, 1),
                                                              (1, 1, 0, 0), (0, 0, 1, 0)])
def test_tile_order_is_bit_identical(sv, orc, sizes, graph, persistent, xcd_remap, split):
    g, cams, targets = _scene(sv, orc, sizes)
    cfg = dict(use_graph=graph, persistent=persistent, xcd_remap=xcd_remap, split_backward=split)
    ref = _run(sv, g, cams, targets, tile_order=0, **cfg)
    got = _run(sv, g, cams, targets, tile_order=1, **cfg)
    for a, b in zip(got, ref):
        for k in KEYS:
            assert np.array_equal(a[k], b[k]), k
    print(f"[order] {len(sizes)} views, graph {graph}, persistent {persistent}, xcd {xcd_remap}, split {split}: "
          "identical", flush=True)


def test_tile_order_forward_outputs(sv, orc):
    sizes = [(200, 136), (96, 72), (330, 250)]
    g, cams, targets = _scene(sv, orc, sizes)
    out = {}
    for order in (0, 1):
        with sv.Context(capacity=g["pos"].shape[0]) as c:
            c.set_config(sh_deg=3, tile_order=order)
            c.set_gaussians(g)
            c.set_views(cams)
            c.render_batched(list(range(len(cams))))
            out[order] = [(c.image(v), c.alpha(v), c.last_ids(v)) for v in range(len(cams))]
    for v in range(len(cams)):
        for x, y in zip(out[0][v], out[1][v]):
            assert np.array_equal(x, y)
