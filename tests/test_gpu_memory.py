"""Device-memory accounting (sv_memory_stats): a new context holds at least
its parameter, moment and gradient planes (4 x 59 floats per Gaussian of
capacity); building tiles adds the tile-build buffers; a densify past the
capacity regrows the planes; the peak never falls below the current figure
and resets to it."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_memory_stats(sv):
    cap = 20000
    with sv.Context(capacity=cap) as c:
        now, peak = c.memory_stats()
        assert now >= 4 * 59 * 4 * cap and peak >= now
        g = sv.scene_gaussians(0, cap, 1.0)
        c.set_gaussians(g)
        c.set_views([sv.scene_camera([0.0, 0.0, 3.0], [0, 0, 0], 256, 256, 60.0)])
        c.build_tiles(0)
        built, peak2 = c.memory_stats()
        assert built > now and peak2 >= built
        act = np.zeros(cap, np.uint8)
        act[: cap // 2] = 1
        c.densify(act, 1)
        grown, _ = c.memory_stats()
        assert c.count == cap + cap // 2
        assert grown >= built + 4 * 59 * 4 * (32768 - cap)  # planes regrown to the next power of two
        c.reset_memory_peak()
        a, b = c.memory_stats()
        assert a == b == grown
        print(f"[memory] new {now / 2**20:.1f} MiB, built {built / 2**20:.1f} MiB, grown {grown / 2**20:.1f} MiB",
              flush=True)
