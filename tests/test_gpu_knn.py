"""sv_knn_init on the device (initializeKnn, gaussian-renderer.ts:2131-2146;
knn-init.wgsl:40-90) against the oracle's all-pairs restatement: density
and log-scales bit-exact on clouds that stress the grid search (uniform,
clusters with far outliers, a plane, exact duplicates, n <= 4), other
parameters untouched; at 1M points, sampled points against a numpy
all-pairs scan with the same float32 arithmetic."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def f16(x):
    return np.asarray(x, np.float32).astype(np.float16).astype(np.float32)


def _cloud(kind, n, rng):
    if kind == "uniform":
        return rng.uniform(-1, 1, (n, 3))
    if kind == "outliers":
        c = rng.normal(size=(16, 3)) * 3
        p = c[rng.integers(0, 16, n)] + 0.02 * rng.normal(size=(n, 3))
        p[: n // 500] = rng.normal(size=(n // 500, 3)) * 200  # far outliers
        return p
    if kind == "plane":
        return np.c_[rng.uniform(-3, 3, (n, 2)), np.full(n, 0.25)]
    return np.round(rng.uniform(-1, 1, (n, 3)) * 6) / 6  # duplicates


def _run(sv, pos, extent):
    n = len(pos)
    g = sv.scene_gaussians(1, n, 1.0)
    g["pos"] = np.ascontiguousarray(pos, np.float32)
    c = sv.Context(capacity=n)
    c.set_gaussians(g)
    t0 = time.perf_counter()
    dens = c.knn_init(extent)
    dt = time.perf_counter() - t0
    h = c.get_gaussians()
    c.close()
    return g, h, dens, dt


@pytest.mark.parametrize("kind", ["uniform", "outliers", "plane", "duplicates"])
def test_knn_matches_oracle(sv, orc, kind):
    rng = np.random.default_rng(3)
    pos = f16(_cloud(kind, 20000, rng))
    g, h, dens, dt = _run(sv, pos, 4.0)
    want_d, want_ls = orc.knn_init(pos, 4.0)
    print(f"[knn] {kind}: {len(pos)} points {dt * 1e3:.1f} ms", flush=True)
    assert np.array_equal(dens.view(np.uint32), want_d.view(np.uint32))
    assert np.array_equal(h["logscale"], np.repeat(want_ls[:, None], 3, 1))
    for k in ("pos", "quat", "logit", "sh"):
        assert np.array_equal(h[k], g[k]), k


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5])
def test_knn_tiny(sv, orc, n):
    pos = f16(np.random.default_rng(n).normal(size=(n, 3)))
    _, h, dens, _ = _run(sv, pos, 2.0)
    want_d, want_ls = orc.knn_init(pos, 2.0)
    assert np.array_equal(dens, want_d) and np.array_equal(h["logscale"][:, 0], want_ls)


def test_knn_1m_sampled(sv):
    rng = np.random.default_rng(11)
    n = 1_000_000
    pos = f16(rng.normal(size=(n, 3)) * np.array([2.0, 1.0, 0.5]))
    _, h, dens, dt = _run(sv, pos, 6.0)
    print(f"[knn] 1M points {dt * 1e3:.1f} ms (with upload)", flush=True)
    for i in rng.choice(n, 256, replace=False):
        d = pos - pos[i]
        d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        d2[i] = np.inf
        k3 = np.sort(np.partition(d2, 3)[:3])
        s = np.sqrt(k3.astype(np.float32))
        assert dens[i] == np.float32((s[0] + s[1] + s[2]) / np.float32(3)), i
