"""The oracle's knn_init (knn-init.wgsl:40-90, constants
gaussian-renderer.ts:77-79) pinned without the restatement: 3-NN distances
from scipy's cKDTree in float64, the sigma rule in float64 numpy, and closed
forms for n <= 4 and duplicate points."""
import math

import numpy as np
import pytest
from scipy.spatial import cKDTree

MIN_R, MAX_R, BIAS_R = 0.0007, 0.007, 0.0015


def f16(x):
    return np.asarray(x, np.float32).astype(np.float16).astype(np.float32)


def sigma_rule(mean, extent):
    """Float64 restatement of knn-init.wgsl:66-75."""
    base = np.asarray(mean, np.float64) / 3.0
    s = np.maximum(base, extent * BIAS_R)
    s = np.minimum(np.maximum(s, extent * MIN_R), extent * MAX_R)
    return np.minimum(s, 0.1)


def kdtree_mean(pos):
    p = pos.astype(np.float64)
    n = len(p)
    k = min(4, n)
    d, idx = cKDTree(p).query(p, k=k)
    d = d.reshape(n, k)
    idx = idx.reshape(n, k)
    out = np.empty(n)
    for i in range(n):
        row = [d[i, j] for j in range(k) if idx[i, j] != i]
        row = row[:3] + [1e10] * (3 - len(row[:3]))
        out[i] = sum(row) / 3.0
    return out


@pytest.mark.parametrize("kind", ["normal", "clusters", "plane", "lattice"])
def test_knn_against_kdtree(orc, kind):
    rng = np.random.default_rng(7)
    n = 3000
    if kind == "normal":
        pos = rng.normal(size=(n, 3))
    elif kind == "clusters":
        c = rng.normal(size=(12, 3)) * 4
        pos = c[rng.integers(0, 12, n)] + 0.05 * rng.normal(size=(n, 3))
    elif kind == "plane":
        pos = np.c_[rng.uniform(-2, 2, (n, 2)), np.zeros(n)]
    else:  # many exact duplicates
        pos = np.round(rng.uniform(-1, 1, (n, 3)) * 4) / 4
    pos = f16(pos)
    extent = 5.0
    dens, ls = orc.knn_init(pos, extent)
    want = kdtree_mean(pos)
    assert np.abs(dens - want).max() <= 1e-6 * np.maximum(want, 1e-30).max() + 1e-7
    lsw = np.log(np.maximum(sigma_rule(want, extent), 1e-6))
    # the f16 rounding of log(sigma): within half an f16 ulp of the float64 value
    ulp = np.spacing(np.abs(lsw).astype(np.float16)).astype(np.float64)
    assert np.all(np.abs(ls - lsw) <= 0.5 * ulp + 1e-6 * np.abs(lsw))
    assert np.array_equal(ls, f16(ls))


def test_knn_closed_forms(orc):
    # four points at distances 1, 2, 3 from the origin (scaled by 1/64, exact in f16)
    s = 1.0 / 64
    pos = f16(np.array([[0, 0, 0], [s, 0, 0], [0, 2 * s, 0], [0, 0, 3 * s]]))
    dens, ls = orc.knn_init(pos, 1.0)
    assert dens[0] == np.float32((np.float32(s) + np.float32(2 * s) + np.float32(3 * s)) / np.float32(3))
    d1 = sorted([s, math.sqrt(5) * s, math.sqrt(10) * s])
    assert abs(dens[1] - sum(d1) / 3) <= 1e-7
    # origin: mean 2/64, sigma = mean / 3 = 1/96 > 0.007 e -> 0.007
    assert ls[0] == flog(np.float32(1.0) * np.float32(MAX_R))
    # one point: every slot 1e10 -> sigma = min(0.007 e, 0.1)
    d, l1 = orc.knn_init(f16([[0.25, 0.5, 0.75]]), 20.0)
    assert d[0] == np.float32(1e10) and l1[0] == flog(0.1)
    # two points: (dist + 1e10 + 1e10) / 3
    d, _ = orc.knn_init(f16([[0, 0, 0], [0, 0.5, 0]]), 1.0)
    assert d[0] == d[1] == np.float32((np.float32(0.5) + np.float32(1e10) + np.float32(1e10)) / np.float32(3))
    # duplicates count as distance 0 (the shader skips only its own index)
    d, l2 = orc.knn_init(f16([[0, 0, 0], [0, 0, 0], [0, 0, 0], [1, 0, 0]]), 1.0)
    assert d[0] == np.float32(1) / np.float32(3) and d[3] == np.float32(1.0)
    # sigma floor: mean 0 -> max(0, 0.0015 e) = 0.0015 e, inside [0.0007 e, 0.007 e]
    d, l3 = orc.knn_init(f16([[0.5, 0.5, 0.5]] * 4 + [[2, 0, 0]]), 1.0)
    assert d[0] == 0.0
    assert l3[0] == flog(np.float32(1.0) * np.float32(BIAS_R))


def flog(x):
    """f16(log(x)) via float64 log rounded to float32."""
    return f16(np.float32(np.log(np.float64(np.float32(x)))))
