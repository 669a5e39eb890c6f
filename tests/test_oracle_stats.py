"""The oracle's per-step statistics (orc_accumulate_importance:
importance-accum.wgsl:19-40; orc_accumulate_grads: densify-accum.wgsl:19-46)
against closed forms: the decay is applied before the add, so k steps of a
constant tile count c sum to c (1 - d^k) / (1 - d); a splat that leaves the
view only decays; the radius is a running max of the decayed maximum; the
gradient norms of Pythagorean triples are exact, the 1e-10 threshold
decides whether anything (count included) changes, and the unused second
slot is zeroed only on an accumulating step."""
import numpy as np


def _proj(orc, counts, radii):
    p = np.zeros(len(counts), orc.PROJ_DTYPE)
    p["tiles_count"] = np.asarray(counts, np.uint32)
    p["radius"] = np.asarray(radii, np.float32)
    return p


def test_importance_geometric_sums(orc):
    d = 0.98
    p = _proj(orc, [3, 0, 7, 1], [4.0, 9.0, 2.5, 1.0])
    st = np.zeros((4, 4), np.float32)
    k = 40
    for _ in range(k):
        st = orc.accumulate_importance(p, st, d)
    geo = (1.0 - d ** k) / (1.0 - d)
    assert np.allclose(st[:, 0], np.array([3, 0, 7, 1]) * geo, rtol=2e-6, atol=0)
    assert np.allclose(st[:, 1], np.array([1, 0, 1, 1]) * geo, rtol=2e-6, atol=0)
    assert st[1, 0] == 0 and st[1, 1] == 0 and st[1, 2] == 0  # never visible: nothing moves
    # the running max of a constant radius is that radius (r d < r)
    assert np.array_equal(st[[0, 2, 3], 2], np.float32([4.0, 2.5, 1.0]))


def test_importance_decay_before_add(orc):
    st = np.float32([[10.0, 2.0, 5.0, 0.0], [10.0, 2.0, 5.0, 0.0]])
    out = orc.accumulate_importance(_proj(orc, [3, 0], [4.0, 8.0]), st, 0.98)
    want0 = [np.float32(10.0) * np.float32(0.98) + np.float32(3), np.float32(2.0) * np.float32(0.98) + 1,
             max(np.float32(5.0) * np.float32(0.98), np.float32(4.0))]
    assert np.array_equal(out[0, :3], np.float32(want0))
    # not visible this step (count 0): decay only, the radius is not taken
    assert np.array_equal(out[1, :3], np.float32([10.0, 2.0, 5.0]) * np.float32(0.98))


def test_importance_leaves_view(orc):
    d = np.float32(0.98)
    st = np.zeros((1, 4), np.float32)
    st = orc.accumulate_importance(_proj(orc, [5], [3.0]), st, float(d))
    for m in range(1, 25):
        st = orc.accumulate_importance(_proj(orc, [0], [3.0]), st, float(d))
        assert np.isclose(st[0, 0], 5.0 * float(d) ** m, rtol=1e-6)
        assert np.isclose(st[0, 2], 3.0 * float(d) ** m, rtol=1e-6)


def test_grad_accum_norms_and_threshold(orc):
    geom = np.float32([
        [2.0, 3.0, 6.0, 3.0, 4.0],      # |conic| = 7, |mean| = 5
        [0.0, 0.0, 0.0, 0.0, 0.0],      # nothing
        [0.0, 0.0, 0.0, 1e-11, 0.0],    # below 1e-10 in both norms
        [0.0, 0.0, 0.0, 2e-10, 0.0],    # above: counted
        [0.0, 0.0, 0.0, 0.0, 0.0],
    ])
    acc0 = np.float32([[1.0, 7.0, 1.0, 1.0]] * 5)
    acc = orc.accumulate_grads(geom, acc0)
    assert np.array_equal(acc[0], np.float32([6.0, 0.0, 8.0, 2.0]))
    assert np.array_equal(acc[1], acc0[1]) and np.array_equal(acc[2], acc0[2])  # untouched, slot 1 kept
    assert acc[3, 3] == 2.0 and acc[3, 1] == 0.0 and np.isclose(acc[3, 0], 1.0 + 2e-10)
    acc = orc.accumulate_grads(geom, acc)
    assert np.array_equal(acc[0], np.float32([11.0, 0.0, 15.0, 3.0]))


def test_grad_accum_conic_only(orc):
    # only the conic gradient is non-zero: still counted (either norm suffices)
    geom = np.float32([[0.0, 1e-9, 0.0, 0.0, 0.0], [1.0, 2.0, 2.0, 0.0, 0.0]])
    acc = orc.accumulate_grads(geom, np.zeros((2, 4), np.float32))
    assert acc[0, 3] == 1.0 and acc[0, 0] == 0.0 and np.isclose(acc[0, 2], 1e-9)
    assert np.array_equal(acc[1], np.float32([0.0, 0.0, 3.0, 1.0]))
