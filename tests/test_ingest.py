"""On-disk inputs (sv_ply_*, sv_colmap_*, sv_points_to_gaussians; host code
in libsv.so, CPU only): PLY in the 3DGS layout (channel-major f_rest,
rot_0..3 = (w, x, y, z)) and as a plain coloured point cloud, COLMAP
sparse binaries, and the reference's point-cloud initialisation
(load.ts:14-206, plyreader.ts:1-75, colmap-dataset.ts:129-326,
pointcloud-loader.ts:29-100). Files are written here with numpy."""
import math
import struct

import numpy as np
import pytest

SH_C0 = 0.28209479177387814


def f16(x):
    return np.asarray(x, np.float32).astype(np.float16).astype(np.float32)


def _write_ply(path, names, types, rows, fmt="binary_little_endian"):
    head = f"ply\nformat {fmt} 1.0\nelement vertex {len(rows)}\n"
    head += "".join(f"property {t} {n}\n" for n, t in zip(names, types)) + "end_header\n"
    with open(path, "wb") as fh:
        fh.write(head.encode())
        for r in rows:
            for v, t in zip(r, types):
                fh.write(struct.pack("<f", v) if t == "float" else struct.pack("<B", int(v)))


def test_ply_round_trip(sv, tmp_path):
    g = sv.scene_gaussians(3, 500, 1.0)
    p = str(tmp_path / "g.ply")
    sv.ply_save(p, g, 3)
    h, deg, ext = sv.ply_load(p)
    assert deg == 3
    for k in g:
        assert np.array_equal(h[k], g[k]), k
    assert abs(ext - float((g["pos"].max(0) - g["pos"].min(0)).max())) < 1e-6


@pytest.mark.parametrize("deg", [0, 1, 3])
def test_ply_3dgs_layout(sv, tmp_path, deg):
    rng = np.random.default_rng(deg)
    n = 40
    per = (deg + 1) ** 2 - 1
    names = ["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"]
    names += [f"f_rest_{i}" for i in range(3 * per)] + ["opacity", "scale_0", "scale_1", "scale_2",
                                                         "rot_0", "rot_1", "rot_2", "rot_3"]
    vals = rng.normal(size=(n, len(names))).astype(np.float32)
    p = str(tmp_path / "a.ply")
    _write_ply(p, names, ["float"] * len(names), vals.tolist())
    g, d, _ = sv.ply_load(p)
    assert d == deg
    col = {nm: vals[:, i] for i, nm in enumerate(names)}
    assert np.array_equal(g["pos"], f16(np.stack([col["x"], col["y"], col["z"]], 1)))
    assert np.array_equal(g["logit"], f16(col["opacity"]))
    assert np.array_equal(g["quat"], f16(np.stack([col["rot_1"], col["rot_2"], col["rot_3"], col["rot_0"]], 1)))
    assert np.array_equal(g["logscale"], f16(np.stack([col["scale_0"], col["scale_1"], col["scale_2"]], 1)))
    want = np.zeros((n, 48), np.float32)
    for c in range(3):
        want[:, c] = col[f"f_dc_{c}"]
        for i in range(per):
            want[:, 3 * (i + 1) + c] = col[f"f_rest_{c * per + i}"]
    assert np.array_equal(g["sh"], f16(want))


def test_ply_plain_point_cloud_defaults(sv, tmp_path):
    rows = [[0.5, -1.0, 2.0, 255, 128, 0], [1.0, 1.0, 1.0, 10, 20, 30]]
    p = str(tmp_path / "pc.ply")
    _write_ply(p, ["x", "y", "z", "red", "green", "blue"], ["float"] * 3 + ["uchar"] * 3, rows)
    g, deg, ext = sv.ply_load(p)
    assert deg == 0 and abs(ext - 2.0) < 1e-6
    for i, r in enumerate(rows):
        # the reader divides uchar by 255 and the loader divides again (load.ts:117-126)
        want = [((v / 255.0) / 255.0 - 0.5) / SH_C0 for v in r[3:]]
        assert np.array_equal(g["sh"][i, :3], f16(want))
        assert not g["sh"][i, 3:].any()
    assert np.all(g["logit"] == f16(-2.197))
    assert np.all(g["logscale"] == f16(math.log(0.01)))
    assert np.all(g["quat"] == np.array([0, 0, 1, 0], np.float32))


def test_ply_errors(sv, tmp_path):
    p = str(tmp_path / "ascii.ply")
    _write_ply(p, ["x", "y", "z"], ["float"] * 3, [[0, 0, 0]], fmt="ascii")
    with pytest.raises(sv.SvError, match="binary_little_endian"):
        sv.ply_load(p)
    q = str(tmp_path / "short.ply")
    _write_ply(q, ["x", "y", "z"], ["float"] * 3, [[0, 0, 0], [1, 1, 1]])
    data = open(q, "rb").read()
    open(q, "wb").write(data[:-5])
    with pytest.raises(sv.SvError, match="truncated"):
        sv.ply_load(q)


def _qvec(R):
    """Rotation matrix -> (w, x, y, z)."""
    w = math.sqrt(max(0.0, 1.0 + R[0, 0] + R[1, 1] + R[2, 2])) / 2.0
    return np.array([w, (R[2, 1] - R[1, 2]) / (4 * w), (R[0, 2] - R[2, 0]) / (4 * w), (R[1, 0] - R[0, 1]) / (4 * w)])


def _write_colmap(d, cams, images, points):
    with open(d / "cameras.bin", "wb") as fh:
        fh.write(struct.pack("<Q", len(cams)))
        for cid, model, w, h, prm in cams:
            fh.write(struct.pack("<iiQQ", cid, model, w, h) + struct.pack(f"<{len(prm)}d", *prm))
    with open(d / "images.bin", "wb") as fh:
        fh.write(struct.pack("<Q", len(images)))
        for iid, q, t, cid, name, n2 in images:
            fh.write(struct.pack("<i4d3di", iid, *q, *t, cid) + name.encode() + b"\0")
            fh.write(struct.pack("<Q", n2) + b"\1" * (24 * n2))
    with open(d / "points3D.bin", "wb") as fh:
        fh.write(struct.pack("<Q", len(points)))
        for pid, xyz, rgb, track in points:
            fh.write(struct.pack("<Q3d3Bd", pid, *xyz, *rgb, 0.5) + struct.pack("<Q", track) + b"\2" * (8 * track))


def test_colmap_model(sv, tmp_path):
    # the scene generator's look-at camera with y pointing down (COLMAP's
    # convention: a proper rotation, x right, y down, z forward)
    ref = sv.camera_arrays(sv.scene_camera([2.0, 1.0, 3.0], [0, 0, 0], 640, 360, 60.0))
    flip = np.diag([1.0, -1.0, 1.0])
    V = ref["view"].astype(np.float64).reshape(4, 4).T
    R, t = flip @ V[:3, :3], flip @ V[:3, 3]
    assert abs(np.linalg.det(R) - 1.0) < 1e-6
    want_view = np.eye(4)
    want_view[:3, :3], want_view[:3, 3] = R, t
    fx = float(ref["focal"][0])
    cams = [(7, 1, 640, 360, [fx, fx, 320.0, 180.0]), (9, 0, 100, 80, [50.0, 40.0, 30.0])]
    images = [(1, _qvec(R), t, 7, "a.png", 3), (2, [1.0, 0, 0, 0], [0.0, 0.0, 5.0], 9, "dir/b.jpg", 0)]
    points = [(1, [0.1, 0.2, 0.3], [255, 0, 10], 2), (5, [-1.0, 2.0, 0.5], [1, 2, 3], 0)]
    _write_colmap(tmp_path, cams, images, points)
    cs, names, xyz, rgb = sv.colmap_load(str(tmp_path))
    assert names == ["a.png", "dir/b.jpg"]
    assert np.allclose(xyz, [[0.1, 0.2, 0.3], [-1.0, 2.0, 0.5]]) and rgb.tolist() == [[255, 0, 10], [1, 2, 3]]
    a = sv.camera_arrays(cs[0])
    assert np.abs(np.array(a["view"], np.float64).reshape(4, 4).T - want_view).max() < 1e-6
    assert np.array_equal(a["proj"], ref["proj"]) and np.array_equal(a["focal"], ref["focal"])
    Vi = np.array(a["view_inv"], np.float64).reshape(4, 4).T
    assert np.abs(Vi @ np.array(a["view"], np.float64).reshape(4, 4).T - np.eye(4)).max() < 1e-5
    b = sv.camera_arrays(cs[1])
    assert b["viewport"].tolist() == [100, 80] and b["focal"].tolist() == [50, 50]
    P = np.array(b["proj"], np.float64).reshape(4, 4).T
    assert np.allclose(P[0], [1.0, 0, 2 * 40.0 / 100 - 1, 0]) and np.allclose(P[1], [0, -1.25, 1 - 2 * 30.0 / 80, 0])
    g, ext = sv.points_to_gaussians(xyz, rgb)
    assert np.array_equal(g["sh"][:, :3], f16((rgb / 255.0 - 0.5) / SH_C0))
    assert np.all(g["quat"] == np.array([0, 0, 1, 0], np.float32)) and abs(ext - 1.8) < 1e-6


def test_colmap_rejects_non_pinhole(sv, tmp_path):
    _write_colmap(tmp_path, [(1, 4, 10, 10, [1.0] * 8)], [(1, [1.0, 0, 0, 0], [0, 0, 1], 1, "x", 0)], [])
    with pytest.raises(sv.SvError, match="not a pinhole"):
        sv.colmap_load(str(tmp_path))
