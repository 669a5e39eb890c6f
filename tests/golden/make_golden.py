#!/usr/bin/env python3
"""Write the golden fixtures of this directory from the CPU oracle.

    python tests/golden/make_golden.py

Inputs come from the seeded scene generator, restated here in numpy
(`gen_gaussians`, `gen_jitter`, `gen_camera`) and checked bit for bit
against the library's C++ generator by tests/test_golden.py. Outputs are
the oracle's (oracle/liboracle.so): the full tile build, image and per-pixel
state of the adjoint-check scene (4,096 Gaussians, 64x64), gradients of 64
sampled Gaussians in full, and SHA-256 digests of every array of both the
adjoint-check scene and the small scene (16,384 Gaussians, 256x256), so a
change of one bit anywhere shows up.
"""
from __future__ import annotations

import hashlib
import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

M64 = (1 << 64) - 1


def _mix64(seed: int, ctr: np.ndarray) -> np.ndarray:
    """SplitMix64 at counter ctr (uint64 arithmetic wraps)."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (ctr.astype(np.uint64) + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, ctr) -> np.ndarray:
    return (_mix64(seed, np.asarray(ctr)) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed: int, ctr) -> np.ndarray:
    ctr = np.asarray(ctr, np.uint64)
    u1 = uniform(seed, np.uint64(2) * ctr)
    u2 = uniform(seed, np.uint64(2) * ctr + np.uint64(1))
    return np.sqrt(-2.0 * np.log(1.0 - u1)) * np.cos(2.0 * math.pi * u2)


def f16(x) -> np.ndarray:
    with np.errstate(over="ignore"):
        return np.asarray(x, np.float32).astype(np.float16).astype(np.float32)


def gen_gaussians(seed: int, n: int, half_extent: float = 1.0) -> dict:
    b = np.arange(n, dtype=np.uint64)[:, None] * np.uint64(64)
    pos = f16(((2.0 * uniform(seed, b + np.arange(3, dtype=np.uint64)) - 1.0) * half_extent).astype(np.float32))
    ls = f16((-4.5 + 0.3 * normal(seed, b + np.uint64(3) + np.arange(3, dtype=np.uint64))).astype(np.float32))
    q = normal(seed, b + np.uint64(6) + np.arange(4, dtype=np.uint64))
    qn = np.sqrt((q * q).sum(1, keepdims=True))
    quat = f16((q / qn).astype(np.float32))
    logit = f16(normal(seed, b[:, 0] + np.uint64(10)).astype(np.float32))
    z = normal(seed, b + np.uint64(11) + np.arange(48, dtype=np.uint64))
    sh = np.where(np.arange(48)[None, :] < 3, 0.5 + 0.2 * z, 0.05 * z)
    return {"pos": pos, "logscale": ls, "quat": quat, "logit": logit, "sh": f16(sh.astype(np.float32))}


def gen_jitter(seed: int, pos: np.ndarray, sigma: float) -> np.ndarray:
    n = pos.shape[0]
    ctr = np.uint64(3) * np.arange(n, dtype=np.uint64)[:, None] + np.arange(3, dtype=np.uint64)
    return f16((pos.astype(np.float64) + np.float64(np.float32(sigma)) * normal(seed, ctr)).astype(np.float32))


def gen_camera(eye, target, w: int, h: int, fovy_deg: float = 60.0) -> dict:
    """Look-at view (rows right, up, forward; +z forward) and the pinhole
    projection of training-camera.ts:29-52, column-major float32[16]."""
    e = np.asarray(eye, np.float32).astype(np.float64)
    f = np.asarray(target, np.float32).astype(np.float64) - e
    f /= math.sqrt(f @ f)
    up = np.array([1.0, 0.0, 0.0]) if abs(f[1]) > 0.999 else np.array([0.0, 1.0, 0.0])
    r = np.cross(f, up)
    r /= math.sqrt(r @ r)
    u = np.cross(r, f)
    V = np.zeros(16)
    for row, vec in enumerate((r, u, f)):
        V[[0 + row, 4 + row, 8 + row]] = vec
        V[12 + row] = -(vec[0] * e[0] + vec[1] * e[1] + vec[2] * e[2])
    V[15] = 1.0
    fy = (0.5 * h) / math.tan(0.5 * fovy_deg * math.pi / 180.0)
    P = np.zeros(16)
    P[0], P[5] = 2.0 * fy / w, -2.0 * fy / h
    P[8], P[9] = 2.0 * (0.5 * w) / w - 1.0, 1.0 - 2.0 * (0.5 * h) / h
    P[10] = P[11] = 1.0
    return {"view": V.astype(np.float32), "proj": P.astype(np.float32),
            "viewport": np.array([w, h], np.float32), "focal": np.array([fy, fy], np.float32)}


class _Cam:
    """Minimal camera object with the fields the oracle binding copies."""

    def __init__(self, d: dict):
        import ctypes as C

        import oracle as orc

        self.s = orc.OrcCamera()
        for k in ("view", "proj", "viewport", "focal"):
            arr = getattr(self.s, k)
            for i, v in enumerate(d[k]):
                arr[i] = float(v)
        V = d["view"].astype(np.float64).reshape(4, 4).T
        Vi = np.linalg.inv(V).T.ravel().astype(np.float32)
        for i, v in enumerate(Vi):
            self.s.view_inv[i] = float(v)
        self._C = C


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


SCENES = {
    # name: (n, w, h, eye)
    "small256": (16384, 256, 256, (0.0, 0.0, 3.0)),
    "adjoint64": (4096, 64, 64, (0.0, 0.0, 3.0)),
}
ARRAYS = ("keys", "vals", "offsets", "color", "alpha", "last_ids", "residual")
GRADS = ("pos", "logscale", "quat", "logit", "sh", "geom")


def run_oracle(name: str):
    import oracle as orc

    n, w, h, eye = SCENES[name]
    g = gen_gaussians(0, n, 1.0)
    tg = dict(g)
    tg["pos"] = gen_jitter(1, g["pos"], 0.02)
    cam = _Cam(gen_camera(eye, (0, 0, 0), w, h, 60.0)).s
    view = orc.view_params(w, h, sh_deg=3)
    target = orc.render_and_grad(tg, cam, view)["color"]
    r = orc.render_and_grad(g, cam, view, target=target, w_l2=0.8, w_l1=0.2)
    return g, target, r


def main():
    manifest = {}
    for name in SCENES:
        g, target, r = run_oracle(name)
        ent = {"n_isect": r["n_isect"], "inputs": {k: digest(v) for k, v in g.items()},
               "target": digest(target), "outputs": {k: digest(r[k]) for k in ARRAYS},
               "grads": {k: digest(r["grads"][k]) for k in GRADS}}
        manifest[name] = ent
        if name == "adjoint64":
            pick = np.random.default_rng(3).choice(np.flatnonzero(r["proj"]["tiles_count"] > 0), 64, replace=False)
            np.savez_compressed(os.path.join(HERE, "adjoint64_4096_64x64.npz"),
                                target=target, pick=pick.astype(np.int64),
                                **{k: r[k] for k in ARRAYS},
                                **{"grad_" + k: r["grads"][k][pick] for k in GRADS})
    with open(os.path.join(HERE, "digests.json"), "w") as fh:
        json.dump(manifest, fh, indent=1, sort_keys=True)
        fh.write("\n")
    print("wrote", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    main()
