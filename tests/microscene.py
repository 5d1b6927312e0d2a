"""Deterministic micro-scenes for parity tests.

Restates the reference's scene generators so parity tests run on the same
scenes the reference's own tests use:
  - ``make_scene`` of crates/gs-cpu-ref/tests/gradcheck.rs:23-85 (focal 30)
    and crates/gs-kernels/tests/common/mod.rs:20-82 (focal 0.9*size, scales
    U[0.1, 0.5]) -- selected by ``flavor``;
  - xorshift64 (shifts 13/7/17, (x >> 11) / 2^53) of common/mod.rs:7-18;
  - xorshift32 (shifts 13/17/5) of crates/gs-wgpu/tests/sort.rs:18-25.
"""

from __future__ import annotations

import numpy as np

from oracle import Camera, Scene, quat_to_mat

_M64 = (1 << 64) - 1
_M32 = (1 << 32) - 1


class XorShift64:
    """common/mod.rs:7-18."""

    def __init__(self, seed: int):
        self.s = seed & _M64

    def next(self) -> float:
        x = self.s
        x ^= (x << 13) & _M64
        x ^= x >> 7
        x ^= (x << 17) & _M64
        self.s = x
        return (x >> 11) / float(1 << 53)

    def uniform(self, lo: float, hi: float) -> float:
        return lo + (hi - lo) * self.next()


def xorshift32_stream(seed: int, n: int) -> np.ndarray:
    """sort.rs:18-25, vectorised over n draws (exact u32 arithmetic)."""
    out = np.empty(n, dtype=np.uint32)
    x = seed & _M32
    for i in range(n):
        x ^= (x << 13) & _M32
        x ^= x >> 17
        x ^= (x << 5) & _M32
        out[i] = x
    return out


def make_scene(seed: int, n_surfels: int, sh_degree: int, size: int, flavor: str = "gradcheck"):
    """gradcheck.rs:23-85 (flavor 'gradcheck') / common/mod.rs:20-82 ('parity')."""
    rng = XorShift64(seed)
    center = np.array([rng.uniform(-0.3, 0.3) for _ in range(3)])
    quat = np.array([rng.uniform(-0.08, 0.08), rng.uniform(-0.08, 0.08),
                     rng.uniform(-0.08, 0.08), 1.0])
    focal = 30.0 if flavor == "gradcheck" else size * 0.9
    smax = 0.4 if flavor == "gradcheck" else 0.5
    cam = Camera(center, quat, focal, size, size)
    r = quat_to_mat(quat)
    ncoef = (sh_degree + 1) ** 2
    pos = np.zeros((n_surfels, 3))
    scales = np.zeros((n_surfels, 2))
    quats = np.zeros((n_surfels, 4))
    opacity = np.zeros(n_surfels)
    sh = np.zeros((n_surfels, ncoef, 3))
    for i in range(n_surfels):
        z = -rng.uniform(2.0, 6.0)
        extent = -z * (size * 0.5) / focal * 0.7
        p_cam = np.array([rng.uniform(-extent, extent), rng.uniform(-extent, extent), z])
        sh[i, 0] = [rng.uniform(-0.8, 0.8) for _ in range(3)]
        for k in range(1, ncoef):
            sh[i, k] = [rng.uniform(-0.25, 0.25) for _ in range(3)]
        pos[i] = center + r @ p_cam
        scales[i] = [rng.uniform(0.1, smax), rng.uniform(0.1, smax)]
        quats[i] = [rng.uniform(-1.0, 1.0), rng.uniform(-1.0, 1.0), rng.uniform(-1.0, 1.0),
                    rng.uniform(-1.0, 1.0) + 1.5]
        opacity[i] = rng.uniform(0.25, 0.85)
    return Scene(pos, scales, quats, opacity, sh, sh_degree), cam


def make_weights(seed: int, n: int) -> np.ndarray:
    """gradcheck.rs:87-98: per-pixel rgb weights U[-1, 1]."""
    rng = XorShift64(seed)
    return np.array([[rng.uniform(-1.0, 1.0) for _ in range(3)] for _ in range(n)])


def make_uniform(seed: int, n: int, lo: float, hi: float) -> np.ndarray:
    rng = XorShift64(seed)
    return np.array([rng.uniform(lo, hi) for _ in range(n)])


def scene_to_f32(scene: Scene):
    """common/mod.rs:95-134 (to_gpu): f32 arrays with the reference GPU layout
    (pos vec4 w=1, SH padded to 48 floats)."""
    n = scene.n
    pos4 = np.zeros((n, 4), dtype=np.float32)
    pos4[:, :3] = scene.pos
    pos4[:, 3] = 1.0
    scales = scene.scales.astype(np.float32)
    quats = scene.quat.astype(np.float32)
    opac = scene.opacity.astype(np.float32)
    sh48 = np.zeros((n, 48), dtype=np.float32)
    flat = scene.sh.reshape(n, -1).astype(np.float32)
    sh48[:, :flat.shape[1]] = flat
    return pos4, scales, quats, opac, sh48


def scene_from_f32(scene: Scene) -> Scene:
    """The same scene rounded to f32 and promoted back: what the GPU sees."""
    return Scene(scene.pos.astype(np.float32).astype(np.float64),
                 scene.scales.astype(np.float32).astype(np.float64),
                 scene.quat.astype(np.float32).astype(np.float64),
                 scene.opacity.astype(np.float32).astype(np.float64),
                 scene.sh.astype(np.float32).astype(np.float64), scene.sh_degree)


def camera_from_f32(cam: Camera) -> Camera:
    return Camera(cam.center.astype(np.float32).astype(np.float64),
                  cam.quat.astype(np.float32).astype(np.float64),
                  float(np.float32(cam.focal)), cam.width, cam.height)
