"""Synthetic surfel scenes and cameras for the benchmark configurations.

Distributions follow the benchmark definitions: centres U[-1, 1]^3, two
log-normal scales per surfel, quaternions uniform on S^3 (normalised
N(0, 1)^4), opacity sigmoid(N(0, 1)), colour RGB U[0, 1] stored as the SH DC
coefficient (rgb - 0.5) / C0 (surfel_prep_fwd.wgsl colour = 0.5 + C0 * dc),
higher SH bands N(0, 0.2). Cameras use the reference conventions:
right-handed, y up, looking down -z (crates/gs-core/src/camera.rs:40-49
look_at) with focal 0.5 H / tan(30 deg) (the 60-degree default FOV,
camera.rs:57, focal_px :28-32).

Seeds: numpy PCG64 seeded 0x5EED0000 + config index; the target scene of a
config jitters the centres with seed ^ 1.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

SH_C0 = 0.28209479177387814


@dataclass
class SceneArrays:
    positions: np.ndarray   # (n, 3) f32
    scales: np.ndarray      # (n, 2) f32, activated
    quats: np.ndarray       # (n, 4) f32, xyzw
    opacities: np.ndarray   # (n,) f32, activated
    sh: np.ndarray          # (n, (deg+1)^2 * 3) f32, coefficient-major rgb
    sh_degree: int

    @property
    def n(self) -> int:
        return int(self.positions.shape[0])

    def as_init(self) -> dict:
        return dict(positions=self.positions, scales=self.scales, quats=self.quats,
                    opacities=self.opacities, sh=self.sh)


def make_scene(n: int, sh_degree: int, log_scale_mu: float, log_scale_sigma: float, seed: int,
               rest_sigma: float = 0.2) -> SceneArrays:
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-1.0, 1.0, (n, 3))
    scales = np.exp(rng.normal(log_scale_mu, log_scale_sigma, (n, 2)))
    q = rng.normal(0.0, 1.0, (n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    op = 1.0 / (1.0 + np.exp(-rng.normal(0.0, 1.0, n)))
    ncoef = (sh_degree + 1) ** 2
    sh = np.zeros((n, ncoef, 3))
    rgb = rng.uniform(0.0, 1.0, (n, 3))
    sh[:, 0, :] = (rgb - 0.5) / SH_C0
    if ncoef > 1:
        sh[:, 1:, :] = rng.normal(0.0, rest_sigma, (n, ncoef - 1, 3))
    f = np.float32
    return SceneArrays(pos.astype(f), scales.astype(f), q.astype(f), op.astype(f),
                       sh.reshape(n, ncoef * 3).astype(f), sh_degree)


def jittered(scene: SceneArrays, sigma: float, seed: int) -> SceneArrays:
    rng = np.random.default_rng(seed)
    pos = scene.positions.astype(np.float64) + rng.normal(0.0, sigma, scene.positions.shape)
    return SceneArrays(pos.astype(np.float32), scene.scales, scene.quats, scene.opacities, scene.sh,
                       scene.sh_degree)


def quat_from_mat(m: np.ndarray) -> np.ndarray:
    """Unit quaternion (xyzw) of a rotation matrix (Shepperd's method)."""
    t = np.trace(m)
    if t > 0:
        s = np.sqrt(t + 1.0) * 2
        w = 0.25 * s
        x = (m[2, 1] - m[1, 2]) / s
        y = (m[0, 2] - m[2, 0]) / s
        z = (m[1, 0] - m[0, 1]) / s
    elif m[0, 0] > m[1, 1] and m[0, 0] > m[2, 2]:
        s = np.sqrt(1.0 + m[0, 0] - m[1, 1] - m[2, 2]) * 2
        w = (m[2, 1] - m[1, 2]) / s
        x = 0.25 * s
        y = (m[0, 1] + m[1, 0]) / s
        z = (m[0, 2] + m[2, 0]) / s
    elif m[1, 1] > m[2, 2]:
        s = np.sqrt(1.0 + m[1, 1] - m[0, 0] - m[2, 2]) * 2
        w = (m[0, 2] - m[2, 0]) / s
        x = (m[0, 1] + m[1, 0]) / s
        y = 0.25 * s
        z = (m[1, 2] + m[2, 1]) / s
    else:
        s = np.sqrt(1.0 + m[2, 2] - m[0, 0] - m[1, 1]) * 2
        w = (m[1, 0] - m[0, 1]) / s
        x = (m[0, 2] + m[2, 0]) / s
        y = (m[1, 2] + m[2, 1]) / s
        z = 0.25 * s
    q = np.array([x, y, z, w])
    return q / np.linalg.norm(q)


def look_at(eye, target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0)) -> np.ndarray:
    """camera.rs:40-49: camera-to-world quaternion looking from eye to target."""
    eye = np.asarray(eye, float)
    fwd = np.asarray(target, float) - eye
    fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, np.asarray(up, float))
    right /= np.linalg.norm(right)
    up2 = np.cross(right, fwd)
    return quat_from_mat(np.stack([right, up2, -fwd], axis=1))


def focal_for(height: int, fov_y_deg: float = 60.0) -> float:
    """camera.rs:28-32 focal_px with the 60-degree default FOV."""
    return 0.5 * height / np.tan(np.deg2rad(fov_y_deg) * 0.5)


@dataclass
class ViewSpec:
    width: int
    height: int
    center: np.ndarray
    quat: np.ndarray
    focal: float


def front_view(width: int, height: int, dist: float = 3.0) -> ViewSpec:
    eye = np.array([0.0, 0.0, -dist])
    return ViewSpec(width, height, eye, look_at(eye), focal_for(height))


def sphere_views(count: int, seed: int, sizes, radius: float = 3.0):
    """Cameras on a sphere of radius 3 looking at the origin, yaw U[0, 2pi),
    pitch U[-pi/3, pi/3]; sizes is a list of (w, h) (cycled)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        yaw = rng.uniform(0.0, 2 * np.pi)
        pitch = rng.uniform(-np.pi / 3, np.pi / 3)
        eye = radius * np.array([np.cos(pitch) * np.sin(yaw), np.sin(pitch), np.cos(pitch) * np.cos(yaw)])
        w, h = sizes[k % len(sizes)]
        out.append(ViewSpec(w, h, eye, look_at(eye), focal_for(h)))
    return out


def _quat_mul(a, b):
    """Hamilton product of xyzw quaternions."""
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    return np.array([aw * bx + ax * bw + ay * bz - az * by,
                     aw * by - ax * bz + ay * bw + az * bx,
                     aw * bz + ax * by - ay * bx + az * bw,
                     aw * bw - ax * bx - ay * by - az * bz])


def walk_views(count: int, seed: int, size, step_m: float = 0.01, step_deg: float = 0.5,
               radius: float = 3.0):
    """A smoothly moving camera: a start pose on the sphere of `radius`
    looking at the origin, then an SE(3) random walk, each frame moving
    `step_m` along a random direction and turning `step_deg` about a random
    camera-frame axis (the streaming-frame configuration: 1 cm and 0.5 deg
    per frame). size = (w, h) of every frame."""
    rng = np.random.default_rng(seed)
    yaw = rng.uniform(0.0, 2 * np.pi)
    pitch = rng.uniform(-np.pi / 6, np.pi / 6)
    eye = radius * np.array([np.cos(pitch) * np.sin(yaw), np.sin(pitch), np.cos(pitch) * np.cos(yaw)])
    q = look_at(eye)
    w, h = size
    out = []
    for _ in range(count):
        out.append(ViewSpec(w, h, eye.copy(), q.copy(), focal_for(h)))
        d = rng.normal(size=3)
        eye = eye + step_m * d / np.linalg.norm(d)
        axis = rng.normal(size=3)
        axis /= np.linalg.norm(axis)
        half = 0.5 * np.deg2rad(step_deg)
        q = _quat_mul(q, np.concatenate([np.sin(half) * axis, [np.cos(half)]]))
        q /= np.linalg.norm(q)
    return out


def frame_surfel_counts(count: int, base: int, growth: float = 0.1):
    """Surfels seen by frame k of a growing scene: base plus
    round(growth * base) newly densified surfels per frame (frame k renders
    surfels [0, base + k round(growth base)))."""
    step = int(round(growth * base))
    return [base + k * step for k in range(count)]


def uneven_sizes(count: int, seed: int):
    """16:9 sizes between 640x360 and 1920x1080."""
    rng = np.random.default_rng(seed)
    hs = rng.integers(360, 1081, count)
    return [(int(round(h * 16 / 9)), int(h)) for h in hs]


CONFIG_SEED = 0x5EED0000


def config_scene(index: int, n: int | None = None) -> SceneArrays:
    """Scene of benchmark configuration `index` (0..4); n overrides the count."""
    seed = CONFIG_SEED + index
    if index == 0:
        return make_scene(n or 10_000, 0, -4.0, 0.5, seed)
    if index in (1, 2):
        return make_scene(n or 100_000, 3, -5.0, 0.5, seed)
    if index == 3:
        return make_scene(n or 1_000_000, 3, -5.0, 1.0, seed)
    if index == 4:
        # The base scene plus 15 frames of 10 % densification (the scene at
        # the last of 16 frames); frame k renders a prefix of it.
        return make_scene(n or frame_surfel_counts(16, 500_000)[-1], 3, -5.0, 0.5, seed)
    raise ValueError(index)


def _swatch(face: int, i: int, j: int):
    """A stable pseudo-random mid-bright colour per (face, i, j)
    (crates/gs-map/src/synthetic.rs:68-82)."""
    h = ((face * 0x9E3779B9) ^ (i * 0x85EBCA6B) ^ (j * 0xC2B2AE35)) & 0xFFFFFFFF
    out = []
    for _ in range(3):
        h ^= h >> 15
        h = (h * 0x2545F491) & 0xFFFFFFFF
        h ^= h >> 13
        out.append(np.float32(0.25) + np.float32(0.6) * np.float32((h >> 8) / float(1 << 24)))
    return out


def _quat_arc_from_z(n):
    """Unit quaternion (xyzw) rotating +z onto the unit vector n (glam
    Quat::from_rotation_arc); for n = -z, a half turn about x."""
    z = np.array([0.0, 0.0, 1.0])
    d = float(np.dot(z, n))
    if d < -1.0 + 1e-6:
        return np.array([1.0, 0.0, 0.0, 0.0])
    c = np.cross(z, n)
    q = np.array([c[0], c[1], c[2], 1.0 + d])
    return q / np.linalg.norm(q)


def room(per_side: int = 24, size: float = 3.0) -> SceneArrays:
    """The tracking tests' synthetic room (crates/gs-map/src/synthetic.rs:22-66):
    six walls of per_side x per_side round surfels (radius 0.75 x the spacing,
    opacity 0.95, normals facing inward, hashed colours), SH degree 0."""
    n = max(per_side, 2)
    step = 2.0 * size / (n - 1)
    radius = step * 0.75
    pos, quats, cols = [], [], []
    faces = [((1, 0, 0), 1.0), ((1, 0, 0), -1.0), ((0, 1, 0), 1.0), ((0, 1, 0), -1.0), ((0, 0, 1), 1.0),
             ((0, 0, 1), -1.0)]
    for fi, (axis, sign) in enumerate(faces):
        axis = np.array(axis, np.float64)
        u = np.array([0.0, 1.0, 0.0]) if axis[0] != 0.0 else np.array([1.0, 0.0, 0.0])
        v = np.cross(axis, u)
        v /= np.linalg.norm(v)
        q = _quat_arc_from_z(-axis * sign)
        for i in range(n):
            for j in range(n):
                a = -size + i * step
                b = -size + j * step
                pos.append(axis * sign * size + u * a + v * b)
                quats.append(q)
                cols.append(_swatch(fi, i, j))
    m = len(pos)
    rgb = np.array(cols, np.float64)
    return SceneArrays(positions=np.array(pos, np.float32), scales=np.full((m, 2), radius, np.float32),
                       quats=np.array(quats, np.float32), opacities=np.full(m, 0.95, np.float32),
                       sh=((rgb - 0.5) / SH_C0).astype(np.float32), sh_degree=0)


def quat_from_rotation_y(angle: float) -> np.ndarray:
    return np.array([0.0, np.sin(angle / 2), 0.0, np.cos(angle / 2)])


def quat_from_rotation_x(angle: float) -> np.ndarray:
    return np.array([np.sin(angle / 2), 0.0, 0.0, np.cos(angle / 2)])


def quat_mul(a, b) -> np.ndarray:
    return _quat_mul(a, b)
