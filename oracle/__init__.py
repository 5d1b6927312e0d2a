"""CPU oracle for the 2DGS surfel training step -- test infrastructure only.

A ctypes view of ``oracle/liboracle.so`` (built from ``cpuref.cpp`` by
``oracle/Makefile``). Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package; the product
(``splattar_amd``) never does.

The C++ restates, function by function, the reference oracle
``crates/gs-cpu-ref/src/{forward,backward,math,scene}.rs`` in FP64, plus an
fp32 mirror of the integer stages, the L1 + D-SSIM loss of
``first_attempt/crates/gs-train/src/shaders/ssim.wgsl`` and the Adam step of
``first_attempt/crates/gs-train/src/shaders/adam.wgsl``. See ``cpuref.h`` for
the per-function citations and the checks that pin each of them.
"""

from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_i64p = C.POINTER(C.c_int64)


class _Scene(C.Structure):
    _fields_ = [("n", C.c_int64), ("sh_degree", C.c_int32), ("pos", _dp), ("scales", _dp),
                ("quat", _dp), ("opacity", _dp), ("sh", _dp)]


class _Camera(C.Structure):
    _fields_ = [("center", C.c_double * 3), ("quat", C.c_double * 4), ("focal", C.c_double),
                ("width", C.c_int32), ("height", C.c_int32), ("crop_x0", C.c_int32),
                ("crop_y0", C.c_int32), ("crop_w", C.c_int32), ("crop_h", C.c_int32)]


class _Grads(C.Structure):
    _fields_ = [("pos", _dp), ("scales", _dp), ("quat", _dp), ("opacity", _dp), ("sh", _dp),
                ("cam_center", C.c_double * 3), ("cam_quat", C.c_double * 4), ("focal", C.c_double)]


class _Cam32(C.Structure):
    _fields_ = [("rot", C.c_float * 9), ("center", C.c_float * 3), ("focal", C.c_float),
                ("width", C.c_uint32), ("height", C.c_uint32), ("tiles_x", C.c_uint32),
                ("tiles_y", C.c_uint32)]


def use_library(path: str):
    """Load the oracle from `path` (a build of the same cpuref.cpp, e.g. one
    compiled for the host's own CPU) instead of oracle/liboracle.so; only
    before the first call into the oracle."""
    global _LIB_PATH
    if _lib is not None:
        raise RuntimeError("oracle library already loaded")
    _LIB_PATH = path


def build_info() -> str:
    return lib().oref_build_info().decode()


def lib():
    """Load (once) and return the oracle shared library."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run `make -C oracle`)")
        L = C.CDLL(_LIB_PATH)
        L.oref_quat_to_mat.argtypes = [_dp, _dp]
        L.oref_quat_grad.argtypes = [_dp, _dp, _dp]
        L.oref_sh_basis.argtypes = [C.c_int, _dp, _dp]
        L.oref_sh_basis_grad.argtypes = [C.c_int, _dp, _dp]
        L.oref_render.argtypes = [C.POINTER(_Scene), C.POINTER(_Camera), C.c_int, _dp, _dp, _dp,
                                  _dp, _i64p]
        L.oref_distortion_loss.argtypes = [C.POINTER(_Scene), C.POINTER(_Camera), C.c_int]
        L.oref_distortion_loss.restype = C.c_double
        L.oref_gradients.argtypes = [C.POINTER(_Scene), C.POINTER(_Camera), C.c_int, _dp, _dp,
                                     _dp, C.c_double, C.POINTER(_Grads)]
        L.oref_marginal_decisions.argtypes = [C.POINTER(_Scene), C.POINTER(_Camera), C.c_double,
                                              C.POINTER(C.c_uint8)]
        L.oref_marginal_decisions.restype = C.c_int64
        L.oref_pixel_contributors.argtypes = [C.POINTER(_Scene), C.POINTER(_Camera), C.c_int64,
                                              C.POINTER(C.c_int32), C.POINTER(C.c_uint8)]
        L.oref_prepare_rects.argtypes = [C.POINTER(_Scene), C.POINTER(_Camera), C.POINTER(C.c_int32),
                                         C.POINTER(C.c_uint8), C.POINTER(C.c_uint8)]
        L.oref_normal_loss.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, _dp, _dp, _dp,
                                       _dp, _dp, _dp]
        L.oref_trace.argtypes = [C.POINTER(_Scene), C.POINTER(_Camera), C.c_int64, _u32p, _u32p]
        L.oref_replay.argtypes = [C.POINTER(_Scene), C.POINTER(_Camera), _u32p, _u32p, C.c_int64, _dp,
                                  _dp, _dp, _dp, _dp, C.c_double, C.POINTER(_Grads)]
        L.oref_prepare_order.argtypes = [C.POINTER(_Scene), C.POINTER(_Camera), _i64p]
        L.oref_prepare_order.restype = C.c_int64
        L.oref_prep_f32.argtypes = [C.c_int64, _fp, _fp, _fp, C.POINTER(_Cam32), _u32p, _u32p,
                                    _u32p]
        L.oref_exclusive_scan_u32.argtypes = [C.c_int64, _u32p, _u32p]
        L.oref_exclusive_scan_u32.restype = C.c_uint32
        L.oref_emit_keys.argtypes = [C.c_int64, _u32p, _u32p, C.c_uint32, C.c_uint32, _u64p,
                                     _u32p, C.c_int64]
        L.oref_emit_keys.restype = C.c_int64
        L.oref_stable_sort_u64.argtypes = [C.c_int64, _u64p, _u32p]
        L.oref_tile_ranges.argtypes = [C.c_int64, _u64p, C.c_uint32, C.c_uint32, _u32p]
        L.oref_l1_ssim.argtypes = [C.c_int, C.c_int, _dp, _dp, C.c_double, C.c_double, _dp, _dp,
                                   _dp]
        L.oref_l1_ssim.restype = C.c_double
        L.oref_blur.argtypes = [C.c_int, C.c_int, _dp, _dp]
        L.oref_adam_step_f32.argtypes = [C.c_int64, _fp, _fp, _fp, _fp, _fp, C.c_float, C.c_float,
                                         C.c_int, C.c_int]
        L.oref_num_threads.restype = C.c_int
        L.oref_build_info.restype = C.c_char_p
        L.oref_xorshift32_fill.argtypes = [C.c_uint32, C.c_int64, C.c_uint32, _u32p]
        _lib = L
    return _lib


def _ptr(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# --------------------------------------------------------------- scene types

@dataclass
class Scene:
    """FP64 micro-scene (crates/gs-cpu-ref/src/scene.rs:23-52).

    sh is (n, ncoef, 3), coefficient-major rgb.
    """
    pos: np.ndarray
    scales: np.ndarray
    quat: np.ndarray
    opacity: np.ndarray
    sh: np.ndarray
    sh_degree: int

    @property
    def n(self) -> int:
        return int(self.pos.shape[0])

    def copy(self) -> "Scene":
        return Scene(self.pos.copy(), self.scales.copy(), self.quat.copy(), self.opacity.copy(),
                     self.sh.copy(), self.sh_degree)

    def _c(self):
        keep = (_d(self.pos), _d(self.scales), _d(self.quat), _d(self.opacity), _d(self.sh))
        s = _Scene(self.n, self.sh_degree, *[_ptr(k, _dp) for k in keep])
        return s, keep


@dataclass
class Camera:
    """FP64 camera (scene.rs:36-45): camera-to-world quat xyzw, center, focal."""
    center: np.ndarray
    quat: np.ndarray
    focal: float
    width: int
    height: int
    crop: tuple | None = None   # (x0, y0, w, h) pixel window; None = whole image

    def copy(self) -> "Camera":
        return Camera(self.center.copy(), self.quat.copy(), float(self.focal), self.width,
                      self.height, self.crop)

    @property
    def out_size(self) -> tuple[int, int]:
        """(width, height) of the evaluated window."""
        if self.crop is None:
            return self.width, self.height
        x0, y0, w, h = self.crop
        return min(w, self.width - x0), min(h, self.height - y0)

    def _c(self):
        c = _Camera()
        for k in range(3):
            c.center[k] = float(self.center[k])
        for k in range(4):
            c.quat[k] = float(self.quat[k])
        c.focal = float(self.focal)
        c.width = int(self.width)
        c.height = int(self.height)
        if self.crop is not None:
            c.crop_x0, c.crop_y0, c.crop_w, c.crop_h = (int(v) for v in self.crop)
        return c


@dataclass
class Gradients:
    """Gradients of every parameter class (scene.rs:66-97)."""
    pos: np.ndarray
    scales: np.ndarray
    quat: np.ndarray
    opacity: np.ndarray
    sh: np.ndarray
    cam_center: np.ndarray = field(default_factory=lambda: np.zeros(3))
    cam_quat: np.ndarray = field(default_factory=lambda: np.zeros(4))
    focal: float = 0.0


# ------------------------------------------------------------------ math.rs

def quat_to_mat(q) -> np.ndarray:
    """math.rs:14-34; returns R (3x3, R[:, k] = image of e_k)."""
    out = np.zeros(9)
    lib().oref_quat_to_mat(_ptr(_d(q), _dp), _ptr(out, _dp))
    return out.reshape(3, 3).T.copy()


def quat_grad(q, dl_dr: np.ndarray) -> np.ndarray:
    """math.rs:38-88; dl_dr is 3x3 with dl_dr[j, k] = dL/dR[j][k]."""
    cm = _d(np.asarray(dl_dr, dtype=np.float64).T.reshape(9))
    out = np.zeros(4)
    lib().oref_quat_grad(_ptr(_d(q), _dp), _ptr(cm, _dp), _ptr(out, _dp))
    return out


def sh_basis(degree: int, d) -> np.ndarray:
    out = np.zeros(16)
    k = lib().oref_sh_basis(degree, _ptr(_d(d), _dp), _ptr(out, _dp))
    return out[:k]


def sh_basis_grad(degree: int, d) -> np.ndarray:
    out = np.zeros(48)
    k = lib().oref_sh_basis_grad(degree, _ptr(_d(d), _dp), _ptr(out, _dp))
    return out[:3 * k].reshape(k, 3)


# -------------------------------------------------------- forward / backward

def render(scene: Scene, cam: Camera, mode: int = 1, with_hits: bool = False):
    """forward.rs:207-246. Returns dict color (H,W,3), alpha, depth, normal."""
    W, H = cam.out_size
    P = W * H
    color = np.zeros(P * 3)
    alpha = np.zeros(P)
    depth = np.zeros(P)
    normal = np.zeros(P * 3)
    hits = np.zeros(1, dtype=np.int64)
    s, keep = scene._c()
    c = cam._c()
    rc = lib().oref_render(C.byref(s), C.byref(c), mode, _ptr(color, _dp), _ptr(alpha, _dp),
                           _ptr(depth, _dp), _ptr(normal, _dp), _ptr(hits, _i64p))
    if rc != 0:
        raise RuntimeError("oref_render failed")
    out = dict(color=color.reshape(H, W, 3), alpha=alpha.reshape(H, W),
               depth=depth.reshape(H, W), normal=normal.reshape(H, W, 3))
    if with_hits:
        out["hits"] = int(hits[0])
    return out


def distortion_loss(scene: Scene, cam: Camera, mode: int = 1) -> float:
    s, keep = scene._c()
    c = cam._c()
    return float(lib().oref_distortion_loss(C.byref(s), C.byref(c), mode))


def gradients(scene: Scene, cam: Camera, dl_dcolor, dl_ddepth=None, dl_dnormal=None,
              lambda_dist: float = 0.0, mode: int = 1) -> Gradients:
    """backward.rs:57-249 (gradients_full)."""
    n = scene.n
    ncoef = (scene.sh_degree + 1) ** 2
    g = Gradients(np.zeros((n, 3)), np.zeros((n, 2)), np.zeros((n, 4)), np.zeros(n),
                  np.zeros((n, ncoef, 3)))
    cg = _Grads(_ptr(g.pos, _dp), _ptr(g.scales, _dp), _ptr(g.quat, _dp), _ptr(g.opacity, _dp),
                _ptr(g.sh, _dp))
    s, keep = scene._c()
    c = cam._c()
    dlc = _d(dl_dcolor).reshape(-1)
    dld = _d(dl_ddepth).reshape(-1) if dl_ddepth is not None else None
    dln = _d(dl_dnormal).reshape(-1) if dl_dnormal is not None else None
    rc = lib().oref_gradients(C.byref(s), C.byref(c), mode, _ptr(dlc, _dp), _ptr(dld, _dp),
                              _ptr(dln, _dp), float(lambda_dist), C.byref(cg))
    if rc != 0:
        raise RuntimeError("oref_gradients failed")
    g.cam_center = np.array(cg.cam_center[:])
    g.cam_quat = np.array(cg.cam_quat[:])
    g.focal = float(cg.focal)
    return g


def marginal_decisions(scene: Scene, cam: Camera, rel: float = 1e-5):
    """Test aid with no reference counterpart: (count, flags) of the forward
    walk's discrete decisions within relative band `rel` of a threshold
    (alpha skip / clamp, ray vs screen branch, near plane, transmittance stop,
    16-px rect truncation, on-screen, near cull, f32 depth-key ties), flags[i]
    = 1 for surfels involved. With count 0, an f32 evaluation whose relative
    error stays below rel makes the same decisions as this FP64 one."""
    s, keep = scene._c()
    c = cam._c()
    flags = np.zeros(scene.n, dtype=np.uint8)
    k = lib().oref_marginal_decisions(C.byref(s), C.byref(c), float(rel),
                                      flags.ctypes.data_as(C.POINTER(C.c_uint8)))
    if k < 0:
        raise RuntimeError("oref_marginal_decisions failed")
    return int(k), flags


def pixel_contributors(scene: Scene, cam: Camera, pixels) -> np.ndarray:
    """Test aid: flags[i] = 1 for surfels that composite at any of the
    (x, y) pixels (alpha >= 1/255 before the transmittance stop)."""
    xy = np.ascontiguousarray(np.asarray(pixels, dtype=np.int32).reshape(-1, 2))
    s, keep = scene._c()
    c = cam._c()
    flags = np.zeros(scene.n, dtype=np.uint8)
    rc = lib().oref_pixel_contributors(C.byref(s), C.byref(c), xy.shape[0],
                                       xy.ctypes.data_as(C.POINTER(C.c_int32)),
                                       flags.ctypes.data_as(C.POINTER(C.c_uint8)))
    if rc != 0:
        raise RuntimeError("oref_pixel_contributors failed")
    return flags


def trace(scene: Scene, cam: Camera, cap: int = 1024):
    """The faithful FP64 walk's contribution lists over the camera's window,
    in replay()'s format: (lists (P, cap) u32, counts (P,) u32)."""
    W, H = cam.out_size
    lists = np.zeros((W * H, cap), np.uint32)
    counts = np.zeros(W * H, np.uint32)
    s, keep = scene._c()
    c = cam._c()
    if lib().oref_trace(C.byref(s), C.byref(c), cap, _ptr(lists, _u32p), _ptr(counts, _u32p)) != 0:
        raise RuntimeError("oref_trace failed")
    return lists, counts


def replay(scene: Scene, cam: Camera, lists, counts, dl_dcolor=None, dl_ddepth=None, dl_dnormal=None,
           lambda_dist: float = 0.0):
    """Decision replay (test aid, no forward.rs counterpart): the FP64
    forward, and with dl_dcolor the backward, along another evaluation's
    per-pixel contribution lists over the camera's window (lists (P, cap) of
    surfel | ray << 30 | clamped << 31, counts (P,)). Returns (color (H, W, 3),
    alpha (H, W), Gradients or None)."""
    W, H = cam.out_size
    P = W * H
    lists = np.ascontiguousarray(lists, dtype=np.uint32).reshape(P, -1)
    counts = np.ascontiguousarray(counts, dtype=np.uint32).reshape(P)
    color = np.zeros(P * 3)
    alpha = np.zeros(P)
    s, keep = scene._c()
    c = cam._c()
    g, cg, dlc = None, None, None
    if dl_dcolor is not None:
        n = scene.n
        ncoef = (scene.sh_degree + 1) ** 2
        g = Gradients(np.zeros((n, 3)), np.zeros((n, 2)), np.zeros((n, 4)), np.zeros(n), np.zeros((n, ncoef, 3)))
        cg = _Grads(_ptr(g.pos, _dp), _ptr(g.scales, _dp), _ptr(g.quat, _dp), _ptr(g.opacity, _dp),
                    _ptr(g.sh, _dp))
        dlc = _d(dl_dcolor).reshape(-1)
        if dlc.size != 3 * P:
            raise ValueError("dl_dcolor must be (H, W, 3) over the window")
    dld = _d(dl_ddepth).reshape(-1) if dl_ddepth is not None else None
    dln = _d(dl_dnormal).reshape(-1) if dl_dnormal is not None else None
    rc = lib().oref_replay(C.byref(s), C.byref(c), _ptr(lists, _u32p), _ptr(counts, _u32p), lists.shape[1],
                           _ptr(color, _dp), _ptr(alpha, _dp), _ptr(dlc, _dp), _ptr(dld, _dp), _ptr(dln, _dp),
                           float(lambda_dist), C.byref(cg) if cg is not None else None)
    if rc != 0:
        raise RuntimeError(f"oref_replay failed ({rc}: a count above cap or a surfel index out of range)")
    if g is not None:
        g.cam_center = np.array(cg.cam_center[:])
        g.cam_quat = np.array(cg.cam_quat[:])
        g.focal = float(cg.focal)
    return color.reshape(H, W, 3), alpha.reshape(H, W), g


def prepare_rects(scene: Scene, cam: Camera):
    """forward.rs:104-123 in FP64: (rects (n, 4) int32 tile x0 y0 x1 y1,
    on_screen (n,) bool, culled (n,) bool: behind the near plane)."""
    n = scene.n
    rects = np.zeros((n, 4), np.int32)
    on = np.zeros(n, np.uint8)
    cul = np.zeros(n, np.uint8)
    s, keep = scene._c()
    c = cam._c()
    if lib().oref_prepare_rects(C.byref(s), C.byref(c), rects.ctypes.data_as(C.POINTER(C.c_int32)),
                                on.ctypes.data_as(C.POINTER(C.c_uint8)),
                                cul.ctypes.data_as(C.POINTER(C.c_uint8))) != 0:
        raise RuntimeError("oref_prepare_rects failed")
    return rects, on.astype(bool), cul.astype(bool)


def prepare_order(scene: Scene, cam: Camera) -> np.ndarray:
    s, keep = scene._c()
    c = cam._c()
    order = np.zeros(scene.n, dtype=np.int64)
    k = lib().oref_prepare_order(C.byref(s), C.byref(c), _ptr(order, _i64p))
    return order[:k]


# ------------------------------------------------------ fp32 integer mirror

def cam32(rot: np.ndarray, center, focal: float, width: int, height: int) -> _Cam32:
    """Build the fp32 camera the mirror and the GPU share. rot is R (3x3)."""
    c = _Cam32()
    r = np.asarray(rot, dtype=np.float32)
    for k in range(3):
        for j in range(3):
            c.rot[3 * k + j] = float(r[j, k])
    for k in range(3):
        c.center[k] = float(np.float32(center[k]))
    c.focal = float(np.float32(focal))
    c.width, c.height = int(width), int(height)
    c.tiles_x, c.tiles_y = (width + 15) // 16, (height + 15) // 16
    return c


def prep_f32(pos4, scales2, quat4, cam: _Cam32):
    """fp32 mirror of surfel_prep_fwd.wgsl:64-132 -> (rects n*4, keys, counts)."""
    pos4 = np.ascontiguousarray(pos4, dtype=np.float32)
    scales2 = np.ascontiguousarray(scales2, dtype=np.float32)
    quat4 = np.ascontiguousarray(quat4, dtype=np.float32)
    n = pos4.shape[0]
    rects = np.zeros((n, 4), dtype=np.uint32)
    keys = np.zeros(n, dtype=np.uint32)
    counts = np.zeros(n, dtype=np.uint32)
    lib().oref_prep_f32(n, _ptr(pos4, _fp), _ptr(scales2, _fp), _ptr(quat4, _fp), C.byref(cam),
                        _ptr(rects, _u32p), _ptr(keys, _u32p), _ptr(counts, _u32p))
    return rects, keys, counts


def exclusive_scan(values) -> tuple[np.ndarray, int]:
    v = np.ascontiguousarray(values, dtype=np.uint32)
    out = np.zeros_like(v)
    total = lib().oref_exclusive_scan_u32(v.shape[0], _ptr(v, _u32p), _ptr(out, _u32p))
    return out, int(total)


def emit_keys(rects, depth_keys, tiles_x: int, tile_base: int = 0, max_entries: int | None = None):
    rects = np.ascontiguousarray(rects, dtype=np.uint32)
    depth_keys = np.ascontiguousarray(depth_keys, dtype=np.uint32)
    n = depth_keys.shape[0]
    r = rects.reshape(n, 4).astype(np.int64)
    valid = (r[:, 2] >= r[:, 0]) & (r[:, 3] >= r[:, 1])
    total = int(((r[:, 2] - r[:, 0] + 1) * (r[:, 3] - r[:, 1] + 1) * valid).sum())
    cap = total if max_entries is None else int(max_entries)
    keys = np.zeros(max(cap, 1), dtype=np.uint64)
    payloads = np.zeros(max(cap, 1), dtype=np.uint32)
    e = lib().oref_emit_keys(n, _ptr(rects, _u32p), _ptr(depth_keys, _u32p), tiles_x, tile_base,
                             _ptr(keys, _u64p), _ptr(payloads, _u32p), cap)
    m = min(e, cap)
    return keys[:m].copy(), payloads[:m].copy(), int(e)


def stable_sort(keys, payloads):
    k = np.ascontiguousarray(keys, dtype=np.uint64).copy()
    p = np.ascontiguousarray(payloads, dtype=np.uint32).copy()
    lib().oref_stable_sort_u64(k.shape[0], _ptr(k, _u64p), _ptr(p, _u32p))
    return k, p


def tile_ranges(sorted_keys, num_tiles: int, tile_base: int = 0) -> np.ndarray:
    k = np.ascontiguousarray(sorted_keys, dtype=np.uint64)
    r = np.zeros((num_tiles, 2), dtype=np.uint32)
    lib().oref_tile_ranges(k.shape[0], _ptr(k, _u64p), num_tiles, tile_base, _ptr(r, _u32p))
    return r


# -------------------------------------------------------------- loss / Adam

def l1_ssim(img, target, lam: float = 0.2, dssim_scale: float = 0.5, grad: bool = True):
    """ssim.wgsl:41-143. img/target (H, W, 3). Returns (loss, dl_dimg, l1, ssim)."""
    img = _d(img)
    target = _d(target)
    H, W = img.shape[:2]
    dl = np.zeros(H * W * 3) if grad else None
    l1 = np.zeros(1)
    ss = np.zeros(1)
    loss = lib().oref_l1_ssim(W, H, _ptr(img.reshape(-1), _dp), _ptr(target.reshape(-1), _dp),
                              lam, dssim_scale, _ptr(dl, _dp), _ptr(l1, _dp), _ptr(ss, _dp))
    return (float(loss), dl.reshape(H, W, 3) if grad else None, float(l1[0]), float(ss[0]))


def normal_loss(color4, aux4, normal4, focal: float, lam: float, alpha_min: float = 0.2):
    """normal_loss.wgsl:52-145 in FP64. Inputs (H, W, 4); lam is the kernel
    weight (raw lambda / (W H)). Returns (loss_map (H, W), dl_dnormal
    (H, W, 3), dl_ddepth (H, W)); the loss is lam * loss_map.sum()."""
    H, W = np.shape(color4)[:2]
    c4, a4, n4 = _d(color4), _d(aux4), _d(normal4)
    lm = np.zeros(H * W)
    dn = np.zeros(H * W * 3)
    dd = np.zeros(H * W)
    lib().oref_normal_loss(W, H, float(focal), float(lam), float(alpha_min), _ptr(c4, _dp), _ptr(a4, _dp),
                           _ptr(n4, _dp), _ptr(lm, _dp), _ptr(dn, _dp), _ptr(dd, _dp))
    return lm.reshape(H, W), dn.reshape(H, W, 3), dd.reshape(H, W)


# ------------------------------------------------------------ MCMC

def _xorshift64(x: int) -> int:
    m = (1 << 64) - 1
    x ^= (x << 13) & m
    x ^= x >> 7
    x ^= (x << 17) & m
    return x


def _total_key(f) -> int:
    """Rust f32::total_cmp as an integer key."""
    b = int(np.float32(f).view(np.int32))
    return b ^ (((b >> 31) & 0xFFFFFFFF) >> 1)


def mcmc_sample_pairs(opacity, scales, dead: float, split_ratio: float, rng: int):
    """first_attempt/crates/gs-train/src/trainer.rs:1558-1650 (the host half
    of Trainer::mcmc_relocate), step by step in Python: (pairs [(dead,
    alive, mode, 0)], rng)."""
    op = np.asarray(opacity, np.float32).ravel()
    n = op.size
    dead_t = np.float32(dead)
    dead_idx = [i for i in range(n) if op[i] < dead_t]
    if not dead_idx:
        return [], rng
    cum, total = [], 0.0
    for o in op:
        if o >= dead_t:
            total += float(o)
        cum.append(total)
    if total <= 0.0:
        return [], rng
    used, spent, pairs = [False] * n, [False] * len(dead_idx), []
    for di, d in enumerate(dead_idx):
        rng = _xorshift64(rng)
        r = float(rng >> 11) / float(1 << 53) * total
        a = min(int(np.searchsorted(np.asarray(cum), r, side="left")), n - 1)
        if a == d or used[a]:
            continue
        used[a] = True
        spent[di] = True
        pairs.append((d, a, 0, 0))
    if split_ratio > 0.0:
        sc = np.asarray(scales, np.float32).reshape(n, 2)
        size = [max(sc[i, 0], sc[i, 1]) for i in range(n)]
        alive = sorted((size[i] for i in range(n) if op[i] >= dead_t), key=_total_key)
        if len(alive) >= 16:
            cap = np.float32(split_ratio) * np.float32(alive[len(alive) // 2])
            big = [i for i in range(n) if op[i] >= dead_t and size[i] > cap and not used[i]]
            big.sort(key=lambda i: -_total_key(size[i]))  # stable, largest first
            left = [d for di, d in enumerate(dead_idx) if not spent[di]]
            for b, slot in zip(big, left):
                used[b] = True
                pairs.append((slot, b, 1, 0))
    return pairs, rng


def mcmc_relocate_arrays(pairs, opac_act, raw, m, v):
    """relocate.wgsl:39-125 in f32 on host copies: raw / m / v are dicts of the
    five classes (pos (n,4), scales (n,2), quat (n,4), opacity (n,), sh
    (n,48)), changed in place. sqrt / log / exp and the normalization are
    f32 operations here (numpy), correctly rounded or within an ulp."""
    f = np.float32
    for d, a, mode, _ in pairs:
        alpha_a = np.clip(f(opac_act[a]), f(1e-4), f(0.999))
        alpha_new = np.clip(f(1.0) - np.sqrt(f(1.0) - alpha_a), f(1e-4), f(0.999))
        logit = np.float32(np.log(np.float64(alpha_new / (f(1.0) - alpha_new))))
        raw["opacity"][d] = raw["opacity"][a] = logit
        q = raw["quat"][a].copy()
        raw["quat"][d] = q
        sa = raw["scales"][a].copy()
        if mode == 1:
            qn = np.sqrt((q * q).sum(dtype=np.float32))
            x, y, z, w = q / qn
            col0 = np.array([1 - 2 * (y * y + z * z), 2 * (x * y + w * z), 2 * (x * z - w * y)], np.float32)
            col1 = np.array([2 * (x * y - w * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w * x)], np.float32)
            sw = np.exp(sa.astype(np.float64)).astype(np.float32)
            axis = col1 * sw[1] if sw[1] > sw[0] else col0 * sw[0]
            c = raw["pos"][a].copy()
            off = f(0.6) * axis
            raw["pos"][d] = np.append(c[:3] + off, c[3])
            raw["pos"][a] = np.append(c[:3] - off, c[3])
            s = sa - f(0.51)
        else:
            raw["pos"][d] = raw["pos"][a]
            s = sa - f(0.16)
        raw["scales"][a] = s
        raw["scales"][d] = s
        raw["sh"][d] = raw["sh"][a]
        for k in raw:
            m[k][d] = 0.0
            m[k][a] = 0.0
            v[k][d] = 0.0
            v[k][a] = 0.0


def _pcg(v):
    v = np.asarray(v, np.uint64)
    state = (v * np.uint64(747796405) + np.uint64(2891336453)) & np.uint64(0xFFFFFFFF)
    word = (((state >> ((state >> np.uint64(28)) + np.uint64(4))) ^ state) * np.uint64(277803737)) & np.uint64(0xFFFFFFFF)
    return ((word >> np.uint64(22)) ^ word) & np.uint64(0xFFFFFFFF)


def mcmc_noise(pos_raw, scales_act, opac_act, seed: int, sigma: float, opacity_k: float = 5.0):
    """shaders/mcmc.wgsl:25-58 (add_noise) in f32: the new raw positions (n, 4).
    The PCG stream is integer-exact; exp is the correctly rounded f32 value."""
    p = np.asarray(pos_raw, np.float32).reshape(-1, 4).copy()
    n = p.shape[0]
    s = np.asarray(scales_act, np.float32).reshape(n, 2)
    mean_scale = np.float32(0.5) * (s[:, 0] + s[:, 1])
    gate = np.exp((np.float32(-opacity_k) * np.asarray(opac_act, np.float32)).astype(np.float64)).astype(np.float32)
    amp = (np.float32(sigma) * mean_scale * gate).astype(np.float32)
    base = _pcg(np.arange(n, dtype=np.uint64) ^ np.uint64(seed & 0xFFFFFFFF))
    for c, salt in enumerate((0x68bc21eb, 0x2c1b3c6d, 0x1f123bb5)):
        st = base ^ np.uint64(salt)
        acc = np.zeros(n, np.float32)
        for _ in range(4):
            st = _pcg(st)
            acc = acc + (st & np.uint64(0x00ffffff)).astype(np.float32) / np.float32(16777216.0)
        nz = (acc - np.float32(2.0)).astype(np.float32)
        upd = amp > 0
        p[upd, c] = p[upd, c] + amp[upd] * nz[upd]
    return p


# ------------------------------------------------------------ tracking

def photometric(render4, frame4, gain: float = 1.0, bias: float = 0.0):
    """crates/gs-track/src/shaders/photometric.wgsl:47-73 in f32 (sums in
    float64): (dl_dcolor (P, 4), residual, coverage)."""
    r = np.asarray(render4, np.float32).reshape(-1, 4)
    t = np.asarray(frame4, np.float32).reshape(-1, 4)
    P = r.shape[0]
    d = r[:, :3] - (np.float32(gain) * t[:, :3] + np.float32(bias))
    dl = np.zeros((P, 4), np.float32)
    dl[:, :3] = np.clip(d / np.float32(1e-4), -1.0, 1.0) * (np.float32(1.0) / np.float32(3.0 * P))
    residual = float(np.abs(d).astype(np.float64).sum() / (3.0 * P))
    coverage = float((r[:, 3] > 0.5).sum() / P)
    return dl, residual, coverage


def _mat3_from_quat32(q):
    x, y, z, w = (np.float32(v) for v in q)
    x2, y2, z2 = x + x, y + y, z + z
    xx, xy, xz, yy, yz, zz = x * x2, x * y2, x * z2, y * y2, y * z2, z * z2
    wx, wy, wz = w * x2, w * y2, w * z2
    one = np.float32(1.0)
    # Rows of R (R[i][j]); glam's columns are (1-(yy+zz), xy+wz, xz-wy), ...
    return np.array([[one - (yy + zz), xy - wz, xz + wy],
                     [xy + wz, one - (xx + zz), yz - wx],
                     [xz - wy, yz + wx, one - (xx + yy)]], np.float32)


def camera_step_grad(grad_cam, quat):
    """crates/gs-track/src/lib.rs:276-297: [rotation triple (antisymmetric part
    of R^T dL/dR), dL/dC] in f32; None if non-finite."""
    g = np.asarray(grad_cam, np.float32)[:12]
    if not np.all(np.isfinite(g)):
        return None
    A = g[:9].reshape(3, 3).T  # columns g[0..2], g[3..5], g[6..8]
    B = (_mat3_from_quat32(quat).T @ A).astype(np.float32)
    return np.array([B[2, 1] - B[1, 2], B[0, 2] - B[2, 0], B[1, 0] - B[0, 1], g[9], g[10], g[11]], np.float32)


def adam6(m, v, t: int, grad, lr_rot: float, lr_center: float):
    """lib.rs:300-319 in f32 (m, v updated in place); returns the step."""
    f = np.float32
    bc1 = f(1.0) - f(0.9) ** f(t)
    bc2 = f(1.0) - f(0.999) ** f(t)
    step = np.zeros(6, np.float32)
    for k in range(6):
        m[k] = f(0.9) * m[k] + f(0.1) * grad[k]
        v[k] = f(0.999) * v[k] + f(0.001) * grad[k] * grad[k]
        lr = f(lr_rot) if k < 3 else f(lr_center)
        step[k] = lr * (m[k] / bc1) / (np.sqrt(v[k] / bc2) + f(1e-10))
    return step


def apply_camera_step(quat, center, step):
    """lib.rs:321-329: quat' = normalize(quat * exp(-step[0:3] / 2)), centre' =
    centre - step[3:6], in f32."""
    d = -np.asarray(step[:3], np.float32)
    ang = np.float32(np.sqrt((d.astype(np.float64) ** 2).sum()))
    if ang > 0:
        e = np.append(d / ang * np.float32(np.sin(ang * 0.5)), np.float32(np.cos(ang * 0.5))).astype(np.float32)
    else:
        e = np.array([0, 0, 0, 1], np.float32)
    q = np.asarray(quat, np.float32)
    p = np.array([q[3] * e[0] + q[0] * e[3] + q[1] * e[2] - q[2] * e[1],
                  q[3] * e[1] - q[0] * e[2] + q[1] * e[3] + q[2] * e[0],
                  q[3] * e[2] + q[0] * e[1] - q[1] * e[0] + q[2] * e[3],
                  q[3] * e[3] - q[0] * e[0] - q[1] * e[1] - q[2] * e[2]], np.float32)
    p = p / np.float32(np.sqrt((p.astype(np.float64) ** 2).sum()))
    return p.astype(np.float32), (np.asarray(center, np.float32) - np.asarray(step[3:], np.float32)).astype(np.float32)


# ------------------------------------------------------------ 3DGS .ply

def ply_header(count: int) -> bytes:
    """first_attempt/crates/gs-io/src/ply/writer.rs:31-51."""
    lines = ["ply", "format binary_little_endian 1.0", "comment splattar baked export (surfels flattened to 3DGS)",
             f"element vertex {count}"]
    names = (["x", "y", "z"] + [f"f_dc_{i}" for i in range(3)] + [f"f_rest_{i}" for i in range(45)] + ["opacity"]
             + [f"scale_{i}" for i in range(3)] + [f"rot_{i}" for i in range(4)])
    lines += [f"property float {nm}" for nm in names] + ["end_header"]
    return ("\n".join(lines) + "\n").encode()


def ply_rows(positions4, scales2, quats4, opacities, sh48) -> np.ndarray:
    """writer.rs:53-77 in f32: (n, 59) rows. The three logarithms are the
    correctly rounded f32 log of the f32 argument (the FP64 log rounded once;
    the reference's f32::ln is libm's logf)."""
    p = np.asarray(positions4, np.float32).reshape(-1, 4)
    n = p.shape[0]
    sh = np.asarray(sh48, np.float32).reshape(n, 16, 3)
    rows = np.zeros((n, 59), np.float32)
    rows[:, 0:3] = p[:, :3]
    rows[:, 3:6] = sh[:, 0, :]
    rows[:, 6:51] = sh[:, 1:, :].transpose(0, 2, 1).reshape(n, 45)  # channel-major over coeffs 1..15
    o = np.clip(np.asarray(opacities, np.float32).reshape(n), np.float32(1e-6), np.float32(1.0 - 1e-6))
    rows[:, 51] = np.log((o / (np.float32(1.0) - o)).astype(np.float64)).astype(np.float32)
    s = np.asarray(scales2, np.float32).reshape(n, 2)
    rows[:, 52:54] = np.log(np.maximum(s, np.float32(1e-9)).astype(np.float64)).astype(np.float32)
    rows[:, 54] = np.float32(-10.0)
    q = np.asarray(quats4, np.float32).reshape(n, 4)
    rows[:, 55] = q[:, 3]
    rows[:, 56:59] = q[:, :3]
    return rows


def ply_read_canonical(data: bytes):
    """gs-io/src/ply/reader.rs:77-147 (push_row_fast) on a canonical INRIA
    degree-3 file: (positions (n, 3), sh (n, 16, 3) coefficient-major,
    opacity (sigmoid), scales (n, 3) (exp), quats xyzw normalized)."""
    end = data.index(b"end_header\n") + len(b"end_header\n")
    head = data[:end].decode().split("\n")
    n = int([h for h in head if h.startswith("element vertex")][0].split()[-1])
    rows = np.frombuffer(data[end:], dtype="<f4").reshape(n, 59).astype(np.float32)
    pos = rows[:, 0:3]
    sh = np.zeros((n, 16, 3), np.float32)
    sh[:, 0, :] = rows[:, 3:6]
    sh[:, 1:, :] = rows[:, 6:51].reshape(n, 3, 15).transpose(0, 2, 1)
    opac = (np.float32(1.0) / (np.float32(1.0) + np.exp(-rows[:, 51]))).astype(np.float32)
    scales = np.exp(rows[:, 52:55])
    q = np.stack([rows[:, 56], rows[:, 57], rows[:, 58], rows[:, 55]], 1)
    nrm = np.sqrt((q.astype(np.float64) ** 2).sum(1, keepdims=True))
    q = np.where(nrm ** 2 > 1e-12, q / np.maximum(nrm, 1e-30), np.array([0, 0, 0, 1.0]))
    return pos, sh, opac, scales, q


def blur(channel: np.ndarray) -> np.ndarray:
    a = _d(channel)
    H, W = a.shape
    out = np.zeros_like(a)
    lib().oref_blur(W, H, _ptr(a.reshape(-1), _dp), _ptr(out.reshape(-1), _dp))
    return out


def adam_step_f32(raw, grad_act, m, v, lr: float, activation: int, t: int, reg: float = 0.0):
    """adam.wgsl:44-74 in f32 (in place). Returns the activated copy."""
    act = np.zeros_like(raw)
    lib().oref_adam_step_f32(raw.shape[0], _ptr(raw, _fp), _ptr(grad_act, _fp), _ptr(m, _fp),
                             _ptr(v, _fp), _ptr(act, _fp), lr, reg, activation, t)
    return act


def xorshift32(seed: int, n: int, modulus: int = 0) -> np.ndarray:
    """xorshift32 stream of crates/gs-wgpu/tests/sort.rs:18-25 (mod modulus)."""
    out = np.empty(n, dtype=np.uint32)
    lib().oref_xorshift32_fill(seed & 0xFFFFFFFF, n, modulus, _ptr(out, _u32p))
    return out


def num_threads() -> int:
    return int(lib().oref_num_threads())
