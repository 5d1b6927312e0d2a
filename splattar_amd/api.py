"""Host-side mirror of the reference op API over the HIP C ABI.

Same names, argument meaning and error behaviour as the reference's Rust
types, so callers (and the parity tests) read like the reference's own code:

  GpuContext      crates/gs-wgpu/src/context.rs:7-96
  PrefixSum       crates/gs-wgpu/src/prefix_sum.rs:19-164
  RadixSorter     crates/gs-wgpu/src/sort.rs:17-327
  TileBinner      crates/gs-kernels/src/binning.rs:22-238
  Rasterizer      crates/gs-kernels/src/rasterizer.rs:56-448
  SsimLoss        first_attempt/crates/gs-train/src/ssim.rs:25-298
  Optimizer       first_attempt/crates/gs-train/src/optim.rs:50-241
  Trainer         first_attempt/crates/gs-train/src/trainer.rs:1139-1332 (+ batching)

The reference's encode(encoder) calls become enqueue-on-stream calls; device
buffers are exposed as DeviceBuffer views (pointer + dtype + shape) with
write() (queue.write_buffer) and read() (buffers::readback).
"""

from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from ._lib import check, lib


class DeviceBuffer:
    """A typed view of device memory owned by a handle (or by itself)."""

    def __init__(self, ctx: "GpuContext", ptr: int, dtype, shape, owned: bool = False):
        self.ctx = ctx
        self.ptr = int(ptr or 0)
        self.dtype = np.dtype(dtype)
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.owned = owned

    @property
    def nbytes(self) -> int:
        return int(np.prod(self.shape)) * self.dtype.itemsize

    def write(self, arr, offset_elems: int = 0):
        a = np.ascontiguousarray(arr, dtype=self.dtype)
        if a.size + offset_elems > int(np.prod(self.shape)):
            raise ValueError("write exceeds buffer size")
        if a.size:
            check(lib().gs_write_buffer(self.ctx.h, self.ptr + offset_elems * self.dtype.itemsize,
                                        a.ctypes.data, a.nbytes))

    def read(self, count: int | None = None) -> np.ndarray:
        n = int(np.prod(self.shape)) if count is None else int(count)
        out = np.empty(n, dtype=self.dtype)
        if n:
            check(lib().gs_read_buffer(self.ctx.h, out.ctypes.data, self.ptr, out.nbytes))
        if count is None:
            return out.reshape(self.shape)
        return out

    def clear(self, stream=None):
        check(lib().gs_clear_buffer(self.ctx.h, stream, self.ptr, self.nbytes))

    def free(self):
        if self.owned and self.ptr:
            check(lib().gs_buffer_free(self.ctx.h, self.ptr))
            self.ptr = 0


class GpuContext:
    """GpuContext (context.rs:7-96): one device, one stream."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().gs_ctx_create(device, C.byref(h)))
        self.h = h.value

    @property
    def stream(self):
        return lib().gs_ctx_stream(self.h)

    def synchronize(self):
        check(lib().gs_ctx_synchronize(self.h))

    def device_info(self):
        name = C.create_string_buffer(256)
        cu = C.c_int()
        check(lib().gs_ctx_device_info(self.h, name, 256, C.byref(cu)))
        return name.value.decode(), cu.value

    def alloc(self, dtype, shape) -> DeviceBuffer:
        n = int(np.prod(shape)) * np.dtype(dtype).itemsize
        p = C.c_void_p()
        check(lib().gs_buffer_alloc(self.h, n, C.byref(p)))
        return DeviceBuffer(self, p.value, dtype, shape, owned=True)

    def close(self):
        if self.h:
            lib().gs_ctx_destroy(self.h)
            self.h = None


class PrefixSum:
    """PrefixSum (prefix_sum.rs:19-164)."""

    def __init__(self, ctx: GpuContext, capacity: int):
        self.ctx = ctx
        h = C.c_void_p()
        check(lib().gs_scan_create(ctx.h, capacity, C.byref(h)))
        self.h = h.value
        self.capacity = capacity

    def data(self) -> DeviceBuffer:
        return DeviceBuffer(self.ctx, lib().gs_scan_data(self.h), np.uint32, self.capacity)

    def total(self) -> DeviceBuffer:
        return DeviceBuffer(self.ctx, lib().gs_scan_total(self.h), np.uint32, 1)

    def encode(self, n: int, stream=None):
        check(lib().gs_scan_encode(self.h, stream, n))

    def close(self):
        if self.h:
            lib().gs_scan_destroy(self.h)
            self.h = None


class RadixSorter:
    """RadixSorter (sort.rs:17-327), 32- or 64-bit keys."""

    def __init__(self, ctx: GpuContext, capacity: int, key_bits: int = 32):
        self.ctx = ctx
        self.capacity = capacity
        self.key_bits = key_bits
        h = C.c_void_p()
        check(lib().gs_sorter_create(ctx.h, capacity, key_bits, C.byref(h)))
        self.h = h.value

    def keys(self) -> DeviceBuffer:
        return DeviceBuffer(self.ctx, lib().gs_sorter_keys(self.h),
                            np.uint64 if self.key_bits == 64 else np.uint32, self.capacity)

    def payloads(self) -> DeviceBuffer:
        return DeviceBuffer(self.ctx, lib().gs_sorter_payloads(self.h), np.uint32, self.capacity)

    def counts(self) -> DeviceBuffer:
        return DeviceBuffer(self.ctx, lib().gs_sorter_counts(self.h), np.uint32, 1)

    def encode(self, stream=None, begin_bit: int = 0, end_bit: int | None = None, digit_bits: int = 8):
        check(lib().gs_sorter_encode(self.h, stream, begin_bit,
                                     self.key_bits if end_bit is None else end_bit, digit_bits))

    def close(self):
        if self.h:
            lib().gs_sorter_destroy(self.h)
            self.h = None


class TileBinner:
    """TileBinner (binning.rs:22-238)."""

    def __init__(self, ctx: GpuContext, max_items: int, max_entries: int, num_tiles: int):
        self.ctx = ctx
        self.max_items, self.max_entries, self.num_tiles = max_items, max_entries, num_tiles
        h = C.c_void_p()
        check(lib().gs_binner_create(ctx.h, max_items, max_entries, num_tiles, C.byref(h)))
        self.h = h.value
        self.rects = DeviceBuffer(ctx, lib().gs_binner_rects(self.h), np.uint32, (max_items, 4))
        self.depths = DeviceBuffer(ctx, lib().gs_binner_depths(self.h), np.uint32, max_items)
        self.ranges = DeviceBuffer(ctx, lib().gs_binner_ranges(self.h), np.uint32, (num_tiles, 2))

    def sorted_entries(self) -> DeviceBuffer:
        return DeviceBuffer(self.ctx, lib().gs_binner_sorted_entries(self.h), np.uint32, self.max_entries)

    def sorted_keys(self) -> DeviceBuffer:
        return DeviceBuffer(self.ctx, lib().gs_binner_sorted_keys(self.h), np.uint64, self.max_entries)

    def entry_count(self) -> DeviceBuffer:
        return DeviceBuffer(self.ctx, lib().gs_binner_entry_count(self.h), np.uint32, 1)

    def encode(self, num_items: int, tiles_x: int, stream=None, digit_bits: int = 8):
        check(lib().gs_binner_encode(self.h, stream, num_items, tiles_x, digit_bits))

    def overflowed(self) -> bool:
        f = C.c_int()
        check(lib().gs_binner_overflowed(self.h, C.byref(f)))
        return bool(f.value)

    def close(self):
        if self.h:
            lib().gs_binner_destroy(self.h)
            self.h = None


@dataclass
class Camera:
    """RasterCamera (rasterizer.rs:34-41): camera-to-world quat xyzw."""
    center: np.ndarray
    quat: np.ndarray
    focal: float
    sh_degree: int = 0

    def c(self) -> L.RasterCamera:
        rc = L.RasterCamera()
        for k in range(3):
            rc.center[k] = float(self.center[k])
        for k in range(4):
            rc.quat[k] = float(self.quat[k])
        rc.focal = float(self.focal)
        rc.sh_degree = int(self.sh_degree)
        return rc


def scene_input(positions, scales, quats, opacities, sh, keep: list) -> L.SceneInput:
    """SceneInput (rasterizer.rs:43-54) from host arrays; sh is (n, coeffs*3)."""
    p = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
    n = p.shape[0]
    s = np.ascontiguousarray(scales, np.float32).reshape(n, 2)
    q = np.ascontiguousarray(quats, np.float32).reshape(n, 4)
    o = np.ascontiguousarray(opacities, np.float32).reshape(n)
    h = np.ascontiguousarray(sh, np.float32).reshape(n, -1)
    if h.shape[1] % 3:
        raise ValueError("sh must hold 3 floats per coefficient")
    keep.extend([p, s, q, o, h])
    return L.SceneInput(p.ctypes.data, s.ctypes.data, q.ctypes.data, o.ctypes.data, h.ctypes.data,
                        h.shape[1] // 3, n)


class Rasterizer:
    """Rasterizer (rasterizer.rs:56-448)."""

    def __init__(self, ctx: GpuContext, capacity: int, width: int, height: int, max_entries: int,
                 max_views: int = 1):
        """max_views > 1 makes a batched rasterizer (gs_raster_create_batch):
        per-pixel buffers then hold max_views images of width x height (a
        batch's views back to back) and grad_cam 16 floats per view."""
        self.ctx = ctx
        self.capacity, self.width, self.height, self.max_entries = capacity, width, height, max_entries
        self.max_views = max_views
        h = C.c_void_p()
        check(lib().gs_raster_create_batch(ctx.h, capacity, width, height, max_entries, max_views,
                                           C.byref(h)))
        self.h = h.value
        b = L.RasterBuffers()
        check(lib().gs_raster_get_buffers(self.h, C.byref(b)))
        P = width * height
        f, u = np.float32, np.uint32
        self.positions = DeviceBuffer(ctx, b.positions, f, (capacity, 4))
        self.scales = DeviceBuffer(ctx, b.scales, f, (capacity, 2))
        self.quats = DeviceBuffer(ctx, b.quats, f, (capacity, 4))
        self.opacities = DeviceBuffer(ctx, b.opacities, f, capacity)
        self.sh = DeviceBuffer(ctx, b.sh, f, (capacity, 48))
        if max_views > 1:
            # Flat per-pixel views (a batch packs its views back to back).
            Pb = P * max_views
            self.out_color = DeviceBuffer(ctx, b.out_color, f, (Pb, 4))
            self.out_aux = DeviceBuffer(ctx, b.out_aux, f, (Pb, 4))
            self.out_normal = DeviceBuffer(ctx, b.out_normal, f, (Pb, 4))
            self.out_ncontrib = DeviceBuffer(ctx, b.out_ncontrib, u, Pb)
            self.dl_dcolor = DeviceBuffer(ctx, b.dl_dcolor, f, (Pb, 4))
            self.dl_dnormal = DeviceBuffer(ctx, b.dl_dnormal, f, (Pb, 4))
        else:
            self.out_color = DeviceBuffer(ctx, b.out_color, f, (height, width, 4))
            self.out_aux = DeviceBuffer(ctx, b.out_aux, f, (height, width, 4))
            self.out_normal = DeviceBuffer(ctx, b.out_normal, f, (height, width, 4))
            self.out_ncontrib = DeviceBuffer(ctx, b.out_ncontrib, u, (height, width))
            self.dl_dcolor = DeviceBuffer(ctx, b.dl_dcolor, f, (height, width, 4))
            self.dl_dnormal = DeviceBuffer(ctx, b.dl_dnormal, f, (height, width, 4))
        self.grad_pos = DeviceBuffer(ctx, b.grad_pos, f, (capacity, 4))
        self.grad_scales = DeviceBuffer(ctx, b.grad_scales, f, (capacity, 2))
        self.grad_quat = DeviceBuffer(ctx, b.grad_quat, f, (capacity, 4))
        self.grad_opacity = DeviceBuffer(ctx, b.grad_opacity, f, capacity)
        self.grad_sh = DeviceBuffer(ctx, b.grad_sh, f, (capacity, 48))
        self.grad_cam = DeviceBuffer(ctx, b.grad_cam, f, 16 if max_views == 1 else (max_views, 16))
        self._P = P

    def upload_scene(self, positions, scales, quats, opacities, sh):
        keep: list = []
        si = scene_input(positions, scales, quats, opacities, sh, keep)
        check(lib().gs_raster_upload_scene(self.h, C.byref(si)))

    def set_lambda_dist(self, lam: float):
        check(lib().gs_raster_set_lambda_dist(self.h, lam))

    def set_digit_bits(self, bits: int):
        check(lib().gs_raster_set_digit_bits(self.h, bits))

    def set_binning(self, mode: int):
        """L.GS_BINNING_PRESORT (default), L.GS_BINNING_SORT64 or L.GS_BINNING_TIGHT."""
        check(lib().gs_raster_set_binning(self.h, mode))

    def set_culling(self, enable: bool):
        """Per-quadrant footprint culling in the rasterizer kernels (default on)."""
        check(lib().gs_raster_set_culling(self.h, 1 if enable else 0))

    def set_trace(self, x0: int, y0: int, w: int, h: int, cap: int):
        """Record the next forwards' contributions over a pixel window
        (verification aid; cap 0 turns it off)."""
        check(lib().gs_raster_set_trace(self.h, x0, y0, w, h, cap))
        self._trace = (w, h, cap) if w * h * cap else None

    def trace(self):
        """(lists (h*w, cap) u32, counts (h*w,) u32) of the last traced forward."""
        w, h, cap = self._trace
        lp, cp = C.c_void_p(), C.c_void_p()
        check(lib().gs_raster_trace_buffers(self.h, C.byref(lp), C.byref(cp)))
        lists = DeviceBuffer(self.ctx, lp.value, np.uint32, (h * w, cap)).read()
        counts = DeviceBuffer(self.ctx, cp.value, np.uint32, h * w).read()
        return lists, counts

    def forward(self, camera: Camera, num_surfels: int, stream=None):
        rc = camera.c()
        check(lib().gs_raster_forward(self.h, stream, C.byref(rc), num_surfels))

    def backward(self, num_surfels: int, stream=None, flags: int = 0):
        check(lib().gs_raster_backward_ex(self.h, stream, num_surfels, flags))

    def forward_batch(self, views, stream=None):
        """views: [(Camera, width, height, num_surfels)]; view j's pixels follow
        view j-1's in the per-pixel buffers."""
        arr = (L.View * len(views))()
        for k, (cam, w, h, m) in enumerate(views):
            arr[k].camera = cam.c()
            arr[k].width, arr[k].height, arr[k].num_surfels = w, h, m
        check(lib().gs_raster_forward_batch(self.h, stream, arr, len(views)))

    def backward_batch(self, stream=None, flags: int = 0):
        check(lib().gs_raster_backward_batch(self.h, stream, flags))

    def entry_count(self) -> int:
        e = C.c_uint32()
        check(lib().gs_raster_entry_count(self.h, C.byref(e)))
        return int(e.value)

    def debug(self) -> dict:
        d = L.RasterDebug()
        check(lib().gs_raster_debug_buffers(self.h, C.byref(d)))
        e = C.c_uint32()
        lib().gs_raster_entry_count(self.h, C.byref(e))
        n = min(int(e.value), self.max_entries)
        ctx = self.ctx
        return dict(
            sorted_keys=DeviceBuffer(ctx, d.sorted_keys, np.uint64, self.max_entries).read(n),
            sorted_payloads=DeviceBuffer(ctx, d.sorted_payloads, np.uint32, self.max_entries).read(n),
            ranges=DeviceBuffer(ctx, d.ranges, np.uint32, (d.num_tiles, 2)).read(),
            rects=DeviceBuffer(ctx, d.rects, np.uint32, (self.capacity, 4)).read(),
            depth_keys=DeviceBuffer(ctx, d.depth_keys, np.uint32, self.capacity).read(),
            entries=int(e.value))

    def close(self):
        if self.h:
            lib().gs_raster_destroy(self.h)
            self.h = None


class SsimLoss:
    """SsimLoss (ssim.rs:25-298) bound to an image and a dL/d(image) buffer."""

    def __init__(self, ctx: GpuContext, width: int, height: int, lam: float, img: DeviceBuffer,
                 dl_out: DeviceBuffer, dssim_scale: float = 0.5):
        self.ctx = ctx
        h = C.c_void_p()
        check(lib().gs_loss_create(ctx.h, width, height, lam, img.ptr, dl_out.ptr, C.byref(h)))
        self.h = h.value
        self.target = DeviceBuffer(ctx, lib().gs_loss_target(self.h), np.float32, (height, width, 4))
        check(lib().gs_loss_set_dssim_scale(self.h, dssim_scale))

    def encode(self, stream=None):
        check(lib().gs_loss_encode(self.h, stream))

    def read_losses(self):
        a, b = C.c_double(), C.c_double()
        check(lib().gs_loss_read(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def close(self):
        if self.h:
            lib().gs_loss_destroy(self.h)
            self.h = None


def write_3dgs_ply(ctx: "GpuContext", path: str, count: int, positions: DeviceBuffer, scales: DeviceBuffer,
                   quats: DeviceBuffer, opacities: DeviceBuffer, sh: DeviceBuffer):
    """gs_io::write_3dgs_ply (writer.rs:17-79) over device buffers."""
    check(lib().gs_write_3dgs_ply(ctx.h, os.fsencode(path), count, positions.ptr, scales.ptr, quats.ptr,
                                  opacities.ptr, sh.ptr))


def mcmc_sample_pairs(opacity, scales, dead: float, split_ratio: float, rng: int):
    """The host sampling of Trainer::mcmc_relocate: ((n_pairs, 4) u32 pairs,
    advanced rng state). Runs on the host only."""
    op = np.ascontiguousarray(opacity, dtype=np.float32).ravel()
    sc = np.ascontiguousarray(scales, dtype=np.float32).ravel() if scales is not None else None
    n = op.size
    pairs = np.zeros((max(n, 1), 4), np.uint32)
    r = C.c_uint64(rng)
    npairs = C.c_uint32()
    check(lib().gs_mcmc_sample_pairs(op.ctypes.data, sc.ctypes.data if sc is not None else None, n, dead,
                                     split_ratio, C.byref(r), pairs.ctypes.data, C.byref(npairs)))
    return pairs[:npairs.value].copy(), int(r.value)


def mcmc_add_noise(ctx: "GpuContext", n: int, seed: int, sigma: float, opacity_k: float, pos_raw: DeviceBuffer,
                   scales_act: DeviceBuffer, opac_act: DeviceBuffer, stream=None):
    """mcmc.wgsl add_noise over device buffers."""
    check(lib().gs_mcmc_add_noise(ctx.h, stream, n, seed, sigma, opacity_k, pos_raw.ptr, scales_act.ptr,
                                  opac_act.ptr))


def ply_pack_rows(ctx: "GpuContext", count: int, positions: DeviceBuffer, scales: DeviceBuffer,
                  quats: DeviceBuffer, opacities: DeviceBuffer, sh: DeviceBuffer, rows: DeviceBuffer, stream=None):
    """The export's rows (count x 59 floats) into device memory."""
    check(lib().gs_ply_pack_rows(ctx.h, stream, count, positions.ptr, scales.ptr, quats.ptr, opacities.ptr,
                                 sh.ptr, rows.ptr))


@dataclass
class TrackConfig:
    """TrackConfig (crates/gs-track/src/lib.rs:46-87) with its defaults."""
    iterations: int = 60
    lr_rot: float = 2e-3
    lr_center: float = 2e-3
    min_improvement: float = 1e-4
    patience: int = 10


@dataclass
class TrackReport:
    """TrackReport (lib.rs:17-44) plus the last evaluated pose."""
    residuals: list
    coverage: float
    iterations: int
    best: float
    last_quat: np.ndarray
    last_center: np.ndarray

    def final_residual(self) -> float:
        return self.best

    def improved(self) -> bool:
        return bool(self.residuals) and self.best < self.residuals[0]


class Tracker:
    """Tracker (crates/gs-track/src/lib.rs:89-337) over a Rasterizer holding the map."""

    def __init__(self, ctx: GpuContext, raster: "Rasterizer", width: int, height: int):
        self.ctx = ctx
        self.raster = raster
        self.width, self.height = width, height
        h = C.c_void_p()
        check(lib().gs_tracker_create(ctx.h, raster.h, width, height, C.byref(h)))
        self.h = h.value
        self.target = DeviceBuffer(ctx, lib().gs_tracker_target(self.h), np.float32, (height, width, 4))

    def set_target(self, frame):
        f = np.ascontiguousarray(frame, np.float32).reshape(self.height, self.width, 4)
        check(lib().gs_tracker_set_target(self.h, f.ctypes.data))

    def track(self, surfels: int, camera: Camera, depth: float, cfg: TrackConfig | None = None):
        """Recover `camera` (updated in place to the best pose); returns the report."""
        cfg = cfg or TrackConfig()
        rc = camera.c()
        tc = L.TrackConfig(cfg.iterations, cfg.lr_rot, cfg.lr_center, cfg.min_improvement, cfg.patience)
        res = np.zeros(cfg.iterations, np.float32)
        rep = L.TrackReport()
        check(lib().gs_tracker_track(self.h, surfels, C.byref(rc), depth, C.byref(tc), res.ctypes.data,
                                     C.byref(rep)))
        camera.center = np.array(rc.center[:], np.float32)
        camera.quat = np.array(rc.quat[:], np.float32)
        return TrackReport(res[:rep.iterations].tolist(), rep.coverage, rep.iterations, rep.best,
                           np.array(rep.last_quat[:], np.float32), np.array(rep.last_center[:], np.float32))

    def photometric(self, gain: float = 1.0, bias: float = 0.0, stream=None):
        """The photometric pass on the rasterizer's current render: (residual, coverage)."""
        e, c = C.c_double(), C.c_double()
        check(lib().gs_tracker_photometric(self.h, stream, gain, bias, C.byref(e), C.byref(c)))
        return e.value, c.value

    def close(self):
        if self.h:
            lib().gs_tracker_destroy(self.h)
            self.h = None


class NormalLoss:
    """NormalLoss (normal_loss.rs) bound to a rasterizer's forward outputs and
    backward inputs (out_color, out_aux, out_normal -> dl_dnormal, dl_dcolor.w)."""

    def __init__(self, ctx: GpuContext, width: int, height: int, out_color: DeviceBuffer,
                 out_aux: DeviceBuffer, out_normal: DeviceBuffer, dl_dnormal: DeviceBuffer,
                 dl_dcolor: DeviceBuffer):
        self.ctx = ctx
        h = C.c_void_p()
        check(lib().gs_normal_loss_create(ctx.h, width, height, out_color.ptr, out_aux.ptr, out_normal.ptr,
                                          dl_dnormal.ptr, dl_dcolor.ptr, C.byref(h)))
        self.h = h.value
        self.loss_map = DeviceBuffer(ctx, lib().gs_normal_loss_loss_map(self.h), np.float32, (height, width, 4))

    def set_lambda(self, lam: float, focal: float):
        check(lib().gs_normal_loss_set_lambda(self.h, lam, focal))

    def encode(self, stream=None):
        check(lib().gs_normal_loss_encode(self.h, stream))

    def read_loss(self) -> float:
        v = C.c_double()
        check(lib().gs_normal_loss_read(self.h, C.byref(v)))
        return v.value

    def close(self):
        if self.h:
            lib().gs_normal_loss_destroy(self.h)
            self.h = None


class Optimizer:
    """Optimizer (optim.rs:50-241)."""

    IDENTITY, EXP, SIGMOID = L.GS_ACT_IDENTITY, L.GS_ACT_EXP, L.GS_ACT_SIGMOID

    def __init__(self, ctx: GpuContext):
        self.ctx = ctx
        h = C.c_void_p()
        check(lib().gs_optim_create(ctx.h, C.byref(h)))
        self.h = h.value
        self._owned = True
        self.sizes: dict[str, int] = {}

    def add_class(self, name: str, n: int, activation: int, lr: float, grads: DeviceBuffer,
                  activated: DeviceBuffer):
        check(lib().gs_optim_add_class(self.h, name.encode(), n, activation, lr, grads.ptr, activated.ptr))
        self.sizes[name] = n

    def raw(self, name: str) -> DeviceBuffer:
        p = lib().gs_optim_class_raw(self.h, name.encode())
        if not p:
            raise KeyError(name)
        return DeviceBuffer(self.ctx, p, np.float32, self.sizes[name])

    def set_lr(self, name: str, lr: float):
        check(lib().gs_optim_set_lr(self.h, name.encode(), lr))

    def set_reg(self, name: str, reg: float):
        check(lib().gs_optim_set_reg(self.h, name.encode(), reg))

    def set_lr_schedule(self, name: str, base: float, final_factor: float, iters: int):
        """lr of the step with 0-based index i = base * final_factor**(i / iters)
        (trainer.rs:1149-1151), evaluated on the device by the step."""
        check(lib().gs_optim_set_lr_schedule(self.h, name.encode(), base, final_factor, iters))

    def current_lr(self, name: str) -> float:
        lr = C.c_float()
        check(lib().gs_optim_current_lr(self.h, name.encode(), C.byref(lr)))
        return float(lr.value)

    def export_state(self) -> bytes:
        """The exact optimizer state (raw, m, v, step count, rates) as bytes."""
        return _export_state(lib().gs_optim_state_bytes, lib().gs_optim_export_state, self.h)

    def import_state(self, blob: bytes):
        buf = C.create_string_buffer(bytes(blob), len(blob))
        check(lib().gs_optim_import_state(self.h, buf, len(blob)))

    def encode_step(self, stream=None):
        check(lib().gs_optim_encode_step(self.h, stream))

    def encode_activate(self, stream=None):
        check(lib().gs_optim_encode_activate(self.h, stream))

    def step_count(self) -> int:
        t = C.c_uint32()
        check(lib().gs_optim_step_count(self.h, C.byref(t)))
        return int(t.value)

    def snapshot_layout(self) -> list[tuple[str, int]]:
        """[(class name, raw bytes)] in add order (optim.rs:148-153)."""
        n = C.c_uint32()
        check(lib().gs_optim_class_count(self.h, C.byref(n)))
        out = []
        for k in range(n.value):
            name, nbytes = C.c_char_p(), C.c_uint64()
            check(lib().gs_optim_class_info(self.h, k, C.byref(name), C.byref(nbytes)))
            out.append((name.value.decode(), int(nbytes.value)))
        return out

    def encode_snapshot(self, dst: list, stream=None):
        """Copy every class's raw parameters into dst[k] (optim.rs:155-159)."""
        arr = (C.c_void_p * len(dst))(*[b.ptr for b in dst])
        check(lib().gs_optim_encode_snapshot(self.h, stream, arr, len(dst)))

    def encode_restore(self, src: list, stream=None):
        """Copy every class's raw parameters back from src[k] (optim.rs:161-165);
        the activated copies change only with encode_activate."""
        arr = (C.c_void_p * len(src))(*[b.ptr for b in src])
        check(lib().gs_optim_encode_restore(self.h, stream, arr, len(src)))

    def close(self):
        if self.h and self._owned:
            lib().gs_optim_destroy(self.h)
        self.h = None


def _export_state(size_fn, export_fn, h) -> bytes:
    n = C.c_uint64()
    check(size_fn(h, C.byref(n)))
    buf = C.create_string_buffer(int(n.value))
    check(export_fn(h, buf, n.value))
    return buf.raw


@dataclass
class TrainConfig:
    """TrainConfig subset on the path (trainer.rs:19-111, defaults :459-499)."""
    lam: float = 0.2
    dssim_scale: float = 0.5
    lr_pos: float = 1.6e-4
    lr_scales: float = 5e-3
    lr_quat: float = 1e-3
    lr_opacity: float = 5e-2
    lr_sh: float = 2.5e-3
    max_views_per_step: int = 4
    entries_per_view: int = 0
    digit_bits: int = 8
    binning: int = L.GS_BINNING_TIGHT
    pos_lr_final_factor: float = 0.0   # trainer.rs:477 default 0.01; 0 = no decay
    iters: int = 0                     # decay horizon (trainer.rs:460); 0 = no decay


class Trainer:
    """Trainer (trainer.rs:1139-1332) with plain / batched / graph steps."""

    PLAIN, BATCHED, GRAPH = L.GS_STEP_PLAIN, L.GS_STEP_BATCHED, L.GS_STEP_GRAPH

    def __init__(self, ctx: GpuContext, cfg: TrainConfig, init: dict, views: list):
        """init: activated arrays (positions, scales, quats, opacities, sh);
        views: list of (Camera, target_rgba (H, W, 4)) or (Camera, target,
        num_surfels): a view renders surfels [0, num_surfels) only (0 = all)."""
        self.ctx = ctx
        keep: list = []
        si = scene_input(init["positions"], init["scales"], init["quats"], init["opacities"], init["sh"],
                         keep)
        self.n = si.n
        tv = (L.TrainView * len(views))()
        self._targets = []
        self.sizes = []
        for k, view in enumerate(views):
            cam, tgt = view[0], view[1]
            tv[k].num_surfels = int(view[2]) if len(view) > 2 else 0
            t = np.ascontiguousarray(tgt, np.float32)
            H, W = t.shape[:2]
            self._targets.append(t)
            self.sizes.append((W, H))
            tv[k].width, tv[k].height = W, H
            tv[k].camera = cam.c()
            tv[k].target_rgba = t.ctypes.data
        c = L.TrainConfig(cfg.lam, cfg.dssim_scale, cfg.lr_pos, cfg.lr_scales, cfg.lr_quat, cfg.lr_opacity,
                          cfg.lr_sh, cfg.max_views_per_step, cfg.entries_per_view, cfg.digit_bits,
                          cfg.binning, cfg.pos_lr_final_factor, cfg.iters)
        h = C.c_void_p()
        check(lib().gs_trainer_create(ctx.h, C.byref(c), C.byref(si), len(views), tv, C.byref(h)))
        self.h = h.value
        self._targets = None  # uploaded; host copies no longer needed
        b = L.RasterBuffers()
        check(lib().gs_trainer_params(self.h, C.byref(b)))
        n = self.n
        f = np.float32
        self.positions = DeviceBuffer(ctx, b.positions, f, (n, 4))
        self.scales = DeviceBuffer(ctx, b.scales, f, (n, 2))
        self.quats = DeviceBuffer(ctx, b.quats, f, (n, 4))
        self.opacities = DeviceBuffer(ctx, b.opacities, f, n)
        self.sh = DeviceBuffer(ctx, b.sh, f, (n, 48))
        self.grad_pos = DeviceBuffer(ctx, b.grad_pos, f, (n, 4))
        self.grad_scales = DeviceBuffer(ctx, b.grad_scales, f, (n, 2))
        self.grad_quat = DeviceBuffer(ctx, b.grad_quat, f, (n, 4))
        self.grad_opacity = DeviceBuffer(ctx, b.grad_opacity, f, n)
        self.grad_sh = DeviceBuffer(ctx, b.grad_sh, f, (n, 48))

    def step(self, view_ids, mode: int = 0, stream=None):
        ids = (C.c_uint32 * len(view_ids))(*[int(v) for v in view_ids])
        check(lib().gs_trainer_step(self.h, stream, ids, len(view_ids), mode))

    def set_schedule(self, digit_bits: int, blocks_per_cu: int):
        check(lib().gs_trainer_set_schedule(self.h, digit_bits, blocks_per_cu))

    def optim(self) -> "Optimizer":
        """The trainer's optimizer (classes pos, scales, quat, opacity, sh), not owned."""
        h = C.c_void_p()
        check(lib().gs_trainer_optim(self.h, C.byref(h)))
        o = Optimizer.__new__(Optimizer)
        o.ctx, o.h, o._owned = self.ctx, h.value, False
        o.sizes = {"pos": 4 * self.n, "scales": 2 * self.n, "quat": 4 * self.n, "opacity": self.n,
                   "sh": 48 * self.n}
        return o

    def export_state(self) -> bytes:
        """Exact training state after the last step (gs_trainer_export_state)."""
        return _export_state(lib().gs_trainer_state_bytes, lib().gs_trainer_export_state, self.h)

    def import_state(self, blob: bytes):
        buf = C.create_string_buffer(bytes(blob), len(blob))
        check(lib().gs_trainer_import_state(self.h, buf, len(blob)))

    def set_mcmc_noise(self, noise: float):
        """MCMC exploration noise after every step (TrainConfig::mcmc_noise); 0 = off."""
        check(lib().gs_trainer_set_mcmc_noise(self.h, noise))

    def mcmc_relocate(self, dead: float = 0.005, split_ratio: float = 8.0, rng: int = 0x9E3779B97F4A7C15):
        """Trainer::mcmc_relocate; returns (pairs moved, advanced rng state)."""
        r = C.c_uint64(rng)
        npairs = C.c_uint32()
        check(lib().gs_trainer_mcmc_relocate(self.h, dead, split_ratio, C.byref(r), C.byref(npairs)))
        return int(npairs.value), int(r.value)

    def relocate_pairs(self, pairs):
        p = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 4)
        check(lib().gs_trainer_relocate_pairs(self.h, p.ctypes.data, p.shape[0]))

    def write_ply(self, path: str):
        """Baked 3DGS .ply of the current parameters (Trainer::read_scene +
        write_3dgs_ply)."""
        check(lib().gs_trainer_write_ply(self.h, os.fsencode(path)))

    def set_tiling(self, tile_size: int, fwd_batch: int, bwd_batch: int):
        check(lib().gs_trainer_set_tiling(self.h, tile_size, fwd_batch, bwd_batch))

    def set_overlap(self, enable: bool):
        check(lib().gs_trainer_set_overlap(self.h, 1 if enable else 0))

    def set_keep_grads(self, enable: bool):
        """Batched steps also store the summed parameter gradients in
        grad_* (plain steps always do); off by default."""
        check(lib().gs_trainer_set_keep_grads(self.h, 1 if enable else 0))

    def autotune(self, view_ids, reps: int = 2) -> dict:
        """Time the candidate schedules on real batched steps and keep the
        fastest; returns {"digit_bits", "blocks_per_cu" (0 = one workgroup
        per tile), "tile_size", "fwd_batch", "bwd_batch"}."""
        ids = (C.c_uint32 * len(view_ids))(*[int(v) for v in view_ids])
        chosen = (C.c_uint32 * 5)()
        check(lib().gs_trainer_autotune(self.h, ids, len(view_ids), reps, chosen))
        keys = ("digit_bits", "blocks_per_cu", "tile_size", "fwd_batch", "bwd_batch")
        return {k: int(v) for k, v in zip(keys, chosen)}

    def last_losses(self, k: int) -> np.ndarray:
        out = (C.c_double * (2 * k))()
        check(lib().gs_trainer_last_losses(self.h, out, k))
        return np.array(out[:]).reshape(k, 2)

    def render(self, view: int) -> np.ndarray:
        W, H = self.sizes[view]
        out = np.empty((H, W, 4), np.float32)
        check(lib().gs_trainer_render(self.h, view, out.ctypes.data))
        return out

    def entry_count(self) -> int:
        e = C.c_uint32()
        check(lib().gs_trainer_entry_count(self.h, C.byref(e)))
        return int(e.value)

    def digit_bits(self) -> int:
        b = C.c_uint32()
        check(lib().gs_trainer_digit_bits(self.h, C.byref(b)))
        return int(b.value)

    def set_profiling(self, enable: bool):
        check(lib().gs_trainer_set_profiling(self.h, 1 if enable else 0))

    def stage_times(self) -> tuple[dict, int]:
        """Summed ms per stage since the last call, and the profiled step count."""
        ms = (C.c_double * L.GS_STAGE_COUNT)()
        steps = C.c_uint32()
        check(lib().gs_trainer_stage_times(self.h, ms, C.byref(steps)))
        names = [lib().gs_stage_name(i).decode() for i in range(L.GS_STAGE_COUNT)]
        return dict(zip(names, ms[:])), int(steps.value)

    def close(self):
        if self.h:
            lib().gs_trainer_destroy(self.h)
            self.h = None
