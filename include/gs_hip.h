/*
 * gs_hip.h -- C ABI of the MI355X (gfx950) differentiable 2D-Gaussian-surfel
 * training step.
 *
 * Drop-in boundary for splattar's GPU op layer: every entry point below
 * replaces one method of the reference Rust API (crates/gs-wgpu,
 * crates/gs-kernels, first_attempt/crates/gs-train), cited per declaration.
 * The reference records deferred work on a wgpu CommandEncoder; here an
 * encode call enqueues asynchronously on a HIP stream (passed as void*, a
 * hipStream_t; NULL = the context's stream) and is capturable into a
 * hipGraph. Device buffers are exposed as raw pointers exactly where the
 * reference exposes wgpu::Buffer fields; the caller wires objects together
 * by passing one object's pointer to another, as in the reference.
 *
 * Errors: every call returns a gs_status; no abort crosses the ABI. The
 * reference panics on misuse (assert!) -- here that is GS_ERR_INVALID /
 * GS_ERR_CAPACITY with a message in gs_last_error() (thread-local).
 * Threading: a handle is not thread-safe; distinct handles on distinct
 * streams are independent. All device buffers stay allocated at capacity
 * for the handle's lifetime (no reallocation mid-training).
 */
#ifndef GS_HIP_H
#define GS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GS_ABI_VERSION 1

typedef enum gs_status {
    GS_OK = 0,
    GS_ERR_INVALID = 1,   /* bad argument / misuse (reference: assert! panic) */
    GS_ERR_CAPACITY = 2,  /* request exceeds a capacity fixed at creation */
    GS_ERR_OOM = 3,       /* device allocation failed */
    GS_ERR_HIP = 4,       /* HIP runtime error (message names the call) */
    GS_ERR_OVERFLOW = 5,  /* tile entries exceeded max_entries (reference: unchecked) */
    GS_ERR_NO_DEVICE = 6, /* no HIP device (reference: GpuError::NoAdapter, gs-wgpu/src/lib.rs:19-27) */
    GS_ERR_IO = 7         /* file i/o failed (reference: std::io::Error, gs-io/src/ply/writer.rs:17) */
} gs_status;

/* Message of the last failing call on this thread ("" if none). */
const char* gs_last_error(void);
/* ABI version (GS_ABI_VERSION) and build string. */
int gs_abi_version(void);
const char* gs_build_info(void);

/* ------------------------------------------------------------------ context
 * GpuContext (crates/gs-wgpu/src/context.rs:7-96): one device, one queue. */
typedef struct gs_context* gs_ctx;
gs_status gs_ctx_create(int device_index, gs_ctx* out);
void gs_ctx_destroy(gs_ctx ctx);
/* The context's stream (hipStream_t): the reference's single queue. */
void* gs_ctx_stream(gs_ctx ctx);
gs_status gs_ctx_synchronize(gs_ctx ctx);
/* Device name and compute-unit count (for logs / the autotuner key). */
gs_status gs_ctx_device_info(gs_ctx ctx, char* name, size_t name_len, int* compute_units);

/* ------------------------------------------------------------------ buffers
 * crates/gs-wgpu/src/buffers.rs: storage_empty / queue.write_buffer /
 * readback / encoder.clear_buffer. */
gs_status gs_buffer_alloc(gs_ctx ctx, size_t bytes, void** out); /* zero-filled */
gs_status gs_buffer_free(gs_ctx ctx, void* ptr);
/* queue.write_buffer: host -> device, ordered on the context stream. */
gs_status gs_write_buffer(gs_ctx ctx, void* dst, const void* src, size_t bytes);
/* buffers::readback (buffers.rs:24-51): blocking device -> host. */
gs_status gs_read_buffer(gs_ctx ctx, void* dst, const void* src, size_t bytes);
/* encoder.clear_buffer: zero `bytes` (multiple of 4) with a kernel
 * (graph-capturable; no memset node is recorded). */
gs_status gs_clear_buffer(gs_ctx ctx, void* stream, void* dst, size_t bytes);

/* ---------------------------------------------------------------- PrefixSum
 * crates/gs-wgpu/src/prefix_sum.rs:19-164: exclusive u32 scan in place,
 * wrapping mod 2^32, grand total into a 1-element device buffer. */
typedef struct gs_scan_s* gs_scan;
gs_status gs_scan_create(gs_ctx ctx, uint32_t capacity, gs_scan* out); /* PrefixSum::new */
void gs_scan_destroy(gs_scan scan);
uint32_t* gs_scan_data(gs_scan scan);   /* PrefixSum::data  (prefix_sum.rs) */
uint32_t* gs_scan_total(gs_scan scan);  /* PrefixSum::total */
/* PrefixSum::encode(ctx, encoder, n) (prefix_sum.rs:134-164); n <= capacity. */
gs_status gs_scan_encode(gs_scan scan, void* stream, uint32_t n);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the exclusive scan of gs_scan_encode (same kernels, same
 * results) over the caller's own buffers -- data[0..n) in place, the grand
 * total into *total (both device pointers on ctx's device; stream null = the
 * context stream). Calls no allocating routine: the partition sums live in a
 * fixed module-scope device array of ~256 KiB per device, held until the
 * process ends. n > GS_SCAN_NOALLOC_MAX_ITEMS is GS_ERR_CAPACITY. Every call
 * on one device shares that array, so calls that overlap on one device (two
 * streams, two threads) race; serialize them. */
#define GS_SCAN_NOALLOC_MAX_ITEMS (1u << 28)
gs_status gs_scan_encode_noalloc(gs_ctx ctx, void* stream, uint32_t* data, uint32_t n, uint32_t* total);

/* -------------------------------------------------------------- RadixSorter
 * crates/gs-wgpu/src/sort.rs:17-327: stable LSD sort of the first counts[0]
 * (key, u32 payload) pairs in place; the count lives in device memory.
 * Keys are 32- or 64-bit (key_bits); the reference sorts 32-bit keys. */
typedef struct gs_sorter_s* gs_sorter;
gs_status gs_sorter_create(gs_ctx ctx, uint32_t capacity, uint32_t key_bits, gs_sorter* out); /* RadixSorter::new */
void gs_sorter_destroy(gs_sorter s);
void* gs_sorter_keys(gs_sorter s);          /* RadixSorter::keys     (sort.rs:101-104) */
uint32_t* gs_sorter_payloads(gs_sorter s);  /* RadixSorter::payloads (sort.rs:106-109) */
uint32_t* gs_sorter_counts(gs_sorter s);    /* RadixSorter::counts   (sort.rs:111-114) */
/* RadixSorter::encode (sort.rs:117-119): sorts key bits [begin_bit,
 * end_bit) with 4- or 8-bit digits; the result is back in keys()/payloads().
 * A device count above the capacity sorts the first `capacity` pairs. */
gs_status gs_sorter_encode(gs_sorter s, void* stream, uint32_t begin_bit, uint32_t end_bit,
                           uint32_t digit_bits);

/* --------------------------------------------------------------- TileBinner
 * crates/gs-kernels/src/binning.rs:22-238. Items carry an inclusive
 * tile-space rect (x1 < x0 or y1 < y0 = culled) and a u32 depth key; encode
 * counts, scans, expands into 64-bit (tile << 32 | depth) keys with item-id
 * payloads, sorts stably and extracts per-tile [start, end) ranges. The
 * reference's payload is an entry index mapped through entry_items; here the
 * payload is the item id itself (same order: an item's entries lie in
 * distinct tiles). */
typedef struct gs_binner_s* gs_binner;
gs_status gs_binner_create(gs_ctx ctx, uint32_t max_items, uint32_t max_entries, uint32_t num_tiles,
                           gs_binner* out);                 /* TileBinner::new (binning.rs:46) */
void gs_binner_destroy(gs_binner b);
uint32_t* gs_binner_rects(gs_binner b);           /* pub rects  (binning.rs:27), 4 u32 per item */
uint32_t* gs_binner_depths(gs_binner b);          /* pub depths (binning.rs:28) */
uint32_t* gs_binner_ranges(gs_binner b);          /* pub ranges (binning.rs:32), 2 u32 per tile */
uint32_t* gs_binner_sorted_entries(gs_binner b);  /* sorted_entries() (binning.rs:148) */
uint64_t* gs_binner_sorted_keys(gs_binner b);     /* sorted 64-bit keys (no reference counterpart) */
uint32_t* gs_binner_entry_count(gs_binner b);     /* entry_count() (binning.rs:153) */
/* TileBinner::encode (binning.rs:159-238); digit_bits 4 or 8. */
gs_status gs_binner_encode(gs_binner b, void* stream, uint32_t num_items, uint32_t tiles_x,
                           uint32_t digit_bits);
/* Blocking: *overflowed = 1 if any encode since the last call produced more
 * entries than max_entries (entries beyond it were dropped, memory-safe). */
gs_status gs_binner_overflowed(gs_binner b, int* overflowed);

/* --------------------------------------------------------------- Rasterizer
 * crates/gs-kernels/src/rasterizer.rs:34-448. */
typedef struct gs_raster_s* gs_raster;

/* RasterCamera (rasterizer.rs:34-41): camera-to-world quat (xyzw, normalized
 * on use), centre, focal in pixels, active SH degree. */
typedef struct gs_raster_camera {
    float center[3];
    float quat[4];
    float focal;
    uint32_t sh_degree;
} gs_raster_camera;

/* SceneInput (rasterizer.rs:43-54): ACTIVATED parameters, host memory.
 * positions n*3, scales n*2 (world units), quats n*4 (xyzw, unnormalized),
 * opacities n (0..1), sh n*sh_coeffs*3 coefficient-major rgb (padded to 48
 * floats per surfel on upload). */
typedef struct gs_scene_input {
    const float* positions;
    const float* scales;
    const float* quats;
    const float* opacities;
    const float* sh;
    uint32_t sh_coeffs;
    uint32_t n;
} gs_scene_input;

/* The public device buffers of Rasterizer (rasterizer.rs:56-98). P = W*H. */
typedef struct gs_raster_buffers {
    float* positions;      /* capacity * 4 (xyz, w = 1) */
    float* scales;         /* capacity * 2 */
    float* quats;          /* capacity * 4 */
    float* opacities;      /* capacity */
    float* sh;             /* capacity * 48 */
    float* out_color;      /* P * 4: composited rgb, accumulated alpha */
    float* out_aux;        /* P * 4: final transmittance, alpha-weighted depth, 0, 0 */
    float* out_normal;     /* P * 4: alpha-weighted normal, 0 */
    uint32_t* out_ncontrib;/* P: tile-list index just past the last contributor */
    float* dl_dcolor;      /* P * 4: dL/d(rgb), w = dL/d(depth image) -- backward input */
    float* dl_dnormal;     /* P * 4: dL/d(normal image) -- backward input */
    float* grad_pos;       /* capacity * 4 */
    float* grad_scales;    /* capacity * 2 */
    float* grad_quat;      /* capacity * 4 */
    float* grad_opacity;   /* capacity */
    float* grad_sh;        /* capacity * 48 */
    float* grad_cam;       /* 16: dL/dR (9, column-major), dL/dC (3), dL/df, pad 3 */
} gs_raster_buffers;

/* Rasterizer::new (rasterizer.rs:108). */
gs_status gs_raster_create(gs_ctx ctx, uint32_t capacity, uint32_t width, uint32_t height,
                           uint32_t max_entries, gs_raster* out);
void gs_raster_destroy(gs_raster r);
gs_status gs_raster_get_buffers(gs_raster r, gs_raster_buffers* out);
/* Rasterizer::upload_scene (rasterizer.rs:359-386). */
gs_status gs_raster_upload_scene(gs_raster r, const gs_scene_input* scene);
/* pub lambda_dist (rasterizer.rs:60): depth-distortion weight for backward. */
gs_status gs_raster_set_lambda_dist(gs_raster r, float lambda_dist);
/* Radix digit width for the binning sort (4 or 8; autotuned by default). */
gs_status gs_raster_set_digit_bits(gs_raster r, uint32_t digit_bits);
/* Rasterizer::forward (rasterizer.rs:389-448): project, bin, sort, ranges,
 * composite. */
gs_status gs_raster_forward(gs_raster r, void* stream, const gs_raster_camera* camera,
                            uint32_t num_surfels);
/* Rasterizer::backward (rasterizer.rs:320-356): clears the gradient buffers
 * and differentiates colour, depth (dl_dcolor.w), normal and distortion
 * losses into grad_* and grad_cam. */
gs_status gs_raster_backward(gs_raster r, void* stream, uint32_t num_surfels);
#define GS_BWD_ACCUMULATE 1u  /* add into grad_* instead of overwriting */
#define GS_BWD_COLOR_ONLY 2u  /* only dL/d(rgb) is non-zero (faster kernel) */
#define GS_BWD_NO_CAMERA 4u   /* skip grad_cam */
gs_status gs_raster_backward_ex(gs_raster r, void* stream, uint32_t num_surfels, uint32_t flags);
/* Batched multi-view extension (no counterpart in the reference, whose
 * Rasterizer renders one view per forward; for callers that render or
 * differentiate several views of one scene at once: frame batches, multi-view
 * training loops). gs_raster_create_batch makes a rasterizer with room for
 * max_views views of at most width x height pixels and max_entries tile
 * entries each; its per-pixel buffers (out_*, dl_dcolor, dl_dnormal) hold a
 * batch's views back to back in view order (view j starts at pixel
 * sum_{k<j} w_k h_k) and grad_cam holds 16 floats per view. The single-view
 * calls work on it as on a gs_raster_create(width, height) rasterizer, which
 * is gs_raster_create_batch with max_views = 1. */
typedef struct gs_view {
    gs_raster_camera camera;
    uint32_t width, height;  /* <= the rasterizer's width x height */
    uint32_t num_surfels;    /* the view renders surfels [0, num_surfels) */
} gs_view;
gs_status gs_raster_create_batch(gs_ctx ctx, uint32_t capacity, uint32_t width, uint32_t height,
                                 uint32_t max_entries, uint32_t max_views, gs_raster* out);
/* Forward of n_views views: one launch per stage for the whole batch. */
gs_status gs_raster_forward_batch(gs_raster r, void* stream, const gs_view* views, uint32_t n_views);
/* Backward of the last gs_raster_forward_batch (flags as
 * gs_raster_backward_ex): grad_* receive the gradients summed over its views,
 * grad_cam[16 j ..] view j's camera gradients. */
gs_status gs_raster_backward_batch(gs_raster r, void* stream, uint32_t flags);
/* Blocking: entries produced by the last forward; GS_ERR_OVERFLOW if it
 * exceeded max_entries (the reference does not check, binning.rs:186). */
gs_status gs_raster_entry_count(gs_raster r, uint32_t* entries);

/* Binning internals of the last forward, for parity tests and tooling. */
typedef struct gs_raster_debug {
    const uint64_t* sorted_keys;    /* entry_count: (tile << 32) | f32 depth bits */
    const uint32_t* sorted_payloads;/* entry_count: surfel ids */
    const uint32_t* ranges;         /* tiles * 2 */
    const uint32_t* rects;          /* capacity * 4 */
    const uint32_t* depth_keys;     /* capacity */
    const uint32_t* offsets;        /* capacity: per-surfel tile counts (presort binning) or
                                       their exclusive scan (64-bit binning) */
    uint32_t num_tiles;
} gs_raster_debug;
/* Blocking (composes the 64-bit keys when the presort binning ran). */
gs_status gs_raster_debug_buffers(gs_raster r, gs_raster_debug* out);

/* Binning strategy. SORT64 and PRESORT give the reference's entry list
 * exactly (same order, ranges and 64-bit keys):
 *   GS_BINNING_SORT64:  one stable sort of 64-bit (tile << 32 | depth) keys
 *                       over every varying bit (the reference's composite
 *                       of two stable 32-bit sorts, binning.rs:5-8);
 *   GS_BINNING_PRESORT: stable sort of the N surfel depth keys, emission in
 *                       depth order, stable sort of 32-bit tile ids (default);
 *   GS_BINNING_TIGHT:   PRESORT with each surfel's tile rect cut to the tiles
 *                       its conservative alpha >= 1/255 footprint box reaches.
 *                       It drops only entries that contribute to no pixel, so
 *                       images and gradients equal PRESORT's; rects, entry
 *                       counts and ncontrib (list positions) differ. */
#define GS_BINNING_SORT64 0u
#define GS_BINNING_PRESORT 1u
#define GS_BINNING_TIGHT 2u
gs_status gs_raster_set_binning(gs_raster r, uint32_t mode);

/* Per-quadrant footprint culling in the rasterizer kernels (default on).
 * No counterpart in the reference: a verification switch. Culling only skips
 * records that reach no pixel of a wave's 8x8 block, so images, ncontrib and
 * gradients are the same with it off. */
gs_status gs_raster_set_culling(gs_raster r, int enable);

/* Decision trace of the next single-view forwards (no counterpart in the
 * reference: a verification aid, so that an FP64 evaluation can replay the
 * f32 path's discrete decisions). For each pixel of the window [x0, x0 + w) x
 * [y0, y0 + h), in row-major window order, the forward records its
 * contributions in compositing order: lists[p * cap + k] = surfel index
 * (bits 0..29) | ray branch taken (bit 30) | alpha clamped at 0.995 (bit 31),
 * and counts[p] = the number of contributions (those past cap are counted,
 * not stored). Needs the default 16-px tiles; the trace forward is a
 * separate kernel instantiation, so untraced forwards are unaffected.
 * w * h * cap = 0 turns the trace off. Allocates the trace buffers. */
gs_status gs_raster_set_trace(gs_raster r, uint32_t x0, uint32_t y0, uint32_t w, uint32_t h, uint32_t cap);
/* Blocking: device pointers of the trace (null when off). */
gs_status gs_raster_trace_buffers(gs_raster r, const uint32_t** lists, const uint32_t** counts);

/* ----------------------------------------------------------------- SsimLoss
 * first_attempt/crates/gs-train/src/ssim.rs:25-298: fused L1 + D-SSIM,
 * writes dL/d(image) rgb (w = 0) into dl_out. Loss =
 * (1-lambda) mean|x-y| + lambda * dssim_scale * mean(1 - SSIM); the
 * reference's dssim_scale is 0.5 (ssim.rs:4), the default here too. */
typedef struct gs_loss_s* gs_loss;
/* SsimLoss::new(ctx, w, h, lambda, img, dl_out) (ssim.rs:63-70). */
gs_status gs_loss_create(gs_ctx ctx, uint32_t width, uint32_t height, float lambda,
                         const float* img, float* dl_out, gs_loss* out);
void gs_loss_destroy(gs_loss l);
float* gs_loss_target(gs_loss l);   /* pub target (ssim.rs:31): P * 4, rgb used */
gs_status gs_loss_set_dssim_scale(gs_loss l, float scale);
gs_status gs_loss_encode(gs_loss l, void* stream);   /* SsimLoss::encode (ssim.rs:231) */
/* SsimLoss::read_losses (ssim.rs:282-298): blocking (l1_mean, dssim_mean),
 * dssim_mean = dssim_scale * mean(1 - SSIM). */
gs_status gs_loss_read(gs_loss l, double* l1_mean, double* dssim_mean);

/* -------------------------------------------------------- 3DGS .ply export
 * first_attempt/crates/gs-io/src/ply/writer.rs:17-79 write_3dgs_ply(path,
 * count, positions, scales, quats, opacities, sh) over device buffers (the
 * activated layout: positions float4, scales float2, quats xyzw, opacity,
 * SH 48 floats coefficient-major): header, then per surfel 59 floats
 * x y z | f_dc rgb | f_rest channel-major | logit(clamp(o, 1e-6, 1-1e-6)) |
 * ln(max(s, 1e-9)) x2, -10 | w x y z. The rows are built on the device;
 * blocking; GS_ERR_IO when the file cannot be written. */
gs_status gs_write_3dgs_ply(gs_ctx ctx, const char* path, uint32_t count, const float* positions,
                            const float* scales, const float* quats, const float* opacities, const float* sh);
/* -------------------------------------------------------------------- MCMC
 * shaders/mcmc.wgsl:39-58 add_noise over device buffers: pos_raw float4
 * (raw positions), scales_act float2, opac_act (activated); amp = sigma *
 * mean scale * exp(-opacity_k * opacity). */
gs_status gs_mcmc_add_noise(gs_ctx ctx, void* stream, uint32_t n, uint32_t seed, float sigma, float opacity_k,
                           float* pos_raw, const float* scales_act, const float* opac_act);
/* Host part of Trainer::mcmc_relocate (trainer.rs:1558-1650) on host arrays:
 * dead = activated opacity < dead_threshold; one xorshift64 draw per dead
 * surfel picks an alive target in proportion to opacity, without
 * replacement (a colliding pick is skipped); then, when split_ratio > 0 and
 * at least 16 surfels are alive, alive surfels larger than split_ratio x the
 * median alive size (max of the two scales) take the remaining dead slots,
 * largest first (mode 1). pairs receives (dead, alive, mode, 0) x n_pairs
 * (room for 4 n). No device work. */
gs_status gs_mcmc_sample_pairs(const float* opacity, const float* scales, uint32_t n, float dead_threshold,
                               float split_ratio, uint64_t* rng_state, uint32_t* pairs, uint32_t* n_pairs);

/* The rows alone: count x 59 floats into device memory `rows`, enqueued on
 * the stream (no reference counterpart: callers streaming the export). */
gs_status gs_ply_pack_rows(gs_ctx ctx, void* stream, uint32_t count, const float* positions, const float* scales,
                           const float* quats, const float* opacities, const float* sh, float* rows);

/* ------------------------------------------------------------------ Tracker
 * crates/gs-track/src/lib.rs:89-337 + shaders/photometric.wgsl: photometric
 * camera tracking against a frozen map held in a rasterizer. Per iteration:
 * forward at the current pose, softened L1 against the target (dL/d(render) =
 * clamp(d / 1e-4, -1, 1) / 3P), backward, then the Riemannian gradient
 * (antisymmetric part of R^T dL/dR; dL/dC), Adam over the six (lr_rot,
 * lr_center * depth; betas 0.9 / 0.999, eps 1e-10) and the right
 * perturbation R' = R exp([-step]x), C' = C - step. The whole loop runs on
 * the device (no per-iteration readback); the pose returned is the best one
 * seen. Uses the rasterizer's images and gradient buffers. */
typedef struct gs_tracker_s* gs_tracker;
/* TrackConfig (lib.rs:46-87); defaults 60, 2e-3, 2e-3, 1e-4, 10. */
typedef struct gs_track_config {
    uint32_t iterations;
    float lr_rot;
    float lr_center;
    float min_improvement;
    uint32_t patience;
} gs_track_config;
/* TrackReport (lib.rs:17-44): iterations run, coverage of the last
 * iteration, best residual; plus the pose of the last evaluated iteration
 * (not part of the reference report). */
typedef struct gs_track_report {
    uint32_t iterations;
    float coverage;
    float best;
    float last_quat[4];
    float last_center[3];
} gs_track_report;
/* Tracker::new(ctx, raster, width, height); width x height = the raster's. */
gs_status gs_tracker_create(gs_ctx ctx, gs_raster raster, uint32_t width, uint32_t height, gs_tracker* out);
void gs_tracker_destroy(gs_tracker t);
/* Tracker::set_target (lib.rs:159-165): host RGBA frame, W x H x 4 floats. */
gs_status gs_tracker_set_target(gs_tracker t, const float* frame_rgba);
/* Device pointer of the target (for frames already in device memory). */
float* gs_tracker_target(gs_tracker t);
/* Tracker::track (lib.rs:167-222): camera in/out (returns the best pose);
 * residuals (optional, cfg.iterations floats) receives the trace. Blocking. */
gs_status gs_tracker_track(gs_tracker t, uint32_t surfels, gs_raster_camera* camera, float depth,
                           const gs_track_config* cfg, float* residuals, gs_track_report* report);
/* The photometric pass alone over the rasterizer's current render
 * (photometric.wgsl:47-73 with gain / bias on the target): writes the
 * rasterizer's dl_dcolor; residual / coverage (optional) make it blocking. */
gs_status gs_tracker_photometric(gs_tracker t, void* stream, float gain, float bias, double* residual,
                                 double* coverage);

/* --------------------------------------------------------------- NormalLoss
 * first_attempt/crates/gs-train/src/normal_loss.rs + shaders/normal_loss.wgsl:
 * 52-145: 2DGS normal consistency, L = lambda/(W H) * sum_q alpha_q
 * (1 - nhat_q . N_d(q)) with N_d from forward differences of the mean-depth
 * points (alpha and the orientation sign detached). Reads the rasterizer's
 * out_color (.w = alpha), out_aux (.y = depth_acc) and out_normal; writes
 * dl_dnormal and dl_dcolor.w (dL/d depth_acc; rgb untouched). Encode after
 * the colour loss and before the rasterizer backward (without
 * GS_BWD_COLOR_ONLY). */
typedef struct gs_normal_loss_s* gs_normal_loss;
/* NormalLoss::new(ctx, w, h, out_color, out_aux, out_normal, dl_dnormal,
 * dl_dcolor) (normal_loss.rs:40-120); all float4 per pixel. lambda starts 0. */
gs_status gs_normal_loss_create(gs_ctx ctx, uint32_t width, uint32_t height, const float* out_color,
                                const float* out_aux, const float* out_normal, float* dl_dnormal,
                                float* dl_dcolor, gs_normal_loss* out);
void gs_normal_loss_destroy(gs_normal_loss nl);
/* pub loss_map (normal_loss.rs:33): P float4, x = alpha_q (1 - nhat . N_d). */
float* gs_normal_loss_loss_map(gs_normal_loss nl);
/* NormalLoss::set_lambda(ctx, lambda, focal) (normal_loss.rs:123-135): the
 * kernel weight is lambda / (W H); alpha_min 0.2. */
gs_status gs_normal_loss_set_lambda(gs_normal_loss nl, float lambda, float focal);
gs_status gs_normal_loss_encode(gs_normal_loss nl, void* stream);  /* NormalLoss::encode */
/* Blocking: the loss lambda/(W H) * sum of the loss map, summed in double
 * over per-tile f32 partials (no reference counterpart: the reference
 * leaves the loss map to the caller). */
gs_status gs_normal_loss_read(gs_normal_loss nl, double* loss);

/* ---------------------------------------------------------------- Optimizer
 * first_attempt/crates/gs-train/src/optim.rs:50-241. */
typedef struct gs_optim_s* gs_optim;
#define GS_ACT_IDENTITY 0
#define GS_ACT_EXP 1      /* raw = log(activated): scales */
#define GS_ACT_SIGMOID 2  /* raw = logit(activated): opacity */
gs_status gs_optim_create(gs_ctx ctx, gs_optim* out);   /* Optimizer::new (optim.rs:58) */
void gs_optim_destroy(gs_optim o);
/* Optimizer::add_class (optim.rs:87-141): grads = dL/d(activated). */
gs_status gs_optim_add_class(gs_optim o, const char* name, uint32_t n, int activation, float lr,
                             const float* grads, float* activated);
float* gs_optim_class_raw(gs_optim o, const char* name);  /* class(name).raw (optim.rs:41) */
/* Optimizer::set_lr / set_reg (optim.rs:171-177): take effect at the next
 * encode, whose stream carries the upload into device memory (the rates,
 * schedules and regularisers live there, so a captured step graph replays
 * whatever was set before its launch; changing them while the encoding
 * stream is being captured is GS_ERR_INVALID). set_lr clears a schedule. */
gs_status gs_optim_set_lr(gs_optim o, const char* name, float lr);    /* optim.rs:171 */
gs_status gs_optim_set_reg(gs_optim o, const char* name, float reg);  /* optim.rs:175 */
/* The exponential decay Trainer::step applies to the position rate
 * (trainer.rs:1149-1151, schedule_t trainer.rs:407-416): the step with
 * 0-based index iter (the optimizer's step count before it) uses
 * lr = base * final_factor^(iter / iters), computed on the device by the
 * step itself. The reference sets it from the host every iteration. */
gs_status gs_optim_set_lr_schedule(gs_optim o, const char* name, float base, float final_factor,
                                   uint32_t iters);
/* Blocking: the learning rate the last step used for the class. */
gs_status gs_optim_current_lr(gs_optim o, const char* name, float* lr);
/* Optimizer::encode_step (optim.rs:200-227): one Adam step on every class,
 * then the activated copies. One fused launch. */
gs_status gs_optim_encode_step(gs_optim o, void* stream);
gs_status gs_optim_encode_activate(gs_optim o, void* stream);  /* optim.rs:230-241 */
gs_status gs_optim_step_count(gs_optim o, uint32_t* t);        /* blocking */
/* Optimizer::snapshot_layout (optim.rs:148-153): the classes in add order,
 * each with its name and raw-parameter size in bytes (n * 4). */
gs_status gs_optim_class_count(gs_optim o, uint32_t* count);
gs_status gs_optim_class_info(gs_optim o, uint32_t index, const char** name, uint64_t* bytes);
/* Optimizer::encode_snapshot / encode_restore (optim.rs:155-165): copy every
 * class's raw parameters into bufs[k] (device buffers of at least the class's
 * bytes, add order; count = number of classes), or back from them, as
 * kernels on the stream (graph-capturable). Restore leaves the activated
 * copies untouched: follow it with gs_optim_encode_activate, as the
 * reference's trainer does (trainer.rs:1891-1893). */
gs_status gs_optim_encode_snapshot(gs_optim o, void* stream, float* const* bufs, uint32_t count);
gs_status gs_optim_encode_restore(gs_optim o, void* stream, const float* const* bufs, uint32_t count);
/* Exact resume (extends optim.rs:148-165, which snapshots raw parameters
 * only): the whole optimizer state -- class descriptors, step count, bias
 * products, learning rates, schedules, regularisers and every class's raw,
 * m and v -- as one host blob of gs_optim_state_bytes bytes. Blocking.
 * Import checks the class names, sizes and activations, restores the state
 * bit for bit and re-activates, so training continues exactly where the
 * export left it. */
gs_status gs_optim_state_bytes(gs_optim o, uint64_t* bytes);
gs_status gs_optim_export_state(gs_optim o, void* dst, uint64_t bytes);
gs_status gs_optim_import_state(gs_optim o, const void* src, uint64_t bytes);

/* ------------------------------------------------------------------ Trainer
 * Trainer::step (first_attempt/crates/gs-train/src/trainer.rs:1139-1332):
 * render -> L1 + D-SSIM -> backward -> Adam with raw -> activated chains.
 * A step here covers a set of views whose losses are summed (each view's
 * loss normalised by its own pixel count, ssim.rs:84) before ONE Adam step.
 *   GS_STEP_PLAIN:   every kernel launched per view (the reference's shape);
 *   GS_STEP_BATCHED: every stage launched once for all views of the step;
 *   GS_STEP_GRAPH:   GS_STEP_BATCHED captured once into a hipGraph and
 *                    replayed (view descriptors live in device memory). */
typedef struct gs_trainer_s* gs_trainer;

typedef struct gs_train_view {
    uint32_t width, height;
    gs_raster_camera camera;
    const float* target_rgba;   /* host, width*height*4 (rgb used) */
    uint32_t num_surfels;       /* surfels [0, num_surfels) render in this view (0 = all):
                                   frames of a growing scene; no reference counterpart */
} gs_train_view;

typedef struct gs_train_config {
    float lambda;               /* D-SSIM weight (trainer.rs:470: 0.2) */
    float dssim_scale;          /* 0.5 = reference (1-SSIM)/2, 1.0 = (1-SSIM) */
    float lr_pos, lr_scales, lr_quat, lr_opacity, lr_sh;   /* trainer.rs:471-476 */
    uint32_t max_views_per_step;
    uint32_t entries_per_view;  /* binning capacity per view (0 = 64 * n) */
    uint32_t digit_bits;        /* 4 or 8; 0 = 8 until gs_trainer_autotune picks */
    uint32_t binning;           /* GS_BINNING_SORT64 (0), GS_BINNING_PRESORT or GS_BINNING_TIGHT */
    float pos_lr_final_factor;  /* trainer.rs:477 (0.01); with iters > 0 the position rate */
    uint32_t iters;             /* decays as lr_pos * factor^(step / iters) (trainer.rs:1149-1151); */
                                /* iters = 0 keeps it constant */
} gs_train_config;

#define GS_STEP_PLAIN 0u
#define GS_STEP_BATCHED 1u
#define GS_STEP_GRAPH 2u

/* Trainer::new (trainer.rs:722-792): `init` holds ACTIVATED values; raw
 * parameters are log(scales), logit(opacity), identity otherwise. Targets
 * are uploaded once and stay resident. */
gs_status gs_trainer_create(gs_ctx ctx, const gs_train_config* cfg, const gs_scene_input* init,
                            uint32_t n_views, const gs_train_view* views, gs_trainer* out);
void gs_trainer_destroy(gs_trainer t);
gs_status gs_trainer_step(gs_trainer t, void* stream, const uint32_t* view_ids, uint32_t n_ids,
                          uint32_t mode);
/* Blocking. No counterpart in the reference (its tile size, LDS batch,
 * digit width and dispatch shape are fixed). Stage 1 tries every work tiling
 * (tile_size, fwd_batch, bwd_batch) in {(16, 256, 64), (16, 256, 128),
 * (16, 128, 64), (32, 256, 128), (32, 512, 256), (8, 128, 128), (8, 64, 64)} (16-px only
 * unless binning is GS_BINNING_TIGHT) x raster scheduling {persistent tile
 * queue with 8 or 4 workgroups per CU, one workgroup per tile} at 8-bit
 * digits; stage 2 tries radix digit widths {8, 4} at the stage-1 winner.
 * Each candidate runs one settling batched step on `view_ids` and `reps`
 * more timed one by one, over two interleaved rounds; the fastest step wins,
 * candidates within 2 % of it counting as ties that go to the earliest in the
 * order above (16-px tiles, batches 256 / 64, 8 workgroups per CU, 8-bit
 * digits first). A candidate whose binning overflows the entry capacity is skipped
 * (GS_ERR_OVERFLOW if all do). The steps are real training steps (no
 * candidate changes a step's result beyond float-atomic order), but the
 * parameters, Adam moments and step count are saved first and restored
 * afterwards, so tuning leaves the training state as it found it (and
 * gradients / losses of the last step are those of a tuning step). chosen (optional, 5
 * entries) = digit bits, workgroups per CU (0 = one per tile), tile_size,
 * fwd_batch, bwd_batch. A captured step graph is recaptured afterwards. */
gs_status gs_trainer_autotune(gs_trainer t, const uint32_t* view_ids, uint32_t n_ids, uint32_t reps,
                              uint32_t* chosen);
/* The schedule gs_trainer_autotune chooses among, set directly (no
 * counterpart in the reference): radix digit bits 4 or 8; blocks_per_cu
 * workgroups per CU for the persistent raster passes, 0 = one per tile. */
gs_status gs_trainer_set_schedule(gs_trainer t, uint32_t digit_bits, uint32_t blocks_per_cu);
/* Work tiling, set directly (no counterpart in the reference): the binning
 * and raster passes work on tile_size x tile_size pixel tiles (8, 16 or 32;
 * coverage stays the reference's 16-px rects, so images do not change) and
 * stage fwd_batch / bwd_batch records per LDS batch. Supported:
 * tile 8: fwd 64 | 128, bwd 64 | 128; tile 16: fwd 128 | 256, bwd 64 | 128;
 * tile 32: fwd 256 | 512, bwd 128 | 256. Tiles other than 16 need
 * GS_BINNING_TIGHT. Default (16, 256, 64). */
gs_status gs_trainer_set_tiling(gs_trainer t, uint32_t tile_size, uint32_t fwd_batch, uint32_t bwd_batch);
/* Overlapped batched steps (default off; no counterpart in the reference):
 * with two or more views the step splits them in halves and runs each
 * half's L1 + D-SSIM on a second stream beside the other half's raster pass
 * (forward of the second half, backward of the first). Same result as the
 * in-order step up to float-atomic order; profiled steps always run in
 * order so their stage times stay separable. On MI355X the persistent raster
 * passes already fill every CU, so the split costs more than it hides on the
 * 1M-surfel 4 x 1080p step and gains nothing on the 16-frame 720p batch. */
gs_status gs_trainer_set_overlap(gs_trainer t, int enable);
/* Parameter gradients of batched steps (no counterpart in the reference,
 * whose trainer keeps them in internal buffers): a batched or graph step runs
 * the geometry backward and the Adam update as one kernel and, by default,
 * never stores the summed parameter gradients. With enable != 0 it also
 * writes them to the grad_* buffers of gs_trainer_buffers, as plain steps
 * always do (callers that inspect or check gradients). Default 0. */
gs_status gs_trainer_set_keep_grads(gs_trainer t, int enable);
/* The trainer's optimizer (Trainer::optim, trainer.rs): its five classes
 * "pos", "scales", "quat", "opacity", "sh" in that order. Use it to change
 * rates and schedules (they take effect at the next step, graph replays
 * included) or read the current rate. Owned by the trainer; a step
 * enqueued on a caller's stream must be complete before the optimizer's
 * blocking calls read it (the gs_trainer_*_state calls below wait for it). */
gs_status gs_trainer_optim(gs_trainer t, gs_optim* out);
/* Exact resume of a trainer: gs_optim_export_state / import_state of its
 * optimizer after the last step (on any stream) has finished. */
gs_status gs_trainer_state_bytes(gs_trainer t, uint64_t* bytes);
gs_status gs_trainer_export_state(gs_trainer t, void* dst, uint64_t bytes);
gs_status gs_trainer_import_state(gs_trainer t, const void* src, uint64_t bytes);
/* MCMC exploration noise (TrainConfig::mcmc_noise, trainer.rs:58, 1256-1272;
 * shaders/mcmc.wgsl): after every step's optimizer update the raw positions
 * receive noise of amplitude noise * the step's position rate * mean scale *
 * exp(-5 opacity), seeded by the step index; graph replays included. 0 =
 * off (the default). */
gs_status gs_trainer_set_mcmc_noise(gs_trainer t, float noise);
/* Trainer::mcmc_relocate (trainer.rs:1558-1690): one blocking call -- reads
 * the activated opacities (and scales when split_ratio > 0), samples the
 * (dead, alive, mode) pairs as gs_mcmc_sample_pairs, then moves the data on
 * the device (shaders/relocate.wgsl: parameters, split opacity, scale
 * shrink, moments zeroed) and re-activates. rng_state is the reference's
 * xorshift64 state (Trainer::rng), advanced in place. */
gs_status gs_trainer_mcmc_relocate(gs_trainer t, float dead_threshold, float split_ratio, uint64_t* rng_state,
                                   uint32_t* n_pairs);
/* The device half alone for host pairs (dead, alive, mode, 0) x n_pairs (no
 * reference counterpart as a separate call); GS_ERR_INVALID unless every
 * index is in range and used at most once. Blocking. */
gs_status gs_trainer_relocate_pairs(gs_trainer t, const uint32_t* pairs, uint32_t n_pairs);
/* Trainer::read_scene + gs_io::write_3dgs_ply (trainer.rs:1938-1950,
 * writer.rs:17-79): the baked .ply of the current activated parameters. */
gs_status gs_trainer_write_ply(gs_trainer t, const char* path);
/* Blocking: per-view (l1_mean, dssim_mean) of the last step, 2 per view. */
gs_status gs_trainer_last_losses(gs_trainer t, double* out, uint32_t n_ids);
/* Blocking: render view `view` with the current parameters into rgba (W*H*4). */
gs_status gs_trainer_render(gs_trainer t, uint32_t view, float* rgba);
/* Activated parameters (positions n*4, scales n*2, quats n*4, opacities n,
 * sh n*48) and parameter gradients of the last step. The step's losses are
 * colour-only, so its forward writes out_color, out_aux = (T_final, 0, 0, 0)
 * and out_ncontrib, and leaves out_normal untouched. */
gs_status gs_trainer_params(gs_trainer t, gs_raster_buffers* out);
/* Blocking: entries of the last step (sum over its views); GS_ERR_OVERFLOW
 * if any view exceeded its binning capacity. */
gs_status gs_trainer_entry_count(gs_trainer t, uint32_t* entries);
/* Selected radix digit width after autotuning (4 or 8). */
gs_status gs_trainer_digit_bits(gs_trainer t, uint32_t* bits);

/* Per-stage timing (no reference counterpart; the reference's GpuTimer
 * timestamp scopes, crates/gs-wgpu/src/profile.rs:16-133): with profiling
 * on, plain and batched steps record a HIP event on their stream before
 * each stage; gs_trainer_stage_times (blocking) returns the summed
 * milliseconds per stage since the last call and the number of profiled
 * steps, then resets. Graph steps are not instrumented. */
#define GS_STAGE_PROJECT 0u
#define GS_STAGE_SCAN 1u
#define GS_STAGE_EMIT 2u
#define GS_STAGE_SORT 3u
#define GS_STAGE_RANGES 4u
#define GS_STAGE_RASTER_FWD 5u
#define GS_STAGE_LOSS 6u
#define GS_STAGE_GRAD_CLEAR 7u
#define GS_STAGE_RASTER_BWD 8u
#define GS_STAGE_GEOM_BWD 9u
#define GS_STAGE_ADAM 10u
#define GS_STAGE_COUNT 11u
const char* gs_stage_name(uint32_t stage);
gs_status gs_trainer_set_profiling(gs_trainer t, int enable);
gs_status gs_trainer_stage_times(gs_trainer t, double* ms, uint32_t* steps);

#ifdef __cplusplus
}
#endif
#endif /* GS_HIP_H */
