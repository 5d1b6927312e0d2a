/*
 * CPU oracle for the differentiable 2D-Gaussian-surfel training step.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker the HIP path is
 * compared against; nothing in the product (splattar_amd/, include/) links,
 * loads or calls it. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may use it.
 *
 * Contents (every function cites the reference code it restates):
 *   1. FP64 restatement of crates/gs-cpu-ref (forward.rs, backward.rs,
 *      math.rs, scene.rs): prepare / covers_pixel / evaluate / render /
 *      distortion_loss / gradients_full, quat_to_mat / quat_grad /
 *      sh_basis / sh_basis_grad. Literal O(P*N) mode and a tile-binned mode
 *      that walks exactly the same surfels in the same order.
 *   2. An fp32 mirror of the integer stages of the GPU path
 *      (surfel_prep_fwd.wgsl:64-132 rect + depth key, tile_binning.wgsl
 *      count/expand/ranges, the stable (tile||depth) sort of binning.rs:5-8)
 *      with the operation order pinned, for bit-exact key/offset/range checks.
 *   3. FP64 L1 + D-SSIM loss and its gradient (first_attempt/crates/gs-train/
 *      src/shaders/ssim.wgsl:1-143, ssim.rs:1-4,84).
 *   4. f32 Adam with the raw->activated chain (first_attempt/crates/gs-train/
 *      src/shaders/adam.wgsl:33-74, optim.rs:179-197).
 *
 * Parity pins (what checks each part; see tests/):
 *   - quat_grad / sh_basis_grad / triple products: central FD at the
 *     reference's own points (math.rs:200-267).
 *   - sh_basis: known answers of crates/gs-core/src/sh.rs:80-126.
 *   - render / gradients_full: central FD over every parameter class at the
 *     reference gradcheck seeds (crates/gs-cpu-ref/tests/gradcheck.rs:214-326),
 *     closed forms (single fronto-parallel surfel, two stacked surfels,
 *     T_final = prod(1-a), sum w = 1 - T_final), literal == binned.
 *   - integer mirror: std::stable_sort / exclusive scan / brute-force
 *     range extraction, the reference property-test case lists.
 *   - SSIM: FD in FP64 on small images, SSIM(x,x) = 1, blur self-adjointness.
 *   - Adam: the reference's own CPU loop (gs-train/tests/adam.rs:37-98).
 */
#ifndef SPLATTAR_ORACLE_CPUREF_H
#define SPLATTAR_ORACLE_CPUREF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Scene in FP64, flat arrays. sh is coefficient-major rgb:
 * sh[(i*ncoef + k)*3 + ch], ncoef = (deg+1)^2. */
typedef struct {
    int64_t n;
    int32_t sh_degree;
    const double* pos;      /* n*3 */
    const double* scales;   /* n*2, activated */
    const double* quat;     /* n*4, xyzw, unnormalized */
    const double* opacity;  /* n, activated */
    const double* sh;       /* n*ncoef*3 */
} oref_scene;

typedef struct {
    double center[3];
    double quat[4];  /* camera-to-world, xyzw, unnormalized */
    double focal;
    int32_t width;
    int32_t height;
    /* Optional pixel window [x0, x0+w) x [y0, y0+h) of the full image (all 0
     * = whole image). Outputs and dL/d(image) inputs are window-sized; the
     * model (principal point, tile rects) stays that of the full image. Used
     * to time a bounded sample for the CPU baseline. */
    int32_t crop_x0, crop_y0, crop_w, crop_h;
} oref_camera;

typedef struct {
    double* pos;        /* n*3 */
    double* scales;     /* n*2 */
    double* quat;       /* n*4 */
    double* opacity;    /* n */
    double* sh;         /* n*ncoef*3 */
    double cam_center[3];
    double cam_quat[4];
    double focal;
} oref_grads;

/* ---- math.rs ---- */
void oref_quat_to_mat(const double q[4], double r_colmajor[9]);
void oref_quat_grad(const double q[4], const double dl_dr_colmajor[9], double out[4]);
int oref_sh_basis(int degree, const double d[3], double* out /* 16 */);
int oref_sh_basis_grad(int degree, const double d[3], double* out /* 16*3 */);

/* ---- forward.rs ---- */
/* mode 0 = literal O(P*N) walk, 1 = tile-binned walk (identical results).
 * Any output pointer may be NULL. color/normal are P*3, alpha/depth P. */
int oref_render(const oref_scene* s, const oref_camera* cam, int mode,
                double* color, double* alpha, double* depth, double* normal,
                int64_t* n_hits /* optional: total evaluated (pixel, surfel) pairs */);
double oref_distortion_loss(const oref_scene* s, const oref_camera* cam, int mode);

/* ---- backward.rs ---- */
/* dl_ddepth / dl_dnormal may be NULL (zero). Grad arrays are overwritten. */
int oref_gradients(const oref_scene* s, const oref_camera* cam, int mode,
                   const double* dl_dcolor, const double* dl_ddepth,
                   const double* dl_dnormal, double lambda_dist, oref_grads* g);

/* Test aid (no reference counterpart): the forward walk's discrete
 * decisions within relative band `rel` of a threshold (alpha skip / clamp,
 * ray vs screen branch, near plane, transmittance stop, rect truncation,
 * on-screen, near cull, f32 depth-key ties). flags (n, optional) marks the
 * surfels involved; returns their count of decisions. */
int64_t oref_marginal_decisions(const oref_scene* s, const oref_camera* cam, double rel, uint8_t* flags);

/* Test aid (no reference counterpart): flags[i] = 1 for the surfels that
 * composite at any of the npix pixels (x, y) = xy[2k], xy[2k+1]. */
int oref_pixel_contributors(const oref_scene* s, const oref_camera* cam, int64_t npix, const int32_t* xy,
                            uint8_t* flags);

/* Number of surfels that survive the near cull and their prepare() order
 * (indices sorted by (f32 depth bits, idx)). order may be NULL. */
/* forward.rs:104-123 tile rects (FP64) of every surfel; a test aid. */
int oref_prepare_rects(const oref_scene* s, const oref_camera* cam, int32_t* rects, uint8_t* on_screen,
                       uint8_t* culled);
/* normal_loss.wgsl:52-145 (2DGS normal consistency) in FP64. */
void oref_normal_loss(int w, int h, double focal, double lambda, double alpha_min, const double* color4,
                      const double* aux4, const double* normal4, double* loss_map, double* dl_dnormal,
                      double* dl_ddepth);
/* The faithful FP64 walk's contribution lists in oref_replay's format. */
int oref_trace(const oref_scene* s, const oref_camera* cam, int64_t cap, uint32_t* lists, uint32_t* counts);
/* Decision replay: FP64 forward (+ backward) along given per-pixel
 * contribution lists (surfel | ray << 30 | clamped << 31), see cpuref.cpp. */
int oref_replay(const oref_scene* s, const oref_camera* cam, const uint32_t* lists, const uint32_t* counts,
                int64_t cap, double* color, double* alpha, const double* dl_dcolor, const double* dl_ddepth,
                const double* dl_dnormal, double lambda_dist, oref_grads* g);

int64_t oref_prepare_order(const oref_scene* s, const oref_camera* cam, int64_t* order);

/* ---- fp32 mirror of the integer stages ---- */
typedef struct {
    float rot[9];     /* camera-to-world rotation, column-major (rot[3*k+j] = R[j][k]) */
    float center[3];
    float focal;
    uint32_t width, height;
    uint32_t tiles_x, tiles_y;
} oref_cam32;

/* surfel_prep_fwd.wgsl:64-132 with the operation order of the HIP project
 * kernel. rects n*4 (x0,y0,x1,y1; culled = 1,1,0,0), keys n, counts n. */
void oref_prep_f32(int64_t n, const float* pos4, const float* scales2, const float* quat4,
                   const oref_cam32* cam, uint32_t* rects, uint32_t* depth_keys,
                   uint32_t* counts);
/* Exclusive scan (mod 2^32); returns the total. */
uint32_t oref_exclusive_scan_u32(int64_t n, const uint32_t* in, uint32_t* out);
/* Expansion (tile_binning.wgsl:46-66) into 64-bit (tile<<32 | depth) keys with
 * surfel-id payloads, entries in surfel order, row-major within each rect.
 * tile_base is added to every tile id (multi-view batches). Returns E. */
int64_t oref_emit_keys(int64_t n, const uint32_t* rects, const uint32_t* depth_keys,
                       uint32_t tiles_x, uint32_t tile_base, uint64_t* keys,
                       uint32_t* payloads, int64_t max_entries);
/* Stable sort of (key, payload) by key. */
void oref_stable_sort_u64(int64_t n, uint64_t* keys, uint32_t* payloads);
/* tile_binning.wgsl:89-103 over sorted 64-bit keys: ranges[2*t] = start,
 * ranges[2*t+1] = end, tile t = key >> 32 (minus tile_base); empty tiles 0,0. */
void oref_tile_ranges(int64_t e, const uint64_t* keys, uint32_t num_tiles,
                      uint32_t tile_base, uint32_t* ranges);

/* ---- ssim.wgsl / ssim.rs ---- */
/* img/target are P*3 (rgb). dl_dimg P*3 (may be NULL). Returns the loss
 * (1-lambda)*mean|x-y| + lambda*dssim_scale*mean(1-SSIM); l1_mean and
 * ssim_mean are optional outputs. */
double oref_l1_ssim(int w, int h, const double* img, const double* target, double lambda,
                    double dssim_scale, double* dl_dimg, double* l1_mean, double* ssim_mean);
/* Zero-padded separable 11-tap blur of one P-sized channel map. */
void oref_blur(int w, int h, const double* in, double* out);

/* ---- adam.wgsl / optim.rs ---- */
/* One f32 Adam step (t is the 1-based step index) with activation mode
 * 0 = identity, 1 = exp, 2 = sigmoid. Updates raw/m/v in place and writes
 * activated (may be NULL). */
void oref_adam_step_f32(int64_t n, float* raw, const float* grad_act, float* m, float* v,
                        float* activated, float lr, float reg, int activation, int t);

/* ---- test-input generators ---- */
/* xorshift32 (shifts 13/17/5) stream of crates/gs-wgpu/tests/sort.rs:18-25;
 * out[i] = (i-th draw) % modulus (modulus 0 = full range). */
void oref_xorshift32_fill(uint32_t seed, int64_t n, uint32_t modulus, uint32_t* out);

/* ---- timing helper for the CPU baseline ---- */
int oref_num_threads(void);
/* Compiler and target features the library was built with. */
const char* oref_build_info(void);

#ifdef __cplusplus
}
#endif
#endif
