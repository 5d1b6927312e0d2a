// Internal declarations shared by the gfx950 kernels and the C-ABI runtime.
// Everything here is MI355X-only HIP: wave64, hardware FP32 atomics, no
// alternative backends.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <string>

#include "gs_hip.h"

namespace gs {

// ------------------------------------------------------------ model constants
// crates/gs-cpu-ref/src/scene.rs:17-21 and raster_common.wgsl:34-39.
constexpr float kAlphaSkip = 1.0f / 255.0f;
constexpr float kAlphaClamp = 0.995f;
constexpr float kTTerminate = 1e-4f;
constexpr float kNearDepth = 0.05f;
constexpr int kTile = 16;        // coverage granularity: part of the forward model
// Work-tile sizes of the binning and raster passes (pixels). Independent of
// kTile: a record's footprint box is clipped to the pixels of its kTile rect,
// so lists of any work-tile size hold exactly the records the 16-px model
// lets reach each pixel.
constexpr bool valid_work_tile(uint32_t t) { return t == 8 || t == 16 || t == 32; }
constexpr int kShFloats = 48;    // SH storage padded to degree 3 (rasterizer.rs:13-15)
// Depth-presort payload with packed tile counts: item << kPackBits | count.
// A record's 16-px rect has radius <= 64 px, so it spans at most 9 x 9 work
// tiles of 16 px (5 x 5 of 32 px): counts stay <= 81 < 2^kPackBits, and
// packing needs work tiles >= 16 px and fewer than 2^(32 - kPackBits) items.
constexpr uint32_t kPackBits = 7;
constexpr uint32_t kPackMask = (1u << kPackBits) - 1u;
constexpr int kWave = 64;

// Per-surfel, per-view raster record written by the project kernel: the
// ray-splat setup in the form the per-pixel evaluation needs. With
// s = (-(x+0.5-cx)/f, (y+0.5-cy)/f, 1) the negated pixel ray, the Cramer
// numerators of forward.rs:136-147 are linear in s:
//   D  = det[tu, tv, s] = M.s,  M = tu x tv
//   Nu = det[-c, tv, s] = A.s,  A = tv x c
//   Nv = det[tu, -c, s] = B.s,  B = c x tu
//   Nt = det[tu, tv, -c] = -M.c (per surfel constant)
// so u = A.s/D, v = B.s/D, t = Nt/D.
struct __align__(16) Record {
    float4 m_nt;     // M.xyz, Nt
    float4 a_px;     // A.xyz, projected center x (pixels)
    float4 b_py;     // B.xyz, projected center y (pixels)
    float4 col_op;   // clamped SH color rgb, opacity
    float4 nrm_dep;  // camera-facing unit normal xyz, center depth
    float4 box;      // conservative alpha >= 1/255 footprint (x0, y0, x1, y1), pixel centres
    float4 ell0;     // footprint ellipse: centre (X, Y) in pixels, quadratic form h00, h01
    float4 ell1;     // h11, 1/h00, 1/h11, screen-disk radius (pixels)
};
static_assert(sizeof(Record) == 128, "record layout");

// Coefficient-space gradient accumulator per (view, surfel), summed over the
// pixels of every tile the surfel reaches (raster_bwd), chained to parameter
// space once per surfel (geom_bwd). Slots:
enum GradSlot : int {
    kGA = 0,        // dL/dA (3)
    kGB = 3,        // dL/dB (3)
    kGM = 6,        // dL/dM (3)
    kGCol = 9,      // dL/dcolor (3)
    kGOp = 12,      // dL/dopacity
    kGPx = 13,      // dL/dpx
    kGPy = 14,      // dL/dpy
    kGNt = 15,      // dL/dNt        (depth / distortion losses only)
    kGDep = 16,     // dL/ddepth     (screen-branch depth, depth losses only)
    kGNrm = 17,     // dL/dnormal (3) (normal loss only)
    kGradStride = 20,
};
// Slots per (view, surfel) in memory: colour-only backward passes (no depth,
// distortion or normal terms) keep the first 16.
constexpr int grad_stride(bool aux) { return aux ? (int)kGradStride : 16; }

// Camera as the kernels see it (crates/gs-kernels/src/shaders/
// raster_common.wgsl:5-20 RasterCamera), one per view.
struct __align__(16) CamParams {
    float rot[9];          // camera-to-world R, column-major: rot[3k + j] = R[j][k]
    float center[3];
    float focal;
    float cx, cy;          // principal point = (W/2, H/2)
    uint32_t width, height;
    uint32_t tiles_x, tiles_y;  // work tiles (tile_size px) of this view
    uint32_t tile_base;    // first global tile id of this view (batched views)
    uint32_t pix_base;     // first pixel of this view in batched image buffers
    uint32_t rec_base;     // first record (= view * N) of this view
    uint32_t sh_degree;
    float lambda_dist;
    uint32_t target_base;  // first pixel of this view's loss target (trainer)
    uint32_t tile_size;    // work-tile size in pixels (8, 16 or 32)
    uint32_t rtiles_x, rtiles_y;  // kTile (16-px) tiles of the forward model's rects
    uint32_t n_surfels;    // surfels [0, n_surfels) take part in this view
    uint32_t pad[3];
};

// ------------------------------------------------------------- launch shape

// The dispatcher deals a grid's workgroups round-robin over the 8 XCDs, each
// with its own L2: workgroup b runs on XCD b % 8. Blocks that share data
// (neighbouring tiles, the views of one surfel range) should share an L2, so
// kernels map the hardware index b to a logical block with this bijection of
// [0, n): XCD x takes the contiguous logical range of the x-th 1/8 of n.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_swizzle(uint32_t b, uint32_t n) {
    const uint32_t q = n / kXcds, r = n % kXcds;
    const uint32_t x = b % kXcds, j = b / kXcds;
    return x < r ? x * (q + 1) + j : r * (q + 1) + (x - r) * q + j;
}

// ------------------------------------------------------------- device math

// The projection quantities that decide tile rects and depth keys are
// computed with a pinned operation order and no FMA contraction, so the CPU
// fp32 mirror (oracle/cpuref.cpp oref_prep_f32) reproduces them bit for bit.
struct CamFrame {
    float cx_, cy_, cz_;     // c = R^T (p - C)
    float tux, tuy, tuz;     // tu = R^T (R(q) e0 s_u)
    float tvx, tvy, tvz;     // tv = R^T (R(q) e1 s_v)
    float xx, xy, xz;        // p - C (world)
    float qx, qy, qz, qw;    // normalized quaternion
    float qn;                // |q|
};

__device__ __forceinline__ float pinned_dot3(const float* a, float x, float y, float z) {
#pragma clang fp contract(off)
    float p0 = a[0] * x;
    float p1 = a[1] * y;
    float p2 = a[2] * z;
    return (p0 + p1) + p2;
}

// surfel_prep_fwd.wgsl:168-181 (camera-space centre and tangent axes).
__device__ __forceinline__ CamFrame cam_frame(const CamParams& cam, float4 p, float2 s, float4 q) {
#pragma clang fp contract(off)
    CamFrame f;
    f.xx = p.x - cam.center[0];
    f.xy = p.y - cam.center[1];
    f.xz = p.z - cam.center[2];
    f.cx_ = pinned_dot3(cam.rot + 0, f.xx, f.xy, f.xz);
    f.cy_ = pinned_dot3(cam.rot + 3, f.xx, f.xy, f.xz);
    f.cz_ = pinned_dot3(cam.rot + 6, f.xx, f.xy, f.xz);
    float qq = ((q.x * q.x + q.y * q.y) + q.z * q.z) + q.w * q.w;
    f.qn = __fsqrt_rn(qq);
    f.qx = __fdiv_rn(q.x, f.qn);
    f.qy = __fdiv_rn(q.y, f.qn);
    f.qz = __fdiv_rn(q.z, f.qn);
    f.qw = __fdiv_rn(q.w, f.qn);
    const float x = f.qx, y = f.qy, z = f.qz, w = f.qw;
    float a0 = 1.0f - 2.0f * (y * y + z * z);
    float a1 = 2.0f * (x * y + w * z);
    float a2 = 2.0f * (x * z - w * y);
    float b0 = 2.0f * (x * y - w * z);
    float b1 = 1.0f - 2.0f * (x * x + z * z);
    float b2 = 2.0f * (y * z + w * x);
    float ux = a0 * s.x, uy = a1 * s.x, uz = a2 * s.x;
    float vx = b0 * s.y, vy = b1 * s.y, vz = b2 * s.y;
    f.tux = pinned_dot3(cam.rot + 0, ux, uy, uz);
    f.tuy = pinned_dot3(cam.rot + 3, ux, uy, uz);
    f.tuz = pinned_dot3(cam.rot + 6, ux, uy, uz);
    f.tvx = pinned_dot3(cam.rot + 0, vx, vy, vz);
    f.tvy = pinned_dot3(cam.rot + 3, vx, vy, vz);
    f.tvz = pinned_dot3(cam.rot + 6, vx, vy, vz);
    return f;
}

// surfel_prep_fwd.wgsl:196-224: projected centre, conservative 16-px tile rect
// (radius min(3 f (|tu|+|tv|)/depth + 3, 64)), depth key. Returns false if
// the rect is degenerate (off screen); depth >= near is the caller's check.
struct RectOut {
    float px, py;
    uint32_t x0, y0, x1, y1;
    bool on_screen;
};

__device__ __forceinline__ RectOut tile_rect(const CamParams& cam, const CamFrame& f, float depth) {
#pragma clang fp contract(off)
    RectOut r;
    float lu = __fsqrt_rn((f.tux * f.tux + f.tuy * f.tuy) + f.tuz * f.tuz);
    float lv = __fsqrt_rn((f.tvx * f.tvx + f.tvy * f.tvy) + f.tvz * f.tvz);
    r.px = cam.cx + __fdiv_rn(cam.focal * f.cx_, depth);
    r.py = cam.cy - __fdiv_rn(cam.focal * f.cy_, depth);
    float rad = __fdiv_rn((3.0f * cam.focal) * (lu + lv), depth) + 3.0f;
    rad = rad < 64.0f ? rad : 64.0f;
    const float fw = (float)cam.width, fh = (float)cam.height;
    r.on_screen = r.px + rad >= 0.0f && r.px - rad < fw && r.py + rad >= 0.0f && r.py - rad < fh;
    const float mx = (float)(cam.rtiles_x - 1), my = (float)(cam.rtiles_y - 1);
    auto cl = [](float v, float hi) {
        float c = v > 0.0f ? v : 0.0f;
        c = c < hi ? c : hi;
        return (uint32_t)c;
    };
    r.x0 = cl(__fdiv_rn(r.px - rad, 16.0f), mx);
    r.y0 = cl(__fdiv_rn(r.py - rad, 16.0f), my);
    r.x1 = cl(__fdiv_rn(r.px + rad, 16.0f), mx);
    r.y1 = cl(__fdiv_rn(r.py + rad, 16.0f), my);
    return r;
}

// SH constants, crates/gs-core/src/sh.rs:6-18 (f32 values).
constexpr float kShC0 = 0.28209479f;
constexpr float kShC1 = 0.48860252f;
constexpr float kShC2_0 = 1.0925484f, kShC2_1 = -1.0925484f, kShC2_2 = 0.31539157f,
                kShC2_3 = -1.0925484f, kShC2_4 = 0.5462742f;
constexpr float kShC3_0 = -0.5900436f, kShC3_1 = 2.8906114f, kShC3_2 = -0.4570458f,
                kShC3_3 = 0.37317634f, kShC3_4 = -0.4570458f, kShC3_5 = 1.4453057f,
                kShC3_6 = -0.5900436f;

// Basis values b_k(dir) for degree <= 3 (math.rs:127-156); returns count.
__device__ __forceinline__ int sh_basis(int degree, float x, float y, float z, float* b) {
    b[0] = kShC0;
    if (degree < 1) return 1;
    b[1] = -kShC1 * y;
    b[2] = kShC1 * z;
    b[3] = -kShC1 * x;
    if (degree < 2) return 4;
    const float xx = x * x, yy = y * y, zz = z * z;
    b[4] = kShC2_0 * x * y;
    b[5] = kShC2_1 * y * z;
    b[6] = kShC2_2 * (2.0f * zz - xx - yy);
    b[7] = kShC2_3 * x * z;
    b[8] = kShC2_4 * (xx - yy);
    if (degree < 3) return 9;
    b[9] = kShC3_0 * y * (3.0f * xx - yy);
    b[10] = kShC3_1 * x * y * z;
    b[11] = kShC3_2 * y * (4.0f * zz - xx - yy);
    b[12] = kShC3_3 * z * (2.0f * zz - 3.0f * xx - 3.0f * yy);
    b[13] = kShC3_4 * x * (4.0f * zz - xx - yy);
    b[14] = kShC3_5 * z * (xx - yy);
    b[15] = kShC3_6 * x * (xx - 3.0f * yy);
    return 16;
}

// sum_k w_k d b_k / d dir (math.rs:159-191) for the basis functions of the
// degree, accumulated term by term in k order (no 48-float table of
// gradients held at once).
__device__ __forceinline__ float3 sh_dir_grad(int degree, float x, float y, float z, const float* w) {
    float3 d = make_float3(0.0f, 0.0f, 0.0f);
    auto add = [&](float gx, float gy, float gz, float s) {
        d.x += gx * s;
        d.y += gy * s;
        d.z += gz * s;
    };
    if (degree < 1) return d;
    add(0.0f, -kShC1, 0.0f, w[1]);
    add(0.0f, 0.0f, kShC1, w[2]);
    add(-kShC1, 0.0f, 0.0f, w[3]);
    if (degree < 2) return d;
    const float xx = x * x, yy = y * y, zz = z * z;
    add(kShC2_0 * y, kShC2_0 * x, 0.0f, w[4]);
    add(0.0f, kShC2_1 * z, kShC2_1 * y, w[5]);
    add(kShC2_2 * (-2.0f * x), kShC2_2 * (-2.0f * y), kShC2_2 * (4.0f * z), w[6]);
    add(kShC2_3 * z, 0.0f, kShC2_3 * x, w[7]);
    add(kShC2_4 * (2.0f * x), kShC2_4 * (-2.0f * y), 0.0f, w[8]);
    if (degree < 3) return d;
    add(kShC3_0 * (6.0f * x * y), kShC3_0 * (3.0f * xx - 3.0f * yy), 0.0f, w[9]);
    add(kShC3_1 * (y * z), kShC3_1 * (x * z), kShC3_1 * (x * y), w[10]);
    add(kShC3_2 * (-2.0f * x * y), kShC3_2 * (4.0f * zz - xx - 3.0f * yy), kShC3_2 * (8.0f * y * z), w[11]);
    add(kShC3_3 * (-6.0f * x * z), kShC3_3 * (-6.0f * y * z), kShC3_3 * (6.0f * zz - 3.0f * xx - 3.0f * yy),
        w[12]);
    add(kShC3_4 * (4.0f * zz - 3.0f * xx - yy), kShC3_4 * (-2.0f * x * y), kShC3_4 * (8.0f * x * z), w[13]);
    add(kShC3_5 * (2.0f * x * z), kShC3_5 * (-2.0f * y * z), kShC3_5 * (xx - yy), w[14]);
    add(kShC3_6 * (3.0f * xx - 3.0f * yy), kShC3_6 * (-6.0f * x * y), 0.0f, w[15]);
    return d;
}

__device__ __forceinline__ float3 cross3(float3 a, float3 b) {
    return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// Negated pixel ray s for pixel (x, y) (forward.rs:32-40 with s = -d).
__device__ __forceinline__ void pixel_s(const CamParams& cam, int x, int y, float& sx, float& sy) {
    sx = -(((float)x + 0.5f - cam.cx) / cam.focal);
    sy = ((float)y + 0.5f - cam.cy) / cam.focal;
}

// One pixel-surfel evaluation (forward.rs:136-169), shared verbatim by the
// forward and backward kernels so the backward replays the forward's alpha
// decisions bit for bit. Returns alpha (0 when skipped).
struct HitEval {
    float alpha;     // min(o * ghat, 0.995), 0 if below 1/255
    float ghat;
    float u, v, t;   // ray-branch coordinates (t = centre depth on the screen branch)
    float inv_d;     // 1 / D (ray branch)
    bool ray;
    bool clamped;    // o * ghat > 0.995
};

// Exponent threshold below which a record cannot reach alpha >= 1/255:
// o * exp(e) >= 1/255 needs e >= -ln(255 o); the 1e-3 slack (a 0.1 % factor)
// dwarfs the error of __expf, so skipping e < lthr never drops a
// contribution the exact test would keep.
__device__ __forceinline__ float record_lthr(float opacity) {
    return opacity * 255.0f > 1.0f ? -(logf(opacity * 255.0f) + 1e-3f) : 1.0f;  // 1 > any e: never
}

// Conservative footprint of a record: the pixels at which it can reach
// alpha >= 1/255 lie in the screen low-pass disk dx^2 + dy^2 <= L around
// (px, py) united with the image of the ray-branch disk u^2 + v^2 <= 2L,
// L = ln(255 opacity) + 2e-3. With s = (sx, sy, 1), the latter is the conic
// s^T Q s <= 0, Q = A A^T + B B^T - 2L M M^T (u = A.s/D, v = B.s/D, D = M.s).
// When its quadratic part Q2 is positive definite it is an ellipse with
// centre c = -Q2^-1 (Q02, Q12) and minimum q_c = Q22 + (Q02, Q12).c; in pixel
// offsets d from the centre it reads d^T H d <= 1 with
// H = diag(-1/f, 1/f) Q2 diag(-1/f, 1/f) / (-q_c). Its bounding box comes
// from the dual conic adj(Q): the tangent lines sx = k solve
// adj00 - 2k adj02 + k^2 adj22 = 0 (likewise for sy). Otherwise the ray
// part is unbounded and the footprint is the whole image. Computed in double
// once per (view, surfel) at projection; tests add 1 px of margin.
constexpr float kFar = 1e18f;

__device__ inline void record_footprint(Record& r, const CamParams& cam) {
    const float op = r.col_op.w;
    if (!(op * 255.0f > 1.0f)) {  // never contributes: empty box, empty ellipse, no disk
        r.box = make_float4(kFar, kFar, -kFar, -kFar);
        r.ell0 = make_float4(kFar, kFar, 1.0f, 0.0f);
        r.ell1 = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
        return;
    }
    const double L = log((double)op * 255.0) + 2e-3;
    const double rs = sqrt(L) + 0.5;
    double x0 = r.a_px.w - rs, x1 = r.a_px.w + rs, y0 = r.b_py.w - rs, y1 = r.b_py.w + rs;
    const double k = 2.0 * L;
    const double A[3] = {r.a_px.x, r.a_px.y, r.a_px.z};
    const double B[3] = {r.b_py.x, r.b_py.y, r.b_py.z};
    const double M[3] = {r.m_nt.x, r.m_nt.y, r.m_nt.z};
    double Q[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) Q[i][j] = A[i] * A[j] + B[i] * B[j] - k * M[i] * M[j];
    // The per-pixel test evaluates D = M.s, Nu = A.s, Nv = B.s in f32 (two
    // FMAs each, on f32 s), so the conic value it sees, Nu^2 + Nv^2 - 2L D^2,
    // is off by up to a few f32 roundings of its terms: |err| <= eps
    // |s|^T Qa |s| with Qa_ij = |A_i A_j| + |B_i B_j| + 2L |M_i M_j|. A
    // Laplacian argument gives |s|^T Qa |s| <= s^T diag(rowsum Qa) s, so the
    // conic Q' = Q - eps diag(rowsum Qa) contains every pixel the f32 test
    // can accept. For well-conditioned records this grows the ellipse by a
    // fraction of a percent; a near-degenerate one (thin surfel seen edge-on)
    // becomes a larger ellipse or unbounded instead of a region that rounding
    // has made arbitrary.
    constexpr double kConicEps = 2.5e-7;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        double row = 0.0;
#pragma unroll
        for (int j = 0; j < 3; ++j) row += fabs(A[i] * A[j]) + fabs(B[i] * B[j]) + k * fabs(M[i] * M[j]);
        Q[i][i] -= kConicEps * row;
    }
    const double f = cam.focal;
    const double det2 = Q[0][0] * Q[1][1] - Q[0][1] * Q[0][1];
    const double mag = Q[0][0] * Q[1][1] + Q[0][1] * Q[0][1];
    // det2 / mag ~ 1 / cond(H): the ellipse test reads H in f32, which keeps
    // the long axis (the small eigenvalue) only while cond(H) stays well
    // below 1 / (f32 unit roundoff); past 1e5 the ray part is treated as
    // unbounded.
    bool bounded = Q[0][0] > 0.0 && det2 > 1e-5 * mag;
    double ex = kFar, ey = kFar, h00 = 1.0, h01 = 0.0, h11 = 1.0;  // empty ellipse by default
    if (bounded) {
        const double a00 = Q[1][1] * Q[2][2] - Q[1][2] * Q[1][2];
        const double a11 = Q[0][0] * Q[2][2] - Q[0][2] * Q[0][2];
        const double a02 = Q[0][1] * Q[1][2] - Q[0][2] * Q[1][1];
        const double a12 = Q[0][1] * Q[0][2] - Q[0][0] * Q[1][2];
        const double dx = a02 * a02 - a00 * det2, dy = a12 * a12 - a11 * det2;
        bounded = dx >= 0.0 && dy >= 0.0;
        if (bounded) {
            const double sdx = sqrt(dx), sdy = sqrt(dy);
            const double sx_lo = (a02 - sdx) / det2, sx_hi = (a02 + sdx) / det2;
            const double sy_lo = (a12 - sdy) / det2, sy_hi = (a12 + sdy) / det2;
            // X = cx - f sx (pixel centre), Y = cy + f sy.
            x0 = fmin(x0, cam.cx - f * sx_hi - 1.0);
            x1 = fmax(x1, cam.cx - f * sx_lo + 1.0);
            y0 = fmin(y0, cam.cy + f * sy_lo - 1.0);
            y1 = fmax(y1, cam.cy + f * sy_hi + 1.0);
        }
        const double csx = -(Q[1][1] * Q[0][2] - Q[0][1] * Q[1][2]) / det2;
        const double csy = -(Q[0][0] * Q[1][2] - Q[0][1] * Q[0][2]) / det2;
        const double qc = Q[2][2] + Q[0][2] * csx + Q[1][2] * csy;
        if (bounded && qc < 0.0) {
            const double sc = 1.0 / (-qc * f * f);
            ex = cam.cx - f * csx;
            ey = cam.cy + f * csy;
            h00 = Q[0][0] * sc;
            h01 = -Q[0][1] * sc;
            h11 = Q[1][1] * sc;
        }
    }
    const bool finite = x0 == x0 && y0 == y0 && x1 == x1 && y1 == y1 && ex == ex && ey == ey &&
                        h00 == h00 && h01 == h01 && h11 == h11 && h00 > 0.0 && h11 > 0.0;
    if (!bounded || !finite) {  // whole image: box unbounded, H = 0 accepts every block
        r.box = make_float4(-kFar, -kFar, kFar, kFar);
        r.ell0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        r.ell1 = make_float4(0.0f, 0.0f, 0.0f, (float)rs);
        return;
    }
    r.box = make_float4((float)x0, (float)y0, (float)x1, (float)y1);
    r.ell0 = make_float4((float)ex, (float)ey, (float)h00, (float)h01);
    r.ell1 = make_float4((float)h11, (float)(1.0 / h00), (float)(1.0 / h11), (float)rs);
}

// True when the record's footprint can reach the pixel-centre rectangle
// [x0, x1] x [y0, y1]: box overlap, then the screen disk or the ellipse
// against the rectangle grown by 1 px. For the ellipse, the rectangle
// contains the centre or the minimum of d^T H d over its four edges is <= 1
// (a convex quadratic is minimised on the boundary when the centre is
// outside).
__device__ __forceinline__ bool footprint_hits(const Record& r, float x0, float y0, float x1, float y1) {
    const float4 b = r.box;
    if (!(b.z >= x0 && b.x <= x1 && b.w >= y0 && b.y <= y1)) return false;
    x0 -= 1.0f;
    y0 -= 1.0f;
    x1 += 1.0f;
    y1 += 1.0f;
    const float px = r.a_px.w, py = r.b_py.w, rd = r.ell1.w;
    const float qx = fminf(fmaxf(px, x0), x1) - px, qy = fminf(fmaxf(py, y0), y1) - py;
    if (qx * qx + qy * qy <= rd * rd) return true;
    const float dx0 = x0 - r.ell0.x, dx1 = x1 - r.ell0.x, dy0 = y0 - r.ell0.y, dy1 = y1 - r.ell0.y;
    if (dx0 <= 0.0f && dx1 >= 0.0f && dy0 <= 0.0f && dy1 >= 0.0f) return true;
    const float h00 = r.ell0.z, h01 = r.ell0.w, h11 = r.ell1.x, i00 = r.ell1.y, i11 = r.ell1.z;
    auto quad = [&](float dx, float dy) { return (h00 * dx + 2.0f * h01 * dy) * dx + h11 * dy * dy; };
    const float ya = fminf(fmaxf(-h01 * dx0 * i11, dy0), dy1);
    const float yb = fminf(fmaxf(-h01 * dx1 * i11, dy0), dy1);
    const float xa = fminf(fmaxf(-h01 * dy0 * i00, dx0), dx1);
    const float xb = fminf(fmaxf(-h01 * dy1 * i00, dx0), dx1);
    const float m = fminf(fminf(quad(dx0, ya), quad(dx1, yb)), fminf(quad(xa, dy0), quad(xb, dy1)));
    return m <= 1.0f;
}

// Bit by * (TW / 8) + bx is set when the record's footprint reaches 8x8
// pixel block (bx, by) of the TW x TW work tile whose top-left pixel is
// (x0, y0).
template <int TW>
__device__ __forceinline__ uint32_t block_mask(const Record& r, uint32_t x0, uint32_t y0) {
    constexpr int NB = TW / 8;
    uint32_t m = 0;
#pragma unroll
    for (int q = 0; q < NB * NB; ++q) {
        const float qx = (float)(x0 + (q % NB) * 8), qy = (float)(y0 + (q / NB) * 8);
        if (footprint_hits(r, qx + 0.5f, qy + 0.5f, qx + 7.5f, qy + 7.5f)) m |= 1u << q;
    }
    return m;
}

// Cuts a footprint box (pixel centres) to the pixels of the kTile rect
// (x0, y0, x1, y1, inclusive): pixels outside the rect are not reached in the
// forward model (forward.rs:104-123), whatever the footprint.
__device__ __forceinline__ float4 clip_box_to_rect(float4 b, uint4 rect) {
    b.x = fmaxf(b.x, (float)(rect.x * kTile) + 0.5f);
    b.y = fmaxf(b.y, (float)(rect.y * kTile) + 0.5f);
    b.z = fminf(b.z, (float)(rect.z * kTile) + (kTile - 0.5f));
    b.w = fminf(b.w, (float)(rect.w * kTile) + (kTile - 0.5f));
    return b;
}

// The work tiles (size ts px, tiles_x x tiles_y of them) whose pixel
// centres ts t + 0.5 .. ts t + ts - 0.5 meet the box; x0 > x1 (or y0 > y1)
// when none do.
__device__ __forceinline__ uint4 box_tiles(const float4 b, uint32_t ts, uint32_t tiles_x, uint32_t tiles_y) {
    const float bx0 = fminf(fmaxf(b.x, -1e6f), 1e6f), bx1 = fminf(fmaxf(b.z, -1e6f), 1e6f);
    const float by0 = fminf(fmaxf(b.y, -1e6f), 1e6f), by1 = fminf(fmaxf(b.w, -1e6f), 1e6f);
    const float inv = 1.0f / (float)ts, hi = (float)ts - 0.5f;
    const int tx0 = max(0, (int)ceilf((bx0 - hi) * inv));
    const int tx1 = min((int)tiles_x - 1, (int)floorf((bx1 - 0.5f) * inv));
    const int ty0 = max(0, (int)ceilf((by0 - hi) * inv));
    const int ty1 = min((int)tiles_y - 1, (int)floorf((by1 - 0.5f) * inv));
    if (tx0 > tx1 || ty0 > ty1) return make_uint4(1u, 1u, 0u, 0u);
    return make_uint4((uint32_t)tx0, (uint32_t)ty0, (uint32_t)tx1, (uint32_t)ty1);
}

// 1/x as v_rcp_f32 (1 ulp) refined by one Newton step r += r (1 - x r),
// instead of the IEEE division sequence; exact for powers of two (1/1 = 1).
__device__ __forceinline__ float rcp_refined(float x) {
    const float r0 = __builtin_amdgcn_rcpf(x);
    return fmaf(r0, fmaf(-x, r0, 1.0f), r0);
}

__device__ __forceinline__ HitEval eval_hit(const Record& r, float sx, float sy, float pxc, float pyc,
                                            float lthr = -INFINITY) {
    HitEval h;
    // D = M.s, Nu = A.s, Nv = B.s with s = (sx, sy, 1): two FMAs each.
    const float D = fmaf(r.m_nt.x, sx, fmaf(r.m_nt.y, sy, r.m_nt.z));
    const float Nu = fmaf(r.a_px.x, sx, fmaf(r.a_px.y, sy, r.a_px.z));
    const float Nv = fmaf(r.b_py.x, sx, fmaf(r.b_py.y, sy, r.b_py.z));
    // Branch-free: |D| <= 1e-12 gives inv = 0, hence t = 0 and no ray hit.
    // v_rcp_f32 (<= 1 ulp): the error stays in this hit (the backward's
    // running 1 / (1 - alpha) keeps the Newton step: it compounds along a
    // list).
    const float inv = fabsf(D) > 1e-12f ? __builtin_amdgcn_rcpf(D) : 0.0f;
    h.u = Nu * inv;
    h.v = Nv * inv;
    h.t = r.m_nt.w * inv;
    h.inv_d = inv;
    const float e_ray = h.t > kNearDepth ? -(h.u * h.u + h.v * h.v) * 0.5f : -INFINITY;
    const float dx = pxc - r.a_px.w, dy = pyc - r.b_py.w;
    const float e_scr = -(dx * dx + dy * dy);   // 2 sigma^2 = 1 (LOWPASS_SIGMA2 = 0.5)
    h.ray = e_ray >= e_scr;
    const float e = h.ray ? e_ray : e_scr;
    if (!h.ray) h.t = r.nrm_dep.w;
    // e < lthr cannot reach 1/255 (see record_lthr): ghat = 0, hence alpha
    // = 0 and no clamp. A select, not a branch: exp runs on every lane (e is
    // finite, the screen term bounds it below), and no lane's state forks.
    const float g = __expf(e);
    h.ghat = e < lthr ? 0.0f : g;
    const float raw = r.col_op.w * h.ghat;
    h.clamped = raw > kAlphaClamp;
    float a = fminf(raw, kAlphaClamp);
    h.alpha = a < kAlphaSkip ? 0.0f : a;
    return h;
}

// ------------------------------------------------------------ runtime glue

void set_error(const std::string& msg);
gs_status fail(gs_status st, const std::string& msg);
gs_status hip_fail(hipError_t e, const char* what);

#define GS_HIP(call)                                          \
    do {                                                      \
        hipError_t e_ = (call);                               \
        if (e_ != hipSuccess) return ::gs::hip_fail(e_, #call); \
    } while (0)

inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// ------------------------------------------------------------- launchers
// Each returns hipSuccess or the launch error. Shapes are validated by the
// callers in runtime.cpp before any launch.

// project.hip
// presort_keys / presort_vals (optional) receive a copy of the depth key and
// the item index for the depth presort; item_count (optional) receives
// n * n_views; rects / depth_keys / counts may be null; write_normal false
// leaves Record::nrm_dep unwritten (colour-only consumers).
hipError_t launch_project(hipStream_t st, const CamParams* cams_dev, uint32_t n_views, uint32_t n,
                          const float4* pos, const float2* scales, const float4* quats,
                          const float* opac, const float* sh, Record* recs, uint4* rects,
                          uint32_t* depth_keys, uint32_t* counts, uint32_t* ranges_to_clear,
                          uint32_t num_ranges, uint32_t* presort_keys = nullptr,
                          uint32_t* presort_vals = nullptr, uint32_t* item_count = nullptr,
                          bool clip_to_box = false, uint4* emit_info = nullptr,
                          bool pack_counts = false, bool write_normal = true);

// scan.hip
// Items per scan partition; the scratch holds a 256-B header and one u32 sum
// per partition plus one.
constexpr uint32_t kScanPartItems = 4096;
constexpr size_t scan_scratch_size(uint32_t capacity) {
    return 256 + ((size_t)(capacity + kScanPartItems - 1) / kScanPartItems + 1) * sizeof(uint32_t);
}
size_t scan_scratch_bytes(uint32_t capacity);
// Exclusive scan of in[0..n) (in place allowed), of in[perm[i]] (gather),
// or, with in == nullptr, of perm[i] & kPackMask (packed tile counts).
hipError_t launch_exclusive_scan(hipStream_t st, const uint32_t* in, uint32_t* out, uint32_t n,
                                 uint32_t* total_out, void* scratch, const uint32_t* perm = nullptr);

// binning.hip
hipError_t launch_emit_keys(hipStream_t st, uint32_t n_items, const uint4* rects,
                            const uint32_t* depth_keys, const uint32_t* offsets,
                            const CamParams* cams_dev, uint32_t items_per_view,
                            uint32_t tiles_x_fixed, uint64_t* keys, uint32_t* payloads,
                            uint32_t max_entries, const uint32_t* total, uint32_t* sort_count,
                            uint32_t* overflow_flag);
hipError_t launch_tile_ranges(hipStream_t st, const uint64_t* keys, const uint32_t* count,
                              uint32_t capacity, uint32_t* ranges);
hipError_t launch_fill_u32(hipStream_t st, uint32_t* dst, uint32_t value, size_t count);
// dst[0..count) = src[0..count) as a kernel on the stream (device to device).
hipError_t launch_copy_u32(hipStream_t st, uint32_t* dst, const uint32_t* src, size_t count);
hipError_t launch_count_tiles(hipStream_t st, uint32_t n, const uint4* rects, uint32_t* counts);
// Depth-ordered emission: order[k] is the k-th item in depth order (shifted
// left by kPackBits over its tile count when packed), emit_info the per-item record
// written by launch_project.
hipError_t launch_emit_tile_keys(hipStream_t st, uint32_t n_items, const uint32_t* order, bool packed,
                                 const uint4* emit_info, const uint32_t* offsets_sorted,
                                 const CamParams* cams_dev, uint32_t items_per_view, uint32_t* tkeys,
                                 uint32_t* payloads, uint32_t max_entries, const uint32_t* total,
                                 uint32_t* sort_count, uint32_t* overflow_flag);
// Longest-list-first order of the tiles [tile_begin, tile_begin + total_tiles):
// keys[t] = 0xFFFF - min(list length, 0xFFFF), vals[t] = tile_begin + t,
// *count = total_tiles and the 2 kTileQueues cursors = 0 (sort the pairs by
// 16 key bits, stably, to get the order).
hipError_t launch_tile_order_keys(hipStream_t st, const uint32_t* ranges, uint32_t tile_begin,
                                  uint32_t total_tiles, uint32_t* keys, uint32_t* vals, uint32_t* count,
                                  uint32_t* cursors);
hipError_t launch_tile_ranges_u32(hipStream_t st, const uint32_t* keys, const uint32_t* count,
                                  uint32_t capacity, uint32_t* ranges);
hipError_t launch_compose_keys(hipStream_t st, const uint32_t* tkeys, const uint32_t* payloads,
                               const uint32_t* depth_keys, const uint32_t* count, uint32_t capacity,
                               uint64_t* keys);

// sort.hip
size_t sort_scratch_bytes(uint32_t capacity);
hipError_t launch_radix_sort_u64(hipStream_t st, uint64_t* keys0, uint32_t* vals0, uint64_t* keys1,
                                 uint32_t* vals1, const uint32_t* count, uint32_t capacity,
                                 uint32_t begin_bit, uint32_t end_bit, uint32_t digit_bits,
                                 void* scratch, int* result_in_alt);
hipError_t launch_copy_pairs(hipStream_t st, const void* src_keys, const uint32_t* src_vals,
                             void* dst_keys, uint32_t* dst_vals, uint32_t key_bytes,
                             const uint32_t* count, uint32_t capacity);
hipError_t launch_radix_sort_u32(hipStream_t st, uint32_t* keys0, uint32_t* vals0, uint32_t* keys1,
                                 uint32_t* vals1, const uint32_t* count, uint32_t capacity,
                                 uint32_t begin_bit, uint32_t end_bit, uint32_t digit_bits,
                                 void* scratch, int* result_in_alt);

// raster_fwd.hip
struct FwdOut {
    float4* color;     // rgb, sum w
    float4* aux;       // T_final, sum w t, 0, 0
    float4* normal;    // sum w n
    uint32_t* ncontrib;
};
// Decision trace of the forward (a verification aid): for each pixel of the
// window [x0, x0 + w) x [y0, y0 + h) of view `view`, the contributing
// records in compositing order, lists[p * cap + k] = record index (bits
// 0..29) | ray branch (bit 30) | alpha clamped (bit 31), and counts[p] = the
// number of contributions (entries past cap are counted, not stored).
// lists == nullptr: no trace.
struct FwdTrace {
    uint32_t* lists = nullptr;
    uint32_t* counts = nullptr;
    uint32_t x0 = 0, y0 = 0, w = 0, h = 0, cap = 0, view = 0;
};
constexpr uint32_t kTraceRecordMask = (1u << 30) - 1u;
// Tile scheduling for the raster passes: with a queue (cursor != nullptr),
// grid_blocks resident workgroups pull tile positions from kTileQueues
// cursors (which must be 0 at launch): workgroup b uses cursor
// x = b % q, q = min(kTileQueues, grid), and claim j of it is position
// k = j q + x; the workgroup processes tile order[k] (longest list first)
// until k >= total_tiles. Several cursors spread the atomics over several
// addresses. Without a queue, workgroup b processes tile b.
constexpr uint32_t kTileQueues = 8;
struct TileQueue {
    const uint32_t* order;
    uint32_t* cursor;  // kTileQueues counters
    uint32_t grid_blocks;
};
// Work-tile size (px; must equal every view's CamParams::tile_size) and the
// records staged through LDS per batch by the forward and backward kernels.
// Instantiated: forward (tile, batch) in {(8, 64), (8, 128), (16, 128),
// (16, 256), (32, 256), (32, 512)}, backward in {(8, 64), (8, 128),
// (16, 64), (16, 128), (32, 128), (32, 256)}; the depth / normal (aux)
// variants on 16-px tiles only (forward batch 256; the aux backward stages
// 128 records whatever bwd_batch says). Others return hipErrorInvalidValue.
struct RasterShape {
    uint32_t tile = 16;
    uint32_t fwd_batch = 256;
    uint32_t bwd_batch = 64;
};
// The pass covers total_tiles tiles: queue.order[0 .. total_tiles) with a
// queue, else the global tiles [tile_begin, tile_begin + total_tiles).
hipError_t launch_raster_fwd(hipStream_t st, const CamParams* cams_dev, const CamParams* cams_host,
                             uint32_t n_views, uint32_t total_tiles, const uint32_t* ranges,
                             const uint32_t* payloads, const Record* recs, FwdOut out,
                             bool cull = true, TileQueue queue = {nullptr, nullptr, 0},
                             bool aux_outputs = true, RasterShape shape = {}, uint32_t tile_begin = 0,
                             FwdTrace trace = {});

// raster_bwd.hip
struct BwdIn {
    const float4* dl_dcolor;   // rgb, w = depth-loss gradient
    const float4* dl_dnormal;  // may be null
    const float4* out_color;   // forward sum w in .w
    const float4* out_aux;     // T_final, sum w t
    const uint32_t* ncontrib;
};
hipError_t launch_raster_bwd(hipStream_t st, const CamParams* cams_dev, const CamParams* cams_host,
                             uint32_t n_views, uint32_t total_tiles, const uint32_t* ranges,
                             const uint32_t* payloads, const Record* recs, BwdIn in,
                             float* grad_coef, bool aux, bool cull = true,
                             TileQueue queue = {nullptr, nullptr, 0}, RasterShape shape = {},
                             uint32_t tile_begin = 0);

// geom_bwd.hip
struct ParamGrads {
    float4* pos;
    float2* scales;
    float4* quat;
    float* opacity;
    float* sh;
    float* cam;   // 16 floats per view, or null
};
size_t geom_cam_partials_bytes(uint32_t n, uint32_t n_views);
hipError_t launch_geom_bwd(hipStream_t st, const CamParams* cams_dev, uint32_t n_views, uint32_t n,
                           const float4* pos, const float2* scales, const float4* quats,
                           const float* sh, const float* grad_coef, bool aux, ParamGrads g, bool accumulate,
                           float* cam_partials);

// loss.hip
// With `view` set, the target is read at target + view->target_base (the
// per-view descriptor lives in device memory, so a captured graph replays
// with whatever views the descriptor table holds). partial_sums receives
// l1_ssim_partials(w, h) pairs (sum |x - y|, sum SSIM); dl_out may be null
// (loss only). tstats, when set, holds the target's window mean and second
// moment per channel (launch_target_stats), at (target pixel) * 3 + channel
// from the same base offset as the target.
hipError_t launch_l1_ssim(hipStream_t st, uint32_t w, uint32_t h, float lambda, float dssim_scale,
                          const float4* img, const float4* target, float4* dl_out, double* partial_sums,
                          const CamParams* view = nullptr, const float2* tstats = nullptr);
uint32_t l1_ssim_partials(uint32_t w, uint32_t h);
// The 11 x 11 window mean and second moment of each channel of a w x h target
// (zero padding), as float2 (mean, E[y^2]) at pixel * 3 + channel: the target
// half of the SSIM statistics, for targets that stay fixed over many steps.
hipError_t launch_target_stats(hipStream_t st, uint32_t w, uint32_t h, const float4* target, float2* out);
// The losses of view slots [first, first + k) of a step in one launch: view
// j's image and dL/dimage start at img / dl_out + pix_base[j], its target at
// target + cams_dev[first + j].target_base, its partial sums at
// partials + 2 (first + j) partials_stride. At most 64 views.
hipError_t launch_l1_ssim_views(hipStream_t st, uint32_t first, uint32_t k, const uint32_t* widths,
                                const uint32_t* heights, const uint32_t* pix_base, float lambda,
                                float dssim_scale, const float4* img, const float4* target, float4* dl_out,
                                double* partials, uint32_t partials_stride, const CamParams* cams_dev,
                                const float2* tstats = nullptr);

// mcmc.hip: exploration noise on raw positions (mcmc.wgsl) and the relocation
// of dead surfels (relocate.wgsl). With state set, the noise takes its seed
// ((t - 1) * 0x9e3779b9) and sigma (noise_scale * lr[lr_class]) from the
// device-resident Adam state of the step just taken; else seed and sigma.
struct AdamState;
struct NoiseArgs {
    uint32_t n = 0, seed = 0;
    float sigma = 0.0f, opacity_k = 5.0f;
    const AdamState* state = nullptr;
    float noise_scale = 0.0f;
    uint32_t lr_class = 0;
    float4* pos_raw = nullptr;
    const float2* scales_act = nullptr;
    const float* opac_act = nullptr;
};
hipError_t launch_add_noise(hipStream_t st, const NoiseArgs& a);
// pairs[i] = (dead, alive, mode, 0): mode 0 relocation, 1 split. Class c of
// the five (pos, scales, quat, opacity, SH) holds class_width[c] floats per
// surfel (4, 2, 4, 1, 48) in m[c] / v[c].
struct RelocateArgs {
    const uint4* pairs = nullptr;
    uint32_t n_pairs = 0;
    const float* opac_act = nullptr;
    float4* pos_raw = nullptr;
    float2* scales_raw = nullptr;
    float4* quat_raw = nullptr;
    float* opac_raw = nullptr;
    float4* sh_raw = nullptr;
    float* m[5] = {};
    float* v[5] = {};
    uint32_t class_width[5] = {4, 2, 4, 1, 48};
};
hipError_t launch_relocate(hipStream_t st, const RelocateArgs& a);

// track.hip: photometric tracking (crates/gs-track). Device state of one
// track call: Adam moments of the six pose parameters, the residual
// bookkeeping and the current / best pose (quat xyzw, centre).
struct TrackState {
    float m[6], v[6];
    float best, coverage;
    uint32_t stale, stopped, iters, pad;
    float cur_q[4], cur_c[3];
    float best_q[4], best_c[3];
};
struct TrackArgs {
    TrackState* state;
    const float2* partials;
    uint32_t n_partials, npix;
    const float* grad_cam;  // [dL/dR col-major 9, dL/dC 3, dL/df, pad]
    CamParams* cam;         // the camera the next forward reads
    float* residuals;       // one per iteration
    uint32_t iterations, patience;
    float lr_rot, lr_center, min_improvement;
};
uint32_t photometric_partials(uint32_t npix);
hipError_t launch_photometric(hipStream_t st, uint32_t npix, float gain, float bias, const float4* render,
                              const float4* frame, float4* dl_dcolor, float2* partials);
hipError_t launch_track_update(hipStream_t st, const TrackArgs& a, uint32_t t);

// ply.hip: rows of the baked 3DGS .ply export (writer.rs:17-79), 59 floats
// per surfel, for surfels [0, n) of the activated SoA.
constexpr int kPlyRowFloats = 59;
hipError_t launch_ply_rows(hipStream_t st, uint32_t n, const float4* pos, const float2* scales,
                           const float4* quats, const float* opac, const float* sh, float* rows);

// normal_loss.hip: normal-consistency loss over one W x H view (the
// reference's NormalLoss, normal_loss.wgsl:52-145). lambda carries the 1/N.
struct NormalLossArgs {
    uint32_t width, height;
    float focal, lambda, alpha_min;
    const float4* color;   // .w = alpha
    const float4* aux;     // .y = depth_acc
    const float4* normal;  // sum w n
    float4* dl_dnormal;
    float* dl_dcolor_w;    // &dl_dcolor[0].w, stride 4 floats
    float4* loss_map;      // may be null
    float* partials;       // one per 16x16 tile
};
hipError_t launch_normal_loss(hipStream_t st, const NormalLossArgs& a);
uint32_t normal_loss_partials(uint32_t width, uint32_t height);

// adam.hip
struct AdamClass {
    float* raw;
    const float* grad;
    float* m;
    float* v;
    float* activated;
    uint32_t n;
    int activation;   // 0 identity, 1 exp, 2 sigmoid
    uint32_t block_begin;  // first block of this class in the fused grid
};
constexpr int kMaxAdamClasses = 8;
// Learning-rate schedule of one class: lr = base while iters == 0, else
// lr = base * final_factor^(iter / iters) with iter the 0-based index of the
// step being taken (the exponential decay of trainer.rs:1149-1151 with
// schedule_t = iter / iters, trainer.rs:407-416).
struct AdamSchedule {
    float base;
    float final_factor;
    uint32_t iters;
    uint32_t pad;
};
// Host-written hyperparameters (uploaded on the step's stream when changed).
struct AdamHyper {
    AdamSchedule sched[kMaxAdamClasses];
    float reg[kMaxAdamClasses];
};
// Device-resident optimizer state: the step counter, the running f32
// products beta1^t, beta2^t and their bias corrections, and this step's
// learning rates -- all advanced by adam_tick on the device, so a captured
// graph replays the bias corrections and the learning-rate schedule.
struct AdamState {
    uint32_t t;
    float p1, p2;          // beta1^t, beta2^t (running f32 products)
    float bc1_inv, bc2_inv;
    float lr[kMaxAdamClasses];  // learning rate of the current step, per class
    AdamHyper hyper;
};
struct AdamTable {
    AdamClass c[kMaxAdamClasses];
    uint32_t n_classes;
};

// Activation of a raw parameter (0 identity, 1 exp, 2 sigmoid) and its
// derivative, as adam.wgsl's activate / the chain factor of adam_step.
__device__ __forceinline__ float adam_activate(int act, float x) {
    if (act == 1) return __expf(x);
    if (act == 2) return 1.0f / (1.0f + __expf(-x));
    return x;
}
__device__ __forceinline__ float adam_dact(int act, float x) {
    if (act == 1) return __expf(x);
    if (act == 2) {
        const float s = 1.0f / (1.0f + __expf(-x));
        return s * (1.0f - s);
    }
    return 1.0f;
}
// One Adam element update (adam.wgsl:44-74 with the host bias corrections of
// optim.rs:179-197); shared by the class-table kernel and the fused geometry
// backward so both give the same bits.
__device__ __forceinline__ void adam_element(int act, float lr, float reg, float bc1, float bc2, float& raw,
                                             float grad, float& m, float& v) {
    const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    const float g = (grad + reg) * adam_dact(act, raw);
    m = beta1 * m + (1.0f - beta1) * g;
    v = beta2 * v + (1.0f - beta2) * g * g;
    const float mhat = m * bc1;
    const float vhat = v * bc2;
    raw -= lr * mhat / (sqrtf(vhat) + eps);
}

uint32_t adam_blocks_for(uint32_t n);
hipError_t launch_adam_tick(hipStream_t st, AdamState* state);
// The class table (buffers and sizes) travels by value in the kernel
// arguments; learning rates and regularisers are read from the state.
hipError_t launch_adam(hipStream_t st, const AdamTable& table, uint32_t total_blocks,
                       const AdamState* state, bool step);
// Geometry backward of a batch fused with the Adam step (after adam_tick) of
// the trainer's five classes (pos float4, scales float2 exp, quat float4,
// opacity sigmoid, SH 48): reads raw parameters, writes raw, m, v and the
// activated copies; keep.pos..sh (all or none) also receive the summed
// parameter gradients. grad_coef holds grad_stride(false) slots per (view,
// surfel) (colour-only backward). hipErrorInvalidValue for any other class
// table.
hipError_t launch_geom_adam(hipStream_t st, const CamParams* cams_dev, uint32_t n_views, uint32_t n,
                            const float* grad_coef, const AdamTable& table, const AdamState* state,
                            ParamGrads keep);

}  // namespace gs
