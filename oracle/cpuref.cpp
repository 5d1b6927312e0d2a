// CPU oracle for the 2DGS surfel training step -- TEST INFRASTRUCTURE ONLY.
// The product never links or calls this file; see cpuref.h for the contract
// and the list of parity pins. Reference paths are relative to the
// jimmejardine/splattar checkout.
//
// Build: g++ -O2 -ffp-contract=off -fopenmp -shared -fPIC (oracle/Makefile).
// -ffp-contract=off matters for the fp32 mirror (bit-exact keys/rects) and
// keeps the FP64 path free of fused multiply-adds the reference never used.

#include "cpuref.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ------------------------------------------------------------------ vectors

struct V3 {
    double x = 0, y = 0, z = 0;
    V3() = default;
    V3(double a, double b, double c) : x(a), y(b), z(c) {}
    double& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(double s, V3 a) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator/(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline V3& operator+=(V3& a, V3 b) { a = a + b; return a; }
inline V3& operator-=(V3& a, V3 b) { a = a - b; return a; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double length(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 normalize(V3 a) { return a / length(a); }
inline bool is_zero(V3 a) { return a.x == 0.0 && a.y == 0.0 && a.z == 0.0; }

// Column-major 3x3: col[k] is the image of e_k.
struct M3 {
    V3 col[3];
    V3 mul(V3 v) const { return col[0] * v.x + col[1] * v.y + col[2] * v.z; }
    V3 tmul(V3 v) const { return {dot(col[0], v), dot(col[1], v), dot(col[2], v)}; }
};

// ------------------------------------------------- constants (scene.rs:17-21)

constexpr double LOWPASS_SIGMA2 = 0.5;
constexpr double ALPHA_SKIP = 1.0 / 255.0;
constexpr double ALPHA_CLAMP = 0.995;
constexpr double T_TERMINATE = 1e-4;
constexpr double NEAR_DEPTH = 0.05;
constexpr int TILE = 16;  // forward.rs:105 -- part of the forward model

// ---------------------------------------------------- math.rs (FP64 oracle)

// math.rs:14-34
M3 quat_to_mat(const double* q) {
    double n = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    double x = q[0] / n, y = q[1] / n, z = q[2] / n, w = q[3] / n;
    M3 r;
    r.col[0] = {1.0 - 2.0 * (y * y + z * z), 2.0 * (x * y + w * z), 2.0 * (x * z - w * y)};
    r.col[1] = {2.0 * (x * y - w * z), 1.0 - 2.0 * (x * x + z * z), 2.0 * (y * z + w * x)};
    r.col[2] = {2.0 * (x * z + w * y), 2.0 * (y * z - w * x), 1.0 - 2.0 * (x * x + y * y)};
    return r;
}

// math.rs:38-88: dL/dq for the unnormalized quaternion, through the
// normalization Jacobian.
void quat_grad(const double* q, const M3& dl_dr, double* out) {
    double n = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    double qh[4] = {q[0] / n, q[1] / n, q[2] / n, q[3] / n};
    double x = qh[0], y = qh[1], z = qh[2], w = qh[3];
    M3 dr[4];
    dr[0].col[0] = {0.0, 2.0 * y, 2.0 * z};
    dr[0].col[1] = {2.0 * y, -4.0 * x, 2.0 * w};
    dr[0].col[2] = {2.0 * z, -2.0 * w, -4.0 * x};
    dr[1].col[0] = {-4.0 * y, 2.0 * x, -2.0 * w};
    dr[1].col[1] = {2.0 * x, 0.0, 2.0 * z};
    dr[1].col[2] = {2.0 * w, 2.0 * z, -4.0 * y};
    dr[2].col[0] = {-4.0 * z, 2.0 * w, 2.0 * x};
    dr[2].col[1] = {-2.0 * w, -4.0 * z, 2.0 * y};
    dr[2].col[2] = {2.0 * x, 2.0 * y, 0.0};
    dr[3].col[0] = {0.0, 2.0 * z, -2.0 * y};
    dr[3].col[1] = {-2.0 * z, 0.0, 2.0 * x};
    dr[3].col[2] = {2.0 * y, -2.0 * x, 0.0};
    double dqh[4] = {0, 0, 0, 0};
    for (int k = 0; k < 4; ++k)
        for (int c = 0; c < 3; ++c) dqh[k] += dot(dl_dr.col[c], dr[k].col[c]);
    double d = 0.0;
    for (int i = 0; i < 4; ++i) d += dqh[i] * qh[i];
    for (int i = 0; i < 4; ++i) out[i] = (dqh[i] - qh[i] * d) / n;
}

// math.rs:93-99
inline double triple(V3 a, V3 b, V3 c) { return dot(a, cross(b, c)); }
inline void triple_grads(V3 a, V3 b, V3 c, V3& ga, V3& gb, V3& gc) {
    ga = cross(b, c);
    gb = cross(c, a);
    gc = cross(a, b);
}

// math.rs:103-120
constexpr double SH_C0 = 0.28209479177387814;
constexpr double SH_C1 = 0.4886025119029199;
constexpr double SH_C2[5] = {1.0925484305920792, -1.0925484305920792, 0.31539156525252005,
                             -1.0925484305920792, 0.5462742152960396};
constexpr double SH_C3[7] = {-0.5900435899266435, 2.890611442640554, -0.4570457994644658,
                             0.3731763325901154,  -0.4570457994644658, 1.445305721320277,
                             -0.5900435899266435};

inline int num_coeffs(int degree) { return (degree + 1) * (degree + 1); }

// math.rs:127-156
int sh_basis(int degree, V3 d, double* b) {
    double x = d.x, y = d.y, z = d.z;
    int k = 0;
    b[k++] = SH_C0;
    if (degree >= 1) {
        b[k++] = -SH_C1 * y;
        b[k++] = SH_C1 * z;
        b[k++] = -SH_C1 * x;
    }
    if (degree >= 2) {
        double xx = x * x, yy = y * y, zz = z * z;
        b[k++] = SH_C2[0] * x * y;
        b[k++] = SH_C2[1] * y * z;
        b[k++] = SH_C2[2] * (2.0 * zz - xx - yy);
        b[k++] = SH_C2[3] * x * z;
        b[k++] = SH_C2[4] * (xx - yy);
    }
    if (degree >= 3) {
        double xx = x * x, yy = y * y, zz = z * z;
        b[k++] = SH_C3[0] * y * (3.0 * xx - yy);
        b[k++] = SH_C3[1] * x * y * z;
        b[k++] = SH_C3[2] * y * (4.0 * zz - xx - yy);
        b[k++] = SH_C3[3] * z * (2.0 * zz - 3.0 * xx - 3.0 * yy);
        b[k++] = SH_C3[4] * x * (4.0 * zz - xx - yy);
        b[k++] = SH_C3[5] * z * (xx - yy);
        b[k++] = SH_C3[6] * x * (xx - 3.0 * yy);
    }
    return k;
}

// math.rs:159-191
int sh_basis_grad(int degree, V3 d, V3* g) {
    double x = d.x, y = d.y, z = d.z;
    int k = 0;
    g[k++] = V3(0, 0, 0);
    if (degree >= 1) {
        g[k++] = V3(0.0, -SH_C1, 0.0);
        g[k++] = V3(0.0, 0.0, SH_C1);
        g[k++] = V3(-SH_C1, 0.0, 0.0);
    }
    if (degree >= 2) {
        g[k++] = SH_C2[0] * V3(y, x, 0.0);
        g[k++] = SH_C2[1] * V3(0.0, z, y);
        g[k++] = SH_C2[2] * V3(-2.0 * x, -2.0 * y, 4.0 * z);
        g[k++] = SH_C2[3] * V3(z, 0.0, x);
        g[k++] = SH_C2[4] * V3(2.0 * x, -2.0 * y, 0.0);
    }
    if (degree >= 3) {
        double xx = x * x, yy = y * y, zz = z * z;
        g[k++] = SH_C3[0] * V3(6.0 * x * y, 3.0 * xx - 3.0 * yy, 0.0);
        g[k++] = SH_C3[1] * V3(y * z, x * z, x * y);
        g[k++] = SH_C3[2] * V3(-2.0 * x * y, 4.0 * zz - xx - 3.0 * yy, 8.0 * y * z);
        g[k++] = SH_C3[3] * V3(-6.0 * x * z, -6.0 * y * z, 6.0 * zz - 3.0 * xx - 3.0 * yy);
        g[k++] = SH_C3[4] * V3(4.0 * zz - 3.0 * xx - yy, -2.0 * x * y, 8.0 * x * z);
        g[k++] = SH_C3[5] * V3(2.0 * x * z, -2.0 * y * z, xx - yy);
        g[k++] = SH_C3[6] * V3(3.0 * xx - 3.0 * yy, -6.0 * x * y, 0.0);
    }
    return k;
}

// ------------------------------------------------------- forward.rs (FP64)

// forward.rs:13-29
struct SurfelCam {
    int64_t idx;
    V3 c, tu, tv, color;
    bool clamped[3];
    double opacity;
    V3 normal;
    double depth, px, py;
    // Tile rect of covers_pixel (forward.rs:104-123), precomputed once.
    bool on_screen;
    int x0, y0, x1, y1;
};

struct Scene {
    const oref_scene* s;
    int ncoef;
    V3 pos(int64_t i) const { return {s->pos[3 * i], s->pos[3 * i + 1], s->pos[3 * i + 2]}; }
    V3 sh(int64_t i, int k) const {
        const double* p = s->sh + (i * ncoef + k) * 3;
        return {p[0], p[1], p[2]};
    }
};

// forward.rs:32-40
inline V3 pixel_ray(const oref_camera* cam, int x, int y) {
    double cx = cam->width * 0.5, cy = cam->height * 0.5;
    return {(x + 0.5 - cx) / cam->focal, -(y + 0.5 - cy) / cam->focal, -1.0};
}

inline V3 cam_center(const oref_camera* cam) {
    return {cam->center[0], cam->center[1], cam->center[2]};
}

// forward.rs:104-123, evaluated once per surfel (the rect does not depend on
// the pixel; membership of the pixel's tile is the per-pixel test).
void compute_rect(SurfelCam& sc, const oref_camera* cam) {
    double r_px = std::min(3.0 * cam->focal * (length(sc.tu) + length(sc.tv)) / sc.depth + 3.0, 64.0);
    sc.on_screen = sc.px + r_px >= 0.0 && sc.px - r_px < (double)cam->width &&
                   sc.py + r_px >= 0.0 && sc.py - r_px < (double)cam->height;
    double max_tx = (double)((cam->width + TILE - 1) / TILE - 1);
    double max_ty = (double)((cam->height + TILE - 1) / TILE - 1);
    auto cl = [](double v, double lo, double hi) {
        // Rust f64::clamp then `as usize` (truncation toward zero).
        double c = v < lo ? lo : (v > hi ? hi : v);
        return (int)c;
    };
    sc.x0 = cl((sc.px - r_px) / TILE, 0.0, max_tx);
    sc.y0 = cl((sc.py - r_px) / TILE, 0.0, max_ty);
    sc.x1 = cl((sc.px + r_px) / TILE, 0.0, max_tx);
    sc.y1 = cl((sc.py + r_px) / TILE, 0.0, max_ty);
}

inline bool covers_pixel(const SurfelCam& sc, int x, int y) {
    if (!sc.on_screen) return false;
    int tx = x / TILE, ty = y / TILE;
    return tx >= sc.x0 && tx <= sc.x1 && ty >= sc.y0 && ty <= sc.y1;
}

// forward.rs:50-92 for surfel idx (no near cull; `prepare` applies it).
SurfelCam make_surfel_cam(const Scene& sc, const oref_camera* cam, const M3& r_cam, int64_t idx) {
    double cx = cam->width * 0.5, cy = cam->height * 0.5;
    V3 C = cam_center(cam);
    double basis[16];
    V3 p = sc.pos(idx);
    V3 x = p - C;
    V3 c = r_cam.tmul(x);
    double depth = -c.z;
    M3 rs = quat_to_mat(sc.s->quat + 4 * idx);
    V3 tu = r_cam.tmul(rs.col[0] * sc.s->scales[2 * idx]);
    V3 tv = r_cam.tmul(rs.col[1] * sc.s->scales[2 * idx + 1]);
    V3 dir = normalize(p - C);
    int nb = sh_basis(sc.s->sh_degree, dir, basis);
    V3 raw(0.5, 0.5, 0.5);
    for (int k = 0; k < nb; ++k) raw += basis[k] * sc.sh(idx, k);
    SurfelCam s;
    s.idx = idx;
    s.c = c;
    s.tu = tu;
    s.tv = tv;
    s.clamped[0] = raw.x < 0.0;
    s.clamped[1] = raw.y < 0.0;
    s.clamped[2] = raw.z < 0.0;
    s.color = V3(std::max(raw.x, 0.0), std::max(raw.y, 0.0), std::max(raw.z, 0.0));
    V3 normal = cross(tu, tv);
    if (dot(normal, normal) > 0.0) {
        normal = normalize(normal);
        if (dot(normal, c) > 0.0) normal = -normal;
    }
    s.normal = normal;
    s.opacity = sc.s->opacity[idx];
    s.depth = depth;
    double w = -c.z;
    s.px = cx + cam->focal * c.x / w;
    s.py = cy - cam->focal * c.y / w;
    compute_rect(s, cam);
    return s;
}

// forward.rs:44-97
std::vector<SurfelCam> prepare(const Scene& sc, const oref_camera* cam) {
    M3 r_cam = quat_to_mat(cam->quat);
    V3 C = cam_center(cam);
    std::vector<SurfelCam> out;
    out.reserve(sc.s->n);
    for (int64_t idx = 0; idx < sc.s->n; ++idx) {
        V3 c = r_cam.tmul(sc.pos(idx) - C);
        if (-c.z < NEAR_DEPTH) continue;  // forward.rs:56-58
        out.push_back(make_surfel_cam(sc, cam, r_cam, idx));
    }
    // forward.rs:95: ascending f32 depth-key bits, stable by original index.
    auto key = [](const SurfelCam& s) {
        float f = (float)s.depth;
        uint32_t b;
        std::memcpy(&b, &f, 4);
        return b;
    };
    std::stable_sort(out.begin(), out.end(), [&](const SurfelCam& a, const SurfelCam& b) {
        uint32_t ka = key(a), kb = key(b);
        if (ka != kb) return ka < kb;
        return a.idx < b.idx;
    });
    return out;
}

// forward.rs:126-134
struct Hit {
    double ghat;
    bool ray_branch;
    double u, v, t;
};

// forward.rs:136-169
inline Hit evaluate(const SurfelCam& sc, V3 d, double pix_x, double pix_y) {
    V3 s = -d;
    double det = triple(sc.tu, sc.tv, s);
    double g_ray = 0.0, u = 0.0, v = 0.0, t = 0.0;
    if (std::fabs(det) > 1e-12) {
        u = triple(-sc.c, sc.tv, s) / det;
        v = triple(sc.tu, -sc.c, s) / det;
        t = triple(sc.tu, sc.tv, -sc.c) / det;
        if (t > NEAR_DEPTH) g_ray = std::exp(-(u * u + v * v) * 0.5);
    }
    double dx = pix_x - sc.px, dy = pix_y - sc.py;
    double g_scr = std::exp(-(dx * dx + dy * dy) / (2.0 * LOWPASS_SIGMA2));
    if (g_ray >= g_scr) return {g_ray, true, u, v, t};
    return {g_scr, false, u, v, sc.depth};
}

// Per-tile surfel lists in prepare() order: the binned walk visits exactly
// the surfels covers_pixel() admits, in the same order as the literal walk.
struct TileLists {
    int tiles_x = 0, tiles_y = 0;
    std::vector<int64_t> start;  // tiles+1
    std::vector<int32_t> items;
};

TileLists build_tile_lists(const std::vector<SurfelCam>& surfels, const oref_camera* cam) {
    TileLists tl;
    tl.tiles_x = (cam->width + TILE - 1) / TILE;
    tl.tiles_y = (cam->height + TILE - 1) / TILE;
    int64_t nt = (int64_t)tl.tiles_x * tl.tiles_y;
    std::vector<int64_t> count(nt + 1, 0);
    for (const auto& s : surfels) {
        if (!s.on_screen) continue;
        for (int ty = s.y0; ty <= s.y1; ++ty)
            for (int tx = s.x0; tx <= s.x1; ++tx) count[(int64_t)ty * tl.tiles_x + tx]++;
    }
    tl.start.assign(nt + 1, 0);
    for (int64_t t = 0; t < nt; ++t) tl.start[t + 1] = tl.start[t] + count[t];
    tl.items.resize(tl.start[nt]);
    std::vector<int64_t> fill(tl.start.begin(), tl.start.end() - 1);
    for (size_t i = 0; i < surfels.size(); ++i) {
        const auto& s = surfels[i];
        if (!s.on_screen) continue;
        for (int ty = s.y0; ty <= s.y1; ++ty)
            for (int tx = s.x0; tx <= s.x1; ++tx)
                tl.items[fill[(int64_t)ty * tl.tiles_x + tx]++] = (int32_t)i;
    }
    return tl;
}

// Pixel window evaluated by render / gradients (the whole image unless the
// camera carries a crop).
struct Window {
    int x0, y0, w, h;
};
inline Window window_of(const oref_camera* cam) {
    if (cam->crop_w <= 0 || cam->crop_h <= 0) return {0, 0, cam->width, cam->height};
    int x0 = std::max(0, std::min(cam->crop_x0, cam->width - 1));
    int y0 = std::max(0, std::min(cam->crop_y0, cam->height - 1));
    return {x0, y0, std::min(cam->crop_w, cam->width - x0), std::min(cam->crop_h, cam->height - y0)};
}

// Calls f(list_idx) for every surfel covers_pixel() admits at (x, y), in
// prepare() order. Literal mode tests every surfel (forward.rs:224-227).
template <class F>
inline void walk_pixel(const std::vector<SurfelCam>& surfels, const TileLists* tl, int x, int y,
                       F&& f) {
    if (tl == nullptr) {
        for (size_t i = 0; i < surfels.size(); ++i) {
            if (!covers_pixel(surfels[i], x, y)) continue;
            if (!f((int64_t)i)) return;
        }
    } else {
        int64_t t = (int64_t)(y / TILE) * tl->tiles_x + (x / TILE);
        for (int64_t k = tl->start[t]; k < tl->start[t + 1]; ++k)
            if (!f((int64_t)tl->items[k])) return;
    }
}

// ------------------------------------------------------ backward.rs (FP64)

// backward.rs:15-23
struct CamSpaceAcc {
    V3 dc, dtu, dtv, dcolor;
    double dopacity = 0.0;
    V3 dnormal;
};

// backward.rs:36-42
struct Contribution {
    int64_t list_idx;
    double alpha;
    bool clamped_alpha;
    double t_before;
    Hit hit;
};

// backward.rs:252-257: for a = R^T y, dL/dR[j][k] += y_j * dL/da_k.
inline void add_transpose_chain(M3& dl_dr, V3 y, V3 dl_da) {
    for (int k = 0; k < 3; ++k) dl_dr.col[k] += y * dl_da[k];
}

// backward.rs:262-308
void chain_ray_branch(const SurfelCam& sc, V3 d, const Hit& hit, double dl_dghat, double dl_dt,
                      CamSpaceAcc& a, double& dfocal, double focal) {
    V3 s_vec = -d;
    double det = triple(sc.tu, sc.tv, s_vec);
    double u = hit.u, v = hit.v, t = hit.t;
    double dl_du = -u * hit.ghat * dl_dghat;
    double dl_dv = -v * hit.ghat * dl_dghat;
    double dl_dnu = dl_du / det;
    double dl_dnv = dl_dv / det;
    double dl_dnt = dl_dt / det;
    double dl_ddet = -(u * dl_du + v * dl_dv + t * dl_dt) / det;
    V3 ds_vec, ga, gb, gc;
    triple_grads(-sc.c, sc.tv, s_vec, ga, gb, gc);  // Nu = det[-c, tv, S]
    a.dc -= ga * dl_dnu;
    a.dtv += gb * dl_dnu;
    ds_vec += gc * dl_dnu;
    triple_grads(sc.tu, -sc.c, s_vec, ga, gb, gc);  // Nv = det[tu, -c, S]
    a.dtu += ga * dl_dnv;
    a.dc -= gb * dl_dnv;
    ds_vec += gc * dl_dnv;
    triple_grads(sc.tu, sc.tv, -sc.c, ga, gb, gc);  // Nt = det[tu, tv, -c]
    a.dtu += ga * dl_dnt;
    a.dtv += gb * dl_dnt;
    a.dc -= gc * dl_dnt;
    triple_grads(sc.tu, sc.tv, s_vec, ga, gb, gc);  // D = det[tu, tv, S]
    a.dtu += ga * dl_ddet;
    a.dtv += gb * dl_ddet;
    ds_vec += gc * dl_ddet;
    V3 dd = -ds_vec;
    dfocal += dd.x * (-d.x / focal) + dd.y * (-d.y / focal);
}

// backward.rs:312-337
void chain_screen_branch(const SurfelCam& sc, double pix_x, double pix_y, double dl_dghat,
                         CamSpaceAcc& a, double& dfocal, const oref_camera* cam) {
    double dx = pix_x - sc.px, dy = pix_y - sc.py;
    double g = std::exp(-(dx * dx + dy * dy) / (2.0 * LOWPASS_SIGMA2));
    double dl_ddx = -(dx / LOWPASS_SIGMA2) * g * dl_dghat;
    double dl_ddy = -(dy / LOWPASS_SIGMA2) * g * dl_dghat;
    double dpx = -dl_ddx, dpy = -dl_ddy;
    double f = cam->focal;
    double w = -sc.c.z;
    a.dc.x += dpx * f / w;
    a.dc.z += dpx * f * sc.c.x / (w * w);
    a.dc.y += dpy * (-f / w);
    a.dc.z += dpy * (-f * sc.c.y / (w * w));
    dfocal += dpx * sc.c.x / w + dpy * (-sc.c.y / w);
}

struct Workspace {
    std::vector<CamSpaceAcc> acc;
    double dfocal = 0.0;
    std::vector<Contribution> contribs;
};

void reverse_walk(const std::vector<SurfelCam>& surfels, const oref_camera* cam, V3 d, double pix_x,
                  double pix_y, double w_total, double wt_total, V3 dl_dc_pix, double dl_dd_pix,
                  V3 dl_dn_pix, double lambda_dist, Workspace& ws);

// backward.rs:67-173 for one pixel.
void backward_pixel(const std::vector<SurfelCam>& surfels, const TileLists* tl,
                    const oref_camera* cam, int x, int y, V3 dl_dc_pix, double dl_dd_pix,
                    V3 dl_dn_pix, double lambda_dist, Workspace& ws) {
    bool any_image_loss = !is_zero(dl_dc_pix) || dl_dd_pix != 0.0 || !is_zero(dl_dn_pix);
    if (!any_image_loss && lambda_dist == 0.0) return;
    V3 d = pixel_ray(cam, x, y);
    double pix_x = x + 0.5, pix_y = y + 0.5;
    // Forward replay (backward.rs:81-110).
    ws.contribs.clear();
    double transmittance = 1.0, w_total = 0.0, wt_total = 0.0;
    walk_pixel(surfels, tl, x, y, [&](int64_t li) {
        const SurfelCam& sc = surfels[li];
        Hit hit = evaluate(sc, d, pix_x, pix_y);
        double raw_alpha = sc.opacity * hit.ghat;
        double alpha = std::min(raw_alpha, ALPHA_CLAMP);
        if (alpha < ALPHA_SKIP) return true;
        double w = transmittance * alpha;
        w_total += w;
        wt_total += w * hit.t;
        ws.contribs.push_back({li, alpha, raw_alpha > ALPHA_CLAMP, transmittance, hit});
        transmittance *= 1.0 - alpha;
        return !(transmittance < T_TERMINATE);
    });
    reverse_walk(surfels, cam, d, pix_x, pix_y, w_total, wt_total, dl_dc_pix, dl_dd_pix, dl_dn_pix,
                 lambda_dist, ws);
}

// backward.rs:112-172: the reverse walk over the pixel's contributions
// (ws.contribs, front to back), accumulating into ws.acc[list_idx].
void reverse_walk(const std::vector<SurfelCam>& surfels, const oref_camera* cam, V3 d, double pix_x,
                  double pix_y, double w_total, double wt_total, V3 dl_dc_pix, double dl_dd_pix,
                  V3 dl_dn_pix, double lambda_dist, Workspace& ws) {
    V3 suffix_color, suffix_normal;
    double suffix_depth = 0.0, suffix_w = 0.0, suffix_cw = 0.0;
    for (auto it = ws.contribs.rbegin(); it != ws.contribs.rend(); ++it) {
        const SurfelCam& sc = surfels[it->list_idx];
        CamSpaceAcc& a = ws.acc[it->list_idx];
        double t_i = it->t_before, alpha = it->alpha;
        double w = t_i * alpha;
        double t_hit = it->hit.t;
        double dt_dist = 0.0, coef = 0.0;
        if (lambda_dist != 0.0) {
            double a_pre = w_total - suffix_w - w;
            double b_pre = wt_total - suffix_depth - w * t_hit;
            dt_dist = lambda_dist * 2.0 * w * (a_pre - suffix_w);
            coef = lambda_dist * (2.0 * (t_hit * a_pre - b_pre) + 2.0 * (suffix_depth - t_hit * suffix_w));
        }
        a.dcolor += dl_dc_pix * w;
        a.dnormal += dl_dn_pix * w;
        double dl_dt = dl_dd_pix * w + dt_dist;
        double dl_dalpha = dot(dl_dc_pix, sc.color) * t_i - dot(dl_dc_pix, suffix_color) / (1.0 - alpha) +
                           dl_dd_pix * (t_hit * t_i - suffix_depth / (1.0 - alpha)) +
                           dot(dl_dn_pix, sc.normal) * t_i - dot(dl_dn_pix, suffix_normal) / (1.0 - alpha) +
                           coef * t_i - suffix_cw / (1.0 - alpha);
        suffix_color += sc.color * w;
        suffix_depth += t_hit * w;
        suffix_normal += sc.normal * w;
        suffix_w += w;
        suffix_cw += coef * w;
        if (it->clamped_alpha) continue;
        a.dopacity += dl_dalpha * it->hit.ghat;
        double dl_dghat = dl_dalpha * sc.opacity;
        if (it->hit.ray_branch) {
            chain_ray_branch(sc, d, it->hit, dl_dghat, dl_dt, a, ws.dfocal, cam->focal);
        } else {
            a.dc.z -= dl_dt;  // t fell back to the center depth (-c.z)
            chain_screen_branch(sc, pix_x, pix_y, dl_dghat, a, ws.dfocal, cam);
        }
    }
}

void add_acc(CamSpaceAcc& a, const CamSpaceAcc& b) {
    a.dc += b.dc;
    a.dtu += b.dtu;
    a.dtv += b.dtv;
    a.dcolor += b.dcolor;
    a.dopacity += b.dopacity;
    a.dnormal += b.dnormal;
}

int threads_available() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// ----------------------------------------------------- fp32 mirror helpers
// Operation order pinned to the HIP project kernel (splattar_amd/csrc/
// project.hip): every product and sum is a separately rounded f32 op.
inline float dot3f(const float* a, float x, float y, float z) {
    float p0 = a[0] * x;
    float p1 = a[1] * y;
    float p2 = a[2] * z;
    return (p0 + p1) + p2;
}

}  // namespace

// =================================================================== C API

extern "C" {

void oref_quat_to_mat(const double q[4], double r[9]) {
    M3 m = quat_to_mat(q);
    for (int k = 0; k < 3; ++k)
        for (int j = 0; j < 3; ++j) r[3 * k + j] = m.col[k][j];
}

void oref_quat_grad(const double q[4], const double dl_dr[9], double out[4]) {
    M3 m;
    for (int k = 0; k < 3; ++k) m.col[k] = {dl_dr[3 * k], dl_dr[3 * k + 1], dl_dr[3 * k + 2]};
    quat_grad(q, m, out);
}

int oref_sh_basis(int degree, const double d[3], double* out) {
    return sh_basis(degree, V3(d[0], d[1], d[2]), out);
}

int oref_sh_basis_grad(int degree, const double d[3], double* out) {
    V3 g[16];
    int k = sh_basis_grad(degree, V3(d[0], d[1], d[2]), g);
    for (int i = 0; i < k; ++i) {
        out[3 * i] = g[i].x;
        out[3 * i + 1] = g[i].y;
        out[3 * i + 2] = g[i].z;
    }
    return k;
}

int64_t oref_prepare_order(const oref_scene* s, const oref_camera* cam, int64_t* order) {
    Scene sc{s, num_coeffs(s->sh_degree)};
    auto surfels = prepare(sc, cam);
    if (order)
        for (size_t i = 0; i < surfels.size(); ++i) order[i] = surfels[i].idx;
    return (int64_t)surfels.size();
}

// forward.rs:104-123 per surfel (FP64): rects[4 i] = (x0, y0, x1, y1) tile
// rect and on_screen[i] for every surfel in front of the near plane;
// culled[i] = 1 for those behind it (rect left zero). A test aid.
int oref_prepare_rects(const oref_scene* s, const oref_camera* cam, int32_t* rects, uint8_t* on_screen,
                       uint8_t* culled) {
    if (!s || !cam || !rects || !on_screen || !culled) return -1;
    Scene sc{s, num_coeffs(s->sh_degree)};
    M3 r_cam = quat_to_mat(cam->quat);
    V3 C = cam_center(cam);
    for (int64_t i = 0; i < s->n; ++i) {
        V3 c = r_cam.tmul(sc.pos(i) - C);
        culled[i] = -c.z < NEAR_DEPTH;
        on_screen[i] = 0;
        rects[4 * i] = rects[4 * i + 1] = rects[4 * i + 2] = rects[4 * i + 3] = 0;
        if (culled[i]) continue;
        SurfelCam sf = make_surfel_cam(sc, cam, r_cam, i);
        on_screen[i] = sf.on_screen;
        rects[4 * i] = sf.x0;
        rects[4 * i + 1] = sf.y0;
        rects[4 * i + 2] = sf.x1;
        rects[4 * i + 3] = sf.y1;
    }
    return 0;
}

// forward.rs:207-246
int oref_render(const oref_scene* s, const oref_camera* cam, int mode, double* color,
                double* alpha, double* depth, double* normal, int64_t* n_hits) {
    if (!s || !cam || cam->width <= 0 || cam->height <= 0) return -1;
    Scene sc{s, num_coeffs(s->sh_degree)};
    auto surfels = prepare(sc, cam);
    TileLists tl;
    const TileLists* tlp = nullptr;
    if (mode == 1) {
        tl = build_tile_lists(surfels, cam);
        tlp = &tl;
    }
    const Window win = window_of(cam);
    int64_t hits = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : hits) if (mode == 1)
    for (int yy = 0; yy < win.h; ++yy) {
        for (int xx = 0; xx < win.w; ++xx) {
            const int x = win.x0 + xx, y = win.y0 + yy;
            V3 d = pixel_ray(cam, x, y);
            double pix_x = x + 0.5, pix_y = y + 0.5;
            int64_t i = (int64_t)yy * win.w + xx;
            double transmittance = 1.0, a_acc = 0.0, d_acc = 0.0;
            V3 c_acc, n_acc;
            walk_pixel(surfels, tlp, x, y, [&](int64_t li) {
                const SurfelCam& sf = surfels[li];
                ++hits;
                Hit hit = evaluate(sf, d, pix_x, pix_y);
                double a = std::min(sf.opacity * hit.ghat, ALPHA_CLAMP);
                if (a < ALPHA_SKIP) return true;
                double w = transmittance * a;
                c_acc += sf.color * w;
                a_acc += w;
                d_acc += hit.t * w;
                n_acc += sf.normal * w;
                transmittance *= 1.0 - a;
                return !(transmittance < T_TERMINATE);
            });
            if (color) {
                color[3 * i] = c_acc.x;
                color[3 * i + 1] = c_acc.y;
                color[3 * i + 2] = c_acc.z;
            }
            if (alpha) alpha[i] = a_acc;
            if (depth) depth[i] = d_acc;
            if (normal) {
                normal[3 * i] = n_acc.x;
                normal[3 * i + 1] = n_acc.y;
                normal[3 * i + 2] = n_acc.z;
            }
        }
    }
    if (n_hits) *n_hits = hits;
    return 0;
}

// Decisions of the forward walk that sit within a relative band `rel` of a
// threshold, where an evaluation in f32 (or any other rounding) may decide
// the other way: alpha vs the 1/255 skip and the 0.995 clamp, the ray vs
// screen branch (g_ray vs g_scr) and the ray's near plane, transmittance vs
// the 1e-4 stop (flags the surfel that reached it and the next contributor),
// the tile-rect truncation (px -+ r) / 16 near an integer, the on-screen test,
// the near cull, and equal-to-two-ulp f32 depth keys of consecutive surfels
// whose rects overlap (their order may swap). Not part of forward.rs: a
// test aid that tells whether an f32 path and this FP64 evaluation must make
// the same discrete decisions. flags (n, optional) gets 1 for every surfel in
// a marginal decision; returns the number of marginal decisions.
int64_t oref_marginal_decisions(const oref_scene* s, const oref_camera* cam, double rel, uint8_t* flags) {
    if (!s || !cam || cam->width <= 0 || cam->height <= 0) return -1;
    Scene sc{s, num_coeffs(s->sh_degree)};
    const int64_t n = s->n;
    std::vector<uint8_t> fl(n, 0);
    int64_t count = 0;
    auto near = [&](double v, double thr) { return std::fabs(v - thr) <= rel * std::max(std::fabs(thr), 1e-300); };
    // Near cull (on every surfel, before prepare drops the culled ones).
    {
        M3 r_cam = quat_to_mat(cam->quat);
        V3 C = cam_center(cam);
        for (int64_t i = 0; i < n; ++i) {
            const double depth = -r_cam.tmul(sc.pos(i) - C).z;
            if (near(depth, NEAR_DEPTH)) {
                fl[i] = 1;
                ++count;
            }
        }
    }
    auto surfels = prepare(sc, cam);
    // Rect truncation and the on-screen test (forward.rs:104-123).
    for (const auto& sf : surfels) {
        if (sf.depth < NEAR_DEPTH) continue;
        const double r_raw = 3.0 * cam->focal * (length(sf.tu) + length(sf.tv)) / sf.depth + 3.0;
        const double r_px = std::min(r_raw, 64.0);
        bool m = near(r_raw, 64.0);
        const double edges[4] = {sf.px - r_px, sf.px + r_px, sf.py - r_px, sf.py + r_px};
        for (double e : edges) {
            const double v = e / TILE;
            const double k = std::round(v);
            if (std::fabs(v - k) <= rel * std::max(std::fabs(v), 1.0)) m = true;
        }
        m = m || near(sf.px + r_px, 0.0) || near(sf.px - r_px, (double)cam->width) ||
            near(sf.py + r_px, 0.0) || near(sf.py - r_px, (double)cam->height);
        if (m) {
            fl[sf.idx] = 1;
            ++count;
        }
    }
    // Depth-order ties.
    for (size_t i = 1; i < surfels.size(); ++i) {
        const SurfelCam &a = surfels[i - 1], &b = surfels[i];
        float fa = (float)a.depth, fb = (float)b.depth;
        uint32_t ka, kb;
        std::memcpy(&ka, &fa, 4);
        std::memcpy(&kb, &fb, 4);
        const bool overlap = a.on_screen && b.on_screen && a.x0 <= b.x1 && b.x0 <= a.x1 && a.y0 <= b.y1 &&
                             b.y0 <= a.y1;
        if (overlap && kb - ka <= 2u) {
            fl[a.idx] = fl[b.idx] = 1;
            ++count;
        }
    }
    // The per-pixel walk (rows in parallel; a flag is only ever set to 1).
    TileLists tl = build_tile_lists(surfels, cam);
    const Window win = window_of(cam);
    uint8_t* flp = fl.data();
    auto mark = [flp](int64_t i) {
#pragma omp atomic write
        flp[i] = 1;
    };
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : count)
    for (int yy = 0; yy < win.h; ++yy) {
        for (int xx = 0; xx < win.w; ++xx) {
            const int x = win.x0 + xx, y = win.y0 + yy;
            V3 d = pixel_ray(cam, x, y);
            const double pix_x = x + 0.5, pix_y = y + 0.5;
            double transmittance = 1.0;
            bool flag_next = false, stopped = false;
            walk_pixel(surfels, &tl, x, y, [&](int64_t li) {
                const SurfelCam& sf = surfels[li];
                // evaluate() with the branch quantities kept.
                V3 sv = -d;
                const double det = triple(sf.tu, sf.tv, sv);
                double g_ray = 0.0, t = 0.0;
                bool m = false;
                if (std::fabs(det) > 1e-12) {
                    const double u = triple(-sf.c, sf.tv, sv) / det;
                    const double v = triple(sf.tu, -sf.c, sv) / det;
                    t = triple(sf.tu, sf.tv, -sf.c) / det;
                    const double g = std::exp(-(u * u + v * v) * 0.5);
                    if (t > NEAR_DEPTH) g_ray = g;
                    m = m || (near(t, NEAR_DEPTH) && sf.opacity * g >= ALPHA_SKIP * (1.0 - rel));
                }
                const double dx = pix_x - sf.px, dy = pix_y - sf.py;
                const double g_scr = std::exp(-(dx * dx + dy * dy) / (2.0 * LOWPASS_SIGMA2));
                const double ghat = std::max(g_ray, g_scr);
                const double raw = sf.opacity * ghat;
                if (raw >= ALPHA_SKIP * (1.0 - rel))
                    m = m || std::fabs(g_ray - g_scr) <= rel * ghat;  // branch matters only if it can contribute
                m = m || near(raw, ALPHA_SKIP) || near(raw, ALPHA_CLAMP);
                const double alpha = std::min(raw, ALPHA_CLAMP);
                const bool contributes = alpha >= ALPHA_SKIP;
                if (stopped) {  // past a marginal stop: the next contributor
                    if (contributes) {
                        mark(sf.idx);
                        return false;
                    }
                    return true;
                }
                if (m) {
                    mark(sf.idx);
                    ++count;
                }
                if (!contributes) return true;
                if (flag_next) {
                    mark(sf.idx);
                    flag_next = false;
                }
                transmittance *= 1.0 - alpha;
                if (near(transmittance, T_TERMINATE)) {
                    mark(sf.idx);
                    ++count;
                    flag_next = true;
                    if (transmittance < T_TERMINATE) stopped = true;
                    return true;
                }
                return !(transmittance < T_TERMINATE);
            });
        }
    }
    if (flags) std::memcpy(flags, fl.data(), (size_t)n);
    return count;
}

// The surfels that composite (alpha >= 1/255, before the transmittance stop)
// at each of the npix pixels xy[2k], xy[2k + 1]: flags[i] = 1. A test aid (no
// forward.rs counterpart) for masking surfels whose gradient depends on a
// per-pixel decision made elsewhere (the L1 sign where render and target
// nearly agree).
int oref_pixel_contributors(const oref_scene* s, const oref_camera* cam, int64_t npix, const int32_t* xy,
                            uint8_t* flags) {
    if (!s || !cam || !flags || cam->width <= 0 || cam->height <= 0) return -1;
    Scene sc{s, num_coeffs(s->sh_degree)};
    auto surfels = prepare(sc, cam);
    TileLists tl = build_tile_lists(surfels, cam);
    std::memset(flags, 0, (size_t)s->n);
    for (int64_t k = 0; k < npix; ++k) {
        const int x = xy[2 * k], y = xy[2 * k + 1];
        if (x < 0 || y < 0 || x >= cam->width || y >= cam->height) continue;
        V3 d = pixel_ray(cam, x, y);
        double transmittance = 1.0;
        walk_pixel(surfels, &tl, x, y, [&](int64_t li) {
            const SurfelCam& sf = surfels[li];
            Hit hit = evaluate(sf, d, x + 0.5, y + 0.5);
            const double a = std::min(sf.opacity * hit.ghat, ALPHA_CLAMP);
            if (a < ALPHA_SKIP) return true;
            flags[sf.idx] = 1;
            transmittance *= 1.0 - a;
            return !(transmittance < T_TERMINATE);
        });
    }
    return 0;
}

// forward.rs:174-205
double oref_distortion_loss(const oref_scene* s, const oref_camera* cam, int mode) {
    Scene sc{s, num_coeffs(s->sh_degree)};
    auto surfels = prepare(sc, cam);
    TileLists tl;
    const TileLists* tlp = nullptr;
    if (mode == 1) {
        tl = build_tile_lists(surfels, cam);
        tlp = &tl;
    }
    double total = 0.0;
    for (int y = 0; y < cam->height; ++y) {
        for (int x = 0; x < cam->width; ++x) {
            V3 d = pixel_ray(cam, x, y);
            double pix_x = x + 0.5, pix_y = y + 0.5;
            double transmittance = 1.0, a_pre = 0.0, b_pre = 0.0;
            walk_pixel(surfels, tlp, x, y, [&](int64_t li) {
                const SurfelCam& sf = surfels[li];
                Hit hit = evaluate(sf, d, pix_x, pix_y);
                double a = std::min(sf.opacity * hit.ghat, ALPHA_CLAMP);
                if (a < ALPHA_SKIP) return true;
                double w = transmittance * a;
                total += 2.0 * w * (hit.t * a_pre - b_pre);
                a_pre += w;
                b_pre += w * hit.t;
                transmittance *= 1.0 - a;
                return !(transmittance < T_TERMINATE);
            });
        }
    }
    return total;
}

// backward.rs:176-245: camera-space accumulators acc[li] of surfels[li] ->
// parameter and camera gradients (the focal term is added by the caller).
void chain_to_params(const std::vector<SurfelCam>& surfels, const std::vector<CamSpaceAcc>& acc,
                     const Scene& sc, const oref_camera* cam, oref_grads* g) {
    const oref_scene* s = sc.s;
    const int64_t n = s->n;

    const int ncoef = sc.ncoef;
    std::fill(g->pos, g->pos + 3 * n, 0.0);
    std::fill(g->scales, g->scales + 2 * n, 0.0);
    std::fill(g->quat, g->quat + 4 * n, 0.0);
    std::fill(g->opacity, g->opacity + n, 0.0);
    std::fill(g->sh, g->sh + n * ncoef * 3, 0.0);
    V3 gcenter;
    M3 r_cam = quat_to_mat(cam->quat);
    M3 dl_dr_cam;
    V3 C = cam_center(cam);
    double basis[16];
    V3 bgrad[16];
    // backward.rs:176-245: camera-space accumulators -> parameter gradients.
    for (size_t li = 0; li < surfels.size(); ++li) {
        const SurfelCam& sf = surfels[li];
        const CamSpaceAcc& a = acc[li];
        const int64_t i = sf.idx;
        M3 rs = quat_to_mat(s->quat + 4 * i);
        V3 a_dtu = a.dtu, a_dtv = a.dtv;
        if (!is_zero(a.dnormal)) {  // backward.rs:186-196
            V3 m = cross(sf.tu, sf.tv);
            double len = length(m);
            if (len > 1e-12) {
                V3 mhat = m / len;
                double sign = dot(mhat, sf.c) > 0.0 ? -1.0 : 1.0;
                V3 dl_dm = (a.dnormal - mhat * dot(mhat, a.dnormal)) * (sign / len);
                a_dtu += cross(sf.tv, dl_dm);
                a_dtv += cross(dl_dm, sf.tu);
            }
        }
        V3 p = sc.pos(i);
        V3 x = p - C;  // backward.rs:198-203
        V3 dx = r_cam.mul(a.dc);
        V3 gp(g->pos[3 * i], g->pos[3 * i + 1], g->pos[3 * i + 2]);
        gp += dx;
        gcenter -= dx;
        add_transpose_chain(dl_dr_cam, x, a.dc);
        for (int axis = 0; axis < 2; ++axis) {  // backward.rs:206-219
            V3 dtau = axis == 0 ? a_dtu : a_dtv;
            double scale = s->scales[2 * i + axis];
            V3 axis_world = rs.col[axis];
            V3 y_vec = axis_world * scale;
            V3 dy = r_cam.mul(dtau);
            g->scales[2 * i + axis] += dot(dy, axis_world);
            M3 dl_drs;
            dl_drs.col[axis] = dy * scale;
            double dq[4];
            quat_grad(s->quat + 4 * i, dl_drs, dq);
            for (int k = 0; k < 4; ++k) g->quat[4 * i + k] += dq[k];
            add_transpose_chain(dl_dr_cam, y_vec, dtau);
        }
        V3 dlc = a.dcolor;  // backward.rs:221-242
        for (int ch = 0; ch < 3; ++ch)
            if (sf.clamped[ch]) dlc[ch] = 0.0;
        if (!is_zero(dlc)) {
            V3 v = p - C;
            double len = length(v);
            V3 dir = v / len;
            int nb = sh_basis(s->sh_degree, dir, basis);
            sh_basis_grad(s->sh_degree, dir, bgrad);
            V3 ddir;
            for (int k = 0; k < nb; ++k) {
                double* gs = g->sh + (i * ncoef + k) * 3;
                gs[0] += dlc.x * basis[k];
                gs[1] += dlc.y * basis[k];
                gs[2] += dlc.z * basis[k];
                ddir += bgrad[k] * dot(sc.sh(i, k), dlc);
            }
            V3 dv = (ddir - dir * dot(dir, ddir)) / len;
            gp += dv;
            gcenter -= dv;
        }
        g->pos[3 * i] = gp.x;
        g->pos[3 * i + 1] = gp.y;
        g->pos[3 * i + 2] = gp.z;
        g->opacity[i] += a.dopacity;
    }
    g->cam_center[0] = gcenter.x;
    g->cam_center[1] = gcenter.y;
    g->cam_center[2] = gcenter.z;
    quat_grad(cam->quat, dl_dr_cam, g->cam_quat);
}

// backward.rs:57-249
int oref_gradients(const oref_scene* s, const oref_camera* cam, int mode,
                   const double* dl_dcolor, const double* dl_ddepth, const double* dl_dnormal,
                   double lambda_dist, oref_grads* g) {
    if (!s || !cam || !dl_dcolor || !g) return -1;
    Scene sc{s, num_coeffs(s->sh_degree)};
    auto surfels = prepare(sc, cam);
    TileLists tl;
    const TileLists* tlp = nullptr;
    if (mode == 1) {
        tl = build_tile_lists(surfels, cam);
        tlp = &tl;
    }
    const Window win = window_of(cam);
    // Per-thread accumulators, reduced in thread order (literal mode is
    // single-threaded and so accumulates in exactly the reference's order).
    int nthreads = mode == 1 ? threads_available() : 1;
    std::vector<Workspace> ws(nthreads);
    for (auto& w : ws) w.acc.assign(surfels.size(), CamSpaceAcc());
#pragma omp parallel for schedule(static, 1) num_threads(nthreads) if (mode == 1)
    for (int yy = 0; yy < win.h; ++yy) {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        for (int xx = 0; xx < win.w; ++xx) {
            const int x = win.x0 + xx, y = win.y0 + yy;
            int64_t pi = (int64_t)yy * win.w + xx;
            V3 dlc(dl_dcolor[3 * pi], dl_dcolor[3 * pi + 1], dl_dcolor[3 * pi + 2]);
            double dld = dl_ddepth ? dl_ddepth[pi] : 0.0;
            V3 dln = dl_dnormal ? V3(dl_dnormal[3 * pi], dl_dnormal[3 * pi + 1], dl_dnormal[3 * pi + 2])
                                : V3();
            backward_pixel(surfels, tlp, cam, x, y, dlc, dld, dln, lambda_dist, ws[tid]);
        }
    }
    for (int t = 1; t < nthreads; ++t) {
        for (size_t i = 0; i < surfels.size(); ++i) add_acc(ws[0].acc[i], ws[t].acc[i]);
        ws[0].dfocal += ws[t].dfocal;
    }
    chain_to_params(surfels, ws[0].acc, sc, cam, g);
    g->focal = ws[0].dfocal;
    return 0;
}

// The contribution lists of the faithful FP64 walk (forward.rs:207-246),
// in oref_replay's format, over the camera's window: what an evaluation
// with exactly FP64's decisions would record. counts[p] may exceed cap
// (entries past cap are counted, not stored). A test aid.
int oref_trace(const oref_scene* s, const oref_camera* cam, int64_t cap, uint32_t* lists, uint32_t* counts) {
    if (!s || !cam || !lists || !counts || cap <= 0) return -1;
    Scene sc{s, num_coeffs(s->sh_degree)};
    auto surfels = prepare(sc, cam);
    TileLists tl = build_tile_lists(surfels, cam);
    const Window win = window_of(cam);
#pragma omp parallel for schedule(dynamic, 1)
    for (int yy = 0; yy < win.h; ++yy) {
        for (int xx = 0; xx < win.w; ++xx) {
            const int x = win.x0 + xx, y = win.y0 + yy;
            const int64_t pi = (int64_t)yy * win.w + xx;
            V3 d = pixel_ray(cam, x, y);
            double pix_x = x + 0.5, pix_y = y + 0.5, transmittance = 1.0;
            uint32_t cnt = 0;
            walk_pixel(surfels, &tl, x, y, [&](int64_t li) {
                const SurfelCam& sf = surfels[li];
                Hit hit = evaluate(sf, d, pix_x, pix_y);
                double raw = sf.opacity * hit.ghat;
                double a = std::min(raw, ALPHA_CLAMP);
                if (a < ALPHA_SKIP) return true;
                if ((int64_t)cnt < cap)
                    lists[pi * cap + cnt] = (uint32_t)sf.idx | (hit.ray_branch ? 1u << 30 : 0u) |
                                            (raw > ALPHA_CLAMP ? 1u << 31 : 0u);
                ++cnt;
                transmittance *= 1.0 - a;
                return !(transmittance < T_TERMINATE);
            });
            counts[pi] = cnt;
        }
    }
    return 0;
}

// Decision replay (a test aid with no forward.rs counterpart: it separates
// FP32 rounding from the discrete decisions that FP32 and FP64 can take
// differently near the alpha skip, clamp, stop and branch thresholds): the
// FP64 forward and backward along the discrete decisions another evaluation
// took. For window pixel p (row-major over the camera's crop window),
// lists[p * cap + k], k < counts[p], is its k-th contribution in compositing
// order: surfel index (bits 0..29) | ray branch (bit 30) | alpha clamped at
// 0.995 (bit 31). Each listed hit is evaluated in FP64 with the given branch
// (no near-plane or det test: the other evaluation took the branch) and the
// given clamp (alpha = 0.995 if clamped, else opacity * ghat with no skip
// test); the walk ends with the list (the transmittance stop is replayed
// too). The surfel's colour clamp stays FP64's. dl_ddepth (P), dl_dnormal
// (3P) and lambda_dist as in oref_gradients. Gradients follow the
// backward.rs:112-245 chain along the replayed walk, so they are the exact
// derivative of the function the decisions define. color (P x 3) and alpha
// (P) optional; dl_dcolor + g optional (P x 3). Returns -2 if a count
// exceeds cap or a surfel index is out of range.
int oref_replay(const oref_scene* s, const oref_camera* cam, const uint32_t* lists, const uint32_t* counts,
                int64_t cap, double* color, double* alpha, const double* dl_dcolor, const double* dl_ddepth,
                const double* dl_dnormal, double lambda_dist, oref_grads* g) {
    if (!s || !cam || !lists || !counts || cap <= 0) return -1;
    Scene sc{s, num_coeffs(s->sh_degree)};
    const int64_t n = s->n;
    M3 r_cam = quat_to_mat(cam->quat);
    std::vector<SurfelCam> all(n);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) all[i] = make_surfel_cam(sc, cam, r_cam, i);
    const Window win = window_of(cam);
    const int64_t P = (int64_t)win.w * win.h;
    for (int64_t p = 0; p < P; ++p) {
        if ((int64_t)counts[p] > cap) return -2;
        for (uint32_t k = 0; k < counts[p]; ++k)
            if ((int64_t)(lists[p * cap + k] & 0x3FFFFFFFu) >= n) return -2;
    }
    const bool grads = dl_dcolor != nullptr && g != nullptr;
    int nthreads = threads_available();
    std::vector<Workspace> ws(nthreads);
    if (grads)
        for (auto& w : ws) w.acc.assign(n, CamSpaceAcc());
#pragma omp parallel for schedule(static, 1) num_threads(nthreads)
    for (int yy = 0; yy < win.h; ++yy) {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        Workspace& w = ws[tid];
        for (int xx = 0; xx < win.w; ++xx) {
            const int x = win.x0 + xx, y = win.y0 + yy;
            const int64_t pi = (int64_t)yy * win.w + xx;
            V3 d = pixel_ray(cam, x, y);
            V3 sv = -d;
            double pix_x = x + 0.5, pix_y = y + 0.5;
            w.contribs.clear();
            double transmittance = 1.0, w_total = 0.0, wt_total = 0.0;
            V3 c_acc;
            for (uint32_t k = 0; k < counts[pi]; ++k) {
                const uint32_t e = lists[pi * cap + k];
                const int64_t id = (int64_t)(e & 0x3FFFFFFFu);
                const bool ray = (e >> 30) & 1u, clamped = (e >> 31) & 1u;
                const SurfelCam& sf = all[id];
                // forward.rs:136-169 with the branch given.
                Hit hit;
                if (ray) {
                    double det = triple(sf.tu, sf.tv, sv);
                    double u = 0.0, v = 0.0, t = 0.0, gr = 0.0;
                    if (det != 0.0) {
                        u = triple(-sf.c, sf.tv, sv) / det;
                        v = triple(sf.tu, -sf.c, sv) / det;
                        t = triple(sf.tu, sf.tv, -sf.c) / det;
                        gr = std::exp(-(u * u + v * v) * 0.5);
                    }
                    hit = {gr, true, u, v, t};
                } else {
                    double dx = pix_x - sf.px, dy = pix_y - sf.py;
                    hit = {std::exp(-(dx * dx + dy * dy) / (2.0 * LOWPASS_SIGMA2)), false, 0.0, 0.0, sf.depth};
                }
                const double a = clamped ? ALPHA_CLAMP : sf.opacity * hit.ghat;
                const double wgt = transmittance * a;
                c_acc += sf.color * wgt;
                w_total += wgt;
                wt_total += wgt * hit.t;
                w.contribs.push_back({id, a, clamped, transmittance, hit});
                transmittance *= 1.0 - a;
            }
            if (color) {
                color[3 * pi] = c_acc.x;
                color[3 * pi + 1] = c_acc.y;
                color[3 * pi + 2] = c_acc.z;
            }
            if (alpha) alpha[pi] = w_total;
            if (grads) {
                V3 dlc(dl_dcolor[3 * pi], dl_dcolor[3 * pi + 1], dl_dcolor[3 * pi + 2]);
                double dld = dl_ddepth ? dl_ddepth[pi] : 0.0;
                V3 dln = dl_dnormal ? V3(dl_dnormal[3 * pi], dl_dnormal[3 * pi + 1], dl_dnormal[3 * pi + 2]) : V3();
                if (!is_zero(dlc) || dld != 0.0 || !is_zero(dln) || lambda_dist != 0.0)
                    reverse_walk(all, cam, d, pix_x, pix_y, w_total, wt_total, dlc, dld, dln, lambda_dist, w);
            }
        }
    }
    if (!grads) return 0;
    for (int t = 1; t < nthreads; ++t) {
        for (int64_t i = 0; i < n; ++i) add_acc(ws[0].acc[i], ws[t].acc[i]);
        ws[0].dfocal += ws[t].dfocal;
    }
    chain_to_params(all, ws[0].acc, sc, cam, g);
    g->focal = ws[0].dfocal;
    return 0;
}

// first_attempt/crates/gs-train/src/shaders/normal_loss.wgsl:52-145 (pass1 +
// pass2) in FP64: inputs per pixel color4 (alpha in [3]), aux4 (depth_acc in
// [1]), normal4 (rows of 4, row-major W x H); lambda is the kernel weight
// (the reference's set_lambda divides by W H before it reaches the kernel).
// Outputs: loss_map[P] = alpha_q (1 - nhat . N_d), dl_dnormal[3P],
// dl_ddepth[P] = dL/d(depth_acc) (the value pass2 writes to dl_dcolor.w).
void oref_normal_loss(int w, int h, double focal, double lambda, double alpha_min, const double* color4,
                      const double* aux4, const double* normal4, double* loss_map, double* dl_dnormal,
                      double* dl_ddepth) {
    const int64_t P = (int64_t)w * h;
    std::vector<V3> cpx(P), cpy(P);
    const double cx = w * 0.5, cy = h * 0.5;
    auto ray = [&](int x, int y) { return V3((x + 0.5 - cx) / focal, -(y + 0.5 - cy) / focal, -1.0); };
    auto alpha = [&](int64_t i) { return color4[4 * i + 3]; };
    auto tbar = [&](int64_t i) { return alpha(i) < alpha_min ? 0.0 : aux4[4 * i + 1] / alpha(i); };
    for (int64_t i = 0; i < P; ++i) {
        loss_map[i] = 0.0;
        dl_dnormal[3 * i] = dl_dnormal[3 * i + 1] = dl_dnormal[3 * i + 2] = 0.0;
        if (lambda == 0.0) continue;
        const int x = (int)(i % w), y = (int)(i / w);
        if (x + 1 >= w || y + 1 >= h) continue;
        const int64_t ix = i + 1, iy = i + w;
        const double a_q = alpha(i);
        if (a_q < alpha_min || alpha(ix) < alpha_min || alpha(iy) < alpha_min) continue;
        V3 p = ray(x, y) * tbar(i);
        V3 dpx = ray(x + 1, y) * tbar(ix) - p;
        V3 dpy = ray(x, y + 1) * tbar(iy) - p;
        V3 m = cross(dpx, dpy);
        double len2 = dot(m, m);
        if (len2 < 1e-18) continue;
        double len = std::sqrt(len2);
        V3 mhat = m / len;
        double sign = dot(mhat, p) > 0.0 ? -1.0 : 1.0;
        V3 nd = mhat * sign;
        V3 av(normal4[4 * i], normal4[4 * i + 1], normal4[4 * i + 2]);
        double la2 = dot(av, av);
        if (la2 < 1e-18) continue;
        double la = std::sqrt(la2);
        V3 nhat = av / la;
        loss_map[i] = (1.0 - dot(nhat, nd)) * a_q;
        V3 dl_dnd = nhat * (-lambda * a_q);
        V3 dl_dnhat = nd * (-lambda * a_q);
        V3 dl_da = (dl_dnhat - nhat * dot(nhat, dl_dnhat)) / la;
        dl_dnormal[3 * i] = dl_da.x;
        dl_dnormal[3 * i + 1] = dl_da.y;
        dl_dnormal[3 * i + 2] = dl_da.z;
        V3 dl_dm = (dl_dnd - mhat * dot(mhat, dl_dnd)) * (sign / len);
        cpx[i] = cross(dpy, dl_dm);
        cpy[i] = cross(dl_dm, dpx);
    }
    for (int64_t i = 0; i < P; ++i) {
        dl_ddepth[i] = 0.0;
        if (lambda == 0.0) continue;
        const int x = (int)(i % w), y = (int)(i / w);
        V3 acc = -cpx[i] - cpy[i];
        if (x > 0) acc += cpx[i - 1];
        if (y > 0) acc += cpy[i - w];
        if (alpha(i) >= alpha_min) dl_ddepth[i] = dot(ray(x, y), acc) / alpha(i);
    }
}

// ------------------------------------------------ fp32 mirror: integer stages

// surfel_prep_fwd.wgsl:64-132 (rect + depth key + count; count_tiles of
// tile_binning.wgsl:30-43 fused). Operation order = splattar_amd/csrc
// project kernel, every op a separately rounded f32 operation.
void oref_prep_f32(int64_t n, const float* pos4, const float* scales2, const float* quat4,
                   const oref_cam32* cam, uint32_t* rects, uint32_t* depth_keys,
                   uint32_t* counts) {
    const float* r0 = cam->rot + 0;
    const float* r1 = cam->rot + 3;
    const float* r2 = cam->rot + 6;
    const float fw = (float)cam->width, fh = (float)cam->height;
    const float cx = fw * 0.5f, cy = fh * 0.5f;
    const float max_tx = (float)(cam->tiles_x - 1), max_ty = (float)(cam->tiles_y - 1);
    for (int64_t i = 0; i < n; ++i) {
        uint32_t* rect = rects + 4 * i;
        rect[0] = 1;
        rect[1] = 1;
        rect[2] = 0;
        rect[3] = 0;
        depth_keys[i] = 0;
        counts[i] = 0;
        float xx = pos4[4 * i] - cam->center[0];
        float xy = pos4[4 * i + 1] - cam->center[1];
        float xz = pos4[4 * i + 2] - cam->center[2];
        float ccx = dot3f(r0, xx, xy, xz);
        float ccy = dot3f(r1, xx, xy, xz);
        float ccz = dot3f(r2, xx, xy, xz);
        float depth = -ccz;
        if (!(depth >= 0.05f)) continue;  // near cull (NaN culls too)
        // quat_mat (surfel_prep_fwd.wgsl:54-62): q / |q|.
        const float* q = quat4 + 4 * i;
        float qq = ((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]) + q[3] * q[3];
        float qn = std::sqrt(qq);
        float x = q[0] / qn, y = q[1] / qn, z = q[2] / qn, w = q[3] / qn;
        // Columns 0 and 1 of R(q) (math.rs:14-34 layout).
        float a0 = 1.0f - 2.0f * (y * y + z * z);
        float a1 = 2.0f * (x * y + w * z);
        float a2 = 2.0f * (x * z - w * y);
        float b0 = 2.0f * (x * y - w * z);
        float b1 = 1.0f - 2.0f * (x * x + z * z);
        float b2 = 2.0f * (y * z + w * x);
        float su = scales2[2 * i], sv = scales2[2 * i + 1];
        float ux = a0 * su, uy = a1 * su, uz = a2 * su;
        float vx = b0 * sv, vy = b1 * sv, vz = b2 * sv;
        float tux = dot3f(r0, ux, uy, uz), tuy = dot3f(r1, ux, uy, uz), tuz = dot3f(r2, ux, uy, uz);
        float tvx = dot3f(r0, vx, vy, vz), tvy = dot3f(r1, vx, vy, vz), tvz = dot3f(r2, vx, vy, vz);
        float lu = std::sqrt((tux * tux + tuy * tuy) + tuz * tuz);
        float lv = std::sqrt((tvx * tvx + tvy * tvy) + tvz * tvz);
        float px = cx + (cam->focal * ccx) / depth;
        float py = cy - (cam->focal * ccy) / depth;
        float r = ((3.0f * cam->focal) * (lu + lv)) / depth + 3.0f;
        r = r < 64.0f ? r : 64.0f;
        uint32_t key;
        std::memcpy(&key, &depth, 4);
        depth_keys[i] = key;
        if (px + r >= 0.0f && px - r < fw && py + r >= 0.0f && py - r < fh) {
            auto cl = [](float v, float hi) {
                float c = v > 0.0f ? v : 0.0f;  // max(v, 0)
                c = c < hi ? c : hi;             // min(., hi)
                return (uint32_t)c;
            };
            uint32_t x0 = cl((px - r) / 16.0f, max_tx);
            uint32_t y0 = cl((py - r) / 16.0f, max_ty);
            uint32_t x1 = cl((px + r) / 16.0f, max_tx);
            uint32_t y1 = cl((py + r) / 16.0f, max_ty);
            rect[0] = x0;
            rect[1] = y0;
            rect[2] = x1;
            rect[3] = y1;
            counts[i] = (x1 - x0 + 1) * (y1 - y0 + 1);
        }
    }
}

uint32_t oref_exclusive_scan_u32(int64_t n, const uint32_t* in, uint32_t* out) {
    uint32_t acc = 0;
    for (int64_t i = 0; i < n; ++i) {
        uint32_t v = in[i];
        out[i] = acc;
        acc += v;
    }
    return acc;
}

// tile_binning.wgsl:46-66 with (tile << 32 | depth) keys and surfel payloads.
int64_t oref_emit_keys(int64_t n, const uint32_t* rects, const uint32_t* depth_keys,
                       uint32_t tiles_x, uint32_t tile_base, uint64_t* keys,
                       uint32_t* payloads, int64_t max_entries) {
    int64_t e = 0;
    for (int64_t i = 0; i < n; ++i) {
        const uint32_t* r = rects + 4 * i;
        if (r[2] < r[0] || r[3] < r[1]) continue;
        for (uint32_t ty = r[1]; ty <= r[3]; ++ty)
            for (uint32_t tx = r[0]; tx <= r[2]; ++tx) {
                if (e < max_entries) {
                    uint64_t tile = (uint64_t)(tile_base + ty * tiles_x + tx);
                    keys[e] = (tile << 32) | depth_keys[i];
                    payloads[e] = (uint32_t)i;
                }
                ++e;
            }
    }
    return e;
}

void oref_stable_sort_u64(int64_t n, uint64_t* keys, uint32_t* payloads) {
    std::vector<int64_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) { return keys[a] < keys[b]; });
    std::vector<uint64_t> k2(n);
    std::vector<uint32_t> p2(n);
    for (int64_t i = 0; i < n; ++i) {
        k2[i] = keys[idx[i]];
        p2[i] = payloads[idx[i]];
    }
    std::memcpy(keys, k2.data(), n * 8);
    std::memcpy(payloads, p2.data(), n * 4);
}

// tile_binning.wgsl:89-103 (ranges cleared first, binning.rs:228).
void oref_tile_ranges(int64_t e, const uint64_t* keys, uint32_t num_tiles, uint32_t tile_base,
                      uint32_t* ranges) {
    std::memset(ranges, 0, sizeof(uint32_t) * 2 * (size_t)num_tiles);
    for (int64_t i = 0; i < e; ++i) {
        uint32_t tile = (uint32_t)(keys[i] >> 32) - tile_base;
        if (tile >= num_tiles) continue;
        if (i == 0 || (keys[i - 1] >> 32) != (keys[i] >> 32)) ranges[2 * tile] = (uint32_t)i;
        if (i == e - 1 || (keys[i + 1] >> 32) != (keys[i] >> 32)) ranges[2 * tile + 1] = (uint32_t)(i + 1);
    }
}

// ----------------------------------------------------------- L1 + D-SSIM

// ssim.wgsl:32-34 (the reference's f32 window constants, held in f64).
static const double GAUSS_W[6] = {0.26601818, 0.21300237, 0.10936069,
                                  0.03599398, 0.00759866, 0.00102838};

// ssim.wgsl:54-88: zero-padded separable blur (horizontal, then vertical).
void oref_blur(int w, int h, const double* in, double* out) {
    std::vector<double> tmp((size_t)w * h);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            double acc = in[(size_t)y * w + x] * GAUSS_W[0];
            for (int k = 1; k <= 5; ++k) {
                if (x >= k) acc += in[(size_t)y * w + x - k] * GAUSS_W[k];
                if (x + k < w) acc += in[(size_t)y * w + x + k] * GAUSS_W[k];
            }
            tmp[(size_t)y * w + x] = acc;
        }
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            double acc = tmp[(size_t)y * w + x] * GAUSS_W[0];
            for (int k = 1; k <= 5; ++k) {
                if (y >= k) acc += tmp[(size_t)(y - k) * w + x] * GAUSS_W[k];
                if (y + k < h) acc += tmp[(size_t)(y + k) * w + x] * GAUSS_W[k];
            }
            out[(size_t)y * w + x] = acc;
        }
}

// ssim.wgsl:41-143, ssim.rs:1-4,84, with D-SSIM = dssim_scale*(1-SSIM)
// (the reference uses dssim_scale = 0.5).
double oref_l1_ssim(int w, int h, const double* img, const double* target, double lambda,
                    double dssim_scale, double* dl_dimg, double* l1_mean_out,
                    double* ssim_mean_out) {
    const size_t P = (size_t)w * h;
    const double C1 = 0.0001, C2 = 0.0009;
    const double inv_n3 = 1.0 / ((double)P * 3.0);
    double l1_sum = 0.0, ssim_sum = 0.0;
    std::vector<double> x(P), y(P), xx(P), yy(P), xy(P), mux(P), muy(P), bx2(P), by2(P), bxy(P);
    std::vector<double> cmu(P), cx2(P), cxy(P), bcmu(P), bcx2(P), bcxy(P);
    for (int ch = 0; ch < 3; ++ch) {
        for (size_t i = 0; i < P; ++i) {
            x[i] = img[3 * i + ch];
            y[i] = target[3 * i + ch];
            xx[i] = x[i] * x[i];
            yy[i] = y[i] * y[i];
            xy[i] = x[i] * y[i];
        }
        oref_blur(w, h, x.data(), mux.data());
        oref_blur(w, h, y.data(), muy.data());
        oref_blur(w, h, xx.data(), bx2.data());
        oref_blur(w, h, yy.data(), by2.data());
        oref_blur(w, h, xy.data(), bxy.data());
        const double dw = -lambda * dssim_scale * inv_n3;
        for (size_t i = 0; i < P; ++i) {  // ssim.wgsl:95-126 (partials)
            double mx = mux[i], my = muy[i];
            double sx2 = bx2[i] - mx * mx, sy2 = by2[i] - my * my, sxy = bxy[i] - mx * my;
            double p = mx * mx + my * my + C1;
            double q = sx2 + sy2 + C2;
            double l = (2.0 * mx * my + C1) / p;
            double cs = (2.0 * sxy + C2) / q;
            ssim_sum += l * cs;
            double ds_dmux = cs * (2.0 * my - 2.0 * mx * l) / p + l * (-2.0 * my + 2.0 * mx * cs) / q;
            cmu[i] = ds_dmux * dw;
            cx2[i] = (-l * cs / q) * dw;
            cxy[i] = (2.0 * l / q) * dw;
            l1_sum += std::fabs(x[i] - y[i]);
        }
        if (dl_dimg) {  // ssim.wgsl:128-143 (combine)
            oref_blur(w, h, cmu.data(), bcmu.data());
            oref_blur(w, h, cx2.data(), bcx2.data());
            oref_blur(w, h, cxy.data(), bcxy.data());
            for (size_t i = 0; i < P; ++i) {
                double diff = x[i] - y[i];
                double sgn = diff > 0.0 ? 1.0 : (diff < 0.0 ? -1.0 : 0.0);
                dl_dimg[3 * i + ch] =
                    sgn * ((1.0 - lambda) * inv_n3) + bcmu[i] + 2.0 * x[i] * bcx2[i] + y[i] * bcxy[i];
            }
        }
    }
    double l1_mean = l1_sum * inv_n3, ssim_mean = ssim_sum * inv_n3;
    if (l1_mean_out) *l1_mean_out = l1_mean;
    if (ssim_mean_out) *ssim_mean_out = ssim_mean;
    return (1.0 - lambda) * l1_mean + lambda * dssim_scale * (1.0 - ssim_mean);
}

// ----------------------------------------------------------------- Adam

// adam.wgsl:33-74 with the host bias corrections of optim.rs:179-197
// (f32 powi by repeated multiplication, as Rust's powi lowers to).
void oref_adam_step_f32(int64_t n, float* raw, const float* grad_act, float* m, float* v,
                        float* activated, float lr, float reg, int activation, int t) {
    const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    float p1 = 1.0f, p2 = 1.0f;
    for (int k = 0; k < t; ++k) {
        p1 *= beta1;
        p2 *= beta2;
    }
    const float bc1_inv = 1.0f / (1.0f - p1), bc2_inv = 1.0f / (1.0f - p2);
    for (int64_t i = 0; i < n; ++i) {
        float x = raw[i];
        float da = 1.0f;
        if (activation == 1) {
            da = std::exp(x);
        } else if (activation == 2) {
            float s = 1.0f / (1.0f + std::exp(-x));
            da = s * (1.0f - s);
        }
        float g = (grad_act[i] + reg) * da;
        float mi = beta1 * m[i] + (1.0f - beta1) * g;
        float vi = beta2 * v[i] + (1.0f - beta2) * g * g;
        m[i] = mi;
        v[i] = vi;
        float mhat = mi * bc1_inv;
        float vhat = vi * bc2_inv;
        x -= lr * mhat / (std::sqrt(vhat) + eps);
        raw[i] = x;
        if (activated) {
            if (activation == 1)
                activated[i] = std::exp(x);
            else if (activation == 2)
                activated[i] = 1.0f / (1.0f + std::exp(-x));
            else
                activated[i] = x;
        }
    }
}

void oref_xorshift32_fill(uint32_t seed, int64_t n, uint32_t modulus, uint32_t* out) {
    uint32_t x = seed;
    for (int64_t i = 0; i < n; ++i) {
        x ^= x << 13;
        x ^= x >> 17;
        x ^= x << 5;
        out[i] = modulus ? x % modulus : x;
    }
}

int oref_num_threads(void) { return threads_available(); }

const char* oref_build_info(void) {
    return "g++ " __VERSION__
#ifdef __OPTIMIZE__
           " optimized"
#endif
#ifdef __AVX512F__
           " avx512f"
#endif
#ifdef __AVX2__
           " avx2"
#endif
#ifdef __FMA__
           " fma"
#endif
#ifdef _OPENMP
           " openmp"
#endif
        ;
}

}  // extern "C"
