// Geometry backward: one thread per surfel, all views of a batch.
//
// Restates crates/gs-kernels/src/shaders/surfel_prep_bwd.wgsl:129-284 (the
// oracle's backward.rs:176-245 and quat_grad of math.rs:38-88). The
// rasterizer backward hands over gradients w.r.t. the record coefficients
// (A = tv x c, B = c x tu, M = tu x tv, Nt = -M.c, px, py, depth, colour,
// opacity, normal); this kernel first maps them to the camera-space
// (dc, dtu, dtv) of the reference by the transposed cross-product rules, then
// chains to position, scales, quaternion, opacity and SH exactly as the
// reference does. Gradients of every view in the batch are summed per surfel
// before one write, so the parameters are read once per step, not per view.
// With cam_partials set, the per-view camera gradients [dL/dR (9, column
// major), dL/dC (3), dL/df, 0, 0, 0] are reduced per block into partials that
// cam_reduce folds (no contended atomics on the 16 hot addresses).
//
// geom_adam_kernel is the trainer's batched form: the same per-view chain,
// then the Adam update of all five parameter classes in the same pass (the
// reference's separate adam.wgsl dispatches), so the parameter gradients are
// never written to and re-read from HBM. It reads the raw parameters and
// activates them in registers (identity / exp, exactly as the Adam kernel
// writes the activated copies), and moves the 48 SH gradients of a block
// through LDS so the SH update reads and writes whole contiguous rows.

#include "gs_internal.h"

namespace gs {

namespace {

constexpr int kGeomThreads = 256;
constexpr int kShVec = kShFloats / 4;  // float4 per SH row
constexpr int kShRow = kShVec + 1;     // LDS row stride in float4 (one vector of padding)

// Stages the SH rows of the block's surfels [blockIdx.x * T, + T) in LDS
// (row stride kShRow float4) with lane-consecutive float4 loads; every
// view's colour -> direction chain then reads a surfel's 48 coefficients
// from LDS instead of with 48 scattered global loads. Ends with a barrier.
// (Worth it over several views: for the one-view launches of geom_bwd_kernel
// the staging cost more than the scattered loads, 0.54 -> 0.65 ms per
// step at 1M surfels, so that kernel reads SH from global memory.)
template <int T>
__device__ __forceinline__ void stage_sh_rows(const float* __restrict__ sh, uint32_t n, float4* s_sh) {
    const uint32_t base = blockIdx.x * T * kShVec;
    const uint32_t total = n * kShVec;
#pragma unroll 4
    for (int r = 0; r < kShVec; ++r) {
        const uint32_t e = (uint32_t)r * T + threadIdx.x;
        if (base + e >= total) break;
        const uint32_t j = e / kShVec, k = e - j * kShVec;
        s_sh[j * kShRow + k] = reinterpret_cast<const float4*>(sh)[base + e];
    }
    __syncthreads();
}

struct V3f {
    float x, y, z;
};
__device__ __forceinline__ V3f v3(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ V3f operator+(V3f a, V3f b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3f operator-(V3f a, V3f b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3f operator*(V3f a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3f crs(V3f a, V3f b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float dt(V3f a, V3f b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// quat_grad (math.rs:38-88 / surfel_prep_bwd.wgsl:75-112) for dL/dR with only
// columns 0 and 1 non-zero (the tangent axes).
__device__ __forceinline__ void quat_grad_cols01(float qx, float qy, float qz, float qw, float n,
                                                 V3f c0, V3f c1, float out[4]) {
    const float x = qx, y = qy, z = qz, w = qw;
    // <dL/dR, dR/dq_k> over columns 0 and 1.
    const float g0 = dt(c0, v3(0.f, 2.f * y, 2.f * z)) + dt(c1, v3(2.f * y, -4.f * x, 2.f * w));
    const float g1 = dt(c0, v3(-4.f * y, 2.f * x, -2.f * w)) + dt(c1, v3(2.f * x, 0.f, 2.f * z));
    const float g2 = dt(c0, v3(-4.f * z, 2.f * w, 2.f * x)) + dt(c1, v3(-2.f * w, -4.f * z, 2.f * y));
    const float g3 = dt(c0, v3(0.f, 2.f * z, -2.f * y)) + dt(c1, v3(-2.f * z, 0.f, 2.f * x));
    const float d = g0 * x + g1 * y + g2 * z + g3 * w;
    const float inv_n = 1.0f / n;
    out[0] = (g0 - x * d) * inv_n;
    out[1] = (g1 - y * d) * inv_n;
    out[2] = (g2 - z * d) * inv_n;
    out[3] = (g3 - w * d) * inv_n;
}

// Per-surfel gradient sums over the views of a batch.
struct SurfelAcc {
    float pos[3], sc[2], q[4], op;
    float sh[kShFloats];
};

__device__ __forceinline__ void zero_acc(SurfelAcc& a) {
#pragma unroll
    for (int k = 0; k < 3; ++k) a.pos[k] = 0.f;
    a.sc[0] = a.sc[1] = 0.f;
#pragma unroll
    for (int k = 0; k < 4; ++k) a.q[k] = 0.f;
    a.op = 0.f;
#pragma unroll
    for (int k = 0; k < kShFloats; ++k) a.sh[k] = 0.f;
}

// Loads one view's coefficient gradients of surfel i (GS slots stored per
// (view, surfel); the rest are 0); false when all are 0 (the surfel did not
// reach a pixel of the view).
template <int GS>
__device__ __forceinline__ bool load_coef_grads(const float* __restrict__ grad_coef, uint32_t v, uint32_t n,
                                                uint32_t i, float gc[kGradStride]) {
    const float4* src = reinterpret_cast<const float4*>(grad_coef + ((size_t)v * n + i) * GS);
#pragma unroll
    for (int k = GS; k < kGradStride; ++k) gc[k] = 0.0f;
#pragma unroll
    for (int k = 0; k < GS / 4; ++k) {
        const float4 t = src[k];
        gc[4 * k + 0] = t.x;
        gc[4 * k + 1] = t.y;
        gc[4 * k + 2] = t.z;
        gc[4 * k + 3] = t.w;
    }
    bool any = false;
#pragma unroll
    for (int k = 0; k < GS; ++k) any |= gc[k] != 0.0f;
    return any;
}

// Adds one view's contribution of one surfel (activated position p, scales s,
// quaternion q, SH row shp) to a; with want_cam also fills the view's 16
// camera gradient slots cg.
// VEC: the SH row is in LDS (row stride kShRow float4) and is read as
// float4; otherwise it is read from global memory one coefficient at a time
// (fewer registers in flight).
template <bool VEC>
__device__ __forceinline__ void add_view(const CamParams& cam, const float gc[kGradStride], float4 p, float2 s,
                                         float4 q, const float* __restrict__ shp, SurfelAcc& a, bool want_cam,
                                         float cg[16]) {
    const CamFrame f = cam_frame(cam, p, s, q);
    const V3f c = v3(f.cx_, f.cy_, f.cz_);
    const V3f tu = v3(f.tux, f.tuy, f.tuz);
    const V3f tv = v3(f.tvx, f.tvy, f.tvz);
    const V3f A = crs(tv, c), B = crs(c, tu), M = crs(tu, tv);
    const V3f gA = v3(gc[kGA], gc[kGA + 1], gc[kGA + 2]);
    const V3f gB = v3(gc[kGB], gc[kGB + 1], gc[kGB + 2]);
    const V3f gMs = v3(gc[kGM], gc[kGM + 1], gc[kGM + 2]);
    const float gNt = gc[kGNt];
    const float w = -c.z;
    const float fo = cam.focal;
    // A = tv x c, B = c x tu, M = tu x tv, Nt = -M.c (transposed rules).
    const V3f gM = gMs - c * gNt;
    V3f dc = crs(gA, tv) + crs(tu, gB) - M * gNt;
    V3f dtu = crs(gB, c) + crs(tv, gM);
    V3f dtv = crs(c, gA) + crs(gM, tu);
    // px = cx + f c.x / w, py = cy - f c.y / w, w = -c.z (backward.rs:329-336).
    const float gpx = gc[kGPx], gpy = gc[kGPy];
    dc.x += gpx * fo / w;
    dc.z += gpx * fo * c.x / (w * w);
    dc.y += -gpy * fo / w;
    dc.z += -gpy * fo * c.y / (w * w);
    dc.z -= gc[kGDep];  // screen-branch t = centre depth = -c.z
    const float dfocal = -(A.x * gA.x + A.y * gA.y + B.x * gB.x + B.y * gB.y + M.x * gMs.x + M.y * gMs.y) / fo +
                         gpx * c.x / w - gpy * c.y / w;
    // Normal chain (backward.rs:186-196).
    const V3f gn = v3(gc[kGNrm], gc[kGNrm + 1], gc[kGNrm + 2]);
    if (gn.x != 0.f || gn.y != 0.f || gn.z != 0.f) {
        const float len = sqrtf(dt(M, M));
        if (len > 1e-12f) {
            const V3f mh = M * (1.0f / len);
            const float sign = dt(mh, c) > 0.f ? -1.f : 1.f;
            const V3f dm = (gn - mh * dt(mh, gn)) * (sign / len);
            dtu = dtu + crs(tv, dm);
            dtv = dtv + crs(dm, tu);
        }
    }
    // c = R^T (p - C): dpos += R dc (backward.rs:198-203).
    const float* R = cam.rot;
    auto rmul = [&](V3f x) {
        return v3(R[0] * x.x + R[3] * x.y + R[6] * x.z, R[1] * x.x + R[4] * x.y + R[7] * x.z,
                  R[2] * x.x + R[5] * x.y + R[8] * x.z);
    };
    V3f dpos = rmul(dc);
    // tangent axes: tau = R^T (s * Rq e_axis) (backward.rs:206-219).
    const float x = f.qx, y = f.qy, z = f.qz, ww = f.qw;
    const V3f ax_u = v3(1.f - 2.f * (y * y + z * z), 2.f * (x * y + ww * z), 2.f * (x * z - ww * y));
    const V3f ax_v = v3(2.f * (x * y - ww * z), 1.f - 2.f * (x * x + z * z), 2.f * (y * z + ww * x));
    const V3f dy_u = rmul(dtu), dy_v = rmul(dtv);
    a.sc[0] += dt(dy_u, ax_u);
    a.sc[1] += dt(dy_v, ax_v);
    float dq[4];
    quat_grad_cols01(x, y, z, ww, f.qn, dy_u * s.x, dy_v * s.y, dq);
    a.q[0] += dq[0];
    a.q[1] += dq[1];
    a.q[2] += dq[2];
    a.q[3] += dq[3];
    // Colour -> SH (+ view direction). The clamp mask is already applied: the
    // rasterizer backward drops the colour gradient of a channel clamped at 0.
    const V3f dcol = v3(gc[kGCol], gc[kGCol + 1], gc[kGCol + 2]);
    V3f dcenter = v3(0.f, 0.f, 0.f) - dpos;
    if (dcol.x != 0.f || dcol.y != 0.f || dcol.z != 0.f) {
        const float len = sqrtf(f.xx * f.xx + f.xy * f.xy + f.xz * f.xz);
        const float il = 1.0f / len;
        const V3f dir = v3(f.xx * il, f.xy * il, f.xz * il);
        float basis[16], sdot[16];
        const int nb = sh_basis((int)cam.sh_degree, dir.x, dir.y, dir.z, basis);
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            sdot[k] = 0.f;
            if (k < nb) {
                a.sh[3 * k + 0] += dcol.x * basis[k];
                a.sh[3 * k + 1] += dcol.y * basis[k];
                a.sh[3 * k + 2] += dcol.z * basis[k];
                if (!VEC) sdot[k] = shp[3 * k] * dcol.x + shp[3 * k + 1] * dcol.y + shp[3 * k + 2] * dcol.z;
            }
        }
        if (VEC) {
            // sdot[k] = SH_k . dcol from the LDS row read as float4, one vector
            // at a time: at a row stride of 13 float4 these reads are free of
            // bank conflicts, where scalar reads would be 4-way.
            const float4* shp4 = reinterpret_cast<const float4*>(shp);
            const float dc[3] = {dcol.x, dcol.y, dcol.z};
#pragma unroll
            for (int c = 0; c < kShFloats / 4; ++c) {
                if (4 * c >= 3 * nb) break;
                const float4 v = shp4[c];
                const float f[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int m = 4 * c + j;
                    if (m < 3 * nb) sdot[m / 3] += f[j] * dc[m % 3];
                }
            }
        }
        const float3 dd = sh_dir_grad((int)cam.sh_degree, dir.x, dir.y, dir.z, sdot);
        const V3f ddir = v3(dd.x, dd.y, dd.z);
        const V3f dv = (ddir - dir * dt(dir, ddir)) * il;
        dpos = dpos + dv;
        dcenter = dcenter - dv;
    }
    a.pos[0] += dpos.x;
    a.pos[1] += dpos.y;
    a.pos[2] += dpos.z;
    a.op += gc[kGOp];
    if (want_cam) {
        // dL/dR_cam[j][k] += y_j da_k for a = R^T y (backward.rs:252-257).
        const V3f xw = v3(f.xx, f.xy, f.xz);
        const V3f yu = ax_u * s.x, yv = ax_v * s.y;
        const float das[3][3] = {{dc.x, dc.y, dc.z}, {dtu.x, dtu.y, dtu.z}, {dtv.x, dtv.y, dtv.z}};
        const V3f ys[3] = {xw, yu, yv};
#pragma unroll
        for (int kk = 0; kk < 3; ++kk)
#pragma unroll
            for (int src = 0; src < 3; ++src) {
                cg[3 * kk + 0] += ys[src].x * das[src][kk];
                cg[3 * kk + 1] += ys[src].y * das[src][kk];
                cg[3 * kk + 2] += ys[src].z * das[src][kk];
            }
        cg[9] = dcenter.x;
        cg[10] = dcenter.y;
        cg[11] = dcenter.z;
        cg[12] = dfocal;
    }
}

template <int GS>
__global__ __launch_bounds__(kGeomThreads) void geom_bwd_kernel(
    const CamParams* __restrict__ cams, uint32_t n_views, uint32_t n, const float4* __restrict__ pos,
    const float2* __restrict__ scales, const float4* __restrict__ quats, const float* __restrict__ sh,
    const float* __restrict__ grad_coef, ParamGrads g,
    int accumulate, float* __restrict__ cam_partials) {
    __shared__ float s_cam[kGeomThreads / 64][16];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const bool valid = i < n;

    float4 p = make_float4(0.f, 0.f, 0.f, 0.f), q = make_float4(0.f, 0.f, 0.f, 1.f);
    float2 s = make_float2(0.f, 0.f);
    if (valid) {
        p = pos[i];
        s = scales[i];
        q = quats[i];
    }
    SurfelAcc a;
    zero_acc(a);
    const bool want_cam = cam_partials != nullptr;

    for (uint32_t v = 0; v < n_views; ++v) {
        float cg[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) cg[k] = 0.f;
        float gc[kGradStride];
        const bool any = valid && load_coef_grads<GS>(grad_coef, v, n, i, gc);
        if (any) add_view<false>(cams[v], gc, p, s, q, sh + (size_t)i * kShFloats, a, want_cam, cg);
        if (want_cam) {
            // Block reduction of the 16 camera slots (every thread takes part).
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                float r = cg[k];
                for (int o = 32; o >= 1; o >>= 1) r += __shfl_xor(r, o);
                if (lane == 0) s_cam[wave][k] = r;
            }
            __syncthreads();
            if (threadIdx.x < 16) {
                float r = 0.f;
                for (int w2 = 0; w2 < kGeomThreads / 64; ++w2) r += s_cam[w2][threadIdx.x];
                cam_partials[((size_t)v * gridDim.x + blockIdx.x) * 16 + threadIdx.x] = r;
            }
            __syncthreads();
        }
    }
    if (!valid) return;
    float4 gp = make_float4(a.pos[0], a.pos[1], a.pos[2], 0.f);
    float2 gs = make_float2(a.sc[0], a.sc[1]);
    float4 gq = make_float4(a.q[0], a.q[1], a.q[2], a.q[3]);
    float go = a.op;
    float* gsh = g.sh + (size_t)i * kShFloats;
    if (accumulate) {
        const float4 op = g.pos[i];
        gp.x += op.x;
        gp.y += op.y;
        gp.z += op.z;
        const float2 os = g.scales[i];
        gs.x += os.x;
        gs.y += os.y;
        const float4 oq = g.quat[i];
        gq.x += oq.x;
        gq.y += oq.y;
        gq.z += oq.z;
        gq.w += oq.w;
        go += g.opacity[i];
#pragma unroll
        for (int k = 0; k < kShFloats; ++k) a.sh[k] += gsh[k];
    }
    g.pos[i] = gp;
    g.scales[i] = gs;
    g.quat[i] = gq;
    g.opacity[i] = go;
    float4* gsh4 = reinterpret_cast<float4*>(gsh);
#pragma unroll
    for (int k = 0; k < kShFloats / 4; ++k)
        gsh4[k] = make_float4(a.sh[4 * k], a.sh[4 * k + 1], a.sh[4 * k + 2], a.sh[4 * k + 3]);
}

constexpr int kFuseThreads = 128;

// The optimizer state and parameters written here are next read a whole
// step later: non-temporal stores keep them from displacing the step's
// working set in L2 / MALL.
typedef float f4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void store_nt(float4* p, float4 v) {
    __builtin_nontemporal_store(f4v{v.x, v.y, v.z, v.w}, reinterpret_cast<f4v*>(p));
}

// Adam over 4 consecutive floats of class c at float4 index j (raw given).
__device__ __forceinline__ void adam4(const AdamClass& c, float lr, float reg, float bc1, float bc2, uint32_t j,
                                      float4 raw, float4 g) {
    float4 m = reinterpret_cast<const float4*>(c.m)[j];
    float4 v = reinterpret_cast<const float4*>(c.v)[j];
    adam_element(c.activation, lr, reg, bc1, bc2, raw.x, g.x, m.x, v.x);
    adam_element(c.activation, lr, reg, bc1, bc2, raw.y, g.y, m.y, v.y);
    adam_element(c.activation, lr, reg, bc1, bc2, raw.z, g.z, m.z, v.z);
    adam_element(c.activation, lr, reg, bc1, bc2, raw.w, g.w, m.w, v.w);
    store_nt(reinterpret_cast<float4*>(c.raw) + j, raw);
    store_nt(reinterpret_cast<float4*>(c.m) + j, m);
    store_nt(reinterpret_cast<float4*>(c.v) + j, v);
    store_nt(reinterpret_cast<float4*>(c.activated) + j,
             make_float4(adam_activate(c.activation, raw.x), adam_activate(c.activation, raw.y),
                         adam_activate(c.activation, raw.z), adam_activate(c.activation, raw.w)));
}

// Table classes: 0 position (float4 rows), 1 scales (float2), 2 quaternion
// (float4), 3 opacity, 4 SH (48 floats); checked by the launcher.
__global__ __launch_bounds__(kFuseThreads, 3) void geom_adam_kernel(
    const CamParams* __restrict__ cams, uint32_t n_views, uint32_t n, const float* __restrict__ grad_coef,
    const AdamTable table, const AdamState* __restrict__ state, ParamGrads keep) {
    __shared__ float4 s_sh[kFuseThreads * kShRow];
    const uint32_t i = blockIdx.x * kFuseThreads + threadIdx.x;
    const bool valid = i < n;
    const AdamClass& cp = table.c[0];
    const AdamClass& cs = table.c[1];
    const AdamClass& cq = table.c[2];
    const AdamClass& co = table.c[3];
    const AdamClass& ch = table.c[4];

    float4 praw = make_float4(0.f, 0.f, 0.f, 0.f), qraw = make_float4(0.f, 0.f, 0.f, 1.f);
    float2 sraw = make_float2(0.f, 0.f);
    if (valid) {
        praw = reinterpret_cast<const float4*>(cp.raw)[i];
        sraw = reinterpret_cast<const float2*>(cs.raw)[i];
        qraw = reinterpret_cast<const float4*>(cq.raw)[i];
    }
    // Activated values, as the Adam kernel writes them.
    const float4 p = make_float4(adam_activate(cp.activation, praw.x), adam_activate(cp.activation, praw.y),
                                 adam_activate(cp.activation, praw.z), adam_activate(cp.activation, praw.w));
    const float2 s = make_float2(adam_activate(cs.activation, sraw.x), adam_activate(cs.activation, sraw.y));
    const float4 q = make_float4(adam_activate(cq.activation, qraw.x), adam_activate(cq.activation, qraw.y),
                                 adam_activate(cq.activation, qraw.z), adam_activate(cq.activation, qraw.w));
    stage_sh_rows<kFuseThreads>(ch.raw, n, s_sh);
    const uint32_t base = blockIdx.x * kFuseThreads * kShVec;
    const uint32_t total = n * kShVec;
    float4* row = s_sh + threadIdx.x * kShRow;
    SurfelAcc a;
    zero_acc(a);
    for (uint32_t v = 0; v < n_views; ++v) {
        float gc[kGradStride];
        const bool any = valid && load_coef_grads<grad_stride(false)>(grad_coef, v, n, i, gc);
        if (any) add_view<true>(cams[v], gc, p, s, q, reinterpret_cast<const float*>(row), a, false, nullptr);
    }
    const float bc1 = state->bc1_inv, bc2 = state->bc2_inv;
    const float* lr = state->lr;
    const float* reg = state->hyper.reg;
    if (valid) {
        const float4 gp = make_float4(a.pos[0], a.pos[1], a.pos[2], 0.f);
        const float4 gq = make_float4(a.q[0], a.q[1], a.q[2], a.q[3]);
        if (keep.pos != nullptr) {
            keep.pos[i] = gp;
            keep.scales[i] = make_float2(a.sc[0], a.sc[1]);
            keep.quat[i] = gq;
            keep.opacity[i] = a.op;
        }
        adam4(cp, lr[0], reg[0], bc1, bc2, i, praw, gp);
        adam4(cq, lr[2], reg[2], bc1, bc2, i, qraw, gq);
        {
            float2 m = reinterpret_cast<const float2*>(cs.m)[i];
            float2 vv = reinterpret_cast<const float2*>(cs.v)[i];
            adam_element(cs.activation, lr[1], reg[1], bc1, bc2, sraw.x, a.sc[0], m.x, vv.x);
            adam_element(cs.activation, lr[1], reg[1], bc1, bc2, sraw.y, a.sc[1], m.y, vv.y);
            reinterpret_cast<float2*>(cs.raw)[i] = sraw;
            reinterpret_cast<float2*>(cs.m)[i] = m;
            reinterpret_cast<float2*>(cs.v)[i] = vv;
            reinterpret_cast<float2*>(cs.activated)[i] =
                make_float2(adam_activate(cs.activation, sraw.x), adam_activate(cs.activation, sraw.y));
        }
        {
            float r = co.raw[i], m = co.m[i], vv = co.v[i];
            adam_element(co.activation, lr[3], reg[3], bc1, bc2, r, a.op, m, vv);
            co.raw[i] = r;
            co.m[i] = m;
            co.v[i] = vv;
            co.activated[i] = adam_activate(co.activation, r);
        }
    }
    // SH: gradients through LDS (each thread overwrites only its own row,
    // which no other thread read), then the block's rows updated as one
    // contiguous range (lane-consecutive float4).
#pragma unroll
    for (int k = 0; k < kShVec; ++k) row[k] = make_float4(a.sh[4 * k], a.sh[4 * k + 1], a.sh[4 * k + 2], a.sh[4 * k + 3]);
    __syncthreads();
#pragma unroll 4
    for (int r = 0; r < kShVec; ++r) {
        const uint32_t e = (uint32_t)r * kFuseThreads + threadIdx.x;
        if (base + e >= total) break;
        const uint32_t j = e / kShVec, k = e - j * kShVec;
        const float4 g4 = s_sh[j * kShRow + k];
        if (keep.sh != nullptr) reinterpret_cast<float4*>(keep.sh)[base + e] = g4;
        adam4(ch, lr[4], reg[4], bc1, bc2, base + e, reinterpret_cast<const float4*>(ch.raw)[base + e], g4);
    }
}

// Folds per-block camera partials into 16 floats per view (accumulating into
// g_cam when accumulate != 0).
__global__ __launch_bounds__(256) void cam_reduce_kernel(const float* __restrict__ partials,
                                                         uint32_t nblocks, float* __restrict__ g_cam,
                                                         int accumulate) {
    __shared__ float s[256 / 64][16];
    const uint32_t v = blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    float acc[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) acc[k] = 0.f;
    for (uint32_t b = threadIdx.x; b < nblocks; b += blockDim.x) {
        const float* src = partials + ((size_t)v * nblocks + b) * 16;
#pragma unroll
        for (int k = 0; k < 16; ++k) acc[k] += src[k];
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        float r = acc[k];
        for (int o = 32; o >= 1; o >>= 1) r += __shfl_xor(r, o);
        if (lane == 0) s[wave][k] = r;
    }
    __syncthreads();
    if (threadIdx.x < 16) {
        float r = 0.f;
        for (int w = 0; w < 4; ++w) r += s[w][threadIdx.x];
        float* dst = g_cam + (size_t)v * 16 + threadIdx.x;
        *dst = accumulate ? *dst + r : r;
    }
}

}  // namespace

hipError_t launch_geom_adam(hipStream_t st, const CamParams* cams_dev, uint32_t n_views, uint32_t n,
                            const float* grad_coef, const AdamTable& table,
                            const AdamState* state, ParamGrads keep) {
    if (n == 0) return hipSuccess;
    // The kernel's fixed class layout and row shapes.
    const uint32_t rows[5] = {4, 2, 4, 1, (uint32_t)kShFloats};
    const int acts[5] = {0, 1, 0, 2, 0};
    if (table.n_classes != 5 || (uint64_t)n * kShFloats >= (1ull << 32)) return hipErrorInvalidValue;
    for (int k = 0; k < 5; ++k) {
        const AdamClass& c = table.c[k];
        if (c.n != n * rows[k] || c.activation != acts[k]) return hipErrorInvalidValue;
        if (((uintptr_t)c.raw | (uintptr_t)c.m | (uintptr_t)c.v | (uintptr_t)c.activated) & 15u)
            return hipErrorInvalidValue;
    }
    if (keep.pos != nullptr && (((uintptr_t)keep.sh & 15u) || !keep.scales || !keep.quat || !keep.opacity))
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(geom_adam_kernel, dim3(div_up(n, kFuseThreads)), dim3(kFuseThreads), 0, st, cams_dev,
                       n_views, n, grad_coef, table, state, keep);
    return hipGetLastError();
}

size_t geom_cam_partials_bytes(uint32_t n, uint32_t n_views) {
    return (size_t)div_up(n ? n : 1, kGeomThreads) * n_views * 16 * sizeof(float);
}

hipError_t launch_geom_bwd(hipStream_t st, const CamParams* cams_dev, uint32_t n_views, uint32_t n,
                           const float4* pos, const float2* scales, const float4* quats,
                           const float* sh, const float* grad_coef, bool aux,
                           ParamGrads g, bool accumulate, float* cam_partials) {
    if (n_views == 0) return hipSuccess;
    if (n == 0) {
        // No surfel: a pass that does not accumulate still sets the camera
        // gradients to 0 (the reduction over no partials reads nothing).
        if (g.cam == nullptr || accumulate) return hipSuccess;
        hipLaunchKernelGGL(cam_reduce_kernel, dim3(n_views), dim3(256), 0, st, cam_partials, 0u, g.cam, 0);
        return hipGetLastError();
    }
    const uint32_t blocks = div_up(n, kGeomThreads);
    float* parts = g.cam ? cam_partials : nullptr;
    if (aux)
        hipLaunchKernelGGL(geom_bwd_kernel<grad_stride(true)>, dim3(blocks), dim3(kGeomThreads), 0, st, cams_dev,
                           n_views, n, pos, scales, quats, sh, grad_coef, g, accumulate ? 1 : 0, parts);
    else
        hipLaunchKernelGGL(geom_bwd_kernel<grad_stride(false)>, dim3(blocks), dim3(kGeomThreads), 0, st, cams_dev,
                           n_views, n, pos, scales, quats, sh, grad_coef, g, accumulate ? 1 : 0, parts);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !parts) return e;
    hipLaunchKernelGGL(cam_reduce_kernel, dim3(n_views), dim3(256), 0, st, parts, blocks, g.cam,
                       accumulate ? 1 : 0);
    return hipGetLastError();
}

}  // namespace gs
