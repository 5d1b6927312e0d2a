// Photometric camera tracking against a frozen map, iterated on the device.
//
// Restates crates/gs-track/src/shaders/photometric.wgsl:47-73 and the pose
// descent of crates/gs-track/src/lib.rs:173-337 (Tracker::track, step,
// camera_step_grad, adam6, apply_camera_step).
// The reference reads the residual and the camera gradient back to the host
// after every iteration and runs the 6-parameter Adam there. Here the
// iteration is forward -> photometric -> backward -> track_update, all on the
// stream: track_update reduces the photometric partials, keeps the residual
// trace and the best pose, applies the Riemannian Adam step, and writes the
// next pose into the device camera table the next forward reads. The host
// only enqueues iterations (and looks at the stop flag every few).
//
// photometric: per pixel, target' = gain target + bias; d = render - target';
// dL/d(render) = clamp(d / 1e-4, -1, 1) / (3 P) (a softened sign), w = 0;
// the residual sum |d| / (3 P) and the coverage (render alpha > 0.5) / P are
// summed per block (the reference adds them with f32 atomics; fixed-order
// partials here).

#include "gs_internal.h"

namespace gs {

namespace {

constexpr int kPhotoBlock = 256;

__global__ __launch_bounds__(kPhotoBlock) void photometric_kernel(uint32_t npix, float gain, float bias,
                                                                  const float4* __restrict__ render,
                                                                  const float4* __restrict__ frame,
                                                                  float4* __restrict__ dl_dcolor,
                                                                  float2* __restrict__ partials) {
    __shared__ float s_err[kPhotoBlock / 64], s_cov[kPhotoBlock / 64];
    const uint32_t i = blockIdx.x * kPhotoBlock + threadIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    float err = 0.0f, cov = 0.0f;
    if (i < npix) {
        const float4 r = render[i];
        const float4 t = frame[i];
        const float inv3p = 1.0f / (3.0f * (float)npix);
        const float rc[3] = {r.x, r.y, r.z}, tc[3] = {t.x, t.y, t.z};
        float g[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float d = rc[c] - (gain * tc[c] + bias);
            err += fabsf(d);
            g[c] = fminf(fmaxf(d / 1e-4f, -1.0f), 1.0f) * inv3p;
        }
        dl_dcolor[i] = make_float4(g[0], g[1], g[2], 0.0f);
        cov = r.w > 0.5f ? 1.0f : 0.0f;
    }
    for (int o = 32; o >= 1; o >>= 1) {
        err += __shfl_xor(err, o);
        cov += __shfl_xor(cov, o);
    }
    if (lane == 0) {
        s_err[wave] = err;
        s_cov[wave] = cov;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float e = 0.0f, c = 0.0f;
#pragma unroll
        for (int w = 0; w < kPhotoBlock / 64; ++w) {
            e += s_err[w];
            c += s_cov[w];
        }
        partials[blockIdx.x] = make_float2(e, c);
    }
}

// glam Mat3::from_quat (f32): columns of R for a unit quaternion (x, y, z, w).
__device__ void mat3_from_quat(const float q[4], float col[3][3]) {
    const float x = q[0], y = q[1], z = q[2], w = q[3];
    const float x2 = x + x, y2 = y + y, z2 = z + z;
    const float xx = x * x2, xy = x * y2, xz = x * z2;
    const float yy = y * y2, yz = y * z2, zz = z * z2;
    const float wx = w * x2, wy = w * y2, wz = w * z2;
    col[0][0] = 1.0f - (yy + zz);
    col[0][1] = xy + wz;
    col[0][2] = xz - wy;
    col[1][0] = xy - wz;
    col[1][1] = 1.0f - (xx + zz);
    col[1][2] = yz + wx;
    col[2][0] = xz + wy;
    col[2][1] = yz - wx;
    col[2][2] = 1.0f - (xx + yy);
}

// Camera-to-world rotation written into the camera table, as the host's
// make_cam: evaluated in double from the f32 quaternion, rounded once.
__device__ void write_cam_rot(const float q[4], float* rot) {
    const double n = sqrt((double)q[0] * q[0] + (double)q[1] * q[1] + (double)q[2] * q[2] + (double)q[3] * q[3]);
    const double x = q[0] / n, y = q[1] / n, z = q[2] / n, w = q[3] / n;
    const double c[3][3] = {{1 - 2 * (y * y + z * z), 2 * (x * y + w * z), 2 * (x * z - w * y)},
                            {2 * (x * y - w * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w * x)},
                            {2 * (x * z + w * y), 2 * (y * z - w * x), 1 - 2 * (x * x + y * y)}};
    for (int k = 0; k < 3; ++k)
        for (int j = 0; j < 3; ++j) rot[3 * k + j] = (float)c[k][j];
}

__global__ __launch_bounds__(64) void track_update_kernel(TrackArgs a, uint32_t t) {
    TrackState* s = a.state;
    // Fixed-order reduction of the photometric partials (double).
    double e = 0.0, c = 0.0;
    for (uint32_t k = threadIdx.x; k < a.n_partials; k += 64) {
        const float2 p = a.partials[k];
        e += p.x;
        c += p.y;
    }
    for (int o = 32; o >= 1; o >>= 1) {
        e += __shfl_xor(e, o);
        c += __shfl_xor(c, o);
    }
    if (threadIdx.x != 0 || s->stopped) return;
    const float residual = (float)(e / (3.0 * (double)a.npix));
    const float coverage = (float)(c / (double)a.npix);
    a.residuals[t - 1] = residual;
    s->iters = t;
    s->coverage = coverage;
    // lib.rs:188-207: keep the best pose; stop after `patience` stale steps.
    auto keep_best = [&]() {
        s->best = residual;
        for (int k = 0; k < 4; ++k) s->best_q[k] = s->cur_q[k];
        for (int k = 0; k < 3; ++k) s->best_c[k] = s->cur_c[k];
    };
    if (residual < s->best * (1.0f - a.min_improvement)) {
        keep_best();
        s->stale = 0;
    } else {
        if (residual < s->best) keep_best();
        s->stale += 1;
    }
    if (s->stale >= a.patience || t >= a.iterations) {
        s->stopped = 1;
        return;
    }
    // lib.rs:276-297: Riemannian gradient, antisymmetric part of R^T dL/dR.
    const float* g = a.grad_cam;
    for (int k = 0; k < 12; ++k)
        if (!isfinite(g[k])) {
            s->stopped = 1;  // lib.rs:209-215: non-finite gradient, keep the best pose
            return;
        }
    float r[3][3];
    mat3_from_quat(s->cur_q, r);
    // b = R^T A with A's columns g[0..2], g[3..5], g[6..8]: b(row i, col j) = r_i . A_j.
    auto b = [&](int i, int j) {
        return r[i][0] * g[3 * j + 0] + r[i][1] * g[3 * j + 1] + r[i][2] * g[3 * j + 2];
    };
    const float grad[6] = {b(2, 1) - b(1, 2), b(0, 2) - b(2, 0), b(1, 0) - b(0, 1), g[9], g[10], g[11]};
    // lib.rs:300-319: Adam over the six, separate rates.
    const float bc1 = 1.0f - powf(0.9f, (float)t);
    const float bc2 = 1.0f - powf(0.999f, (float)t);
    float step[6];
    for (int k = 0; k < 6; ++k) {
        s->m[k] = 0.9f * s->m[k] + 0.1f * grad[k];
        s->v[k] = 0.999f * s->v[k] + 0.001f * grad[k] * grad[k];
        const float lr = k < 3 ? a.lr_rot : a.lr_center;
        step[k] = lr * (s->m[k] / bc1) / (sqrtf(s->v[k] / bc2) + 1e-10f);
    }
    // lib.rs:321-329: right perturbation R' = R exp([-step]x), centre -= step.
    const float dx = -step[0], dy = -step[1], dz = -step[2];
    const float ang = sqrtf(dx * dx + dy * dy + dz * dz);
    float e4[4] = {0.0f, 0.0f, 0.0f, 1.0f};
    if (ang > 0.0f) {
        const float sh = sinf(ang * 0.5f), ch = cosf(ang * 0.5f);
        e4[0] = dx / ang * sh;
        e4[1] = dy / ang * sh;
        e4[2] = dz / ang * sh;
        e4[3] = ch;
    }
    const float* q = s->cur_q;
    float p[4];
    p[0] = q[3] * e4[0] + q[0] * e4[3] + q[1] * e4[2] - q[2] * e4[1];
    p[1] = q[3] * e4[1] - q[0] * e4[2] + q[1] * e4[3] + q[2] * e4[0];
    p[2] = q[3] * e4[2] + q[0] * e4[1] - q[1] * e4[0] + q[2] * e4[3];
    p[3] = q[3] * e4[3] - q[0] * e4[0] - q[1] * e4[1] - q[2] * e4[2];
    const float pn = sqrtf(p[0] * p[0] + p[1] * p[1] + p[2] * p[2] + p[3] * p[3]);
    for (int k = 0; k < 4; ++k) s->cur_q[k] = p[k] / pn;
    for (int k = 0; k < 3; ++k) s->cur_c[k] -= step[3 + k];
    // The next forward reads the new pose from the camera table.
    write_cam_rot(s->cur_q, a.cam->rot);
    for (int k = 0; k < 3; ++k) a.cam->center[k] = s->cur_c[k];
}

}  // namespace

uint32_t photometric_partials(uint32_t npix) { return (npix + kPhotoBlock - 1) / kPhotoBlock; }

hipError_t launch_photometric(hipStream_t st, uint32_t npix, float gain, float bias, const float4* render,
                              const float4* frame, float4* dl_dcolor, float2* partials) {
    if (npix == 0) return hipSuccess;
    hipLaunchKernelGGL(photometric_kernel, dim3(photometric_partials(npix)), dim3(kPhotoBlock), 0, st, npix, gain,
                       bias, render, frame, dl_dcolor, partials);
    return hipGetLastError();
}

hipError_t launch_track_update(hipStream_t st, const TrackArgs& a, uint32_t t) {
    hipLaunchKernelGGL(track_update_kernel, dim3(1), dim3(64), 0, st, a, t);
    return hipGetLastError();
}

}  // namespace gs
