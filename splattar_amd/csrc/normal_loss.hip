// Normal-consistency loss (2DGS) and its gradient, one LDS-tiled kernel.
//
// Restates first_attempt/crates/gs-train/src/shaders/normal_loss.wgsl:52-145
// (NormalLoss, normal_loss.rs): per pixel q with mean depth
// tbar = depth_acc / alpha (0 when alpha < alpha_min), P(q) = tbar * ray(q);
// the depth normal N_d = +-normalize(dPx x dPy) from forward differences,
// oriented toward the camera; L = lambda * sum_q alpha_q (1 - nhat . N_d),
// with alpha and the orientation sign detached. Outputs: dL/d(normal image)
// into dl_dnormal, dL/d(depth_acc) into dl_dcolor.w (rgb untouched; .w is
// zeroed where no stencil reaches, as the reference's pass2), the per-pixel
// loss map (x = alpha_q (1 - nhat . N_d), unweighted) and per-tile partial
// sums of the loss map.
//
// The reference runs two dispatches and round-trips the stencil
// coefficients cpx = dL/d(dPx), cpy = dL/d(dPy) through two P-sized vec4
// buffers. Here one workgroup owns a 16x16 tile: it stages tbar and alpha of
// the 18x18 region the tile's stencils and their left / upper neighbours
// read, evaluates the 17x17 stencils (the tile plus the column to its left
// and the row above, whose coefficients the adjoint gathers) into LDS, and
// applies the adjoint of the forward differences from LDS. The halo stencils
// are evaluated twice (by this tile and their owner) instead of stored.

#include "gs_internal.h"

namespace gs {

namespace {

constexpr int kNT = 16;            // output tile
constexpr int kNR = kNT + 2;       // staged region: x0 - 1 .. x0 + 16
constexpr int kNS = kNT + 1;       // stencil region: x0 - 1 .. x0 + 15

__global__ __launch_bounds__(256) void normal_loss_kernel(NormalLossArgs a) {
    __shared__ float s_t[kNR][kNR];    // tbar
    __shared__ float s_al[kNR][kNR];   // alpha
    __shared__ float3 s_cx[kNS][kNS];  // cpx
    __shared__ float3 s_cy[kNS][kNS];  // cpy
    __shared__ float s_red[4];
    const int tid = threadIdx.x;
    const uint32_t tiles_x = (a.width + kNT - 1) / kNT;
    const uint32_t tile = xcd_swizzle(blockIdx.x, gridDim.x);
    const int x0 = (int)(tile % tiles_x) * kNT, y0 = (int)(tile / tiles_x) * kNT;
    const int W = (int)a.width, H = (int)a.height;
    const float cx = (float)a.width * 0.5f, cy = (float)a.height * 0.5f;
    auto ray = [&](int x, int y) {
        return make_float3(((float)x + 0.5f - cx) / a.focal, -((float)y + 0.5f - cy) / a.focal, -1.0f);
    };
    // Stage tbar and alpha of the region (pixels outside the image: alpha 0).
    for (int k = tid; k < kNR * kNR; k += 256) {
        const int rx = k % kNR, ry = k / kNR;
        const int x = x0 - 1 + rx, y = y0 - 1 + ry;
        float al = 0.0f, tb = 0.0f;
        if (x >= 0 && y >= 0 && x < W && y < H) {
            const size_t i = (size_t)y * W + x;
            al = a.color[i].w;
            if (al >= a.alpha_min) tb = a.aux[i].y / al;  // normal_loss.wgsl:47-53
        }
        s_t[ry][rx] = tb;
        s_al[ry][rx] = al;
    }
    __syncthreads();
    // Stencils of the 17x17 region (normal_loss.wgsl:56-118).
    float lsum = 0.0f;
    for (int k = tid; k < kNS * kNS; k += 256) {
        const int sx = k % kNS, sy = k / kNS;  // region coordinates: pixel (x0 - 1 + sx, y0 - 1 + sy)
        const int x = x0 - 1 + sx, y = y0 - 1 + sy;
        const bool own = sx >= 1 && sy >= 1;  // a pixel of this tile
        float3 cpx = make_float3(0.0f, 0.0f, 0.0f), cpy = cpx, dln = cpx;
        float lval = 0.0f;
        const bool valid = a.lambda != 0.0f && x >= 0 && y >= 0 && x + 1 < W && y + 1 < H &&
                           s_al[sy][sx] >= a.alpha_min && s_al[sy][sx + 1] >= a.alpha_min &&
                           s_al[sy + 1][sx] >= a.alpha_min;
        if (valid) {
            const float a_q = s_al[sy][sx];
            const float3 r0 = ray(x, y), rx1 = ray(x + 1, y), ry1 = ray(x, y + 1);
            const float t0 = s_t[sy][sx], tx1 = s_t[sy][sx + 1], ty1 = s_t[sy + 1][sx];
            const float3 p = make_float3(t0 * r0.x, t0 * r0.y, t0 * r0.z);
            const float3 dpx = make_float3(tx1 * rx1.x - p.x, tx1 * rx1.y - p.y, tx1 * rx1.z - p.z);
            const float3 dpy = make_float3(ty1 * ry1.x - p.x, ty1 * ry1.y - p.y, ty1 * ry1.z - p.z);
            const float3 m = cross3(dpx, dpy);
            const float len2 = dot3(m, m);
            const size_t i = (size_t)y * W + x;
            const float4 nv = a.normal[i];
            const float3 av = make_float3(nv.x, nv.y, nv.z);
            const float la2 = dot3(av, av);
            if (len2 >= 1e-18f && la2 >= 1e-18f) {
                const float len = sqrtf(len2);
                const float3 mh = make_float3(m.x / len, m.y / len, m.z / len);
                const float sign = dot3(mh, p) > 0.0f ? -1.0f : 1.0f;
                const float3 nd = make_float3(sign * mh.x, sign * mh.y, sign * mh.z);
                const float la = sqrtf(la2);
                const float3 nh = make_float3(av.x / la, av.y / la, av.z / la);
                lval = (1.0f - dot3(nh, nd)) * a_q;
                // Gradients (alpha weight and orientation sign detached).
                const float wq = -a.lambda * a_q;
                const float3 dl_dnd = make_float3(wq * nh.x, wq * nh.y, wq * nh.z);
                const float3 dl_dnh = make_float3(wq * nd.x, wq * nd.y, wq * nd.z);
                const float pn = dot3(nh, dl_dnh);
                dln = make_float3((dl_dnh.x - nh.x * pn) / la, (dl_dnh.y - nh.y * pn) / la,
                                  (dl_dnh.z - nh.z * pn) / la);
                const float pm = dot3(mh, dl_dnd);
                const float3 dl_dm = make_float3(sign * (dl_dnd.x - mh.x * pm) / len,
                                                 sign * (dl_dnd.y - mh.y * pm) / len,
                                                 sign * (dl_dnd.z - mh.z * pm) / len);
                cpx = cross3(dpy, dl_dm);
                cpy = cross3(dl_dm, dpx);
            }
        }
        s_cx[sy][sx] = cpx;
        s_cy[sy][sx] = cpy;
        if (own && x < W && y < H) {
            const size_t i = (size_t)y * W + x;
            a.dl_dnormal[i] = make_float4(dln.x, dln.y, dln.z, 0.0f);
            if (a.loss_map) a.loss_map[i] = make_float4(lval, 0.0f, 0.0f, 0.0f);
            lsum += lval;
        }
    }
    __syncthreads();
    // Adjoint of the forward differences (normal_loss.wgsl:121-145): the
    // centre contributes -cpx - cpy, the left neighbour's cpx and the upper
    // neighbour's cpy contribute +; dL/d(depth_acc) = ray . acc / alpha.
    {
        const int tx = tid & 15, ty = tid >> 4;
        const int x = x0 + tx, y = y0 + ty;
        if (x < W && y < H) {
            float w = 0.0f;
            if (a.lambda != 0.0f) {
                const int sx = tx + 1, sy = ty + 1;
                float3 acc = make_float3(-s_cx[sy][sx].x - s_cy[sy][sx].x, -s_cx[sy][sx].y - s_cy[sy][sx].y,
                                         -s_cx[sy][sx].z - s_cy[sy][sx].z);
                if (x > 0) {
                    acc.x += s_cx[sy][sx - 1].x;
                    acc.y += s_cx[sy][sx - 1].y;
                    acc.z += s_cx[sy][sx - 1].z;
                }
                if (y > 0) {
                    acc.x += s_cy[sy - 1][sx].x;
                    acc.y += s_cy[sy - 1][sx].y;
                    acc.z += s_cy[sy - 1][sx].z;
                }
                const float a_q = s_al[sy][sx];
                if (a_q >= a.alpha_min) w = dot3(ray(x, y), acc) / a_q;
            }
            a.dl_dcolor_w[4 * ((size_t)y * W + x)] = w;
        }
    }
    // Per-tile loss-map sum, in a fixed order (deterministic).
    for (int o = 32; o >= 1; o >>= 1) lsum += __shfl_xor(lsum, o);
    if ((tid & 63) == 0) s_red[tid >> 6] = lsum;
    __syncthreads();
    if (tid == 0) a.partials[tile] = (s_red[0] + s_red[1]) + (s_red[2] + s_red[3]);
}

}  // namespace

hipError_t launch_normal_loss(hipStream_t st, const NormalLossArgs& a) {
    const uint32_t tiles = ((a.width + kNT - 1) / kNT) * ((a.height + kNT - 1) / kNT);
    if (tiles == 0) return hipSuccess;
    hipLaunchKernelGGL(normal_loss_kernel, dim3(tiles), dim3(256), 0, st, a);
    return hipGetLastError();
}

uint32_t normal_loss_partials(uint32_t width, uint32_t height) {
    return ((width + kNT - 1) / kNT) * ((height + kNT - 1) / kNT);
}

}  // namespace gs
