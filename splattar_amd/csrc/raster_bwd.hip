// Tile rasterizer backward (A10): the hand-written adjoint of raster_fwd.
//
// Restates crates/gs-kernels/src/shaders/rasterize_bwd.wgsl:113-312 (the
// oracle's backward.rs:67-173): per pixel, walk the tile list back to front
// from the stored final transmittance and last-contributor index, recover
// T_before = T / (1 - alpha) instead of storing per-fragment state, keep
// suffix sums, and differentiate the compositing and the ray-splat
// intersection. Differences from the reference, all MI355X-driven:
//   * gradients are taken w.r.t. the per-surfel record coefficients (A, B, M,
//     Nt, px, py, depth, colour, opacity, normal); the linear chain to
//     camera-space (c, tu, tv) happens once per surfel in geom_bwd instead of
//     four triple-product gradients per pixel-hit;
//   * per record, each wave64 reduces its 64 pixels' contributions with a
//     transposed butterfly (17 lane exchanges for 16 values instead of 96,
//     on permlane swaps and DPP rather than LDS permutes), adds them into an
//     LDS slot (ds_add_f32), and the batch's slots are
//     flushed with hardware global FP32 atomics (no CAS loops).
// AUX = true adds the depth / normal image losses and the in-walk depth
// distortion loss (backward.rs:126-153).

#include "gs_internal.h"

namespace gs {

namespace {

constexpr int kBatch = 256;

__device__ __forceinline__ uint32_t view_of_tile(const CamParams* cams, uint32_t n_views, uint32_t t) {
    uint32_t v = 0;
    for (uint32_t k = 1; k < n_views; ++k)
        if (cams[k].tile_base <= t) v = k;
    return v;
}

// Cross-lane exchange without the LDS crossbar: DPP for partners inside a
// 16-lane row, the gfx950 permlane swaps across rows. xlane<OFF>(x) returns x
// of a partner lane whose index differs from this lane's in bit OFF (and
// possibly in lower bits): row_mirror (i <-> 15-i) for 8, row_half_mirror
// (i <-> 7-i) for 4, quad permutations (xor) for 2 and 1.
template <int OFF>
__device__ __forceinline__ float xlane(float x) {
    static_assert(OFF == 8 || OFF == 4 || OFF == 2 || OFF == 1, "row-local exchange only");
    constexpr int ctrl = OFF == 8 ? 0x140 : OFF == 4 ? 0x141 : OFF == 2 ? 0x4E : 0xB1;
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), ctrl, 0xF, 0xF, false));
}

// One transposition step at lane bit OFF: lanes with the bit clear end up
// with lo(self) + lo(partner), lanes with it set with hi(self) + hi(partner).
// For OFF = 32 / 16 one v_permlane{32,16}_swap moves the upper half of lo
// and the lower half of hi across in a single VALU op.
template <int OFF>
__device__ __forceinline__ float pair_sum(float lo, float hi, int lane) {
    if constexpr (OFF == 32) {
        const auto p = __builtin_amdgcn_permlane32_swap(__float_as_uint(lo), __float_as_uint(hi), false, false);
        return __uint_as_float(p[0]) + __uint_as_float(p[1]);
    } else if constexpr (OFF == 16) {
        const auto p = __builtin_amdgcn_permlane16_swap(__float_as_uint(lo), __float_as_uint(hi), false, false);
        return __uint_as_float(p[0]) + __uint_as_float(p[1]);
    } else {
        const bool upper = (lane & OFF) != 0;
        const float send = upper ? lo : hi;
        const float keep = upper ? hi : lo;
        return keep + xlane<OFF>(send);
    }
}

template <int OFF>
__device__ __forceinline__ float finish_sum(float r) {
    if constexpr (OFF >= 1) return finish_sum<OFF / 2>(r + xlane<OFF>(r));
    else return r;
}

template <int COUNT, int OFF, int NV>
__device__ __forceinline__ float transpose_reduce(float (&v)[NV], int lane) {
    if constexpr (COUNT > 1) {
#pragma unroll
        for (int i = 0; i < COUNT / 2; ++i) v[i] = pair_sum<OFF>(v[i], v[COUNT / 2 + i], lane);
        return transpose_reduce<COUNT / 2, OFF / 2, NV>(v, lane);
    } else {
        return finish_sum<OFF>(v[0]);
    }
}

// Sum NV values (power of two, 4 <= NV <= 16) over the 64 lanes of a wave:
// NV - 1 exchanges halve the value count per step, then the remaining lane
// bits are summed. On return lane l holds the total of value index
// (l >> log2(64/NV)) & (NV-1); lanes with (l & (64/NV - 1)) == 0 are the
// designated writers.
template <int NV>
__device__ __forceinline__ float wave_reduce_transpose(float (&v)[NV], int lane) {
    static_assert(NV >= 4 && NV <= 16 && (NV & (NV - 1)) == 0, "4, 8 or 16 values");
    return transpose_reduce<NV, 32, NV>(v, lane);
}

template <bool AUX>
__global__ __launch_bounds__(256) void raster_bwd_kernel(const CamParams* __restrict__ cams,
                                                         uint32_t n_views,
                                                         const uint32_t* __restrict__ ranges,
                                                         const uint32_t* __restrict__ payloads,
                                                         const Record* __restrict__ recs, BwdIn in,
                                                         float* __restrict__ grad_coef, bool cull,
                                                         uint32_t total_tiles, TileQueue queue) {
    constexpr int NS = AUX ? 20 : 16;  // LDS slots per record
    __shared__ float4 s_m[kBatch], s_a[kBatch], s_b[kBatch], s_c[kBatch], s_n[kBatch];
    __shared__ float s_lthr[kBatch];
    __shared__ uint16_t s_list[4][kBatch];
    __shared__ uint32_t s_qcnt[4][4];
    __shared__ float s_grad[kBatch * NS];
    __shared__ uint32_t s_view, s_maxlast, s_next;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (uint32_t iter = 0;; ++iter) {
        uint32_t pos;
        if (queue.cursor != nullptr) {
            if (tid == 0) {
                const uint32_t q = min(kTileQueues, gridDim.x), x = blockIdx.x % q;
                s_next = atomicAdd(queue.cursor + x, 1u) * q + x;
            }
            __syncthreads();
            pos = s_next;
            __syncthreads();  // every thread has pos before s_next is rewritten
        } else {
            pos = iter == 0 ? xcd_swizzle(blockIdx.x, gridDim.x) : total_tiles;
        }
        if (pos >= total_tiles) break;
        const uint32_t gtile = queue.order != nullptr ? queue.order[pos] : pos;
        const uint32_t start = ranges[2 * (size_t)gtile], end = ranges[2 * (size_t)gtile + 1];
        if (end <= start) continue;
        if (tid == 0) {
            s_view = view_of_tile(cams, n_views, gtile);
            s_maxlast = 0u;
        }
        __syncthreads();
        const CamParams& cam = cams[s_view];
        const uint32_t local = gtile - cam.tile_base;
        const uint32_t tx = local % cam.tiles_x, ty = local / cam.tiles_x;
        const int px = (int)(tx * kTile) + (wave & 1) * 8 + (lane & 7);
        const int py = (int)(ty * kTile) + (wave >> 1) * 8 + (lane >> 3);
        const bool inside = px < (int)cam.width && py < (int)cam.height;
        float sx, sy;
        pixel_s(cam, px, py, sx, sy);
        const float pxc = (float)px + 0.5f, pyc = (float)py + 0.5f;
        const size_t pi = (size_t)cam.pix_base + (size_t)py * cam.width + (size_t)px;

        float dlr = 0.f, dlg = 0.f, dlb = 0.f, dld = 0.f, dnx = 0.f, dny = 0.f, dnz = 0.f;
        float T = 1.0f, w_total = 0.f, wt_total = 0.f;
        uint32_t last = 0;
        if (inside) {
            const float4 dl = in.dl_dcolor[pi];
            dlr = dl.x;
            dlg = dl.y;
            dlb = dl.z;
            const float4 aux = in.out_aux[pi];
            T = aux.x;
            last = in.ncontrib[pi];
            if (AUX) {
                dld = dl.w;
                if (in.dl_dnormal) {
                    const float4 dn = in.dl_dnormal[pi];
                    dnx = dn.x;
                    dny = dn.y;
                    dnz = dn.z;
                }
                w_total = in.out_color[pi].w;
                wt_total = aux.y;
            }
        }
        const float lambda = AUX ? cam.lambda_dist : 0.0f;
        bool has_loss = inside && (dlr != 0.f || dlg != 0.f || dlb != 0.f);
        if (AUX) has_loss = inside && (has_loss || dld != 0.f || dnx != 0.f || dny != 0.f ||
                                       dnz != 0.f || lambda != 0.f);
        if (!has_loss) last = 0;
        if (last) atomicMax(&s_maxlast, last);
        __syncthreads();
        const uint32_t maxlast = s_maxlast;
        if (maxlast == 0) continue;

        float suf_r = 0.f, suf_g = 0.f, suf_b = 0.f;
        float suf_d = 0.f, suf_nx = 0.f, suf_ny = 0.f, suf_nz = 0.f, suf_w = 0.f, suf_cw = 0.f;

        const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
        const int nbatch = (int)((maxlast + kBatch - 1) / kBatch);
        for (int b = nbatch - 1; b >= 0; --b) {
            const uint32_t base = start + (uint32_t)b * kBatch;
            const uint32_t cnt = min(min((uint32_t)kBatch, end - base), maxlast - (uint32_t)b * kBatch);
            uint32_t qmask = 0;
            if ((uint32_t)tid < cnt) {
                const uint32_t pay = payloads[base + tid];
                Record r = recs[pay];
                s_m[tid] = r.m_nt;
                s_a[tid] = r.a_px;
                s_b[tid] = r.b_py;
                s_c[tid] = r.col_op;
                s_n[tid] = r.nrm_dep;
                s_lthr[tid] = record_lthr(r.col_op.w);
                qmask = cull ? quadrant_mask(r, tx, ty) : 0xFu;
            }
            for (int k = tid; k < (int)cnt * NS; k += 256) s_grad[k] = 0.0f;
            uint32_t qpos[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint64_t bal = __ballot((qmask >> q) & 1u);
                qpos[q] = (uint32_t)__popcll(bal & lt_mask);
                if (lane == 0) s_qcnt[wave][q] = (uint32_t)__popcll(bal);
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if ((qmask >> q) & 1u) {
                    uint32_t off = 0;
                    for (int w = 0; w < wave; ++w) off += s_qcnt[w][q];
                    s_list[q][off + qpos[q]] = (uint16_t)tid;
                }
            }
            const int my_count = (int)(s_qcnt[0][wave] + s_qcnt[1][wave] + s_qcnt[2][wave] + s_qcnt[3][wave]);
            __syncthreads();

            for (int kk = my_count - 1; kk >= 0; --kk) {
                const int j = s_list[wave][kk];
                const uint32_t rel = (uint32_t)b * kBatch + (uint32_t)j;
                if (!__any(rel < last)) continue;
                Record r;
                r.m_nt = s_m[j];
                r.a_px = s_a[j];
                r.b_py = s_b[j];
                r.col_op = s_c[j];
                r.nrm_dep = s_n[j];
                // Branch-free body: a lane that does not contribute gets w = 0 and
                // 1/(1-alpha) = 1 exactly, so its running sums and T are unchanged,
                // and its gradient terms are selected to 0 (no NaN from 0 * inf).
                const HitEval h = eval_hit(r, sx, sy, pxc, pyc, s_lthr[j]);
                const bool act = rel < last && h.alpha > 0.0f;
                const float alpha = act ? h.alpha : 0.0f;
                const float one_m = 1.0f - alpha;
                const float inv_one_m = rcp_refined(one_m);
                const float t_before = T * inv_one_m;
                const float w = t_before * alpha;
                const float t_hit = h.t;
                float g[16];
                float g2[4];
                g[kGCol + 0] = dlr * w;
                g[kGCol + 1] = dlg * w;
                g[kGCol + 2] = dlb * w;
                float dl_dt = 0.0f;
                float dl_dalpha = (dlr * r.col_op.x + dlg * r.col_op.y + dlb * r.col_op.z) * t_before -
                                  (dlr * suf_r + dlg * suf_g + dlb * suf_b) * inv_one_m;
                suf_r += r.col_op.x * w;
                suf_g += r.col_op.y * w;
                suf_b += r.col_op.z * w;
                if (AUX) {
                    float dt_dist = 0.0f, coef = 0.0f;
                    if (lambda != 0.0f) {
                        const float a_pre = w_total - suf_w - w;
                        const float b_pre = wt_total - suf_d - w * t_hit;
                        dt_dist = lambda * 2.0f * w * (a_pre - suf_w);
                        coef = lambda * (2.0f * (t_hit * a_pre - b_pre) + 2.0f * (suf_d - t_hit * suf_w));
                    }
                    g2[kGNrm - 16 + 0] = dnx * w;
                    g2[kGNrm - 16 + 1] = dny * w;
                    g2[kGNrm - 16 + 2] = dnz * w;
                    dl_dt = dld * w + dt_dist;
                    dl_dalpha += dld * (t_hit * t_before - suf_d * inv_one_m) +
                                 (dnx * r.nrm_dep.x + dny * r.nrm_dep.y + dnz * r.nrm_dep.z) * t_before -
                                 (dnx * suf_nx + dny * suf_ny + dnz * suf_nz) * inv_one_m +
                                 coef * t_before - suf_cw * inv_one_m;
                    suf_d += t_hit * w;
                    suf_nx += r.nrm_dep.x * w;
                    suf_ny += r.nrm_dep.y * w;
                    suf_nz += r.nrm_dep.z * w;
                    suf_w += w;
                    suf_cw += coef * w;
                }
                T = t_before;
                // Geometry terms exist only for contributing, unclamped hits.
                const bool geo = act && !h.clamped;
                g[kGOp] = geo ? dl_dalpha * h.ghat : 0.0f;
                const float gd = geo ? h.ghat * (dl_dalpha * r.col_op.w) : 0.0f;  // ghat dL/dghat
                const float dl_du = -h.u * gd;
                const float dl_dv = -h.v * gd;
                const float dnu = h.ray ? dl_du * h.inv_d : 0.0f;
                const float dnv = h.ray ? dl_dv * h.inv_d : 0.0f;
                const float dD = h.ray ? -(h.u * dl_du + h.v * dl_dv + t_hit * dl_dt) * h.inv_d : 0.0f;
                g[kGA + 0] = dnu * sx;
                g[kGA + 1] = dnu * sy;
                g[kGA + 2] = dnu;
                g[kGB + 0] = dnv * sx;
                g[kGB + 1] = dnv * sy;
                g[kGB + 2] = dnv;
                g[kGM + 0] = dD * sx;
                g[kGM + 1] = dD * sy;
                g[kGM + 2] = dD;
                const float gs = h.ray ? 0.0f : 2.0f * gd;
                g[kGPx] = gs * (pxc - r.a_px.w);
                g[kGPy] = gs * (pyc - r.b_py.w);
                g[kGNt] = 0.0f;
                if (AUX) {
                    g[kGNt] = (geo && h.ray) ? dl_dt * h.inv_d : 0.0f;
                    g2[kGDep - 16] = (geo && !h.ray) ? dl_dt : 0.0f;
                }
                const bool contrib = act;
                if (!__any(contrib)) continue;
                const float red = wave_reduce_transpose<16>(g, lane);
                if ((lane & 3) == 0) {
                    if (red != 0.0f) atomicAdd(&s_grad[j * NS + (lane >> 2)], red);
                }
                if (AUX) {
                    const float red2 = wave_reduce_transpose<4>(g2, lane);
                    if ((lane & 15) == 0 && red2 != 0.0f) atomicAdd(&s_grad[j * NS + 16 + (lane >> 4)], red2);
                }
            }
            __syncthreads();
            for (int k = tid; k < (int)cnt * NS; k += 256) {
                const float v = s_grad[k];
                if (v != 0.0f) {
                    const int e = k / NS, slot = k - e * NS;
                    atomicAdd(&grad_coef[(size_t)payloads[base + e] * kGradStride + slot], v);
                }
            }
            __syncthreads();
        }
    }
}

}  // namespace

hipError_t launch_raster_bwd(hipStream_t st, const CamParams* cams_dev, const CamParams* /*cams_host*/,
                             uint32_t n_views, uint32_t total_tiles, const uint32_t* ranges,
                             const uint32_t* payloads, const Record* recs, BwdIn in,
                             float* grad_coef, bool aux, bool cull, TileQueue queue) {
    if (total_tiles == 0) return hipSuccess;
    const uint32_t grid = queue.cursor != nullptr ? std::min(total_tiles, std::max(queue.grid_blocks, 1u))
                                                  : total_tiles;
    if (aux)
        hipLaunchKernelGGL(raster_bwd_kernel<true>, dim3(grid), dim3(256), 0, st, cams_dev, n_views, ranges,
                           payloads, recs, in, grad_coef, cull, total_tiles, queue);
    else
        hipLaunchKernelGGL(raster_bwd_kernel<false>, dim3(grid), dim3(256), 0, st, cams_dev, n_views, ranges,
                           payloads, recs, in, grad_coef, cull, total_tiles, queue);
    return hipGetLastError();
}

}  // namespace gs
