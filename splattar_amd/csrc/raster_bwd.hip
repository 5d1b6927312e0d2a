// Tile rasterizer backward: the hand-written adjoint of raster_fwd.
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
//   * a 16x16 work tile is one wave64 with four pixels per lane, one in each
//     8x8 block (raster_bwd_wave_kernel): a record is evaluated only in the
//     blocks its footprint reaches, its terms over the tile are summed in
//     registers, reduced once per (record, tile) and added to the record's
//     gradient with hardware global FP32 atomics (no CAS loops). A 32x32
//     tile is eight waves, each owning an 8-wide, 16-tall column of two
//     blocks (two pixels per lane), whose per-record reductions meet in LDS
//     slots (ds_add_f32) flushed once per batch; an 8x8 tile is one wave with
//     one pixel per lane (raster_bwd_kernel). Records are staged BATCH at a
//     time into order-preserving lists that carry the blocks each reaches;
//   * each wave reduces its lanes' 16 values with a transposed butterfly (17
//     lane exchanges instead of 96, on permlane swaps and DPP rather than LDS
//     permutes).
// AUX = true (16-px tiles, raster_bwd_kernel) adds the depth / normal image
// losses and the in-walk depth distortion loss (backward.rs:126-153).

#include "gs_internal.h"

namespace gs {

namespace {

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

// Per-pixel state of the backward walk.
struct PixState {
    float sx, sy, pxc, pyc;
    float dlr, dlg, dlb, dld, dnx, dny, dnz;
    float T, w_total, wt_total;
    float suf_r, suf_g, suf_b, suf_d, suf_nx, suf_ny, suf_nz, suf_w, suf_cw;
    uint32_t last;
};

// Loads a pixel's incoming gradients and forward state; last = 0 when the
// pixel is outside the image or carries no loss gradient (nothing to walk).
template <bool AUX>
__device__ __forceinline__ void load_pixel(PixState& p, const CamParams& cam, int px, int py,
                                           const BwdIn& in, float lambda) {
    pixel_s(cam, px, py, p.sx, p.sy);
    p.pxc = (float)px + 0.5f;
    p.pyc = (float)py + 0.5f;
    p.dlr = p.dlg = p.dlb = p.dld = p.dnx = p.dny = p.dnz = 0.f;
    p.T = 1.0f;
    p.w_total = p.wt_total = 0.f;
    p.suf_r = p.suf_g = p.suf_b = p.suf_d = p.suf_nx = p.suf_ny = p.suf_nz = p.suf_w = p.suf_cw = 0.f;
    p.last = 0;
    const bool inside = px < (int)cam.width && py < (int)cam.height;
    if (!inside) return;
    const size_t pi = (size_t)cam.pix_base + (size_t)py * cam.width + (size_t)px;
    const float4 dl = in.dl_dcolor[pi];
    p.dlr = dl.x;
    p.dlg = dl.y;
    p.dlb = dl.z;
    const float4 aux = in.out_aux[pi];
    p.T = aux.x;
    uint32_t last = in.ncontrib[pi];
    bool has_loss = p.dlr != 0.f || p.dlg != 0.f || p.dlb != 0.f;
    if (AUX) {
        p.dld = dl.w;
        if (in.dl_dnormal) {
            const float4 dn = in.dl_dnormal[pi];
            p.dnx = dn.x;
            p.dny = dn.y;
            p.dnz = dn.z;
        }
        p.w_total = in.out_color[pi].w;
        p.wt_total = aux.y;
        has_loss = has_loss || p.dld != 0.f || p.dnx != 0.f || p.dny != 0.f || p.dnz != 0.f || lambda != 0.f;
    }
    p.last = has_loss ? last : 0u;
}

template <bool AUX, bool ACC>
__device__ __forceinline__ void put(float& dst, float v) {
    if constexpr (ACC) dst += v;
    else dst = v;
}

// Writes (ACC = false) or adds (ACC = true) one pixel-record hit's gradient
// terms into g / g2 and steps the pixel's walk, given the hit evaluation h and
// whether this lane contributes (act). Branch-free: a lane that does not
// contribute gets w = 0 and 1/(1-alpha) = 1 exactly, so its running sums and
// T are unchanged, and its gradient terms are selected to 0 (no NaN from
// 0 * inf); a wave with no contributing lane need not call it at all.
template <bool AUX, bool ACC>
__device__ __forceinline__ void add_hit(PixState& p, const Record& r, const HitEval& h, bool act,
                                        float lambda, float (&g)[16], float (&g2)[4]) {
    const float alpha = act ? h.alpha : 0.0f;
    const float inv_one_m = rcp_refined(1.0f - alpha);
    const float t_before = p.T * inv_one_m;
    const float w = t_before * alpha;
    const float t_hit = h.t;
    put<AUX, ACC>(g[kGCol + 0], p.dlr * w);
    put<AUX, ACC>(g[kGCol + 1], p.dlg * w);
    put<AUX, ACC>(g[kGCol + 2], p.dlb * w);
    float dl_dt = 0.0f;
    float dl_dalpha = (p.dlr * r.col_op.x + p.dlg * r.col_op.y + p.dlb * r.col_op.z) * t_before -
                      (p.dlr * p.suf_r + p.dlg * p.suf_g + p.dlb * p.suf_b) * inv_one_m;
    p.suf_r += r.col_op.x * w;
    p.suf_g += r.col_op.y * w;
    p.suf_b += r.col_op.z * w;
    if (AUX) {
        float dt_dist = 0.0f, coef = 0.0f;
        if (lambda != 0.0f) {
            const float a_pre = p.w_total - p.suf_w - w;
            const float b_pre = p.wt_total - p.suf_d - w * t_hit;
            dt_dist = lambda * 2.0f * w * (a_pre - p.suf_w);
            coef = lambda * (2.0f * (t_hit * a_pre - b_pre) + 2.0f * (p.suf_d - t_hit * p.suf_w));
        }
        put<AUX, ACC>(g2[kGNrm - 16 + 0], p.dnx * w);
        put<AUX, ACC>(g2[kGNrm - 16 + 1], p.dny * w);
        put<AUX, ACC>(g2[kGNrm - 16 + 2], p.dnz * w);
        dl_dt = p.dld * w + dt_dist;
        dl_dalpha += p.dld * (t_hit * t_before - p.suf_d * inv_one_m) +
                     (p.dnx * r.nrm_dep.x + p.dny * r.nrm_dep.y + p.dnz * r.nrm_dep.z) * t_before -
                     (p.dnx * p.suf_nx + p.dny * p.suf_ny + p.dnz * p.suf_nz) * inv_one_m +
                     coef * t_before - p.suf_cw * inv_one_m;
        p.suf_d += t_hit * w;
        p.suf_nx += r.nrm_dep.x * w;
        p.suf_ny += r.nrm_dep.y * w;
        p.suf_nz += r.nrm_dep.z * w;
        p.suf_w += w;
        p.suf_cw += coef * w;
    }
    p.T = t_before;
    // Geometry terms exist only for contributing, unclamped hits, and the
    // plane terms only on the ray branch: the masks are applied once, to ghat
    // and to 1/D, so the terms derived from them are exact zeros (possibly
    // -0, which no sum tells apart) where the terms are absent. dl_dalpha,
    // u and v are finite (|1/D| <= 1e12 and finite record coefficients), so
    // no inf * 0 arises.
    const bool geo = act && !h.clamped;
    const float ghat_g = geo ? h.ghat : 0.0f;
    const float inv_r = h.ray ? h.inv_d : 0.0f;
    put<AUX, ACC>(g[kGOp], dl_dalpha * ghat_g);
    const float gd = ghat_g * (dl_dalpha * r.col_op.w);  // ghat dL/dghat
    const float dl_du = -h.u * gd;
    const float dl_dv = -h.v * gd;
    const float dnu = dl_du * inv_r;
    const float dnv = dl_dv * inv_r;
    // Colour-only (AUX = false): dl_dt = 0, and the record's depth is not
    // staged, so t_hit must not enter (it may hold anything).
    const float dD_num = AUX ? h.u * dl_du + h.v * dl_dv + t_hit * dl_dt : h.u * dl_du + h.v * dl_dv;
    const float dD = -dD_num * inv_r;
    put<AUX, ACC>(g[kGA + 0], dnu * p.sx);
    put<AUX, ACC>(g[kGA + 1], dnu * p.sy);
    put<AUX, ACC>(g[kGA + 2], dnu);
    put<AUX, ACC>(g[kGB + 0], dnv * p.sx);
    put<AUX, ACC>(g[kGB + 1], dnv * p.sy);
    put<AUX, ACC>(g[kGB + 2], dnv);
    put<AUX, ACC>(g[kGM + 0], dD * p.sx);
    put<AUX, ACC>(g[kGM + 1], dD * p.sy);
    put<AUX, ACC>(g[kGM + 2], dD);
    const float gs = h.ray ? 0.0f : 2.0f * gd;
    put<AUX, ACC>(g[kGPx], gs * (p.pxc - r.a_px.w));
    put<AUX, ACC>(g[kGPy], gs * (p.pyc - r.b_py.w));
    if (AUX) {
        put<AUX, ACC>(g[kGNt], (geo && h.ray) ? dl_dt * h.inv_d : 0.0f);
        put<AUX, ACC>(g2[kGDep - 16], (geo && !h.ray) ? dl_dt : 0.0f);
    }
}

template <bool AUX, int TW, int BATCH>
__global__ __launch_bounds__(TW == 8 ? 64 : TW * TW / 2) void raster_bwd_kernel(
    const CamParams* __restrict__ cams, uint32_t n_views, const uint32_t* __restrict__ ranges,
    const uint32_t* __restrict__ payloads, const Record* __restrict__ recs, BwdIn in,
    float* __restrict__ grad_coef, bool cull, uint32_t total_tiles, TileQueue queue, uint32_t tile_begin) {
    constexpr int NS = grad_stride(AUX);     // slots per record (LDS and grad_coef)
    constexpr bool TWO = TW >= 16;           // two pixels per lane (an 8x16 column per wave)
    constexpr int NC = TW / 8;               // 8x8 blocks per tile row
    constexpr int NW = TWO ? TW * TW / 128 : 1;  // waves
    constexpr int NT = NW * 64;
    constexpr int SW = BATCH / 64;           // staging waves per batch
    constexpr int SPT = (BATCH + NT - 1) / NT;
    constexpr int MB = 10;                   // list entry: slot | top << MB | bottom << (MB + 1)
    static_assert(BATCH % 64 == 0 && BATCH <= (1 << MB), "batch of whole waves");
    __shared__ float4 s_m[BATCH], s_a[BATCH], s_b[BATCH], s_c[BATCH], s_n[AUX ? BATCH : 1];
    __shared__ float s_lthr[BATCH];
    __shared__ uint32_t s_pay[BATCH];        // the batch's surfel ids, for the flush
    __shared__ uint16_t s_list[NW][BATCH];
    __shared__ uint32_t s_qcnt[SW][NW];      // [staging wave][list wave]
    __shared__ float s_grad[BATCH * NS];
    __shared__ uint32_t s_view, s_maxlast, s_next;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    // Wave w owns block column c = w % NC, rows 16 (w / NC) .. + 15: blocks
    // top = (c, 2 (w / NC)) and bottom = (c, 2 (w / NC) + 1).
    const int col = wave % NC, hrow = wave / NC;
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
        const uint32_t gtile = queue.order != nullptr ? queue.order[pos] : tile_begin + pos;
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
        const uint32_t x0 = tx * TW, y0 = ty * TW;
        const float lambda = AUX ? cam.lambda_dist : 0.0f;
        const int px = (int)x0 + col * 8 + (lane & 7);
        const int py = (int)y0 + hrow * 16 + (lane >> 3);
        PixState pa, pb;
        load_pixel<AUX>(pa, cam, px, py, in, lambda);
        if constexpr (TWO) load_pixel<AUX>(pb, cam, px, py + 8, in, lambda);
        else pb.last = 0u;
        const uint32_t mylast = max(pa.last, pb.last);
        if (mylast) atomicMax(&s_maxlast, mylast);
        __syncthreads();
        const uint32_t maxlast = s_maxlast;
        if (maxlast == 0) continue;

        const int nbatch = (int)((maxlast + BATCH - 1) / BATCH);
        for (int b = nbatch - 1; b >= 0; --b) {
            const uint32_t base = start + (uint32_t)b * BATCH;
            const uint32_t cnt = min(min((uint32_t)BATCH, end - base), maxlast - (uint32_t)b * BATCH);
            uint32_t memb[SPT][NW], qpos[SPT][NW];
#pragma unroll
            for (int m = 0; m < SPT; ++m) {
                const int slot = tid + m * NT;
                if (slot >= BATCH) continue;  // whole waves (NT, BATCH multiples of 64)
                uint32_t qmask = 0;
                if ((uint32_t)slot < cnt) {
                    const uint32_t pay = payloads[base + slot];
                    Record r = recs[pay];
                    s_pay[slot] = pay;
                    s_m[slot] = r.m_nt;
                    s_a[slot] = r.a_px;
                    s_b[slot] = r.b_py;
                    s_c[slot] = r.col_op;
                    if (AUX) s_n[slot] = r.nrm_dep;
                    s_lthr[slot] = record_lthr(r.col_op.w);
                    qmask = cull ? block_mask<TW>(r, x0, y0) : ~0u;
                }
                // Order-preserving per-wave lists: wave w takes the records
                // that reach its top or bottom block.
                const int sw = slot >> 6;
#pragma unroll
                for (int w = 0; w < NW; ++w) {
                    const int bt = TWO ? (2 * (w / NC)) * NC + w % NC : 0;
                    const uint32_t top = (qmask >> bt) & 1u;
                    const uint32_t bot = TWO ? (qmask >> (bt + NC)) & 1u : 0u;
                    memb[m][w] = top | (bot << 1);
                    const uint64_t bal = __ballot(memb[m][w] != 0u);
                    qpos[m][w] = (uint32_t)__popcll(bal & lt_mask);
                    if (lane == 0) s_qcnt[sw][w] = (uint32_t)__popcll(bal);
                }
            }
            for (int k = tid; k < (int)cnt * NS; k += NT) s_grad[k] = 0.0f;
            __syncthreads();
#pragma unroll
            for (int m = 0; m < SPT; ++m) {
                const int slot = tid + m * NT;
                if (slot >= BATCH) continue;
                const int sw = slot >> 6;
#pragma unroll
                for (int w = 0; w < NW; ++w) {
                    if (memb[m][w] != 0u) {
                        uint32_t off = 0;
                        for (int v = 0; v < sw; ++v) off += s_qcnt[v][w];
                        s_list[w][off + qpos[m][w]] = (uint16_t)(slot | (memb[m][w] << MB));
                    }
                }
            }
            int my_count = 0;
#pragma unroll
            for (int v = 0; v < SW; ++v) my_count += (int)s_qcnt[v][wave];
            __syncthreads();

            for (int kk = my_count - 1; kk >= 0; --kk) {
                // The entry is the same for every lane: make it scalar, so the
                // block branches below are uniform.
                const uint32_t e = __builtin_amdgcn_readfirstlane((uint32_t)s_list[wave][kk]);
                const uint32_t j = e & ((1u << MB) - 1u);
                const bool top = (e >> MB) & 1u, bot = TWO && ((e >> (MB + 1)) & 1u);
                const uint32_t rel = (uint32_t)b * BATCH + j;
                // A block is walked only while some pixel of it is at or before
                // its last contributor (a pixel past it changes nothing).
                const bool need_a = top && __any(rel < pa.last);
                const bool need_b = bot && __any(rel < pb.last);
                if (!need_a && !need_b) continue;
                Record r;
                r.m_nt = s_m[j];
                r.a_px = s_a[j];
                r.b_py = s_b[j];
                r.col_op = s_c[j];
                r.nrm_dep = AUX ? s_n[AUX ? j : 0] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                const float lthr = s_lthr[j];
                // Evaluate first; the gradient terms and the reduction run only
                // for blocks with a contributing lane (for the others the walk
                // state would not change).
                HitEval ha, hb;
                bool act_a = false, act_b = false;
                if (need_a) {
                    ha = eval_hit(r, pa.sx, pa.sy, pa.pxc, pa.pyc, lthr);
                    act_a = rel < pa.last && ha.alpha > 0.0f;
                }
                if (need_b) {
                    hb = eval_hit(r, pb.sx, pb.sy, pb.pxc, pb.pyc, lthr);
                    act_b = rel < pb.last && hb.alpha > 0.0f;
                }
                const bool go_a = need_a && __any(act_a);
                const bool go_b = need_b && __any(act_b);
                if (!go_a && !go_b) continue;
                float g[16];
                float g2[4];
                g[kGNt] = 0.0f;  // written only by the AUX terms
#pragma unroll
                for (int k = 0; k < 4; ++k) g2[k] = 0.0f;
                if (go_a && go_b) {
                    add_hit<AUX, false>(pa, r, ha, act_a, lambda, g, g2);
                    add_hit<AUX, true>(pb, r, hb, act_b, lambda, g, g2);
                } else if (go_a) {
                    add_hit<AUX, false>(pa, r, ha, act_a, lambda, g, g2);
                } else {
                    add_hit<AUX, false>(pb, r, hb, act_b, lambda, g, g2);
                }
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
            for (int k = tid; k < (int)cnt * NS; k += NT) {
                const float v = s_grad[k];
                if (v != 0.0f) {
                    const int e = k / NS, slot = k - e * NS;
                    // A colour channel clamped at 0 in the forward passes no
                    // gradient to SH (surfel_prep_bwd.wgsl's clamp mask; the
                    // record stores such a channel as exactly 0).
                    if (slot >= kGCol && slot < kGCol + 3) {
                        const float4 c = s_c[e];
                        const float cc = slot == kGCol ? c.x : (slot == kGCol + 1 ? c.y : c.z);
                        if (cc == 0.0f) continue;
                    }
                    atomicAdd(&grad_coef[(size_t)s_pay[e] * NS + slot], v);
                }
            }
            __syncthreads();
        }
    }
}

// Colour-only backward over 16-px work tiles with one wave per tile: lane l
// owns pixel (l & 7, l >> 3) of each of the tile's four 8x8 blocks, so a
// record's terms over the whole tile are summed in registers and reduced
// once per (record, tile) instead of once per half tile, and the reduced
// values go straight to the global FP32 atomics (no LDS accumulation slots:
// one wave is the only contributor of its tile). The records a batch stages
// form one order-preserving list carrying each record's 4-bit block mask.
template <int BATCH>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void raster_bwd_wave_kernel(
    const CamParams* __restrict__ cams, uint32_t n_views, const uint32_t* __restrict__ ranges,
    const uint32_t* __restrict__ payloads, const Record* __restrict__ recs, BwdIn in,
    float* __restrict__ grad_coef, bool cull, uint32_t total_tiles, TileQueue queue, uint32_t tile_begin) {
    constexpr int NS = grad_stride(false);
    constexpr int SPT = BATCH / 64;          // staging rounds per batch
    constexpr int MB = 10;                   // list entry: slot | block mask << MB
    static_assert(BATCH % 64 == 0 && BATCH <= (1 << MB), "batch of whole waves");
    __shared__ float4 s_m[BATCH], s_a[BATCH], s_b[BATCH], s_c[BATCH];
    __shared__ float s_lthr[BATCH];
    __shared__ uint32_t s_pay[BATCH];
    __shared__ uint16_t s_list[BATCH];

    const int lane = threadIdx.x;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (uint32_t iter = 0;; ++iter) {
        uint32_t pos;
        if (queue.cursor != nullptr) {
            uint32_t claim = 0;
            if (lane == 0) {
                const uint32_t q = min(kTileQueues, gridDim.x), x = blockIdx.x % q;
                claim = atomicAdd(queue.cursor + x, 1u) * q + x;
            }
            pos = __builtin_amdgcn_readfirstlane(claim);
        } else {
            pos = iter == 0 ? xcd_swizzle(blockIdx.x, gridDim.x) : total_tiles;
        }
        if (pos >= total_tiles) break;
        const uint32_t gtile = queue.order != nullptr ? queue.order[pos] : tile_begin + pos;
        const uint32_t start = ranges[2 * (size_t)gtile], end = ranges[2 * (size_t)gtile + 1];
        if (end <= start) continue;
        const uint32_t v = view_of_tile(cams, n_views, gtile);
        const CamParams& cam = cams[v];
        const uint32_t local = gtile - cam.tile_base;
        const uint32_t tx = local % cam.tiles_x, ty = local / cam.tiles_x;
        const uint32_t x0 = tx * 16, y0 = ty * 16;
        PixState p[4];
        uint32_t mylast = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            load_pixel<false>(p[q], cam, (int)x0 + 8 * (q & 1) + (lane & 7), (int)y0 + 8 * (q >> 1) + (lane >> 3),
                              in, 0.0f);
            mylast = max(mylast, p[q].last);
        }
        for (int o = 32; o >= 1; o >>= 1) mylast = max(mylast, (uint32_t)__shfl_xor((int)mylast, o));
        const uint32_t maxlast = __builtin_amdgcn_readfirstlane(mylast);
        if (maxlast == 0) continue;

        const int nbatch = (int)((maxlast + BATCH - 1) / BATCH);
        for (int b = nbatch - 1; b >= 0; --b) {
            const uint32_t base = start + (uint32_t)b * BATCH;
            const uint32_t cnt = min(min((uint32_t)BATCH, end - base), maxlast - (uint32_t)b * BATCH);
            __syncthreads();  // the previous batch's walk has read its LDS
            uint32_t count = 0;
#pragma unroll
            for (int m = 0; m < SPT; ++m) {
                const int slot = lane + m * 64;
                uint32_t qmask = 0;
                if ((uint32_t)slot < cnt) {
                    const uint32_t pay = payloads[base + slot];
                    Record r = recs[pay];
                    s_pay[slot] = pay;
                    s_m[slot] = r.m_nt;
                    s_a[slot] = r.a_px;
                    s_b[slot] = r.b_py;
                    s_c[slot] = r.col_op;
                    s_lthr[slot] = record_lthr(r.col_op.w);
                    qmask = cull ? block_mask<16>(r, x0, y0) : 0xFu;
                }
                const uint64_t bal = __ballot(qmask != 0u);
                if (qmask != 0u)
                    s_list[count + (uint32_t)__popcll(bal & lt_mask)] = (uint16_t)(slot | (qmask << MB));
                count += (uint32_t)__popcll(bal);
            }
            __syncthreads();

            for (int kk = (int)count - 1; kk >= 0; --kk) {
                const uint32_t e = __builtin_amdgcn_readfirstlane((uint32_t)s_list[kk]);
                const uint32_t j = e & ((1u << MB) - 1u);
                const uint32_t qm = e >> MB;
                const uint32_t rel = (uint32_t)b * BATCH + j;
                bool need[4];
                bool any_need = false;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    need[q] = ((qm >> q) & 1u) && __any(rel < p[q].last);
                    any_need = any_need || need[q];
                }
                if (!any_need) continue;
                Record r;
                r.m_nt = s_m[j];
                r.a_px = s_a[j];
                r.b_py = s_b[j];
                r.col_op = s_c[j];
                r.nrm_dep = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                const float lthr = s_lthr[j];
                float g[16];
                float g2[4];
                g[kGNt] = 0.0f;
                bool first = true;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if (!need[q]) continue;
                    const HitEval h = eval_hit(r, p[q].sx, p[q].sy, p[q].pxc, p[q].pyc, lthr);
                    const bool act = rel < p[q].last && h.alpha > 0.0f;
                    if (!__any(act)) continue;
                    if (first) add_hit<false, false>(p[q], r, h, act, 0.0f, g, g2);
                    else add_hit<false, true>(p[q], r, h, act, 0.0f, g, g2);
                    first = false;
                }
                if (first) continue;
                const float red = wave_reduce_transpose<16>(g, lane);
                if ((lane & 3) == 0 && red != 0.0f) {
                    const int slot = lane >> 2;
                    bool keep = true;
                    if (slot >= kGCol && slot < kGCol + 3) {
                        // A colour channel clamped at 0 in the forward passes no
                        // gradient to SH (the record stores it as exactly 0).
                        const float4 c = s_c[j];
                        keep = (slot == kGCol ? c.x : (slot == kGCol + 1 ? c.y : c.z)) != 0.0f;
                    }
                    if (keep) atomicAdd(&grad_coef[(size_t)s_pay[j] * NS + slot], red);
                }
            }
        }
    }
}

template <bool AUX, int TW, int BATCH>
void launch_bwd(hipStream_t st, uint32_t grid, const CamParams* cams_dev, uint32_t n_views,
                uint32_t total_tiles, const uint32_t* ranges, const uint32_t* payloads, const Record* recs,
                BwdIn in, float* grad_coef, bool cull, TileQueue queue, uint32_t tile_begin) {
    constexpr int NT = TW == 8 ? 64 : TW * TW / 2;
    hipLaunchKernelGGL((raster_bwd_kernel<AUX, TW, BATCH>), dim3(grid), dim3(NT), 0, st, cams_dev, n_views,
                       ranges, payloads, recs, in, grad_coef, cull, total_tiles, queue, tile_begin);
}

}  // namespace

hipError_t launch_raster_bwd(hipStream_t st, const CamParams* cams_dev, const CamParams* /*cams_host*/,
                             uint32_t n_views, uint32_t total_tiles, const uint32_t* ranges,
                             const uint32_t* payloads, const Record* recs, BwdIn in,
                             float* grad_coef, bool aux, bool cull, TileQueue queue, RasterShape shape,
                             uint32_t tile_begin) {
    if (total_tiles == 0) return hipSuccess;
    const uint32_t grid = queue.cursor != nullptr ? std::min(total_tiles, std::max(queue.grid_blocks, 1u))
                                                  : total_tiles;
    const uint32_t t = shape.tile, b = shape.bwd_batch;
    auto go = [&](auto fn) {
        fn(st, grid, cams_dev, n_views, total_tiles, ranges, payloads, recs, in, grad_coef, cull, queue,
           tile_begin);
    };
    if (aux) {
        if (t != 16) return hipErrorInvalidValue;  // depth / normal losses: 16-px tiles
        go(launch_bwd<true, 16, 128>);
    } else if (t == 8 && b == 64) {
        go(launch_bwd<false, 8, 64>);
    } else if (t == 8 && b == 128) {
        go(launch_bwd<false, 8, 128>);
    } else if (t == 16) {
        // One wave per tile: twice the workgroups for the same waves per CU.
        const uint32_t g2 = queue.cursor != nullptr ? std::min(total_tiles, std::max(2 * queue.grid_blocks, 1u))
                                                    : total_tiles;
        if (b == 64)
            hipLaunchKernelGGL(raster_bwd_wave_kernel<64>, dim3(g2), dim3(64), 0, st, cams_dev, n_views, ranges,
                               payloads, recs, in, grad_coef, cull, total_tiles, queue, tile_begin);
        else if (b == 128)
            hipLaunchKernelGGL(raster_bwd_wave_kernel<128>, dim3(g2), dim3(64), 0, st, cams_dev, n_views, ranges,
                               payloads, recs, in, grad_coef, cull, total_tiles, queue, tile_begin);
        else
            return hipErrorInvalidValue;
    } else if (t == 32 && b == 128) {
        go(launch_bwd<false, 32, 128>);
    } else if (t == 32 && b == 256) {
        go(launch_bwd<false, 32, 256>);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace gs
