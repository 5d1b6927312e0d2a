// Surfel projection (A2) with tile counting (A3) fused: one thread per
// (view, surfel). Restates crates/gs-kernels/src/shaders/surfel_prep_fwd.wgsl:
// 64-132 and count_tiles of tile_binning.wgsl:30-43, and additionally emits
// the per-surfel ray-splat setup (M, A, B, Nt) so the rasterizer's per-pixel
// intersection is three 2-FMA dot products instead of four triple products.
//
// HBM traffic per (view, surfel): reads 16+8+16+4+12*(d+1)^2 B of parameters,
// writes a 128-B record, a 16-B rect, a 4-B depth key and a 4-B count.
// The record's footprint box is clipped to the pixels of its 16-px tile rect
// (the forward model's coverage). With clip_to_box the emitted rect is the
// set of work tiles (cam.tile_size px) that clipped box reaches (tight
// binning); without it, the 16-px rect itself (work tiles must be 16 px).
// Surfels i >= cam.n_surfels are culled in that view.

#include "gs_internal.h"

namespace gs {

namespace {

__global__ __launch_bounds__(256) void project_kernel(
    const CamParams* __restrict__ cams, uint32_t n_views, uint32_t n, const float4* __restrict__ pos,
    const float2* __restrict__ scales, const float4* __restrict__ quats,
    const float* __restrict__ opac, const float* __restrict__ sh, Record* __restrict__ recs,
    uint4* __restrict__ rects, uint32_t* __restrict__ depth_keys, uint32_t* __restrict__ counts,
    uint32_t* __restrict__ ranges_to_clear, uint32_t num_ranges, uint32_t* __restrict__ presort_keys,
    uint32_t* __restrict__ presort_vals, uint32_t* __restrict__ item_count, bool clip_to_box,
    uint4* __restrict__ emit_info, bool pack_counts) {
    // Views vary fastest over the logical blocks and each XCD takes a
    // contiguous range of them, so the V blocks that read the same 256
    // surfels run on one XCD and the later ones find the parameters in its L2.
    const uint32_t lb = xcd_swizzle(blockIdx.x, gridDim.x);
    const uint32_t view = lb % n_views;
    const uint32_t sblocks = gridDim.x / n_views;
    const uint32_t i = (lb / n_views) * blockDim.x + threadIdx.x;
    // The tile-range table of the following binning stage starts cleared
    // (binning.rs:228); the first view's blocks do it (stream order makes it
    // visible to tile_ranges).
    if (view == 0 && ranges_to_clear != nullptr) {
        for (uint32_t k = i; k < num_ranges; k += sblocks * blockDim.x) ranges_to_clear[k] = 0u;
    }
    if (view == 0 && i == 0 && item_count != nullptr) *item_count = n * n_views;
    if (i >= n) return;
    const CamParams cam = cams[view];
    const uint32_t g = view * n + i;
    // Per-item binning outputs, written on every path. The presort payload
    // carries the tile count in its low 8 bits when pack_counts (for work
    // tiles of 16 or 32 px a rect spans at most 9 x 9 tiles), so the scan
    // reads counts in depth order without a gather.
    // emit_info = (x0 | y0 << 16, x1 | y1 << 16, mask lo, mask hi): the rect
    // in 16-bit tile coordinates and the mask of its tiles the footprint
    // reaches (all ones = every tile).
    auto finish = [&](uint4 rect, uint32_t key, uint32_t count, uint64_t mask) {
        rects[g] = rect;
        depth_keys[g] = key;
        counts[g] = count;
        if (presort_keys) presort_keys[g] = key;
        if (presort_vals) presort_vals[g] = pack_counts ? (g << 8) | count : g;
        if (emit_info)
            emit_info[g] = make_uint4(rect.x | (rect.y << 16), rect.z | (rect.w << 16), (uint32_t)mask,
                                      (uint32_t)(mask >> 32));
    };

    uint4 rect = make_uint4(1u, 1u, 0u, 0u);
    if (i >= cam.n_surfels) {  // not part of this view
        finish(rect, 0u, 0u, ~0ull);
        return;
    }
    const float4 p = pos[i];
    const float2 s = scales[i];
    const float4 q = quats[i];
    const CamFrame f = cam_frame(cam, p, s, q);
    const float depth = -f.cz_;
    if (!(depth >= kNearDepth)) {  // near cull (surfel_prep_fwd.wgsl:171-176); NaN culls too
        finish(rect, 0u, 0u, ~0ull);
        return;
    }
    const RectOut r = tile_rect(cam, f, depth);

    // Ray-splat setup.
    const float3 c = make_float3(f.cx_, f.cy_, f.cz_);
    const float3 tu = make_float3(f.tux, f.tuy, f.tuz);
    const float3 tv = make_float3(f.tvx, f.tvy, f.tvz);
    const float3 M = cross3(tu, tv);
    const float3 A = cross3(tv, c);
    const float3 B = cross3(c, tu);
    const float Nt = -dot3(M, c);
    Record rec;
    rec.m_nt = make_float4(M.x, M.y, M.z, Nt);
    rec.a_px = make_float4(A.x, A.y, A.z, r.px);
    rec.b_py = make_float4(B.x, B.y, B.z, r.py);
    rec.col_op.w = opac[i];

    uint32_t count = 0u;
    uint64_t mask = ~0ull;
    if (r.on_screen) {
        rect = make_uint4(r.x0, r.y0, r.x1, r.y1);
        record_footprint(rec, cam);
        rec.box = clip_box_to_rect(rec.box, rect);
        // Tight binning keeps only the work tiles the alpha >= 1/255
        // footprint reaches; the dropped entries would contribute nothing.
        const uint32_t ts = cam.tile_size;
        if (clip_to_box) rect = box_tiles(rec.box, ts, cam.tiles_x, cam.tiles_y);
        if (rect.x <= rect.z && rect.y <= rect.w) {
            const uint32_t w = rect.z - rect.x + 1u;
            count = w * (rect.w - rect.y + 1u);
            if (clip_to_box && count <= 64u) {
                // Per-tile footprint test (box, disk, ellipse) for rects of at
                // most 64 tiles: bit t = row-major tile t of the rect. Larger
                // rects keep every tile.
                mask = 0ull;
                const float span = (float)ts - 0.5f;
                for (uint32_t t = 0; t < count; ++t) {
                    const float tx = (float)((rect.x + t % w) * ts);
                    const float ty = (float)((rect.y + t / w) * ts);
                    if (footprint_hits(rec, tx + 0.5f, ty + 0.5f, tx + span, ty + span)) mask |= 1ull << t;
                }
                count = (uint32_t)__popcll(mask);
            }
        }
        if (count == 0u) {
            rect = make_uint4(1u, 1u, 0u, 0u);
            mask = ~0ull;
        }
    }
    finish(rect, __float_as_uint(depth), count, mask);
    if (count == 0u) return;  // no entry will ever read the record

    // View-dependent colour (surfel_prep_fwd.wgsl:113-147, 183-186).
    const float inv_len = 1.0f / sqrtf(f.xx * f.xx + f.xy * f.xy + f.xz * f.xz);
    float basis[16];
    const int nb = sh_basis((int)cam.sh_degree, f.xx * inv_len, f.xy * inv_len, f.xz * inv_len, basis);
    // 16-B loads of the (48-float, 192-B aligned) coefficient row, only as
    // far as the degree needs.
    const float4* sh4 = reinterpret_cast<const float4*>(sh + (size_t)i * kShFloats);
    float shv[kShFloats];
#pragma unroll
    for (int q = 0; q < kShFloats / 4; ++q) {
        const float4 v = 4 * q < 3 * nb ? sh4[q] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        shv[4 * q + 0] = v.x;
        shv[4 * q + 1] = v.y;
        shv[4 * q + 2] = v.z;
        shv[4 * q + 3] = v.w;
    }
    float cr = 0.5f, cg = 0.5f, cb = 0.5f;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        if (k < nb) {
            cr += basis[k] * shv[3 * k + 0];
            cg += basis[k] * shv[3 * k + 1];
            cb += basis[k] * shv[3 * k + 2];
        }
    }
    // Camera-facing unit normal (surfel_prep_fwd.wgsl:188-194).
    float3 nrm = M;
    const float mm = dot3(M, M);
    if (mm > 0.0f) {
        const float inv = 1.0f / sqrtf(mm);
        nrm = make_float3(M.x * inv, M.y * inv, M.z * inv);
        if (dot3(nrm, c) > 0.0f) nrm = make_float3(-nrm.x, -nrm.y, -nrm.z);
    }
    rec.col_op = make_float4(fmaxf(cr, 0.0f), fmaxf(cg, 0.0f), fmaxf(cb, 0.0f), opac[i]);
    rec.nrm_dep = make_float4(nrm.x, nrm.y, nrm.z, depth);
    recs[g] = rec;
}

}  // namespace

hipError_t launch_project(hipStream_t st, const CamParams* cams_dev, uint32_t n_views, uint32_t n,
                          const float4* pos, const float2* scales, const float4* quats,
                          const float* opac, const float* sh, Record* recs, uint4* rects,
                          uint32_t* depth_keys, uint32_t* counts, uint32_t* ranges_to_clear,
                          uint32_t num_ranges, uint32_t* presort_keys, uint32_t* presort_vals,
                          uint32_t* item_count, bool clip_to_box, uint4* emit_info, bool pack_counts) {
    if (n == 0 && num_ranges == 0 && item_count == nullptr) return hipSuccess;
    if (n_views == 0) return hipSuccess;
    const uint32_t blocks = div_up(n > num_ranges ? (n ? n : 1) : num_ranges, 256);
    dim3 grid(blocks * n_views);
    hipLaunchKernelGGL(project_kernel, grid, dim3(256), 0, st, cams_dev, n_views, n, pos, scales, quats, opac,
                       sh, recs, rects, depth_keys, counts, ranges_to_clear, num_ranges, presort_keys,
                       presort_vals, item_count, clip_to_box, emit_info, pack_counts);
    return hipGetLastError();
}

}  // namespace gs
