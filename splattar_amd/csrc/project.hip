// Surfel projection with tile counting fused: one thread per
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
// Surfels i >= cam.n_surfels are culled in that view. rects, depth_keys and
// counts may be null (callers that read only the presort and emit_info
// outputs); with write_normal false the record's normal / depth float4 is
// not written (colour-only consumers never read it).

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
    uint4* __restrict__ emit_info, bool pack_counts, bool write_normal) {
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
    // Every thread of the block runs to the cooperative tile-mask step below
    // (it uses the whole wave); items past n only take part with no tiles.
    const bool in_range = i < n;
    const CamParams cam = cams[view];
    const uint32_t g = view * n + i;
    // Per-item binning outputs, written on every path. The presort payload
    // carries the tile count in its low kPackBits bits when pack_counts (for
    // work tiles of 16 or 32 px a rect spans at most 9 x 9 tiles), so the
    // scan reads counts in depth order without a gather.
    // emit_info = (x0 | y0 << 16, x1 | y1 << 16, mask lo, mask hi): the rect
    // in 16-bit tile coordinates and the mask of its tiles the footprint
    // reaches (all ones = every tile).
    auto finish = [&](uint4 rect, uint32_t key, uint32_t count, uint64_t mask) {
        if (rects) rects[g] = rect;
        if (depth_keys) depth_keys[g] = key;
        if (counts) counts[g] = count;
        if (presort_keys) presort_keys[g] = key;
        if (presort_vals) presort_vals[g] = pack_counts ? (g << kPackBits) | count : g;
        if (emit_info)
            emit_info[g] = make_uint4(rect.x | (rect.y << 16), rect.z | (rect.w << 16), (uint32_t)mask,
                                      (uint32_t)(mask >> 32));
    };

    uint4 rect = make_uint4(1u, 1u, 0u, 0u);
    bool live = in_range && i < cam.n_surfels;  // surfels past n_surfels are not part of this view
    CamFrame f;
    float depth = 0.0f;
    if (live) {
        f = cam_frame(cam, pos[i], scales[i], quats[i]);
        depth = -f.cz_;
        // Near cull (surfel_prep_fwd.wgsl:171-176); NaN culls too.
        live = depth >= kNearDepth;
    }
    Record rec;
    uint32_t count = 0u;
    uint64_t mask = ~0ull;
    bool masked = false;  // rect of at most 64 tiles: per-tile footprint test below
    float3 c = make_float3(0.0f, 0.0f, 0.0f), M = c;
    if (live) {
        const RectOut r = tile_rect(cam, f, depth);
        // Ray-splat setup.
        c = make_float3(f.cx_, f.cy_, f.cz_);
        const float3 tu = make_float3(f.tux, f.tuy, f.tuz);
        const float3 tv = make_float3(f.tvx, f.tvy, f.tvz);
        M = cross3(tu, tv);
        const float3 A = cross3(tv, c);
        const float3 B = cross3(c, tu);
        const float Nt = -dot3(M, c);
        rec.m_nt = make_float4(M.x, M.y, M.z, Nt);
        rec.a_px = make_float4(A.x, A.y, A.z, r.px);
        rec.b_py = make_float4(B.x, B.y, B.z, r.py);
        rec.col_op.w = opac[i];
        if (r.on_screen) {
            rect = make_uint4(r.x0, r.y0, r.x1, r.y1);
            record_footprint(rec, cam);
            rec.box = clip_box_to_rect(rec.box, rect);
            // Tight binning keeps only the work tiles the alpha >= 1/255
            // footprint reaches; the dropped entries would contribute nothing.
            if (clip_to_box) rect = box_tiles(rec.box, cam.tile_size, cam.tiles_x, cam.tiles_y);
            if (rect.x <= rect.z && rect.y <= rect.w) {
                count = (rect.z - rect.x + 1u) * (rect.w - rect.y + 1u);
                // Rects of at most 64 tiles get a per-tile footprint test (box,
                // disk, ellipse); larger ones keep every tile.
                masked = clip_to_box && count <= 64u;
            }
        }
    }

    // Per-tile footprint masks, wave-cooperatively: the (item, tile) pairs of
    // the wave's masked items are dealt 64 at a time over the lanes (owner by
    // binary search over the wave's exclusive scan of tile counts), each lane
    // tests one tile against its owner's footprint read from LDS; an owner's
    // pairs of a round are consecutive lanes, so its bits are a shifted slice
    // of the round's ballot, merged into its mask by the run's first lane.
    // Bit t = row-major tile t of the rect.
    // The cost follows the wave's total tile count, not 64 x its widest item.
    // s_stage holds the footprints for the masks (s_fp), then the block's
    // records for one contiguous store (rows of kRecRow float4, one spare
    // against LDS bank conflicts).
    constexpr int kRecRow = 9;
    __shared__ float4 s_stage[256 * kRecRow];
    __shared__ uint32_t s_bits[256][2];
    __shared__ uint32_t s_write[256];
    // Field-major (lane-consecutive 16-B accesses): box, ell0, ell1, (px, py, rect width, 0).
    float4 (*s_fp)[256] = reinterpret_cast<float4 (*)[256]>(s_stage);
    {
        const int tid = threadIdx.x, lane = tid & 63, wbase = tid & ~63;
        const uint32_t mine = masked ? count : 0u;
        const uint32_t rw = rect.z - rect.x + 1u;
        if (masked) {
            s_fp[0][tid] = rec.box;
            s_fp[1][tid] = rec.ell0;
            s_fp[2][tid] = rec.ell1;
            s_fp[3][tid] = make_float4(rec.a_px.w, rec.b_py.w, (float)rw, 0.0f);
        }
        s_bits[tid][0] = 0u;
        s_bits[tid][1] = 0u;
        uint32_t incl = mine;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t v = __shfl_up(incl, o);
            if (lane >= o) incl += v;
        }
        const uint32_t total = __shfl(incl, 63);
        const uint32_t excl = incl - mine;
        __syncthreads();
        const uint32_t ts = cam.tile_size;
        const float span = (float)ts - 0.5f;
        for (uint32_t k0 = 0; k0 < total; k0 += 64) {
            const uint32_t k = k0 + (uint32_t)lane;
            int owner = 0;
            for (int step = 32; step >= 1; step >>= 1) {
                const int cand = owner + step;
                const uint32_t ex = __shfl(excl, cand & 63);
                if (cand < 64 && ex <= k) owner = cand;
            }
            const uint32_t o_excl = __shfl(excl, owner);
            const uint32_t o_end = o_excl + __shfl(mine, owner);
            const uint32_t o_x0 = __shfl(rect.x, owner), o_y0 = __shfl(rect.y, owner);
            const int prev_owner = __shfl_up(owner, 1);
            const uint32_t t = k - o_excl;
            bool hit = false;
            if (k < total) {
                const float4 q3 = s_fp[3][wbase + owner];
                const uint32_t w = (uint32_t)q3.z;
                // t < 64 and w <= 64: the float quotient's floor is exact.
                const uint32_t ty = (uint32_t)(((float)t + 0.5f) / q3.z);
                const uint32_t tx = t - ty * w;
                Record fr;
                fr.box = s_fp[0][wbase + owner];
                fr.ell0 = s_fp[1][wbase + owner];
                fr.ell1 = s_fp[2][wbase + owner];
                fr.a_px.w = q3.x;
                fr.b_py.w = q3.y;
                const float px0 = (float)((o_x0 + tx) * ts), py0 = (float)((o_y0 + ty) * ts);
                hit = footprint_hits(fr, px0 + 0.5f, py0 + 0.5f, px0 + span, py0 + span);
            }
            const uint64_t bal = __ballot(hit);
            if (k < total && (lane == 0 || prev_owner != owner)) {
                // This lane starts its owner's run of the round: lanes
                // [lane, lane + run) hold tiles t .. t + run - 1 (t + run <= 64).
                const uint32_t run = min(o_end, k0 + 64u) - k;
                const uint64_t seg = run >= 64u ? ~0ull : ((1ull << run) - 1ull);
                const uint64_t bits = ((bal >> lane) & seg) << t;
                s_bits[wbase + owner][0] |= (uint32_t)bits;
                s_bits[wbase + owner][1] |= (uint32_t)(bits >> 32);
            }
        }
        __syncthreads();
        if (masked) {
            mask = (uint64_t)s_bits[tid][0] | ((uint64_t)s_bits[tid][1] << 32);
            count = (uint32_t)__popcll(mask);
        }
    }
    // Binning outputs; a record only for items that reach a tile (no entry
    // ever reads the others).
    bool write_rec = false;
    if (in_range) {
        if (!live) {
            finish(make_uint4(1u, 1u, 0u, 0u), 0u, 0u, ~0ull);
        } else {
            if (count == 0u) {
                rect = make_uint4(1u, 1u, 0u, 0u);
                mask = ~0ull;
            }
            finish(rect, __float_as_uint(depth), count, mask);
            write_rec = count != 0u;
        }
    }
    if (write_rec) {
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
    }
    // The block's items are consecutive (one view, 256 consecutive surfels):
    // stage the records in LDS and store them as one run, lane-consecutive
    // 16-B stores instead of eight 16-B stores at a 128-B stride per lane.
    {
        const int tid = threadIdx.x;
        s_write[tid] = write_rec ? 1u : 0u;
        if (write_rec) {
            float4* row = s_stage + tid * kRecRow;
            row[0] = rec.m_nt;
            row[1] = rec.a_px;
            row[2] = rec.b_py;
            row[3] = rec.col_op;
            row[4] = rec.nrm_dep;
            row[5] = rec.box;
            row[6] = rec.ell0;
            row[7] = rec.ell1;
        }
        __syncthreads();
        const uint32_t i0 = i - (uint32_t)tid;
        const uint32_t nrec = i0 < n ? min(256u, n - i0) : 0u;
        float4* dst = reinterpret_cast<float4*>(recs + (size_t)view * n + i0);
        for (uint32_t q = (uint32_t)tid; q < nrec * 8u; q += 256u) {
            const uint32_t r = q >> 3, fld = q & 7u;
            if (s_write[r] == 0u || (fld == 4u && !write_normal)) continue;
            const float4 v = s_stage[r * kRecRow + fld];
            typedef float f4v __attribute__((ext_vector_type(4)));
            __builtin_nontemporal_store(f4v{v.x, v.y, v.z, v.w}, reinterpret_cast<f4v*>(dst + q));
        }
    }
}

}  // namespace

hipError_t launch_project(hipStream_t st, const CamParams* cams_dev, uint32_t n_views, uint32_t n,
                          const float4* pos, const float2* scales, const float4* quats,
                          const float* opac, const float* sh, Record* recs, uint4* rects,
                          uint32_t* depth_keys, uint32_t* counts, uint32_t* ranges_to_clear,
                          uint32_t num_ranges, uint32_t* presort_keys, uint32_t* presort_vals,
                          uint32_t* item_count, bool clip_to_box, uint4* emit_info, bool pack_counts,
                          bool write_normal) {
    if (n == 0 && num_ranges == 0 && item_count == nullptr) return hipSuccess;
    if (n_views == 0) return hipSuccess;
    const uint32_t blocks = div_up(n > num_ranges ? (n ? n : 1) : num_ranges, 256);
    dim3 grid(blocks * n_views);
    hipLaunchKernelGGL(project_kernel, grid, dim3(256), 0, st, cams_dev, n_views, n, pos, scales, quats, opac,
                       sh, recs, rects, depth_keys, counts, ranges_to_clear, num_ranges, presort_keys,
                       presort_vals, item_count, clip_to_box, emit_info, pack_counts, write_normal);
    return hipGetLastError();
}

}  // namespace gs
