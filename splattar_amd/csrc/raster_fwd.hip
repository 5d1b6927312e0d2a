// Tile rasterizer forward: one workgroup per work tile (of any view in
// a batch), front-to-back alpha compositing.
//
// Restates crates/gs-kernels/src/shaders/rasterize_fwd.wgsl:20-85 and
// evaluate_hit of raster_common.wgsl:69-95 (the oracle's forward.rs:207-246).
// MI355X layout: a TW x TW work tile (TW = 8, 16 or 32) is (TW/8)^2 wave64s,
// each owning one 8x8 pixel block. Surfel records are staged through LDS in
// batches of BATCH (gathered through the sorted payloads); while staging,
// each record's alpha >= 1/255 footprint (box clipped to its 16-px rect,
// screen disk and ellipse, computed at projection) is tested against the
// tile's blocks, and the batch is split into order-preserving per-block
// lists, so a wave only walks the records that can touch its 64 pixels. Per
// pixel, an exponent test rejects non-contributing records before exp(). A
// wave leaves the batch once its pixels are saturated; the workgroup stops
// when every pixel's transmittance is below 1e-4.
//
// Outputs per pixel (same buffers as the reference Rasterizer):
//   color  = (sum w c, sum w),  aux = (T_final, sum w t, 0, 0),
//   normal = (sum w n, 0),      ncontrib = index (within the tile list) just
//   past the last contributing record.

#include "gs_internal.h"

namespace gs {

namespace {

// Locate the view owning a global tile id (views are few; tile_base ascends).
__device__ __forceinline__ uint32_t view_of_tile(const CamParams* cams, uint32_t n_views, uint32_t t) {
    uint32_t v = 0;
    for (uint32_t k = 1; k < n_views; ++k)
        if (cams[k].tile_base <= t) v = k;
    return v;
}

// AUX = false (training with colour losses only) skips the depth and normal
// accumulators: aux = (T_final, 0, 0, 0), normal is not written. TRACE
// additionally records every contribution of the trace window's pixels
// (FwdTrace; a verification variant, never used by training steps).
// The colour-only 16-px variants are held to seven waves per SIMD (at most
// 72 VGPRs; five spilled registers, stored before the tile loop and
// reloaded once per tile, none in the batch or record loops): more
// workgroups per CU (20 KB of LDS each) hide more of the list walk's LDS
// latency. At eight the spills reach the batch loop.
template <bool AUX, int TW, int BATCH, bool TRACE = false>
__global__ __launch_bounds__((TW / 8) * (TW / 8) * 64)
__attribute__((amdgpu_waves_per_eu((!AUX && TW == 16) ? 7 : 1))) void raster_fwd_kernel(
    const CamParams* __restrict__ cams, uint32_t n_views, const uint32_t* __restrict__ ranges,
    const uint32_t* __restrict__ payloads, const Record* __restrict__ recs, FwdOut out, bool cull,
    uint32_t total_tiles, TileQueue queue, uint32_t tile_begin, FwdTrace trace) {
    constexpr int NC = TW / 8;             // blocks per tile row
    constexpr int NB = NC * NC;            // blocks (= waves) per tile
    constexpr int NT = NB * 64;            // threads
    constexpr int SW = BATCH / 64;         // staging waves per batch
    constexpr int SPT = (BATCH + NT - 1) / NT;  // batch slots per thread
    static_assert(BATCH % 64 == 0 && BATCH <= 1024, "batch of whole waves");
    __shared__ float4 s_m[BATCH], s_a[BATCH], s_b[BATCH], s_c[BATCH], s_n[AUX ? BATCH : 1];
    __shared__ float s_lthr[BATCH];
    __shared__ uint16_t s_list[NB][BATCH];
    __shared__ uint32_t s_qcnt[SW][NB];
    __shared__ uint32_t s_view, s_next;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t all_blocks = NB == 32 ? ~0u : (1u << NB) - 1u;
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
        if (tid == 0) s_view = view_of_tile(cams, n_views, gtile);
        __syncthreads();
        const CamParams& cam = cams[s_view];
        const uint32_t local = gtile - cam.tile_base;
        const uint32_t tx = local % cam.tiles_x, ty = local / cam.tiles_x;
        const uint32_t x0 = tx * TW, y0 = ty * TW;
        const int px = (int)x0 + (wave % NC) * 8 + (lane & 7);
        const int py = (int)y0 + (wave / NC) * 8 + (lane >> 3);
        const bool inside = px < (int)cam.width && py < (int)cam.height;
        float sx, sy;
        pixel_s(cam, px, py, sx, sy);
        const float pxc = (float)px + 0.5f, pyc = (float)py + 0.5f;
        const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

        const uint32_t start = ranges[2 * (size_t)gtile], end = ranges[2 * (size_t)gtile + 1];
        // A pixel is done once T < kTTerminate (forward.rs stops there); T
        // only falls, so the test on T itself is the done flag, and pixels
        // outside the image start done (T = 0; nothing of theirs is written).
        float T = inside ? 1.0f : 0.0f;
        float cr = 0.f, cg = 0.f, cb = 0.f, wsum = 0.f, dsum = 0.f, nx = 0.f, ny = 0.f, nz = 0.f;
        uint32_t last = 0;
        // Trace window slot of this pixel (TRACE only).
        const bool traced = TRACE && inside && s_view == trace.view && (uint32_t)px >= trace.x0 &&
                            (uint32_t)px < trace.x0 + trace.w && (uint32_t)py >= trace.y0 &&
                            (uint32_t)py < trace.y0 + trace.h;
        const size_t tslot = traced ? (size_t)(py - trace.y0) * trace.w + (size_t)(px - trace.x0) : 0;
        uint32_t tcount = 0;

        for (uint32_t base = start; base < end; base += BATCH) {
            const uint32_t cnt = min((uint32_t)BATCH, end - base);
            uint32_t qm[SPT];
            uint32_t qpos[SPT][NB];
#pragma unroll
            for (int m = 0; m < SPT; ++m) {
                const int slot = tid + m * NT;
                qm[m] = 0;
                if (slot >= BATCH) continue;  // whole waves (NT, BATCH multiples of 64)
                if ((uint32_t)slot < cnt) {
                    Record r = recs[payloads[base + slot]];
                    s_m[slot] = r.m_nt;
                    s_a[slot] = r.a_px;
                    s_b[slot] = r.b_py;
                    s_c[slot] = r.col_op;
                    if (AUX) s_n[slot] = r.nrm_dep;
                    s_lthr[slot] = record_lthr(r.col_op.w);
                    qm[m] = cull ? block_mask<TW>(r, x0, y0) : all_blocks;
                }
                // Order-preserving split of the batch into per-block lists.
                const int sw = slot >> 6;
#pragma unroll
                for (int q = 0; q < NB; ++q) {
                    const uint64_t bal = __ballot((qm[m] >> q) & 1u);
                    qpos[m][q] = (uint32_t)__popcll(bal & lt_mask);
                    if (lane == 0) s_qcnt[sw][q] = (uint32_t)__popcll(bal);
                }
            }
            __syncthreads();
#pragma unroll
            for (int m = 0; m < SPT; ++m) {
                const int slot = tid + m * NT;
                if (slot >= BATCH) continue;
                const int sw = slot >> 6;
#pragma unroll
                for (int q = 0; q < NB; ++q) {
                    if ((qm[m] >> q) & 1u) {
                        uint32_t off = 0;
                        for (int w = 0; w < sw; ++w) off += s_qcnt[w][q];
                        s_list[q][off + qpos[m][q]] = (uint16_t)slot;
                    }
                }
            }
            uint32_t my_count = 0;
#pragma unroll
            for (int w = 0; w < SW; ++w) my_count += s_qcnt[w][wave];
            __syncthreads();
            if (__any(T >= kTTerminate)) {
                auto load = [&](uint32_t j, Record& r, float& lthr) {
                    r.m_nt = s_m[j];
                    r.a_px = s_a[j];
                    r.b_py = s_b[j];
                    r.col_op = s_c[j];
                    r.nrm_dep = AUX ? s_n[AUX ? j : 0] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    lthr = s_lthr[j];
                };
                auto composite = [&](uint32_t j, const Record& r, float lthr) {
                    const HitEval h = eval_hit(r, sx, sy, pxc, pyc, lthr);
                    if (T >= kTTerminate && h.alpha > 0.0f) {
                        const float w = T * h.alpha;
                        cr += r.col_op.x * w;
                        cg += r.col_op.y * w;
                        cb += r.col_op.z * w;
                        wsum += w;
                        if (AUX) {
                            dsum += h.t * w;
                            nx += r.nrm_dep.x * w;
                            ny += r.nrm_dep.y * w;
                            nz += r.nrm_dep.z * w;
                        }
                        T *= 1.0f - h.alpha;
                        last = base - start + j + 1u;
                        if (TRACE && traced) {
                            if (tcount < trace.cap)
                                trace.lists[tslot * trace.cap + tcount] =
                                    (payloads[base + j] & kTraceRecordMask) | (h.ray ? 1u << 30 : 0u) |
                                    (h.clamped ? 1u << 31 : 0u);
                            ++tcount;
                        }
                    }
                };
                // Two records per round: both list entries and both records
                // are read from LDS together (one wait instead of two
                // dependent round trips each), then composited in list order
                // with the same per-record early exit.
                for (uint32_t k = 0; k < my_count; k += 2) {
                    const bool two = k + 1 < my_count;
                    const uint32_t j0 = s_list[wave][k];
                    const uint32_t j1 = two ? (uint32_t)s_list[wave][k + 1] : j0;
                    Record r0, r1;
                    float l0, l1;
                    load(j0, r0, l0);
                    load(j1, r1, l1);
                    // One saturation test per round: a wave done after j0
                    // walks j1 with every lane masked off (no contribution).
                    composite(j0, r0, l0);
                    if (!two) break;
                    composite(j1, r1, l1);
                    if (!__any(T >= kTTerminate)) break;
                }
            }
            if (__syncthreads_count(T >= kTTerminate) == 0) break;
        }
        if (inside) {
            const size_t pi = (size_t)cam.pix_base + (size_t)py * cam.width + (size_t)px;
            out.color[pi] = make_float4(cr, cg, cb, wsum);
            out.aux[pi] = make_float4(T, dsum, 0.0f, 0.0f);
            if (AUX) out.normal[pi] = make_float4(nx, ny, nz, 0.0f);
            out.ncontrib[pi] = last;
            if (TRACE && traced) trace.counts[tslot] = tcount;
        }
        __syncthreads();  // s_view and the LDS batch are reused by the next tile
    }
}

template <bool AUX, int TW, int BATCH, bool TRACE = false>
void launch_fwd(hipStream_t st, uint32_t grid, const CamParams* cams_dev, uint32_t n_views,
                uint32_t total_tiles, const uint32_t* ranges, const uint32_t* payloads, const Record* recs,
                FwdOut out, bool cull, TileQueue queue, uint32_t tile_begin, FwdTrace trace) {
    constexpr int NT = (TW / 8) * (TW / 8) * 64;
    hipLaunchKernelGGL((raster_fwd_kernel<AUX, TW, BATCH, TRACE>), dim3(grid), dim3(NT), 0, st, cams_dev,
                       n_views, ranges, payloads, recs, out, cull, total_tiles, queue, tile_begin, trace);
}

}  // namespace

hipError_t launch_raster_fwd(hipStream_t st, const CamParams* cams_dev, const CamParams* /*cams_host*/,
                             uint32_t n_views, uint32_t total_tiles, const uint32_t* ranges,
                             const uint32_t* payloads, const Record* recs, FwdOut out, bool cull,
                             TileQueue queue, bool aux_outputs, RasterShape shape, uint32_t tile_begin,
                             FwdTrace trace) {
    if (total_tiles == 0) return hipSuccess;
    const uint32_t grid = queue.cursor != nullptr ? std::min(total_tiles, std::max(queue.grid_blocks, 1u))
                                                  : total_tiles;
    const uint32_t t = shape.tile, b = shape.fwd_batch;
    auto go = [&](auto fn) {
        fn(st, grid, cams_dev, n_views, total_tiles, ranges, payloads, recs, out, cull, queue, tile_begin, trace);
    };
    if (trace.lists != nullptr) {
        // The trace variant: depth / normal outputs, 16-px tiles, 256-record batches.
        if (!aux_outputs || t != 16 || b != 256 || trace.counts == nullptr) return hipErrorInvalidValue;
        go(launch_fwd<true, 16, 256, true>);
    } else if (aux_outputs) {
        if (t != 16 || b != 256) return hipErrorInvalidValue;  // depth / normal outputs: 16-px tiles
        go(launch_fwd<true, 16, 256>);
    } else if (t == 8 && b == 64) {
        go(launch_fwd<false, 8, 64>);
    } else if (t == 8 && b == 128) {
        go(launch_fwd<false, 8, 128>);
    } else if (t == 16 && b == 128) {
        go(launch_fwd<false, 16, 128>);
    } else if (t == 16 && b == 256) {
        go(launch_fwd<false, 16, 256>);
    } else if (t == 32 && b == 256) {
        go(launch_fwd<false, 32, 256>);
    } else if (t == 32 && b == 512) {
        go(launch_fwd<false, 32, 512>);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace gs
