// Tile rasterizer forward (A8): one 256-thread workgroup per 16x16 tile
// (of any view in a batch), front-to-back alpha compositing.
//
// Restates crates/gs-kernels/src/shaders/rasterize_fwd.wgsl:20-85 and
// evaluate_hit of raster_common.wgsl:69-95 (the oracle's forward.rs:207-246).
// MI355X layout: the four wave64s of a workgroup each own an 8x8 quadrant of
// the tile. Surfel records are staged through LDS in batches of 256 (one
// record per thread, gathered through the sorted payloads); while staging,
// each thread tests its record's alpha >= 1/255 footprint (box, screen disk
// and ellipse, computed at projection) against the four quadrants, and the
// batch is split into four order-preserving per-quadrant lists, so a wave
// only walks the records that can touch its 64 pixels. Per pixel, an
// exponent test rejects non-contributing records before exp(). A wave leaves
// the batch once its pixels are saturated; the workgroup stops when every
// pixel's transmittance is below 1e-4.
//
// Outputs per pixel (same buffers as the reference Rasterizer):
//   color  = (sum w c, sum w),  aux = (T_final, sum w t, 0, 0),
//   normal = (sum w n, 0),      ncontrib = index (within the tile list) just
//   past the last contributing record.

#include "gs_internal.h"

namespace gs {

namespace {

constexpr int kBatch = 256;

// Locate the view owning a global tile id (views are few; tile_base ascends).
__device__ __forceinline__ uint32_t view_of_tile(const CamParams* cams, uint32_t n_views, uint32_t t) {
    uint32_t v = 0;
    for (uint32_t k = 1; k < n_views; ++k)
        if (cams[k].tile_base <= t) v = k;
    return v;
}

// AUX = false (training with colour losses only) skips the depth and normal
// accumulators: aux = (T_final, 0, 0, 0), normal is not written.
template <bool AUX>
__global__ __launch_bounds__(256) void raster_fwd_kernel(const CamParams* __restrict__ cams,
                                                         uint32_t n_views,
                                                         const uint32_t* __restrict__ ranges,
                                                         const uint32_t* __restrict__ payloads,
                                                         const Record* __restrict__ recs,
                                                         FwdOut out, bool cull, uint32_t total_tiles,
                                                         TileQueue queue) {
    __shared__ float4 s_m[kBatch], s_a[kBatch], s_b[kBatch], s_c[kBatch], s_n[kBatch];
    __shared__ float s_lthr[kBatch];
    __shared__ uint16_t s_list[4][kBatch];
    __shared__ uint32_t s_qcnt[4][4];
    __shared__ uint32_t s_view, s_next;

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
        if (tid == 0) s_view = view_of_tile(cams, n_views, gtile);
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
        const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

        const uint32_t start = ranges[2 * (size_t)gtile], end = ranges[2 * (size_t)gtile + 1];
        float T = 1.0f;
        float cr = 0.f, cg = 0.f, cb = 0.f, wsum = 0.f, dsum = 0.f, nx = 0.f, ny = 0.f, nz = 0.f;
        uint32_t last = 0;
        bool done = !inside;

        for (uint32_t base = start; base < end; base += kBatch) {
            const uint32_t cnt = min((uint32_t)kBatch, end - base);
            uint32_t qmask = 0;
            if ((uint32_t)tid < cnt) {
                Record r = recs[payloads[base + tid]];
                s_m[tid] = r.m_nt;
                s_a[tid] = r.a_px;
                s_b[tid] = r.b_py;
                s_c[tid] = r.col_op;
                if (AUX) s_n[tid] = r.nrm_dep;
                s_lthr[tid] = record_lthr(r.col_op.w);
                qmask = cull ? quadrant_mask(r, tx, ty) : 0xFu;
            }
            // Order-preserving split of the batch into per-quadrant lists.
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
            const uint32_t my_count = s_qcnt[0][wave] + s_qcnt[1][wave] + s_qcnt[2][wave] + s_qcnt[3][wave];
            __syncthreads();
            if (__any(!done)) {
                for (uint32_t k = 0; k < my_count; ++k) {
                    const uint32_t j = s_list[wave][k];
                    Record r;
                    r.m_nt = s_m[j];
                    r.a_px = s_a[j];
                    r.b_py = s_b[j];
                    r.col_op = s_c[j];
                    r.nrm_dep = AUX ? s_n[j] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    const HitEval h = eval_hit(r, sx, sy, pxc, pyc, s_lthr[j]);
                    if (!done && h.alpha > 0.0f) {
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
                        if (T < kTTerminate) done = true;
                    }
                    if (!__any(!done)) break;
                }
            }
            if (__syncthreads_count(!done) == 0) break;
        }
        if (inside) {
            const size_t pi = (size_t)cam.pix_base + (size_t)py * cam.width + (size_t)px;
            out.color[pi] = make_float4(cr, cg, cb, wsum);
            out.aux[pi] = make_float4(T, dsum, 0.0f, 0.0f);
            if (AUX) out.normal[pi] = make_float4(nx, ny, nz, 0.0f);
            out.ncontrib[pi] = last;
        }
        __syncthreads();  // s_view and the LDS batch are reused by the next tile
    }
}

}  // namespace

hipError_t launch_raster_fwd(hipStream_t st, const CamParams* cams_dev, const CamParams* /*cams_host*/,
                             uint32_t n_views, uint32_t total_tiles, const uint32_t* ranges,
                             const uint32_t* payloads, const Record* recs, FwdOut out, bool cull,
                             TileQueue queue, bool aux_outputs) {
    if (total_tiles == 0) return hipSuccess;
    const uint32_t grid = queue.cursor != nullptr ? std::min(total_tiles, std::max(queue.grid_blocks, 1u))
                                                  : total_tiles;
    if (aux_outputs)
        hipLaunchKernelGGL(raster_fwd_kernel<true>, dim3(grid), dim3(256), 0, st, cams_dev, n_views, ranges,
                           payloads, recs, out, cull, total_tiles, queue);
    else
        hipLaunchKernelGGL(raster_fwd_kernel<false>, dim3(grid), dim3(256), 0, st, cams_dev, n_views, ranges,
                           payloads, recs, out, cull, total_tiles, queue);
    return hipGetLastError();
}

}  // namespace gs
