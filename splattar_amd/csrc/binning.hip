// Tile binning: key emission and tile ranges.
//
// emit_keys restates `expand` of crates/gs-kernels/src/shaders/
// tile_binning.wgsl:46-66, with one 64-bit key per entry instead of the
// reference's separate 32-bit tile and depth keys (two stable 32-bit sorts,
// binning.rs:5-8): key = (global_tile << 32) | f32_bits(depth), payload =
// global record index (view * N + surfel). Entries are written in record
// order and row-major within each rect, so a stable sort reproduces the
// oracle's (depth bits, surfel index) tie-break (forward.rs:95). Unlike the
// reference it bounds-checks against max_entries and raises an overflow flag.
//
// tile_ranges restates tile_binning.wgsl:89-103 over the sorted 64-bit keys.
//
// The training path uses the presorted variant (emit_tile_keys): surfels are
// first stably sorted by depth key (N items, 32-bit keys), entries are
// emitted in that order with 32-bit tile ids only, and a stable sort over the
// tile bits alone finishes the job -- the same entry order, ranges and
// (reconstructible) 64-bit keys as the 64-bit sort, with 2 instead of 6
// passes over the entries.

#include "gs_internal.h"

namespace gs {

namespace {

// One wave per item-group: each lane owns one item; the wave then emits the
// union of its lanes' entries cooperatively so that stores of consecutive
// entries come from consecutive lanes (coalesced 8-B key and 4-B payload
// writes) instead of each lane streaming its own run.
__global__ __launch_bounds__(256) void emit_keys_kernel(
    uint32_t n_items, const uint4* __restrict__ rects, const uint32_t* __restrict__ depth_keys,
    const uint32_t* __restrict__ offsets, const CamParams* __restrict__ cams, uint32_t items_per_view,
    uint64_t* __restrict__ keys, uint32_t* __restrict__ payloads, uint32_t max_entries,
    const uint32_t* __restrict__ total, uint32_t* __restrict__ sort_count,
    uint32_t* __restrict__ overflow_flag) {
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid == 0) {
        const uint32_t t = *total;
        *sort_count = t < max_entries ? t : max_entries;
        if (t > max_entries) atomicOr(overflow_flag, 1u);
    }
    const int lane = threadIdx.x & 63;
    const uint32_t wave_first = gid - (uint32_t)lane;
    if (wave_first >= n_items) return;

    uint32_t cnt = 0, base = 0, w = 0, x0 = 0, y0 = 0, tiles_x = 0, tile_base = 0, dk = 0;
    if (gid < n_items) {
        const uint4 r = rects[gid];
        if (r.z >= r.x && r.w >= r.y) {
            w = r.z - r.x + 1u;
            cnt = w * (r.w - r.y + 1u);
            x0 = r.x;
            y0 = r.y;
            base = offsets[gid];
            dk = depth_keys[gid];
            const CamParams& cam = cams[gid / items_per_view];
            tiles_x = cam.tiles_x;
            tile_base = cam.tile_base;
        }
    }
    // Inclusive prefix of counts across the wave (ordering within the wave
    // follows item order, matching the scan's offsets).
    uint32_t incl = cnt;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t v = __shfl_up(incl, o);
        if (lane >= o) incl += v;
    }
    const uint32_t wave_total = __shfl(incl, 63);
    const uint32_t excl = incl - cnt;
    // Cooperative emission: entry k of the wave's run belongs to the lane
    // whose [excl, incl) interval holds k.
    for (uint32_t k0 = 0; k0 < wave_total; k0 += 64) {
        const uint32_t k = k0 + (uint32_t)lane;
        // Find the owner lane: largest lane with excl <= k (binary search over
        // the wave via shuffles).
        int owner = 0;
        for (int step = 32; step >= 1; step >>= 1) {
            const int cand = owner + step;
            const uint32_t ex = __shfl(excl, cand & 63);
            if (cand < 64 && ex <= k) owner = cand;
        }
        const uint32_t o_excl = __shfl(excl, owner);
        const uint32_t o_cnt = __shfl(cnt, owner);
        const uint32_t o_base = __shfl(base, owner);
        const uint32_t o_w = __shfl(w, owner);
        const uint32_t o_x0 = __shfl(x0, owner);
        const uint32_t o_y0 = __shfl(y0, owner);
        const uint32_t o_tx = __shfl(tiles_x, owner);
        const uint32_t o_tb = __shfl(tile_base, owner);
        const uint32_t o_dk = __shfl(dk, owner);
        if (k < wave_total && o_cnt > 0) {
            const uint32_t local = k - o_excl;
            const uint32_t e = o_base + local;
            if (e < max_entries) {
                const uint32_t ty = o_y0 + local / o_w;
                const uint32_t tx = o_x0 + local % o_w;
                const uint64_t tile = (uint64_t)(o_tb + ty * o_tx + tx);
                keys[e] = (tile << 32) | (uint64_t)o_dk;
                payloads[e] = wave_first + (uint32_t)owner;
            }
        }
    }
}

// Depth-presorted emission: items are visited in the order of a stable sort
// of their depth keys (order[k] = item), so entries come out ordered by
// (depth bits, record index) and a stable sort on the tile id alone yields
// exactly the (tile, depth, index) order of the 64-bit path. Keys are the
// 32-bit global tile ids; payloads the record indices.
// Position of the k-th (0-based) set bit of the 64-bit mask (hi:lo); k is
// below the mask's population count.
__device__ __forceinline__ uint32_t select_set_bit(uint32_t lo, uint32_t hi, uint32_t k) {
    uint32_t pos = 0;
    uint32_t v = lo;
    const uint32_t plo = (uint32_t)__popc(lo);
    if (k >= plo) {
        k -= plo;
        v = hi;
        pos = 32;
    }
#pragma unroll
    for (uint32_t width = 16; width >= 1; width >>= 1) {
        const uint32_t low = v & ((1u << width) - 1u);
        const uint32_t c = (uint32_t)__popc(low);
        if (k >= c) {
            k -= c;
            v >>= width;
            pos += width;
        } else {
            v = low;
        }
    }
    return pos;
}

__global__ __launch_bounds__(256) void emit_tile_keys_kernel(
    uint32_t n_items, const uint32_t* __restrict__ order, bool packed, const uint4* __restrict__ emit_info,
    const uint32_t* __restrict__ offsets_sorted, const CamParams* __restrict__ cams,
    uint32_t items_per_view, uint32_t* __restrict__ tkeys, uint32_t* __restrict__ payloads,
    uint32_t max_entries, const uint32_t* __restrict__ total, uint32_t* __restrict__ sort_count,
    uint32_t* __restrict__ overflow_flag) {
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid == 0) {
        const uint32_t t = *total;
        *sort_count = t < max_entries ? t : max_entries;
        if (t > max_entries) atomicOr(overflow_flag, 1u);
    }
    const int lane = threadIdx.x & 63;
    const uint32_t wave_first = gid - (uint32_t)lane;
    if (wave_first >= n_items) return;

    uint32_t cnt = 0, base = 0, w = 0, x0 = 0, y0 = 0, tiles_x = 0, tile_base = 0, item = 0;
    uint64_t mask = 0;
    bool masked = false;
    if (gid < n_items) {
        item = packed ? order[gid] >> kPackBits : order[gid];
        const uint4 ei = emit_info[item];
        const uint32_t rx0 = ei.x & 0xFFFFu, ry0 = ei.x >> 16, rx1 = ei.y & 0xFFFFu, ry1 = ei.y >> 16;
        if (rx1 >= rx0 && ry1 >= ry0) {
            w = rx1 - rx0 + 1u;
            cnt = w * (ry1 - ry0 + 1u);
            // Tight binning: only the rect tiles whose mask bit is set (an
            // all-ones mask means every tile; project wrote the same count).
            mask = (uint64_t)ei.z | ((uint64_t)ei.w << 32);
            if (mask != ~0ull) {
                masked = true;
                cnt = (uint32_t)__popcll(mask);
            }
            x0 = rx0;
            y0 = ry0;
            base = offsets_sorted[gid];
            const CamParams& cam = cams[item / items_per_view];
            tiles_x = cam.tiles_x;
            tile_base = cam.tile_base;
        }
    }
    uint32_t incl = cnt;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t v = __shfl_up(incl, o);
        if (lane >= o) incl += v;
    }
    const uint32_t wave_total = __shfl(incl, 63);
    const uint32_t excl = incl - cnt;
    for (uint32_t k0 = 0; k0 < wave_total; k0 += 64) {
        const uint32_t k = k0 + (uint32_t)lane;
        int owner = 0;
        for (int step = 32; step >= 1; step >>= 1) {
            const int cand = owner + step;
            const uint32_t ex = __shfl(excl, cand & 63);
            if (cand < 64 && ex <= k) owner = cand;
        }
        const uint32_t o_excl = __shfl(excl, owner);
        const uint32_t o_cnt = __shfl(cnt, owner);
        const uint32_t o_base = __shfl(base, owner);
        const uint32_t o_w = __shfl(w, owner);
        const uint32_t o_x0 = __shfl(x0, owner);
        const uint32_t o_y0 = __shfl(y0, owner);
        const uint32_t o_tx = __shfl(tiles_x, owner);
        const uint32_t o_tb = __shfl(tile_base, owner);
        const uint32_t o_item = __shfl(item, owner);
        const uint32_t o_mlo = __shfl((uint32_t)mask, owner);
        const uint32_t o_mhi = __shfl((uint32_t)(mask >> 32), owner);
        const bool o_masked = __shfl((int)masked, owner) != 0;
        if (k < wave_total && o_cnt > 0) {
            uint32_t local = k - o_excl;
            const uint32_t e = o_base + local;
            if (o_masked) local = select_set_bit(o_mlo, o_mhi, local);
            if (e < max_entries) {
                const uint32_t ty = o_y0 + local / o_w;
                const uint32_t tx = o_x0 + local % o_w;
                tkeys[e] = o_tb + ty * o_tx + tx;
                payloads[e] = o_item;
            }
        }
    }
}

__global__ __launch_bounds__(256) void tile_order_keys_kernel(const uint32_t* __restrict__ ranges,
                                                              uint32_t tile_begin, uint32_t total_tiles,
                                                              uint32_t* __restrict__ keys,
                                                              uint32_t* __restrict__ vals,
                                                              uint32_t* __restrict__ count,
                                                              uint32_t* __restrict__ cursors) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) *count = total_tiles;
    if (t < 2 * kTileQueues) cursors[t] = 0u;
    if (t >= total_tiles) return;
    const size_t g = (size_t)tile_begin + t;
    const uint32_t len = ranges[2 * g + 1] - ranges[2 * g];
    keys[t] = 0xFFFFu - min(len, 0xFFFFu);
    vals[t] = (uint32_t)g;
}

// Four sorted keys per thread (one 16-B load, keys 16-B aligned) and their
// two neighbours: a key that differs from its predecessor opens its tile's
// range, one that differs from its successor closes it.
__global__ __launch_bounds__(256) void tile_ranges_u32_kernel(const uint32_t* __restrict__ keys,
                                                              const uint32_t* __restrict__ count,
                                                              uint32_t* __restrict__ ranges) {
    const uint32_t n = *count;
    const uint32_t nq = (n + 3u) / 4u;
    for (uint32_t qi = blockIdx.x * blockDim.x + threadIdx.x; qi < nq; qi += gridDim.x * blockDim.x) {
        const uint32_t i0 = qi * 4u;
        uint32_t k[6];
        if (i0 + 4u <= n) {
            const uint4 v = reinterpret_cast<const uint4*>(keys)[qi];
            k[1] = v.x;
            k[2] = v.y;
            k[3] = v.z;
            k[4] = v.w;
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) k[1 + j] = i0 + j < n ? keys[i0 + j] : 0u;
        }
        k[0] = i0 > 0u ? keys[i0 - 1u] : 0u;
        k[5] = i0 + 4u < n ? keys[i0 + 4u] : 0u;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t i = i0 + (uint32_t)j;
            if (i >= n) break;
            const uint32_t tile = k[1 + j];
            if (i == 0u || k[j] != tile) ranges[2 * (size_t)tile] = i;
            if (i == n - 1u || k[2 + j] != tile) ranges[2 * (size_t)tile + 1] = i + 1u;
        }
    }
}

// 64-bit (tile << 32 | depth bits) keys of the presorted path, for parity
// checks and tooling.
__global__ __launch_bounds__(256) void compose_keys_kernel(const uint32_t* __restrict__ tkeys,
                                                           const uint32_t* __restrict__ payloads,
                                                           const uint32_t* __restrict__ depth_keys,
                                                           const uint32_t* __restrict__ count,
                                                           uint64_t* __restrict__ keys) {
    const uint32_t n = *count;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        keys[i] = ((uint64_t)tkeys[i] << 32) | (uint64_t)depth_keys[payloads[i]];
}

__global__ __launch_bounds__(256) void tile_ranges_kernel(const uint64_t* __restrict__ keys,
                                                          const uint32_t* __restrict__ count,
                                                          uint32_t* __restrict__ ranges) {
    const uint32_t n = *count;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t tile = (uint32_t)(keys[i] >> 32);
        if (i == 0 || (uint32_t)(keys[i - 1] >> 32) != tile) ranges[2 * (size_t)tile] = i;
        if (i == n - 1 || (uint32_t)(keys[i + 1] >> 32) != tile) ranges[2 * (size_t)tile + 1] = i + 1;
    }
}

// count_tiles of tile_binning.wgsl:30-43 (standalone binner; the raster path
// fuses it into project).
__global__ __launch_bounds__(256) void count_tiles_kernel(uint32_t n, const uint4* __restrict__ rects,
                                                          uint32_t* __restrict__ counts) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint4 r = rects[i];
    counts[i] = (r.z >= r.x && r.w >= r.y) ? (r.z - r.x + 1u) * (r.w - r.y + 1u) : 0u;
}

__global__ __launch_bounds__(256) void fill_u32_kernel(uint32_t* __restrict__ dst, uint32_t value,
                                                       size_t count) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = value;
}

__global__ __launch_bounds__(256) void copy_u32_kernel(uint32_t* __restrict__ dst,
                                                       const uint32_t* __restrict__ src, size_t count) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

}  // namespace

hipError_t launch_emit_keys(hipStream_t st, uint32_t n_items, const uint4* rects,
                            const uint32_t* depth_keys, const uint32_t* offsets,
                            const CamParams* cams_dev, uint32_t items_per_view,
                            uint32_t /*tiles_x_fixed*/, uint64_t* keys, uint32_t* payloads,
                            uint32_t max_entries, const uint32_t* total, uint32_t* sort_count,
                            uint32_t* overflow_flag) {
    uint32_t blocks = div_up(n_items > 0 ? n_items : 1, 256);
    hipLaunchKernelGGL(emit_keys_kernel, dim3(blocks), dim3(256), 0, st, n_items, rects, depth_keys,
                       offsets, cams_dev, items_per_view > 0 ? items_per_view : 1, keys, payloads,
                       max_entries, total, sort_count, overflow_flag);
    return hipGetLastError();
}

hipError_t launch_tile_ranges(hipStream_t st, const uint64_t* keys, const uint32_t* count,
                              uint32_t capacity, uint32_t* ranges) {
    uint32_t blocks = div_up(capacity > 0 ? capacity : 1, 256);
    if (blocks > 8192) blocks = 8192;  // grid-stride beyond this
    hipLaunchKernelGGL(tile_ranges_kernel, dim3(blocks), dim3(256), 0, st, keys, count, ranges);
    return hipGetLastError();
}

hipError_t launch_emit_tile_keys(hipStream_t st, uint32_t n_items, const uint32_t* order, bool packed,
                                 const uint4* emit_info, const uint32_t* offsets_sorted,
                                 const CamParams* cams_dev, uint32_t items_per_view, uint32_t* tkeys,
                                 uint32_t* payloads, uint32_t max_entries, const uint32_t* total,
                                 uint32_t* sort_count, uint32_t* overflow_flag) {
    uint32_t blocks = div_up(n_items > 0 ? n_items : 1, 256);
    hipLaunchKernelGGL(emit_tile_keys_kernel, dim3(blocks), dim3(256), 0, st, n_items, order, packed,
                       emit_info, offsets_sorted, cams_dev, items_per_view > 0 ? items_per_view : 1, tkeys,
                       payloads, max_entries, total, sort_count, overflow_flag);
    return hipGetLastError();
}

hipError_t launch_tile_order_keys(hipStream_t st, const uint32_t* ranges, uint32_t tile_begin,
                                  uint32_t total_tiles, uint32_t* keys, uint32_t* vals, uint32_t* count,
                                  uint32_t* cursors) {
    const uint32_t blocks = div_up(total_tiles > 0 ? total_tiles : 1, 256);
    hipLaunchKernelGGL(tile_order_keys_kernel, dim3(blocks), dim3(256), 0, st, ranges, tile_begin, total_tiles,
                       keys, vals, count, cursors);
    return hipGetLastError();
}

hipError_t launch_tile_ranges_u32(hipStream_t st, const uint32_t* keys, const uint32_t* count,
                                  uint32_t capacity, uint32_t* ranges) {
    if (reinterpret_cast<uintptr_t>(keys) & 15u) return hipErrorInvalidValue;
    uint32_t blocks = div_up(capacity > 0 ? capacity : 1, 4 * 256);
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(tile_ranges_u32_kernel, dim3(blocks), dim3(256), 0, st, keys, count, ranges);
    return hipGetLastError();
}

hipError_t launch_compose_keys(hipStream_t st, const uint32_t* tkeys, const uint32_t* payloads,
                               const uint32_t* depth_keys, const uint32_t* count, uint32_t capacity,
                               uint64_t* keys) {
    uint32_t blocks = div_up(capacity > 0 ? capacity : 1, 256);
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(compose_keys_kernel, dim3(blocks), dim3(256), 0, st, tkeys, payloads, depth_keys,
                       count, keys);
    return hipGetLastError();
}

hipError_t launch_count_tiles(hipStream_t st, uint32_t n, const uint4* rects, uint32_t* counts) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(count_tiles_kernel, dim3(div_up(n, 256)), dim3(256), 0, st, n, rects, counts);
    return hipGetLastError();
}

hipError_t launch_copy_u32(hipStream_t st, uint32_t* dst, const uint32_t* src, size_t count) {
    if (count == 0) return hipSuccess;
    uint64_t blocks = (count + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(copy_u32_kernel, dim3((uint32_t)blocks), dim3(256), 0, st, dst, src, count);
    return hipGetLastError();
}

hipError_t launch_fill_u32(hipStream_t st, uint32_t* dst, uint32_t value, size_t count) {
    if (count == 0) return hipSuccess;
    uint64_t blocks = (count + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(fill_u32_kernel, dim3((uint32_t)blocks), dim3(256), 0, st, dst, value, count);
    return hipGetLastError();
}

}  // namespace gs
