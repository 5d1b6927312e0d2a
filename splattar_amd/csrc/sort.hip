// Stable LSD radix sort of (key, u32 payload) pairs, 64- or 32-bit keys.
//
// Semantics of crates/gs-wgpu/src/sort.rs:17-327 (stable, element count read
// from a device buffer so it never round-trips to the host) but one 64-bit
// (tile || depth) key replaces the reference's two stable 32-bit sorts
// (binning.rs:5-8, 209-227). Each digit pass is reduce-then-scan over
// 4096-key partitions, three kernels:
//   * upsweep: per-partition digit counts (per-wave LDS histograms), written
//     digit-major;
//   * scan: one block per digit turns its row of partition counts into
//     exclusive offsets and writes the digit's total;
//   * downsweep: ranks the partition's keys stably with wave64 ballots (8
//     ballots per key at 8-bit digits), permutes them through LDS into digit
//     order and writes each digit run at global start (exclusive scan of the
//     totals) + partition offset.
// No partition waits for another (a decoupled look-back walked ~6 rounds of
// 16 predecessor statuses per partition here and read ~1.8x the keys' bytes).
// Digit width is 4 or 8 bits; only the bit range [begin_bit, end_bit) is
// sorted, so a key whose high bits are known constant costs fewer passes. A
// count above the capacity sorts the first `capacity` pairs.
//
// Per pass HBM traffic: read (2 sizeof(K) + 4) B and write (sizeof(K) + 4) B
// per pair, plus 8 B per partition and digit of offsets.

#include "gs_internal.h"

namespace gs {

namespace {

constexpr int kSortThreads = 256;
constexpr int kSortWaves = kSortThreads / 64;
constexpr int kSortItems = 16;                       // keys per thread
constexpr uint32_t kSortPart = kSortThreads * kSortItems;  // 4096 keys per partition
constexpr int kMaxPasses = 16;

struct SortScratch {
    uint32_t hist[kMaxPasses * 256];  // per pass: digit totals
    // followed by uint32_t offsets[256][parts] (digit-major partition counts,
    // then exclusive offsets)
};

__device__ __forceinline__ uint32_t* part_offsets(SortScratch* sc) {
    return reinterpret_cast<uint32_t*>(sc + 1);
}

// Wave-level match of the lanes whose digit d (RB bits) equals this lane's,
// among the lanes with ok set.
template <int RB>
__device__ __forceinline__ uint64_t digit_match(uint32_t d, bool ok) {
    uint64_t match = __ballot(ok);
#pragma unroll
    for (int b = 0; b < RB; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t bal = __ballot(bit);
        match &= bit ? bal : ~bal;
    }
    return match;
}

// Exclusive scan over R <= 256 values held one per thread (threads >= R hold
// 0); every thread gets its own exclusive prefix.
template <uint32_t R>
__device__ __forceinline__ uint32_t block_exclusive_scan_digits(uint32_t v, uint32_t* s_tmp) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    uint32_t incl = v;
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(incl, o);
        if (lane >= o) incl += t;
    }
    if (lane == 63) s_tmp[wave] = incl;
    __syncthreads();
    uint32_t add = 0;
    for (int w = 0; w < wave; ++w) add += s_tmp[w];
    __syncthreads();
    return add + incl - v;
}

// Partition `part`'s digit counts of the pass at `shift`: per-wave LDS
// histograms with plain atomics (a round whose live lanes share one digit
// adds its lane count once), summed per digit at the end.
template <typename K, int RB>
__global__ __launch_bounds__(kSortThreads) void upsweep_kernel(const K* __restrict__ kin,
                                                               const uint32_t* __restrict__ count,
                                                               uint32_t cap, uint32_t shift, uint32_t stride,
                                                               SortScratch* __restrict__ sc) {
    constexpr uint32_t R = 1u << RB;
    __shared__ uint32_t s_h[kSortWaves][R];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t n = min(*count, cap);
    const uint32_t part = blockIdx.x;
    if (part >= (n + kSortPart - 1) / kSortPart) return;  // grid sized for the capacity
    for (uint32_t k = tid; k < kSortWaves * R; k += kSortThreads) (&s_h[0][0])[k] = 0u;
    const uint32_t base = part * kSortPart;
    K key[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const uint32_t idx = base + (uint32_t)r * kSortThreads + (uint32_t)tid;
        key[r] = idx < n ? kin[idx] : (K)0;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const uint32_t idx = base + (uint32_t)r * kSortThreads + (uint32_t)tid;
        const bool ok = idx < n;
        const uint32_t d = (uint32_t)(key[r] >> shift) & (R - 1u);
        const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
        const uint64_t live = __ballot(ok);
        if (__ballot(ok && d != d0) == 0ull) {
            if (lane == 0 && live != 0ull) atomicAdd(&s_h[wave][d0], (uint32_t)__popcll(live));
        } else if (ok) {
            atomicAdd(&s_h[wave][d], 1u);
        }
    }
    __syncthreads();
    uint32_t* off = part_offsets(sc);
    for (uint32_t k = tid; k < R; k += kSortThreads) {
        uint32_t c = 0;
#pragma unroll
        for (int w = 0; w < kSortWaves; ++w) c += s_h[w][k];
        off[(size_t)k * stride + part] = c;
    }
}

// One block per digit: row of partition counts -> exclusive offsets; the
// digit's total goes to hist[pass][digit].
constexpr int kScanPer = 8;
__global__ __launch_bounds__(256) void part_scan_kernel(const uint32_t* __restrict__ count, uint32_t cap,
                                                        uint32_t stride, uint32_t pass, uint32_t radix,
                                                        SortScratch* __restrict__ sc) {
    __shared__ uint32_t s_tmp[4];
    __shared__ uint32_t s_total;
    const uint32_t d = blockIdx.x;
    const int tid = threadIdx.x;
    const uint32_t nparts = (min(*count, cap) + kSortPart - 1) / kSortPart;
    uint32_t* row = part_offsets(sc) + (size_t)d * stride;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nparts; base += 256 * kScanPer) {
        uint32_t v[kScanPer], sum = 0;
#pragma unroll
        for (int j = 0; j < kScanPer; ++j) {
            const uint32_t idx = base + (uint32_t)tid * kScanPer + j;
            v[j] = idx < nparts ? row[idx] : 0u;
            sum += v[j];
        }
        const uint32_t excl = block_exclusive_scan_digits<256>(sum, s_tmp);
        if (tid == 255) s_total = excl + sum;
        uint32_t run = carry + excl;
#pragma unroll
        for (int j = 0; j < kScanPer; ++j) {
            const uint32_t idx = base + (uint32_t)tid * kScanPer + j;
            if (idx < nparts) row[idx] = run;
            run += v[j];
        }
        __syncthreads();
        carry += s_total;
        __syncthreads();
    }
    if (tid == 0) sc->hist[pass * radix + d] = carry;
}

template <typename K, int RB>
__global__ __launch_bounds__(kSortThreads) void downsweep_kernel(
    const K* __restrict__ kin, const uint32_t* __restrict__ vin, K* __restrict__ kout,
    uint32_t* __restrict__ vout, const uint32_t* __restrict__ count, uint32_t cap, uint32_t shift,
    uint32_t stride, uint32_t pass, const SortScratch* __restrict__ sc) {
    constexpr uint32_t R = 1u << RB;
    __shared__ uint32_t s_wave_hist[kSortWaves][R];   // running counts, then wave-exclusive
    __shared__ uint32_t s_digit_start[R];             // block-local exclusive digit starts
    __shared__ uint32_t s_out_base[R];                // global dest = s_out_base[d] + local pos
    __shared__ uint32_t s_tmp[kSortWaves];
    __shared__ K s_keys[kSortPart];
    __shared__ uint32_t s_vals[kSortPart];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t n = min(*count, cap);
    const uint32_t part = blockIdx.x;
    if (part >= (n + kSortPart - 1) / kSortPart) return;  // grid sized for the capacity
    for (uint32_t k = tid; k < kSortWaves * R; k += kSortThreads) (&s_wave_hist[0][0])[k] = 0u;
    // This partition's offset of every digit and the digit totals (read
    // early; used after the ranking).
    const uint32_t part_off =
        (uint32_t)tid < R ? reinterpret_cast<const uint32_t*>(sc + 1)[(size_t)tid * stride + part] : 0u;
    const uint32_t hist_d = (uint32_t)tid < R ? sc->hist[pass * R + tid] : 0u;
    __syncthreads();

    // Load: wave w owns keys [base + 1024 w, base + 1024 (w+1)), 16 rounds of
    // 64 consecutive keys -> input order = (wave, round, lane).
    const uint32_t wbase = part * kSortPart + (uint32_t)wave * (64u * kSortItems);
    K key[kSortItems];
    uint32_t val[kSortItems];
    uint32_t rank[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const uint32_t idx = wbase + (uint32_t)r * 64u + (uint32_t)lane;
        const bool ok = idx < n;
        key[r] = ok ? kin[idx] : (K)0;
        val[r] = ok ? vin[idx] : 0u;
    }
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    // Stable wave-level ranking with ballots.
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const uint32_t idx = wbase + (uint32_t)r * 64u + (uint32_t)lane;
        const bool ok = idx < n;
        const uint32_t d = (uint32_t)(key[r] >> shift) & (R - 1u);
        const uint64_t match = digit_match<RB>(d, ok);
        const uint32_t before = (uint32_t)__popcll(match & lt_mask);
        const uint32_t total = (uint32_t)__popcll(match);
        uint32_t prev = 0;
        if (ok) prev = s_wave_hist[wave][d];
        __builtin_amdgcn_wave_barrier();
        rank[r] = prev + before;
        if (ok && before == 0) s_wave_hist[wave][d] = prev + total;
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();

    // Per digit: wave-exclusive prefixes and the block count.
    uint32_t block_count = 0;
    if ((uint32_t)tid < R) {
        uint32_t run = 0;
#pragma unroll
        for (int w = 0; w < kSortWaves; ++w) {
            const uint32_t c = s_wave_hist[w][tid];
            s_wave_hist[w][tid] = run;
            run += c;
        }
        block_count = run;
    }
    const uint32_t digit_start = block_exclusive_scan_digits<R>(block_count, s_tmp);
    const uint32_t global_start = block_exclusive_scan_digits<R>(hist_d, s_tmp);
    if ((uint32_t)tid < R) {
        s_digit_start[tid] = digit_start;
        s_out_base[tid] = global_start + part_off - digit_start;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const uint32_t idx = wbase + (uint32_t)r * 64u + (uint32_t)lane;
        if (idx < n) {
            const uint32_t d = (uint32_t)(key[r] >> shift) & (R - 1u);
            const uint32_t pos = s_digit_start[d] + s_wave_hist[wave][d] + rank[r];
            s_keys[pos] = key[r];
            s_vals[pos] = val[r];
        }
    }
    __syncthreads();
    const uint32_t valid = min(kSortPart, n - part * kSortPart);
    for (uint32_t i = tid; i < valid; i += kSortThreads) {
        const K k = s_keys[i];
        const uint32_t d = (uint32_t)(k >> shift) & (R - 1u);
        const uint32_t o = s_out_base[d] + i;
        kout[o] = k;
        vout[o] = s_vals[i];
    }
}

template <typename K>
__global__ __launch_bounds__(256) void copy_pairs_kernel(const K* __restrict__ sk,
                                                         const uint32_t* __restrict__ sv,
                                                         K* __restrict__ dk, uint32_t* __restrict__ dv,
                                                         const uint32_t* __restrict__ count, uint32_t cap) {
    const uint32_t n = min(*count, cap);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        dk[i] = sk[i];
        dv[i] = sv[i];
    }
}

template <typename K>
hipError_t radix_sort_impl(hipStream_t st, K* keys0, uint32_t* vals0, K* keys1, uint32_t* vals1,
                           const uint32_t* count, uint32_t capacity, uint32_t begin_bit,
                           uint32_t end_bit, uint32_t digit_bits, void* scratch,
                           int* result_in_alt) {
    SortScratch* sc = static_cast<SortScratch*>(scratch);
    if (digit_bits != 4 && digit_bits != 8) return hipErrorInvalidValue;
    if (end_bit <= begin_bit || end_bit > sizeof(K) * 8) return hipErrorInvalidValue;
    const uint32_t passes = (end_bit - begin_bit + digit_bits - 1) / digit_bits;
    if (passes > (uint32_t)kMaxPasses) return hipErrorInvalidValue;
    *result_in_alt = (int)(passes & 1u);
    const uint32_t radix = 1u << digit_bits;
    const uint32_t cap = capacity;
    const uint32_t parts = div_up(capacity ? capacity : 1, kSortPart);
    const uint32_t stride = parts;
    K* kb[2] = {keys0, keys1};
    uint32_t* vb[2] = {vals0, vals1};
    hipError_t e = hipSuccess;
    for (uint32_t p = 0; p < passes; ++p) {
        const uint32_t shift = begin_bit + p * digit_bits;
        const int src = (int)(p & 1u), dst = 1 - src;
        if (digit_bits == 8)
            hipLaunchKernelGGL((upsweep_kernel<K, 8>), dim3(parts), dim3(kSortThreads), 0, st, kb[src], count,
                               cap, shift, stride, sc);
        else
            hipLaunchKernelGGL((upsweep_kernel<K, 4>), dim3(parts), dim3(kSortThreads), 0, st, kb[src], count,
                               cap, shift, stride, sc);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        hipLaunchKernelGGL(part_scan_kernel, dim3(radix), dim3(256), 0, st, count, cap, stride, p, radix, sc);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if (digit_bits == 8)
            hipLaunchKernelGGL((downsweep_kernel<K, 8>), dim3(parts), dim3(kSortThreads), 0, st,
                               kb[src], vb[src], kb[dst], vb[dst], count, cap, shift, stride, p, sc);
        else
            hipLaunchKernelGGL((downsweep_kernel<K, 4>), dim3(parts), dim3(kSortThreads), 0, st,
                               kb[src], vb[src], kb[dst], vb[dst], count, cap, shift, stride, p, sc);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace

hipError_t launch_copy_pairs(hipStream_t st, const void* src_keys, const uint32_t* src_vals,
                             void* dst_keys, uint32_t* dst_vals, uint32_t key_bytes,
                             const uint32_t* count, uint32_t capacity) {
    uint32_t blocks = div_up(capacity ? capacity : 1, 256);
    if (blocks > 8192) blocks = 8192;
    if (key_bytes == 8)
        hipLaunchKernelGGL(copy_pairs_kernel<uint64_t>, dim3(blocks), dim3(256), 0, st,
                           static_cast<const uint64_t*>(src_keys), src_vals,
                           static_cast<uint64_t*>(dst_keys), dst_vals, count, capacity);
    else
        hipLaunchKernelGGL(copy_pairs_kernel<uint32_t>, dim3(blocks), dim3(256), 0, st,
                           static_cast<const uint32_t*>(src_keys), src_vals,
                           static_cast<uint32_t*>(dst_keys), dst_vals, count, capacity);
    return hipGetLastError();
}

size_t sort_scratch_bytes(uint32_t capacity) {
    const size_t parts = div_up(capacity ? capacity : 1, kSortPart);
    return sizeof(SortScratch) + parts * 256 * sizeof(uint32_t);
}

hipError_t launch_radix_sort_u64(hipStream_t st, uint64_t* keys0, uint32_t* vals0, uint64_t* keys1,
                                 uint32_t* vals1, const uint32_t* count, uint32_t capacity,
                                 uint32_t begin_bit, uint32_t end_bit, uint32_t digit_bits,
                                 void* scratch, int* result_in_alt) {
    return radix_sort_impl<uint64_t>(st, keys0, vals0, keys1, vals1, count, capacity, begin_bit,
                                     end_bit, digit_bits, scratch, result_in_alt);
}

hipError_t launch_radix_sort_u32(hipStream_t st, uint32_t* keys0, uint32_t* vals0, uint32_t* keys1,
                                 uint32_t* vals1, const uint32_t* count, uint32_t capacity,
                                 uint32_t begin_bit, uint32_t end_bit, uint32_t digit_bits,
                                 void* scratch, int* result_in_alt) {
    return radix_sort_impl<uint32_t>(st, keys0, vals0, keys1, vals1, count, capacity, begin_bit,
                                     end_bit, digit_bits, scratch, result_in_alt);
}

}  // namespace gs
