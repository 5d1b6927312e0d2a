// Stable LSD radix sort of (key, u32 payload) pairs, 64- or 32-bit keys (A6).
//
// Semantics of crates/gs-wgpu/src/sort.rs:17-327 (stable, element count read
// from a device buffer so it never round-trips to the host) but one 64-bit
// (tile || depth) key replaces the reference's two stable 32-bit sorts
// (binning.rs:5-8, 209-227), and each digit pass is a single "onesweep"
// kernel:
//   * one up-front histogram kernel counts the digits of every pass in one
//     read of the keys (LDS histograms, one global atomic per bin per block);
//   * each pass kernel takes a 4096-key partition (partition id from an
//     atomic counter), ranks its keys stably with wave64 ballots (8 ballots
//     per key at 8-bit digits), publishes its digit counts, permutes the
//     partition through LDS, resolves its global digit offsets by decoupled
//     look-back over 64-bit tagged status words (16 predecessors per round
//     trip) and writes it out in digit runs.
// Digit width is 4 or 8 bits; only the bit range [begin_bit, end_bit) is
// sorted, so a key whose high bits are known constant costs fewer passes.
//
// Per pass HBM traffic: read (8 + 4) B and write (8 + 4) B per pair for 64-bit
// keys, plus one 8-B histogram read per pair per sort.

#include "gs_internal.h"

namespace gs {

namespace {

constexpr int kSortThreads = 256;
constexpr int kSortWaves = kSortThreads / 64;
constexpr int kSortItems = 16;                       // keys per thread
constexpr uint32_t kSortPart = kSortThreads * kSortItems;  // 4096 keys per partition
constexpr int kMaxPasses = 16;
constexpr int kLookback = 16;                        // predecessor statuses per look-back round

struct SortScratch {
    uint32_t counters[kMaxPasses];   // partition counters, one per pass
    uint32_t sort_tag;               // bumped once per sort
    uint32_t pad[15];
    uint32_t hist[kMaxPasses * 256];  // global digit histograms
    // followed by uint64_t status[max_parts * 256]
};

constexpr uint64_t kFlagAgg = 1ull, kFlagPrefix = 2ull;

__device__ __forceinline__ uint64_t pack_status(uint32_t tag, uint64_t flag, uint32_t value) {
    return ((uint64_t)(tag & 0x3FFFFFFFu) << 34) | (flag << 32) | (uint64_t)value;
}

__global__ __launch_bounds__(256) void sort_init_kernel(SortScratch* sc, uint32_t num_passes,
                                                        uint32_t radix) {
    const uint32_t tid = threadIdx.x;
    for (uint32_t k = tid; k < num_passes * radix; k += blockDim.x) sc->hist[k] = 0u;
    if (tid < kMaxPasses) sc->counters[tid] = 0u;
    if (tid == 0) sc->sort_tag = sc->sort_tag + 1u;
}

template <typename K, int RB>
__global__ __launch_bounds__(256) void sort_hist_kernel(const K* __restrict__ keys,
                                                        const uint32_t* __restrict__ count,
                                                        uint32_t begin_bit, uint32_t num_passes,
                                                        SortScratch* __restrict__ sc) {
    constexpr uint32_t R = 1u << RB;
    __shared__ uint32_t s_hist[kMaxPasses * R];
    for (uint32_t k = threadIdx.x; k < num_passes * R; k += blockDim.x) s_hist[k] = 0u;
    __syncthreads();
    const uint32_t n = *count;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const K key = keys[i];
        for (uint32_t p = 0; p < num_passes; ++p) {
            const uint32_t d = (uint32_t)(key >> (begin_bit + p * RB)) & (R - 1u);
            atomicAdd(&s_hist[p * R + d], 1u);
        }
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < num_passes * R; k += blockDim.x) {
        const uint32_t c = s_hist[k];
        if (c) atomicAdd(&sc->hist[k], c);
    }
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

template <typename K, int RB>
__global__ __launch_bounds__(kSortThreads) void onesweep_kernel(
    const K* __restrict__ kin, const uint32_t* __restrict__ vin, K* __restrict__ kout,
    uint32_t* __restrict__ vout, const uint32_t* __restrict__ count, uint32_t shift,
    uint32_t pass, SortScratch* __restrict__ sc) {
    constexpr uint32_t R = 1u << RB;
    __shared__ uint32_t s_part;
    __shared__ uint32_t s_wave_hist[kSortWaves][R];   // running counts, then wave-exclusive
    __shared__ uint32_t s_digit_start[R];             // block-local exclusive digit starts
    __shared__ uint32_t s_out_base[R];                // global dest = s_out_base[d] + local pos
    __shared__ uint32_t s_tmp[kSortWaves];
    __shared__ K s_keys[kSortPart];
    __shared__ uint32_t s_vals[kSortPart];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // The grid is sized for the capacity; blocks beyond the live partition
    // count leave before touching the shared partition counter, so exactly
    // nparts blocks claim the ids 0..nparts-1 (in arrival order, which keeps
    // the look-back waiting only on blocks that are already running).
    const uint32_t n = *count;
    const uint32_t nparts = (n + kSortPart - 1) / kSortPart;
    if (blockIdx.x >= nparts) return;
    if (tid == 0) s_part = atomicAdd(&sc->counters[pass], 1u);
    for (uint32_t k = tid; k < kSortWaves * R; k += kSortThreads) (&s_wave_hist[0][0])[k] = 0u;
    __syncthreads();
    const uint32_t part = s_part;
    const uint32_t tag = sc->sort_tag * (uint32_t)kMaxPasses + pass;
    uint64_t* status = reinterpret_cast<uint64_t*>(sc + 1);

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
        uint64_t match = __ballot(ok);
#pragma unroll
        for (int b = 0; b < RB; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            match &= bit ? bal : ~bal;
        }
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
    // Global offset of this pass's digits (exclusive scan of the histogram).
    const uint32_t hist_d = (uint32_t)tid < R ? sc->hist[pass * R + tid] : 0u;
    const uint32_t global_start = block_exclusive_scan_digits<R>(hist_d, s_tmp);

    // Publish this partition's digit counts first (successors may sum them),
    // then permute the partition through LDS into digit order (only block-
    // local offsets are needed), which frees the key/value registers before
    // the look-back.
    uint64_t* my = status + (size_t)part * R + tid;
    if ((uint32_t)tid < R) {
        s_digit_start[tid] = digit_start;
        __hip_atomic_store(my, pack_status(tag, part == 0 ? kFlagPrefix : kFlagAgg, block_count),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
    if ((uint32_t)tid < R) {
        // Decoupled look-back, one thread per digit, kLookback predecessor
        // statuses per round trip (nearest first). Aggregates are summed up
        // to the first inclusive prefix; an entry not yet published ends the
        // round and the next round resumes there. Partition 0 always
        // publishes a prefix, so the walk ends; no load goes below it.
        uint32_t excl = 0;
        if (part != 0) {
            int64_t q = (int64_t)part - 1;
            const uint32_t want = tag & 0x3FFFFFFFu;
            while (true) {
                uint64_t st[kLookback];
#pragma unroll
                for (int k = 0; k < kLookback; ++k)
                    st[k] = q - k >= 0 ? __hip_atomic_load(status + (size_t)(q - k) * R + tid,
                                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                       : 0ull;
                int used = 0;
                bool found = false, stalled = false;
#pragma unroll
                for (int k = 0; k < kLookback; ++k) {
                    if (found || stalled) continue;
                    const uint64_t s = st[k];
                    const uint64_t flag = (s >> 32) & 3ull;
                    if ((uint32_t)(s >> 34) != want || flag == 0ull) {
                        stalled = true;
                        continue;
                    }
                    excl += (uint32_t)s;
                    ++used;
                    if (flag == kFlagPrefix) found = true;
                }
                if (found) break;
                q -= used;
                if (stalled) __builtin_amdgcn_s_sleep(1);
            }
            __hip_atomic_store(my, pack_status(tag, kFlagPrefix, excl + block_count),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        s_out_base[tid] = global_start + excl - digit_start;
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
                                                         const uint32_t* __restrict__ count) {
    const uint32_t n = *count;
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
    hipLaunchKernelGGL(sort_init_kernel, dim3(1), dim3(256), 0, st, sc, passes, radix);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    uint32_t hblocks = div_up(capacity ? capacity : 1, 256 * 16);
    if (hblocks > 1024) hblocks = 1024;
    if (digit_bits == 8)
        hipLaunchKernelGGL((sort_hist_kernel<K, 8>), dim3(hblocks), dim3(256), 0, st, keys0, count,
                           begin_bit, passes, sc);
    else
        hipLaunchKernelGGL((sort_hist_kernel<K, 4>), dim3(hblocks), dim3(256), 0, st, keys0, count,
                           begin_bit, passes, sc);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const uint32_t parts = div_up(capacity ? capacity : 1, kSortPart);
    K* kb[2] = {keys0, keys1};
    uint32_t* vb[2] = {vals0, vals1};
    for (uint32_t p = 0; p < passes; ++p) {
        const uint32_t shift = begin_bit + p * digit_bits;
        const int src = (int)(p & 1u), dst = 1 - src;
        if (digit_bits == 8)
            hipLaunchKernelGGL((onesweep_kernel<K, 8>), dim3(parts), dim3(kSortThreads), 0, st,
                               kb[src], vb[src], kb[dst], vb[dst], count, shift, p, sc);
        else
            hipLaunchKernelGGL((onesweep_kernel<K, 4>), dim3(parts), dim3(kSortThreads), 0, st,
                               kb[src], vb[src], kb[dst], vb[dst], count, shift, p, sc);
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
                           static_cast<uint64_t*>(dst_keys), dst_vals, count);
    else
        hipLaunchKernelGGL(copy_pairs_kernel<uint32_t>, dim3(blocks), dim3(256), 0, st,
                           static_cast<const uint32_t*>(src_keys), src_vals,
                           static_cast<uint32_t*>(dst_keys), dst_vals, count);
    return hipGetLastError();
}

size_t sort_scratch_bytes(uint32_t capacity) {
    const size_t parts = (capacity + kSortPart - 1) / kSortPart + 1;
    return sizeof(SortScratch) + parts * 256 * sizeof(uint64_t);
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
