// Exclusive prefix sum over u32: reduce-then-scan over 4096-item
// partitions.
//
// Semantics of crates/gs-wgpu/src/prefix_sum.rs:19-164 / prefix_sum.wgsl:
// 17-74 (exclusive scan in place, wrapping mod 2^32, grand total into a
// 1-element device buffer) in three kernels instead of the reference's
// multi-level Hillis-Steele recursion: partition sums; one block scans the
// sums (and writes the grand total); every partition rescans its items from
// its offset. No partition waits for another (a single-pass decoupled
// look-back spent ~0.1 us per partition in a serial prefix chain on 20M
// items). Reads the input twice: 8 B read + 4 B written per item.

#include "gs_internal.h"

namespace gs {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr uint32_t kScanPart = kScanThreads * kScanItems;  // 4096

struct ScanScratch {
    uint32_t pad[64];
    // followed by uint32_t sums[max_parts]
};
static_assert(kScanPart == kScanPartItems && sizeof(ScanScratch) == 256, "scan scratch layout");

__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v, int lane) {
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(v, o);
        if (lane >= o) v += t;
    }
    return v;
}

// This thread's 16 consecutive values from `first`. With `perm`, element i
// of the scanned sequence is in[perm[i]] (a gather); with `perm` and
// in == nullptr it is perm[i] & kPackMask (counts packed into the low bits
// of a sorted payload).
__device__ __forceinline__ void load_run(const uint32_t* in, const uint32_t* perm, uint32_t n, uint32_t first,
                                         uint32_t (&v)[kScanItems]) {
    if ((perm == nullptr || in == nullptr) && first + kScanItems <= n) {
        // Whole run in range, no gather: 16-B loads (first is a multiple of
        // 16 items).
        const uint32_t m = in != nullptr ? 0xFFFFFFFFu : kPackMask;
        const uint4* in4 = reinterpret_cast<const uint4*>((in != nullptr ? in : perm) + first);
#pragma unroll
        for (int k = 0; k < kScanItems / 4; ++k) {
            const uint4 q = in4[k];
            v[4 * k + 0] = q.x & m;
            v[4 * k + 1] = q.y & m;
            v[4 * k + 2] = q.z & m;
            v[4 * k + 3] = q.w & m;
        }
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            const uint32_t idx = first + k;
            v[k] = idx >= n ? 0u : in == nullptr ? (perm[idx] & kPackMask) : in[perm ? perm[idx] : idx];
        }
    }
}

// Sum of partition blockIdx.x. With `stage`, the (gathered) values are also
// stored there contiguously, so the second read needs no gather.
__global__ __launch_bounds__(kScanThreads) void scan_reduce_kernel(const uint32_t* in, uint32_t n,
                                                                   const uint32_t* __restrict__ perm,
                                                                   uint32_t* __restrict__ stage,
                                                                   ScanScratch* __restrict__ sc) {
    __shared__ uint32_t s_wave[kScanThreads / 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    uint32_t v[kScanItems];
    const uint32_t first = blockIdx.x * kScanPart + (uint32_t)tid * kScanItems;
    load_run(in, perm, n, first, v);
    if (stage != nullptr) {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k)
            if (first + k < n) stage[first + k] = v[k];
    }
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) sum += v[k];
    for (int o = 32; o >= 1; o >>= 1) sum += __shfl_xor(sum, o);
    if (lane == 0) s_wave[wave] = sum;
    __syncthreads();
    if (tid == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kScanThreads / 64; ++w) t += s_wave[w];
        reinterpret_cast<uint32_t*>(sc + 1)[blockIdx.x] = t;
    }
}

// One block: partition sums -> exclusive offsets (in place), grand total.
constexpr int kSumThreads = 1024, kSumPer = 8;
__global__ __launch_bounds__(kSumThreads) void scan_sums_kernel(ScanScratch* __restrict__ sc, uint32_t nparts,
                                                                uint32_t* __restrict__ total) {
    __shared__ uint32_t s_wave[kSumThreads / 64];
    __shared__ uint32_t s_total;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    uint32_t* sums = reinterpret_cast<uint32_t*>(sc + 1);
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nparts; base += kSumThreads * kSumPer) {
        uint32_t v[kSumPer], sum = 0;
#pragma unroll
        for (int j = 0; j < kSumPer; ++j) {
            const uint32_t idx = base + (uint32_t)tid * kSumPer + j;
            v[j] = idx < nparts ? sums[idx] : 0u;
            sum += v[j];
        }
        const uint32_t incl = wave_inclusive_scan(sum, lane);
        if (lane == 63) s_wave[wave] = incl;
        __syncthreads();
        uint32_t before = 0, all = 0;
        for (int w = 0; w < kSumThreads / 64; ++w) {
            const uint32_t t = s_wave[w];
            if (w < wave) before += t;
            all += t;
        }
        uint32_t run = carry + before + incl - sum;
#pragma unroll
        for (int j = 0; j < kSumPer; ++j) {
            const uint32_t idx = base + (uint32_t)tid * kSumPer + j;
            if (idx < nparts) sums[idx] = run;
            run += v[j];
        }
        carry += all;
        __syncthreads();
    }
    if (tid == 0) s_total = carry;
    __syncthreads();
    if (tid == 0) *total = s_total;
}

// `in` and `out` may alias (in-place scan, as PrefixSum::encode): every block
// reads its whole partition into registers before writing it back; with a
// gather (perm and in both set) they must not.
__global__ __launch_bounds__(kScanThreads) void scan_down_kernel(const uint32_t* in, uint32_t* out, uint32_t n,
                                                                 const uint32_t* __restrict__ perm,
                                                                 const ScanScratch* __restrict__ sc) {
    __shared__ uint32_t s_wave[kScanThreads / 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t part = blockIdx.x;
    const uint32_t first = part * kScanPart + (uint32_t)tid * kScanItems;
    const uint32_t offset = reinterpret_cast<const uint32_t*>(sc + 1)[part];
    uint32_t v[kScanItems];
    load_run(in, perm, n, first, v);
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) sum += v[k];
    const uint32_t incl = wave_inclusive_scan(sum, lane);
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    uint32_t wave_excl = 0;
#pragma unroll
    for (int w = 0; w < kScanThreads / 64; ++w)
        if (w < wave) wave_excl += s_wave[w];
    uint32_t run = offset + wave_excl + (incl - sum);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint32_t idx = first + k;
        if (idx < n) out[idx] = run;
        run += v[k];
    }
}

}  // namespace

size_t scan_scratch_bytes(uint32_t capacity) { return scan_scratch_size(capacity); }

hipError_t launch_exclusive_scan(hipStream_t st, const uint32_t* in, uint32_t* out, uint32_t n,
                                 uint32_t* total_out, void* scratch, const uint32_t* perm) {
    ScanScratch* sc = static_cast<ScanScratch*>(scratch);
    const uint32_t parts = (n + kScanPart - 1) / kScanPart;
    // A gather (in[perm[i]]; in and out distinct) runs once: the first pass
    // stages the gathered values in `out`, the second scans them in place.
    const bool gather = perm != nullptr && in != nullptr;
    hipError_t e = hipSuccess;
    if (parts > 0) {
        hipLaunchKernelGGL(scan_reduce_kernel, dim3(parts), dim3(kScanThreads), 0, st, in, n, perm,
                           gather ? out : nullptr, sc);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(kSumThreads), 0, st, sc, parts, total_out);
    if ((e = hipGetLastError()) != hipSuccess || parts == 0) return e;
    hipLaunchKernelGGL(scan_down_kernel, dim3(parts), dim3(kScanThreads), 0, st, gather ? out : in, out, n,
                       gather ? nullptr : perm, sc);
    return hipGetLastError();
}

}  // namespace gs
