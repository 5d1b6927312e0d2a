// Single-pass exclusive prefix sum over u32 (A4): decoupled look-back.
//
// Semantics of crates/gs-wgpu/src/prefix_sum.rs:19-164 / prefix_sum.wgsl:
// 17-74 (exclusive scan in place, wrapping mod 2^32, grand total into a
// 1-element device buffer) in one kernel instead of the reference's
// multi-level Hillis-Steele recursion.
//
// Partition ids come from an atomic counter, so a block only ever waits on
// partitions claimed by blocks that are already running: this is what makes
// the look-back safe without forward-progress guarantees between workgroups.
// Status words are 64-bit: [tag:30 | flag:2 | value:32]. The tag is bumped by
// scan_init once per launch, so stale words from an earlier scan read as "not
// ready" and no per-launch clearing of the status array is needed (graph
// replays stay correct).

#include "gs_internal.h"

namespace gs {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr uint32_t kScanPart = kScanThreads * kScanItems;  // 4096

struct ScanScratch {
    uint32_t counter;
    uint32_t tag;
    uint32_t pad[62];
    // followed by uint64_t status[max_parts]
};

constexpr uint64_t kFlagAgg = 1ull, kFlagPrefix = 2ull;

__device__ __forceinline__ uint64_t pack_status(uint32_t tag, uint64_t flag, uint32_t value) {
    return ((uint64_t)(tag & 0x3FFFFFFFu) << 34) | (flag << 32) | (uint64_t)value;
}

__global__ void scan_init_kernel(ScanScratch* sc, uint32_t* total) {
    sc->counter = 0u;
    sc->tag = sc->tag + 1u;
    *total = 0u;
}

__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v, int lane) {
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(v, o);
        if (lane >= o) v += t;
    }
    return v;
}

// `in` and `out` may alias (in-place scan, as PrefixSum::encode): every block
// reads its whole partition into registers before writing it back. With
// `perm`, element i of the scanned sequence is in[perm[i]] (a gather; then
// in and out must not alias); with `perm` and in == nullptr it is
// perm[i] & 0xFF (values packed into the low byte of a sorted payload).
__global__ __launch_bounds__(kScanThreads) void scan_kernel(const uint32_t* in, uint32_t* out,
                                                            uint32_t n,
                                                            uint32_t* __restrict__ total,
                                                            ScanScratch* __restrict__ sc,
                                                            const uint32_t* __restrict__ perm) {
    __shared__ uint32_t s_part;
    __shared__ uint32_t s_wave[kScanThreads / 64];
    __shared__ uint32_t s_prefix;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // Blocks beyond the partition count leave before the shared counter, so
    // exactly nparts blocks claim ids 0..nparts-1 in arrival order.
    const uint32_t nparts = (n + kScanPart - 1) / kScanPart;
    if (blockIdx.x >= nparts) return;
    if (tid == 0) s_part = atomicAdd(&sc->counter, 1u);
    __syncthreads();
    const uint32_t part = s_part;
    const uint32_t tag = sc->tag;
    uint64_t* status = reinterpret_cast<uint64_t*>(sc + 1);

    const uint32_t first = part * kScanPart + (uint32_t)tid * kScanItems;
    uint32_t v[kScanItems];
    uint32_t sum = 0;
    if ((perm == nullptr || in == nullptr) && first + kScanItems <= n) {
        // Whole run in range, no gather: 16-B loads (first is a multiple of
        // 16 items).
        const uint32_t m = in != nullptr ? 0xFFFFFFFFu : 0xFFu;
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
            // in == nullptr: the values are the low 8 bits of perm (packed).
            v[k] = idx >= n ? 0u : in == nullptr ? (perm[idx] & 0xFFu) : in[perm ? perm[idx] : idx];
        }
    }
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) sum += v[k];
    const uint32_t incl = wave_inclusive_scan(sum, lane);
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    uint32_t wave_excl = 0, block_total = 0;
#pragma unroll
    for (int w = 0; w < kScanThreads / 64; ++w) {
        const uint32_t t = s_wave[w];
        if (w < wave) wave_excl += t;
        block_total += t;
    }
    if (wave == 0) {
        // Wave-parallel decoupled look-back: lane l reads the status of
        // partition q - l. The lanes before the first unpublished entry are
        // summed up to and including the first inclusive prefix; if an entry
        // is unpublished first, the ready lanes before it are consumed and
        // the window moves there. Partition 0 always publishes a prefix, so
        // the walk ends, and no lane whose partition index is negative is
        // ever consumed.
        uint32_t excl = 0;
        if (part == 0) {
            if (lane == 0)
                __hip_atomic_store(&status[0], pack_status(tag, kFlagPrefix, block_total),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            if (lane == 0)
                __hip_atomic_store(&status[part], pack_status(tag, kFlagAgg, block_total),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const uint32_t want = tag & 0x3FFFFFFFu;
            int64_t q = (int64_t)part - 1;
            while (true) {
                const int64_t idx = q - lane;
                const uint64_t st = idx >= 0 ? __hip_atomic_load(&status[idx], __ATOMIC_RELAXED,
                                                                 __HIP_MEMORY_SCOPE_AGENT)
                                             : 0ull;
                const uint64_t flag = (st >> 32) & 3ull;
                const bool ready = (uint32_t)(st >> 34) == want && flag != 0ull;
                const uint64_t not_ready = __ballot(!ready);
                const uint64_t prefixes = __ballot(ready && flag == kFlagPrefix);
                const int first_nr = not_ready ? __builtin_ctzll(not_ready) : 64;
                const int first_pf = prefixes ? __builtin_ctzll(prefixes) : 64;
                const bool done = first_pf < first_nr;
                const int take = done ? first_pf + 1 : first_nr;
                uint32_t val = lane < take ? (uint32_t)st : 0u;
#pragma unroll
                for (int o = 32; o >= 1; o >>= 1) val += __shfl_xor(val, o);
                excl += val;
                if (done) break;
                q -= take;
                if (take < 64) __builtin_amdgcn_s_sleep(1);
            }
            if (lane == 0)
                __hip_atomic_store(&status[part], pack_status(tag, kFlagPrefix, excl + block_total),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (lane == 0) {
            s_prefix = excl;
            if (part == nparts - 1) *total = excl + block_total;
        }
    }
    __syncthreads();
    uint32_t run = s_prefix + wave_excl + (incl - sum);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint32_t idx = first + k;
        if (idx < n) out[idx] = run;
        run += v[k];
    }
}

}  // namespace

size_t scan_scratch_bytes(uint32_t capacity) {
    const uint32_t parts = (capacity + kScanPart - 1) / kScanPart + 1;
    return sizeof(ScanScratch) + (size_t)parts * sizeof(uint64_t);
}

hipError_t launch_exclusive_scan(hipStream_t st, const uint32_t* in, uint32_t* out, uint32_t n,
                                 uint32_t* total_out, void* scratch, const uint32_t* perm) {
    ScanScratch* sc = static_cast<ScanScratch*>(scratch);
    hipLaunchKernelGGL(scan_init_kernel, dim3(1), dim3(1), 0, st, sc, total_out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || n == 0) return e;
    const uint32_t parts = (n + kScanPart - 1) / kScanPart;
    hipLaunchKernelGGL(scan_kernel, dim3(parts), dim3(kScanThreads), 0, st, in, out, n, total_out, sc,
                       perm);
    return hipGetLastError();
}

}  // namespace gs
