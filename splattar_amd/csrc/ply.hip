// Baked 3DGS .ply rows from the resident activated surfel SoA.
//
// Restates first_attempt/crates/gs-io/src/ply/writer.rs:17-79 (the row
// layout of write_3dgs_ply): per surfel 59 little-endian floats
//   x y z | f_dc_0..2 = SH coefficient 0 rgb | f_rest_0..44 channel-major
//   (all red coefficients 1..15, then green, then blue) | opacity =
//   ln(o / (1 - o)) with o clamped to [1e-6, 1 - 1e-6] | scale_0, scale_1 =
//   ln(max(s, 1e-9)), scale_2 = -10 (the flattened third axis) | rot_0..3 =
//   w x y z.
// The reference reads the activated buffers back and builds the rows on the
// host; here a kernel builds them where the parameters live and only the
// packed rows cross to the host. A block of 128 surfels reads its SH rows
// (192 B each) with lane-consecutive float4 loads, builds its 128 rows
// (30,208 B, contiguous in the output) in LDS and stores them with
// lane-consecutive dword stores.

#include "gs_internal.h"

namespace gs {

namespace {

constexpr int kPlyBlock = 128;

__global__ __launch_bounds__(kPlyBlock) void ply_rows_kernel(uint32_t n, const float4* __restrict__ pos,
                                                             const float2* __restrict__ scales,
                                                             const float4* __restrict__ quats,
                                                             const float* __restrict__ opac,
                                                             const float* __restrict__ sh,
                                                             float* __restrict__ rows) {
    __shared__ float s_sh[kPlyBlock * kShFloats];
    __shared__ float s_row[kPlyBlock * kPlyRowFloats];
    const int tid = threadIdx.x;
    const uint32_t first = blockIdx.x * kPlyBlock;
    const uint32_t cnt = min((uint32_t)kPlyBlock, n - first);
    // SH rows of the block: contiguous, float4-aligned (48 floats per surfel).
    const float4* sh4 = reinterpret_cast<const float4*>(sh + (size_t)first * kShFloats);
    for (uint32_t k = tid; k < cnt * (kShFloats / 4); k += kPlyBlock)
        reinterpret_cast<float4*>(s_sh)[k] = sh4[k];
    __syncthreads();
    if ((uint32_t)tid < cnt) {
        const uint32_t i = first + tid;
        const float* c = s_sh + tid * kShFloats;
        float* r = s_row + tid * kPlyRowFloats;
        const float4 p = pos[i];
        r[0] = p.x;
        r[1] = p.y;
        r[2] = p.z;
        r[3] = c[0];
        r[4] = c[1];
        r[5] = c[2];
#pragma unroll
        for (int ch = 0; ch < 3; ++ch)
#pragma unroll
            for (int k = 1; k < 16; ++k) r[6 + ch * 15 + (k - 1)] = c[k * 3 + ch];
        const float o = fminf(fmaxf(opac[i], 1e-6f), 1.0f - 1e-6f);
        r[51] = logf(o / (1.0f - o));
        const float2 s = scales[i];
        r[52] = logf(fmaxf(s.x, 1e-9f));
        r[53] = logf(fmaxf(s.y, 1e-9f));
        r[54] = -10.0f;
        const float4 q = quats[i];
        r[55] = q.w;
        r[56] = q.x;
        r[57] = q.y;
        r[58] = q.z;
    }
    __syncthreads();
    float* out = rows + (size_t)first * kPlyRowFloats;
    for (uint32_t k = tid; k < cnt * kPlyRowFloats; k += kPlyBlock) out[k] = s_row[k];
}

}  // namespace

hipError_t launch_ply_rows(hipStream_t st, uint32_t n, const float4* pos, const float2* scales,
                           const float4* quats, const float* opac, const float* sh, float* rows) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(ply_rows_kernel, dim3((n + kPlyBlock - 1) / kPlyBlock), dim3(kPlyBlock), 0, st, n, pos,
                       scales, quats, opac, sh, rows);
    return hipGetLastError();
}

}  // namespace gs
