// MCMC densification kernels on the optimizer's raw parameters and moments.
//
// add_noise restates first_attempt/crates/gs-train/src/shaders/mcmc.wgsl:
// 25-58: raw positions receive near-Gaussian noise (sum of four PCG uniforms
// - 2) of amplitude sigma * mean activated scale * exp(-k * activated
// opacity). In the trainer, sigma = mcmc_noise * the position learning rate
// of the step and seed = iter * 0x9e3779b9 (trainer.rs:1256-1272); both are
// read from the device-resident Adam state, so a captured graph step keeps
// drawing fresh noise with the decayed rate.
//
// relocate restates relocate.wgsl:39-125 (relocate_params +
// relocate_moments) as one kernel over the (dead, alive, mode) pairs the
// host samples: the alive surfel's parameters move onto the dead slot (mode
// 0: co-located, log-scales - 0.16; mode 1: split of an oversized surfel,
// the halves at +-0.6 of its major axis, log-scales - 0.51), both get the
// split opacity 1 - sqrt(1 - alpha) in logit form, and the Adam moments of
// both indices are zeroed. Alive indices are unique within a batch and
// disjoint from the dead ones, so pairs never race.

#include "gs_internal.h"

namespace gs {

namespace {

// mcmc.wgsl:25-45.
__device__ __forceinline__ uint32_t pcg_hash(uint32_t v) {
    const uint32_t state = v * 747796405u + 2891336453u;
    const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
__device__ __forceinline__ float unit_uniform(uint32_t v) { return (float)(v & 0x00ffffffu) / 16777216.0f; }
__device__ __forceinline__ float noise1(uint32_t seed) {
    float acc = 0.0f;
    uint32_t s = seed;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        s = pcg_hash(s);
        acc += unit_uniform(s);
    }
    return acc - 2.0f;
}

__global__ __launch_bounds__(256) void add_noise_kernel(NoiseArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    uint32_t seed = a.seed;
    float sigma = a.sigma;
    if (a.state != nullptr) {
        // The step just taken: Adam's t is 1-based, the reference's iter 0-based.
        seed = (a.state->t - 1u) * 0x9e3779b9u;
        sigma = a.noise_scale * a.state->lr[a.lr_class];
    }
    const float2 s = a.scales_act[i];
    const float mean_scale = 0.5f * (s.x + s.y);
    const float gate = expf(-a.opacity_k * a.opac_act[i]);
    const float amp = sigma * mean_scale * gate;
    if (!(amp > 0.0f)) return;
    const uint32_t base = pcg_hash(i ^ seed);
    float4 p = a.pos_raw[i];
    p.x += amp * noise1(base ^ 0x68bc21ebu);
    p.y += amp * noise1(base ^ 0x2c1b3c6du);
    p.z += amp * noise1(base ^ 0x1f123bb5u);
    a.pos_raw[i] = p;
}

__global__ __launch_bounds__(256) void relocate_kernel(RelocateArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n_pairs) return;
    const uint4 pr = a.pairs[i];
    const uint32_t d = pr.x, al = pr.y;
    const bool split = pr.z == 1u;
    // relocate.wgsl:50-54: split opacity in logit space.
    const float alpha_a = fminf(fmaxf(a.opac_act[al], 1e-4f), 0.999f);
    const float alpha_new = fminf(fmaxf(1.0f - sqrtf(1.0f - alpha_a), 1e-4f), 0.999f);
    const float logit = logf(alpha_new / (1.0f - alpha_new));
    a.opac_raw[d] = logit;
    a.opac_raw[al] = logit;
    const float4 q = a.quat_raw[al];
    a.quat_raw[d] = q;
    const float2 sa = a.scales_raw[al];
    if (split) {
        // relocate.wgsl:57-80.
        const float qn = sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
        const float x = q.x / qn, y = q.y / qn, z = q.z / qn, w = q.w / qn;
        const float3 col0 = make_float3(1.0f - 2.0f * (y * y + z * z), 2.0f * (x * y + w * z), 2.0f * (x * z - w * y));
        const float3 col1 = make_float3(2.0f * (x * y - w * z), 1.0f - 2.0f * (x * x + z * z), 2.0f * (y * z + w * x));
        const float swx = expf(sa.x), swy = expf(sa.y);
        float3 axis = make_float3(col0.x * swx, col0.y * swx, col0.z * swx);
        if (swy > swx) axis = make_float3(col1.x * swy, col1.y * swy, col1.z * swy);
        const float4 c = a.pos_raw[al];
        const float ox = 0.6f * axis.x, oy = 0.6f * axis.y, oz = 0.6f * axis.z;
        a.pos_raw[d] = make_float4(c.x + ox, c.y + oy, c.z + oz, c.w);
        a.pos_raw[al] = make_float4(c.x - ox, c.y - oy, c.z - oz, c.w);
        const float2 s = make_float2(sa.x - 0.51f, sa.y - 0.51f);  // log-space x0.6
        a.scales_raw[al] = s;
        a.scales_raw[d] = s;
    } else {
        a.pos_raw[d] = a.pos_raw[al];
        const float2 s = make_float2(sa.x - 0.16f, sa.y - 0.16f);  // log-space x0.85
        a.scales_raw[al] = s;
        a.scales_raw[d] = s;
    }
    const float4* sh_a = a.sh_raw + (size_t)al * (kShFloats / 4);
    float4* sh_d = a.sh_raw + (size_t)d * (kShFloats / 4);
#pragma unroll
    for (int k = 0; k < kShFloats / 4; ++k) sh_d[k] = sh_a[k];
    // relocate.wgsl:101-125: zero both indices' moments in every class.
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        const uint32_t wd = a.class_width[c];
#pragma unroll 1
        for (int idx = 0; idx < 2; ++idx) {
            const size_t base = (size_t)(idx == 0 ? d : al) * wd;
            for (uint32_t k = 0; k < wd; ++k) {
                a.m[c][base + k] = 0.0f;
                a.v[c][base + k] = 0.0f;
            }
        }
    }
}

}  // namespace

hipError_t launch_add_noise(hipStream_t st, const NoiseArgs& a) {
    if (a.n == 0) return hipSuccess;
    hipLaunchKernelGGL(add_noise_kernel, dim3((a.n + 255) / 256), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_relocate(hipStream_t st, const RelocateArgs& a) {
    if (a.n_pairs == 0) return hipSuccess;
    hipLaunchKernelGGL(relocate_kernel, dim3((a.n_pairs + 255) / 256), dim3(256), 0, st, a);
    return hipGetLastError();
}

}  // namespace gs
