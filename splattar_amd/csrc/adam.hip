// Fused Adam over every parameter class in one launch.
//
// Restates first_attempt/crates/gs-train/src/shaders/adam.wgsl:44-74 (adam_step
// then activate) with the host-side bias corrections of optim.rs:179-197:
//   g = (grad_act + reg) * dact(raw), dact = exp (scales) | s(1-s) (opacity) | 1
//   m = b1 m + (1-b1) g;  v = b2 v + (1-b2) g^2
//   raw -= lr (m bc1) / (sqrt(v bc2) + eps);  activated = act(raw)
// The reference launches 2 dispatches per class (10 per step) and re-reads
// raw for the activation; here one grid covers all classes (block ranges per
// class) and the activated copy is written from registers in the same pass.
// The step counter, the running f32 products b1^t, b2^t and each class's
// learning rate for the step (from its schedule) live in device memory
// (adam_tick), so a captured graph replays correct bias corrections and a
// decaying learning rate.
//
// HBM traffic: 28 B per parameter float (read raw, g, m, v; write raw, m, v,
// activated).

#include "gs_internal.h"

namespace gs {

namespace {

constexpr int kAdamThreads = 256;
constexpr int kAdamItems = 4;   // floats per thread (one float4)

__global__ void adam_tick_kernel(AdamState* st) {
    // One wave: lane k < kMaxAdamClasses computes class k's learning rate.
    const uint32_t k = threadIdx.x;
    const uint32_t t = st->t + 1u;
    if (k < (uint32_t)kMaxAdamClasses) {
        const AdamSchedule sc = st->hyper.sched[k];
        float lr = sc.base;
        if (sc.iters > 0u) lr = sc.base * powf(sc.final_factor, (float)(t - 1u) / (float)sc.iters);
        st->lr[k] = lr;
    }
    if (k == 0) {
        const float p1 = st->p1 * 0.9f, p2 = st->p2 * 0.999f;
        st->t = t;
        st->p1 = p1;
        st->p2 = p2;
        st->bc1_inv = 1.0f / (1.0f - p1);
        st->bc2_inv = 1.0f / (1.0f - p2);
    }
}

__global__ __launch_bounds__(kAdamThreads) void adam_kernel(const AdamTable table,
                                                            const AdamState* __restrict__ state,
                                                            int step) {
    // Which class owns this block (block ranges are contiguous, ascending).
    uint32_t ci = 0;
    for (uint32_t k = 1; k < table.n_classes; ++k)
        if (table.c[k].block_begin <= blockIdx.x) ci = k;
    const AdamClass c = table.c[ci];
    const uint32_t first = (blockIdx.x - c.block_begin) * (kAdamThreads * kAdamItems) + threadIdx.x * kAdamItems;
    if (first >= c.n) return;
    const float bc1 = state->bc1_inv, bc2 = state->bc2_inv;
    const float lr = state->lr[ci], reg = state->hyper.reg[ci];
    const uintptr_t align = (uintptr_t)c.raw | (uintptr_t)c.grad | (uintptr_t)c.m | (uintptr_t)c.v |
                            (uintptr_t)c.activated;
    const bool vec = (first + kAdamItems <= c.n) && ((align & 15u) == 0);
    float raw[kAdamItems], gr[kAdamItems], m[kAdamItems], v[kAdamItems];
    if (vec) {
        const float4 a = *reinterpret_cast<const float4*>(c.raw + first);
        raw[0] = a.x; raw[1] = a.y; raw[2] = a.z; raw[3] = a.w;
        if (step) {
            const float4 b = *reinterpret_cast<const float4*>(c.grad + first);
            const float4 mm = *reinterpret_cast<const float4*>(c.m + first);
            const float4 vv = *reinterpret_cast<const float4*>(c.v + first);
            gr[0] = b.x; gr[1] = b.y; gr[2] = b.z; gr[3] = b.w;
            m[0] = mm.x; m[1] = mm.y; m[2] = mm.z; m[3] = mm.w;
            v[0] = vv.x; v[1] = vv.y; v[2] = vv.z; v[3] = vv.w;
        }
    } else {
        for (int k = 0; k < kAdamItems; ++k) {
            const uint32_t idx = first + k;
            const bool ok = idx < c.n;
            raw[k] = ok ? c.raw[idx] : 0.f;
            gr[k] = (ok && step) ? c.grad[idx] : 0.f;
            m[k] = (ok && step) ? c.m[idx] : 0.f;
            v[k] = (ok && step) ? c.v[idx] : 0.f;
        }
    }
    float act[kAdamItems];
#pragma unroll
    for (int k = 0; k < kAdamItems; ++k) {
        if (step) adam_element(c.activation, lr, reg, bc1, bc2, raw[k], gr[k], m[k], v[k]);
        act[k] = adam_activate(c.activation, raw[k]);
    }
    if (vec) {
        if (step) {
            *reinterpret_cast<float4*>(c.raw + first) = make_float4(raw[0], raw[1], raw[2], raw[3]);
            *reinterpret_cast<float4*>(c.m + first) = make_float4(m[0], m[1], m[2], m[3]);
            *reinterpret_cast<float4*>(c.v + first) = make_float4(v[0], v[1], v[2], v[3]);
        }
        if (c.activated)
            *reinterpret_cast<float4*>(c.activated + first) = make_float4(act[0], act[1], act[2], act[3]);
    } else {
        for (int k = 0; k < kAdamItems; ++k) {
            const uint32_t idx = first + k;
            if (idx >= c.n) break;
            if (step) {
                c.raw[idx] = raw[k];
                c.m[idx] = m[k];
                c.v[idx] = v[k];
            }
            if (c.activated) c.activated[idx] = act[k];
        }
    }
}

}  // namespace

uint32_t adam_blocks_for(uint32_t n) { return div_up(n ? n : 1, kAdamThreads * kAdamItems); }

hipError_t launch_adam_tick(hipStream_t st, AdamState* state) {
    hipLaunchKernelGGL(adam_tick_kernel, dim3(1), dim3(kWave), 0, st, state);
    return hipGetLastError();
}

hipError_t launch_adam(hipStream_t st, const AdamTable& table, uint32_t total_blocks,
                       const AdamState* state, bool step) {
    if (total_blocks == 0 || table.n_classes == 0) return hipSuccess;
    hipLaunchKernelGGL(adam_kernel, dim3(total_blocks), dim3(kAdamThreads), 0, st, table, state,
                       step ? 1 : 0);
    return hipGetLastError();
}

}  // namespace gs
