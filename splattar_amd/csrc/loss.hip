// Fused L1 + D-SSIM loss and its image gradient in one kernel.
//
// Restates first_attempt/crates/gs-train/src/shaders/ssim.wgsl:41-143 and
// ssim.rs:1-4,84: loss = (1-l) mean|x-y| + l * s * mean(1 - SSIM), 11-tap
// sigma 1.5 separable Gaussian with zero padding (the reference's f32 taps),
// C1 = 1e-4, C2 = 9e-4, means over pixels x 3 channels; s = dssim_scale is
// 0.5 in the reference ((1 - SSIM)/2) and selectable here. The gradient
// blurs the per-pixel coefficient maps C_mu = dw dS/dmu_x, C_x2 = dw
// dS/dE[x^2], C_xy = dw dS/dE[xy] (the zero-padded symmetric blur is
// self-adjoint). The reference runs 19 dispatches over 16 P-sized
// intermediates; here nothing but the image, the target and dL/dimage
// touches HBM, plus, for a training target that never changes, its blurred
// mean and second moment per channel (computed once by target_stats_kernel;
// the step then blurs three statistics of the image instead of five).
#include "gs_internal.h"

#include <type_traits>

namespace gs {

namespace {

constexpr int kH = 5;           // blur half-width (11-tap window)
__constant__ float c_w[6] = {0.26601818f, 0.21300237f, 0.10936069f,
                             0.03599398f, 0.00759866f, 0.00102838f};

constexpr float kC1 = 0.0001f, kC2 = 0.0009f;

__device__ __forceinline__ float tap(int d) { return c_w[d < 0 ? -d : d]; }

// Workgroup barrier that waits for this thread's LDS operations only
// (__syncthreads also drains global loads): the next row's loads, issued
// before it, stay in flight across the row's barriers.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// One launch covers the views of a step. A workgroup
// owns a strip of kSW output columns and SH output rows of one view (all
// three channels) and walks its input rows top to bottom, one row per step:
//   load    the input row (kSW + 20 pixels: the blur of the coefficient maps
//           needs them 5 px beyond the strip, and those need 5 px more) as
//           (x, y) per channel into a 13-row LDS ring (each pixel is read
//           from HBM once);
//   stats   thread (channel, stat column) of kAC = kSW + 10 columns blurs the
//           five statistics x, y, x^2, y^2, xy horizontally (three with the
//           target's precomputed) and keeps the last 11 rows in a register
//           ring;
//           once 11 are there, the vertical blur gives the window statistics
//           of the stat row 5 above, then SSIM and the coefficients C_mu,
//           C_x2, C_xy (0 outside the image: zero padding), which go to an
//           LDS row;
//   output  thread (channel, output column) blurs that row horizontally (a
//           step later, into a second 11-row ring) and, 5 rows later,
//           vertically, giving
//           dL/dx = (1-l) sign(x-y)/3P + B(C_mu) + 2x B(C_x2) + y B(C_xy)
//           at the output row 11 above the input row.
// The coefficient maps never leave the chip; a strip re-blurs a 10-px
// horizontal halo and its band a 20-row vertical one. Loss partial sums
// (sum |x - y|, sum SSIM over the workgroup's own outputs) accumulate in
// double per thread.
constexpr int kSW = 74;                 // output columns per strip
constexpr int kAC = kSW + 2 * kH;       // stat columns (84): 3 kAC = 252 threads
constexpr int kIC = kSW + 4 * kH;       // input columns (94)
constexpr int kMaxLossViews = 64;

struct LossViews {
    uint32_t n;                          // views in this launch
    uint32_t first;                      // view slot of the first (camera table, partials)
    uint32_t partials_stride;            // partial-sum pairs per view slot
    uint32_t block_begin[kMaxLossViews + 1];
    uint32_t w[kMaxLossViews], h[kMaxLossViews], pix_base[kMaxLossViews];
    float dw[kMaxLossViews], l1w[kMaxLossViews];
};

// TS: the target's window mean and second moment come from tstats
// ((target pixel) * 3 + channel) instead of being blurred here. Four waves
// per SIMD (at most 128 VGPRs; 35 KB of LDS): a fourth workgroup per CU
// hides the row barrier better than the third's spare registers would. The
// TS variant (register rings of 66) is held to that; the other (rings of
// 88) fits it too when left a looser bound (forced, it spills two).
template <int SH, bool TS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TS ? 4 : 3))) void l1_ssim_fused_kernel(
    LossViews lv, const float4* __restrict__ img_base, const float4* __restrict__ tgt_base,
    float4* __restrict__ dl_base, double* __restrict__ partials, const CamParams* __restrict__ cams,
    const float2* __restrict__ tstats) {
    constexpr int NR = TS ? 3 : 5;  // statistics blurred per row: (x, x^2, xy) or (x, y, x^2, y^2, xy)
    // Channels are the fastest-varying index of every LDS array, as they are
    // of the thread roles (thread 3c + ch), so a wave's lanes read and write
    // consecutive elements: no bank conflicts where a wave spans two channels.
    __shared__ float s_c[2][3][kAC][3];  // C_mu, C_x2, C_xy of a stat row, double-buffered
    // (x, y) of the last input rows, the only staged form of the input: the
    // horizontal statistics blur reads its row in the step that loads it and
    // forms x^2 and xy in registers (8 B of LDS per tap instead of 12 or 20),
    // the output stage reads its row 11 steps later (13 slots: the loads of
    // the next two steps never meet a slot still being read).
    __shared__ float2 s_xy[13][kIC][3];
    __shared__ double s_red[2][4];
    const int tid = threadIdx.x;
    const uint32_t b = xcd_swizzle(blockIdx.x, gridDim.x);
    uint32_t v = 0;
    for (uint32_t k = 1; k < lv.n; ++k)
        if (lv.block_begin[k] <= b) v = k;
    const uint32_t blk = b - lv.block_begin[v];
    const uint32_t W = lv.w[v], H = lv.h[v];
    const float dw = lv.dw[v], l1w = lv.l1w[v];
    const float4* img = img_base + lv.pix_base[v];
    float4* dl_out = dl_base != nullptr ? dl_base + lv.pix_base[v] : nullptr;
    const uint32_t tbase = cams != nullptr ? cams[lv.first + v].target_base : 0u;
    const float4* tgt = tgt_base + tbase;
    const uint32_t strips = (W + kSW - 1) / kSW;
    const int x0 = (int)(blk % strips) * kSW, y0 = (int)(blk / strips) * SH;
    const int iW = (int)W, iH = (int)H;
    // Stats role: thread 3 cA + chA is (stat column x0 - 5 + cA, channel chA);
    // output role: thread 3 cB + chB is (output column x0 + cB, channel chB).
    const bool roleA = tid < 3 * kAC, roleB = tid < 3 * kSW;
    const int cA = roleA ? tid / 3 : 0, chA = roleA ? tid - 3 * cA : 0;
    const int cB = roleB ? tid / 3 : 0, chB = roleB ? tid - 3 * cB : 0;
    const int gxA = x0 - kH + cA, gxB = x0 + cB;
    const bool colA_in = roleA && gxA >= 0 && gxA < iW;
    const bool colA_own = colA_in && cA >= kH && cA < kH + kSW;
    const bool colB_in = roleB && gxB < iW;
    float ra[11][NR], rb[11][3];
    double l1 = 0.0, ss = 0.0;
    constexpr int nrows = SH + 4 * kH;  // input rows y0 - 10 .. y0 + SH + 9
    // Step i (0 <= i <= nrows): input row y0 - 10 + i and stat row
    // y0 - 15 + i (i < nrows); the C-map row of the previous step's stat row
    // is blurred horizontally and output row y0 - 21 + i is finished. In the
    // register rings, step i's statistics row sits in slot i % 11 (the
    // template K) and step i - 1's blurred C-map row in (i - 1) % 11 (KP).
    // One barrier per step: before it, the input row goes into LDS ring slot
    // i % 13 (read by this step's statistics blur and, 11 steps later, by
    // the output) and the previous step's C-map row into buffer (i - 1) & 1;
    // a slot or buffer is rewritten at the earliest two steps after its last
    // read, when every thread has passed the barrier between.
    // The loader threads hold the next input row in registers: its global
    // loads are issued a step ahead and overlap the row's work. They are
    // unconditional (clamped addresses; threads past the loaders repeat the
    // last loader's, which coalesces), so nothing waits on them before the
    // next step uses them; outside the image the values are replaced by 0.
    // TS: the target statistics of the stat row come the same way, a step
    // ahead (the stat row of input row gy is gy - 5).
    const int lgx = x0 - 2 * kH + min(tid, kIC - 1);
    const bool lcol = lgx >= 0 && lgx < iW;
    const int cgx = min(max(lgx, 0), iW - 1);
    const int sgx = min(max(gxA, 0), iW - 1);
    const float2* ts_row = TS ? tstats + (size_t)tbase * 3 + (size_t)sgx * 3 + chA : nullptr;
    float4 a, t;
    float2 ts = make_float2(0.f, 0.f);
    bool a_ok = false;
    auto fetch = [&](int gy) {
        a_ok = lcol && gy >= 0 && gy < iH;
        const size_t q = (size_t)min(max(gy, 0), iH - 1) * W + cgx;
        a = img[q];
        t = tgt[q];
        if constexpr (TS) ts = ts_row[(size_t)min(max(gy - kH, 0), iH - 1) * W * 3];
    };
    fetch(y0 - 2 * kH);
    float cm = 0.f, cx2 = 0.f, cxy = 0.f;  // the C-map values of this thread's last stat row
    auto step = [&](auto kc, int i) {
        constexpr int K = decltype(kc)::value;
        constexpr int KP = (K + 10) % 11;
        const int gy = y0 - 2 * kH + i;
        const int buf = i & 1;
        const float2 tsc = ts;  // target statistics of this step's stat row
        if (tid < kIC && i < nrows) {
            if (!a_ok) {
                a = make_float4(0.f, 0.f, 0.f, 0.f);
                t = a;
            }
            const int slot = i % 13;
            s_xy[slot][tid][0] = make_float2(a.x, t.x);
            s_xy[slot][tid][1] = make_float2(a.y, t.y);
            s_xy[slot][tid][2] = make_float2(a.z, t.z);
        }
        if (roleA && i > 2 * kH) {  // the previous step's C-map row
            s_c[buf ^ 1][0][cA][chA] = cm;
            s_c[buf ^ 1][1][cA][chA] = cx2;
            s_c[buf ^ 1][2][cA][chA] = cxy;
        }
        fetch(gy + 1);  // every thread, every step: no branch for the loads to wait at
        lds_barrier();
        if (roleA && i < nrows) {
            // Per tap: wx = w x, sums wx, wx x, wx y (and wy, wy y without TS;
            // target_stats_kernel forms the target's the same way).
            const int slot = i % 13;
            if constexpr (TS) {
                float h0 = 0.f, h2 = 0.f, h4 = 0.f;
#pragma unroll
                for (int j = 0; j <= 2 * kH; ++j) {
                    const float2 p = s_xy[slot][cA + j][chA];
                    const float wx = tap(j - kH) * p.x;
                    h0 = h0 + wx;
                    h2 = fmaf(wx, p.x, h2);
                    h4 = fmaf(wx, p.y, h4);
                }
                ra[K][0] = h0;
                ra[K][1] = h2;
                ra[K][2] = h4;
            } else {
                float h0 = 0.f, h1 = 0.f, h2 = 0.f, h3 = 0.f, h4 = 0.f;
#pragma unroll
                for (int j = 0; j <= 2 * kH; ++j) {
                    const float2 p = s_xy[slot][cA + j][chA];
                    const float w = tap(j - kH);
                    const float wx = w * p.x, wy = w * p.y;
                    h0 = h0 + wx;
                    h1 = h1 + wy;
                    h2 = fmaf(wx, p.x, h2);
                    h3 = fmaf(wy, p.y, h3);
                    h4 = fmaf(wx, p.y, h4);
                }
                ra[K][0] = h0;
                ra[K][1] = h1;
                ra[K][2] = h2;
                ra[K][3] = h3;
                ra[K][4] = h4;
            }
            if (i >= 2 * kH) {
                const int sy = gy - kH;
                float mu_x = 0.f, mu_y = 0.f, bx2 = 0.f, by2 = 0.f, bxy = 0.f;
#pragma unroll
                for (int d = -kH; d <= kH; ++d) {
                    const int s = (K - kH + d + 22) % 11;
                    const float w = tap(d);
                    if constexpr (TS) {
                        mu_x = fmaf(w, ra[s][0], mu_x);
                        bx2 = fmaf(w, ra[s][1], bx2);
                        bxy = fmaf(w, ra[s][2], bxy);
                    } else {
                        mu_x = fmaf(w, ra[s][0], mu_x);
                        mu_y = fmaf(w, ra[s][1], mu_y);
                        bx2 = fmaf(w, ra[s][2], bx2);
                        by2 = fmaf(w, ra[s][3], by2);
                        bxy = fmaf(w, ra[s][4], bxy);
                    }
                }
                if constexpr (TS) {
                    mu_y = tsc.x;
                    by2 = tsc.y;
                }
                cm = cx2 = cxy = 0.f;
                if (colA_in && sy >= 0 && sy < iH) {
                    // ssim.wgsl:95-126 (partials).
                    // Every multiply-add is an explicit fmaf: the rounding is then
                    // the same in both template variants (contraction left to the
                    // compiler may fuse different products in each).
                    const float sx2 = fmaf(-mu_x, mu_x, bx2), sy2 = fmaf(-mu_y, mu_y, by2);
                    const float sxy_ = fmaf(-mu_x, mu_y, bxy);
                    const float p = fmaf(mu_x, mu_x, fmaf(mu_y, mu_y, kC1));
                    const float q = (sx2 + sy2) + kC2;
                    const float rp = rcp_refined(p), rq = rcp_refined(q);
                    const float l = fmaf(2.0f * mu_x, mu_y, kC1) * rp;
                    const float cs = fmaf(2.0f, sxy_, kC2) * rq;
                    const float ga = fmaf(-2.0f * mu_x, l, 2.0f * mu_y);
                    const float gb = fmaf(2.0f * mu_x, cs, -2.0f * mu_y);
                    const float ds_dmux = fmaf(cs * ga, rp, (l * gb) * rq);
                    cm = ds_dmux * dw;
                    cx2 = (-l * cs * rq) * dw;
                    cxy = (2.0f * l * rq) * dw;
                    if (colA_own && sy >= y0 && sy < y0 + SH) ss += (double)(l * cs);
                }
            }
        }
        if (i > 2 * kH) {
            if (roleB) {
                float b0 = 0.f, b1 = 0.f, b2 = 0.f;
#pragma unroll
                for (int j = 0; j <= 2 * kH; ++j) {
                    const float w = tap(j - kH);
                    b0 = fmaf(w, s_c[buf ^ 1][0][cB + j][chB], b0);
                    b1 = fmaf(w, s_c[buf ^ 1][1][cB + j][chB], b1);
                    b2 = fmaf(w, s_c[buf ^ 1][2][cB + j][chB], b2);
                }
                rb[KP][0] = b0;
                rb[KP][1] = b1;
                rb[KP][2] = b2;
            }
            const int oy = gy - 2 * kH - 1;
            if (i > 4 * kH && roleB && colB_in && oy < iH) {
                float b0 = 0.f, b1 = 0.f, b2 = 0.f;
#pragma unroll
                for (int d = -kH; d <= kH; ++d) {
                    const int s = (KP - kH + d + 22) % 11;
                    const float w = tap(d);
                    b0 = fmaf(w, rb[s][0], b0);
                    b1 = fmaf(w, rb[s][1], b1);
                    b2 = fmaf(w, rb[s][2], b2);
                }
                const size_t pix = (size_t)oy * W + gxB;
                const float2 xy = s_xy[(i - 2 * kH - 1) % 13][cB + 2 * kH][chB];
                const float x = xy.x, y = xy.y;
                l1 += (double)fabsf(x - y);
                if (dl_out != nullptr) {
                    const float sg = x - y > 0.f ? 1.f : (x - y < 0.f ? -1.f : 0.f);
                    float* o = reinterpret_cast<float*>(dl_out) + 4 * pix;
                    o[chB] = fmaf(y, b2, fmaf(2.0f * x, b1, fmaf(sg, l1w, b0)));
                    if (chB == 0) o[3] = 0.0f;
                }
            }
        }
    };
    for (int i0 = 0; i0 <= nrows; i0 += 11) {
        using std::integral_constant;
        step(integral_constant<int, 0>{}, i0);
        if (i0 + 1 <= nrows) step(integral_constant<int, 1>{}, i0 + 1);
        if (i0 + 2 <= nrows) step(integral_constant<int, 2>{}, i0 + 2);
        if (i0 + 3 <= nrows) step(integral_constant<int, 3>{}, i0 + 3);
        if (i0 + 4 <= nrows) step(integral_constant<int, 4>{}, i0 + 4);
        if (i0 + 5 <= nrows) step(integral_constant<int, 5>{}, i0 + 5);
        if (i0 + 6 <= nrows) step(integral_constant<int, 6>{}, i0 + 6);
        if (i0 + 7 <= nrows) step(integral_constant<int, 7>{}, i0 + 7);
        if (i0 + 8 <= nrows) step(integral_constant<int, 8>{}, i0 + 8);
        if (i0 + 9 <= nrows) step(integral_constant<int, 9>{}, i0 + 9);
        if (i0 + 10 <= nrows) step(integral_constant<int, 10>{}, i0 + 10);
    }
    // Deterministic block reduction of the loss sums (double).
    const int lane = tid & 63, wave = tid >> 6;
    for (int o = 32; o >= 1; o >>= 1) {
        l1 += __shfl_xor(l1, o);
        ss += __shfl_xor(ss, o);
    }
    if (lane == 0) {
        s_red[0][wave] = l1;
        s_red[1][wave] = ss;
    }
    __syncthreads();
    if (tid == 0) {
        double* pp = partials + ((size_t)(lv.first + v) * lv.partials_stride + blk) * 2;
        pp[0] = ((s_red[0][0] + s_red[0][1]) + s_red[0][2]) + s_red[0][3];
        pp[1] = ((s_red[1][0] + s_red[1][1]) + s_red[1][2]) + s_red[1][3];
    }
}

constexpr uint32_t kSH = 64;  // output rows per band

hipError_t launch_fused(hipStream_t st, const LossViews& lv, const float4* img, const float4* target,
                        float4* dl_out, double* partials, const CamParams* cams, const float2* tstats) {
    const dim3 g(lv.block_begin[lv.n]);
    if (g.x == 0) return hipSuccess;
    if (tstats != nullptr)
        hipLaunchKernelGGL((l1_ssim_fused_kernel<kSH, true>), g, dim3(256), 0, st, lv, img, target, dl_out,
                           partials, cams, tstats);
    else
        hipLaunchKernelGGL((l1_ssim_fused_kernel<kSH, false>), g, dim3(256), 0, st, lv, img, target, dl_out,
                           partials, cams, tstats);
    return hipGetLastError();
}

// The target's window mean and second moment per channel (the statistics
// l1_ssim_fused_kernel<.., true> reads): the same 11 x 11 zero-padded blur,
// horizontal then vertical, with the taps summed in the same order as the
// fused kernel's, one thread per pixel.
__global__ __launch_bounds__(256) void target_stats_kernel(const float4* __restrict__ tgt, uint32_t W,
                                                           uint32_t H, float2* __restrict__ out) {
    const uint32_t p = blockIdx.x * 256u + threadIdx.x;
    if (p >= W * H) return;
    const int x = (int)(p % W), y = (int)(p / W), iW = (int)W, iH = (int)H;
    float mu[3] = {0.f, 0.f, 0.f}, e2[3] = {0.f, 0.f, 0.f};
    for (int d = -kH; d <= kH; ++d) {
        const int r = y + d;
        float h[3] = {0.f, 0.f, 0.f}, h2[3] = {0.f, 0.f, 0.f};
        if (r >= 0 && r < iH) {
            for (int j = 0; j <= 2 * kH; ++j) {
                const int c = x - kH + j;
                if (c < 0 || c >= iW) continue;
                const float4 v = tgt[(size_t)r * W + c];
                const float w = tap(j - kH);
                const float wx = w * v.x, wy = w * v.y, wz = w * v.z;
                h[0] = h[0] + wx;
                h[1] = h[1] + wy;
                h[2] = h[2] + wz;
                h2[0] = fmaf(wx, v.x, h2[0]);
                h2[1] = fmaf(wy, v.y, h2[1]);
                h2[2] = fmaf(wz, v.z, h2[2]);
            }
        }
        const float w = tap(d);
        for (int k = 0; k < 3; ++k) {
            mu[k] = fmaf(w, h[k], mu[k]);
            e2[k] = fmaf(w, h2[k], e2[k]);
        }
    }
    for (int k = 0; k < 3; ++k) out[(size_t)p * 3 + k] = make_float2(mu[k], e2[k]);
}

// Uniform loss weights of a w x h view: dw = -l s / 3P (D-SSIM term), the L1
// term's (1 - l) / 3P.
void loss_weights(uint32_t w, uint32_t h, float lambda, float dssim_scale, float* dw, float* l1w) {
    const float n3 = (float)w * (float)h * 3.0f;
    *dw = -lambda * dssim_scale * (1.0f / n3);
    *l1w = (1.0f - lambda) / n3;
}

}  // namespace

uint32_t l1_ssim_partials(uint32_t w, uint32_t h) { return div_up(w, kSW) * div_up(h, kSH); }

hipError_t launch_target_stats(hipStream_t st, uint32_t w, uint32_t h, const float4* target, float2* out) {
    const size_t P = (size_t)w * h;
    if (P == 0) return hipSuccess;
    if (P >= (1ull << 32)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(target_stats_kernel, dim3((uint32_t)div_up(P, (size_t)256)), dim3(256), 0, st, target, w, h,
                       out);
    return hipGetLastError();
}

hipError_t launch_l1_ssim(hipStream_t st, uint32_t w, uint32_t h, float lambda, float dssim_scale,
                          const float4* img, const float4* target, float4* dl_out, double* partial_sums,
                          const CamParams* view, const float2* tstats) {
    if (w == 0 || h == 0) return hipSuccess;
    LossViews lv{};
    lv.n = 1;
    lv.w[0] = w;
    lv.h[0] = h;
    loss_weights(w, h, lambda, dssim_scale, &lv.dw[0], &lv.l1w[0]);
    lv.block_begin[1] = l1_ssim_partials(w, h);
    return launch_fused(st, lv, img, target, dl_out, partial_sums, view, tstats);
}

hipError_t launch_l1_ssim_views(hipStream_t st, uint32_t first, uint32_t k, const uint32_t* widths,
                                const uint32_t* heights, const uint32_t* pix_base, float lambda,
                                float dssim_scale, const float4* img, const float4* target, float4* dl_out,
                                double* partials, uint32_t partials_stride, const CamParams* cams_dev,
                                const float2* tstats) {
    if (k == 0) return hipSuccess;
    if (k > (uint32_t)kMaxLossViews) return hipErrorInvalidValue;
    LossViews lv{};
    lv.n = k;
    lv.first = first;
    lv.partials_stride = partials_stride;
    uint32_t blocks = 0;
    for (uint32_t j = 0; j < k; ++j) {
        lv.w[j] = widths[j];
        lv.h[j] = heights[j];
        lv.pix_base[j] = pix_base[j];
        loss_weights(widths[j], heights[j], lambda, dssim_scale, &lv.dw[j], &lv.l1w[j]);
        lv.block_begin[j] = blocks;
        blocks += l1_ssim_partials(widths[j], heights[j]);
    }
    lv.block_begin[k] = blocks;
    return launch_fused(st, lv, img, target, dl_out, partials, cams_dev, tstats);
}

}  // namespace gs
