// Path-guide network on the matrix cores: BF16 MFMA (v_mfma_f32_16x16x16_bf16) with FP32
// accumulation, for inference inside the life cycle and for training.
//
// Network: 9 features -> 64 -> 64 -> 64 -> 40 (ReLU on the hidden layers); the 40 outputs
// are 8 von Mises-Fisher lobes x (mu_x, mu_y, mu_z, kappa_raw, weight_logit).
//
// Layout.  A wavefront evaluates 16 samples per MFMA column tile.  Layers are computed in
// transposed form, H_l^T = W_l H_{l-1}^T: the weights are the A operand (row = output unit,
// 4 consecutive inputs per lane group) and the activations the B operand (column = sample,
// 4 consecutive units per lane group).  The MFMA result layout (lane holds 4 consecutive
// output rows of one sample column) is then exactly the B-operand layout of the next layer,
// so activations stay in registers between layers.  The backward pass uses the same trick
// with transposed weight copies; weight gradients contract over samples and go through LDS.
//
// Parameter vector (FP32, the order of the CPU statement): W0 (64x9), b0, W1 (64x64), b1,
// W2 (64x64), b2, W3 (40x64), b3 -- 11,560 values.  BF16 copies, padded to whole tiles
// (inputs 9 -> 16, outputs 40 -> 48), are rebuilt after every optimiser step.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "philox.hpp"

namespace skg
{
    constexpr int kGuideIn = 9;
    constexpr int kGuideHidden = 64;
    constexpr int kGuideOut = 40;
    constexpr int kGuideLobes = 8;
    constexpr int kGuideParams = 64 * 9 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 40 * 64 + 40;  // 11560
    constexpr double kGuideKMin = 0.2;
    constexpr double kGuideKMax = 400.0;

    // offsets of the FP32 parameter blocks
    constexpr int kOffW0 = 0, kOffB0 = kOffW0 + 64 * 9, kOffW1 = kOffB0 + 64, kOffB1 = kOffW1 + 64 * 64,
                  kOffW2 = kOffB1 + 64, kOffB2 = kOffW2 + 64 * 64, kOffW3 = kOffB2 + 64, kOffB3 = kOffW3 + 40 * 64;

    typedef short bf16x4 __attribute__((ext_vector_type(4)));
    typedef float f32x4 __attribute__((ext_vector_type(4)));

    __host__ __device__ __forceinline__ uint16_t f2bf(float f)
    {
        uint32_t u = __builtin_bit_cast(uint32_t, f);
        u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
        return static_cast<uint16_t>(u >> 16);
    }
    __host__ __device__ __forceinline__ float bf2f(uint16_t b)
    {
        return __builtin_bit_cast(float, static_cast<uint32_t>(b) << 16);
    }

    // BF16 weight tiles: W_l row-major (out_pad x in_pad) and its transpose (in_pad x out_pad)
    struct GuideNet
    {
        const uint16_t* W[4];
        const uint16_t* WT[4];
        const float* b[4];
    };
    // padded sizes per layer
    __host__ __device__ constexpr int guide_in_pad(int l) { return l == 0 ? 16 : 64; }
    __host__ __device__ constexpr int guide_out_pad(int l) { return l == 3 ? 48 : 64; }

    __device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
    }

    __device__ __forceinline__ bf16x4 load_bf4(const uint16_t* p)
    {
        // 8-byte aligned load of 4 consecutive bf16
        uint2 v = *reinterpret_cast<const uint2*>(p);
        bf16x4 r;
        r[0] = static_cast<short>(v.x & 0xFFFFu);
        r[1] = static_cast<short>(v.x >> 16);
        r[2] = static_cast<short>(v.y & 0xFFFFu);
        r[3] = static_cast<short>(v.y >> 16);
        return r;
    }
    __device__ __forceinline__ bf16x4 pack_bf4(f32x4 v)
    {
        bf16x4 r;
        r[0] = static_cast<short>(f2bf(v[0]));
        r[1] = static_cast<short>(f2bf(v[1]));
        r[2] = static_cast<short>(f2bf(v[2]));
        r[3] = static_cast<short>(f2bf(v[3]));
        return r;
    }

    // Forward pass of one 16-sample column tile.  x: B-operand fragment of the input
    // (features 4g..4g+3 of sample l%16, g = lane/16).  Returns the 3 output tiles (48 rows,
    // 40 used) in result layout; h1..h3 (post-ReLU, BF16, result layout) are optional outputs
    // for the backward pass.
    __device__ __forceinline__ void guide_tile_forward(const GuideNet& net, bf16x4 x, f32x4 out[3], bf16x4* h1,
                                                       bf16x4* h2, bf16x4* h3)
    {
        const int lane = __lane_id();
        const int row = lane & 15, grp = lane >> 4;
        bf16x4 cur[4];
        // layer 0: K = 16 (one k tile), 4 output tiles
#pragma unroll
        for (int o = 0; o < 4; ++o)
        {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = mfma16(load_bf4(net.W[0] + (o * 16 + row) * 16 + grp * 4), x, acc);
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[t] = fmaxf(acc[t] + net.b[0][o * 16 + grp * 4 + t], 0.f);
            cur[o] = pack_bf4(acc);
            if (h1) h1[o] = cur[o];
        }
        // layers 1, 2: K = 64 (4 k tiles = previous output tiles), 4 output tiles
#pragma unroll
        for (int l = 1; l <= 2; ++l)
        {
            bf16x4 nxt[4];
#pragma unroll
            for (int o = 0; o < 4; ++o)
            {
                f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int k = 0; k < 4; ++k) acc = mfma16(load_bf4(net.W[l] + (o * 16 + row) * 64 + k * 16 + grp * 4), cur[k], acc);
#pragma unroll
                for (int t = 0; t < 4; ++t) acc[t] = fmaxf(acc[t] + net.b[l][o * 16 + grp * 4 + t], 0.f);
                nxt[o] = pack_bf4(acc);
            }
#pragma unroll
            for (int o = 0; o < 4; ++o)
            {
                cur[o] = nxt[o];
                if (l == 1 && h2) h2[o] = cur[o];
                if (l == 2 && h3) h3[o] = cur[o];
            }
        }
        // output layer: 3 output tiles (48 rows, 40 used)
#pragma unroll
        for (int o = 0; o < 3; ++o)
        {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int k = 0; k < 4; ++k) acc = mfma16(load_bf4(net.W[3] + (o * 16 + row) * 64 + k * 16 + grp * 4), cur[k], acc);
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[t] += net.b[3][o * 16 + grp * 4 + t];
            out[o] = acc;
        }
    }

    // ---------------------------------------------------------------- mixture on one sample
    __device__ __forceinline__ double softplus_d(double x) { return x > 20 ? x : log1p(exp(x)); }

    // The sampler inside the life cycle works in FP32: lobe parameters, mixture density and
    // direction sampling.  The weight factor uses the density of this same sampler, so the
    // estimator stays unbiased up to the FP32 rounding of that density (~1e-7 relative); the
    // sampled direction is renormalised in FP64.
    struct LobesF
    {
        float mu[kGuideLobes][3];
        float kap[kGuideLobes];
        float w[kGuideLobes];
        float norm[kGuideLobes];  // kappa / (2 pi (1 - exp(-2 kappa))), the vMF normalisation
    };

    __device__ __forceinline__ void guide_lobes_f(const float* o, LobesF& L)
    {
        float mx = -3.0e38f;
#pragma unroll
        for (int j = 0; j < kGuideLobes; ++j) mx = fmaxf(mx, o[5 * j + 4]);
        float s = 0.f;
#pragma unroll
        for (int j = 0; j < kGuideLobes; ++j)
        {
            const float a = o[5 * j], b = o[5 * j + 1], c = o[5 * j + 2];
            float n = sqrtf(a * a + b * b + c * c);
            n = n > 1e-12f ? n : 1e-12f;
            L.mu[j][0] = a / n;
            L.mu[j][1] = b / n;
            L.mu[j][2] = c / n;
            const float r = o[5 * j + 3];
            const float sp = r > 20.f ? r : log1pf(expf(r));
            const float kap = fminf(static_cast<float>(kGuideKMin) + sp, static_cast<float>(kGuideKMax));
            L.kap[j] = kap;
            L.norm[j] = kap / (2.f * static_cast<float>(M_PI) * (-expm1f(-2.f * kap)));
            L.w[j] = expf(o[5 * j + 4] - mx);
            s += L.w[j];
        }
        const float inv = 1.f / s;
#pragma unroll
        for (int j = 0; j < kGuideLobes; ++j) L.w[j] *= inv;
    }

    // mixture density per steradian at unit vector k
    __device__ __forceinline__ double guide_pdf_f(const LobesF& L, double kx, double ky, double kz)
    {
        const float x = static_cast<float>(kx), y = static_cast<float>(ky), z = static_cast<float>(kz);
        float q = 0.f;
#pragma unroll
        for (int j = 0; j < kGuideLobes; ++j)
        {
            const float c = L.mu[j][0] * x + L.mu[j][1] * y + L.mu[j][2] * z;
            q += L.w[j] * L.norm[j] * expf(L.kap[j] * (c - 1.f));
        }
        return static_cast<double>(q);
    }

    // samples the mixture: lobe by weight, then the 3-D vMF by w = 1 + ln(u + (1-u) e^{-2 kappa}) / kappa
    template<typename Rng> __device__ __forceinline__ void guide_sample_f(const LobesF& L, Rng& rng, double& kx, double& ky, double& kz)
    {
        const float u = static_cast<float>(rng.uniform());
        // the chosen lobe's parameters by selection (no dynamically indexed registers)
        float kap = L.kap[kGuideLobes - 1], mx = L.mu[kGuideLobes - 1][0], my = L.mu[kGuideLobes - 1][1],
              mz = L.mu[kGuideLobes - 1][2];
        float cum = 0.f;
        bool found = false;
#pragma unroll
        for (int q = 0; q < kGuideLobes - 1; ++q)
        {
            cum += L.w[q];
            if (!found && u < cum)
            {
                found = true;
                kap = L.kap[q];
                mx = L.mu[q][0];
                my = L.mu[q][1];
                mz = L.mu[q][2];
            }
        }
        const float e = expf(-2.f * kap);
        float w = 1.f + logf(static_cast<float>(rng.uniform()) * (1.f - e) + e) / kap;
        w = fminf(1.f, fmaxf(-1.f, w));
        const float phi = 2.f * static_cast<float>(M_PI) * static_cast<float>(rng.uniform());
        const float s = sqrtf(fmaxf(0.f, 1.f - w * w));
        // orthonormal frame around mu
        const float ax = fabsf(mz) < 0.9f ? 0.f : 1.f, az = fabsf(mz) < 0.9f ? 1.f : 0.f;
        float e1x = my * az, e1y = mz * ax - mx * az, e1z = -my * ax;
        const float n1 = 1.f / sqrtf(e1x * e1x + e1y * e1y + e1z * e1z);
        e1x *= n1;
        e1y *= n1;
        e1z *= n1;
        const float e2x = my * e1z - mz * e1y, e2y = mz * e1x - mx * e1z, e2z = mx * e1y - my * e1x;
        float sp, cp;
        sincosf(phi, &sp, &cp);
        const double dx = w * mx + s * (cp * e1x + sp * e2x);
        const double dy = w * my + s * (cp * e1y + sp * e2y);
        const double dz = w * mz + s * (cp * e1z + sp * e2z);
        const double n = sqrt(dx * dx + dy * dy + dz * dz);
        kx = dx / n;
        ky = dy / n;
        kz = dz / n;
    }

    // guide features of a scattering event (9 values): position / half-size, incoming
    // direction, normalised log wavelength, |r| / half-size, |z| / half-size
    struct GuideFrame
    {
        double inv_half;
        double lam_c, lam_inv_s;
    };
    __device__ __forceinline__ void guide_features(const GuideFrame& gf, double x, double y, double z, double kx, double ky,
                                                   double kz, double lam, float* f)
    {
        f[0] = static_cast<float>(x * gf.inv_half);
        f[1] = static_cast<float>(y * gf.inv_half);
        f[2] = static_cast<float>(z * gf.inv_half);
        f[3] = static_cast<float>(kx);
        f[4] = static_cast<float>(ky);
        f[5] = static_cast<float>(kz);
        f[6] = static_cast<float>((log(lam) - gf.lam_c) * gf.lam_inv_s);
        f[7] = static_cast<float>(sqrt(x * x + y * y + z * z) * gf.inv_half);
        f[8] = static_cast<float>(fabs(z) * gf.inv_half);
    }

    // Wave-level inference: every lane supplies its 9 features (lanes without a sample supply
    // anything) and receives its 40 raw outputs.  Every thread of the workgroup must call this
    // together (it synchronises the workgroup around the LDS staging); each wavefront uses its
    // own staging buffers.
    // lds: 64 x 16 bf16 staging for the inputs plus 64 x 48 floats for the outputs.
    __device__ __forceinline__ void guide_wave_infer(const GuideNet& net, const float* f, float* out40, uint16_t* lds_in,
                                                     float* lds_out)
    {
        const int lane = __lane_id();
        const int row = lane & 15, grp = lane >> 4;
        for (int q = 0; q < 16; ++q) lds_in[lane * 16 + q] = q < kGuideIn ? f2bf(f[q]) : 0;
        __syncthreads();
        for (int t = 0; t < 4; ++t)
        {
            const int s = t * 16 + row;  // sample of this lane's column
            bf16x4 x = load_bf4(lds_in + s * 16 + grp * 4);
            f32x4 o[3];
            guide_tile_forward(net, x, o, nullptr, nullptr, nullptr);
            for (int ot = 0; ot < 3; ++ot)
                for (int q = 0; q < 4; ++q) lds_out[s * 48 + ot * 16 + grp * 4 + q] = o[ot][q];
        }
        __syncthreads();
        for (int q = 0; q < kGuideOut; ++q) out40[q] = lds_out[lane * 48 + q];
    }

    // Wave-level inference held in registers: every lane supplies its 9 features and, where
    // `want` is set, receives its 40 raw outputs.  Inputs are gathered into the B-operand
    // layout and outputs returned to their lanes with cross-lane permutes (ds_bpermute), so
    // no LDS is allocated and no workgroup barrier is needed; 16-lane tiles without a wanting
    // lane are skipped.  Every lane of the wavefront must call it (converged control flow).
    // Not inlined: its registers are then not added to those of the traversal state that is live
    // around the call in the persistent kernel.
    __device__ __attribute__((noinline)) void guide_wave_infer_regs(const GuideNet& net, const float* f, bool want, float* out40)
    {
        const int lane = __lane_id();
        const int row = lane & 15, grp = lane >> 4;
        const unsigned long long need = __ballot(want);
        bool act[4];
        bf16x4 x[4];
#pragma unroll
        for (int t = 0; t < 4; ++t)
        {
            act[t] = ((need >> (16 * t)) & 0xFFFFull) != 0;
            // features 4*grp .. 4*grp+3 of sample t*16 + row (zero beyond the 9 features)
            const int src = t * 16 + row;
            float xs[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int k = 0; k < kGuideIn; ++k)
            {
                float v = __shfl(f[k], src);
                if (grp == (k >> 2)) xs[k & 3] = v;
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) x[t][q] = static_cast<short>(f2bf(xs[q]));
        }
        // the layers in turn, each applied to every 16-sample tile with a wanting lane, so that a
        // layer's weights are loaded once per call (the same MFMA sequence per tile as
        // guide_tile_forward)
        bf16x4 cur[4][4];
        {
            bf16x4 w[4];
            float bias[4][4];
#pragma unroll
            for (int o = 0; o < 4; ++o)
            {
                w[o] = load_bf4(net.W[0] + (o * 16 + row) * 16 + grp * 4);
#pragma unroll
                for (int q = 0; q < 4; ++q) bias[o][q] = net.b[0][o * 16 + grp * 4 + q];
            }
#pragma unroll
            for (int t = 0; t < 4; ++t)
            {
                if (!act[t]) continue;
#pragma unroll
                for (int o = 0; o < 4; ++o)
                {
                    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
                    acc = mfma16(w[o], x[t], acc);
#pragma unroll
                    for (int q = 0; q < 4; ++q) acc[q] = fmaxf(acc[q] + bias[o][q], 0.f);
                    cur[t][o] = pack_bf4(acc);
                }
            }
        }
#pragma unroll
        for (int l = 1; l <= 2; ++l)
        {
            bf16x4 w[4][4];
            float bias[4][4];
#pragma unroll
            for (int o = 0; o < 4; ++o)
            {
#pragma unroll
                for (int k = 0; k < 4; ++k) w[o][k] = load_bf4(net.W[l] + (o * 16 + row) * 64 + k * 16 + grp * 4);
#pragma unroll
                for (int q = 0; q < 4; ++q) bias[o][q] = net.b[l][o * 16 + grp * 4 + q];
            }
#pragma unroll
            for (int t = 0; t < 4; ++t)
            {
                if (!act[t]) continue;
                bf16x4 nxt[4];
#pragma unroll
                for (int o = 0; o < 4; ++o)
                {
                    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                    for (int k = 0; k < 4; ++k) acc = mfma16(w[o][k], cur[t][k], acc);
#pragma unroll
                    for (int q = 0; q < 4; ++q) acc[q] = fmaxf(acc[q] + bias[o][q], 0.f);
                    nxt[o] = pack_bf4(acc);
                }
#pragma unroll
                for (int o = 0; o < 4; ++o) cur[t][o] = nxt[o];
            }
        }
        {
            // output layer: 3 output tiles (48 rows, 40 used); outputs back to their lanes
            bf16x4 w[3][4];
            float bias[3][4];
#pragma unroll
            for (int o = 0; o < 3; ++o)
            {
#pragma unroll
                for (int k = 0; k < 4; ++k) w[o][k] = load_bf4(net.W[3] + (o * 16 + row) * 64 + k * 16 + grp * 4);
#pragma unroll
                for (int q = 0; q < 4; ++q) bias[o][q] = net.b[3][o * 16 + grp * 4 + q];
            }
#pragma unroll
            for (int t = 0; t < 4; ++t)
            {
                if (!act[t]) continue;
                f32x4 o3[3];
#pragma unroll
                for (int o = 0; o < 3; ++o)
                {
                    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                    for (int k = 0; k < 4; ++k) acc = mfma16(w[o][k], cur[t][k], acc);
#pragma unroll
                    for (int q = 0; q < 4; ++q) acc[q] += bias[o][q];
                    o3[o] = acc;
                }
                // output 16*ot + 4*g + q of sample t*16 + r is o3[ot][q] of lane r + 16*g
#pragma unroll
                for (int ot = 0; ot < 3; ++ot)
#pragma unroll
                    for (int g = 0; g < 4; ++g)
#pragma unroll
                        for (int q = 0; q < 4; ++q)
                        {
                            const int j = 16 * ot + 4 * g + q;
                            if (j < kGuideOut)
                            {
                                float v = __shfl(o3[ot][q], row + 16 * g);
                                if (grp == t) out40[j] = v;
                            }
                        }
            }
        }
    }

    // ---------------------------------------------------------------- kernels
    // batch inference: features n x 9 (row-major floats) -> outputs n x 40
    __global__ void __launch_bounds__(64) guide_infer_kernel(GuideNet net, const float* feat, int64_t n, float* out)
    {
        __shared__ uint16_t lin[64 * 16];
        __shared__ float lout[64 * 48];
        const int lane = threadIdx.x;
        const int64_t s = blockIdx.x * 64LL + lane;
        float f[kGuideIn];
        for (int q = 0; q < kGuideIn; ++q) f[q] = s < n ? feat[s * kGuideIn + q] : 0.f;
        float o[kGuideOut];
        guide_wave_infer(net, f, o, lin, lout);
        if (s < n)
            for (int q = 0; q < kGuideOut; ++q) out[s * kGuideOut + q] = o[q];
    }

    // ---------------------------------------------------------------- training
    // Workgroup of the gradient kernel: 4 wavefronts, one 16-sample column tile each (64 samples).
    constexpr int kGradWaves = 4;
    constexpr int kGradSamples = 16 * kGradWaves;
    // LDS pitch (in BF16 elements) of the [unit][sample] staging rows: 68 samples = 136 B =
    // 34 banks, so the 16 rows one 8-byte MFMA operand load touches start on distinct bank pairs
    constexpr int kGradPitch = kGradSamples + 4;

    // Two lobes (j0, j0 + 1) of dL/d(raw outputs) of the weighted negative log-likelihood
    // -cw ln q(k) of one sample, FP64 math.  The four lanes that hold one sample (lanes row,
    // row + 16, row + 32, row + 48) take lobes 0-1, 2-3, 4-5 and 6-7; the softmax sum and the
    // mixture density are reduced over the four.  o: the sample's 40 outputs; d10: the 10
    // derivatives of lobes j0, j0 + 1 (outputs 5 j0 .. 5 j0 + 9).  Returns q(k).
    __device__ __forceinline__ double guide_dout_pair(const float* o, int j0, double kx, double ky, double kz, double cw,
                                                      float* d10)
    {
        double mx = -1e300;
#pragma unroll
        for (int j = 0; j < kGuideLobes; ++j) mx = fmax(mx, static_cast<double>(o[5 * j + 4]));
        double e[2], f[2], kap[2], cdot[2], mu[2][3], nrm[2];
        double es = 0., efs = 0.;
#pragma unroll
        for (int h = 0; h < 2; ++h)
        {
            const int j = j0 + h;
            const double a = o[5 * j], b = o[5 * j + 1], c = o[5 * j + 2];
            double n = sqrt(a * a + b * b + c * c);
            n = n > 1e-12 ? n : 1e-12;
            nrm[h] = n;
            mu[h][0] = a / n;
            mu[h][1] = b / n;
            mu[h][2] = c / n;
            kap[h] = fmin(kGuideKMin + softplus_d(o[5 * j + 3]), kGuideKMax);
            cdot[h] = mu[h][0] * kx + mu[h][1] * ky + mu[h][2] * kz;
            f[h] = kap[h] / (2 * M_PI * (-expm1(-2 * kap[h]))) * exp(kap[h] * (cdot[h] - 1));
            e[h] = exp(static_cast<double>(o[5 * j + 4]) - mx);
            es += e[h];
            efs += e[h] * f[h];
        }
        // sums over the sample's four lanes
        es += __shfl_xor(es, 16);
        es += __shfl_xor(es, 32);
        efs += __shfl_xor(efs, 16);
        efs += __shfl_xor(efs, 32);
        const double q = efs / es;
        const double g = -cw;
#pragma unroll
        for (int h = 0; h < 2; ++h)
        {
            const int j = j0 + h;
            const double w = e[h] / es;
            const double r = w * f[h] / q;
            const double sp = kGuideKMin + softplus_d(o[5 * j + 3]);
            const double dlnf = 1 / kap[h] - 2 * exp(-2 * kap[h]) / (-expm1(-2 * kap[h])) + (cdot[h] - 1);
            const double dk = (sp < kGuideKMax) ? g * r * dlnf : 0.;
            const double sig = 1 / (1 + exp(-static_cast<double>(o[5 * j + 3])));
            const double s = g * r * kap[h];
            const double dmx = s * kx, dmy = s * ky, dmz = s * kz;
            const double proj = dmx * mu[h][0] + dmy * mu[h][1] + dmz * mu[h][2];
            d10[5 * h + 0] = static_cast<float>((dmx - proj * mu[h][0]) / nrm[h]);
            d10[5 * h + 1] = static_cast<float>((dmy - proj * mu[h][1]) / nrm[h]);
            d10[5 * h + 2] = static_cast<float>((dmz - proj * mu[h][2]) / nrm[h]);
            d10[5 * h + 3] = static_cast<float>(dk * sig);
            d10[5 * h + 4] = static_cast<float>(g * (r - w));
        }
        return q;
    }

    // Gradient of the training loss over records [0, n), one partial gradient per workgroup of
    // 64 samples.  Records hold 9 features, the sampled direction (3) and the contribution
    // weight, multiplied by cscale (so a batch's weights sum to ~1).  Each wavefront runs one
    // 16-sample tile forward and backward with the activations and deltas in registers (the
    // MFMA result layout of one layer is the B-operand layout of the next); activations and
    // deltas are also staged in LDS as [unit][sample] rows, from which the four wavefronts
    // contract the weight gradients over the workgroup's 64 samples on MFMA, a share of the
    // output tiles each.  The workgroup writes its partial gradient (FP32, parameter order) to
    // part[blockIdx.x * kGuideParams ...] with plain stores: no atomics, and the partials are
    // summed in a fixed order by guide_reduce_step.  lossOut (if set) accumulates the loss.
    __global__ void __launch_bounds__(64 * kGradWaves) guide_grad_part(GuideNet net, const float* rec, int recStride,
                                                                       int64_t n, float cscale, float* part,
                                                                       double* lossOut)
    {
        constexpr int P = kGradPitch;
        __shared__ __align__(16) uint16_t sXT[16 * P];
        __shared__ __align__(16) uint16_t sHT[3][64 * P];
        __shared__ __align__(16) uint16_t sDT[2][64 * P];
        __shared__ float sO[kGradSamples * 48];
        __shared__ double sLoss[kGradWaves];
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        const int row = lane & 15, grp = lane >> 4;
        const int sl = 16 * wave + row;  // this lane's sample (B-operand column) in the workgroup
        const int64_t s = blockIdx.x * static_cast<int64_t>(kGradSamples) + sl;
        const bool valid = s < n;
        float* gp = part + static_cast<int64_t>(blockIdx.x) * kGuideParams;

        // inputs: features 4 grp .. 4 grp + 3 of the sample (zero beyond the 9 features)
        bf16x4 x;
#pragma unroll
        for (int q = 0; q < 4; ++q)
        {
            const int k = 4 * grp + q;
            const uint16_t v = (valid && k < kGuideIn) ? f2bf(rec[s * recStride + k]) : 0;
            x[q] = static_cast<short>(v);
            sXT[k * P + sl] = v;
        }
        // forward, activations kept for the backward pass
        f32x4 o[3];
        bf16x4 h[3][4];
        guide_tile_forward(net, x, o, h[0], h[1], h[2]);
#pragma unroll
        for (int l = 0; l < 3; ++l)
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int q = 0; q < 4; ++q) sHT[l][(u * 16 + 4 * grp + q) * P + sl] = static_cast<uint16_t>(h[l][u][q]);
#pragma unroll
        for (int ot = 0; ot < 3; ++ot)
#pragma unroll
            for (int q = 0; q < 4; ++q) sO[sl * 48 + ot * 16 + 4 * grp + q] = o[ot][q];
        __syncthreads();

        // loss derivative: this lane's two lobes of its sample, staged as the delta of layer 3
        {
            double kx = 0., ky = 0., kz = 1., cw = 0.;
            if (valid)
            {
                const float* r = rec + s * recStride;
                kx = r[9];
                ky = r[10];
                kz = r[11];
                cw = static_cast<double>(r[12]) * cscale;
            }
            float d10[10];
            // every lane takes part (the lobe sums are shuffles); lanes past n contribute zero
            const double q = guide_dout_pair(sO + sl * 48, 2 * grp, kx, ky, kz, cw, d10);
#pragma unroll
            for (int k = 0; k < 10; ++k) sDT[0][(10 * grp + k) * P + sl] = valid ? f2bf(d10[k]) : 0;
            // padding outputs 40..47
            sDT[0][(40 + 2 * grp) * P + sl] = 0;
            sDT[0][(41 + 2 * grp) * P + sl] = 0;
            double lossv = (valid && grp == 0) ? -cw * log(q) : 0.;
            for (int m = 32; m > 0; m >>= 1) lossv += __shfl_xor(lossv, m);
            if (lane == 0) sLoss[wave] = lossv;
        }
        __syncthreads();

        // backward, layer 3 down to 0; the delta of layer l is in sDT[(3 - l) & 1]
        bf16x4 dcur[4];
#pragma unroll
        for (int kt = 0; kt < 3; ++kt)
#pragma unroll
            for (int q = 0; q < 4; ++q) dcur[kt][q] = static_cast<short>(sDT[0][(kt * 16 + 4 * grp + q) * P + sl]);
        dcur[3] = bf16x4{0, 0, 0, 0};
#pragma unroll
        for (int l = 3; l >= 0; --l)
        {
            const uint16_t* D = sDT[(3 - l) & 1];
            const uint16_t* Hin = l == 0 ? sXT : sHT[l - 1];
            const int outp = guide_out_pad(l), inp = guide_in_pad(l);
            const int ot_n = outp / 16, it_n = inp / 16;
            const int wOff = l == 0 ? kOffW0 : l == 1 ? kOffW1 : l == 2 ? kOffW2 : kOffW3;
            const int bOff = l == 0 ? kOffB0 : l == 1 ? kOffB1 : l == 2 ? kOffB2 : kOffB3;
            const int nout = l == 3 ? kGuideOut : kGuideHidden, nin = l == 0 ? kGuideIn : kGuideHidden;
            // dW_l (out x in) = sum over samples of delta[s][o] Hin[s][i]: A = delta^T (row o,
            // k = sample), B = Hin (k = sample, column i); output tiles dealt round the waves
            for (int t = wave; t < ot_n * it_n; t += kGradWaves)
            {
                const int ot = t / it_n, it = t % it_n;
                f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int kb = 0; kb < kGradSamples / 16; ++kb)
                    acc = mfma16(load_bf4(D + (ot * 16 + row) * P + kb * 16 + 4 * grp),
                                 load_bf4(Hin + (it * 16 + row) * P + kb * 16 + 4 * grp), acc);
                // rows o = ot*16 + 4 grp + q, column i = it*16 + row
                const int i = it * 16 + row;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                {
                    const int oo = ot * 16 + 4 * grp + q;
                    if (oo < nout && i < nin) gp[wOff + oo * nin + i] = acc[q];
                }
            }
            // bias gradient: the delta summed over the workgroup's samples
            if (threadIdx.x < nout)
            {
                float sb = 0.f;
#pragma unroll
                for (int k = 0; k < kGradSamples; k += 4)
                {
                    const bf16x4 v = load_bf4(D + threadIdx.x * P + k);
                    sb += (bf2f(static_cast<uint16_t>(v[0])) + bf2f(static_cast<uint16_t>(v[1]))) +
                          (bf2f(static_cast<uint16_t>(v[2])) + bf2f(static_cast<uint16_t>(v[3])));
                }
                gp[bOff + threadIdx.x] = sb;
            }
            if (l == 0) break;
            // delta of layer l-1 for this wave's tile: (W_l^T delta^T) * relu'(h_{l-1}); A = W_l^T
            // tiles (row = input unit, k = output unit), B = the delta in registers
            bf16x4 dn[4];
            uint16_t* Dn = sDT[(4 - l) & 1];
#pragma unroll
            for (int it = 0; it < 4; ++it)
            {
                f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int kt = 0; kt < 4; ++kt)
                    if (kt < ot_n) acc = mfma16(load_bf4(net.WT[l] + (it * 16 + row) * outp + kt * 16 + 4 * grp), dcur[kt], acc);
#pragma unroll
                for (int q = 0; q < 4; ++q)
                {
                    const uint16_t v = h[l - 1][it][q] > 0 ? f2bf(acc[q]) : 0;
                    dn[it][q] = static_cast<short>(v);
                    Dn[(it * 16 + 4 * grp + q) * P + sl] = v;
                }
            }
#pragma unroll
            for (int it = 0; it < 4; ++it) dcur[it] = dn[it];
            __syncthreads();
        }
        if (threadIdx.x == 0 && lossOut)
        {
            double lw = 0.;
            for (int w = 0; w < kGradWaves; ++w) lw += sLoss[w];
            atomicAdd(lossOut, lw);
        }
    }

    // sum of the training weights (record field 12) over n records, one FP64 atomic per wave
    __global__ void guide_weight_sum(const float* rec, int64_t n, double* out)
    {
        double acc = 0.;
        for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
             i += static_cast<int64_t>(gridDim.x) * blockDim.x)
            acc += rec[16 * i + 12];
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
        if ((threadIdx.x & 63) == 0) atomicAdd(out, acc);
    }

    // BF16 tiles and FP32 biases of the network, writable (the storage GuideNet reads)
    struct GuideTilesW
    {
        uint16_t* W[4];
        uint16_t* WT[4];
        float* b[4];
    };

    // Launch shape of guide_reduce_step: 32 parameters per workgroup, the partials of each
    // summed in 8 slices (workgroups b = slice, slice + 8, ...) and the slices added in LDS.
    constexpr int kReduceParams = 32;
    constexpr int kReduceSlices = 8;

    // Per parameter: the sum of the nb workgroups' partial gradients, in a fixed order (the
    // slices' sums in slice order).  Without theta it is written to grad.  With theta, an Adam
    // step is taken on theta (moments m, v; bias corrections c1, c2) and the parameter's BF16
    // tile entries (W and W^T) or FP32 bias are rewritten; the tiles' padding stays as
    // guide_pack_kernel left it (zero).
    __global__ void __launch_bounds__(kReduceParams * kReduceSlices)
        guide_reduce_step(const float* part, int nb, float* grad, float* theta, float* m, float* v, float lr, float b1,
                          float b2, float eps, float c1, float c2, GuideTilesW tw)
    {
        __shared__ float red[kReduceSlices][kReduceParams];
        const int c = threadIdx.x % kReduceParams, r = threadIdx.x / kReduceParams;
        const int i = blockIdx.x * kReduceParams + c;
        float g = 0.f;
        if (i < kGuideParams)
        {
            // four independent sums, so that four loads are in flight per thread
            float a[4] = {0.f, 0.f, 0.f, 0.f};
            int b = r;
            for (; b + 3 * kReduceSlices < nb; b += 4 * kReduceSlices)
#pragma unroll
                for (int k = 0; k < 4; ++k) a[k] += part[static_cast<int64_t>(b + k * kReduceSlices) * kGuideParams + i];
            for (; b < nb; b += kReduceSlices) a[0] += part[static_cast<int64_t>(b) * kGuideParams + i];
            g = (a[0] + a[1]) + (a[2] + a[3]);
        }
        red[r][c] = g;
        __syncthreads();
        if (r != 0 || i >= kGuideParams) return;
        g = red[0][c];
#pragma unroll
        for (int k = 1; k < kReduceSlices; ++k) g += red[k][c];
        if (!theta)
        {
            grad[i] = g;
            return;
        }
        const float mi = b1 * m[i] + (1.f - b1) * g;
        const float vi = b2 * v[i] + (1.f - b2) * g * g;
        m[i] = mi;
        v[i] = vi;
        const float th = theta[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        theta[i] = th;
        const int wOff[4] = {kOffW0, kOffW1, kOffW2, kOffW3};
        const int bOff[4] = {kOffB0, kOffB1, kOffB2, kOffB3};
        int l = 3;
        while (l > 0 && i < wOff[l]) --l;
        if (i >= bOff[l])
        {
            tw.b[l][i - bOff[l]] = th;
            return;
        }
        const int nin = l == 0 ? kGuideIn : kGuideHidden;
        const int j = i - wOff[l], o = j / nin, cc = j % nin;
        const uint16_t w = f2bf(th);
        tw.W[l][o * guide_in_pad(l) + cc] = w;
        tw.WT[l][cc * guide_out_pad(l) + o] = w;
    }

    // rebuilds the padded BF16 weight tiles (and transposes) and FP32 biases from theta
    __global__ void guide_pack_kernel(const float* theta, uint16_t* W0, uint16_t* W1, uint16_t* W2, uint16_t* W3,
                                      uint16_t* T0, uint16_t* T1, uint16_t* T2, uint16_t* T3, float* B0, float* B1,
                                      float* B2, float* B3)
    {
        int i = blockIdx.x * blockDim.x + threadIdx.x;
        // layer l: out_pad x in_pad entries
        const int sizes[4] = {64 * 16, 64 * 64, 64 * 64, 48 * 64};
        int l = 0, base = 0;
        while (l < 4 && i >= base + sizes[l])
        {
            base += sizes[l];
            ++l;
        }
        if (l >= 4) return;
        const int j = i - base;
        const int inp = guide_in_pad(l), outp = guide_out_pad(l);
        const int o = j / inp, c = j % inp;
        const int nin = l == 0 ? kGuideIn : kGuideHidden, nout = l == 3 ? kGuideOut : kGuideHidden;
        const int wOff = l == 0 ? kOffW0 : l == 1 ? kOffW1 : l == 2 ? kOffW2 : kOffW3;
        const int bOff = l == 0 ? kOffB0 : l == 1 ? kOffB1 : l == 2 ? kOffB2 : kOffB3;
        float w = (o < nout && c < nin) ? theta[wOff + o * nin + c] : 0.f;
        uint16_t* W = l == 0 ? W0 : l == 1 ? W1 : l == 2 ? W2 : W3;
        uint16_t* T = l == 0 ? T0 : l == 1 ? T1 : l == 2 ? T2 : T3;
        float* B = l == 0 ? B0 : l == 1 ? B1 : l == 2 ? B2 : B3;
        W[o * inp + c] = f2bf(w);
        T[c * outp + o] = f2bf(w);
        if (c == 0) B[o] = o < nout ? theta[bOff + o] : 0.f;
    }
}
