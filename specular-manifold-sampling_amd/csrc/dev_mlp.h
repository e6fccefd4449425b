// dev_mlp.h -- the learned seed proposal's network on gfx950 matrix cores.
//
// 9 inputs (shading point, shading normal, light sample; positions normalised to the scene
// box) -> 8-frequency Fourier features (9 raw + 9 x 8 x {sin, cos} = 153, zero-padded to 160)
// -> 4 ReLU layers of width 64 -> 48 outputs = 8 Gaussian components x (weight logit, mean u,
// mean v, log sigma_u, log sigma_v, correlation).  Weights and activations are BF16, every
// product accumulates in FP32 on v_mfma_f32_16x16x16_bf16; biases, the mixture head, the loss
// and the optimiser are FP32.  A block of 256 lanes owns a tile of 64 samples; weights and the
// tile's activations live in LDS (about 123 KB), in rows padded by 4 entries so that the 16 rows
// an MFMA operand spans fall on different LDS banks.  Wavefront w computes the 16 samples
// 16 w .. 16 w + 15 of a tile against every output column (one accumulator per 16 columns).
//
// Parameter vector (FP32 master copy, gradients, Adam moments share it): for each layer the
// weights stored output-major W[n][k] (K padded: 160 for layer 1), then the bias.
#pragma once
#include "dev_math.h"

namespace smsdev {

constexpr int MLP_IN = 9, MLP_FREQ = 8, MLP_FEAT = 153, MLP_K1 = 160, MLP_H = 64, MLP_OUT = 48, MLP_COMP = 8;
constexpr int MLP_TILE = 64;
constexpr int MLP_W1 = 0, MLP_B1 = MLP_W1 + MLP_H * MLP_K1;
constexpr int MLP_W2 = MLP_B1 + MLP_H, MLP_B2 = MLP_W2 + MLP_H * MLP_H;
constexpr int MLP_W3 = MLP_B2 + MLP_H, MLP_B3 = MLP_W3 + MLP_H * MLP_H;
constexpr int MLP_W4 = MLP_B3 + MLP_H, MLP_B4 = MLP_W4 + MLP_H * MLP_H;
constexpr int MLP_W5 = MLP_B4 + MLP_H, MLP_B5 = MLP_W5 + MLP_OUT * MLP_H;
constexpr int MLP_PARAMS = MLP_B5 + MLP_OUT;
// GMM head limits
constexpr float MLP_LOGSIG_MIN = -7.f, MLP_LOGSIG_MAX = 0.f, MLP_RHO_MAX = 0.95f;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16 bits, round to nearest even (NaN kept quiet)
SMS_DEV uint16_t to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t) ((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t) (u >> 16);
}
SMS_DEV float from_bf16(uint16_t h) { return __uint_as_float((uint32_t) h << 16); }

// Workgroup barrier for LDS traffic only: LDS writes before it are visible after it, while
// global stores in flight (the gradient rows) need not complete, unlike __syncthreads.
SMS_DEV void lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// LDS row strides (BF16 entries) of 160- and 64-wide rows
constexpr int MLP_S1 = MLP_K1 + 4, MLP_SH = MLP_H + 4;

// v_mfma_f32_16x16x16_bf16: lane l supplies A[l % 16][4 (l / 16) + i] and B[4 (l / 16) + i][l % 16]
// (i = 0..3) and receives C[4 (l / 16) + r][l % 16] in acc[r].
SMS_DEV f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 acc) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0); }
// four consecutive BF16 entries (8-byte aligned) / four entries `stride` apart
SMS_DEV bf16x4 lds4(const uint16_t *p) { return *(const bf16x4 *) p; }
SMS_DEV bf16x4 lds4s(const uint16_t *p, int stride) {
    bf16x4 v;
    v[0] = (short) p[0]; v[1] = (short) p[stride]; v[2] = (short) p[2 * stride]; v[3] = (short) p[3 * stride];
    return v;
}

// Fourier features of one sample (9 inputs in [-1, 1]) into a BF16 row of MLP_K1 entries;
// part q of `parts` writes frequencies f = q, q + parts, ... (part 0 also the raw inputs and
// the zero padding), so several lanes can share a row.  sin / cos of the FP32 argument
// pi 2^f x come from the hardware's native sine and cosine (absolute error about 1e-6 on
// arguments up to 128 pi, far below the BF16 rounding of the result).
SMS_DEV void mlp_features(const float *x9, uint16_t *row, int q = 0, int parts = 1) {
    if (q == 0) {
        for (int i = 0; i < MLP_IN; ++i) row[i] = to_bf16(x9[i]);
        for (int k = MLP_FEAT; k < MLP_K1; ++k) row[k] = 0;
    }
    for (int f = q; f < MLP_FREQ; f += parts) {
        const float w = PI_F * (float) (1 << f);
        for (int i = 0; i < MLP_IN; ++i) {
            const float a = w * x9[i];
            row[MLP_IN + (2 * f) * MLP_IN + i] = to_bf16(__sinf(a));
            row[MLP_IN + (2 * f + 1) * MLP_IN + i] = to_bf16(__cosf(a));
        }
    }
}

SMS_DEV float clampf_(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }

// Negative log-likelihood of uv under the mixture of raw outputs o[48] and its gradient
// w.r.t. the raw outputs (times `scale`), computed by a group of 4 adjacent lanes: lane
// quarter `c4` owns components 2 c4 and 2 c4 + 1 and writes d_o[0..11] = their 12 entries.
// Every lane of the group returns the same loss (all 64 lanes of the wavefront must call).
// Whitened form: z1 = (u - mu_u) / s_u, z2 = ((v - mu_v) / s_v - rho z1) / sqrt(1 - rho^2).
// tanh from the native exponential: sign(x) (1 - 2 / (e^{2|x|} + 1))
SMS_DEV float tanh_native(float x) {
    const float t = 1.f - __fdividef(2.f, __expf(2.f * fabsf(x)) + 1.f);
    return x < 0.f ? -t : t;
}

SMS_DEV float gmm_nll_grad4(const float *o, float u, float v, float scale, int c4, float *d_o) {
    float mx = o[0];
    for (int c = 1; c < MLP_COMP; ++c) mx = fmax_(mx, o[6 * c]);
    float wsum = 0.f;
    for (int c = 0; c < MLP_COMP; ++c) wsum += __expf(o[6 * c] - mx);
    float a[2], z1[2], z2[2], v2[2], q[2], wc[2], mu1[2], mu2[2], sx[2], sy[2], rho[2];
    float amax = -INFINITY;
    for (int j = 0; j < 2; ++j) {
        const float *oc = o + 6 * (2 * c4 + j);
        wc[j] = __fdividef(__expf(oc[0] - mx), wsum);
        mu1[j] = __fdividef(1.f, 1.f + __expf(-oc[1]));
        mu2[j] = __fdividef(1.f, 1.f + __expf(-oc[2]));
        sx[j] = __expf(clampf_(oc[3], MLP_LOGSIG_MIN, MLP_LOGSIG_MAX));
        sy[j] = __expf(clampf_(oc[4], MLP_LOGSIG_MIN, MLP_LOGSIG_MAX));
        rho[j] = MLP_RHO_MAX * tanh_native(oc[5]);
        q[j] = sqrtf(1.f - rho[j] * rho[j]);
        z1[j] = __fdividef(u - mu1[j], sx[j]);
        v2[j] = __fdividef(v - mu2[j], sy[j]);
        z2[j] = __fdividef(v2[j] - rho[j] * z1[j], q[j]);
        const float logn = -0.5f * (z1[j] * z1[j] + z2[j] * z2[j]) - 1.8378770664093453f - __logf(sx[j]) - __logf(sy[j]) - __logf(q[j]);
        a[j] = __logf(wc[j]) + logn;
        amax = fmax_(amax, a[j]);
    }
    amax = fmax_(amax, __shfl_xor(amax, 1, 64));
    amax = fmax_(amax, __shfl_xor(amax, 2, 64));
    float s = __expf(a[0] - amax) + __expf(a[1] - amax);
    s += __shfl_xor(s, 1, 64);
    s += __shfl_xor(s, 2, 64);
    const float loss = -(amax + __logf(s));
    for (int j = 0; j < 2; ++j) {
        const float gam = __fdividef(__expf(a[j] - amax), s);   // responsibility
        const float r = rho[j], qc = q[j];
        const float *oc = o + 6 * (2 * c4 + j);
        float *dc = d_o + 6 * j;
        dc[0] = scale * (wc[j] - gam);                      // softmax over the weight logits
        const float iq = __fdividef(1.f, qc), isx = __fdividef(1.f, sx[j]), isy = __fdividef(1.f, sy[j]);
        const float dmu1 = z1[j] * isx - z2[j] * r * iq * isx;
        const float dmu2 = z2[j] * iq * isy;
        const float dls1 = z1[j] * z1[j] - z2[j] * r * z1[j] * iq - 1.f;
        const float dls2 = z2[j] * v2[j] * iq - 1.f;
        const float dz2_drho = -z1[j] * iq + z2[j] * r * iq * iq;
        const float drho = -z2[j] * dz2_drho + r * iq * iq;
        const float in3 = (oc[3] > MLP_LOGSIG_MIN && oc[3] < MLP_LOGSIG_MAX) ? 1.f : 0.f;
        const float in4 = (oc[4] > MLP_LOGSIG_MIN && oc[4] < MLP_LOGSIG_MAX) ? 1.f : 0.f;
        const float th = r * (1.f / MLP_RHO_MAX);
        dc[1] = scale * (-gam) * dmu1 * (mu1[j] * (1.f - mu1[j]));
        dc[2] = scale * (-gam) * dmu2 * (mu2[j] * (1.f - mu2[j]));
        dc[3] = scale * (-gam) * dls1 * in3;
        dc[4] = scale * (-gam) * dls2 * in4;
        dc[5] = scale * (-gam) * drho * (MLP_RHO_MAX * (1.f - th * th));
    }
    return loss;
}

// LDS of one MLP block: BF16 weights (output-major, padded rows) and biases, the tile's
// activations (sample-major, padded rows).
struct alignas(16) MlpLds {
    uint16_t w1[MLP_H][MLP_S1], w2[MLP_H][MLP_SH], w3[MLP_H][MLP_SH], w4[MLP_H][MLP_SH], w5[MLP_OUT][MLP_SH];
    float b1[MLP_H], b2[MLP_H], b3[MLP_H], b4[MLP_H], b5[MLP_OUT];
    uint16_t x0[MLP_TILE][MLP_S1];
    uint16_t h[4][MLP_TILE][MLP_SH];
    float o[MLP_TILE][MLP_OUT];
};

// copy K-wide weight rows (global, 16-byte aligned, row-contiguous) into LDS rows of stride S
template <int ROWS, int K, int S>
SMS_DEV void mlp_copy_rows(const uint16_t *__restrict__ src, uint16_t *dst) {
    static_assert(K % 8 == 0 && S % 4 == 0, "row layout");
    for (int e = threadIdx.x; e < ROWS * K / 8; e += blockDim.x) {
        const int r = e / (K / 8), c = (e % (K / 8)) * 8;
        const uint4 v = ((const uint4 *) src)[e];
        uint2 *d = (uint2 *) (dst + r * S + c);
        d[0] = make_uint2(v.x, v.y);
        d[1] = make_uint2(v.z, v.w);
    }
}

// copy the BF16 weights and FP32 biases of the parameter vector into LDS (all lanes; every
// layer's weight block starts 16-byte aligned in the parameter vector)
SMS_DEV void mlp_load_weights(const uint16_t *__restrict__ wb, const float *__restrict__ p, MlpLds &L) {
    static_assert(MLP_W2 % 8 == 0 && MLP_W3 % 8 == 0 && MLP_W4 % 8 == 0 && MLP_W5 % 8 == 0, "weight alignment");
    mlp_copy_rows<MLP_H, MLP_K1, MLP_S1>(wb + MLP_W1, &L.w1[0][0]);
    mlp_copy_rows<MLP_H, MLP_H, MLP_SH>(wb + MLP_W2, &L.w2[0][0]);
    mlp_copy_rows<MLP_H, MLP_H, MLP_SH>(wb + MLP_W3, &L.w3[0][0]);
    mlp_copy_rows<MLP_H, MLP_H, MLP_SH>(wb + MLP_W4, &L.w4[0][0]);
    mlp_copy_rows<MLP_OUT, MLP_H, MLP_SH>(wb + MLP_W5, &L.w5[0][0]);
    for (int i = threadIdx.x; i < MLP_H; i += blockDim.x) {
        L.b1[i] = p[MLP_B1 + i]; L.b2[i] = p[MLP_B2 + i]; L.b3[i] = p[MLP_B3 + i]; L.b4[i] = p[MLP_B4 + i];
    }
    for (int i = threadIdx.x; i < MLP_OUT; i += blockDim.x) L.b5[i] = p[MLP_B5 + i];
}

// Y = act(X W^T + b) for the tile: X rows (BF16, stride XS), W output-major (stride WS), both in
// LDS.  Wavefront w owns samples 16 w .. 16 w + 15 and all N outputs (N / 16 accumulators).
// relu: BF16 rows of Y (stride MLP_SH); otherwise FP32 rows to `of` (stride N).
template <int K, int N, int XS, int WS, bool RELU>
SMS_DEV void mlp_layer(const uint16_t *x, const uint16_t *w, const float *b, uint16_t *y, float *of) {
    const int mt = threadIdx.x >> 6, lane = threadIdx.x & 63, r16 = lane & 15, kq = (lane >> 4) * 4;
    constexpr int NT = N / 16;
    const uint16_t *xa = x + (mt * 16 + r16) * XS + kq;
    const uint16_t *wa = w + r16 * WS + kq;
    f32x4 acc[NT];
    for (int j = 0; j < NT; ++j) acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k0 = 0; k0 < K; k0 += 16) {
        const bf16x4 a = lds4(xa + k0);
#pragma unroll
        for (int j = 0; j < NT; ++j) acc[j] = mfma16(a, lds4(wa + j * 16 * WS + k0), acc[j]);
    }
#pragma unroll
    for (int j = 0; j < NT; ++j) {
        const int col = j * 16 + r16;
        for (int r = 0; r < 4; ++r) {
            const int row = mt * 16 + (lane >> 4) * 4 + r;
            const float val = acc[j][r] + b[col];
            if (RELU) y[row * MLP_SH + col] = to_bf16(val > 0.f ? val : 0.f);
            else of[row * N + col] = val;
        }
    }
}

// forward pass of the tile whose features are in L.x0; raw outputs in L.o
SMS_DEV void mlp_forward_tile(MlpLds &L) {
    mlp_layer<MLP_K1, MLP_H, MLP_S1, MLP_S1, true>(&L.x0[0][0], &L.w1[0][0], L.b1, &L.h[0][0][0], nullptr);
    lds_sync();
    mlp_layer<MLP_H, MLP_H, MLP_SH, MLP_SH, true>(&L.h[0][0][0], &L.w2[0][0], L.b2, &L.h[1][0][0], nullptr);
    lds_sync();
    mlp_layer<MLP_H, MLP_H, MLP_SH, MLP_SH, true>(&L.h[1][0][0], &L.w3[0][0], L.b3, &L.h[2][0][0], nullptr);
    lds_sync();
    mlp_layer<MLP_H, MLP_H, MLP_SH, MLP_SH, true>(&L.h[2][0][0], &L.w4[0][0], L.b4, &L.h[3][0][0], nullptr);
    lds_sync();
    mlp_layer<MLP_H, MLP_OUT, MLP_SH, MLP_SH, false>(&L.h[3][0][0], &L.w5[0][0], L.b5, nullptr, &L.o[0][0]);
    lds_sync();
}

} // namespace smsdev
