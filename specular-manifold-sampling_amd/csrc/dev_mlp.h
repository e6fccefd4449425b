// dev_mlp.h -- the learned seed proposal's network on gfx950 matrix cores.
//
// 9 inputs (shading point, shading normal, light sample; positions normalised to the scene
// box) -> 8-frequency Fourier features (9 raw + 9 x 8 x {sin, cos} = 153, zero-padded to 160)
// -> 4 ReLU layers of width 64 -> 48 outputs = 8 Gaussian components x (weight logit, mean u,
// mean v, log sigma_u, log sigma_v, correlation).  Weights and activations are BF16, every
// product accumulates in FP32 on v_mfma_f32_16x16x16_bf16; biases, the mixture head, the loss
// and the optimiser are FP32.  A block of 256 lanes owns a tile of 64 samples; weights and the
// tile's activations live in LDS (about 117 KB).
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

// One 16x16 output tile of C = A[16 x K] * B[K x 16] on the matrix cores.  A(m, k) and B(k, n)
// return BF16 bits.  Lane l supplies A[l % 16][4 (l / 16) + i] and B[4 (l / 16) + i][l % 16]
// (i = 0..3) for each 16-deep step and receives C[4 (l / 16) + r][l % 16] in acc[r].
template <class FA, class FB>
SMS_DEV f32x4 mfma_tile(const FA &A, const FB &B, int K, f32x4 acc) {
    const int lane = threadIdx.x & 63, row = lane & 15, kq = (lane >> 4) * 4;
    for (int k0 = 0; k0 < K; k0 += 16) {
        bf16x4 a, b;
        for (int i = 0; i < 4; ++i) {
            a[i] = (short) A(row, k0 + kq + i);
            b[i] = (short) B(k0 + kq + i, row);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    return acc;
}

// Fourier features of one sample (9 inputs in [-1, 1]) into a BF16 row of MLP_K1 entries
SMS_DEV void mlp_features(const float *x9, uint16_t *row) {
    for (int i = 0; i < MLP_IN; ++i) row[i] = to_bf16(x9[i]);
    for (int f = 0; f < MLP_FREQ; ++f) {
        const float w = PI_F * (float) (1 << f);
        for (int i = 0; i < MLP_IN; ++i) {
            float s, c;
            sincosf(w * x9[i], &s, &c);
            row[MLP_IN + (2 * f) * MLP_IN + i] = to_bf16(s);
            row[MLP_IN + (2 * f + 1) * MLP_IN + i] = to_bf16(c);
        }
    }
    for (int k = MLP_FEAT; k < MLP_K1; ++k) row[k] = 0;
}

// Mixture head of one sample: raw outputs o[48] -> component weights, means, Cholesky factors.
struct GMM {
    float w[MLP_COMP], mu[MLP_COMP][2], sx[MLP_COMP], sy[MLP_COMP], rho[MLP_COMP];
};
SMS_DEV float clampf_(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }
SMS_DEV void gmm_from_raw(const float *o, GMM &g) {
    float mx = o[0];
    for (int c = 1; c < MLP_COMP; ++c) mx = fmax_(mx, o[6 * c]);
    float sum = 0.f;
    for (int c = 0; c < MLP_COMP; ++c) { g.w[c] = expf(o[6 * c] - mx); sum += g.w[c]; }
    for (int c = 0; c < MLP_COMP; ++c) {
        g.w[c] = g.w[c] / sum;
        g.mu[c][0] = 1.f / (1.f + expf(-o[6 * c + 1]));
        g.mu[c][1] = 1.f / (1.f + expf(-o[6 * c + 2]));
        g.sx[c] = expf(clampf_(o[6 * c + 3], MLP_LOGSIG_MIN, MLP_LOGSIG_MAX));
        g.sy[c] = expf(clampf_(o[6 * c + 4], MLP_LOGSIG_MIN, MLP_LOGSIG_MAX));
        g.rho[c] = MLP_RHO_MAX * tanhf(o[6 * c + 5]);
    }
}

// Negative log-likelihood of uv under the mixture and its gradient w.r.t. the raw outputs
// (scaled by `scale`).  Returns the loss.
SMS_DEV float gmm_nll_grad(const float *o, float u, float v, float scale, float *d_o) {
    GMM g;
    gmm_from_raw(o, g);
    float a[MLP_COMP], z1[MLP_COMP], z2[MLP_COMP], v2[MLP_COMP], q[MLP_COMP];
    float amax = -INFINITY;
    for (int c = 0; c < MLP_COMP; ++c) {
        q[c] = sqrtf(1.f - g.rho[c] * g.rho[c]);
        z1[c] = (u - g.mu[c][0]) / g.sx[c];
        v2[c] = (v - g.mu[c][1]) / g.sy[c];
        z2[c] = (v2[c] - g.rho[c] * z1[c]) / q[c];
        float logn = -0.5f * (z1[c] * z1[c] + z2[c] * z2[c]) - 1.8378770664093453f - logf(g.sx[c]) - logf(g.sy[c]) - logf(q[c]);
        a[c] = logf(g.w[c]) + logn;
        amax = fmax_(amax, a[c]);
    }
    float s = 0.f;
    for (int c = 0; c < MLP_COMP; ++c) s += expf(a[c] - amax);
    const float loss = -(amax + logf(s));
    for (int c = 0; c < MLP_COMP; ++c) {
        const float gam = expf(a[c] - amax) / s;           // responsibility
        const float r = g.rho[c], qc = q[c];
        // d loss / d logit (softmax over the weights)
        d_o[6 * c] = scale * (g.w[c] - gam);
        // d log N / d mu, d log sigma, d rho
        const float dmu1 = z1[c] / g.sx[c] - z2[c] * r / (qc * g.sx[c]);
        const float dmu2 = z2[c] / (qc * g.sy[c]);
        const float dls1 = z1[c] * z1[c] - z2[c] * r * z1[c] / qc - 1.f;
        const float dls2 = z2[c] * v2[c] / qc - 1.f;
        const float dz2_drho = -z1[c] / qc + z2[c] * r / (qc * qc);
        const float drho = -z2[c] * dz2_drho + r / (qc * qc);
        const float m1 = g.mu[c][0], m2 = g.mu[c][1];
        const float o3 = o[6 * c + 3], o4 = o[6 * c + 4];
        const float in3 = (o3 > MLP_LOGSIG_MIN && o3 < MLP_LOGSIG_MAX) ? 1.f : 0.f;
        const float in4 = (o4 > MLP_LOGSIG_MIN && o4 < MLP_LOGSIG_MAX) ? 1.f : 0.f;
        const float th = r / MLP_RHO_MAX;
        d_o[6 * c + 1] = scale * (-gam) * dmu1 * (m1 * (1.f - m1));
        d_o[6 * c + 2] = scale * (-gam) * dmu2 * (m2 * (1.f - m2));
        d_o[6 * c + 3] = scale * (-gam) * dls1 * in3;
        d_o[6 * c + 4] = scale * (-gam) * dls2 * in4;
        d_o[6 * c + 5] = scale * (-gam) * drho * (MLP_RHO_MAX * (1.f - th * th));
    }
    return loss;
}

// LDS of one MLP block: BF16 weights (output-major) and biases, the tile's activations.
struct MlpLds {
    uint16_t w1[MLP_H][MLP_K1], w2[MLP_H][MLP_H], w3[MLP_H][MLP_H], w4[MLP_H][MLP_H], w5[MLP_OUT][MLP_H];
    float b1[MLP_H], b2[MLP_H], b3[MLP_H], b4[MLP_H], b5[MLP_OUT];
    uint16_t x0[MLP_TILE][MLP_K1];
    uint16_t h[4][MLP_TILE][MLP_H];
    float o[MLP_TILE][MLP_OUT];
};

// copy the BF16 weights and FP32 biases of the parameter vector into LDS (all lanes)
SMS_DEV void mlp_load_weights(const uint16_t *__restrict__ wb, const float *__restrict__ p, MlpLds &L) {
    for (int i = threadIdx.x; i < MLP_H * MLP_K1; i += blockDim.x) (&L.w1[0][0])[i] = wb[MLP_W1 + i];
    for (int i = threadIdx.x; i < MLP_H * MLP_H; i += blockDim.x) {
        (&L.w2[0][0])[i] = wb[MLP_W2 + i];
        (&L.w3[0][0])[i] = wb[MLP_W3 + i];
        (&L.w4[0][0])[i] = wb[MLP_W4 + i];
    }
    for (int i = threadIdx.x; i < MLP_OUT * MLP_H; i += blockDim.x) (&L.w5[0][0])[i] = wb[MLP_W5 + i];
    for (int i = threadIdx.x; i < MLP_H; i += blockDim.x) {
        L.b1[i] = p[MLP_B1 + i]; L.b2[i] = p[MLP_B2 + i]; L.b3[i] = p[MLP_B3 + i]; L.b4[i] = p[MLP_B4 + i];
    }
    for (int i = threadIdx.x; i < MLP_OUT; i += blockDim.x) L.b5[i] = p[MLP_B5 + i];
}

// Y = act(X W + b) for the tile: X rows in LDS (BF16, row stride xs), W output-major in LDS.
// relu: write BF16 rows of Y (stride MLP_H); otherwise write FP32 rows to `of` (stride n).
template <int K, int N, bool RELU>
SMS_DEV void mlp_layer(const uint16_t *x, int xs, const uint16_t *w, const float *b, uint16_t *y, float *of) {
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    constexpr int NT = N / 16, TILES = 4 * NT;
    for (int t = wv; t < TILES; t += 4) {
        const int mt = t / NT, nt = t % NT;
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma_tile([&](int m, int k) { return x[(mt * 16 + m) * xs + k]; },
                        [&](int k, int n) { return w[(nt * 16 + n) * K + k]; }, K, acc);
        const int col = nt * 16 + (lane & 15);
        for (int r = 0; r < 4; ++r) {
            const int row = mt * 16 + (lane >> 4) * 4 + r;
            float val = acc[r] + b[col];
            if (RELU) y[row * MLP_H + col] = to_bf16(val > 0.f ? val : 0.f);
            else of[row * N + col] = val;
        }
    }
}

// forward pass of the tile whose features are in L.x0; raw outputs in L.o
SMS_DEV void mlp_forward_tile(MlpLds &L) {
    mlp_layer<MLP_K1, MLP_H, true>(&L.x0[0][0], MLP_K1, &L.w1[0][0], L.b1, &L.h[0][0][0], nullptr);
    __syncthreads();
    mlp_layer<MLP_H, MLP_H, true>(&L.h[0][0][0], MLP_H, &L.w2[0][0], L.b2, &L.h[1][0][0], nullptr);
    __syncthreads();
    mlp_layer<MLP_H, MLP_H, true>(&L.h[1][0][0], MLP_H, &L.w3[0][0], L.b3, &L.h[2][0][0], nullptr);
    __syncthreads();
    mlp_layer<MLP_H, MLP_H, true>(&L.h[2][0][0], MLP_H, &L.w4[0][0], L.b4, &L.h[3][0][0], nullptr);
    __syncthreads();
    mlp_layer<MLP_H, MLP_OUT, false>(&L.h[3][0][0], MLP_H, &L.w5[0][0], L.b5, nullptr, &L.o[0][0]);
    __syncthreads();
}

} // namespace smsdev
