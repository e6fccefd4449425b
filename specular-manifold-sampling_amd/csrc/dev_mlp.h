// dev_mlp.h -- the learned seed proposal's network on gfx950 matrix cores.
//
// 9 inputs (shading point, shading normal, light sample; positions normalised to the scene
// box) -> 8-frequency Fourier features (9 raw + 9 x 8 x {sin, cos} = 153, zero-padded to 160)
// -> 4 ReLU layers of width 64 -> 48 outputs = 8 Gaussian components x (weight logit, mean u,
// mean v, log sigma_u, log sigma_v, correlation).  Weights and activations are BF16, every
// product accumulates in FP32 on v_mfma_f32_16x16x32_bf16; biases, the mixture head, the loss
// and the optimiser are FP32.
//
// Weight-stationary wavefronts: a block owns a tile of 64 samples, and each wavefront keeps,
// for a whole tile, the BF16 weights of output rows 16 b .. 16 b + 15 of every layer (column
// block b = wavefront index mod 4) in its VGPRs (52 registers; layer 5's 48 rows on blocks 0-2)
// and, for training, the weights of input columns 16 b .. 16 b + 15 of layers 2-5 (32
// registers, from a transposed BF16 copy); biases sit in LDS.  LDS holds otherwise only the
// tile's activations, sample-major in rows padded to conflict-free strides.  A layer computes
// Y^T = W X^T: the wavefront's 16 outputs against its 16-sample tiles, independent accumulator
// chains of K / 32 products.  Inference: 256-lane blocks (one wavefront per column block, all
// four sample tiles), 52 KB of LDS and 150 registers, three blocks per CU.  Training: 512-lane
// blocks (two wavefronts per column block splitting the sample tiles and the weight gradients'
// k tiles), 81 KB of LDS and 256 registers without scratch: two wavefronts per SIMD.
//
// Parameter vector (FP32 master copy, gradients, Adam moments share it): for each layer the
// weights stored output-major W[n][k] (K padded: 160 for layer 1), then the bias.  The BF16
// copy mirrors it; the transposed BF16 copy holds W^T[k][n] of layers 2-5 back to back.
#pragma once
#include "dev_math.h"

namespace smsdev {

constexpr int MLP_IN = 9, MLP_FREQ = 8, MLP_FEAT = 153, MLP_K1 = 160, MLP_H = 64, MLP_OUT = 48, MLP_COMP = 8;
constexpr int MLP_TILE = 64;
constexpr int MLP_GRAD_THREADS = 512;   // training blocks: two wavefronts per 16-column block
constexpr int MLP_W1 = 0, MLP_B1 = MLP_W1 + MLP_H * MLP_K1;
constexpr int MLP_W2 = MLP_B1 + MLP_H, MLP_B2 = MLP_W2 + MLP_H * MLP_H;
constexpr int MLP_W3 = MLP_B2 + MLP_H, MLP_B3 = MLP_W3 + MLP_H * MLP_H;
constexpr int MLP_W4 = MLP_B3 + MLP_H, MLP_B4 = MLP_W4 + MLP_H * MLP_H;
constexpr int MLP_W5 = MLP_B4 + MLP_H, MLP_B5 = MLP_W5 + MLP_OUT * MLP_H;
constexpr int MLP_PARAMS = MLP_B5 + MLP_OUT;
// transposed BF16 copy: W^T[k][n] of layers 2, 3, 4 (64 x 64) and 5 (64 x 48)
constexpr int MLP_T2 = 0, MLP_T3 = MLP_T2 + MLP_H * MLP_H, MLP_T4 = MLP_T3 + MLP_H * MLP_H;
constexpr int MLP_T5 = MLP_T4 + MLP_H * MLP_H, MLP_TPARAMS = MLP_T5 + MLP_H * MLP_OUT;
// GMM head limits
constexpr float MLP_LOGSIG_MIN = -7.f, MLP_LOGSIG_MAX = 0.f, MLP_RHO_MAX = 0.95f;

typedef short bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16v8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef short s16x4 __attribute__((ext_vector_type(4)));

typedef __bf16 bf16v2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
// FP32 -> BF16 bits, round to nearest even (NaN kept quiet): gfx950's v_cvt_pk_bf16_f32, one or
// two values per instruction
SMS_DEV uint16_t to_bf16(float f) { return __builtin_bit_cast(uint16_t, (__bf16) f); }
SMS_DEV uint32_t to_bf16x2(float lo, float hi) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{lo, hi}, bf16v2));
}
SMS_DEV float from_bf16(uint16_t h) { return __uint_as_float((uint32_t) h << 16); }

// Workgroup barrier for LDS traffic only: LDS writes before it are visible after it, while
// global stores in flight (the gradient rows) need not complete, unlike __syncthreads.
SMS_DEV void lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// LDS row strides (BF16 entries) of the 160-wide feature rows and the 64-wide activation rows:
// multiples of 8 (16-byte rows) whose dword strides (84, 36) put the 16 rows one operand load
// spans on disjoint banks.
constexpr int MLP_XS = 168, MLP_HS = 72;

// v_mfma_f32_16x16x32_bf16: lane l supplies A[l % 16][8 (l / 16) + j] and B[8 (l / 16) + j][l % 16]
// (j = 0..7; the hardware pairs A's and B's entries j of a lane group with the same k, so any
// k order that A and B share gives the same sum) and receives C[4 (l / 16) + r][l % 16] in acc[r].
SMS_DEV f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 acc) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16v8, a), __builtin_bit_cast(bf16v8, b), acc, 0, 0, 0);
}
// eight consecutive BF16 entries (16-byte aligned) / eight entries `stride` apart
SMS_DEV bf16x8 ld8(const uint16_t *p) { return *(const bf16x8 *) p; }
SMS_DEV bf16x8 ld8s(const uint16_t *p, int stride) {
    bf16x8 v;
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = (short) p[j * stride];
    return v;
}
constexpr short BF16_ONE = 0x3F80;

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

// tanh from the native exponential: sign(x) (1 - 2 / (e^{2|x|} + 1))
SMS_DEV float tanh_native(float x) {
    const float t = 1.f - __fdividef(2.f, __expf(2.f * fabsf(x)) + 1.f);
    return x < 0.f ? -t : t;
}

// Negative log-likelihood of uv under the mixture of raw outputs o[48] and its gradient
// w.r.t. the raw outputs (times `scale`), computed by a group of 8 adjacent lanes: lane `c`
// of the group owns component c and writes dc[0..5] = its 6 entries.  Every lane of the group
// returns the same loss (all 64 lanes of the wavefront must call).  Whitened form:
// z1 = (u - mu_u) / s_u, z2 = ((v - mu_v) / s_v - rho z1) / sqrt(1 - rho^2).
SMS_DEV float gmm_nll_grad8(const float *o, float u, float v, float scale, int c, float *dc) {
    const float *oc = o + 6 * c;
    float mx = oc[0];
    mx = fmax_(mx, __shfl_xor(mx, 1, 64));
    mx = fmax_(mx, __shfl_xor(mx, 2, 64));
    mx = fmax_(mx, __shfl_xor(mx, 4, 64));
    const float ew = __expf(oc[0] - mx);
    float wsum = ew;
    wsum += __shfl_xor(wsum, 1, 64);
    wsum += __shfl_xor(wsum, 2, 64);
    wsum += __shfl_xor(wsum, 4, 64);
    const float wc = __fdividef(ew, wsum);
    const float mu1 = __fdividef(1.f, 1.f + __expf(-oc[1]));
    const float mu2 = __fdividef(1.f, 1.f + __expf(-oc[2]));
    const float sx = __expf(clampf_(oc[3], MLP_LOGSIG_MIN, MLP_LOGSIG_MAX));
    const float sy = __expf(clampf_(oc[4], MLP_LOGSIG_MIN, MLP_LOGSIG_MAX));
    const float r = MLP_RHO_MAX * tanh_native(oc[5]);
    const float q = sqrtf(1.f - r * r);
    const float z1 = __fdividef(u - mu1, sx);
    const float v2 = __fdividef(v - mu2, sy);
    const float z2 = __fdividef(v2 - r * z1, q);
    const float logn = -0.5f * (z1 * z1 + z2 * z2) - 1.8378770664093453f - __logf(sx) - __logf(sy) - __logf(q);
    const float a = __logf(wc) + logn;
    float amax = a;
    amax = fmax_(amax, __shfl_xor(amax, 1, 64));
    amax = fmax_(amax, __shfl_xor(amax, 2, 64));
    amax = fmax_(amax, __shfl_xor(amax, 4, 64));
    const float ea = __expf(a - amax);
    float s = ea;
    s += __shfl_xor(s, 1, 64);
    s += __shfl_xor(s, 2, 64);
    s += __shfl_xor(s, 4, 64);
    const float loss = -(amax + __logf(s));
    const float gam = __fdividef(ea, s);                    // responsibility
    dc[0] = scale * (wc - gam);                             // softmax over the weight logits
    const float iq = __fdividef(1.f, q), isx = __fdividef(1.f, sx), isy = __fdividef(1.f, sy);
    const float dmu1 = z1 * isx - z2 * r * iq * isx;
    const float dmu2 = z2 * iq * isy;
    const float dls1 = z1 * z1 - z2 * r * z1 * iq - 1.f;
    const float dls2 = z2 * v2 * iq - 1.f;
    const float dz2_drho = -z1 * iq + z2 * r * iq * iq;
    const float drho = -z2 * dz2_drho + r * iq * iq;
    const float in3 = (oc[3] > MLP_LOGSIG_MIN && oc[3] < MLP_LOGSIG_MAX) ? 1.f : 0.f;
    const float in4 = (oc[4] > MLP_LOGSIG_MIN && oc[4] < MLP_LOGSIG_MAX) ? 1.f : 0.f;
    const float th = r * (1.f / MLP_RHO_MAX);
    dc[1] = scale * (-gam) * dmu1 * (mu1 * (1.f - mu1));
    dc[2] = scale * (-gam) * dmu2 * (mu2 * (1.f - mu2));
    dc[3] = scale * (-gam) * dls1 * in3;
    dc[4] = scale * (-gam) * dls2 * in4;
    dc[5] = scale * (-gam) * drho * (MLP_RHO_MAX * (1.f - th * th));
    return loss;
}

// Weights a wavefront keeps in VGPRs for the forward pass: rows n = 16 w + lane % 16 of each
// layer, entries k = 32 t + 8 (lane / 16) .. + 7 of step t.  Layer 5 (48 rows): wavefronts 0-2
// (wavefront 3 holds zeros).
struct MlpRows {
    bf16x8 w1[MLP_K1 / 32], w2[2], w3[2], w4[2], w5[2];
};
// The five layers' biases (FP32, LDS): layer l at 64 (l - 1)
constexpr int MLP_NBIAS = 4 * MLP_H + MLP_OUT;
SMS_DEV void mlp_load_biases(const float *__restrict__ p, float *bias) {
    for (int e = threadIdx.x; e < MLP_NBIAS; e += blockDim.x) {
        const int l = e / MLP_H, c = e % MLP_H;
        const int src[5] = {MLP_B1, MLP_B2, MLP_B3, MLP_B4, MLP_B5};
        bias[e] = p[src[l] + c];
    }
}
// Weights a wavefront keeps for the input gradients: columns k = 16 w + lane % 16 of layers
// 2-5, entries n = 32 t + 8 (lane / 16) .. + 7 (layer 5's n = 48..63 are zeros).
struct MlpCols {
    bf16x8 t2[2], t3[2], t4[2], t5[2];
};

// A wavefront's column block: wavefront w of a block owns block w % 4 (blocks of 512 lanes run
// two wavefronts per block, which split the samples or the k tiles between them).
SMS_DEV int mlp_wave_block() { return (threadIdx.x >> 6) & 3; }
SMS_DEV int mlp_wave_half() { return threadIdx.x >> 8; }

SMS_DEV void mlp_load_rows(const uint16_t *__restrict__ wb, MlpRows &R) {
    static_assert(MLP_W2 % 8 == 0 && MLP_W3 % 8 == 0 && MLP_W4 % 8 == 0 && MLP_W5 % 8 == 0, "weight alignment");
    const int w = mlp_wave_block(), lane = threadIdx.x & 63, g = lane >> 4, i = lane & 15;
    const int n = 16 * w + i;
#pragma unroll
    for (int t = 0; t < MLP_K1 / 32; ++t) R.w1[t] = ld8(wb + MLP_W1 + n * MLP_K1 + 32 * t + 8 * g);
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        R.w2[t] = ld8(wb + MLP_W2 + n * MLP_H + 32 * t + 8 * g);
        R.w3[t] = ld8(wb + MLP_W3 + n * MLP_H + 32 * t + 8 * g);
        R.w4[t] = ld8(wb + MLP_W4 + n * MLP_H + 32 * t + 8 * g);
        R.w5[t] = w < 3 ? ld8(wb + MLP_W5 + n * MLP_H + 32 * t + 8 * g) : bf16x8{0, 0, 0, 0, 0, 0, 0, 0};
    }
}

SMS_DEV void mlp_load_cols(const uint16_t *__restrict__ wt, MlpCols &C) {
    static_assert(MLP_T3 % 8 == 0 && MLP_T4 % 8 == 0 && MLP_T5 % 8 == 0, "weight alignment");
    const int w = mlp_wave_block(), lane = threadIdx.x & 63, g = lane >> 4, i = lane & 15;
    const int k = 16 * w + i;
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        C.t2[t] = ld8(wt + MLP_T2 + k * MLP_H + 32 * t + 8 * g);
        C.t3[t] = ld8(wt + MLP_T3 + k * MLP_H + 32 * t + 8 * g);
        C.t4[t] = ld8(wt + MLP_T4 + k * MLP_H + 32 * t + 8 * g);
        C.t5[t] = 32 * t + 8 * g < MLP_OUT ? ld8(wt + MLP_T5 + k * MLP_OUT + 32 * t + 8 * g) : bf16x8{0, 0, 0, 0, 0, 0, 0, 0};
    }
}

SMS_DEV uint2 pack_bf16x4(float a, float b, float c, float d) { return make_uint2(to_bf16x2(a, b), to_bf16x2(c, d)); }

// One layer for the wavefront's 16 outputs n = 16 w + ..: Y[s][n] = act(sum_k X[s][k] W[n][k]
// + b[n]) for NST of the tile's four 16-sample tiles, starting at tile NST * half (X sample-major,
// stride XS).  RELU: BF16 rows of y (stride MLP_HS, four outputs per 8-byte store); otherwise
// FP32 rows of `of` (stride MLP_OUT).
template <int KSTEPS, int XS, bool RELU, int NST>
SMS_DEV void mlp_layer(const uint16_t *x, const bf16x8 *w, const float *b, uint16_t *y, float *of, int half) {
    const int wv = mlp_wave_block(), lane = threadIdx.x & 63, g = lane >> 4, i = lane & 15;
    f32x4 acc[NST];
#pragma unroll
    for (int st = 0; st < NST; ++st) acc[st] = f32x4{0.f, 0.f, 0.f, 0.f};
    const uint16_t *xs = x + (16 * NST * half + i) * XS + 8 * g;
#pragma unroll
    for (int t = 0; t < KSTEPS; ++t)
#pragma unroll
        for (int st = 0; st < NST; ++st) acc[st] = mfma32(w[t], ld8(xs + 16 * st * XS + 32 * t), acc[st]);
    const int n = 16 * wv + 4 * g;
    const f32x4 bias = *(const f32x4 *) (b + n);
#pragma unroll
    for (int st = 0; st < NST; ++st) {
        const int s = 16 * (NST * half + st) + i;
        const f32x4 v = acc[st] + bias;
        if (RELU)
            *(uint2 *) (y + s * MLP_HS + n) = pack_bf16x4(v[0] > 0.f ? v[0] : 0.f, v[1] > 0.f ? v[1] : 0.f,
                                                         v[2] > 0.f ? v[2] : 0.f, v[3] > 0.f ? v[3] : 0.f);
        else
            *(f32x4 *) (of + s * MLP_OUT + n) = v;
    }
}

// Activations of a tile in LDS (sample-major BF16 rows): the features, the four hidden layers
// (inference keeps two), the raw outputs (FP32).
struct alignas(16) MlpInferLds {
    float bias[MLP_NBIAS];
    uint16_t x0[MLP_TILE][MLP_XS];
    uint16_t h[2][MLP_TILE][MLP_HS];
    float o[MLP_TILE][MLP_OUT];
};

// forward pass of the tile whose features are in x0 (h: NH hidden buffers used in turn; each
// wavefront computes NST of the four 16-sample tiles); raw outputs in o
template <int NH, int NST>
SMS_DEV void mlp_forward_tile(const MlpRows &R, const float *bias, const uint16_t *x0, uint16_t (*h)[MLP_TILE][MLP_HS],
                              float *o) {
    const int half = NST == 4 ? 0 : mlp_wave_half();
    mlp_layer<MLP_K1 / 32, MLP_XS, true, NST>(x0, R.w1, bias, &h[0][0][0], nullptr, half);
    lds_sync();
    mlp_layer<2, MLP_HS, true, NST>(&h[0][0][0], R.w2, bias + MLP_H, &h[1 % NH][0][0], nullptr, half);
    lds_sync();
    mlp_layer<2, MLP_HS, true, NST>(&h[1 % NH][0][0], R.w3, bias + 2 * MLP_H, &h[2 % NH][0][0], nullptr, half);
    lds_sync();
    mlp_layer<2, MLP_HS, true, NST>(&h[2 % NH][0][0], R.w4, bias + 3 * MLP_H, &h[3 % NH][0][0], nullptr, half);
    lds_sync();
    if (mlp_wave_block() < 3)
        mlp_layer<2, MLP_HS, false, NST>(&h[3 % NH][0][0], R.w5, bias + 4 * MLP_H, nullptr, o, half);
    lds_sync();
}

} // namespace smsdev
