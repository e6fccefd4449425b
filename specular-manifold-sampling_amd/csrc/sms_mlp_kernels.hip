// sms_mlp_kernels.hip -- forward, training and optimiser kernels of the learned seed proposal
// (network layout in dev_mlp.h).
//
//   k_mlp_infer   persistent over 64-sample tiles: weights loaded to LDS once per block, then
//                 features -> 4 ReLU layers -> 48 raw mixture outputs per sample
//   k_mlp_grad    one block per 64-sample tile of a batch: forward, weighted mixture negative
//                 log-likelihood of the target uv, backward through the five layers on the
//                 matrix cores; weight gradients are added to the FP32 gradient vector
//   k_mlp_adam    Adam on the FP32 master parameters, refreshes the BF16 copy, clears the
//                 gradient
#include "dev_mlp.h"

namespace smsdev {

struct MlpGradLds {
    MlpLds f;
    uint16_t dy[MLP_TILE][MLP_H];     // gradient w.r.t. the current layer's pre-activation (BF16)
    uint16_t dx[MLP_TILE][MLP_H];     // gradient w.r.t. its input (BF16)
    float loss[MLP_TILE];
};

SMS_DEV void mlp_tile_features(const float *__restrict__ x9, int64_t base, int64_t n, MlpLds &L) {
    for (int s = threadIdx.x; s < MLP_TILE; s += blockDim.x) {
        const int64_t i = base + s;
        float x[MLP_IN];
        for (int j = 0; j < MLP_IN; ++j) x[j] = i < n ? x9[i * MLP_IN + j] : 0.f;
        mlp_features(x, &L.x0[s][0]);
    }
}

__global__ void __launch_bounds__(256) k_mlp_infer(const float *__restrict__ p, const uint16_t *__restrict__ wb,
                                                   const float *__restrict__ x9, int64_t n, float *__restrict__ out) {
    __shared__ MlpLds L;
    mlp_load_weights(wb, p, L);
    const int64_t n_tiles = (n + MLP_TILE - 1) / MLP_TILE;
    for (int64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        __syncthreads();                                  // previous tile's outputs consumed
        mlp_tile_features(x9, t * MLP_TILE, n, L);
        __syncthreads();
        mlp_forward_tile(L);
        for (int e = threadIdx.x; e < MLP_TILE * MLP_OUT; e += blockDim.x) {
            const int s = e / MLP_OUT, c = e % MLP_OUT;
            const int64_t i = t * MLP_TILE + s;
            if (i < n) out[i * MLP_OUT + c] = L.o[s][c];
        }
    }
}

// gW[n][k] += sum_s dy[s][n] x[s][k] (output-major, row stride K) and db[n] += sum_s dy[s][n]
template <int K, int N>
SMS_DEV void mlp_weight_grad(const uint16_t *dy, int dys, const uint16_t *x, int xs, float *gw, float *gb) {
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    constexpr int MT = N / 16, KT = K / 16, TILES = MT * KT;
    for (int t = wv; t < TILES; t += 4) {
        const int mt = t / KT, kt = t % KT;
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma_tile([&](int m, int s) { return dy[s * dys + mt * 16 + m]; },
                        [&](int s, int k) { return x[s * xs + kt * 16 + k]; }, MLP_TILE, acc);
        const int k = kt * 16 + (lane & 15);
        for (int r = 0; r < 4; ++r) {
            const int nn = mt * 16 + (lane >> 4) * 4 + r;
            atomicAdd(&gw[nn * K + k], acc[r]);
        }
    }
    for (int nn = threadIdx.x; nn < N; nn += blockDim.x) {
        float s = 0.f;
        for (int r = 0; r < MLP_TILE; ++r) s += from_bf16(dy[r * dys + nn]);
        atomicAdd(&gb[nn], s);
    }
}

// dx[s][k] = sum_n dy[s][n] w[n][k], masked by the previous layer's ReLU (h > 0)
template <int N>
SMS_DEV void mlp_input_grad(const uint16_t *dy, int dys, const uint16_t *w, const uint16_t *h_prev, uint16_t *dx) {
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    constexpr int TILES = 4 * (MLP_H / 16);
    for (int t = wv; t < TILES; t += 4) {
        const int mt = t / 4, kt = t % 4;
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma_tile([&](int m, int nn) { return dy[(mt * 16 + m) * dys + nn]; },
                        [&](int nn, int k) { return w[nn * MLP_H + kt * 16 + k]; }, N, acc);
        const int k = kt * 16 + (lane & 15);
        for (int r = 0; r < 4; ++r) {
            const int s = mt * 16 + (lane >> 4) * 4 + r;
            const bool active = from_bf16(h_prev[s * MLP_H + k]) > 0.f;
            dx[s * MLP_H + k] = active ? to_bf16(acc[r]) : (uint16_t) 0;
        }
    }
}

// sum of the batch's sample weights (w null: the sample count), one 64-sample tile per block
__global__ void __launch_bounds__(64) k_mlp_weight_sum(const float *__restrict__ w, int64_t n, float *__restrict__ wsum) {
    const int64_t i = (int64_t) blockIdx.x * MLP_TILE + threadIdx.x;
    float v = i < n ? (w ? w[i] : 1.f) : 0.f;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if (threadIdx.x == 0) atomicAdd(wsum, v);
}

// Weighted negative log-likelihood: sample j's head gradient is scaled by w_j / sum(w);
// loss_sum accumulates sum_j w_j NLL_j.
__global__ void __launch_bounds__(256) k_mlp_grad(const float *__restrict__ p, const uint16_t *__restrict__ wb,
                                                  const float *__restrict__ x9, const float *__restrict__ uv,
                                                  const float *__restrict__ w, int64_t n, const float *__restrict__ wsum,
                                                  float *__restrict__ grad, float *__restrict__ loss_sum) {
    __shared__ MlpGradLds G;
    MlpLds &L = G.f;
    mlp_load_weights(wb, p, L);
    const int64_t base = (int64_t) blockIdx.x * MLP_TILE;
    mlp_tile_features(x9, base, n, L);
    __syncthreads();
    mlp_forward_tile(L);
    // mixture head: loss and d loss / d raw outputs (BF16 rows of dy, stride MLP_H)
    const float inv_wsum = 1.f / *wsum;
    for (int s = threadIdx.x; s < MLP_TILE; s += blockDim.x) {
        const int64_t i = base + s;
        float d_o[MLP_OUT];
        float l = 0.f;
        if (i < n) {
            const float wi = w ? w[i] : 1.f;
            l = wi * gmm_nll_grad(&L.o[s][0], uv[2 * i], uv[2 * i + 1], wi * inv_wsum, d_o);
        } else {
            for (int c = 0; c < MLP_OUT; ++c) d_o[c] = 0.f;
        }
        for (int c = 0; c < MLP_OUT; ++c) G.dy[s][c] = to_bf16(d_o[c]);
        G.loss[s] = l;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float s = 0.f;
        for (int r = 0; r < MLP_TILE; ++r) s += G.loss[r];
        atomicAdd(loss_sum, s);
    }
    // layer 5 (linear): dy = d o
    mlp_weight_grad<MLP_H, MLP_OUT>(&G.dy[0][0], MLP_H, &L.h[3][0][0], MLP_H, grad + MLP_W5, grad + MLP_B5);
    mlp_input_grad<MLP_OUT>(&G.dy[0][0], MLP_H, &L.w5[0][0], &L.h[3][0][0], &G.dx[0][0]);
    __syncthreads();
    // layer 4: dy = dx (already masked by h4 > 0)
    mlp_weight_grad<MLP_H, MLP_H>(&G.dx[0][0], MLP_H, &L.h[2][0][0], MLP_H, grad + MLP_W4, grad + MLP_B4);
    mlp_input_grad<MLP_H>(&G.dx[0][0], MLP_H, &L.w4[0][0], &L.h[2][0][0], &G.dy[0][0]);
    __syncthreads();
    // layer 3
    mlp_weight_grad<MLP_H, MLP_H>(&G.dy[0][0], MLP_H, &L.h[1][0][0], MLP_H, grad + MLP_W3, grad + MLP_B3);
    mlp_input_grad<MLP_H>(&G.dy[0][0], MLP_H, &L.w3[0][0], &L.h[1][0][0], &G.dx[0][0]);
    __syncthreads();
    // layer 2
    mlp_weight_grad<MLP_H, MLP_H>(&G.dx[0][0], MLP_H, &L.h[0][0][0], MLP_H, grad + MLP_W2, grad + MLP_B2);
    mlp_input_grad<MLP_H>(&G.dx[0][0], MLP_H, &L.w2[0][0], &L.h[0][0][0], &G.dy[0][0]);
    __syncthreads();
    // layer 1 (inputs are the features)
    mlp_weight_grad<MLP_K1, MLP_H>(&G.dy[0][0], MLP_H, &L.x0[0][0], MLP_K1, grad + MLP_W1, grad + MLP_B1);
}

__global__ void k_mlp_adam(float *__restrict__ p, float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
                           uint16_t *__restrict__ wb, int n, float lr, float b1, float b2, float eps, float bc1,
                           float bc2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / bc1, vh = vi / bc2;
    const float pi = p[i] - lr * mh / (sqrtf(vh) + eps);
    p[i] = pi;
    wb[i] = to_bf16(pi);
    g[i] = 0.f;
}

} // namespace smsdev
