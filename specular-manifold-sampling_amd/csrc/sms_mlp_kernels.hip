// sms_mlp_kernels.hip -- forward, training and optimiser kernels of the learned seed proposal
// (network layout in dev_mlp.h).
//
//   k_mlp_infer        persistent over 64-sample tiles: weights loaded to LDS once per block,
//                      then features -> 4 ReLU layers -> 48 raw mixture outputs per sample
//   k_mlp_grad         block b runs tiles b, b + grid, ... of a batch: forward, weighted
//                      mixture negative log-likelihood of the target uv, backward through the
//                      five layers on the matrix cores; the weighted gradient sum of its tiles
//                      goes to its own FP32 row of a partial-gradient matrix (no atomics)
//   k_mlp_reduce       fixed-order sum of the rows / sum of weights -> gradient and mean loss
//   k_mlp_reduce_adam  the same sum fused with the Adam step (training loop: two kernels per step)
//   k_mlp_adam         Adam from a stored gradient, refreshes the BF16 copy, clears the gradient
//
// The rows are summed in block order, so a training run is bit-reproducible.
#include "dev_mlp.h"

namespace smsdev {

struct MlpGradLds {
    MlpLds f;
    uint16_t dy[MLP_TILE][MLP_SH];    // gradient w.r.t. the current layer's pre-activation (BF16)
    uint16_t dx[MLP_TILE][MLP_SH];    // gradient w.r.t. its input (BF16)
    float wl[MLP_TILE][2];            // per sample: weight, weight * NLL
};

// Inputs a lane needs for a tile: the 9 network inputs of sample (lane % 64), whose features
// it writes with three other lanes (frequencies lane / 64 + 4 i), and target uv and weight of
// sample lane / 4 of the mixture head.  Loaded early so that their latency overlaps other work.
struct TileInputs {
    float x[MLP_IN], u, v, w;
    bool live;
};
SMS_DEV void load_tile_inputs(const float *__restrict__ x9, const float *__restrict__ uv, const float *__restrict__ w,
                              int64_t base, int64_t n, TileInputs &in) {
    const int64_t i = base + (threadIdx.x & (MLP_TILE - 1));
    for (int j = 0; j < MLP_IN; ++j) in.x[j] = i < n ? x9[i * MLP_IN + j] : 0.f;
    const int64_t h = base + (threadIdx.x >> 2);
    in.live = h < n;
    in.u = in.live && uv ? uv[2 * h] : 0.5f;
    in.v = in.live && uv ? uv[2 * h + 1] : 0.5f;
    in.w = in.live ? (w ? w[h] : 1.f) : 0.f;
}

// Features of a tile: 4 lanes per sample, each writing two of the eight frequencies.
SMS_DEV void mlp_tile_features(const TileInputs &in, MlpLds &L) {
    mlp_features(in.x, &L.x0[threadIdx.x & (MLP_TILE - 1)][0], threadIdx.x >> 6, 4);
}

__global__ void __launch_bounds__(256) k_mlp_infer(const float *__restrict__ p, const uint16_t *__restrict__ wb,
                                                   const float *__restrict__ x9, int64_t n, float *__restrict__ out) {
    __shared__ MlpLds L;
    const int64_t n_tiles = (n + MLP_TILE - 1) / MLP_TILE;
    TileInputs in;
    if ((int64_t) blockIdx.x < n_tiles) load_tile_inputs(x9, nullptr, nullptr, blockIdx.x * MLP_TILE, n, in);
    mlp_load_weights(wb, p, L);
    for (int64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        lds_sync();                                       // previous tile's outputs consumed
        mlp_tile_features(in, L);
        if (t + gridDim.x < n_tiles) load_tile_inputs(x9, nullptr, nullptr, (t + gridDim.x) * MLP_TILE, n, in);
        lds_sync();
        mlp_forward_tile(L);
        for (int e = threadIdx.x; e < MLP_TILE * MLP_OUT; e += blockDim.x) {
            const int s = e / MLP_OUT, c = e % MLP_OUT;
            const int64_t i = t * MLP_TILE + s;
            if (i < n) out[i * MLP_OUT + c] = L.o[s][c];
        }
    }
}

// gW[n][k] (+)= sum_s dy[s][n] x[s][k] (output-major, row stride K) and gb[n] (+)= sum_s dy[s][n]
// on the matrix cores (the bias as a product with a column of ones); dy rows stride MLP_SH, x
// rows stride XS.  Wavefront w owns outputs n = 16 w .. 16 w + 15 (N / 16 wavefronts busy).
// add: accumulate onto the row (each entry belongs to one lane for every tile), else store.
template <int K, int N, int XS>
SMS_DEV void mlp_weight_grad(const uint16_t *dy, const uint16_t *x, float *gw, float *gb, bool add) {
    const int mt = threadIdx.x >> 6, lane = threadIdx.x & 63, r16 = lane & 15, kq = (lane >> 4) * 4;
    constexpr int MT = N / 16, KT = K / 16;
    if (mt >= MT) return;
    const bf16x4 ones = {0x3F80, 0x3F80, 0x3F80, 0x3F80};
    f32x4 acc[KT], accb = {0.f, 0.f, 0.f, 0.f};
    for (int j = 0; j < KT; ++j) acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s0 = 0; s0 < MLP_TILE; s0 += 16) {
        const bf16x4 a = lds4s(dy + (s0 + kq) * MLP_SH + mt * 16 + r16, MLP_SH);
#pragma unroll
        for (int j = 0; j < KT; ++j) acc[j] = mfma16(a, lds4s(x + (s0 + kq) * XS + j * 16 + r16, XS), acc[j]);
        accb = mfma16(a, ones, accb);
    }
#pragma unroll
    for (int j = 0; j < KT; ++j) {
        const int k = j * 16 + r16;
        for (int r = 0; r < 4; ++r) {
            float *d = &gw[(mt * 16 + (lane >> 4) * 4 + r) * K + k];
            *d = add ? *d + acc[j][r] : acc[j][r];
        }
    }
    if (r16 == 0)
        for (int r = 0; r < 4; ++r) {
            float *d = &gb[mt * 16 + (lane >> 4) * 4 + r];
            *d = add ? *d + accb[r] : accb[r];
        }
}

// dx[s][k] = sum_n dy[s][n] w[n][k] (N of the layer's outputs), masked by the previous layer's
// ReLU (h > 0); rows stride MLP_SH.  Wavefront w owns samples 16 w .. 16 w + 15.
template <int N>
SMS_DEV void mlp_input_grad(const uint16_t *dy, const uint16_t *w, const uint16_t *h_prev, uint16_t *dx) {
    const int mt = threadIdx.x >> 6, lane = threadIdx.x & 63, r16 = lane & 15, kq = (lane >> 4) * 4;
    constexpr int KT = MLP_H / 16;
    f32x4 acc[KT];
    for (int j = 0; j < KT; ++j) acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int n0 = 0; n0 < N; n0 += 16) {
        const bf16x4 a = lds4(dy + (mt * 16 + r16) * MLP_SH + n0 + kq);
#pragma unroll
        for (int j = 0; j < KT; ++j) acc[j] = mfma16(a, lds4s(w + (n0 + kq) * MLP_SH + j * 16 + r16, MLP_SH), acc[j]);
    }
#pragma unroll
    for (int j = 0; j < KT; ++j) {
        const int k = j * 16 + r16;
        for (int r = 0; r < 4; ++r) {
            const int e = (mt * 16 + (lane >> 4) * 4 + r) * MLP_SH + k;
            dx[e] = from_bf16(h_prev[e]) > 0.f ? to_bf16(acc[j][r]) : (uint16_t) 0;
        }
    }
}

// Block b: tiles b, b + gridDim.x, ... of the n samples.  Sample j's head gradient is scaled
// by its weight w_j (w null: 1); partial[b] receives sum_j w_j d NLL_j / d params over the
// block's tiles and aux[b] = (sum_j w_j, sum_j w_j NLL_j).
__global__ void __launch_bounds__(256) k_mlp_grad(const float *__restrict__ p, const uint16_t *__restrict__ wb,
                                                  const float *__restrict__ x9, const float *__restrict__ uv,
                                                  const float *__restrict__ w, int64_t n, float *__restrict__ partial,
                                                  float *__restrict__ aux) {
    __shared__ MlpGradLds G;
    MlpLds &L = G.f;
    float *grad = partial + (int64_t) blockIdx.x * MLP_PARAMS;
    const int64_t n_tiles = (n + MLP_TILE - 1) / MLP_TILE;
    TileInputs in;
    if ((int64_t) blockIdx.x < n_tiles) load_tile_inputs(x9, uv, w, blockIdx.x * MLP_TILE, n, in);
    mlp_load_weights(wb, p, L);
    float wsum = 0.f, lsum = 0.f;                     // lane 0's running block totals
    for (int64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        const bool add = t != (int64_t) blockIdx.x;
        lds_sync();                                   // previous tile's LDS rows consumed
        mlp_tile_features(in, L);
        lds_sync();
        mlp_forward_tile(L);
        // mixture head, 4 lanes per sample (BF16 rows of dy)
        {
            const int s = threadIdx.x >> 2, c4 = threadIdx.x & 3;
            float d_o[12];
            const float l = gmm_nll_grad4(&L.o[s][0], in.u, in.v, in.w, c4, d_o);
            for (int c = 0; c < 12; ++c) G.dy[s][12 * c4 + c] = in.live ? to_bf16(d_o[c]) : (uint16_t) 0;
            if (c4 == 0) {
                G.wl[s][0] = in.w;
                G.wl[s][1] = in.live ? in.w * l : 0.f;
            }
        }
        if (t + gridDim.x < n_tiles) load_tile_inputs(x9, uv, w, (t + gridDim.x) * MLP_TILE, n, in);
        lds_sync();
        if (threadIdx.x == 0)
            for (int r = 0; r < MLP_TILE; ++r) { wsum += G.wl[r][0]; lsum += G.wl[r][1]; }
        // layer 5 (linear): dy = d o
        mlp_weight_grad<MLP_H, MLP_OUT, MLP_SH>(&G.dy[0][0], &L.h[3][0][0], grad + MLP_W5, grad + MLP_B5, add);
        mlp_input_grad<MLP_OUT>(&G.dy[0][0], &L.w5[0][0], &L.h[3][0][0], &G.dx[0][0]);
        lds_sync();
        // layer 4: dy = dx (already masked by h4 > 0)
        mlp_weight_grad<MLP_H, MLP_H, MLP_SH>(&G.dx[0][0], &L.h[2][0][0], grad + MLP_W4, grad + MLP_B4, add);
        mlp_input_grad<MLP_H>(&G.dx[0][0], &L.w4[0][0], &L.h[2][0][0], &G.dy[0][0]);
        lds_sync();
        // layer 3
        mlp_weight_grad<MLP_H, MLP_H, MLP_SH>(&G.dy[0][0], &L.h[1][0][0], grad + MLP_W3, grad + MLP_B3, add);
        mlp_input_grad<MLP_H>(&G.dy[0][0], &L.w3[0][0], &L.h[1][0][0], &G.dx[0][0]);
        lds_sync();
        // layer 2
        mlp_weight_grad<MLP_H, MLP_H, MLP_SH>(&G.dx[0][0], &L.h[0][0][0], grad + MLP_W2, grad + MLP_B2, add);
        mlp_input_grad<MLP_H>(&G.dx[0][0], &L.w2[0][0], &L.h[0][0][0], &G.dy[0][0]);
        lds_sync();
        // layer 1 (inputs are the features)
        mlp_weight_grad<MLP_K1, MLP_H, MLP_S1>(&G.dy[0][0], &L.x0[0][0], grad + MLP_W1, grad + MLP_B1, add);
    }
    if (threadIdx.x == 0) {
        aux[2 * blockIdx.x] = wsum;
        aux[2 * blockIdx.x + 1] = lsum;
    }
}

// Reduction of the partial rows.  A block of 256 lanes owns 64 consecutive parameters: lane
// (r, c) = (lane / 64, lane % 64) sums rows r, r + 4, r + 8, ... of parameter c (independent
// loads, 8 in flight), then the four partial sums are added in r order.  The row totals of aux
// are summed by a butterfly within each wavefront and the four wavefront sums in order.  Every
// order is fixed, so the result does not depend on scheduling.
constexpr int MLP_RED_PARAMS = 64;

SMS_DEV void mlp_row_totals(const float *__restrict__ aux, int nb, float *tot) {
    __shared__ float part[2][4];
    const int t = threadIdx.x;
    float a = t < nb ? aux[2 * t] : 0.f, l = t < nb ? aux[2 * t + 1] : 0.f;
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        l += __shfl_xor(l, o, 64);
    }
    if ((t & 63) == 0) { part[0][t >> 6] = a; part[1][t >> 6] = l; }
    lds_sync();
    tot[0] = ((part[0][0] + part[0][1]) + part[0][2]) + part[0][3];
    tot[1] = ((part[1][0] + part[1][1]) + part[1][2]) + part[1][3];
}

// sum over the nb rows of parameter blockIdx.x * 64 + lane % 64 (every lane must call; result
// valid in lanes 0..63)
SMS_DEV float mlp_row_sum(const float *__restrict__ partial, int nb, int i) {
    __shared__ float part[4][MLP_RED_PARAMS];
    const int r = threadIdx.x >> 6, c = threadIdx.x & 63;
    float s = 0.f;
    if (i < MLP_PARAMS) {
        int b = r;
        for (; b + 28 < nb; b += 32) {
            float v[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = partial[(int64_t) (b + 4 * j) * MLP_PARAMS + i];
#pragma unroll
            for (int j = 0; j < 8; ++j) s += v[j];
        }
        for (; b < nb; b += 4) s += partial[(int64_t) b * MLP_PARAMS + i];
    }
    part[r][c] = s;
    lds_sync();
    return ((part[0][c] + part[1][c]) + part[2][c]) + part[3][c];
}

// g = (sum of the nb partial rows) / sum of weights; loss[0] = weighted mean NLL (zero
// weights: zero gradient and loss).  Grid: ceil(params / 64) blocks of 256.
__global__ void __launch_bounds__(256) k_mlp_reduce(const float *__restrict__ partial, const float *__restrict__ aux,
                                                    int nb, float *__restrict__ g, float *__restrict__ loss) {
    float tot[2];
    mlp_row_totals(aux, nb, tot);
    const float inv = tot[0] > 0.f ? 1.f / tot[0] : 0.f;
    const int i = blockIdx.x * MLP_RED_PARAMS + (threadIdx.x & 63);
    const float s = mlp_row_sum(partial, nb, i);
    if (threadIdx.x < 64 && i < MLP_PARAMS) g[i] = s * inv;
    if (blockIdx.x == 0 && threadIdx.x == 0 && loss) loss[0] = tot[1] * inv;
}

// Adam step on the reduced gradient (one kernel after each k_mlp_grad of a training loop);
// loss (optional) receives the step's weighted mean NLL.
__global__ void __launch_bounds__(256) k_mlp_reduce_adam(const float *__restrict__ partial, const float *__restrict__ aux,
                                                         int nb, float *__restrict__ p, float *__restrict__ m,
                                                         float *__restrict__ v, uint16_t *__restrict__ wb, float lr,
                                                         float b1, float b2, float eps, float bc1, float bc2,
                                                         float *__restrict__ loss) {
    float tot[2];
    mlp_row_totals(aux, nb, tot);
    const float inv = tot[0] > 0.f ? 1.f / tot[0] : 0.f;
    const int i = blockIdx.x * MLP_RED_PARAMS + (threadIdx.x & 63);
    const float s = mlp_row_sum(partial, nb, i);
    if (blockIdx.x == 0 && threadIdx.x == 0 && loss) loss[0] = tot[1] * inv;
    if (threadIdx.x >= 64 || i >= MLP_PARAMS) return;
    const float gi = s * inv;
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / bc1, vh = vi / bc2;
    const float pi = p[i] - lr * mh / (sqrtf(vh) + eps);
    p[i] = pi;
    wb[i] = to_bf16(pi);
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
