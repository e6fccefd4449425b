// sms_mlp_kernels.hip -- forward, training and optimiser kernels of the learned seed proposal
// (network layout and the weight-stationary wavefronts in dev_mlp.h).
//
//   k_mlp_infer        persistent over 64-sample tiles: each wavefront loads its weight rows
//                      to VGPRs once, then per tile features -> 4 ReLU layers -> 48 raw mixture
//                      outputs, activations in LDS (52 KB per block)
//   k_mlp_grad         block b runs tiles b, b + grid, ... of a batch: forward, weighted
//                      mixture negative log-likelihood of the target uv, backward through the
//                      five layers on the matrix cores (input gradients from the wavefront's
//                      weight columns in VGPRs); the weighted gradient sum of its tiles goes
//                      to its own FP32 row of a partial-gradient matrix (no atomics)
//   k_mlp_reduce       fixed-order sum of the rows / sum of weights -> gradient and mean loss
//   k_mlp_reduce_adam  the same sum fused with the Adam step (training loop: two kernels per step)
//   k_mlp_adam         Adam from a stored gradient, refreshes the BF16 copies, clears the gradient
//
// The rows are summed in block order, so a training run is bit-reproducible.
#include "dev_mlp.h"

namespace smsdev {

// Inputs a lane needs for a tile: the 9 network inputs of sample (lane % 64), whose features
// it writes with the other lanes of that sample (frequencies lane / 64 + (blockDim / 64) i),
// and target uv and weight of sample lane / 8 of the mixture head.  Loaded early so that their latency overlaps other work.
struct TileInputs {
    float x[MLP_IN], u, v, w;
    bool live;
};
SMS_DEV void load_tile_inputs(const float *__restrict__ x9, const float *__restrict__ uv, const float *__restrict__ w,
                              int64_t base, int64_t n, TileInputs &in) {
    const int64_t i = base + (threadIdx.x & (MLP_TILE - 1));
    for (int j = 0; j < MLP_IN; ++j) in.x[j] = i < n ? x9[i * MLP_IN + j] : 0.f;
    const int64_t h = base + (threadIdx.x >> 3);
    in.live = (threadIdx.x >> 3) < MLP_TILE && h < n;
    in.u = in.live && uv ? uv[2 * h] : 0.5f;
    in.v = in.live && uv ? uv[2 * h + 1] : 0.5f;
    in.w = in.live ? (w ? w[h] : 1.f) : 0.f;
}

// Features of a tile: blockDim / 64 lanes per sample, each writing its share of the eight
// frequencies.
SMS_DEV void mlp_tile_features(const TileInputs &in, uint16_t (*x0)[MLP_XS]) {
    mlp_features(in.x, &x0[threadIdx.x & (MLP_TILE - 1)][0], threadIdx.x >> 6, blockDim.x >> 6);
}

__global__ void __launch_bounds__(256) k_mlp_infer(const float *__restrict__ p, const uint16_t *__restrict__ wb,
                                                   const float *__restrict__ x9, int64_t n, float *__restrict__ out) {
    __shared__ MlpInferLds L;
    const int64_t n_tiles = (n + MLP_TILE - 1) / MLP_TILE;
    TileInputs in;
    if ((int64_t) blockIdx.x < n_tiles) load_tile_inputs(x9, nullptr, nullptr, blockIdx.x * MLP_TILE, n, in);
    MlpRows R;
    mlp_load_rows(wb, R);
    mlp_load_biases(p, L.bias);
    for (int64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        mlp_tile_features(in, L.x0);                      // x0 was last read by the previous tile's layer 1
        if (t + gridDim.x < n_tiles) load_tile_inputs(x9, nullptr, nullptr, (t + gridDim.x) * MLP_TILE, n, in);
        lds_sync();
        mlp_forward_tile<2, 4>(R, L.bias, &L.x0[0][0], L.h, &L.o[0][0]);
        const int64_t rows = n - t * MLP_TILE < MLP_TILE ? n - t * MLP_TILE : MLP_TILE;
        float *dst = out + t * MLP_TILE * MLP_OUT;
        for (int e = threadIdx.x; e < rows * MLP_OUT; e += blockDim.x) dst[e] = (&L.o[0][0])[e];
        // the next tile's layer 5 writes o only after four more barriers
    }
}

// Activations of a training tile: the features, the four hidden layers (kept for the
// backward pass), the output-layer gradient dy and the input gradient dx of the layer being
// differentiated (BF16, sample-major; dy and dx swap roles layer by layer), the raw outputs
// (FP32, dead once the head has read them: they share dx's storage); the block's loss totals.
struct alignas(16) MlpGradLds {
    float bias[MLP_NBIAS];
    uint16_t x0[MLP_TILE][MLP_XS];
    uint16_t h[4][MLP_TILE][MLP_HS];
    uint16_t dy[MLP_TILE][MLP_HS];
    union {
        uint16_t dx[MLP_TILE][MLP_HS];
        float o[MLP_TILE][MLP_OUT];
    };
    float tot[2][MLP_GRAD_THREADS / 64];
};

// gW[n][k] (+)= sum_s dy[s][n] x[s][k] (output-major, row stride K) for the wavefront's 16
// outputs n = 16 w + .. and the k tiles j0 .. j0 + JN - 1 (16 k each), and with j0 = 0 the bias
// gb[n] (+)= sum_s dy[s][n] (a product with a row of ones): the product's k dimension runs
// over the tile's samples, so both operands are gathered down the sample-major rows (stride
// MLP_HS for dy, XS for x).  add: accumulate onto the row (each entry belongs to one lane for
// every tile), else store.
template <int K, int XS, int JN>
SMS_DEV void mlp_weight_grad(const uint16_t *dy, const uint16_t *x, float *gw, float *gb, int j0, bool add) {
    const int wv = mlp_wave_block(), lane = threadIdx.x & 63, g = lane >> 4, i = lane & 15;
    const bf16x8 ones = {BF16_ONE, BF16_ONE, BF16_ONE, BF16_ONE, BF16_ONE, BF16_ONE, BF16_ONE, BF16_ONE};
    f32x4 acc[JN], accb = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int j = 0; j < JN; ++j) acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
    for (int s0 = 0; s0 < MLP_TILE; s0 += 32) {
        const bf16x8 a = ld8s(dy + (s0 + 8 * g) * MLP_HS + 16 * wv + i, MLP_HS);
#pragma unroll
        for (int j = 0; j < JN; ++j) acc[j] = mfma32(a, ld8s(x + (s0 + 8 * g) * XS + 16 * (j0 + j) + i, XS), acc[j]);
        if (j0 == 0) accb = mfma32(a, ones, accb);
    }
    const int n = 16 * wv + 4 * g;
#pragma unroll
    for (int j = 0; j < JN; ++j) {
        const int k = 16 * (j0 + j) + i;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            float *d = &gw[(n + r) * K + k];
            *d = add ? *d + acc[j][r] : acc[j][r];
        }
    }
    if (j0 == 0 && i == 0)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            float *d = &gb[n + r];
            *d = add ? *d + accb[r] : accb[r];
        }
}

// dx[s][k] = (h_prev[s][k] > 0) sum_n dy[s][n] W[n][k] for the wavefront's 16 inputs
// k = 16 w + .. (its weight columns wt) and the two 16-sample tiles of its half; dy's columns
// beyond the layer's outputs are zero (as are the matching weights).
SMS_DEV void mlp_input_grad(const uint16_t *dy, const bf16x8 *wt, const uint16_t *h_prev, uint16_t *dx) {
    const int wv = mlp_wave_block(), lane = threadIdx.x & 63, g = lane >> 4, i = lane & 15;
    const int s_base = 32 * mlp_wave_half() + i;
    f32x4 acc[2];
#pragma unroll
    for (int st = 0; st < 2; ++st) acc[st] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int st = 0; st < 2; ++st)
            acc[st] = mfma32(wt[t], ld8(dy + (s_base + 16 * st) * MLP_HS + 32 * t + 8 * g), acc[st]);
    const int k = 16 * wv + 4 * g;
#pragma unroll
    for (int st = 0; st < 2; ++st) {
        const int e = (s_base + 16 * st) * MLP_HS + k;
        const s16x4 hp = *(const s16x4 *) (h_prev + e);
        const f32x4 v = acc[st];
        *(uint2 *) (dx + e) = pack_bf16x4(from_bf16(hp[0]) > 0.f ? v[0] : 0.f, from_bf16(hp[1]) > 0.f ? v[1] : 0.f,
                                          from_bf16(hp[2]) > 0.f ? v[2] : 0.f, from_bf16(hp[3]) > 0.f ? v[3] : 0.f);
    }
}

// Block b (512 lanes: two wavefronts per column block, splitting the samples of the forward
// pass and the input gradients, and the k tiles of the weight gradients): tiles b, b + grid,
// ... of the n samples.  Sample j's head gradient is scaled by its weight w_j (w null: 1);
// partial[b] receives sum_j w_j d NLL_j / d params over the block's tiles and aux[b] =
// (sum_j w_j, sum_j w_j NLL_j).  The weight rows are live in the forward pass, the weight
// columns in the backward pass only: each tile loads them (from L2), the rows before the
// features, the columns after the mixture head.
__global__ void __launch_bounds__(MLP_GRAD_THREADS) k_mlp_grad(const float *__restrict__ p, const uint16_t *__restrict__ wb,
                                                              const uint16_t *__restrict__ wt, const float *__restrict__ x9,
                                                              const float *__restrict__ uv, const float *__restrict__ w,
                                                              int64_t n, float *__restrict__ partial,
                                                              float *__restrict__ aux) {
    __shared__ MlpGradLds G;
    float *grad = partial + (int64_t) blockIdx.x * MLP_PARAMS;
    const int64_t n_tiles = (n + MLP_TILE - 1) / MLP_TILE;
    TileInputs in;
    if ((int64_t) blockIdx.x < n_tiles) load_tile_inputs(x9, uv, w, blockIdx.x * MLP_TILE, n, in);
    const int blk = mlp_wave_block(), half = mlp_wave_half();
    float wsum = 0.f, lsum = 0.f;                     // the lane's running totals over its samples
    for (int64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        const bool add = t != (int64_t) blockIdx.x;
        {
            MlpRows R;
            mlp_load_rows(wb, R);
            lds_sync();                               // previous tile's LDS rows consumed
            mlp_load_biases(p, G.bias);
            mlp_tile_features(in, G.x0);
            lds_sync();
            mlp_forward_tile<4, 2>(R, G.bias, &G.x0[0][0], G.h, &G.o[0][0]);
        }
        // mixture head, 8 lanes per sample (one component each): BF16 rows of dy (columns 48..63 zero)
        {
            const int s = threadIdx.x >> 3, c = threadIdx.x & 7;
            float dc[6];
            const float l = gmm_nll_grad8(&G.o[s][0], in.u, in.v, in.w, c, dc);
            for (int e = 0; e < 6; ++e) G.dy[s][6 * c + e] = in.live ? to_bf16(dc[e]) : (uint16_t) 0;
            G.dy[s][MLP_OUT + 2 * c] = 0;
            G.dy[s][MLP_OUT + 2 * c + 1] = 0;
            if (c == 0) {
                wsum += in.w;
                lsum += in.live ? in.w * l : 0.f;
            }
        }
        MlpCols C;
        mlp_load_cols(wt, C);
        lds_sync();                                   // o consumed, dy complete
        // layer 5 (linear): dy = d o
        if (blk < 3)
            mlp_weight_grad<MLP_H, MLP_HS, 2>(&G.dy[0][0], &G.h[3][0][0], grad + MLP_W5, grad + MLP_B5, 2 * half, add);
        mlp_input_grad(&G.dy[0][0], C.t5, &G.h[3][0][0], &G.dx[0][0]);
        lds_sync();
        // layer 4: dy = dx (already masked by h4 > 0)
        mlp_weight_grad<MLP_H, MLP_HS, 2>(&G.dx[0][0], &G.h[2][0][0], grad + MLP_W4, grad + MLP_B4, 2 * half, add);
        mlp_input_grad(&G.dx[0][0], C.t4, &G.h[2][0][0], &G.dy[0][0]);
        lds_sync();
        // layer 3
        mlp_weight_grad<MLP_H, MLP_HS, 2>(&G.dy[0][0], &G.h[1][0][0], grad + MLP_W3, grad + MLP_B3, 2 * half, add);
        mlp_input_grad(&G.dy[0][0], C.t3, &G.h[1][0][0], &G.dx[0][0]);
        lds_sync();
        // layer 2
        mlp_weight_grad<MLP_H, MLP_HS, 2>(&G.dx[0][0], &G.h[0][0][0], grad + MLP_W2, grad + MLP_B2, 2 * half, add);
        mlp_input_grad(&G.dx[0][0], C.t2, &G.h[0][0][0], &G.dy[0][0]);
        lds_sync();
        // layer 1 (inputs are the features)
        mlp_weight_grad<MLP_K1, MLP_XS, 5>(&G.dy[0][0], &G.x0[0][0], grad + MLP_W1, grad + MLP_B1, 5 * half, add);
        if (t + gridDim.x < n_tiles) load_tile_inputs(x9, uv, w, (t + gridDim.x) * MLP_TILE, n, in);
    }
    // block totals in a fixed order: butterflies within the wavefronts, then the wavefronts in turn
    for (int o = 32; o > 0; o >>= 1) {
        wsum += __shfl_xor(wsum, o, 64);
        lsum += __shfl_xor(lsum, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        G.tot[0][threadIdx.x >> 6] = wsum;
        G.tot[1][threadIdx.x >> 6] = lsum;
    }
    lds_sync();
    if (threadIdx.x == 0) {
        float a = 0.f, l = 0.f;
        for (int k = 0; k < MLP_GRAD_THREADS / 64; ++k) { a += G.tot[0][k]; l += G.tot[1][k]; }
        aux[2 * blockIdx.x] = a;
        aux[2 * blockIdx.x + 1] = l;
    }
}

// Position of parameter i in the transposed BF16 copy (weights of layers 2-5), or -1.
SMS_DEV int mlp_transposed_index(int i) {
    if (i >= MLP_W2 && i < MLP_B2) { const int e = i - MLP_W2; return MLP_T2 + (e % MLP_H) * MLP_H + e / MLP_H; }
    if (i >= MLP_W3 && i < MLP_B3) { const int e = i - MLP_W3; return MLP_T3 + (e % MLP_H) * MLP_H + e / MLP_H; }
    if (i >= MLP_W4 && i < MLP_B4) { const int e = i - MLP_W4; return MLP_T4 + (e % MLP_H) * MLP_H + e / MLP_H; }
    if (i >= MLP_W5 && i < MLP_B5) { const int e = i - MLP_W5; return MLP_T5 + (e % MLP_H) * MLP_OUT + e / MLP_H; }
    return -1;
}

// Reduction of the partial rows.  A block of 256 lanes owns 64 consecutive parameters: lane
// (r, c) = (lane / 16, lane % 16) sums rows r, r + 16, r + 32, ... of parameters 4 c .. 4 c + 3
// (16-byte loads, eight rows in flight), then lanes 0..15 add the 16 partial sums in r order.
// The row totals of aux are summed by a butterfly within each wavefront and the four
// wavefront sums in order.  Every order is fixed, so the result does not depend on scheduling.
constexpr int MLP_RED_PARAMS = 64, MLP_RED_GROUPS = 16;
static_assert(MLP_PARAMS % 4 == 0, "parameter rows in 16-byte units");

// sums of parameters i0 .. i0 + 3 (i0 = blockIdx.x * 64 + 4 (lane % 16)) over the nb rows,
// valid in lanes 0..15; tot = (sum of weights, sum of weighted NLL) in every lane
SMS_DEV float4 mlp_row_sums(const float *__restrict__ partial, const float *__restrict__ aux, int nb, float *tot) {
    __shared__ float4 part[MLP_RED_GROUPS][16];
    __shared__ float ptot[2][4];
    const int r = threadIdx.x >> 4, c = threadIdx.x & 15, t = threadIdx.x;
    const int i0 = blockIdx.x * MLP_RED_PARAMS + 4 * c;
    float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int b0 = r; b0 < nb; b0 += 8 * MLP_RED_GROUPS) {
        float4 v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int b = b0 + MLP_RED_GROUPS * j;
            v[j] = i0 < MLP_PARAMS && b < nb ? *(const float4 *) (partial + (int64_t) b * MLP_PARAMS + i0)
                                             : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            s.x += v[j].x; s.y += v[j].y; s.z += v[j].z; s.w += v[j].w;
        }
    }
    float a = t < nb ? aux[2 * t] : 0.f, l = t < nb ? aux[2 * t + 1] : 0.f;
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        l += __shfl_xor(l, o, 64);
    }
    part[r][c] = s;
    if ((t & 63) == 0) { ptot[0][t >> 6] = a; ptot[1][t >> 6] = l; }
    lds_sync();
    tot[0] = ((ptot[0][0] + ptot[0][1]) + ptot[0][2]) + ptot[0][3];
    tot[1] = ((ptot[1][0] + ptot[1][1]) + ptot[1][2]) + ptot[1][3];
    float4 g = part[0][c];
    for (int k = 1; k < MLP_RED_GROUPS; ++k) {
        const float4 q = part[k][c];
        g.x += q.x; g.y += q.y; g.z += q.z; g.w += q.w;
    }
    return g;
}

// g = (sum of the nb partial rows) / sum of weights; loss[0] = weighted mean NLL (zero
// weights: zero gradient and loss).  Grid: ceil(params / 64) blocks of 256.
__global__ void __launch_bounds__(256) k_mlp_reduce(const float *__restrict__ partial, const float *__restrict__ aux,
                                                    int nb, float *__restrict__ g, float *__restrict__ loss) {
    float tot[2];
    const float4 s = mlp_row_sums(partial, aux, nb, tot);
    const float inv = tot[0] > 0.f ? 1.f / tot[0] : 0.f;
    const int i0 = blockIdx.x * MLP_RED_PARAMS + 4 * threadIdx.x;
    if (threadIdx.x < 16 && i0 < MLP_PARAMS) *(float4 *) (g + i0) = make_float4(s.x * inv, s.y * inv, s.z * inv, s.w * inv);
    if (blockIdx.x == 0 && threadIdx.x == 0 && loss) loss[0] = tot[1] * inv;
}

// Adam update of a parameter from its gradient gi, moments m, v and value p; stores the new
// moments and value at index i and refreshes both BF16 copies
SMS_DEV void mlp_adam_one(int i, float gi, float mi, float vi, float pi, float *__restrict__ p, float *__restrict__ m,
                          float *__restrict__ v, uint16_t *__restrict__ wb, uint16_t *__restrict__ wt, float lr, float b1,
                          float b2, float eps, float bc1, float bc2) {
    mi = b1 * mi + (1.f - b1) * gi;
    vi = b2 * vi + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / bc1, vh = vi / bc2;
    pi = pi - lr * mh / (sqrtf(vh) + eps);
    p[i] = pi;
    const uint16_t bi = to_bf16(pi);
    wb[i] = bi;
    const int ti = mlp_transposed_index(i);
    if (ti >= 0) wt[ti] = bi;
}

// Adam step on the reduced gradient (one kernel after each k_mlp_grad of a training loop);
// loss (optional) receives the step's weighted mean NLL.  The parameters' moments and values
// are loaded with the partial rows, so the step waits for one round of loads.
__global__ void __launch_bounds__(256) k_mlp_reduce_adam(const float *__restrict__ partial, const float *__restrict__ aux,
                                                         int nb, float *__restrict__ p, float *__restrict__ m,
                                                         float *__restrict__ v, uint16_t *__restrict__ wb,
                                                         uint16_t *__restrict__ wt, float lr,
                                                         float b1, float b2, float eps, float bc1, float bc2,
                                                         float *__restrict__ loss) {
    const int i0 = blockIdx.x * MLP_RED_PARAMS + 4 * threadIdx.x;
    const bool own = threadIdx.x < 16 && i0 < MLP_PARAMS;
    float4 m4 = make_float4(0.f, 0.f, 0.f, 0.f), v4 = m4, p4 = m4;
    if (own) {
        m4 = *(const float4 *) (m + i0);
        v4 = *(const float4 *) (v + i0);
        p4 = *(const float4 *) (p + i0);
    }
    float tot[2];
    const float4 s = mlp_row_sums(partial, aux, nb, tot);
    const float inv = tot[0] > 0.f ? 1.f / tot[0] : 0.f;
    if (blockIdx.x == 0 && threadIdx.x == 0 && loss) loss[0] = tot[1] * inv;
    if (!own) return;
    mlp_adam_one(i0, s.x * inv, m4.x, v4.x, p4.x, p, m, v, wb, wt, lr, b1, b2, eps, bc1, bc2);
    mlp_adam_one(i0 + 1, s.y * inv, m4.y, v4.y, p4.y, p, m, v, wb, wt, lr, b1, b2, eps, bc1, bc2);
    mlp_adam_one(i0 + 2, s.z * inv, m4.z, v4.z, p4.z, p, m, v, wb, wt, lr, b1, b2, eps, bc1, bc2);
    mlp_adam_one(i0 + 3, s.w * inv, m4.w, v4.w, p4.w, p, m, v, wb, wt, lr, b1, b2, eps, bc1, bc2);
}

__global__ void k_mlp_adam(float *__restrict__ p, float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
                           uint16_t *__restrict__ wb, uint16_t *__restrict__ wt, int n, float lr, float b1, float b2,
                           float eps, float bc1, float bc2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    mlp_adam_one(i, g[i], m[i], v[i], p[i], p, m, v, wb, wt, lr, b1, b2, eps, bc1, bc2);
    g[i] = 0.f;
}

} // namespace smsdev
