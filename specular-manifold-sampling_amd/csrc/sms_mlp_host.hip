// sms_mlp_host.hip -- C ABI of the learned seed proposal (include/sms.h): network parameters,
// training steps, the proposal's UV grid, the per-estimate seed preparation and the capture
// of training samples.  No counterpart in the reference, whose seeds are uniform over the
// caster's area (src/librender/manifold_ss.cpp:228-235).
#include "host_ctx.h"

#include <hipcub/device/device_radix_sort.hpp>
#include "dev_mlp.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

using namespace smshost;

// Most blocks of a gradient pass: rows of the partial-gradient matrix (256 x 25,904 floats).
constexpr int MLP_GRAD_ROWS = 256;

struct sms_mlp {
    sms_ctx *ctx = nullptr;
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr, *loss = nullptr;   // device, FP32
    uint16_t *wb = nullptr;                                                          // device, BF16 copy of p
    uint16_t *wt = nullptr;               // device, BF16 W^T of layers 2-5 (the input gradients' operands)
    float *partial = nullptr, *aux = nullptr;   // device: per-block gradient rows, (sum w, sum w nll) per row
    float *loss_hist = nullptr;                 // device: per-step losses of sms_mlp_train
    int64_t loss_cap = 0;
    int64_t step = 0;                                                                // Adam steps taken
};

namespace {

static_assert(MLP_PARAMS == SMS_MLP_PARAMS, "parameter count of the ABI and the kernels differ");
static_assert(MLP_IN == SMS_MLP_INPUTS && MLP_OUT == SMS_MLP_OUTPUTS, "network shape of the ABI and the kernels differ");
static_assert(PROP_RAW == MLP_OUT, "mixture layout");

uint16_t host_bf16(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t) ((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t) (u >> 16);
}

// splitmix64 stream for the initial weights
struct InitRng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double uniform() { return ((double) (next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    double normal() {
        const double u1 = uniform(), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

// He-normal weights (std sqrt(2 / fan_in)); the output layer scaled by 0.1; zero biases
std::vector<float> init_params(uint64_t seed) {
    std::vector<float> p(MLP_PARAMS, 0.f);
    InitRng r{seed ^ 0x5eed5eed12345678ULL};
    auto layer = [&](int w_off, int n, int k_used, int k_stride, double gain) {
        const double sd = gain * std::sqrt(2.0 / (double) k_used);
        for (int o = 0; o < n; ++o)
            for (int k = 0; k < k_used; ++k) p[(size_t) w_off + (size_t) o * k_stride + k] = (float) (sd * r.normal());
    };
    layer(MLP_W1, MLP_H, MLP_FEAT, MLP_K1, 1.0);
    layer(MLP_W2, MLP_H, MLP_H, MLP_H, 1.0);
    layer(MLP_W3, MLP_H, MLP_H, MLP_H, 1.0);
    layer(MLP_W4, MLP_H, MLP_H, MLP_H, 1.0);
    layer(MLP_W5, MLP_OUT, MLP_H, MLP_H, 0.1);
    return p;
}

sms_status upload_params(sms_mlp *mlp, const float *h) {
    sms_ctx *ctx = mlp->ctx;
    std::vector<uint16_t> b(MLP_PARAMS), t(MLP_TPARAMS);
    for (int i = 0; i < MLP_PARAMS; ++i) b[i] = host_bf16(h[i]);
    const int layers[4][3] = {{MLP_W2, MLP_T2, MLP_H}, {MLP_W3, MLP_T3, MLP_H}, {MLP_W4, MLP_T4, MLP_H}, {MLP_W5, MLP_T5, MLP_OUT}};
    for (const auto &l : layers)
        for (int n = 0; n < l[2]; ++n)
            for (int k = 0; k < MLP_H; ++k) t[(size_t) l[1] + (size_t) k * l[2] + n] = b[(size_t) l[0] + (size_t) n * MLP_H + k];
    HIP_TRY(ctx, hipMemcpyAsync(mlp->p, h, MLP_PARAMS * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(mlp->wb, b.data(), MLP_PARAMS * sizeof(uint16_t), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(mlp->wt, t.data(), MLP_TPARAMS * sizeof(uint16_t), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return SMS_OK;
}

void free_mlp(sms_mlp *mlp) {
    for (void *q : {(void *) mlp->p, (void *) mlp->g, (void *) mlp->m, (void *) mlp->v, (void *) mlp->loss,
                    (void *) mlp->wb, (void *) mlp->wt, (void *) mlp->partial, (void *) mlp->aux, (void *) mlp->loss_hist})
        if (q) (void) hipFree(q);
    delete mlp;
}

sms_status launch_infer(sms_mlp *mlp, const float *x, int64_t n, float *out) {
    sms_ctx *ctx = mlp->ctx;
    if (n == 0) return SMS_OK;
    const int64_t tiles = (n + MLP_TILE - 1) / MLP_TILE;
    const unsigned grid = (unsigned) std::min<int64_t>(tiles, (int64_t) ctx->n_cu * 3);   // three resident blocks per CU
    hipLaunchKernelGGL(k_mlp_infer, dim3(grid), dim3(256), 0, ctx->stream, mlp->p, mlp->wb, x, n, out);
    HIP_TRY(ctx, hipGetLastError());
    return SMS_OK;
}

constexpr float ADAM_B1 = 0.9f, ADAM_B2 = 0.999f, ADAM_EPS = 1e-8f;

// next Adam step's bias corrections 1 - beta^t (beta and the power in double, rounded once)
void adam_corrections(sms_mlp *mlp, float *bc1, float *bc2) {
    mlp->step++;
    *bc1 = (float) (1.0 - std::pow(0.9, (double) mlp->step));
    *bc2 = (float) (1.0 - std::pow(0.999, (double) mlp->step));
}

// blocks (partial rows) of a gradient pass over n samples
int grad_rows(int64_t n) {
    return (int) std::min<int64_t>((n + MLP_TILE - 1) / MLP_TILE, MLP_GRAD_ROWS);
}

sms_status ensure_prop(sms_ctx *ctx, size_t n) {
    if (n <= ctx->prop_cap) return SMS_OK;
    ctx->gen++;
    const size_t cap = std::max<size_t>(n, 2 * ctx->prop_cap);
    for (float **b : {&ctx->prop_x, &ctx->prop_raw, &ctx->prop_g, &ctx->prop_anchor})
        if (*b) { (void) hipFree(*b); *b = nullptr; }
    ctx->prop_cap = 0;
    HIP_TRY(ctx, hipMalloc(&ctx->prop_anchor, cap * 2 * sizeof(float)));
    HIP_TRY(ctx, hipMalloc(&ctx->prop_x, cap * MLP_IN * sizeof(float)));
    HIP_TRY(ctx, hipMalloc(&ctx->prop_raw, cap * MLP_OUT * sizeof(float)));
    HIP_TRY(ctx, hipMalloc(&ctx->prop_g, cap * MLP_OUT * sizeof(float)));
    ctx->prop_cap = cap;
    return SMS_OK;
}

// Uniform grid over the UV square of mesh `shape`: G x G cells (G the power of two with
// 2 G^2 >= faces, at most 1024); cell (x, y) lists, in ascending face order, every face whose
// UV bounding box [lo, hi] has floor(lo * G) <= x <= floor(hi * G) (clamped to the grid) and the
// same along y.  G is a power of two, so uv * G is exact and a point of a face's UV triangle
// always falls in a cell that lists the face.
sms_status build_uv_grid(sms_ctx *ctx, int shape) {
    const DShape &sh = ctx->h_shapes[shape];
    int G = 1;
    while (2 * (int64_t) G * G < sh.nf && G < 1024) G *= 2;
    auto cell_of = [&](float t) {
        int c = (int) std::floor(t * (float) G);
        return std::min(std::max(c, 0), G - 1);
    };
    std::vector<int> count((size_t) G * G + 1, 0);
    std::vector<int4> span(sh.nf);
    for (int f = 0; f < sh.nf; ++f) {
        const uint4 fc = ctx->h_faces[sh.fbase + f];
        const float2 a = ctx->h_vuv[fc.x], b = ctx->h_vuv[fc.y], c = ctx->h_vuv[fc.z];
        const float lx = std::min(std::min(a.x, b.x), c.x), hx = std::max(std::max(a.x, b.x), c.x);
        const float ly = std::min(std::min(a.y, b.y), c.y), hy = std::max(std::max(a.y, b.y), c.y);
        span[f] = make_int4(cell_of(lx), cell_of(hx), cell_of(ly), cell_of(hy));
        for (int y = span[f].z; y <= span[f].w; ++y)
            for (int x = span[f].x; x <= span[f].y; ++x) count[(size_t) y * G + x + 1]++;
    }
    for (size_t i = 1; i < count.size(); ++i) count[i] += count[i - 1];
    std::vector<int> fill(count.begin(), count.end() - 1), faces((size_t) count.back());
    for (int f = 0; f < sh.nf; ++f)
        for (int y = span[f].z; y <= span[f].w; ++y)
            for (int x = span[f].x; x <= span[f].y; ++x) faces[(size_t) fill[(size_t) y * G + x]++] = sh.fbase + f;
    if (ctx->grid_start) { (void) hipFree(ctx->grid_start); ctx->grid_start = nullptr; }
    if (ctx->grid_faces) { (void) hipFree(ctx->grid_faces); ctx->grid_faces = nullptr; }
    ctx->grid_shape = -1;
    HIP_TRY(ctx, hipMalloc(&ctx->grid_start, count.size() * sizeof(int)));
    HIP_TRY(ctx, hipMalloc(&ctx->grid_faces, std::max<size_t>(faces.size(), 1) * sizeof(int)));
    HIP_TRY(ctx, hipMemcpy(ctx->grid_start, count.data(), count.size() * sizeof(int), hipMemcpyHostToDevice));
    if (!faces.empty())
        HIP_TRY(ctx, hipMemcpy(ctx->grid_faces, faces.data(), faces.size() * sizeof(int), hipMemcpyHostToDevice));
    ctx->grid_g = G;
    ctx->grid_shape = shape;
    return SMS_OK;
}

// anchors of n single-scatter queries with the seed of device slot `seed` on the proposal's mesh
sms_status launch_anchors(sms_ctx *ctx, const sms_ss_query *q, int64_t n, const uint64_t *seed, float *anchor) {
    if (n == 0) return SMS_OK;
    hipLaunchKernelGGL(k_prop_anchor, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, ctx->stream, ctx->scene, q, n,
                       seed, ctx->prop_shape, anchor);
    HIP_TRY(ctx, hipGetLastError());
    return SMS_OK;
}

// network inputs of n queries with the seed of device slot `seed` (ms: multi-scatter emitters)
sms_status launch_inputs(sms_ctx *ctx, const sms_ss_query *q, int64_t n, const uint64_t *seed, int ms, float *x) {
    if (n == 0) return SMS_OK;
    hipLaunchKernelGGL(k_prop_inputs, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, ctx->stream, ctx->scene, q, n,
                       seed, ms, x);
    HIP_TRY(ctx, hipGetLastError());
    return SMS_OK;
}

} // namespace

// anchored single-scatter mixtures in use: a radius, a single-scatter caster mesh as the
// proposal's shape, and a network or a training capture that reads the anchors
bool anchored(const sms_ctx *ctx) {
    if (!(ctx->prop_radius > 0.f) || ctx->prop_shape < 0 || !(ctx->prop_mlp || ctx->cap_sub)) return false;
    const DShape &sh = ctx->h_shapes[ctx->prop_shape];
    return sh.type == SMS_SHAPE_MESH && (sh.flags & SMS_FLAG_CAUSTIC_CASTER_SS);
}

// shared with sms_host.hip
Proposal proposal_of(sms_ctx *ctx) {
    Proposal P;
    memset(&P, 0, sizeof(P));
    P.gmm = ctx->prop_mlp && ctx->prop_shape >= 0 ? ctx->prop_g : nullptr;
    P.shape = ctx->prop_shape;
    P.defensive = ctx->prop_defensive;
    P.anchor = anchored(ctx) ? ctx->prop_anchor : nullptr;
    P.radius = P.anchor ? ctx->prop_radius : 0.f;
    P.grid.g = ctx->grid_g; P.grid.start = ctx->grid_start; P.grid.faces = ctx->grid_faces;
    P.cap = ctx->cap_sub ? ctx->cap : nullptr;
    P.cap_counts = ctx->cap_counts;
    P.cap_sub = ctx->cap_sub;
    P.cap_mode = ctx->cap_mode;
    return P;
}

// buffers of the proposal's per-query mixtures (and anchors) for estimates of n queries
sms_status proposal_alloc(sms_ctx *ctx, int64_t n) {
    if (!(ctx->prop_mlp || anchored(ctx)) || ctx->prop_shape < 0 || n == 0) return SMS_OK;
    return ensure_prop(ctx, (size_t) n);
}

// network inputs -> raw outputs -> mixtures of the n queries of an estimate (prop_g), with the
// estimator's seed slot (ms: the multi-scatter estimator's light samples)
sms_status launch_glint_inputs(sms_ctx *ctx, const GlintQuerySource &q, int64_t n, const uint64_t *seed, float *x) {
    if (n == 0) return SMS_OK;
    hipLaunchKernelGGL(k_glint_prop_inputs, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, ctx->stream, ctx->scene, q,
                       n, seed, x);
    HIP_TRY(ctx, hipGetLastError());
    return SMS_OK;
}

// the glint estimator's mixtures: inputs, network, mixture preparation, when the proposal's
// shape is glinty
sms_status proposal_enqueue_glint(sms_ctx *ctx, const GlintQuerySource &q, int64_t n) {
    if (!ctx->prop_mlp || ctx->prop_shape < 0 || n == 0) return SMS_OK;
    if (!(ctx->h_shapes[ctx->prop_shape].flags & SMS_FLAG_GLINTY)) return SMS_OK;
    if ((size_t) n > ctx->prop_cap) return fail(ctx, SMS_E_INVALID, "proposal buffers not prepared");
    sms_status st = launch_glint_inputs(ctx, q, n, ctx->seeds + SEED_EST, ctx->prop_x);
    if (!st) st = launch_infer(ctx->prop_mlp, ctx->prop_x, n, ctx->prop_raw);
    if (st) return st;
    hipLaunchKernelGGL(k_gmm_prepare, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, ctx->stream, ctx->prop_raw, n,
                       ctx->prop_g);
    HIP_TRY(ctx, hipGetLastError());
    return SMS_OK;
}

sms_status proposal_enqueue(sms_ctx *ctx, const sms_ss_query *q, int64_t n, int ms) {
    if (!ms && n > 0 && anchored(ctx)) {
        if ((size_t) n > ctx->prop_cap) return fail(ctx, SMS_E_INVALID, "proposal buffers not prepared");
        sms_status st = launch_anchors(ctx, q, n, ctx->seeds + SEED_EST, ctx->prop_anchor);
        if (st) return st;
    }
    if (!ctx->prop_mlp || ctx->prop_shape < 0 || n == 0) return SMS_OK;
    // only an estimator whose casters include the proposal's shape uses the mixtures
    if (!(ctx->h_shapes[ctx->prop_shape].flags & (ms ? SMS_FLAG_CAUSTIC_CASTER_MS : SMS_FLAG_CAUSTIC_CASTER_SS)))
        return SMS_OK;
    if ((size_t) n > ctx->prop_cap) return fail(ctx, SMS_E_INVALID, "proposal buffers not prepared");
    sms_status st = launch_inputs(ctx, q, n, ctx->seeds + SEED_EST, ms, ctx->prop_x);
    if (!st) st = launch_infer(ctx->prop_mlp, ctx->prop_x, n, ctx->prop_raw);
    if (st) return st;
    hipLaunchKernelGGL(k_gmm_prepare, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, ctx->stream, ctx->prop_raw, n,
                       ctx->prop_g);
    HIP_TRY(ctx, hipGetLastError());
    return SMS_OK;
}

void proposal_release(sms_ctx *ctx) {
    for (void *q : {(void *) ctx->grid_start, (void *) ctx->grid_faces, (void *) ctx->prop_x, (void *) ctx->prop_raw,
                    (void *) ctx->prop_g, (void *) ctx->prop_anchor, (void *) ctx->cap, (void *) ctx->cap_counts})
        if (q) (void) hipFree(q);
    if (ctx->prop_mlp) ctx->prop_mlp = nullptr;
}

extern "C" {

sms_status sms_mlp_create(sms_ctx *ctx, uint64_t seed, sms_mlp **out) {
    return guarded(ctx, [&]() -> sms_status {
        if (!ctx || !out) return fail(ctx, SMS_E_INVALID, "sms_mlp_create: null argument");
        *out = nullptr;
        sms_mlp *mlp = new (std::nothrow) sms_mlp();
        if (!mlp) return fail(ctx, SMS_E_OOM, "sms_mlp_create: host allocation");
        mlp->ctx = ctx;
        const size_t fb = MLP_PARAMS * sizeof(float);
        hipError_t e = hipMalloc(&mlp->p, fb);
        if (e == hipSuccess) e = hipMalloc(&mlp->g, fb);
        if (e == hipSuccess) e = hipMalloc(&mlp->m, fb);
        if (e == hipSuccess) e = hipMalloc(&mlp->v, fb);
        if (e == hipSuccess) e = hipMalloc(&mlp->loss, 2 * sizeof(float));
        if (e == hipSuccess) e = hipMalloc(&mlp->partial, (size_t) MLP_GRAD_ROWS * fb);
        if (e == hipSuccess) e = hipMalloc(&mlp->aux, 2 * MLP_GRAD_ROWS * sizeof(float));
        if (e == hipSuccess) e = hipMalloc(&mlp->wb, MLP_PARAMS * sizeof(uint16_t));
        if (e == hipSuccess) e = hipMalloc(&mlp->wt, MLP_TPARAMS * sizeof(uint16_t));
        if (e == hipSuccess) e = hipMemsetAsync(mlp->g, 0, fb, ctx->stream);
        if (e == hipSuccess) e = hipMemsetAsync(mlp->m, 0, fb, ctx->stream);
        if (e == hipSuccess) e = hipMemsetAsync(mlp->v, 0, fb, ctx->stream);
        if (e != hipSuccess) {
            free_mlp(mlp);
            return fail(ctx, e == hipErrorOutOfMemory ? SMS_E_OOM : SMS_E_HIP, std::string("sms_mlp_create: ") + hipGetErrorString(e));
        }
        std::vector<float> h = init_params(seed);
        sms_status st = upload_params(mlp, h.data());
        if (st) { free_mlp(mlp); return st; }
        *out = mlp;
        return SMS_OK;
    });
}

sms_status sms_mlp_destroy(sms_mlp *mlp) {
    return guarded(mlp ? mlp->ctx : nullptr, [&]() -> sms_status {
        if (!mlp) return SMS_E_INVALID;
        sms_ctx *ctx = mlp->ctx;
        (void) hipStreamSynchronize(ctx->stream);
        if (ctx->prop_mlp == mlp) {
            ctx->prop_mlp = nullptr;
            ctx->gen++;
        }
        free_mlp(mlp);
        return SMS_OK;
    });
}

sms_status sms_mlp_get_params(sms_mlp *mlp, float *params) {
    return guarded(mlp ? mlp->ctx : nullptr, [&]() -> sms_status {
        if (!mlp || !params) return SMS_E_INVALID;
        sms_ctx *ctx = mlp->ctx;
        HIP_TRY(ctx, hipMemcpyAsync(params, mlp->p, MLP_PARAMS * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        return SMS_OK;
    });
}

sms_status sms_mlp_set_params(sms_mlp *mlp, const float *params) {
    return guarded(mlp ? mlp->ctx : nullptr, [&]() -> sms_status {
        if (!mlp || !params) return SMS_E_INVALID;
        return upload_params(mlp, params);
    });
}

sms_status sms_mlp_infer(sms_mlp *mlp, const float *x, int64_t n, float *out) {
    return guarded(mlp ? mlp->ctx : nullptr, [&]() -> sms_status {
        if (!mlp) return SMS_E_INVALID;
        if (n < 0 || (n > 0 && (!x || !out))) return fail(mlp->ctx, SMS_E_INVALID, "sms_mlp_infer: bad arguments");
        return launch_infer(mlp, x, n, out);
    });
}

sms_status sms_mlp_grad(sms_mlp *mlp, const float *x, const float *uv, const float *w, int64_t n, float *grad,
                        float *loss) {
    return guarded(mlp ? mlp->ctx : nullptr, [&]() -> sms_status {
        if (!mlp) return SMS_E_INVALID;
        sms_ctx *ctx = mlp->ctx;
        if (n <= 0 || !x || !uv) return fail(ctx, SMS_E_INVALID, "sms_mlp_grad: bad arguments");
        const int rows = grad_rows(n);
        hipLaunchKernelGGL(k_mlp_grad, dim3(rows), dim3(MLP_GRAD_THREADS), 0, ctx->stream, mlp->p, mlp->wb, mlp->wt, x, uv, w, n,
                           mlp->partial, mlp->aux);
        HIP_TRY(ctx, hipGetLastError());
        hipLaunchKernelGGL(k_mlp_reduce, dim3((MLP_PARAMS + 63) / 64), dim3(256), 0, ctx->stream,
                           (const float *) mlp->partial, (const float *) mlp->aux, rows, mlp->g, mlp->loss);
        HIP_TRY(ctx, hipGetLastError());
        if (grad)
            HIP_TRY(ctx, hipMemcpyAsync(grad, mlp->g, MLP_PARAMS * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
        if (loss) HIP_TRY(ctx, hipMemcpyAsync(loss, mlp->loss, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
        if (grad || loss) HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        return SMS_OK;
    });
}

sms_status sms_mlp_adam(sms_mlp *mlp, float lr) {
    return guarded(mlp ? mlp->ctx : nullptr, [&]() -> sms_status {
        if (!mlp) return SMS_E_INVALID;
        sms_ctx *ctx = mlp->ctx;
        if (!(lr > 0.f)) return fail(ctx, SMS_E_INVALID, "sms_mlp_adam: learning rate must be positive");
        float bc1, bc2;
        adam_corrections(mlp, &bc1, &bc2);
        hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_PARAMS + 255) / 256), dim3(256), 0, ctx->stream, mlp->p, mlp->g, mlp->m,
                           mlp->v, mlp->wb, mlp->wt, MLP_PARAMS, lr, ADAM_B1, ADAM_B2, ADAM_EPS, bc1, bc2);
        HIP_TRY(ctx, hipGetLastError());
        return SMS_OK;
    });
}

sms_status sms_mlp_train(sms_mlp *mlp, const float *x, const float *uv, const float *w, int32_t batch,
                         const int64_t *offsets, int32_t n_steps, float lr, float *losses) {
    return guarded(mlp ? mlp->ctx : nullptr, [&]() -> sms_status {
        if (!mlp) return SMS_E_INVALID;
        sms_ctx *ctx = mlp->ctx;
        if (n_steps < 0 || (n_steps > 0 && (!x || !uv || !offsets || batch <= 0)))
            return fail(ctx, SMS_E_INVALID, "sms_mlp_train: bad arguments");
        if (!(lr > 0.f)) return fail(ctx, SMS_E_INVALID, "sms_mlp_train: learning rate must be positive");
        for (int32_t i = 0; i < n_steps; ++i)
            if (offsets[i] < 0) return fail(ctx, SMS_E_INVALID, "sms_mlp_train: negative batch offset");
        if (losses && n_steps > mlp->loss_cap) {
            if (mlp->loss_hist) HIP_TRY(ctx, hipFree(mlp->loss_hist));
            mlp->loss_hist = nullptr;
            mlp->loss_cap = 0;
            HIP_TRY(ctx, hipMalloc(&mlp->loss_hist, (size_t) n_steps * sizeof(float)));
            mlp->loss_cap = n_steps;
        }
        const int rows = grad_rows(batch);
        for (int32_t i = 0; i < n_steps; ++i) {
            const int64_t o = offsets[i];
            hipLaunchKernelGGL(k_mlp_grad, dim3(rows), dim3(MLP_GRAD_THREADS), 0, ctx->stream, mlp->p, mlp->wb, mlp->wt, x + MLP_IN * o,
                               uv + 2 * o, w ? w + o : nullptr, (int64_t) batch, mlp->partial, mlp->aux);
            HIP_TRY(ctx, hipGetLastError());
            float bc1, bc2;
            adam_corrections(mlp, &bc1, &bc2);
            hipLaunchKernelGGL(k_mlp_reduce_adam, dim3((MLP_PARAMS + 63) / 64), dim3(256), 0, ctx->stream,
                               (const float *) mlp->partial, (const float *) mlp->aux, rows, mlp->p, mlp->m, mlp->v,
                               mlp->wb, mlp->wt, lr, ADAM_B1, ADAM_B2, ADAM_EPS, bc1, bc2, losses ? mlp->loss_hist + i : nullptr);
            HIP_TRY(ctx, hipGetLastError());
        }
        if (losses && n_steps > 0) {
            HIP_TRY(ctx, hipMemcpyAsync(losses, mlp->loss_hist, (size_t) n_steps * sizeof(float), hipMemcpyDeviceToHost,
                                        ctx->stream));
            HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        }
        return SMS_OK;
    });
}

sms_status sms_ctx_set_proposal(sms_ctx *ctx, sms_mlp *mlp, int32_t shape, float defensive) {
    return guarded(ctx, [&]() -> sms_status {
        if (!ctx) return SMS_E_INVALID;
        if (shape < 0) {
            if (mlp) return fail(ctx, SMS_E_INVALID, "sms_ctx_set_proposal: a network needs a caster shape");
            ctx->prop_mlp = nullptr; ctx->prop_shape = -1; ctx->prop_defensive = 1.f;
            ctx->gen++;
            return SMS_OK;
        }
        if (shape >= (int) ctx->h_shapes.size()) return fail(ctx, SMS_E_INVALID, "sms_ctx_set_proposal: shape out of range");
        const DShape &sh = ctx->h_shapes[shape];
        const bool glinty = (sh.flags & SMS_FLAG_GLINTY) && sh.bsdf == SMS_BSDF_NORMALMAP;
        if (!glinty &&
            (sh.type != SMS_SHAPE_MESH || !sh.has_uv || !(sh.flags & (SMS_FLAG_CAUSTIC_CASTER_SS | SMS_FLAG_CAUSTIC_CASTER_MS))))
            return fail(ctx, SMS_E_INVALID, "sms_ctx_set_proposal: the shape must be a caster mesh with uvs or a glinty shape");
        if (mlp && mlp->ctx != ctx) return fail(ctx, SMS_E_INVALID, "sms_ctx_set_proposal: network of another context");
        if (mlp && !(defensive > 0.f && defensive <= 1.f))
            return fail(ctx, SMS_E_INVALID, "sms_ctx_set_proposal: defensive weight must be in (0, 1]");
        if (!glinty && ctx->grid_shape != shape) {
            sms_status st = build_uv_grid(ctx, shape);
            if (st) return st;
        }
        ctx->prop_mlp = mlp;
        ctx->prop_shape = shape;
        ctx->prop_defensive = mlp ? defensive : 1.f;
        ctx->gen++;
        return SMS_OK;
    });
}

sms_status sms_ctx_set_proposal_anchor(sms_ctx *ctx, float radius) {
    return guarded(ctx, [&]() -> sms_status {
        if (!ctx || !(radius >= 0.f && radius <= 0.5f))
            return fail(ctx, SMS_E_INVALID, "sms_ctx_set_proposal_anchor: radius must be in [0, 1/2]");
        ctx->prop_radius = radius;
        ctx->gen++;
        return SMS_OK;
    });
}

sms_status sms_ss_proposal_anchors(sms_ctx *ctx, const sms_ss_query *q, int64_t n, uint64_t seed, float *anchor) {
    return guarded(ctx, [&]() -> sms_status {
        if (!ctx || n < 0 || (n > 0 && (!q || !anchor))) return fail(ctx, SMS_E_INVALID, "proposal anchors: bad arguments");
        if (ctx->prop_shape < 0) return fail(ctx, SMS_E_INVALID, "proposal anchors: no proposal shape set");
        hipLaunchKernelGGL(k_set_seed, dim3(1), dim3(64), 0, ctx->stream, ctx->seeds + SEED_API, seed);
        HIP_TRY(ctx, hipGetLastError());
        return launch_anchors(ctx, q, n, ctx->seeds + SEED_API, anchor);
    });
}

sms_status sms_ctx_set_training_capture(sms_ctx *ctx, int64_t capacity, int32_t mode) {
    return guarded(ctx, [&]() -> sms_status {
        if (!ctx || capacity < 0) return fail(ctx, SMS_E_INVALID, "sms_ctx_set_training_capture: bad capacity");
        if (mode != SMS_CAPTURE_SOLUTIONS && mode != SMS_CAPTURE_ESTIMATES)
            return fail(ctx, SMS_E_INVALID, "sms_ctx_set_training_capture: unknown mode");
        ctx->cap_mode = mode;
        ctx->gen++;
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        if (ctx->cap) { (void) hipFree(ctx->cap); ctx->cap = nullptr; }
        if (ctx->cap_counts) { (void) hipFree(ctx->cap_counts); ctx->cap_counts = nullptr; }
        ctx->cap_sub = 0;
        if (capacity == 0) return SMS_OK;
        const size_t sub = (size_t) ((capacity + SPREAD - 1) / SPREAD);
        HIP_TRY(ctx, hipMalloc(&ctx->cap, (size_t) SPREAD * sub * 12 * sizeof(float)));
        HIP_TRY(ctx, hipMalloc(&ctx->cap_counts, (size_t) SPREAD * SPREAD_STRIDE * sizeof(unsigned long long)));
        HIP_TRY(ctx, hipMemset(ctx->cap_counts, 0, (size_t) SPREAD * SPREAD_STRIDE * sizeof(unsigned long long)));
        ctx->cap_sub = sub;
        return SMS_OK;
    });
}

sms_status sms_ctx_training_samples(sms_ctx *ctx, int64_t *count, float *x, float *uv, float *w) {
    return guarded(ctx, [&]() -> sms_status {
        if (!ctx || !count) return SMS_E_INVALID;
        *count = 0;
        if (!ctx->cap_sub) return SMS_OK;
        std::vector<unsigned long long> h((size_t) SPREAD * SPREAD_STRIDE);
        HIP_TRY(ctx, hipMemcpyAsync(h.data(), ctx->cap_counts, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                    ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        int64_t total = 0;
        for (int s = 0; s < SPREAD; ++s) total += (int64_t) std::min<unsigned long long>(h[(size_t) s * SPREAD_STRIDE], ctx->cap_sub);
        *count = total;
        if (x && uv && total > 0) {
            // records in an order that does not depend on scheduling: sorted by a hash of their words
            float *rec = nullptr;
            uint64_t *keys = nullptr;
            uint32_t *idx = nullptr;
            void *tmp = nullptr;
            size_t tmp_bytes = 0;
            const int nn = (int) total;
            hipError_t e = hipMalloc(&rec, (size_t) total * 12 * sizeof(float));
            if (e == hipSuccess) e = hipMalloc(&keys, (size_t) total * 2 * sizeof(uint64_t));
            if (e == hipSuccess) e = hipMalloc(&idx, (size_t) total * 2 * sizeof(uint32_t));
            if (e == hipSuccess)
                e = hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, keys, keys + total, idx, idx + total, nn, 0, 64,
                                                       ctx->stream);
            if (e == hipSuccess) e = hipMalloc(&tmp, std::max<size_t>(tmp_bytes, 1));
            if (e == hipSuccess) {
                hipLaunchKernelGGL(k_capture_keys, dim3(SPREAD), dim3(256), 0, ctx->stream, ctx->cap, ctx->cap_counts,
                                   (unsigned long long) ctx->cap_sub, rec, keys, idx);
                e = hipGetLastError();
            }
            if (e == hipSuccess)
                e = hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, keys, keys + total, idx, idx + total, nn, 0, 64,
                                                       ctx->stream);
            if (e == hipSuccess) {
                hipLaunchKernelGGL(k_capture_scatter, dim3((unsigned) ((total + 255) / 256)), dim3(256), 0, ctx->stream, rec,
                                   idx + total, (int64_t) total, x, uv, w);
                e = hipGetLastError();
            }
            if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
            for (void *b : {(void *) rec, (void *) keys, (void *) idx, tmp})
                if (b) (void) hipFree(b);
            if (e != hipSuccess)
                return fail(ctx, e == hipErrorOutOfMemory ? SMS_E_OOM : SMS_E_HIP,
                            std::string("sms_ctx_training_samples: ") + hipGetErrorString(e));
        }
        return SMS_OK;
    });
}

sms_status sms_ctx_clear_training(sms_ctx *ctx) {
    return guarded(ctx, [&]() -> sms_status {
        if (!ctx) return SMS_E_INVALID;
        if (!ctx->cap_counts) return SMS_OK;
        HIP_TRY(ctx, hipMemsetAsync(ctx->cap_counts, 0, (size_t) SPREAD * SPREAD_STRIDE * sizeof(unsigned long long),
                                    ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        return SMS_OK;
    });
}

namespace {

sms_status proposal_io(sms_ctx *ctx, const sms_ss_query *q, int64_t n, uint64_t seed, int ms, float *x, float *out) {
    return guarded(ctx, [&]() -> sms_status {
        if (!ctx || n < 0 || (n > 0 && (!q || !(x || out)))) return fail(ctx, SMS_E_INVALID, "proposal inputs: bad arguments");
        if (out && !ctx->prop_mlp) return fail(ctx, SMS_E_INVALID, "proposal outputs: no proposal network set");
        sms_status st = out ? ensure_prop(ctx, (size_t) std::max<int64_t>(n, 1)) : SMS_OK;
        if (st) return st;
        hipLaunchKernelGGL(k_set_seed, dim3(1), dim3(64), 0, ctx->stream, ctx->seeds + SEED_API, seed);
        HIP_TRY(ctx, hipGetLastError());
        if (!out) return launch_inputs(ctx, q, n, ctx->seeds + SEED_API, ms, x);
        st = launch_inputs(ctx, q, n, ctx->seeds + SEED_API, ms, ctx->prop_x);
        if (!st) st = launch_infer(ctx->prop_mlp, ctx->prop_x, n, out);
        return st;
    });
}

} // namespace

sms_status glint_proposal_io(sms_ctx *ctx, const sms_glint_query *q, int64_t n, uint64_t seed, float *x, float *out) {
    if (!ctx || n < 0 || (n > 0 && (!q || !(x || out)))) return fail(ctx, SMS_E_INVALID, "proposal inputs: bad arguments");
    if (out && !ctx->prop_mlp) return fail(ctx, SMS_E_INVALID, "proposal outputs: no proposal network set");
    sms_status st = out ? ensure_prop(ctx, (size_t) std::max<int64_t>(n, 1)) : SMS_OK;
    if (st) return st;
    hipLaunchKernelGGL(k_set_seed, dim3(1), dim3(64), 0, ctx->stream, ctx->seeds + SEED_API, seed);
    HIP_TRY(ctx, hipGetLastError());
    if (!out) return launch_glint_inputs(ctx, glint_queries_of(q), n, ctx->seeds + SEED_API, x);
    st = launch_glint_inputs(ctx, glint_queries_of(q), n, ctx->seeds + SEED_API, ctx->prop_x);
    if (!st) st = launch_infer(ctx->prop_mlp, ctx->prop_x, n, out);
    return st;
}

sms_status sms_glint_proposal_inputs(sms_ctx *ctx, const sms_glint_query *q, int64_t n, uint64_t seed, float *x) {
    return guarded(ctx, [&]() -> sms_status { return glint_proposal_io(ctx, q, n, seed, x, nullptr); });
}

sms_status sms_glint_proposal_outputs(sms_ctx *ctx, const sms_glint_query *q, int64_t n, uint64_t seed, float *out) {
    return guarded(ctx, [&]() -> sms_status { return glint_proposal_io(ctx, q, n, seed, nullptr, out); });
}

sms_status sms_ss_proposal_inputs(sms_ctx *ctx, const sms_ss_query *q, int64_t n, uint64_t seed, float *x) {
    return guarded(ctx, [&]() -> sms_status {
        return proposal_io(ctx, q, n, seed, 0, x, nullptr);
    });
}

sms_status sms_ss_proposal_outputs(sms_ctx *ctx, const sms_ss_query *q, int64_t n, uint64_t seed, float *out) {
    return guarded(ctx, [&]() -> sms_status {
        return proposal_io(ctx, q, n, seed, 0, nullptr, out);
    });
}

sms_status sms_ms_proposal_inputs(sms_ctx *ctx, const sms_ss_query *q, int64_t n, uint64_t seed, float *x) {
    return guarded(ctx, [&]() -> sms_status {
        return proposal_io(ctx, q, n, seed, 1, x, nullptr);
    });
}

sms_status sms_ms_proposal_outputs(sms_ctx *ctx, const sms_ss_query *q, int64_t n, uint64_t seed, float *out) {
    return guarded(ctx, [&]() -> sms_status {
        return proposal_io(ctx, q, n, seed, 1, nullptr, out);
    });
}

} // extern "C"
