// sms_graph_host.hip -- frame launch sequences captured into HIP graphs (include/sms.h).
//
// A graph holds [k_frame_begin (seed, counters), the estimator's kernels, k_stats_fold]; the
// kernels read the seed from device memory, so a launch with a new seed updates the arguments
// of the first node only and replays the rest as captured.  The graph keeps the context's
// buffers, configuration and schedule as they were at capture: a change of any of them (the
// context's generation counter) makes sms_graph_launch fail until the graph is recreated.
// The reference renders through per-thread scalar calls and has no counterpart.
#include "host_ctx.h"

#include <cstring>
#include <new>
#include <string>

using namespace smshost;

struct sms_graph {
    sms_ctx *ctx = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipGraphNode_t begin = nullptr;           // the frame-begin kernel node
    hipKernelNodeParams begin_params{};       // its launch shape as captured
    uint64_t gen = 0;                         // context generation at capture
    unsigned long long *total = nullptr;      // device: folded statistics since the last reset
    // arguments of the frame-begin node
    unsigned long long *a_counters = nullptr;
    int a_n = 0;
    uint64_t *a_seeds = nullptr;
    uint64_t a_seed = 0;
};

namespace {

void free_graph(sms_graph *g) {
    if (g->exec) (void) hipGraphExecDestroy(g->exec);
    if (g->graph) (void) hipGraphDestroy(g->graph);
    if (g->total) (void) hipFree(g->total);
    delete g;
}

sms_status graph_fail(sms_ctx *ctx, sms_graph *g, hipError_t e, const char *what) {
    free_graph(g);
    return fail(ctx, e == hipErrorOutOfMemory ? SMS_E_OOM : SMS_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// Capture frame begin + body() + statistics fold from the context stream.  body() only launches
// kernels (every buffer it uses was allocated before).
template <class Body>
sms_status capture(sms_ctx *ctx, Body &&body, sms_graph **out) {
    *out = nullptr;
    sms_graph *g = new (std::nothrow) sms_graph();
    if (!g) return fail(ctx, SMS_E_OOM, "graph: host allocation");
    g->ctx = ctx;
    hipError_t e = hipMalloc((void **) &g->total, N_FOLDED * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemset(g->total, 0, N_FOLDED * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return graph_fail(ctx, g, e, "graph: statistics buffer");
    e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return graph_fail(ctx, g, e, "hipStreamBeginCapture");
    ctx->capturing = true;
    sms_status st = frame_begin(ctx, 0);
    if (!st) st = body();
    if (!st) {
        hipLaunchKernelGGL(k_stats_fold, dim3(1), dim3(64), 0, ctx->stream, ctx->counters, g->total);
        if (hipGetLastError() != hipSuccess) st = fail(ctx, SMS_E_HIP, "graph: statistics kernel");
    }
    ctx->capturing = false;
    e = hipStreamEndCapture(ctx->stream, &g->graph);
    if (st) {
        free_graph(g);
        return st;
    }
    if (e != hipSuccess) return graph_fail(ctx, g, e, "hipStreamEndCapture");
    // the frame-begin kernel is the only root of the sequence
    size_t n_roots = 0;
    e = hipGraphGetRootNodes(g->graph, nullptr, &n_roots);
    if (e != hipSuccess) return graph_fail(ctx, g, e, "hipGraphGetRootNodes");
    if (n_roots != 1) {
        free_graph(g);
        return fail(ctx, SMS_E_HIP, "graph: captured sequence does not start with one node");
    }
    e = hipGraphGetRootNodes(g->graph, &g->begin, &n_roots);
    hipGraphNodeType type = hipGraphNodeTypeEmpty;
    if (e == hipSuccess) e = hipGraphNodeGetType(g->begin, &type);
    if (e == hipSuccess && type != hipGraphNodeTypeKernel) {
        free_graph(g);
        return fail(ctx, SMS_E_HIP, "graph: first node is not the frame-begin kernel");
    }
    if (e == hipSuccess) e = hipGraphKernelNodeGetParams(g->begin, &g->begin_params);
    if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return graph_fail(ctx, g, e, "graph: instantiate");
    g->a_counters = ctx->counters;
    g->a_n = N_COUNTERS;
    g->a_seeds = ctx->seeds;
    g->gen = ctx->gen;
    *out = g;
    return SMS_OK;
}

} // namespace

extern "C" {

sms_status sms_glint_graph_create(sms_ctx *ctx, const sms_glint_query *q, int64_t n, float *out_rgb, sms_graph **out) {
    return guarded(ctx, [&]() -> sms_status {
        if (!ctx || !out) return fail(ctx, SMS_E_INVALID, "sms_glint_graph_create: null argument");
        sms_status st = glint_prepare(ctx, q, n, out_rgb, false);
        if (st) return st;
        return capture(ctx, [&] { return glint_enqueue(ctx, glint_queries_of(q), n, out_rgb); }, out);
    });
}

sms_status sms_glint_frame_graph_create(sms_ctx *ctx, int32_t spp, float *out_rgb, sms_graph **out) {
    return guarded(ctx, [&]() -> sms_status {
        if (!ctx || !out || spp <= 0) return fail(ctx, SMS_E_INVALID, "sms_glint_frame_graph_create: bad arguments");
        const int64_t n = (int64_t) ctx->scene.cam.w * ctx->scene.cam.h * spp;
        sms_status st = glint_prepare(ctx, nullptr, n, out_rgb, true);
        if (st) return st;
        return capture(ctx, [&] { return glint_enqueue(ctx, glint_camera_queries(ctx, spp), n, out_rgb); }, out);
    });
}

sms_status sms_ss_graph_create(sms_ctx *ctx, const sms_ss_query *q, int64_t n, float *out_rgb, sms_graph **out) {
    return guarded(ctx, [&]() -> sms_status {
        if (!ctx || !out) return fail(ctx, SMS_E_INVALID, "sms_ss_graph_create: null argument");
        sms_status st = ss_prepare(ctx, q, n, out_rgb);
        if (st) return st;
        return capture(ctx, [&] { return ss_enqueue(ctx, q, n, out_rgb); }, out);
    });
}

sms_status sms_ms_graph_create(sms_ctx *ctx, const sms_ss_query *q, int64_t n, float *out_rgb, sms_graph **out) {
    return guarded(ctx, [&]() -> sms_status {
        if (!ctx || !out) return fail(ctx, SMS_E_INVALID, "sms_ms_graph_create: null argument");
        sms_status st = ms_prepare(ctx, q, n, out_rgb);
        if (st) return st;
        return capture(ctx, [&] { return ms_enqueue(ctx, q, n, out_rgb); }, out);
    });
}

sms_status sms_ss_pass_graph_create(sms_ctx *ctx, int32_t spp, sms_ss_query *q, float *weight, float *out_rgb,
                                    float *image, float *image_sq, float *image_n, sms_graph **out) {
    return guarded(ctx, [&]() -> sms_status {
        if (!ctx || !out) return fail(ctx, SMS_E_INVALID, "sms_ss_pass_graph_create: null argument");
        if (spp <= 0 || !q || !weight || !out_rgb || !image)
            return fail(ctx, SMS_E_INVALID, "sms_ss_pass_graph_create: bad arguments");
        if (!ctx->scene.has_cam) return fail(ctx, SMS_E_INVALID, "scene has no camera");
        const int64_t pixels = (int64_t) ctx->scene.cam.w * ctx->scene.cam.h, n = pixels * spp;
        sms_status st = ss_prepare(ctx, q, n, out_rgb);
        if (st) return st;
        return capture(ctx, [&] {
            sms_status s2 = ss_queries_enqueue(ctx, spp, q, weight);
            if (!s2) s2 = ss_enqueue(ctx, q, n, out_rgb);
            if (!s2) s2 = film_enqueue(ctx, out_rgb, weight, pixels, spp, image, image_sq, image_n);
            return s2;
        }, out);
    });
}

sms_status sms_graph_launch(sms_graph *g, uint64_t seed) {
    return guarded(g ? g->ctx : nullptr, [&]() -> sms_status {
        if (!g) return SMS_E_INVALID;
        sms_ctx *ctx = g->ctx;
        if (g->gen != ctx->gen)
            return fail(ctx, SMS_E_INVALID, "sms_graph_launch: the context's buffers or settings changed since the capture");
        g->a_seed = seed;
        void *args[4] = {(void *) &g->a_counters, (void *) &g->a_n, (void *) &g->a_seeds, (void *) &g->a_seed};
        hipKernelNodeParams p = g->begin_params;
        p.kernelParams = args;
        p.extra = nullptr;
        HIP_TRY(ctx, hipGraphExecKernelNodeSetParams(g->exec, g->begin, &p));
        HIP_TRY(ctx, hipGraphLaunch(g->exec, ctx->stream));
        return SMS_OK;
    });
}

sms_status sms_graph_stats(sms_graph *g, sms_stats *stats, int32_t reset) {
    return guarded(g ? g->ctx : nullptr, [&]() -> sms_status {
        if (!g) return SMS_E_INVALID;
        sms_ctx *ctx = g->ctx;
        unsigned long long h[N_FOLDED];
        HIP_TRY(ctx, hipMemcpyAsync(h, g->total, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        if (stats) {
            memset(stats, 0, sizeof(*stats));
            stats->walks = h[0]; stats->walks_converged = h[1]; stats->newton_iterations = h[2];
            stats->bernoulli_calls = h[3]; stats->bernoulli_trials = h[4]; stats->bernoulli_max = h[5];
            stats->tail_iterations = h[6]; stats->overflow_queries = h[8];
        }
        if (reset) {
            HIP_TRY(ctx, hipMemsetAsync(g->total, 0, sizeof(h), ctx->stream));
            HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        }
        return SMS_OK;
    });
}

sms_status sms_graph_destroy(sms_graph *g) {
    return guarded(g ? g->ctx : nullptr, [&]() -> sms_status {
        if (!g) return SMS_E_INVALID;
        (void) hipStreamSynchronize(g->ctx->stream);
        free_graph(g);
        return SMS_OK;
    });
}

} // extern "C"
