// sms_path_host.hip -- host side of the path integrators (path_sms_ss / path_filtered_ss) and
// of the film reconstruction: validation, per-batch buffers, the launch sequence of a pass.
#include "host_ctx.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>

using namespace smshost;

// per-batch buffers of a pass: slots of SMS_PATH_CHUNK samples x depth_slots, the dense lists
// and the scan's scratch; grown on demand, kept by the context
struct PathBuffers {
    size_t slot_rows = 0;            // samples x depth slots the slot arrays hold
    size_t rows = 0;                 // samples the per-sample arrays hold
    sms_ss_query *q = nullptr; float *q_thr = nullptr;
    sms_ss_walk_in *m = nullptr; float *m_val = nullptr, *m_dir = nullptr, *m_ref = nullptr;
    size_t m_rows = 0;
    int *nq = nullptr, *nm = nullptr, *q_off = nullptr, *m_off = nullptr;
    sms_ss_query *dq = nullptr; float *dq_rgb = nullptr; size_t dq_cap = 0;
    sms_ss_walk_in *dm = nullptr; sms_ss_walk_out *dm_out = nullptr; size_t dm_cap = 0;
    // multi-scatter MNEE checks: slots (walks, the path's chain, its length) and dense lists
    sms_ms_walk_in *mm = nullptr; float *mm_pos = nullptr; int *mm_npos = nullptr; size_t mm_rows = 0;
    sms_ms_walk_in *dmm = nullptr; sms_ms_walk_out *dmm_out = nullptr; size_t dmm_cap = 0;
    void *scan_tmp = nullptr; size_t scan_bytes = 0;
    int *totals = nullptr;           // pinned host: [0] queries, [1] MNEE walks
};

namespace {

template <typename T>
sms_status grow(sms_ctx *ctx, T **p, size_t count) {
    if (*p) (void) hipFree(*p);
    *p = nullptr;
    HIP_TRY(ctx, hipMalloc((void **) p, std::max<size_t>(count, 1) * sizeof(T)));
    return SMS_OK;
}

// slot and per-sample arrays for `rows` samples of `depth_slots` slots; MNEE slots when asked
// (multi-scatter ones with ms)
sms_status path_alloc(sms_ctx *ctx, size_t rows, int depth_slots, bool mnee, bool ms) {
    if (!ctx->path) ctx->path = new PathBuffers();
    PathBuffers &B = *ctx->path;
    const size_t slot_rows = rows * (size_t) std::max(depth_slots, 1);
    sms_status st = SMS_OK;
    if (slot_rows > B.slot_rows) {
        if (!st) st = grow(ctx, &B.q, slot_rows);
        if (!st) st = grow(ctx, &B.q_thr, 3 * slot_rows);
        if (!st) B.slot_rows = slot_rows;
    }
    if (!st && mnee && ms && slot_rows > B.mm_rows) {
        if (!st) st = grow(ctx, &B.mm, slot_rows);
        if (!st) st = grow(ctx, &B.mm_pos, 3 * SMS_MS_MAX_BOUNCES * slot_rows);
        if (!st) st = grow(ctx, &B.mm_npos, slot_rows);
        if (!st) st = grow(ctx, &B.m_val, 3 * slot_rows);
        if (!st) { B.mm_rows = slot_rows; B.m_rows = 0; }
    }
    if (!st && mnee && !ms && slot_rows > B.m_rows) {
        if (!st) st = grow(ctx, &B.m, slot_rows);
        if (!st) st = grow(ctx, &B.m_val, 3 * slot_rows);
        if (!st) st = grow(ctx, &B.m_dir, 3 * slot_rows);
        if (!st) st = grow(ctx, &B.m_ref, 3 * slot_rows);
        if (!st) { B.m_rows = slot_rows; B.mm_rows = 0; }
    }
    if (!st && rows > B.rows) {
        // counts and offsets hold one entry past the batch: the scan's last offset is the total
        if (!st) st = grow(ctx, &B.nq, rows + 1);
        if (!st) st = grow(ctx, &B.nm, rows + 1);
        if (!st) st = grow(ctx, &B.q_off, rows + 1);
        if (!st) st = grow(ctx, &B.m_off, rows + 1);
        size_t bytes = 0;
        if (!st) HIP_TRY(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, B.nq, B.q_off, (int) (rows + 1), ctx->stream));
        if (!st && bytes > B.scan_bytes) {
            if (B.scan_tmp) (void) hipFree(B.scan_tmp);
            B.scan_tmp = nullptr;
            HIP_TRY(ctx, hipMalloc(&B.scan_tmp, bytes));
            B.scan_bytes = bytes;
        }
        if (!st) B.rows = rows;
    }
    if (!st && !B.totals) HIP_TRY(ctx, hipHostMalloc((void **) &B.totals, 2 * sizeof(int)));
    return st;
}

sms_status scan(sms_ctx *ctx, const int *count, int *off, int64_t rows) {
    PathBuffers &B = *ctx->path;
    size_t bytes = B.scan_bytes;
    HIP_TRY(ctx, hipcub::DeviceScan::ExclusiveSum(B.scan_tmp, bytes, count, off, (int) (rows + 1), ctx->stream));
    return SMS_OK;
}

// the scene features the path integrators evaluate
sms_status check_path_scene(sms_ctx *ctx, const sms_path_config *pc) {
    if (!pc) return fail(ctx, SMS_E_INVALID, "path integrator: config is NULL");
    if (pc->integrator < SMS_PATH_SMS_SS || pc->integrator > SMS_PATH_FILTERED_MS)
        return fail(ctx, SMS_E_INVALID, "path integrator: unknown integrator");
    if ((pc->integrator == SMS_PATH_SMS_MS || pc->integrator == SMS_PATH_FILTERED_MS) &&
        (ctx->scene.cfg.bounces < 1 || ctx->scene.cfg.bounces > SMS_MS_MAX_BOUNCES))
        return fail(ctx, SMS_E_UNSUPPORTED, "path integrator: multi-scatter chains of 1..4 vertices are built");
    if (pc->max_depth < -1 || pc->max_depth > SMS_PATH_MAX_DEPTH)
        return fail(ctx, SMS_E_INVALID, "path integrator: max_depth must be -1 or in [0, SMS_PATH_MAX_DEPTH]");
    if (!ctx->scene.has_cam) return fail(ctx, SMS_E_INVALID, "path integrator: the scene has no camera");
    for (const DShape &sh : ctx->h_shapes) {
        if (sh.alpha > 0.f)
            return fail(ctx, SMS_E_UNSUPPORTED, "path integrator: rough BSDFs are not evaluated by the path tracer");
    }
    return SMS_OK;
}

uint64_t mix64_host(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// estimator seed of batch c of a pass
uint64_t path_chunk_seed(uint64_t seed, uint64_t c) { return mix64_host(seed ^ mix64_host(c + 0x632BE59BD9B4E019ULL)); }

void add_stats(sms_stats &tot, const sms_stats &s) {
    tot.walks += s.walks; tot.walks_converged += s.walks_converged; tot.newton_iterations += s.newton_iterations;
    tot.bernoulli_calls += s.bernoulli_calls; tot.bernoulli_trials += s.bernoulli_trials;
    tot.bernoulli_max = std::max(tot.bernoulli_max, s.bernoulli_max);
    tot.overflow_queries += s.overflow_queries; tot.tail_iterations += s.tail_iterations;
}

sms_status path_pass(sms_ctx *ctx, const sms_path_config *pc, int32_t spp, uint64_t seed, float *rgb, float *pos,
                     sms_stats *stats) {
    sms_status st = check_path_scene(ctx, pc);
    if (st) return st;
    if (spp <= 0 || !rgb || !pos) return fail(ctx, SMS_E_INVALID, "path integrator: bad arguments");
    if (ctx->capturing) return fail(ctx, SMS_E_INVALID, "path integrator: the context is capturing a graph");
    const sms_config &cfg = ctx->scene.cfg;
    const bool ms = pc->integrator == SMS_PATH_SMS_MS || pc->integrator == SMS_PATH_FILTERED_MS;
    const bool sms = pc->integrator == SMS_PATH_SMS_SS || pc->integrator == SMS_PATH_SMS_MS;
    const bool mnee = sms && cfg.mnee_init && !pc->biased_mnee;
    const int depth_slots = pc->max_depth < 0 ? SMS_PATH_MAX_DEPTH : pc->max_depth;
    const int64_t n = (int64_t) ctx->scene.cam.w * ctx->scene.cam.h * spp;
    const int64_t batch = std::min<int64_t>(n, SMS_PATH_CHUNK);
    sms_stats tot;
    memset(&tot, 0, sizeof(tot));
    st = path_alloc(ctx, (size_t) batch, depth_slots, mnee, ms);
    if (st) return st;
    PathBuffers &B = *ctx->path;
    for (int64_t c0 = 0; c0 < n; c0 += SMS_PATH_CHUNK) {
        const int64_t nb = std::min<int64_t>(SMS_PATH_CHUNK, n - c0);
        PathSlots P;
        P.depth_slots = depth_slots;
        P.q = B.q; P.q_thr = B.q_thr; P.nq = B.nq;
        P.m = mnee && !ms ? B.m : nullptr; P.m_val = B.m_val; P.m_dir = B.m_dir; P.m_ref = B.m_ref;
        P.mm = mnee && ms ? B.mm : nullptr; P.mm_pos = B.mm_pos; P.mm_npos = B.mm_npos;
        P.nm = mnee ? B.nm : nullptr;
        P.L = rgb + 3 * c0; P.pos = pos + 2 * c0;
        // the scan's extra entry is zero
        HIP_TRY(ctx, hipMemsetAsync(B.nq + nb, 0, sizeof(int), ctx->stream));
        HIP_TRY(ctx, hipMemsetAsync(B.nm + nb, 0, sizeof(int), ctx->stream));
        const unsigned blocks = (unsigned) ((nb + 255) / 256);
        hipLaunchKernelGGL(k_path_trace, dim3(blocks), dim3(256), 0, ctx->stream, ctx->scene, *pc, seed, c0, nb,
                           (int) spp, P);
        HIP_TRY(ctx, hipGetLastError());
        st = scan(ctx, B.nq, B.q_off, nb);
        if (!st && mnee) st = scan(ctx, B.nm, B.m_off, nb);
        if (st) return st;
        HIP_TRY(ctx, hipMemcpyAsync(&B.totals[0], B.q_off + nb, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        if (mnee) HIP_TRY(ctx, hipMemcpyAsync(&B.totals[1], B.m_off + nb, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        const int64_t nq = B.totals[0], nm = mnee ? B.totals[1] : 0;
        if ((size_t) nq > B.dq_cap) {
            st = grow(ctx, &B.dq, (size_t) nq);
            if (!st) st = grow(ctx, &B.dq_rgb, 3 * (size_t) nq);
            if (st) return st;
            B.dq_cap = (size_t) nq;
        }
        if (!ms && (size_t) nm > B.dm_cap) {
            st = grow(ctx, &B.dm, (size_t) nm);
            if (!st) st = grow(ctx, &B.dm_out, (size_t) nm);
            if (st) return st;
            B.dm_cap = (size_t) nm;
        }
        if (ms && (size_t) nm > B.dmm_cap) {
            st = grow(ctx, &B.dmm, (size_t) nm);
            if (!st) st = grow(ctx, &B.dmm_out, (size_t) nm);
            if (st) return st;
            B.dmm_cap = (size_t) nm;
        }
        if (nq > 0 || nm > 0) {
            hipLaunchKernelGGL(k_path_compact, dim3(blocks), dim3(256), 0, ctx->stream, P, nb, B.q_off,
                               nq > 0 ? B.dq : (sms_ss_query *) nullptr, B.m_off,
                               nm > 0 && !ms ? B.dm : (sms_ss_walk_in *) nullptr,
                               nm > 0 && ms ? B.dmm : (sms_ms_walk_in *) nullptr);
            HIP_TRY(ctx, hipGetLastError());
        }
        if (nq > 0) {
            // the batch's SMS estimates: the single- or multi-scatter estimator with the batch's seed
            st = ms ? ms_prepare(ctx, B.dq, nq, B.dq_rgb) : ss_prepare(ctx, B.dq, nq, B.dq_rgb);
            if (!st) st = frame_begin(ctx, path_chunk_seed(seed, (uint64_t) (c0 / SMS_PATH_CHUNK)));
            if (!st) st = ms ? ms_enqueue(ctx, B.dq, nq, B.dq_rgb) : ss_enqueue(ctx, B.dq, nq, B.dq_rgb);
            sms_stats s;
            if (!st) st = estimator_stats(ctx, &s);
            if (st) return st;
            add_stats(tot, s);
        }
        if (nm > 0 && !ms) {
            hipLaunchKernelGGL(k_ss_walk, dim3((unsigned) ((nm + 255) / 256)), dim3(256), 0, ctx->stream, ctx->scene,
                               B.dm, nm, B.dm_out);
            HIP_TRY(ctx, hipGetLastError());
        }
        if (nm > 0 && ms) {
            const dim3 g((unsigned) ((nm + 255) / 256));
            if (cfg.bounces <= 2) hipLaunchKernelGGL(k_ms_walk_2, g, dim3(256), 0, ctx->stream, ctx->scene, B.dmm, nm, B.dmm_out);
            else hipLaunchKernelGGL(k_ms_walk_4, g, dim3(256), 0, ctx->stream, ctx->scene, B.dmm, nm, B.dmm_out);
            HIP_TRY(ctx, hipGetLastError());
        }
        hipLaunchKernelGGL(k_path_gather, dim3(blocks), dim3(256), 0, ctx->stream, P, nb, B.q_off,
                           nq > 0 ? B.dq_rgb : (const float *) nullptr, B.m_off,
                           nm > 0 && !ms ? B.dm_out : (const sms_ss_walk_out *) nullptr,
                           nm > 0 && ms ? B.dmm_out : (const sms_ms_walk_out *) nullptr, cfg.uniqueness_threshold);
        HIP_TRY(ctx, hipGetLastError());
    }
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (stats) *stats = tot;
    return SMS_OK;
}

// the discretized filter (GaussianFilter, src/rfilters/gaussian.cpp:34-48, and
// ReconstructionFilter::init_discretization, src/libcore/rfilter.cpp:9-20); stddev 0: box
FilmFilter film_filter(float stddev) {
    FilmFilter F;
    memset(&F, 0, sizeof(F));
    // the box branch of ImageBlock::put (imageblock.cpp:157-163) also takes filters of radius <= 1
    if (!(stddev > 0.f) || !(4.f * stddev > 1.f)) { F.box = 1; return F; }
    const float radius = 4.f * stddev, alpha = -1.f / (2.f * stddev * stddev), bias = std::exp(alpha * (radius * radius));
    for (int i = 0; i < 31; ++i) {
        const float x = (radius * (float) i) / 31.f;
        F.values[i] = std::max(0.f, std::exp(alpha * (x * x)) - bias);
    }
    F.values[31] = 0.f;
    F.radius = radius;
    F.scale = 31.f / radius;
    F.n = (int) std::ceil((radius - 2.f * RAY_EPS_F) * 2.f);
    return F;
}

} // namespace

void path_release(sms_ctx *ctx) {
    if (!ctx || !ctx->path) return;
    PathBuffers &B = *ctx->path;
    void *dev[] = {B.q, B.q_thr, B.m, B.m_val, B.m_dir, B.m_ref, B.nq, B.nm, B.q_off, B.m_off,
                   B.dq, B.dq_rgb, B.dm, B.dm_out, B.scan_tmp, B.mm, B.mm_pos, B.mm_npos, B.dmm, B.dmm_out};
    for (void *p : dev)
        if (p) (void) hipFree(p);
    if (B.totals) (void) hipHostFree(B.totals);
    delete ctx->path;
    ctx->path = nullptr;
}

extern "C" {

sms_status sms_path_config_default(sms_path_config *pc) {
    if (!pc) return SMS_E_INVALID;
    memset(pc, 0, sizeof(*pc));
    pc->integrator = SMS_PATH_SMS_SS;
    pc->max_depth = -1;
    pc->rr_depth = 5;
    pc->filter_stddev = 0.5f;
    return SMS_OK;
}

sms_status sms_path_samples(sms_ctx *ctx, const sms_path_config *pc, int32_t spp, uint64_t seed, float *rgb,
                            float *pos, sms_stats *stats) {
    return guarded(ctx, [&]() -> sms_status {
        if (!ctx) return SMS_E_INVALID;
        return path_pass(ctx, pc, spp, seed, rgb, pos, stats);
    });
}

sms_status sms_film_splat(sms_ctx *ctx, const float *rgb, const float *pos, int32_t spp, float stddev, float *image,
                          float *weight) {
    return guarded(ctx, [&]() -> sms_status {
        if (!ctx) return SMS_E_INVALID;
        if (!ctx->scene.has_cam) return fail(ctx, SMS_E_INVALID, "film: the scene has no camera");
        if (spp <= 0 || !rgb || !pos || !image || !weight || !(stddev >= 0.f) || stddev > 16.f)
            return fail(ctx, SMS_E_INVALID, "film: bad arguments");
        const int w = ctx->scene.cam.w, h = ctx->scene.cam.h;
        const FilmFilter F = film_filter(stddev);
        // a pixel gathers the samples of its 5 x 5 neighbourhood: taps reach 2 pixels
        if (!F.box && F.radius > 2.f)
            return fail(ctx, SMS_E_UNSUPPORTED, "film: filter radius above 2 pixels (stddev above 0.5)");
        const int64_t px = (int64_t) w * h;
        hipLaunchKernelGGL(k_film_splat, dim3((unsigned) ((px + 255) / 256)), dim3(256), 0, ctx->stream, F, w, h,
                           (int) spp, rgb, pos, image, weight);
        HIP_TRY(ctx, hipGetLastError());
        return SMS_OK;
    });
}

} // extern "C"
