// sms_noalloc_host.hip -- sms_ss_estimate_batch_noalloc: the single-scatter estimator without
// any allocation, for callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops).  Its scratch (work counters, seed slots, per-(query, caster) partial
// results, the spread overflow list of the Bernoulli tail) is one file-scope device array of
// NOALLOC_SCRATCH_BYTES, held by every device the library runs on until the process ends.  The
// array's device address is looked up once per device when a context is created on it.
#include "host_ctx.h"

#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

using namespace smsdev;
using namespace smshost;

namespace {

// (query, caster) pairs one call may hold
constexpr size_t NOALLOC_MAX_ITEMS = (size_t) 1 << 18;
// the spread overflow list: SPREAD sub-lists, each ceil(items / SPREAD) + 256 x casters records
constexpr size_t NOALLOC_OVER_RECS = NOALLOC_MAX_ITEMS + (size_t) SPREAD * 256 * 4;

// scratch layout (byte offsets, each 256-aligned)
constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t) 255; }
constexpr size_t OFF_COUNTERS = 0;
constexpr size_t OFF_SEEDS = align256(OFF_COUNTERS + N_COUNTERS * sizeof(unsigned long long));
constexpr size_t OFF_PART = align256(OFF_SEEDS + SEED_SLOTS * sizeof(uint64_t));
constexpr size_t OFF_INVP = align256(OFF_PART + 3 * NOALLOC_MAX_ITEMS * sizeof(float));
constexpr size_t OFF_OVER = align256(OFF_INVP + NOALLOC_MAX_ITEMS * sizeof(float));
constexpr size_t NOALLOC_SCRATCH_BYTES = align256(OFF_OVER + NOALLOC_OVER_RECS * sizeof(OverflowRec));

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[MAX_DEVICES] = {nullptr};

// the work buffers of the context, saved while a call runs on the scratch
struct SavedWork {
    unsigned long long *counters;
    uint64_t *seeds;
    float *part, *invp;
    size_t part_cap;
    OverflowRec *gover;
    size_t gover_cap, over_sub;
};

SavedWork swap_in(sms_ctx *ctx, unsigned char *base, size_t sub_cap) {
    SavedWork s = {ctx->counters, ctx->seeds, ctx->part, ctx->invp, ctx->part_cap, ctx->gover, ctx->gover_cap,
                   ctx->over_sub};
    ctx->counters = (unsigned long long *) (base + OFF_COUNTERS);
    ctx->seeds = (uint64_t *) (base + OFF_SEEDS);
    ctx->part = (float *) (base + OFF_PART);
    ctx->invp = (float *) (base + OFF_INVP);
    ctx->part_cap = NOALLOC_MAX_ITEMS;
    ctx->gover = (OverflowRec *) (base + OFF_OVER);
    ctx->gover_cap = NOALLOC_OVER_RECS;
    ctx->over_sub = sub_cap;
    return s;
}

void swap_out(sms_ctx *ctx, const SavedWork &s) {
    ctx->counters = s.counters; ctx->seeds = s.seeds; ctx->part = s.part; ctx->invp = s.invp;
    ctx->part_cap = s.part_cap; ctx->gover = s.gover; ctx->gover_cap = s.gover_cap; ctx->over_sub = s.over_sub;
}

// statistics of the call from the scratch counters, into caller-owned stack memory
sms_status stats_from(sms_ctx *ctx, const unsigned long long *dev_counters, sms_stats *stats) {
    unsigned long long h[N_COUNTERS];
    HIP_TRY(ctx, hipMemcpyAsync(h, dev_counters, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (stats) {
        unsigned long long v[8] = {0, 0, 0, 0, 0, 0, 0, 0}, over = 0;
        for (int s = 0; s < SPREAD; ++s) {
            const unsigned long long *cp = &h[STATS_OFF + s * SPREAD_STRIDE];
            for (int j = 0; j < 8; ++j) v[j] = j == 5 ? (cp[j] > v[j] ? cp[j] : v[j]) : v[j] + cp[j];
            over += h[OVER_OFF + s * SPREAD_STRIDE];
        }
        memset(stats, 0, sizeof(*stats));
        stats->walks = v[0]; stats->walks_converged = v[1]; stats->newton_iterations = v[2];
        stats->bernoulli_calls = v[3]; stats->bernoulli_trials = v[4]; stats->bernoulli_max = v[5];
        stats->overflow_queries = h[1] + over;
        stats->tail_iterations = v[6];
    }
    return SMS_OK;
}

} // namespace

// Called by sms_ctx_create with the context's device current: looks the scratch array's device
// address up once per device (each device holds its own copy of the array).
sms_status noalloc_init_device(sms_ctx *ctx) {
    if (ctx->device < 0 || ctx->device >= MAX_DEVICES) return fail(ctx, SMS_E_UNSUPPORTED, "device index above 63");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[ctx->device]) {
        void *p = nullptr;
        HIP_TRY(ctx, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[ctx->device] = p;
    }
    ctx->scratch = g_scratch_ptr[ctx->device];
    return SMS_OK;
}

extern "C" {

sms_status sms_ss_estimate_batch_noalloc(sms_ctx *ctx, const sms_ss_query *q, int64_t n, uint64_t seed,
                                         float *out_rgb, sms_stats *stats) {
    return guarded(ctx, [&]() -> sms_status {
        if (!ctx || n < 0 || (n > 0 && (!q || !out_rgb))) return fail_noexcept(ctx, SMS_E_INVALID, "bad estimator arguments");
        if (!ctx->scratch) return fail_noexcept(ctx, SMS_E_INVALID, "scratch of the device not set up");
        if (ctx->capturing) return fail_noexcept(ctx, SMS_E_INVALID, "context is capturing a graph");
        const sms_config &cfg = ctx->scene.cfg;
        const int nc = std::max(ctx->scene.n_casters_ss, 1);
        if (cfg.biased && cfg.max_trials > 128) return fail_noexcept(ctx, SMS_E_UNSUPPORTED, "biased trial sets above 128 are not built");
        if (ctx->prop_mlp && ctx->prop_shape >= 0)
            return fail_noexcept(ctx, SMS_E_UNSUPPORTED, "the allocation-free estimator runs without a learned proposal");
        if (ctx->cap && ctx->cap_sub > 0)
            return fail_noexcept(ctx, SMS_E_UNSUPPORTED, "the allocation-free estimator runs without a training capture");
        const size_t items = (size_t) n * (size_t) nc;
        const size_t sub_cap = ((size_t) std::max<int64_t>(n, 1) * nc + SPREAD - 1) / SPREAD + 256 * (size_t) nc;
        if (items > NOALLOC_MAX_ITEMS || (size_t) SPREAD * sub_cap > NOALLOC_OVER_RECS)
            return fail_noexcept(ctx, SMS_E_INVALID,
                                 "sms_ss_estimate_batch_noalloc: queries x casters above 262144 (the scratch size)");
        unsigned char *base = (unsigned char *) ctx->scratch;
        SavedWork saved = swap_in(ctx, base, sub_cap);
        sms_status st = frame_begin(ctx, seed);
        if (!st) st = ss_enqueue(ctx, q, n, out_rgb);
        swap_out(ctx, saved);
        if (st) return st;
        st = stats_from(ctx, (const unsigned long long *) (base + OFF_COUNTERS), stats);
        if (st) return st;
        HIP_TRY(ctx, hipEventElapsedTime(&ctx->last_ms, ctx->ev0, ctx->ev1));
        HIP_TRY(ctx, hipEventElapsedTime(&ctx->last_ms2, ctx->ev1, ctx->ev2));
        return SMS_OK;
    });
}

} // extern "C"
