// kernels.h -- kernel declarations and shared device helpers of the kernel and host
// translation units.
#pragma once
#include "dev_ss.h"
#include "dev_glint.h"

namespace smsdev {

// A (query, caster) whose Bernoulli trial loop outgrew the per-lane budget.
struct OverflowRec {
    int64_t q;            // glints: record index; single scatter: query*n_casters + caster
    int64_t k;            // next trial index to run
    float sol[3];         // solution: uv (glints) or direction (single / multi scatter)
    int32_t shapes[4];    // multi scatter: shapes of the first walk's seed chain
    int32_t pad;
};

// Compacted glint footprint (one per query whose setup succeeded), 40 bytes.
struct GlintRec {
    int32_t q;            // query index
    float target[2];      // target slope (manifold_glints.cpp:241-247)
    float uv[2], duv_dx[2], duv_dy[2];
    int32_t pad;
};

// Per-record outcome of the glint walks: status 0 = first walk failed, 1 = resolved,
// 2 = handed to the tail pass (inv_prob written there).
struct GlintRes {
    float uv_sol[2];
    float inv_prob;
    int32_t status;
};

// Device work-queue state of one estimator call.
struct EstimatorWork {
    unsigned long long *cursor;       // next item of the persistent pass
    unsigned long long *n_overflow;   // overflow list length
    unsigned long long *coop_cursor;  // next overflow record (tail pass)
    unsigned long long *n_rec;        // compacted glint records
    unsigned long long *stats;        // sms_stats counters (walks, converged, iterations, calls, trials, max, tail its)
    OverflowRec *overflow;
    GlintRec *rec;
    GlintRes *res;
    int lane_trials;                  // trials a lane runs before handing a query to the tail pass
};

SMS_DEV void wave_add_u64(unsigned long long *dst, unsigned long long v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(dst, v);
}
SMS_DEV void wave_max_u64(unsigned long long *dst, unsigned long long v) {
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    if ((threadIdx.x & 63) == 0 && v) atomicMax(dst, v);
}
SMS_DEV void write_rgb(float *out, int64_t q, V3 v) { out[3 * q] = v.x; out[3 * q + 1] = v.y; out[3 * q + 2] = v.z; }

SMS_DEV SI glint_query_si(const sms_glint_query &g) {
    SI si = si_invalid();
    si.valid = true; si.p = ld3(g.p); si.n = ld3(g.n); si.sh_n = ld3(g.sh_n); si.sh_s = ld3(g.sh_s); si.sh_t = ld3(g.sh_t);
    si.dp_du = ld3(g.dp_du); si.dp_dv = ld3(g.dp_dv); si.uv = v2(g.uv[0], g.uv[1]); si.shape = g.shape;
    return si;
}

// SMS-strategy result of a resolved glint query: (f * (val * inv_prob)) * ew per channel
// (manifold_glints.cpp:108,179)
SMS_DEV V3 glint_result(float val, float inv_prob, float f, V3 ew) {
    float sw = val * inv_prob;
    return v3((f * sw) * ew.x, (f * sw) * ew.y, (f * sw) * ew.z);
}

__global__ void k_ss_walk(DScene S, const sms_ss_walk_in *in, int64_t n, sms_ss_walk_out *out);
__global__ void k_glint_walk(DScene S, const sms_glint_walk_in *in, int64_t n, sms_glint_walk_out *out);
__global__ void k_glint_queries(DScene S, int spp, uint64_t seed, sms_glint_query *q, int64_t n);
__global__ void k_glint_setup(DScene S, const sms_glint_query *Q, int64_t n, uint64_t seed, float *out, EstimatorWork W);
__global__ void k_glint_walks(DScene S, uint64_t seed, EstimatorWork W);
__global__ void k_glint_coop(DScene S, uint64_t seed, EstimatorWork W);
__global__ void k_glint_finalize(DScene S, const sms_glint_query *Q, uint64_t seed, float *out, EstimatorWork W);
__global__ void k_ss_estimate(DScene S, const sms_ss_query *Q, int64_t n, uint64_t seed, float *part, float *invp, EstimatorWork W);
__global__ void k_ss_coop(DScene S, const sms_ss_query *Q, uint64_t seed, float *invp, EstimatorWork W);
__global__ void k_ss_finalize(const float *part, const float *invp, int nc, int64_t n, float *out);
__global__ void k_ss_biased(DScene S, const sms_ss_query *Q, int64_t n, uint64_t seed, float *out, EstimatorWork W);
#define SMS_MS_KERNEL_DECLS(B)                                                                                   \
    __global__ void k_ms_walk_##B(DScene S, const sms_ms_walk_in *in, int64_t n, sms_ms_walk_out *out);         \
    __global__ void k_ms_estimate_##B(DScene S, const sms_ss_query *Q, int64_t n, uint64_t seed, float *part,   \
                                      float *invp, EstimatorWork W);                                             \
    __global__ void k_ms_coop_##B(DScene S, const sms_ss_query *Q, uint64_t seed, float *invp, EstimatorWork W); \
    __global__ void k_ms_biased_##B(DScene S, const sms_ss_query *Q, int64_t n, uint64_t seed, float *out,      \
                                    EstimatorWork W);
SMS_MS_KERNEL_DECLS(2)
SMS_MS_KERNEL_DECLS(4)

} // namespace smsdev
