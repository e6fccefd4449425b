// kernels.h -- kernel declarations shared by the kernel and host translation units.
#pragma once
#include "dev_ss.h"
#include "dev_glint.h"

namespace smsdev {

// A (query, caster) whose Bernoulli trial loop outgrew the per-lane budget.
struct OverflowRec {
    int64_t q;            // query index (glints) or query*n_casters + caster (single scatter)
    int64_t k;            // next trial index to run
    float sol[3];         // solution: uv (glints) or direction (single scatter)
    float val[3];         // glints: F*G of the solution
    float f;              // glints: mis * We
    float ew[3];          // glints: emitter weight
};

// Device work-queue state of one estimator call.
struct EstimatorWork {
    unsigned long long *cursor;       // next query (pass 1)
    unsigned long long *n_overflow;   // overflow list length
    unsigned long long *coop_cursor;  // next overflow record (pass 2)
    unsigned long long *stats;        // sms_stats counters (walks, converged, iterations, calls, trials, max)
    OverflowRec *overflow;
    int lane_trials;                  // trials a lane runs before handing a query to pass 2
};

__global__ void k_ss_walk(DScene S, const sms_ss_walk_in *in, int64_t n, sms_ss_walk_out *out);
__global__ void k_glint_walk(DScene S, const sms_glint_walk_in *in, int64_t n, sms_glint_walk_out *out);
__global__ void k_glint_queries(DScene S, int spp, uint64_t seed, sms_glint_query *q, int64_t n);
__global__ void k_glint_estimate(DScene S, const sms_glint_query *Q, int64_t n, uint64_t seed, float *out, EstimatorWork W);
__global__ void k_glint_coop(DScene S, const sms_glint_query *Q, uint64_t seed, float *out, EstimatorWork W);
__global__ void k_ss_estimate(DScene S, const sms_ss_query *Q, int64_t n, uint64_t seed, float *part, float *invp, EstimatorWork W);
__global__ void k_ss_coop(DScene S, const sms_ss_query *Q, uint64_t seed, float *invp, EstimatorWork W);
__global__ void k_ss_finalize(const float *part, const float *invp, int nc, int64_t n, float *out);
__global__ void k_ss_biased(DScene S, const sms_ss_query *Q, int64_t n, uint64_t seed, float *out, EstimatorWork W);

} // namespace smsdev
