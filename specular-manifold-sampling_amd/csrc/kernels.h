// kernels.h -- kernel declarations and shared device helpers of the kernel and host
// translation units.
#pragma once
#include "dev_ss.h"
#include "dev_glint.h"
#include "dev_lean.h"
#include "dev_proposal.h"
#include "dev_path.h"

namespace smsdev {

// A (query, caster) whose Bernoulli trial loop outgrew the per-lane budget.
struct OverflowRec {
    int64_t q;            // glints: record index; single scatter: query*n_casters + caster
    int64_t k;            // next trial index to run
    float sol[3];         // solution: uv (glints) or direction (single / multi scatter)
    int32_t shapes[4];    // multi scatter: shapes of the first walk's seed chain
    int32_t pad;
};

// Glint footprint of a query that reaches the solve (block-local compaction in LDS), 40 bytes.
struct GlintRec {
    int32_t q;            // query index
    float target[2];      // target slope (manifold_glints.cpp:241-247)
    float uv[2], duv_dx[2], duv_dy[2];
    int32_t shape;
};

// A glint query whose first walk converged: the footprint and the solution its Bernoulli
// trials must find again, 48 bytes.
struct TrialRec {
    GlintRec f;
    float uv_sol[2];
};

// Outcome of a trial item: status 1 = resolved (inv_prob final), 2 = handed to the tail pass.
struct GlintRes {
    float inv_prob;
    int32_t status;
};

// Hot device counters exist in SPREAD copies SPREAD_STRIDE u64 apart (separate 128-byte
// lines): a wavefront updates the copy picked by its block and wave index, so thousands of
// wavefronts finishing together do not serialise on one address.  The host sums the copies.
constexpr int SPREAD = 64;
constexpr int SPREAD_STRIDE = 16;
SMS_DEV int spread_slot() { return (int) (((blockIdx.x << 2) + (threadIdx.x >> 6)) & (SPREAD - 1)); }

// stats counter j of copy 0 is stats[j]; add / max a wavefront's values into its copy
SMS_DEV void stat_add(unsigned long long *dst, unsigned long long v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(dst + spread_slot() * SPREAD_STRIDE, v);
}
SMS_DEV void stat_max(unsigned long long *dst, unsigned long long v) {
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    if ((threadIdx.x & 63) == 0 && v) atomicMax(dst + spread_slot() * SPREAD_STRIDE, v);
}

// A list split into SPREAD sub-lists of sub_cap items, each with its own counter; producers
// append to a sub-list they are assigned, consumers walk the sub-lists by a static mapping
// (no claim counter).  Items are addressed by slot = sub-list * sub_cap + position.
template <class T>
struct SpreadList {
    T *items;
    unsigned long long *counts;       // SPREAD counters, SPREAD_STRIDE apart
    unsigned long long sub_cap;
};
template <class T>
SMS_DEV unsigned long long spread_append(const SpreadList<T> &L, int sub, unsigned long long n) {
    return atomicAdd(L.counts + sub * SPREAD_STRIDE, n);
}
// Static consumer mapping: unit u of n_units (n_units >= SPREAD) walks sub-list u % SPREAD at
// positions u / SPREAD, u / SPREAD + m, ... where m = units on that sub-list.
struct SpreadWalk { int sub; unsigned long long pos, step, count; };
template <class T>
SMS_DEV SpreadWalk spread_walk(const SpreadList<T> &L, unsigned long long u, unsigned long long n_units) {
    SpreadWalk w;
    w.sub = (int) (u % SPREAD);
    w.pos = u / SPREAD;
    w.step = (n_units - (unsigned long long) w.sub + SPREAD - 1) / SPREAD;
    w.count = L.counts[w.sub * SPREAD_STRIDE];
    return w;
}

// A work list of OverflowRec items with its length and claim cursor.
struct WorkList {
    OverflowRec *items;
    unsigned long long *count;
    unsigned long long *cursor;
};

// Device work-queue state of one estimator call.
struct EstimatorWork {
    unsigned long long *cursor;       // next item of the persistent pass
    unsigned long long *n_overflow;   // overflow list length
    unsigned long long *coop_cursor;  // next overflow record (tail pass)
    unsigned long long *stats;        // copy 0 of the spread sms_stats counters (walks, converged, iterations,
                                      // calls, trials, max, tail iterations)
    OverflowRec *overflow;
    GlintRes *res;                    // per trial slot
    SpreadList<TrialRec> trials;      // glints: queries whose first walk converged
    SpreadList<OverflowRec> over;     // glints: trial items left after the grouped trial pass
    int lane_trials;                  // trials a lane (or lane group) runs before handing a query on
    int group;                        // glints: lanes per item in the first trial pass
    int tail_rounds;                  // glints: 64-trial rounds of the one-wavefront tail pass
    // glints, multi-wavefront stage for long trial chains
    WorkList long_list;               // items left after the one-wavefront rounds
    WorkList seq_list;                // items left after the multi-wavefront stage
    long long *mw_found;              // first matching trial per long item (atomic min)
    unsigned long long *mw_next;      // next 64-trial chunk per long item
    uint16_t *mw_log;                 // per trial: iterations | converged << 15
    int mw_items;                     // long items the stage holds (the rest go to seq_list)
    int mw_trials;                    // trials per long item the stage may run
    Proposal prop;                    // single scatter: learned seed proposal (prop.gmm null: none)
};

SMS_DEV void write_rgb(float *out, int64_t q, V3 v) { out[3 * q] = v.x; out[3 * q + 1] = v.y; out[3 * q + 2] = v.z; }

SMS_DEV SI glint_query_si(const sms_glint_query &g) {
    SI si = si_invalid();
    si.valid = true; si.p = ld3(g.p); si.n = ld3(g.n); si.sh_n = ld3(g.sh_n); si.sh_s = ld3(g.sh_s); si.sh_t = ld3(g.sh_t);
    si.dp_du = ld3(g.dp_du); si.dp_dv = ld3(g.dp_dv); si.uv = v2(g.uv[0], g.uv[1]); si.shape = g.shape;
    return si;
}

// Camera query of sample idx of a w x h x spp glint frame (integrator.cpp:243-262,
// perspective.cpp:191-221, SurfaceInteraction::compute_partials interaction.h:216-248): the
// pixel-sample ray differential, its hit and, on a glinty shape, the footprint; shape -1 when the
// first hit is not glinty.  BVH false skips the triangle tree (scenes without meshes).
template <bool BVH>
SMS_DEV sms_glint_query make_glint_query(const DScene &S, uint64_t seed, int spp, int64_t idx) {
    const DCamera &c = S.cam;
    sms_glint_query o;
    float *z = (float *) &o;
    for (int k = 0; k < (int) (sizeof(o) / 4); ++k) z[k] = 0.f;
    o.shape = -1;
    const int64_t pixel = idx / spp;
    const int px = (int) (pixel % c.w), py = (int) (pixel / c.w);
    o.pixel = (int32_t) pixel;
    Pcg32 r;
    stream_seed(r, query_key(seed, (uint64_t) idx), 0xFFFFFFFFull);
    const float sx = pcg32_float(r), sy = pcg32_float(r);
    const V2 pos = v2(((float) px + sx) / (float) c.w, ((float) py + sy) / (float) c.h);
    const V3 near_p = v3((1.f - 2.f * pos.x) * c.tan_half, (1.f - 2.f * pos.y) * c.rect_y, 1.f);
    const V3 dloc = normalize(near_p);
    const V3 L = ld3(c.left), U = ld3(c.up), D = ld3(c.dir), O = ld3(c.o);
    const V3 d = to_world(L, U, D, dloc);
    V3 d_x = to_world(L, U, D, normalize(near_p + ld3(c.dx))), d_y = to_world(L, U, D, normalize(near_p + ld3(c.dy)));
    const float scale = 1.f / sqrtf((float) spp);
    d_x = fmadd3(d_x - d, scale, d); d_y = fmadd3(d_y - d, scale, d);
    const Ray ray = make_ray(O, d, 1e-2f * (1.f / dloc.z), 1e4f * (1.f / dloc.z));
    const SI si = make_si(S, ray, scene_hit_t<BVH>(S, ray, false));
    st3(o.sensor_p, O);
    if (si.valid && (S.shapes[si.shape].flags & SMS_FLAG_GLINTY)) {
        const float dd = dot(si.n, si.p);
        const float t_x = (dd - dot(si.n, O)) / dot(si.n, d_x), t_y = (dd - dot(si.n, O)) / dot(si.n, d_y);
        const V3 dp_dx = fmadd3(d_x, t_x, O) - si.p, dp_dy = fmadd3(d_y, t_y, O) - si.p;
        const float a00 = dot(si.dp_du, si.dp_du), a01 = dot(si.dp_du, si.dp_dv), a11 = dot(si.dp_dv, si.dp_dv);
        float inv_det = 1.f / (a00 * a11 - a01 * a01);
        const float b0x = dot(si.dp_du, dp_dx), b1x = dot(si.dp_dv, dp_dx), b0y = dot(si.dp_du, dp_dy),
                    b1y = dot(si.dp_dv, dp_dy);
        if (!isfinite(inv_det)) inv_det = 0.f;
        o.duv_dx[0] = fmaf(a11, b0x, -(a01 * b1x)) * inv_det; o.duv_dx[1] = fmaf(a00, b1x, -(a01 * b0x)) * inv_det;
        o.duv_dy[0] = fmaf(a11, b0y, -(a01 * b1y)) * inv_det; o.duv_dy[1] = fmaf(a00, b1y, -(a01 * b0y)) * inv_det;
        o.shape = si.shape;
        st3(o.p, si.p); st3(o.n, si.n); st3(o.sh_n, si.sh_n); st3(o.sh_s, si.sh_s); st3(o.sh_t, si.sh_t);
        st3(o.dp_du, si.dp_du); st3(o.dp_dv, si.dp_dv);
        const V3 mw = -ray.d;
        st3(o.wi, v3(dot(mw, si.sh_s), dot(mw, si.sh_t), dot(mw, si.sh_n)));
        o.uv[0] = si.uv.x; o.uv[1] = si.uv.y;
    }
    return o;
}

// Where the glint estimator's queries come from: an array (Q), or the camera -- query q is
// make_glint_query of the frame's sample q, generated where it is read, with the camera seed
// at cam_seed (the estimate of a whole frame without a query buffer).
struct GlintQuerySource { const sms_glint_query *Q; const uint64_t *cam_seed; int spp; };

template <bool BVH = true>
SMS_DEV sms_glint_query glint_query_fetch(const DScene &S, const GlintQuerySource &src, int64_t q) {
    if (src.Q) return src.Q[q];
    return make_glint_query<BVH>(S, *src.cam_seed, src.spp, q);
}

// SMS-strategy result of a resolved glint query: (f * (val * inv_prob)) * ew per channel
// (manifold_glints.cpp:108,179)
SMS_DEV V3 glint_result(float val, float inv_prob, float f, V3 ew) {
    float sw = val * inv_prob;
    return v3((f * sw) * ew.x, (f * sw) * ew.y, (f * sw) * ew.z);
}

__global__ void k_ss_walk(DScene S, const sms_ss_walk_in *in, int64_t n, sms_ss_walk_out *out);
__global__ void k_glint_walk(DScene S, const sms_glint_walk_in *in, int64_t n, sms_glint_walk_out *out);
__global__ void k_glint_queries(DScene S, int spp, const uint64_t *seedp, sms_glint_query *q, int64_t n);
__global__ void k_glint_first_bvh(DScene S, GlintQuerySource Q, int64_t n, const uint64_t *seedp, float *out, EstimatorWork W);
__global__ void k_glint_biased_bvh(DScene S, GlintQuerySource Q, int64_t n, const uint64_t *seedp, float *out, EstimatorWork W);
__global__ void k_glint_biased_nobvh(DScene S, GlintQuerySource Q, int64_t n, const uint64_t *seedp, float *out, EstimatorWork W);
__global__ void k_glint_bsdf(DScene S, GlintQuerySource Q, int64_t n, const uint64_t *seedp, float *out);
__global__ void k_glint_first_nobvh(DScene S, GlintQuerySource Q, int64_t n, const uint64_t *seedp, float *out, EstimatorWork W);
__global__ void k_glint_trials_g1(DScene S, const uint64_t *seedp, EstimatorWork W);
__global__ void k_glint_trials_g2(DScene S, const uint64_t *seedp, EstimatorWork W);
__global__ void k_glint_trials_g4(DScene S, const uint64_t *seedp, EstimatorWork W);
__global__ void k_glint_trials_g8(DScene S, const uint64_t *seedp, EstimatorWork W);
__global__ void k_glint_trials_g16(DScene S, const uint64_t *seedp, EstimatorWork W);
__global__ void k_glint_coop_rounds(DScene S, const uint64_t *seedp, EstimatorWork W);
__global__ void k_glint_multiwave(DScene S, const uint64_t *seedp, EstimatorWork W);
__global__ void k_glint_multiwave_resolve(DScene S, EstimatorWork W);
__global__ void k_glint_coop_seq(DScene S, const uint64_t *seedp, EstimatorWork W);
__global__ void k_glint_finalize(DScene S, GlintQuerySource Q, const uint64_t *seedp, float *out, EstimatorWork W);
__global__ void k_ss_estimate(DScene S, const sms_ss_query *Q, int64_t n, const uint64_t *seedp, float *part, float *invp, EstimatorWork W);
// LDS traversal-stack entries per lane of the persistent walker; scenes whose BVH is deeper run
// the lane-per-query first pass instead
constexpr int WALKER_STACK = 32;
__global__ void k_ss_walker(DScene S, const sms_ss_query *Q, int64_t n, const uint64_t *seedp, float *part, float *invp,
                            EstimatorWork W);
__global__ void k_ss_coop(DScene S, const sms_ss_query *Q, const uint64_t *seedp, const float *part, float *invp, EstimatorWork W);
__global__ void k_glint_prop_inputs(DScene S, GlintQuerySource Q, int64_t n, const uint64_t *seedp, float *x);
__global__ void k_ss_finalize(DScene S, const uint64_t *seedp, const float *part, const float *invp, int nc, int64_t n,
                              int offsets, float *out);
__global__ void k_ss_queries(DScene S, int spp, const uint64_t *seedp, sms_ss_query *q, float *w, int64_t n);
__global__ void k_film_accumulate(const float *rgb, const float *w, int64_t pixels, int spp, float *image, float *image_sq,
                                  float *image_n);
__global__ void k_prop_inputs(DScene S, const sms_ss_query *Q, int64_t n, const uint64_t *seedp, int ms, float *x);
__global__ void k_prop_anchor(DScene S, const sms_ss_query *Q, int64_t n, const uint64_t *seedp, int shape, float *anchor);
__global__ void k_gmm_prepare(const float *raw, int64_t n, float *g);
__global__ void k_capture_keys(const float *cap, const unsigned long long *counts, unsigned long long cap_sub, float *rec,
                               uint64_t *keys, uint32_t *idx);
__global__ void k_capture_scatter(const float *rec, const uint32_t *idx, int64_t n, float *x, float *uv, float *w);
__global__ void k_capture_gather(const float *cap, const unsigned long long *counts, unsigned long long cap_sub,
                                 float *x, float *uv, float *w);
__global__ void k_mlp_infer(const float *p, const uint16_t *wb, const float *x9, int64_t n, float *out);
__global__ void k_mlp_grad(const float *p, const uint16_t *wb, const uint16_t *wt, const float *x9, const float *uv,
                           const float *w, int64_t n, float *partial, float *aux);
__global__ void k_mlp_reduce(const float *partial, const float *aux, int nb, float *g, float *loss);
__global__ void k_mlp_reduce_adam(const float *partial, const float *aux, int nb, float *p, float *m, float *v,
                                  uint16_t *wb, uint16_t *wt, float lr, float b1, float b2, float eps, float bc1, float bc2,
                                  float *loss);
__global__ void k_mlp_adam(float *p, float *g, float *m, float *v, uint16_t *wb, uint16_t *wt, int n, float lr, float b1,
                           float b2, float eps, float bc1, float bc2);
__global__ void k_ss_biased(DScene S, const sms_ss_query *Q, int64_t n, const uint64_t *seedp, float *out, EstimatorWork W);
#define SMS_MS_KERNEL_DECLS(B)                                                                                        \
    __global__ void k_ms_walk_##B(DScene S, const sms_ms_walk_in *in, int64_t n, sms_ms_walk_out *out);               \
    __global__ void k_ms_estimate_##B(DScene S, const sms_ss_query *Q, int64_t n, const uint64_t *seedp, float *part, \
                                      float *invp, EstimatorWork W);                                                  \
    __global__ void k_ms_persist_##B(DScene S, const sms_ss_query *Q, int64_t n, const uint64_t *seedp, float *part, \
                                     float *invp, EstimatorWork W);                                                  \
    __global__ void k_ms_coop_##B(DScene S, const sms_ss_query *Q, const uint64_t *seedp, const float *part,          \
                                  float *invp, EstimatorWork W);                                                      \
    __global__ void k_ms_biased_##B(DScene S, const sms_ss_query *Q, int64_t n, const uint64_t *seedp, float *out,    \
                                    EstimatorWork W);
SMS_MS_KERNEL_DECLS(2)
SMS_MS_KERNEL_DECLS(4)

// path integrators and film reconstruction (sms_path_kernels.hip)
__global__ void k_path_trace(DScene S, sms_path_config pc, uint64_t seed, int64_t begin, int64_t n, int spp, PathSlots P);
__global__ void k_path_compact(PathSlots P, int64_t n, const int *q_off, sms_ss_query *dq, const int *m_off,
                               sms_ss_walk_in *dm, sms_ms_walk_in *dmm);
__global__ void k_path_gather(PathSlots P, int64_t n, const int *q_off, const float *est, const int *m_off,
                              const sms_ss_walk_out *mout, const sms_ms_walk_out *mmout, float uniqueness_threshold);
__global__ void k_film_splat(FilmFilter F, int w, int h, int spp, const float *rgb, const float *pos, float *image,
                             float *weight);

} // namespace smsdev
