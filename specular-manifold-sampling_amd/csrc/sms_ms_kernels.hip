// sms_ms_kernels.hip -- multi-scatter SMS (SpecularManifoldMultiScatter,
// src/librender/manifold_ms.cpp) on gfx950.
//
//   k_ms_walk_<B>      parity entry point: one lane per walk record
//   k_ms_estimate_<B>  unbiased estimator, lane per query: for every multi-scatter caster the
//                      first walk (which fixes the chain's shape sequence), its contribution and
//                      up to `lane_trials` Bernoulli trials; longer loops go to the overflow list
//   k_ms_coop_<B>      overflow pass: one wavefront per (query, caster), 64 trials per round,
//                      the first match found by a ballot
//   k_ms_biased_<B>    biased estimator: one wavefront per query, lanes = trial-set members;
//                      the earliest member whose seed chain is valid fixes the shape sequence
//                      of all later members (manifold_ms.cpp:150-163), solutions de-duplicated
//                      in member order
//
// <B> is the largest chain the kernel holds (2 or 4 specular vertices); the host picks the
// smallest that fits `bounces`.  Walk k of caster c at query q draws from PCG32 stream
// (q, (c << 32) + k + 1); the emitter sample comes from stream (q, 0).  With a learned
// proposal on caster c, the seed point of every unbiased-estimator walk on it (first walk and
// trials) comes from the defensive mixture, exactly as for single scatter; the biased
// estimator keeps the reference's uniform seeds.
#include "kernels.h"
#include "dev_ms.h"
#include "dev_ss_est.h"

namespace smsdev {

SMS_DEV EI ei_from_abi(const sms_emitter_interaction &e) {
    EI ei;
    ei.p = ld3(e.p); ei.n = ld3(e.n); ei.d = ld3(e.d); ei.weight = ld3(e.weight);
    ei.pdf = e.pdf; ei.kind = e.kind; ei.emitter = e.emitter;
    return ei;
}

SMS_DEV SI ms_query_si(const sms_ss_query &g) {
    SI si = si_invalid();
    si.valid = true; si.t = 1.f;
    si.p = ld3(g.p); si.n = ld3(g.n); si.sh_n = ld3(g.sh_n); si.sh_s = ld3(g.sh_s); si.sh_t = ld3(g.sh_t);
    si.dp_du = si.sh_s; si.dp_dv = si.sh_t; si.shape = g.shape;
    return si;
}

// Diffuse::eval at the receiver (src/bsdfs/diffuse.cpp:104-120)
SMS_DEV V3 ms_receiver_bsdf(const DScene &S, const SI &si, V3 wi_local, V3 wo_local) {
    const DShape &sh = S.shapes[si.shape];
    if (sh.bsdf != SMS_BSDF_DIFFUSE) return v3(0.f, 0.f, 0.f);
    if (!(wi_local.z > 0.f && wo_local.z > 0.f)) return v3(0.f, 0.f, 0.f);
    return (ld3(sh.refl) * INV_PI_F) * wo_local.z;
}

// Walk k of caster c at query q.  g: the query's prepared mixture when the proposal seeds this
// caster, else null (uniform seed); a mixture seed outside the caster's UV domain is a failed
// walk with 0 iterations.
template <int MAXB>
SMS_DEV void ms_walk_k(const DScene &S, const Proposal &P, const float *g, uint64_t qkey, int c, int shape, uint64_t k,
                       V3 si_p, const EI &ei, bool first, const int *shapes, MSWalk<MAXB> &w) {
    Pcg32 rw;
    stream_seed(rw, qkey, ((uint64_t) c << 32) + k + 1);
    V2 s = v2(0.f, 0.f);
    V3 p_start = v3(0.f, 0.f, 0.f);
    if (g) {
        if (!proposal_seed(S, P, g, nullptr, rw, s, p_start)) {
            w.ok = false; w.seed_ok = false; w.iterations = 0; w.path.n = 0;
            for (int i = 0; i < MAXB; ++i) w.seed_shapes[i] = shapes ? shapes[i] : -1;
            return;
        }
    } else {
        s.x = pcg32_float(rw); s.y = pcg32_float(rw);
    }
    ms_sample_path<MAXB>(S, shape, si_p, ei, s, first, shapes, p_start, w);
}

// mixture of query q when the proposal seeds `shape`, else null
SMS_DEV const float *ms_mixture(const Proposal &P, int64_t q, int shape) {
    return P.gmm && shape == P.shape ? P.gmm + q * PROP_RAW : nullptr;
}

// uv of the chain's vertex on `shape` (the training target of the proposal): the first hit of
// the chain's segment into that vertex, traced as the reprojection traces it
template <int MAXB>
SMS_DEV bool ms_caster_uv(const DScene &S, V3 si_p, const MSPath<MAXB> &path, int shape, V2 &uv) {
    for (int i = 0; i < path.n && i < MAXB; ++i) {
        if (path.shapes[i] != shape) continue;
        const V3 o = i == 0 ? si_p : path.v[i > 0 ? i - 1 : 0].v.p;
        const V3 d = normalize(path.v[i].v.p - o);
        const SI h = ray_intersect(S, i == 0 ? make_ray(o, d) : spawn_ray(o, d));
        if (!h.valid || h.shape != shape) return false;
        uv = h.uv;
        return true;
    }
    return false;
}

template <int MAXB>
SMS_DEV void ms_capture(const DScene &S, const Proposal &P, const SI &si, const EI &ei, const MSPath<MAXB> &path,
                        int shape, float w) {
    V2 uv;
    if (ms_caster_uv<MAXB>(S, si.p, path, shape, uv)) capture_sample(S, P, si, ei, uv, w);
}

// ================================================================ parity walks
template <int MAXB>
SMS_DEV void ms_walk_body(const DScene &S, const sms_ms_walk_in *__restrict__ in, int64_t n, sms_ms_walk_out *__restrict__ out) {
    int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const sms_ms_walk_in w = in[i];
    EI ei = ei_from_abi(w.ei);
    MSWalk<MAXB> r;
    ms_sample_path<MAXB>(S, w.shape, ld3(w.si_p), ei, v2(w.sample[0], w.sample[1]), w.first_path != 0, w.seed_shapes,
                         ld3(w.p_start), r);
    sms_ms_walk_out o;
    o.success = r.ok; o.iterations = r.iterations; o.seed_valid = r.seed_ok; o.n_vertices = r.path.n;
    for (int k = 0; k < SMS_MS_MAX_BOUNCES; ++k) {
        int sk = k < MAXB ? r.seed_shapes[k < MAXB ? k : 0] : w.seed_shapes[k];
        o.shapes[k] = r.seed_ok ? sk : -1;
        V3 p = v3(0.f, 0.f, 0.f);
        if (k < r.path.n) p = r.path.v[k < MAXB ? k : 0].v.p;
        st3(o.p[k], p);
    }
    out[i] = o;
}

// ================================================================ unbiased estimator
template <int MAXB>
SMS_DEV void ms_estimate_body(const DScene &S, const sms_ss_query *__restrict__ Q, int64_t n, uint64_t seed,
                              float *__restrict__ part, float *__restrict__ invp, EstimatorWork W) {
    const sms_config &cfg = S.cfg;
    unsigned long long st_walks = 0, st_conv = 0, st_iters = 0, st_bcalls = 0, st_btrials = 0, st_bmax = 0;
    const int nc = S.n_casters_ms;
    // one lane per query; the launch covers n with one block per 256 queries, so a block's
    // overflow items fit its sub-list of the spread overflow list
    const int sub = (int) (blockIdx.x % SPREAD);
    for (int64_t q = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t) gridDim.x * blockDim.x) {
        const sms_ss_query g = Q[q];
        SI si = ms_query_si(g);
        V3 wi_local = ld3(g.wi);
        uint64_t qkey = query_key(seed, (uint64_t) q);
        Pcg32 rq;
        stream_seed(rq, qkey, 0);
        EI ei = sample_emitter_interaction(S, si.p, S.emitters_ms, S.n_emitters_ms, rq);
        for (int c = 0; c < nc; ++c) {
            int64_t slot = q * nc + c;
            part[3 * slot] = 0.f; part[3 * slot + 1] = 0.f; part[3 * slot + 2] = 0.f;
            invp[slot] = 0.f;
            if (g.shape < 0) continue;                 // no receiver: zero estimate, no walks
            const int shape = S.casters_ms[c];
            const float *gq = ms_mixture(W.prop, q, shape);
            MSWalk<MAXB> w0;
            ms_walk_k<MAXB>(S, W.prop, gq, qkey, c, shape, 0, si.p, ei, true, nullptr, w0);
            st_walks++; st_iters += (unsigned long long) w0.iterations;
            if (!w0.ok) continue;
            st_conv++;
            if (capturing(W.prop, shape, SMS_CAPTURE_SOLUTIONS)) ms_capture<MAXB>(S, W.prop, si, ei, w0.path, shape, 1.f);
            V3 direction = normalize(w0.path.v[0].v.p - si.p);
            V3 spec = ms_path_contribution<MAXB>(S, si, ei, w0.path);
            V3 dl = v3(dot(direction, si.sh_s), dot(direction, si.sh_t), dot(direction, si.sh_n));
            V3 bsdf = ms_receiver_bsdf(S, si, wi_local, dl);
            part[3 * slot] = bsdf.x * spec.x; part[3 * slot + 1] = bsdf.y * spec.y; part[3 * slot + 2] = bsdf.z * spec.z;
            st_bcalls++;
            int64_t k = 1;
            float inv_prob = -1.f;
            while (true) {
                MSWalk<MAXB> t;
                ms_walk_k<MAXB>(S, W.prop, gq, qkey, c, shape, (uint64_t) k, si.p, ei, false, w0.seed_shapes, t);
                st_walks++; st_iters += (unsigned long long) t.iterations;
                if (t.ok) {
                    st_conv++;
                    if (capturing(W.prop, shape, SMS_CAPTURE_SOLUTIONS))
                        ms_capture<MAXB>(S, W.prop, si, ei, t.path, shape, 1.f);
                    V3 dtr = normalize(t.path.v[0].v.p - si.p);
                    if (fabsf(dot(direction, dtr) - 1.f) < cfg.uniqueness_threshold) { inv_prob = (float) k; break; }
                }
                k++;
                if (cfg.max_trials > 0 && k > cfg.max_trials) { inv_prob = 0.f; break; }
                if (k > W.lane_trials) break;
            }
            if (inv_prob >= 0.f) {
                unsigned long long tr = (unsigned long long) k;
                st_btrials += tr; st_bmax = st_bmax > tr ? st_bmax : tr;
                invp[slot] = inv_prob;
                const float wgt = mean3(&part[3 * slot]) * inv_prob;
                if (wgt > 0.f && capturing(W.prop, shape, SMS_CAPTURE_ESTIMATES))
                    ms_capture<MAXB>(S, W.prop, si, ei, w0.path, shape, wgt);
            } else {
                unsigned long long o = spread_append(W.over, sub, 1ull);
                OverflowRec rec;
                rec.q = slot; rec.k = k;
                rec.sol[0] = direction.x; rec.sol[1] = direction.y; rec.sol[2] = direction.z;
                for (int b = 0; b < 4; ++b) rec.shapes[b] = b < MAXB ? w0.seed_shapes[b < MAXB ? b : 0] : -1;
                rec.pad = 0;
                W.over.items[(size_t) sub * W.over.sub_cap + o] = rec;
            }
        }
    }
    stat_add(W.stats + 0, st_walks); stat_add(W.stats + 1, st_conv); stat_add(W.stats + 2, st_iters);
    stat_add(W.stats + 3, st_bcalls); stat_add(W.stats + 4, st_btrials); stat_max(W.stats + 5, st_bmax);
}

// Persistent first pass: resident lanes take (query, caster) items i = q * nc + c from the
// device cursor and run one walk per loop trip -- the item's first walk, then its Bernoulli
// trials -- from one walk call site, so a lane whose chain ends takes the next item at once
// instead of idling until its wavefront's slowest chain ends.  Streams, results, statistics
// and overflow items equal ms_estimate_body's: item i goes to the overflow sub-list of
// (q / 256) % SPREAD, the sub-list of the lane pass's block for q, so the capacities hold.
template <int MAXB>
SMS_DEV void ms_persistent_body(const DScene &S, const sms_ss_query *__restrict__ Q, int64_t n, uint64_t seed,
                                float *__restrict__ part, float *__restrict__ invp, EstimatorWork W) {
    const sms_config &cfg = S.cfg;
    unsigned long long st_walks = 0, st_conv = 0, st_iters = 0, st_bcalls = 0, st_btrials = 0, st_bmax = 0;
    const int nc = S.n_casters_ms;
    const unsigned long long n_items = (unsigned long long) n * (unsigned long long) nc;
    int64_t item = -1, q = 0, k = 0;
    int c = 0, shape = -1;
    SI si;
    EI ei;
    V3 wi_local = v3(0.f, 0.f, 0.f), direction = v3(0.f, 0.f, 0.f);
    uint64_t qkey = 0;
    const float *gq = nullptr;
    int seed_shapes[MAXB];
    for (int b = 0; b < MAXB; ++b) seed_shapes[b] = -1;
    V2 uv0 = v2(0.f, 0.f);
    bool have_uv0 = false;
    while (true) {
        if (item < 0) {
            const unsigned long long it = atomicAdd(W.cursor, 1ull);
            if (it >= n_items) break;
            item = (int64_t) it;
            q = item / nc;
            c = (int) (item % nc);
            part[3 * item] = 0.f; part[3 * item + 1] = 0.f; part[3 * item + 2] = 0.f;
            invp[item] = 0.f;
            const sms_ss_query g = Q[q];
            if (g.shape < 0) { item = -1; continue; }     // no receiver: zero estimate, no walks
            si = ms_query_si(g);
            wi_local = ld3(g.wi);
            qkey = query_key(seed, (uint64_t) q);
            Pcg32 rq;
            stream_seed(rq, qkey, 0);
            ei = sample_emitter_interaction(S, si.p, S.emitters_ms, S.n_emitters_ms, rq);
            shape = S.casters_ms[c];
            gq = ms_mixture(W.prop, q, shape);
            k = 0;
        }
        MSWalk<MAXB> t;
        ms_walk_k<MAXB>(S, W.prop, gq, qkey, c, shape, (uint64_t) k, si.p, ei, k == 0, k == 0 ? nullptr : seed_shapes, t);
        st_walks++; st_iters += (unsigned long long) t.iterations;
        if (k == 0) {
            if (!t.ok) { item = -1; continue; }
            st_conv++;
            if (capturing(W.prop, shape, SMS_CAPTURE_SOLUTIONS)) ms_capture<MAXB>(S, W.prop, si, ei, t.path, shape, 1.f);
            have_uv0 = capturing(W.prop, shape, SMS_CAPTURE_ESTIMATES) && ms_caster_uv<MAXB>(S, si.p, t.path, shape, uv0);
            direction = normalize(t.path.v[0].v.p - si.p);
            V3 spec = ms_path_contribution<MAXB>(S, si, ei, t.path);
            V3 dl = v3(dot(direction, si.sh_s), dot(direction, si.sh_t), dot(direction, si.sh_n));
            V3 bsdf = ms_receiver_bsdf(S, si, wi_local, dl);
            part[3 * item] = bsdf.x * spec.x; part[3 * item + 1] = bsdf.y * spec.y; part[3 * item + 2] = bsdf.z * spec.z;
            st_bcalls++;
            for (int b = 0; b < MAXB; ++b) seed_shapes[b] = t.seed_shapes[b];
            k = 1;
            continue;
        }
        bool match = false;
        if (t.ok) {
            st_conv++;
            if (capturing(W.prop, shape, SMS_CAPTURE_SOLUTIONS)) ms_capture<MAXB>(S, W.prop, si, ei, t.path, shape, 1.f);
            V3 dtr = normalize(t.path.v[0].v.p - si.p);
            match = fabsf(dot(direction, dtr) - 1.f) < cfg.uniqueness_threshold;
        }
        float inv_prob;
        if (match) {
            inv_prob = (float) k;
        } else {
            k++;
            if (cfg.max_trials > 0 && k > cfg.max_trials) {
                inv_prob = 0.f;
            } else if (k > W.lane_trials) {
                const int sub = (int) ((q >> 8) % SPREAD);
                unsigned long long o = spread_append(W.over, sub, 1ull);
                OverflowRec rec;
                rec.q = item; rec.k = k;
                rec.sol[0] = direction.x; rec.sol[1] = direction.y; rec.sol[2] = direction.z;
                for (int b = 0; b < 4; ++b) rec.shapes[b] = b < MAXB ? seed_shapes[b < MAXB ? b : 0] : -1;
                rec.pad = 0;
                W.over.items[(size_t) sub * W.over.sub_cap + o] = rec;
                item = -1;
                continue;
            } else {
                continue;
            }
        }
        const unsigned long long tr = (unsigned long long) k;
        st_btrials += tr; st_bmax = st_bmax > tr ? st_bmax : tr;
        invp[item] = inv_prob;
        const float wgt = mean3(&part[3 * item]) * inv_prob;
        if (wgt > 0.f && have_uv0) capture_sample(S, W.prop, si, ei, uv0, wgt);
        item = -1;
    }
    stat_add(W.stats + 0, st_walks); stat_add(W.stats + 1, st_conv); stat_add(W.stats + 2, st_iters);
    stat_add(W.stats + 3, st_bcalls); stat_add(W.stats + 4, st_btrials); stat_max(W.stats + 5, st_bmax);
}

template <int MAXB>
SMS_DEV void ms_coop_body(const DScene &S, const sms_ss_query *__restrict__ Q, uint64_t seed,
                          const float *__restrict__ part, float *__restrict__ invp, EstimatorWork W) {
    const sms_config &cfg = S.cfg;
    const int lane = threadIdx.x & 63;
    const int nc = S.n_casters_ms;
    unsigned long long st_walks = 0, st_conv = 0, st_iters = 0, st_btrials = 0, st_bmax = 0, st_tail = 0;
    const unsigned long long wave = ((unsigned long long) blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const SpreadWalk sw = spread_walk(W.over, wave, ((unsigned long long) gridDim.x * blockDim.x) >> 6);
    for (unsigned long long p = sw.pos; p < sw.count; p += sw.step) {
        const OverflowRec rec = W.over.items[(size_t) sw.sub * W.over.sub_cap + p];
        int64_t q = rec.q / nc;
        int c = (int) (rec.q % nc);
        const sms_ss_query g = Q[q];
        SI si = ms_query_si(g);
        uint64_t qkey = query_key(seed, (uint64_t) q);
        Pcg32 rq;
        stream_seed(rq, qkey, 0);
        EI ei = sample_emitter_interaction(S, si.p, S.emitters_ms, S.n_emitters_ms, rq);
        V3 direction = v3(rec.sol[0], rec.sol[1], rec.sol[2]);
        int shapes[4] = {rec.shapes[0], rec.shapes[1], rec.shapes[2], rec.shapes[3]};
        const int shape = S.casters_ms[c];
        const float *gq = ms_mixture(W.prop, q, shape);
        int64_t kcap = cfg.max_trials > 0 ? (int64_t) cfg.max_trials : (int64_t) 1 << 40;
        int64_t found = -1;
        for (int64_t base = rec.k; base <= kcap && found < 0; base += 64) {
            int64_t k = base + lane;
            bool match = false, ran = false, conv = false;
            int its = 0;
            MSWalk<MAXB> t;
            t.ok = false;
            if (k <= kcap) {
                ms_walk_k<MAXB>(S, W.prop, gq, qkey, c, shape, (uint64_t) k, si.p, ei, false, shapes, t);
                ran = true; its = t.iterations; conv = t.ok;
                if (t.ok) {
                    V3 dtr = normalize(t.path.v[0].v.p - si.p);
                    match = fabsf(dot(direction, dtr) - 1.f) < cfg.uniqueness_threshold;
                }
            }
            unsigned long long ballot = __ballot(match);
            if (ballot) found = base + (int64_t) __ffsll((long long) ballot) - 1;
            // the trials the sequential loop runs: up to the first match
            if (ran && (found < 0 || k <= found)) {
                st_walks++; st_iters += (unsigned long long) its; st_tail += (unsigned long long) its;
                if (conv) {
                    st_conv++;
                    if (capturing(W.prop, shape, SMS_CAPTURE_SOLUTIONS))
                        ms_capture<MAXB>(S, W.prop, si, ei, t.path, shape, 1.f);
                }
            }
        }
        if (lane == 0) {
            invp[rec.q] = found >= 0 ? (float) found : 0.f;
            unsigned long long tr = found >= 0 ? (unsigned long long) found : (unsigned long long) kcap + 1ull;
            st_btrials += tr; st_bmax = st_bmax > tr ? st_bmax : tr;
            const float wgt = found > 0 ? mean3(&part[3 * rec.q]) * (float) found : 0.f;
            if (wgt > 0.f && capturing(W.prop, shape, SMS_CAPTURE_ESTIMATES)) {
                // the first walk again (same stream), for the uv of its caster vertex
                MSWalk<MAXB> w0;
                ms_walk_k<MAXB>(S, W.prop, gq, qkey, c, shape, 0, si.p, ei, true, nullptr, w0);
                if (w0.ok) ms_capture<MAXB>(S, W.prop, si, ei, w0.path, shape, wgt);
            }
        }
    }
    stat_add(W.stats + 0, st_walks); stat_add(W.stats + 1, st_conv); stat_add(W.stats + 2, st_iters);
    stat_add(W.stats + 4, st_btrials); stat_max(W.stats + 5, st_bmax); stat_add(W.stats + 6, st_tail);
}

// ================================================================ biased estimator
constexpr int MS_BIASED_MAX_SOL = 128;

template <int MAXB>
SMS_DEV void ms_biased_body(const DScene &S, const sms_ss_query *__restrict__ Q, int64_t n, uint64_t seed,
                            float *__restrict__ out, EstimatorWork W, float (*sol_lds)[3]) {
    const sms_config &cfg = S.cfg;
    const int lane = threadIdx.x & 63;
    const int nc = S.n_casters_ms;
    unsigned long long st_walks = 0, st_conv = 0, st_iters = 0;
    const int64_t n_waves = ((int64_t) gridDim.x * blockDim.x) >> 6;
    for (int64_t q = ((int64_t) blockIdx.x * blockDim.x + threadIdx.x) >> 6; q < n; q += n_waves) {
        const sms_ss_query g = Q[q];
        SI si = ms_query_si(g);
        V3 wi_local = ld3(g.wi);
        uint64_t qkey = query_key(seed, (uint64_t) q);
        Pcg32 rq;
        stream_seed(rq, qkey, 0);
        EI ei = sample_emitter_interaction(S, si.p, S.emitters_ms, S.n_emitters_ms, rq);
        V3 result = v3(0.f, 0.f, 0.f);
        const int ncq = g.shape < 0 ? 0 : nc;        // no receiver: zero estimate, no walks
        for (int c = 0; c < ncq; ++c) {
            const int shape = S.casters_ms[c];
            V3 value = v3(0.f, 0.f, 0.f);
            int n_sol = 0;
            bool have_seed = false;
            int seed_shapes[MAXB];
            for (int b = 0; b < MAXB; ++b) seed_shapes[b] = -1;
            for (int base = 0; base < cfg.max_trials; base += 64) {
                const int m = base + lane;
                const bool active = m < cfg.max_trials;
                // members before the first valid seed chain run as first paths (and fail);
                // the first valid one fixes the shape sequence of all later members
                int j0 = -1;
                if (!have_seed) {
                    bool sok = false;
                    MSPath<MAXB> P;
                    P.n = 0;
                    if (active) {
                        Pcg32 rw;
                        stream_seed(rw, qkey, ((uint64_t) c << 32) + (uint64_t) m + 1);
                        V2 s; s.x = pcg32_float(rw); s.y = pcg32_float(rw);
                        sok = ms_seed_path<MAXB>(S, shape, si.p, ei, s, true, seed_shapes, v3(0.f, 0.f, 0.f), P);
                    }
                    unsigned long long sb = __ballot(sok);
                    if (sb) {
                        j0 = __ffsll((long long) sb) - 1;
                        for (int b = 0; b < MAXB; ++b) {
                            int sh = (sok && b < P.n) ? P.shapes[b] : -1;
                            seed_shapes[b] = __shfl(sh, j0, 64);
                        }
                    }
                }
                const bool first = !have_seed && (j0 < 0 || lane <= j0);
                MSWalk<MAXB> r;
                r.ok = false; r.path.n = 0;
                bool cand = false;
                V3 direction = v3(0.f, 0.f, 0.f);
                if (active) {
                    ms_walk_k<MAXB>(S, W.prop, nullptr, qkey, c, shape, (uint64_t) m, si.p, ei, first,
                                    first ? nullptr : seed_shapes, r);
                    st_walks++; st_iters += (unsigned long long) r.iterations;
                    if (r.ok) {
                        st_conv++;
                        cand = true;
                        direction = normalize(r.path.v[0].v.p - si.p);
                    }
                }
                if (j0 >= 0) have_seed = true;
                bool unique = false;
                unsigned long long cmask = __ballot(cand);
                while (cmask) {
                    int j = __ffsll((long long) cmask) - 1;
                    cmask &= cmask - 1ull;
                    float dx = __shfl(direction.x, j, 64), dy = __shfl(direction.y, j, 64), dz = __shfl(direction.z, j, 64);
                    V3 dj = v3(dx, dy, dz);
                    bool dup = false;
                    for (int s0 = 0; s0 < n_sol; s0 += 64) {
                        int s = s0 + lane;
                        bool d = false;
                        if (s < n_sol) d = fabsf(dot(dj, v3(sol_lds[s][0], sol_lds[s][1], sol_lds[s][2])) - 1.f)
                                           < cfg.uniqueness_threshold;
                        if (__ballot(d)) { dup = true; break; }
                    }
                    if (!dup) {
                        if (n_sol < MS_BIASED_MAX_SOL && lane == 0) {
                            sol_lds[n_sol][0] = dx; sol_lds[n_sol][1] = dy; sol_lds[n_sol][2] = dz;
                        }
                        if (n_sol < MS_BIASED_MAX_SOL) n_sol++;
                        if (lane == j) unique = true;
                    }
                    __builtin_amdgcn_wave_barrier();
                }
                V3 contrib = v3(0.f, 0.f, 0.f);
                if (unique) {
                    V3 dl = v3(dot(direction, si.sh_s), dot(direction, si.sh_t), dot(direction, si.sh_n));
                    V3 bsdf = ms_receiver_bsdf(S, si, wi_local, dl);
                    V3 spec = ms_path_contribution<MAXB>(S, si, ei, r.path);
                    contrib = v3(bsdf.x * spec.x, bsdf.y * spec.y, bsdf.z * spec.z);
                }
                unsigned long long umask = __ballot(unique);
                while (umask) {
                    int j = __ffsll((long long) umask) - 1;
                    umask &= umask - 1ull;
                    value = value + v3(__shfl(contrib.x, j, 64), __shfl(contrib.y, j, 64), __shfl(contrib.z, j, 64));
                }
            }
            result = result + value / 1.f;
        }
        if (lane == 0) write_rgb(out, q, result);
    }
    stat_add(W.stats + 0, st_walks); stat_add(W.stats + 1, st_conv); stat_add(W.stats + 2, st_iters);
}

#define SMS_MS_KERNELS(B)                                                                                       \
    __global__ void __launch_bounds__(256) k_ms_walk_##B(DScene S, const sms_ms_walk_in *in, int64_t n,         \
                                                         sms_ms_walk_out *out) {                                \
        ms_walk_body<B>(S, in, n, out);                                                                         \
    }                                                                                                           \
    __global__ void __launch_bounds__(256) k_ms_estimate_##B(DScene S, const sms_ss_query *Q, int64_t n,        \
                                                             const uint64_t *__restrict__ seedp, float *part,   \
                                                             float *invp, EstimatorWork W) {                    \
        ms_estimate_body<B>(S, Q, n, *seedp, part, invp, W);                                                    \
    }                                                                                                           \
    __global__ void __launch_bounds__(256) k_ms_persist_##B(DScene S, const sms_ss_query *Q, int64_t n,         \
                                                            const uint64_t *__restrict__ seedp, float *part,    \
                                                            float *invp, EstimatorWork W) {                     \
        ms_persistent_body<B>(S, Q, n, *seedp, part, invp, W);                                                  \
    }                                                                                                           \
    __global__ void __launch_bounds__(256) k_ms_coop_##B(DScene S, const sms_ss_query *Q,                       \
                                                         const uint64_t *__restrict__ seedp, const float *part, \
                                                         float *invp, EstimatorWork W) {                        \
        ms_coop_body<B>(S, Q, *seedp, part, invp, W);                                                           \
    }                                                                                                           \
    __global__ void __launch_bounds__(256) k_ms_biased_##B(DScene S, const sms_ss_query *Q, int64_t n,          \
                                                           const uint64_t *__restrict__ seedp, float *out,      \
                                                           EstimatorWork W) {                                   \
        __shared__ float sol_lds[4][MS_BIASED_MAX_SOL][3];                                                      \
        ms_biased_body<B>(S, Q, n, *seedp, out, W, sol_lds[threadIdx.x >> 6]);                                  \
    }

SMS_MS_KERNELS(2)
SMS_MS_KERNELS(4)

} // namespace smsdev
