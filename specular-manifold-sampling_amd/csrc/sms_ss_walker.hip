// sms_ss_walker.hip -- first pass of the unbiased single-scatter estimator
// (SpecularManifoldSingleScatter::specular_manifold_sampling, manifold_ss.cpp:28-141) as a
// persistent walker.
//
// Every lane holds one walk and advances it by exactly one ray cast per loop trip: the seed
// ray of sample_path (manifold_ss.cpp:243-249), a Newton reprojection (:369-371) or the
// shadow ray of the first solution (:96-103).  A lane whose walk ends takes its next walk in
// the same trip -- the chain's next Bernoulli trial, the next caster, or the next query from
// the wavefront's share of a device queue -- so the 64 lanes stay busy while walks of 0 to 20
// iterations and chains of geometric length end at different times.  All lanes share one
// ray-cast site, one surface-interaction build and one Newton-step evaluation per trip.
//
// Walks, random streams and arithmetic are those of sample_path_ss / k_ss_estimate, so every
// estimate, solution and counter is bit-identical; only the order in which walks run differs.
// Chains still unresolved after `lane_trials` trials go to the spread overflow list for the
// cooperative pass, with the same record layout.
#include "kernels.h"
#include "dev_ss_est.h"

namespace smsdev {

namespace {

enum : int { WK_NONE = 0, WK_SEED = 1, WK_NEWTON = 2, WK_VIS = 3 };
constexpr int WK_CHUNK = 64;   // queries a wavefront claims from the queue at a time

// Per-lane state, kept small so that it is all a lane carries through the ray cast: of the
// query only its index, random-stream key and shading point (the rest is re-read from the query
// array when a walk ends), of the walk's manifold vertex only the hit that produced it -- the
// vertex is rebuilt from that record (make_si + mv_from_si, deterministic) at every Newton step,
// which is where the old code built it after each accepted reprojection.  The emitter
// interaction lives in LDS.  A walk can only converge right after an accepted reprojection (the
// constraint depends on the vertex alone), so on success the rebuilt interaction is the
// solution, valid exactly when at least one iteration ran.
struct WalkerLane {
    // query
    int64_t q;
    V3 si_p;
    uint64_t qkey;
    // chain (query, caster)
    int c, shape;
    int64_t slot, k;
    V3 direction;
    V3 part_first;               // bsdf * contribution of the first solution, kept until its shadow ray
    V3 p_first;
    V2 uv_first;
    // walk: the record of the hit that made the current vertex (a ray from si_p along key_d)
    int stage;
    V3 n_off;                    // the chain's offset normal (rough casters; (0, 0, 1) otherwise)
    Hit key;
    V3 key_d;
    float beta;
    int its;
    // counters (per lane and call; far below 2^32)
    unsigned walks, conv, iters, bcalls, btrials, bmax;
};

} // namespace

__device__ __forceinline__ void ss_walker_body(const DScene &S, const sms_ss_query *__restrict__ Q, int64_t n,
                                               uint64_t seed, float *__restrict__ part, float *__restrict__ invp,
                                               const EstimatorWork &W) {
    const sms_config &cfg = S.cfg;
    const int nc = S.n_casters_ss;
    const int lane = threadIdx.x & 63;
    const Proposal &P = W.prop;
    __shared__ EI ei_lds[256];
    EI &ei_lane = ei_lds[threadIdx.x];
    WalkerLane L;
    L.stage = WK_NONE; L.q = -1; L.c = 0; L.shape = -1; L.slot = 0; L.k = 0; L.its = 0; L.beta = 1.f;
    L.walks = L.conv = L.iters = L.bcalls = L.btrials = L.bmax = 0u;
    L.qkey = 0;
    L.si_p = v3(0.f, 0.f, 0.f);
    L.n_off = v3(0.f, 0.f, 1.f);
    L.key_d = v3(0.f, 0.f, 0.f);
    L.key.hit = false; L.key.t = 0.f; L.key.shape = L.key.prim = 0; L.key.b1 = L.key.b2 = L.key.lx = L.key.ly = 0.f;
    // wavefront-uniform query pool
    int64_t pool_next = 0, pool_end = 0;
    bool exhausted = false;

    // ---- chain / walk transitions (no ray casts)
    auto begin_chain = [&]() {
        L.slot = L.q * nc + L.c;
        part[3 * L.slot] = 0.f; part[3 * L.slot + 1] = 0.f; part[3 * L.slot + 2] = 0.f;
        invp[L.slot] = 0.f;
        L.shape = S.casters_ss[L.c];
        (void) caster_offset(S, L.qkey, L.c, L.n_off);
        L.k = 0;
        L.stage = WK_SEED;
    };
    auto end_chain = [&]() {
        L.c++;
        if (L.c < nc) begin_chain();
        else L.stage = WK_NONE;
    };
    auto resolve = [&](float inv_prob) {
        const unsigned tr = (unsigned) L.k;
        L.btrials += tr; L.bmax = L.bmax > tr ? L.bmax : tr;
        invp[L.slot] = inv_prob;
        const float wgt = mean3(&part[3 * L.slot]) * inv_prob;
        if (wgt > 0.f && capturing(P, L.shape, SMS_CAPTURE_ESTIMATES))
            capture_sample(S, P, ss_query_si(Q[L.q]), ei_lane, L.uv_first, wgt, proposal_anchor(P, L.q));
        end_chain();
    };
    // the chain's trial k did not re-find its solution: the next trial, or the cap, or the
    // overflow list for the cooperative pass
    auto next_trial = [&]() {
        L.k++;
        if (cfg.max_trials > 0 && L.k > cfg.max_trials) { resolve(0.f); return; }
        if (L.k > W.lane_trials) {
            const int sub = (int) ((L.q / 256) % SPREAD);
            unsigned long long o = spread_append(W.over, sub, 1ull);
            OverflowRec rec;
            rec.q = L.slot; rec.k = L.k;
            rec.sol[0] = L.direction.x; rec.sol[1] = L.direction.y; rec.sol[2] = L.direction.z;
            rec.shapes[0] = __float_as_int(L.uv_first.x); rec.shapes[1] = __float_as_int(L.uv_first.y);
            rec.shapes[2] = rec.shapes[3] = 0;
            W.over.items[(size_t) sub * W.over.sub_cap + o] = rec;
            end_chain();
            return;
        }
        L.stage = WK_SEED;
    };
    // a walk ended without a solution after `iterations` Newton iterations
    auto walk_failed = [&](int iterations) {
        L.walks++; L.iters += (unsigned) iterations;
        if (L.k == 0) { end_chain(); return; }
        next_trial();
    };
    // a walk converged: its solution is vtx, made by the surface interaction hs
    auto walk_converged = [&](int iterations, const MV &vtx, const SI &hs) {
        L.walks++; L.iters += (unsigned) iterations;
        L.conv++;
        if (capturing(P, L.shape, SMS_CAPTURE_SOLUTIONS))
            capture_sample(S, P, ss_query_si(Q[L.q]), ei_lane, hs.uv, 1.f, proposal_anchor(P, L.q));
        if (L.k == 0) {
            L.direction = normalize(vtx.p - L.si_p);
            // the contribution is evaluated now (before the shadow ray) so the solution's
            // interaction need not be kept; it is discarded if the shadow ray is blocked
            const sms_ss_query g = Q[L.q];
            const SI si = ss_query_si(g);
            V3 spec = path_contribution_ss_mv<true, WALKER_STACK>(S, si, ei_lane, vtx, hs.sh_n, hs.wi);
            V3 dl = v3(dot(L.direction, si.sh_s), dot(L.direction, si.sh_t), dot(L.direction, si.sh_n));
            V3 bsdf = receiver_bsdf_eval(S, si, ld3(g.wi), dl);
            L.part_first = v3(bsdf.x * spec.x, bsdf.y * spec.y, bsdf.z * spec.z);
            L.p_first = vtx.p;
            L.uv_first = hs.uv;
            L.stage = WK_VIS;
            return;
        }
        V3 dtr = normalize(vtx.p - L.si_p);
        if (fabsf(dot(L.direction, dtr) - 1.f) < cfg.uniqueness_threshold) { resolve((float) L.k); return; }
        next_trial();
    };
    auto start_query = [&](int64_t q) {
        L.q = q;
        const sms_ss_query g = Q[q];
        if (g.shape < 0 || nc == 0) {          // no receiver: zero estimate, no walks
            for (int c = 0; c < nc; ++c) {
                const int64_t s = q * nc + c;
                part[3 * s] = 0.f; part[3 * s + 1] = 0.f; part[3 * s + 2] = 0.f;
                invp[s] = 0.f;
            }
            L.stage = WK_NONE;
            return;
        }
        L.si_p = ld3(g.p);
        L.qkey = query_key(seed, (uint64_t) q);
        Pcg32 rq;
        stream_seed(rq, L.qkey, 0);
        ei_lane = sample_emitter_interaction(S, L.si_p, S.emitters_ss, S.n_emitters_ss, rq);
        L.c = 0;
        begin_chain();
    };

    while (true) {
        Ray ray;
        ray.o = ray.d = ray.d_rcp = v3(0.f, 0.f, 0.f); ray.mint = 0.f; ray.maxt = 0.f;
        bool has_ray = false, any_hit = false;
        while (true) {
            // advance this lane's state machine to its next ray (or to needing a query)
            while (!has_ray && L.stage != WK_NONE) {
                if (L.stage == WK_SEED) {
                    // ss_walk + the seed of sample_path_ss
                    Pcg32 rw;
                    stream_seed(rw, L.qkey, ((uint64_t) L.c << 32) + (uint64_t) L.k + 1);
                    V2 s = v2(0.f, 0.f);
                    V3 p_start = v3(0.f, 0.f, 0.f);
                    if (P.gmm && L.shape == P.shape) {
                        if (!proposal_seed(S, P, P.gmm + L.q * PROP_RAW, proposal_anchor(P, L.q), rw, s, p_start)) {
                            walk_failed(0);
                            continue;
                        }
                    } else {
                        s.x = pcg32_float(rw); s.y = pcg32_float(rw);
                    }
                    PS ps = sample_position(S, L.shape, s);
                    if (p_start.x != 0.f || p_start.y != 0.f || p_start.z != 0.f) ps.p = p_start;
                    V3 d_tmp = normalize(ps.p - L.si_p);
                    if (cfg.mnee_init) {
                        const DShape &sh = S.shapes[L.shape];
                        if (sh.k != 0.f) d_tmp = normalize(ld3(sh.bbox_center) - L.si_p);
                        else d_tmp = normalize(ei_lane.p - L.si_p);
                    }
                    ray = make_ray(L.si_p, d_tmp);
                    has_ray = true; any_hit = false;
                } else if (L.stage == WK_NEWTON) {
                    // one iteration of newton_solver (manifold_ss.cpp:329-405) at the vertex the
                    // last accepted hit made
                    const SI hs = make_si(S, make_ray(L.si_p, L.key_d), L.key);
                    const MV vtx = mv_from_si(S, hs);
                    if (L.its >= cfg.max_iterations) { walk_failed(L.its); continue; }
                    const EI ei = ei_lane;
                    Step st = cfg.halfvector_constraints ? step_halfvector(L.si_p, vtx, ei, L.n_off)
                                                         : step_anglediff(L.si_p, vtx, ei, L.n_off);
                    if (!st.ok) { walk_failed(L.its); continue; }
                    if (norm2(st.C) < cfg.solver_threshold) {
                        V3 wx = normalize(L.si_p - vtx.p);
                        V3 wy = ei.kind == SMS_EI_DIRECTIONAL ? ei.d : normalize(ei.p - vtx.p);
                        float cos_theta_x = dot(vtx.gn, wx), cos_theta_y = dot(vtx.gn, wy);
                        bool refraction = cos_theta_x * cos_theta_y < 0.f, reflection = !refraction;
                        bool type_ok = !((vtx.eta == 1.f && !reflection) || (vtx.eta != 1.f && !refraction));
                        // converged at iteration 0 reports failure (no reprojection has run yet)
                        if (type_ok && L.its > 0) walk_converged(L.its, vtx, hs);
                        else walk_failed(L.its);
                        continue;
                    }
                    V3 p_prop = vtx.p - cfg.step_scale * L.beta * (vtx.dp_du * st.dX.x + vtx.dp_dv * st.dX.y);
                    V3 d_prop = normalize(p_prop - L.si_p);
                    ray = make_ray(L.si_p, d_prop);
                    has_ray = true; any_hit = false;
                } else {
                    // shadow ray of the first solution (ss_visible)
                    if (ei_lane.kind == SMS_EI_DIRECTIONAL) {
                        ray = make_ray(L.p_first, ei_lane.d);
                    } else {
                        const V3 ep = ei_lane.p;
                        V3 p = L.p_first;
                        V3 d = p - ep;
                        float dist = norm(d);
                        d = d * (1.f / dist);
                        ray = make_ray(ep, d, RAY_EPS_F * (1.f + hmax_abs(ep)), dist * (1.f - RAY_EPS_F));
                    }
                    has_ray = true; any_hit = true;
                }
            }
            // lanes without work take queries from the wavefront's pool (wavefront-uniform)
            const unsigned long long req = __ballot(L.stage == WK_NONE && !has_ray);
            if (!req || exhausted) break;
            if (pool_next >= pool_end) {
                unsigned long long base = 0;
                if (lane == 0) base = atomicAdd(W.cursor, (unsigned long long) WK_CHUNK);
                base = __shfl(base, 0, 64);
                if ((int64_t) base >= n) { exhausted = true; break; }
                pool_next = (int64_t) base;
                pool_end = (int64_t) base + WK_CHUNK < n ? (int64_t) base + WK_CHUNK : n;
            }
            const int rank = __popcll(req & ((1ull << lane) - 1ull));
            const int64_t avail = pool_end - pool_next;
            if (((req >> lane) & 1ull) && rank < avail) start_query(pool_next + rank);
            const int64_t taken = (int64_t) __popcll(req) < avail ? (int64_t) __popcll(req) : avail;
            pool_next += taken;
        }
        if (!__ballot(has_ray)) break;
        // one ray cast per lane
        Hit h;
        h.hit = false; h.t = __builtin_inff(); h.shape = 1 << 30; h.prim = 1 << 30; h.b1 = h.b2 = h.lx = h.ly = 0.f;
        if (has_ray) h = scene_hit_t<true, true, WALKER_STACK>(S, ray, any_hit);
        if (!has_ray) continue;
        if (L.stage == WK_VIS) {
            if (h.hit) { end_chain(); continue; }
            // visible: the first solution's bsdf * contribution (manifold_ss.cpp:104-118)
            part[3 * L.slot] = L.part_first.x; part[3 * L.slot + 1] = L.part_first.y;
            part[3 * L.slot + 2] = L.part_first.z;
            L.bcalls++;
            L.k = 1;
            L.stage = WK_SEED;
            continue;
        }
        // seed ray or reprojection: a hit on the walk's shape becomes the walk's vertex (kept as
        // its hit record); the seed ray's target is the chain's shape, a reprojection's the shape
        // of the current vertex, which is the same shape
        const bool on_shape = h.hit && h.shape == L.shape;
        if (on_shape) { L.key = h; L.key_d = ray.d; }
        if (L.stage == WK_SEED) {
            if (!on_shape) { walk_failed(0); continue; }
            L.beta = 1.f; L.its = 0;
            L.stage = WK_NEWTON;
        } else {
            if (!on_shape) L.beta = 0.5f * L.beta;
            else L.beta = fmin_(1.f, 2.f * L.beta);
            L.its += 1;
        }
    }
    stat_add(W.stats + 0, L.walks); stat_add(W.stats + 1, L.conv); stat_add(W.stats + 2, L.iters);
    stat_add(W.stats + 3, L.bcalls); stat_add(W.stats + 4, L.btrials); stat_max(W.stats + 5, (unsigned long long) L.bmax);
}

__global__ void __launch_bounds__(256, 3) k_ss_walker(DScene S, const sms_ss_query *__restrict__ Q, int64_t n,
                                                   const uint64_t *__restrict__ seedp, float *__restrict__ part, float *__restrict__ invp,
                                                   EstimatorWork W) {
    const uint64_t seed = *seedp;
    ss_walker_body(S, Q, n, seed, part, invp, W);
}

} // namespace smsdev
