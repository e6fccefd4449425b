// sms_kernels.hip -- gfx950 kernels of the SMS solver path.
//
//   k_ss_walk / k_glint_walk      parity entry points: one lane per walk record
//   k_glint_queries               camera ray differentials -> glint queries (footprints)
//   (the unbiased glint estimator lives in sms_glint_kernels.hip)
//   k_ss_estimate                 unbiased single-scatter SMS, lane per query: first walk,
//                                 contribution and up to `lane_trials` Bernoulli trials;
//                                 longer trial loops go to an overflow list
//   k_ss_coop                     overflow pass: one wavefront per (query, caster), 64 trials
//                                 per round, the first match found by a ballot
//   k_ss_biased                   biased single-scatter SMS: one wavefront per query,
//                                 lanes = trial-set members, in-order de-duplication
//
// Random numbers: query q draws its emitter sample from PCG32 stream (q, 0) and walk k of
// caster c from stream (q, (c << 32) + k + 1), so every result is independent of which
// lane or pass ran the walk.
#include "kernels.h"
#include "dev_ss_est.h"

namespace smsdev {

// ================================================================ parity walks
__global__ void k_ss_walk(DScene S, const sms_ss_walk_in *__restrict__ in, int64_t n, sms_ss_walk_out *__restrict__ out) {
    int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const sms_ss_walk_in w = in[i];
    EI ei;
    ei.p = ld3(w.ei.p); ei.n = ld3(w.ei.n); ei.d = ld3(w.ei.d); ei.weight = ld3(w.ei.weight);
    ei.pdf = w.ei.pdf; ei.kind = w.ei.kind; ei.emitter = w.ei.emitter;
    PathResult r = sample_path_ss(S, w.shape, ld3(w.si_p), ei, v2(w.sample[0], w.sample[1]), ld3(w.n_offset), ld3(w.p_start),
                                  v2(w.lean_sample[0], w.lean_sample[1]));
    sms_ss_walk_out o;
    o.success = r.ok; o.iterations = r.iterations;
    o.final_shape = r.si_final.valid ? r.si_final.shape : -1;
    o.init_valid = r.init_valid;
    st3(o.p, r.si_final.p); st3(o.n, r.si_final.sh_n); st3(o.p_init, r.p_init);
    out[i] = o;
}

__global__ void k_glint_walk(DScene S, const sms_glint_walk_in *__restrict__ in, int64_t n, sms_glint_walk_out *__restrict__ out) {
    int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const sms_glint_walk_in w = in[i];
    sms_glint_walk_out o;
    o.success = 0; o.iterations = 0; o.uv[0] = o.uv[1] = 0.f; o.uv_init[0] = o.uv_init[1] = 0.f;
    GlintFootprint fp;
    if (!glint_target(ld3(w.sensor_p), ld3(w.ei_p), ld3(w.si_p), ld3(w.frame_s), ld3(w.frame_t), ld3(w.frame_n),
                      ld3(w.n_offset), fp.target)) { out[i] = o; return; }
    const DShape &sh = S.shapes[w.shape];
    const DNormalmap nm = S.nmaps[sh.normalmap];
    fp.uv = v2(w.uv[0], w.uv[1]); fp.duv_dx = v2(w.duv_dx[0], w.duv_dx[1]); fp.duv_dy = v2(w.duv_dy[0], w.duv_dy[1]);
    fp.slope = nm.slope; fp.res = nm.res; fp.stride = nm.stride; fp.fres = (float) nm.res; fp.tiles = sh.tiles;
    footprint_setup(fp);
    V2 xi = v2(w.sample[0], w.sample[1]);
    if (w.xi_start[0] != 1.f || w.xi_start[1] != 1.f) xi = v2(w.xi_start[0], w.xi_start[1]);
    GlintWalk gw;
    glint_walk_start(fp, xi, gw);
    o.uv_init[0] = gw.uv.x; o.uv_init[1] = gw.uv.y;
    int st;
    while ((st = glint_walk_iter(fp, S.cfg, gw)) == WALK_RUNNING) {}
    o.success = st == WALK_CONVERGED; o.iterations = gw.it; o.uv[0] = gw.uv.x; o.uv[1] = gw.uv.y;
    out[i] = o;
}

// ================================================================ glint queries (camera)
__global__ void k_glint_queries(DScene S, int spp, const uint64_t *__restrict__ seedp, sms_glint_query *__restrict__ q, int64_t n) {
    const int64_t idx = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    q[idx] = make_glint_query<true>(S, *seedp, spp, idx);
}

// ================================================================ single scatter, unbiased
// Pass 1: lane per query; for every caster the first walk, visibility, contribution and up
// to lane_trials Bernoulli trials.  Per (query, caster) the product bsdf*specular (bxs) and
// the trial count are stored; unresolved (query, caster) pairs go to the overflow list.
__global__ void __launch_bounds__(256) k_ss_estimate(DScene S, const sms_ss_query *__restrict__ Q, int64_t n, const uint64_t *__restrict__ seedp,
                                                     float *__restrict__ part, float *__restrict__ invp, EstimatorWork W) {
    const uint64_t seed = *seedp;
    const sms_config &cfg = S.cfg;
    unsigned long long st_walks = 0, st_conv = 0, st_iters = 0, st_bcalls = 0, st_btrials = 0, st_bmax = 0;
    const int nc = S.n_casters_ss;
    // one lane per query; the launch covers n with one block per 256 queries, so a block's
    // overflow items fit its sub-list of the spread overflow list
    const int sub = (int) (blockIdx.x % SPREAD);
    for (int64_t q = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t) gridDim.x * blockDim.x) {
        const sms_ss_query g = Q[q];
        SI si = ss_query_si(g);
        V3 wi_local = ld3(g.wi);
        uint64_t qkey = query_key(seed, (uint64_t) q);
        Pcg32 rq;
        stream_seed(rq, qkey, 0);
        EI ei = sample_emitter_interaction(S, si.p, S.emitters_ss, S.n_emitters_ss, rq);
        for (int c = 0; c < nc; ++c) {
            int64_t slot = q * nc + c;
            part[3 * slot] = 0.f; part[3 * slot + 1] = 0.f; part[3 * slot + 2] = 0.f;
            invp[slot] = 0.f;
            if (g.shape < 0) continue;                 // no receiver: zero estimate, no walks
            int shape = S.casters_ss[c];
            V3 n_off;
            (void) caster_offset(S, qkey, c, n_off);
            PathResult r = ss_walk(S, W.prop, q, qkey, c, shape, 0, si.p, ei, n_off);
            st_walks++; st_iters += (unsigned long long) r.iterations;
            if (!r.ok) continue;
            st_conv++;
            V3 direction = normalize(r.si_final.p - si.p);
            if (capturing(W.prop, shape, SMS_CAPTURE_SOLUTIONS))
                capture_sample(S, W.prop, si, ei, r.si_final.uv, 1.f, proposal_anchor(W.prop, q));
            if (!ss_visible(S, ei, r.si_final.p)) continue;
            V3 spec = path_contribution_ss(S, si, ei, r.si_final);
            V3 dl = v3(dot(direction, si.sh_s), dot(direction, si.sh_t), dot(direction, si.sh_n));
            V3 bsdf = receiver_bsdf_eval(S, si, wi_local, dl);
            part[3 * slot] = bsdf.x * spec.x; part[3 * slot + 1] = bsdf.y * spec.y; part[3 * slot + 2] = bsdf.z * spec.z;
            st_bcalls++;
            int64_t k = 1;
            float inv_prob = -1.f;
            while (true) {
                PathResult t = ss_walk(S, W.prop, q, qkey, c, shape, (uint64_t) k, si.p, ei, n_off);
                st_walks++; st_iters += (unsigned long long) t.iterations;
                if (t.ok) st_conv++;
                if (t.ok && capturing(W.prop, shape, SMS_CAPTURE_SOLUTIONS))
                    capture_sample(S, W.prop, si, ei, t.si_final.uv, 1.f, proposal_anchor(W.prop, q));
                V3 dtr = normalize(t.si_final.p - si.p);
                if (t.ok && fabsf(dot(direction, dtr) - 1.f) < cfg.uniqueness_threshold) { inv_prob = (float) k; break; }
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
                    capture_sample(S, W.prop, si, ei, r.si_final.uv, wgt, proposal_anchor(W.prop, q));
            } else {
                unsigned long long o = spread_append(W.over, sub, 1ull);
                OverflowRec rec;
                rec.q = slot; rec.k = k;
                rec.sol[0] = direction.x; rec.sol[1] = direction.y; rec.sol[2] = direction.z;
                // the first walk's uv, for a training capture at the chain's end
                rec.shapes[0] = __float_as_int(r.si_final.uv.x); rec.shapes[1] = __float_as_int(r.si_final.uv.y);
                rec.shapes[2] = rec.shapes[3] = 0;
                W.over.items[(size_t) sub * W.over.sub_cap + o] = rec;
            }
        }
    }
    stat_add(W.stats + 0, st_walks); stat_add(W.stats + 1, st_conv); stat_add(W.stats + 2, st_iters);
    stat_add(W.stats + 3, st_bcalls); stat_add(W.stats + 4, st_btrials); stat_max(W.stats + 5, st_bmax);
}

__global__ void __launch_bounds__(256) k_ss_coop(DScene S, const sms_ss_query *__restrict__ Q, const uint64_t *__restrict__ seedp,
                                                 const float *__restrict__ part, float *__restrict__ invp,
                                                 EstimatorWork W) {
    const uint64_t seed = *seedp;
    const sms_config &cfg = S.cfg;
    const int lane = threadIdx.x & 63;
    const int nc = S.n_casters_ss;
    unsigned long long st_walks = 0, st_conv = 0, st_iters = 0, st_btrials = 0, st_bmax = 0, st_tail = 0;
    const unsigned long long wave = ((unsigned long long) blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const SpreadWalk sw = spread_walk(W.over, wave, ((unsigned long long) gridDim.x * blockDim.x) >> 6);
    for (unsigned long long p = sw.pos; p < sw.count; p += sw.step) {
        const OverflowRec rec = W.over.items[(size_t) sw.sub * W.over.sub_cap + p];
        int64_t q = rec.q / nc;
        int c = (int) (rec.q % nc);
        const sms_ss_query g = Q[q];
        SI si = ss_query_si(g);
        uint64_t qkey = query_key(seed, (uint64_t) q);
        Pcg32 rq;
        stream_seed(rq, qkey, 0);
        EI ei = sample_emitter_interaction(S, si.p, S.emitters_ss, S.n_emitters_ss, rq);
        V3 direction = v3(rec.sol[0], rec.sol[1], rec.sol[2]);
        int shape = S.casters_ss[c];
        V3 n_off;
        (void) caster_offset(S, qkey, c, n_off);
        int64_t kcap = cfg.max_trials > 0 ? (int64_t) cfg.max_trials : (int64_t) 1 << 40;
        int64_t found = -1;
        for (int64_t base = rec.k; base <= kcap && found < 0; base += 64) {
            int64_t k = base + lane;
            bool match = false, ran = false, conv = false;
            int its = 0;
            V2 uv_t = v2(0.f, 0.f);
            if (k <= kcap) {
                PathResult t = ss_walk(S, W.prop, q, qkey, c, shape, (uint64_t) k, si.p, ei, n_off);
                ran = true; its = t.iterations; conv = t.ok;
                uv_t = t.si_final.uv;
                V3 dtr = normalize(t.si_final.p - si.p);
                match = t.ok && fabsf(dot(direction, dtr) - 1.f) < cfg.uniqueness_threshold;
            }
            unsigned long long ballot = __ballot(match);
            if (ballot) found = base + (int64_t) __ffsll((long long) ballot) - 1;
            // the trials the sequential loop runs: up to the first match
            if (ran && (found < 0 || k <= found)) {
                st_walks++; st_iters += (unsigned long long) its; st_tail += (unsigned long long) its;
                if (conv) {
                    st_conv++;
                    if (capturing(W.prop, shape, SMS_CAPTURE_SOLUTIONS))
                        capture_sample(S, W.prop, si, ei, uv_t, 1.f, proposal_anchor(W.prop, q));
                }
            }
        }
        if (lane == 0) {
            invp[rec.q] = found >= 0 ? (float) found : 0.f;
            unsigned long long tr = found >= 0 ? (unsigned long long) found : (unsigned long long) kcap + 1ull;
            st_btrials += tr; st_bmax = st_bmax > tr ? st_bmax : tr;
            const float wgt = found > 0 ? mean3(&part[3 * rec.q]) * (float) found : 0.f;
            if (wgt > 0.f && capturing(W.prop, shape, SMS_CAPTURE_ESTIMATES))
                capture_sample(S, W.prop, si, ei, v2(__int_as_float(rec.shapes[0]), __int_as_float(rec.shapes[1])), wgt,
                               proposal_anchor(W.prop, q));
        }
    }
    stat_add(W.stats + 0, st_walks); stat_add(W.stats + 1, st_conv); stat_add(W.stats + 2, st_iters);
    stat_add(W.stats + 4, st_btrials); stat_max(W.stats + 5, st_bmax); stat_add(W.stats + 6, st_tail);
}

// result = sum over casters (in order) of bxs * inv_prob / p_offset  (manifold_ss.cpp:139-140,198);
// offsets = 0 (the multi-scatter estimator, whose casters take no offset normals): p_offset = 1
__global__ void k_ss_finalize(DScene S, const uint64_t *__restrict__ seedp, const float *__restrict__ part,
                              const float *__restrict__ invp, int nc, int64_t n, int offsets, float *__restrict__ out) {
    int64_t q = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const uint64_t qkey = query_key(*seedp, (uint64_t) q);
    V3 r = v3(0.f, 0.f, 0.f);
    for (int c = 0; c < nc; ++c) {
        int64_t s = q * nc + c;
        float ip = invp[s];
        V3 n_off;
        const float p_off = offsets ? caster_offset(S, qkey, c, n_off) : 1.f;
        r = r + v3(part[3 * s] * ip, part[3 * s + 1] * ip, part[3 * s + 2] * ip) / p_off;
    }
    write_rgb(out, q, r);
}

// ================================================================ single scatter, biased
// One wavefront per query: lane j runs trial-set member m = chunk*64 + j; solutions are
// de-duplicated in m order through an LDS list and their contributions summed in m order
// (manifold_ss.cpp:142-197).
constexpr int BIASED_MAX_SOL = 128;
__global__ void __launch_bounds__(256) k_ss_biased(DScene S, const sms_ss_query *__restrict__ Q, int64_t n, const uint64_t *__restrict__ seedp,
                                                   float *__restrict__ out, EstimatorWork W) {
    const uint64_t seed = *seedp;
    __shared__ float sol_lds[4][BIASED_MAX_SOL][3];
    const sms_config &cfg = S.cfg;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int nc = S.n_casters_ss;
    unsigned long long st_walks = 0, st_conv = 0, st_iters = 0;
    const int64_t n_waves = ((int64_t) gridDim.x * blockDim.x) >> 6;
    for (int64_t q = ((int64_t) blockIdx.x * blockDim.x + threadIdx.x) >> 6; q < n; q += n_waves) {
        const sms_ss_query g = Q[q];
        SI si = ss_query_si(g);
        V3 wi_local = ld3(g.wi);
        uint64_t qkey = query_key(seed, (uint64_t) q);
        Pcg32 rq;
        stream_seed(rq, qkey, 0);
        EI ei = sample_emitter_interaction(S, si.p, S.emitters_ss, S.n_emitters_ss, rq);
        V3 result = v3(0.f, 0.f, 0.f);
        const int ncq = g.shape < 0 ? 0 : nc;        // no receiver: zero estimate, no walks
        for (int c = 0; c < ncq; ++c) {
            int shape = S.casters_ss[c];
            V3 n_off;
            const float p_off = caster_offset(S, qkey, c, n_off);
            V3 value = v3(0.f, 0.f, 0.f);
            int n_sol = 0;
            for (int base = 0; base < cfg.max_trials; base += 64) {
                int m = base + lane;
                bool cand = false;
                V3 direction = v3(0.f, 0.f, 0.f);
                PathResult r;
                r.ok = false;
                if (m < cfg.max_trials) {
                    r = ss_walk(S, W.prop, q, qkey, c, shape, (uint64_t) m, si.p, ei, n_off);
                    st_walks++; st_iters += (unsigned long long) r.iterations;
                    if (r.ok) {
                        st_conv++;
                        direction = normalize(r.si_final.p - si.p);
                        cand = ss_visible(S, ei, r.si_final.p);
                    }
                }
                // in-order de-duplication against the solution list
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
                        if (s < n_sol) d = fabsf(dot(dj, v3(sol_lds[wv][s][0], sol_lds[wv][s][1], sol_lds[wv][s][2])) - 1.f)
                                           < cfg.uniqueness_threshold;
                        if (__ballot(d)) { dup = true; break; }
                    }
                    if (!dup) {
                        if (n_sol < BIASED_MAX_SOL && lane == 0) {
                            sol_lds[wv][n_sol][0] = dx; sol_lds[wv][n_sol][1] = dy; sol_lds[wv][n_sol][2] = dz;
                        }
                        if (n_sol < BIASED_MAX_SOL) n_sol++;
                        if (lane == j) unique = true;
                    }
                    __builtin_amdgcn_wave_barrier();
                }
                V3 contrib = v3(0.f, 0.f, 0.f);
                if (unique) {
                    V3 dl = v3(dot(direction, si.sh_s), dot(direction, si.sh_t), dot(direction, si.sh_n));
                    V3 bsdf = receiver_bsdf_eval(S, si, wi_local, dl);
                    V3 spec = path_contribution_ss(S, si, ei, r.si_final);
                    contrib = v3(bsdf.x * spec.x, bsdf.y * spec.y, bsdf.z * spec.z);
                }
                unsigned long long umask = __ballot(unique);
                while (umask) {
                    int j = __ffsll((long long) umask) - 1;
                    umask &= umask - 1ull;
                    value = value + v3(__shfl(contrib.x, j, 64), __shfl(contrib.y, j, 64), __shfl(contrib.z, j, 64));
                }
            }
            result = result + value / p_off;
        }
        if (lane == 0) write_rgb(out, q, result);
    }
    stat_add(W.stats + 0, st_walks); stat_add(W.stats + 1, st_conv); stat_add(W.stats + 2, st_iters);
}

// ================================================================ single-scatter camera pass
// Camera pass of the single-scatter integrator (path_sms_ss.cpp:94-251), SMS term only: the
// pixel sample's ray (camera model and jitter stream of k_glint_queries) follows smooth
// dielectric interfaces -- reflection with probability F, else refraction
// (dielectric.cpp:201-313: weight 1 for reflection, eta_ti^2 for refraction) -- for up to
// SS_CAMERA_DEPTH vertices; the first caustic receiver that is not a caster becomes the query
// (path_sms_ss.cpp:118-124) with the path throughput as its weight.  Paths that leave the
// scene or stop on another surface give shape -1 and weight 0.
constexpr int SS_CAMERA_DEPTH = 8;
__global__ void k_ss_queries(DScene S, int spp, const uint64_t *__restrict__ seedp, sms_ss_query *__restrict__ q, float *__restrict__ w,
                             int64_t n) {
    const uint64_t seed = *seedp;
    const int64_t idx = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const DCamera &c = S.cam;
    sms_ss_query o;
    float *z = (float *) &o;
    for (int k = 0; k < (int) (sizeof(o) / 4); ++k) z[k] = 0.f;
    o.shape = -1;
    float weight = 0.f;
    const int64_t pixel = idx / spp;
    const int px = (int) (pixel % c.w), py = (int) (pixel / c.w);
    Pcg32 r;
    stream_seed(r, query_key(seed, (uint64_t) idx), 0xFFFFFFFFull);
    float sx = pcg32_float(r), sy = pcg32_float(r);
    V2 pos = v2(((float) px + sx) / (float) c.w, ((float) py + sy) / (float) c.h);
    V3 near_p = v3((1.f - 2.f * pos.x) * c.tan_half, (1.f - 2.f * pos.y) * c.rect_y, 1.f);
    V3 dloc = normalize(near_p);
    V3 d = to_world(ld3(c.left), ld3(c.up), ld3(c.dir), dloc);
    Ray ray = make_ray(ld3(c.o), d, 1e-2f * (1.f / dloc.z), 1e4f * (1.f / dloc.z));
    float throughput = 1.f;
    for (int depth = 0; depth < SS_CAMERA_DEPTH; ++depth) {
        SI si = ray_intersect(S, ray);
        if (!si.valid) break;
        const DShape &sh = S.shapes[si.shape];
        V3 wi = v3(-dot(ray.d, si.sh_s), -dot(ray.d, si.sh_t), -dot(ray.d, si.sh_n));
        if ((sh.flags & SMS_FLAG_CAUSTIC_RECEIVER) && !(sh.flags & SMS_FLAG_CAUSTIC_CASTER_SS)) {
            o.shape = si.shape;
            st3(o.p, si.p); st3(o.n, si.n); st3(o.sh_n, si.sh_n); st3(o.sh_s, si.sh_s); st3(o.sh_t, si.sh_t);
            st3(o.wi, wi);
            weight = throughput;
            break;
        }
        if (sh.bsdf != SMS_BSDF_DIELECTRIC) break;
        // fresnel (fresnel.h:34-80) with the transmitted cosine and relative IOR
        const float cos_i = wi.z, eta = sh.eta_r;
        const bool outside = cos_i >= 0.f;
        const float rcp_eta = 1.f / eta, eta_ti = outside ? rcp_eta : eta;
        const float cos_t_sqr = fmaf(-fmaf(-cos_i, cos_i, 1.f), eta_ti * eta_ti, 1.f);
        const float cos_t_abs = safe_sqrt(cos_t_sqr);
        const float F = fresnel(cos_i, eta);
        V3 wo;
        if (pcg32_float(r) <= F) {
            wo = v3(-wi.x, -wi.y, wi.z);
        } else {
            const float cos_t = outside ? -cos_t_abs : cos_t_abs;
            wo = v3(-eta_ti * wi.x, -eta_ti * wi.y, cos_t);
            throughput = throughput * (eta_ti * eta_ti);
        }
        ray = spawn_ray(si.p, to_world(si.sh_s, si.sh_t, si.sh_n, wo));
    }
    q[idx] = o;
    w[idx] = weight;
}

// Film: per pixel, the sums over its spp samples (in sample order) of weight * rgb and of its
// square, added to image / image_sq (box filter; the caller divides by the samples taken).
// A sample whose weighted value has a non-finite or negative (below -1e-5) channel is dropped,
// as ImageBlock::put drops it (src/librender/imageblock.cpp:85-107); image_n counts the rest.
__global__ void k_film_accumulate(const float *__restrict__ rgb, const float *__restrict__ w, int64_t pixels, int spp,
                                  float *__restrict__ image, float *__restrict__ image_sq, float *__restrict__ image_n) {
    const int64_t p = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= pixels) return;
    float s[3] = {0.f, 0.f, 0.f}, s2[3] = {0.f, 0.f, 0.f};
    int valid = 0;
    for (int k = 0; k < spp; ++k) {
        const int64_t i = p * spp + k;
        const float wi = w[i];
        float v[3];
        bool ok = true;
        for (int a = 0; a < 3; ++a) {
            v[a] = wi * rgb[3 * i + a];
            ok = ok && isfinite(v[a]) && v[a] >= -1e-5f;
        }
        if (!ok) continue;
        ++valid;
        for (int a = 0; a < 3; ++a) {
            s[a] = s[a] + v[a];
            s2[a] = s2[a] + v[a] * v[a];
        }
    }
    for (int a = 0; a < 3; ++a) {
        image[3 * p + a] += s[a];
        if (image_sq) image_sq[3 * p + a] += s2[a];
    }
    if (image_n) image_n[p] += (float) valid;
}

// ================================================================ learned seed proposal
// network inputs of every query of an estimate: the shading point, its shading normal and the
// light sample the estimator draws for it (same stream, same arithmetic as k_ss_estimate; ms:
// the multi-scatter estimator's emitters)
__global__ void k_prop_inputs(DScene S, const sms_ss_query *__restrict__ Q, int64_t n, const uint64_t *__restrict__ seedp,
                              int ms, float *__restrict__ x) {
    const uint64_t seed = *seedp;
    const int64_t q = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const sms_ss_query g = Q[q];
    SI si = ss_query_si(g);
    Pcg32 rq;
    stream_seed(rq, query_key(seed, (uint64_t) q), 0);
    EI ei = ms ? sample_emitter_interaction(S, si.p, S.emitters_ms, S.n_emitters_ms, rq)
               : sample_emitter_interaction(S, si.p, S.emitters_ss, S.n_emitters_ss, rq);
    float r[9];
    proposal_inputs(S, si.p, si.sh_n, ei.p, r);
    for (int j = 0; j < 9; ++j) x[q * 9 + j] = r[j];
}

// Anchor of each single-scatter query on mesh `shape`: the uv of the first surface the ray from
// the shading point toward the estimator's light sample meets, if that surface is `shape`
// (the segment manifold next-event estimation starts its walk from), else NaN.
__global__ void k_prop_anchor(DScene S, const sms_ss_query *__restrict__ Q, int64_t n, const uint64_t *__restrict__ seedp,
                              int shape, float *__restrict__ anchor) {
    const uint64_t seed = *seedp;
    const int64_t q = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const sms_ss_query g = Q[q];
    float a0 = __int_as_float(0x7fc00000), a1 = a0;
    if (g.shape >= 0) {
        const SI si = ss_query_si(g);
        Pcg32 rq;
        stream_seed(rq, query_key(seed, (uint64_t) q), 0);
        const EI ei = sample_emitter_interaction(S, si.p, S.emitters_ss, S.n_emitters_ss, rq);
        const SI h = ray_intersect(S, make_ray(si.p, normalize(ei.p - si.p)));
        if (h.valid && h.shape == shape) { a0 = h.uv.x; a1 = h.uv.y; }
    }
    anchor[2 * q] = a0;
    anchor[2 * q + 1] = a1;
}

__global__ void k_gmm_prepare(const float *__restrict__ raw, int64_t n, float *__restrict__ g) {
    const int64_t q = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    float o[PROP_RAW], r[PROP_RAW];
    for (int j = 0; j < PROP_RAW; ++j) o[j] = raw[q * PROP_RAW + j];
    gmm_prepare(o, r);
    for (int j = 0; j < PROP_RAW; ++j) g[q * PROP_RAW + j] = r[j];
}

// compact the capture sub-lists (sub s: min(count_s, cap_sub) records) into x (n x 9), uv
// (n x 2) and, if w is not null, the weights (n), sub-lists in order; one block per sub-list
__global__ void k_capture_gather(const float *__restrict__ cap, const unsigned long long *__restrict__ counts,
                                 unsigned long long cap_sub, float *__restrict__ x, float *__restrict__ uv,
                                 float *__restrict__ w) {
    const int sub = blockIdx.x;
    unsigned long long off = 0;
    for (int s = 0; s < sub; ++s) {
        const unsigned long long c = counts[s * SPREAD_STRIDE];
        off += c < cap_sub ? c : cap_sub;
    }
    const unsigned long long c0 = counts[sub * SPREAD_STRIDE], cnt = c0 < cap_sub ? c0 : cap_sub;
    for (unsigned long long i = threadIdx.x; i < cnt; i += blockDim.x) {
        const float *r = cap + ((size_t) sub * cap_sub + i) * 12;
        for (int j = 0; j < 9; ++j) x[(off + i) * 9 + j] = r[j];
        uv[(off + i) * 2] = r[9];
        uv[(off + i) * 2 + 1] = r[10];
        if (w) w[off + i] = r[11];
    }
}

// Deterministic order of the captured records: lanes append them in scheduling order, so
// k_capture_keys copies them into one array with a 64-bit hash of each record's twelve words,
// the records are sorted by that key (radix sort, stable), and k_capture_scatter writes them
// out in key order.  Equal records are interchangeable, so the output depends only on the set
// of records captured.
__global__ void k_capture_keys(const float *__restrict__ cap, const unsigned long long *__restrict__ counts,
                               unsigned long long cap_sub, float *__restrict__ rec, uint64_t *__restrict__ keys,
                               uint32_t *__restrict__ idx) {
    const int sub = blockIdx.x;
    unsigned long long off = 0;
    for (int s = 0; s < sub; ++s) {
        const unsigned long long c = counts[s * SPREAD_STRIDE];
        off += c < cap_sub ? c : cap_sub;
    }
    const unsigned long long c0 = counts[sub * SPREAD_STRIDE], cnt = c0 < cap_sub ? c0 : cap_sub;
    for (unsigned long long i = threadIdx.x; i < cnt; i += blockDim.x) {
        const float *r = cap + ((size_t) sub * cap_sub + i) * 12;
        uint64_t h = 0x9E3779B97F4A7C15ull;
        for (int j = 0; j < 12; ++j) {
            rec[(off + i) * 12 + j] = r[j];
            h = mix64(h ^ (uint64_t) __float_as_uint(r[j]));
        }
        keys[off + i] = h;
        idx[off + i] = (uint32_t) (off + i);
    }
}
__global__ void k_capture_scatter(const float *__restrict__ rec, const uint32_t *__restrict__ idx, int64_t n,
                                  float *__restrict__ x, float *__restrict__ uv, float *__restrict__ w) {
    const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *r = rec + (size_t) idx[i] * 12;
    for (int j = 0; j < 9; ++j) x[i * 9 + j] = r[j];
    uv[i * 2] = r[9];
    uv[i * 2 + 1] = r[10];
    if (w) w[i] = r[11];
}

} // namespace smsdev
