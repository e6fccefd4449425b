// sms_glint_kernels.hip -- unbiased glint SMS estimator (SpecularManifoldGlints::
// specular_manifold_sampling, src/librender/manifold_glints.cpp:28-181) as a pass sequence
// in which every Bernoulli trial chain, however long, finishes within a few kernel-lengths:
//
//   k_glint_first      one lane per query, 256 consecutive queries per block: zero the output,
//                      sample the emitter, test its visibility, compute the footprint's target
//                      slope; the queries that reach the solve are compacted inside the block
//                      (LDS) so whole wavefronts run first walks and empty ones retire; a
//                      converged first walk becomes a trial item (one atomic per block, on the
//                      block's sub-list of the spread trial list)
//   k_glint_trials_gG  G lanes per trial item run G trials side by side, `lane_trials` trials
//                      per item; most chains end here (the mean chain is about 3 trials)
//   k_glint_coop_rounds one wavefront per remaining item, 64 trials per round, `tail_rounds`
//                      rounds, the footprint's texels staged in LDS
//   k_glint_multiwave  the few long chains left get up to 64 wavefronts each, which claim
//                      64-trial chunks of their item; the first match is an atomic min and a
//                      per-trial log keeps the counters exact (k_glint_multiwave_resolve)
//   k_glint_coop_seq   one wavefront per item for chains longer than the multi-wave range
//   k_glint_finalize   one lane per trial item: contribution (Fresnel x generalized geometric
//                      term), sensor importance, output (f * (val * inv_prob)) * ew
//
// Work is assigned by static mappings over spread lists and counters are spread over copies
// (kernels.h), so no pass serialises thousands of wavefronts on one atomic address.  Walk k
// of query q draws from PCG32 stream (q, k + 1), so the result of every query equals the
// sequential trial loop whatever lane, group, wavefront or pass runs which trial.
#include "kernels.h"
#include "dev_ss_est.h"

namespace smsdev {

SMS_DEV void footprint_from_rec(const GlintRec &r, GlintFootprint &fp) {
    fp.target = v2(r.target[0], r.target[1]);
    fp.uv = v2(r.uv[0], r.uv[1]);
    fp.duv_dx = v2(r.duv_dx[0], r.duv_dx[1]);
    fp.duv_dy = v2(r.duv_dy[0], r.duv_dy[1]);
    footprint_setup(fp);
}

SMS_DEV void footprint_texels(const DScene &S, int shape, GlintFootprint &fp) {
    const DShape &gs = S.shapes[shape];
    const DNormalmap nm = S.nmaps[gs.normalmap];
    fp.slope = nm.slope; fp.res = nm.res; fp.stride = nm.stride; fp.fres = (float) nm.res; fp.tiles = gs.tiles;
    fp.gmm = nullptr; fp.defensive = 1.f;
}

// the learned proposal of query q on this footprint's shape, if one is set for it
SMS_DEV void footprint_proposal(const Proposal &P, int64_t q, int shape, GlintFootprint &fp) {
    if (P.gmm && shape == P.shape) { fp.gmm = P.gmm + q * PROP_RAW; fp.defensive = P.defensive; }
}

// The footprint sample xi of a glint walk (manifold_glints.cpp:252-259): two draws of the walk's
// stream, or with a learned proposal a draw of the query's defensive mixture over the footprint
// (the same mixture for the first walk and every Bernoulli trial); a mixture sample outside the
// footprint is a failed walk with 0 iterations (false).
SMS_DEV bool glint_seed(const GlintFootprint &fp, Pcg32 &rw, V2 &xi) {
    if (!fp.gmm) {
        xi.x = pcg32_float(rw);
        xi.y = pcg32_float(rw);
        return true;
    }
    return proposal_draw(fp.gmm, fp.defensive, rw, xi) == 0 ||
           (xi.x >= 0.f && xi.x < 1.f && xi.y >= 0.f && xi.y < 1.f);
}

SMS_DEV unsigned long long wave_sum_u64(unsigned long long v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}


// ---------------------------------------------------------------- pass 1: setup + first walks
template <bool BVH>
SMS_DEV void glint_first_body(const DScene &S, const GlintQuerySource &Q, int64_t n, uint64_t seed,
                              float *__restrict__ out, EstimatorWork W, GlintRec *lrec, int *wcount,
                              unsigned long long *lbase, unsigned long long (*lstats)[4]) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int64_t q = (int64_t) blockIdx.x * blockDim.x + tid;
    bool valid = false;
    unsigned long long st_walks = 0, st_conv = 0, st_iters = 0;
    GlintRec rec;
    if (q < n) {
        const sms_glint_query g = glint_query_fetch<BVH>(S, Q, q);
        write_rgb(out, q, v3(0.f, 0.f, 0.f));
        if (g.shape >= 0 && g.shape < S.n_shapes && S.shapes[g.shape].bsdf == SMS_BSDF_NORMALMAP) {
            SI si = glint_query_si(g);
            Pcg32 rq;
            stream_seed(rq, query_key(seed, (uint64_t) q), 0);
            EI ei = sample_emitter_interaction(S, si.p, S.emitters_all, S.n_emitters_all, rq);
            V3 edir = ei.p - si.p;
            float edist = norm(edir);
            edir = edir * (1.f / edist);
            bool occluded = ray_test_t<BVH>(S, make_ray(si.p, edir, RAY_EPS_F * (1.f + hmax_abs(si.p)), edist * (1.f - RAY_EPS_F)));
            if (!occluded) {
                V2 target;
                if (glint_target(ld3(g.sensor_p), ei.p, si.p, si.sh_s, si.sh_t, si.sh_n, v3(0.f, 0.f, 1.f), target)) {
                    valid = true;
                    rec.q = (int32_t) q;
                    rec.target[0] = target.x; rec.target[1] = target.y;
                    rec.uv[0] = g.uv[0]; rec.uv[1] = g.uv[1];
                    rec.duv_dx[0] = g.duv_dx[0]; rec.duv_dx[1] = g.duv_dx[1];
                    rec.duv_dy[0] = g.duv_dy[0]; rec.duv_dy[1] = g.duv_dy[1];
                    rec.shape = g.shape;
                } else {
                    st_walks = 1;          // sample_glint fails its side test before the solve
                }
            }
        }
    }
    // block-local compaction of the solvable queries
    unsigned long long ballot = __ballot(valid);
    if (lane == 0) wcount[wv] = (int) __popcll(ballot);
    __syncthreads();
    int base = 0, total = 0;
    for (int i = 0; i < 4; ++i) {
        int c = wcount[i];
        base += i < wv ? c : 0;
        total += c;
    }
    if (valid) lrec[base + (int) __popcll(ballot & ((1ull << lane) - 1ull))] = rec;
    __syncthreads();
    // first walks: slot tid of the compacted block
    bool conv = false;
    GlintRec r;
    V2 uv_sol = v2(0.f, 0.f);
    if (tid < total) {
        r = lrec[tid];
        GlintFootprint fp;
        footprint_texels(S, r.shape, fp);
        footprint_from_rec(r, fp);
        footprint_proposal(W.prop, r.q, r.shape, fp);
        Pcg32 rw;
        stream_seed(rw, query_key(seed, (uint64_t) r.q), 1);
        V2 xi;
        const bool seeded = glint_seed(fp, rw, xi);
        GlintWalk w;
        glint_walk_start(fp, xi, w);
        int st = WALK_FAILED;
        if (seeded)
            while ((st = glint_walk_iter(fp, S.cfg, w)) == WALK_RUNNING) {}
        st_walks += 1;
        st_iters += (unsigned long long) w.it;
        conv = st == WALK_CONVERGED;
        uv_sol = w.uv;
        if (conv && capturing(W.prop, r.shape, SMS_CAPTURE_SOLUTIONS)) {
            // training record: the solution as a footprint sample xi (uv = uv0 + (xi - 1/2) [duv_dx duv_dy])
            const sms_glint_query g = glint_query_fetch<BVH>(S, Q, r.q);
            const SI si = glint_query_si(g);
            Pcg32 rq;
            stream_seed(rq, query_key(seed, (uint64_t) r.q), 0);
            const EI ei = sample_emitter_interaction(S, si.p, S.emitters_all, S.n_emitters_all, rq);
            const V2 d = uv_sol - fp.uv;
            const float det = fp.duv_dx.x * fp.duv_dy.y - fp.duv_dy.x * fp.duv_dx.y;
            const V2 xi_sol = v2((d.x * fp.duv_dy.y - fp.duv_dy.x * d.y) / det + 0.5f,
                                 (fp.duv_dx.x * d.y - d.x * fp.duv_dx.y) / det + 0.5f);
            if (det != 0.f) capture_sample(S, W.prop, si, ei, xi_sol, 1.f);
        }
    }
    if (conv) st_conv = 1;
    // append converged first walks as trial items: one atomic per block, on the block's own
    // sub-list of the spread trial list; the block's counters go to its copy of the stats
    unsigned long long cb = __ballot(conv);
    __syncthreads();                                   // wcount reuse
    if (lane == 0) wcount[wv] = (int) __popcll(cb);
    unsigned long long sw = wave_sum_u64(st_walks), sc = wave_sum_u64(st_conv), si_ = wave_sum_u64(st_iters);
    if (lane == 0) { lstats[wv][0] = sw; lstats[wv][1] = sc; lstats[wv][2] = si_; }
    __syncthreads();
    const int sub = (int) (blockIdx.x % SPREAD);
    if (tid == 0) {
        int nb = wcount[0] + wcount[1] + wcount[2] + wcount[3];
        *lbase = nb ? spread_append(W.trials, sub, (unsigned long long) nb) : 0ull;
        unsigned long long *st = W.stats + sub * SPREAD_STRIDE;
        unsigned long long v0 = (lstats[0][0] + lstats[1][0]) + (lstats[2][0] + lstats[3][0]);
        unsigned long long v1 = (lstats[0][1] + lstats[1][1]) + (lstats[2][1] + lstats[3][1]);
        unsigned long long v2_ = (lstats[0][2] + lstats[1][2]) + (lstats[2][2] + lstats[3][2]);
        if (v0) atomicAdd(st + 0, v0);
        if (v1) atomicAdd(st + 1, v1);
        if (v2_) atomicAdd(st + 2, v2_);
        if (nb) atomicAdd(st + 3, (unsigned long long) nb);   // Bernoulli loops started (bernoulli_calls)
    }
    __syncthreads();
    if (conv) {
        int wb = 0;
        for (int i = 0; i < 4; ++i) wb += i < wv ? wcount[i] : 0;
        unsigned long long pos = *lbase + (unsigned long long) wb + (unsigned long long) __popcll(cb & ((1ull << lane) - 1ull));
        TrialRec t;
        t.f = r;
        t.uv_sol[0] = uv_sol.x; t.uv_sol[1] = uv_sol.y;
        W.trials.items[(size_t) sub * W.trials.sub_cap + pos] = t;
    }
}

#define SMS_GLINT_FIRST_KERNEL(NAME, BVH)                                                                    \
    __global__ void __launch_bounds__(256) NAME(DScene S, GlintQuerySource Q, int64_t n,                     \
                                                const uint64_t *__restrict__ seedp, float *__restrict__ out, \
                                                EstimatorWork W) {                                           \
        __shared__ GlintRec lrec[256];                                                                       \
        __shared__ int wcount[4];                                                                            \
        __shared__ unsigned long long lbase;                                                                 \
        __shared__ unsigned long long lstats[4][4];                                                          \
        glint_first_body<BVH>(S, Q, n, *seedp, out, W, lrec, wcount, &lbase, lstats);                        \
    }
SMS_GLINT_FIRST_KERNEL(k_glint_first_bvh, true)
SMS_GLINT_FIRST_KERNEL(k_glint_first_nobvh, false)

// ---------------------------------------------------------------- trial walks
// Trial k (>= 1) of a trial item: the walk from stream (q, k + 1).  Returns the walk status.
template <class Fetch>
SMS_DEV int glint_trial(const GlintFootprint &fp, const sms_config &cfg, uint64_t qkey, int64_t k, GlintWalk &w,
                        const Fetch &fetch) {
    Pcg32 rw;
    stream_seed(rw, qkey, (uint64_t) k + 1);
    V2 xi;
    const bool seeded = glint_seed(fp, rw, xi);
    glint_walk_start(fp, xi, w);
    if (!seeded) return WALK_FAILED;
    int s;
    while ((s = glint_walk_iter_f(fp, cfg, w, fetch)) == WALK_RUNNING) {}
    return s;
}

SMS_DEV int64_t trial_cap(const sms_config &cfg) {
    return cfg.max_trials > 0 ? (int64_t) cfg.max_trials : (int64_t) 1 << 40;
}

SMS_DEV void append_item(const WorkList &L, const OverflowRec &r) {
    unsigned long long slot = atomicAdd(L.count, 1ull);
    L.items[slot] = r;
}

SMS_DEV OverflowRec make_item(int64_t item, int64_t k, V2 uv_sol) {
    OverflowRec ov;
    ov.q = item; ov.k = k; ov.sol[0] = uv_sol.x; ov.sol[1] = uv_sol.y; ov.sol[2] = 0.f;
    ov.shapes[0] = ov.shapes[1] = ov.shapes[2] = ov.shapes[3] = -1; ov.pad = 0;
    return ov;
}


// ---------------------------------------------------------------- pass 2: grouped trials
// G lanes per trial item run trials k, k+1, ..., k+G-1 side by side; the first match of a
// round comes from the group's ballot bits.  A group runs ceil(lane_trials / G) rounds; items
// still open go to the overflow list (the sub-list of their trial slot).  Groups walk the
// trial list by a static mapping (no claim counter).
template <int G>
SMS_DEV void glint_trials_body(const DScene &S, uint64_t seed, EstimatorWork W) {
    const sms_config &cfg = S.cfg;
    const int lane = threadIdx.x & 63;
    const int gl = lane & (G - 1), gfirst = lane & ~(G - 1);
    const unsigned long long gmask = G == 64 ? ~0ull : ((1ull << G) - 1ull);
    const unsigned long long gid = ((unsigned long long) blockIdx.x * blockDim.x + threadIdx.x) / G;
    const unsigned long long n_groups = ((unsigned long long) gridDim.x * blockDim.x) / G;
    const int64_t kcap = trial_cap(cfg);
    const int rounds = (W.lane_trials + G - 1) / G;
    unsigned long long st_walks = 0, st_conv = 0, st_iters = 0, st_btrials = 0, st_bmax = 0;
    const SpreadWalk sw = spread_walk(W.trials, gid, n_groups);
    for (unsigned long long p = sw.pos; p < sw.count; p += sw.step) {
        const unsigned long long slot = (unsigned long long) sw.sub * W.trials.sub_cap + p;
        const TrialRec t = W.trials.items[slot];
        GlintFootprint fp;
        footprint_texels(S, t.f.shape, fp);
        footprint_from_rec(t.f, fp);
        footprint_proposal(W.prop, t.f.q, t.f.shape, fp);
        GlobalTexels tex; tex.slope = fp.slope; tex.res = fp.res; tex.stride = fp.stride;
        const uint64_t qkey = query_key(seed, (uint64_t) t.f.q);
        const V2 uv_sol = v2(t.uv_sol[0], t.uv_sol[1]);
        int64_t found = -1, base = 1;
        for (int r = 0; r < rounds && found < 0 && base <= kcap; ++r, base += G) {
            const int64_t k = base + gl;
            bool match = false, ran = false, conv = false;
            int its = 0;
            if (k <= kcap) {
                GlintWalk w;
                int s = glint_trial(fp, cfg, qkey, k, w, tex);
                ran = true; its = w.it; conv = s == WALK_CONVERGED;
                match = conv && norm2(w.uv - uv_sol) < cfg.uniqueness_threshold;
            }
            unsigned long long b = (__ballot(match) >> gfirst) & gmask;
            if (b) found = base + (int64_t) __ffsll((long long) b) - 1;
            if (ran && (found < 0 || k <= found)) {
                st_walks++; st_iters += (unsigned long long) its;
                if (conv) st_conv++;
            }
        }
        if (gl == 0) {
            if (found >= 0 || base > kcap) {
                unsigned long long tr = found >= 0 ? (unsigned long long) found : (unsigned long long) kcap + 1ull;
                st_btrials += tr; st_bmax = st_bmax > tr ? st_bmax : tr;
                GlintRes o; o.inv_prob = found >= 0 ? (float) found : 0.f; o.status = 1;
                W.res[slot] = o;
            } else {
                unsigned long long pos = spread_append(W.over, sw.sub, 1ull);
                W.over.items[(size_t) sw.sub * W.over.sub_cap + pos] = make_item((int64_t) slot, base, uv_sol);
                GlintRes o; o.inv_prob = 0.f; o.status = 2;
                W.res[slot] = o;
            }
        }
    }
    stat_add(W.stats + 0, st_walks); stat_add(W.stats + 1, st_conv); stat_add(W.stats + 2, st_iters);
    stat_add(W.stats + 4, st_btrials); stat_max(W.stats + 5, st_bmax); stat_add(W.stats + 6, st_iters);
}

#define SMS_GLINT_TRIALS_KERNEL(G)                                                                                              \
    __global__ void __launch_bounds__(256) k_glint_trials_g##G(DScene S, const uint64_t *__restrict__ seedp, EstimatorWork W) { \
        glint_trials_body<G>(S, *seedp, W);                                                                                     \
    }
SMS_GLINT_TRIALS_KERNEL(1)
SMS_GLINT_TRIALS_KERNEL(2)
SMS_GLINT_TRIALS_KERNEL(4)
SMS_GLINT_TRIALS_KERNEL(8)
SMS_GLINT_TRIALS_KERNEL(16)

// ---------------------------------------------------------------- one-wavefront tail passes
// The 64 lanes of a wavefront iterate inside one footprint, so its texels (plus the bilinear
// margin) are staged in LDS when they fit; otherwise the walk reads the global slope array.
constexpr int COOP_TILE = 24;                 // texels per side of the staged window

// Stage the texel window of a footprint into this wavefront's tile; false if it does not fit.
SMS_DEV bool stage_window(const GlintFootprint &fp, float tiles, float2 (*tile)[COOP_TILE + 1], int lane,
                          LdsTexels<COOP_TILE> &lds) {
    lds.tile = tile;
    lds.g.slope = fp.slope; lds.g.res = fp.res; lds.g.stride = fp.stride;
    bool staged = false;
    if (tiles == 1.f) {
        float ex = 0.5f * (fabsf(fp.duv_dx.x) + fabsf(fp.duv_dy.x)), ey = 0.5f * (fabsf(fp.duv_dx.y) + fabsf(fp.duv_dy.y));
        float u0 = fp.uv.x - ex, u1 = fp.uv.x + ex, v0 = fp.uv.y - ey, v1 = fp.uv.y + ey;
        if (u0 > 0.f && v0 > 0.f && u1 < 1.f && v1 < 1.f) {
            int x0 = (int) floorf(u0 * fp.fres - 0.5f) - 1, y0 = (int) floorf(v0 * fp.fres - 0.5f) - 1;
            int x1 = (int) floorf(u1 * fp.fres - 0.5f) + 2, y1 = (int) floorf(v1 * fp.fres - 0.5f) + 2;
            x0 = x0 < 0 ? 0 : x0; y0 = y0 < 0 ? 0 : y0;
            if (x1 - x0 < COOP_TILE && y1 - y0 < COOP_TILE) {
                for (int t = lane; t < (COOP_TILE + 1) * (COOP_TILE + 1); t += 64) {
                    int ty = t / (COOP_TILE + 1), tx = t % (COOP_TILE + 1);
                    int gx = x0 + tx, gy = y0 + ty;
                    gx = gx > fp.res ? fp.res : gx; gy = gy > fp.res ? fp.res : gy;   // pad column/row
                    tile[ty][tx] = fp.slope[(size_t) gy * fp.stride + gx];
                }
                lds.x0 = x0; lds.y0 = y0;
                staged = true;
            }
        }
    }
    // The window belongs to this wavefront only (waves of a block leave their item loops at
    // different times, so no block barrier): order its LDS writes before the reads.
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    return staged;
}

SMS_DEV void window_done() {
    // every lane is done reading the window before it is overwritten
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}


// One wavefront per overflow item: lane j runs trial k0 + j, at most `max_rounds` rounds of 64
// (negative = until resolved).  Unresolved items go to `out` with the next trial index.
SMS_DEV void glint_coop_item(const DScene &S, uint64_t seed, EstimatorWork W, const OverflowRec &ov, int max_rounds,
                             const WorkList &out, bool init_long, float2 (*tile)[COOP_TILE + 1],
                             unsigned long long st[5]) {
    const sms_config &cfg = S.cfg;
    const int lane = threadIdx.x & 63;
    const int64_t kcap = trial_cap(cfg);
    const TrialRec tr = W.trials.items[ov.q];
    GlintFootprint fp;
    footprint_texels(S, tr.f.shape, fp);
    footprint_from_rec(tr.f, fp);
    footprint_proposal(W.prop, tr.f.q, tr.f.shape, fp);
    LdsTexels<COOP_TILE> lds;
    bool staged = stage_window(fp, S.shapes[tr.f.shape].tiles, tile, lane, lds);
    GlobalTexels tex; tex.slope = fp.slope; tex.res = fp.res; tex.stride = fp.stride;
    const uint64_t qkey = query_key(seed, (uint64_t) tr.f.q);
    const V2 uv_sol = v2(ov.sol[0], ov.sol[1]);
    int64_t found = -1, base = ov.k;
    for (int round = 0; base <= kcap && found < 0 && (max_rounds < 0 || round < max_rounds); base += 64, ++round) {
        const int64_t k = base + lane;
        bool match = false, ran = false, conv = false;
        int its = 0;
        if (k <= kcap) {
            GlintWalk w;
            int s = staged ? glint_trial(fp, cfg, qkey, k, w, lds) : glint_trial(fp, cfg, qkey, k, w, tex);
            ran = true; its = w.it; conv = s == WALK_CONVERGED;
            match = conv && norm2(w.uv - uv_sol) < cfg.uniqueness_threshold;
        }
        unsigned long long ballot = __ballot(match);
        if (ballot) found = base + (int64_t) __ffsll((long long) ballot) - 1;
        // count only the walks the sequential trial loop would have run (k <= first match)
        if (ran && (found < 0 || k <= found)) {
            st[0]++; st[2] += (unsigned long long) its;
            if (conv) st[1]++;
        }
    }
    window_done();
    if (lane == 0) {
        if (found >= 0 || base > kcap) {
            W.res[ov.q].inv_prob = found >= 0 ? (float) found : 0.f;
            W.res[ov.q].status = 1;
            unsigned long long trials = found >= 0 ? (unsigned long long) found : (unsigned long long) kcap + 1ull;
            st[3] += trials;
            st[4] = st[4] > trials ? st[4] : trials;
        } else {
            unsigned long long slot = atomicAdd(out.count, 1ull);
            out.items[slot] = make_item(ov.q, base, uv_sol);
            if (init_long && slot < (unsigned long long) W.mw_items) {
                W.mw_found[slot] = 0x7fffffffffffffffLL;
                W.mw_next[slot] = 0ull;
            }
        }
    }
}

SMS_DEV void coop_stats(EstimatorWork W, const unsigned long long st[5]) {
    stat_add(W.stats + 0, st[0]); stat_add(W.stats + 1, st[1]); stat_add(W.stats + 2, st[2]);
    stat_add(W.stats + 4, st[3]); stat_max(W.stats + 5, st[4]); stat_add(W.stats + 6, st[2]);
}

// overflow items, one wavefront each, by a static mapping over the spread overflow list
__global__ void __launch_bounds__(256) k_glint_coop_rounds(DScene S, const uint64_t *__restrict__ seedp, EstimatorWork W) {
    const uint64_t seed = *seedp;
    __shared__ float2 tile[4][COOP_TILE + 1][COOP_TILE + 1];
    unsigned long long st[5] = {0ull, 0ull, 0ull, 0ull, 0ull};
    const unsigned long long wave = ((unsigned long long) blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const unsigned long long n_waves = ((unsigned long long) gridDim.x * blockDim.x) >> 6;
    const SpreadWalk sw = spread_walk(W.over, wave, n_waves);
    for (unsigned long long p = sw.pos; p < sw.count; p += sw.step) {
        const OverflowRec ov = W.over.items[(size_t) sw.sub * W.over.sub_cap + p];
        glint_coop_item(S, seed, W, ov, W.tail_rounds, W.long_list, true, tile[threadIdx.x >> 6], st);
    }
    coop_stats(W, st);
}

// chains longer than the multi-wavefront range, one wavefront each
__global__ void __launch_bounds__(256) k_glint_coop_seq(DScene S, const uint64_t *__restrict__ seedp, EstimatorWork W) {
    const uint64_t seed = *seedp;
    __shared__ float2 tile[4][COOP_TILE + 1][COOP_TILE + 1];
    unsigned long long st[5] = {0ull, 0ull, 0ull, 0ull, 0ull};
    const unsigned long long n_items = *W.seq_list.count;
    const unsigned long long n_waves = ((unsigned long long) gridDim.x * blockDim.x) >> 6;
    for (unsigned long long i = ((unsigned long long) blockIdx.x * blockDim.x + threadIdx.x) >> 6; i < n_items; i += n_waves)
        glint_coop_item(S, seed, W, W.seq_list.items[i], -1, W.seq_list, false, tile[threadIdx.x >> 6], st);
    coop_stats(W, st);
}

// ---------------------------------------------------------------- multi-wavefront stage
// The few items with long trial chains get several wavefronts each: a wavefront claims the
// item's next 64-trial chunk (per-item counter), runs it and lowers the item's first match
// (atomic min).  Chunks past the current first match are not claimed; chunks that ran past
// the final one are discarded by the resolve pass, which sums the per-trial log only up to
// the first match, so the counters equal the sequential trial loop's.
__global__ void __launch_bounds__(256) k_glint_multiwave(DScene S, const uint64_t *__restrict__ seedp, EstimatorWork W) {
    const uint64_t seed = *seedp;
    __shared__ float2 tile_all[4][COOP_TILE + 1][COOP_TILE + 1];
    const sms_config &cfg = S.cfg;
    const int lane = threadIdx.x & 63;
    float2 (*tile)[COOP_TILE + 1] = tile_all[threadIdx.x >> 6];
    const unsigned long long n_long = *W.long_list.count;
    const int n = (int) (n_long < (unsigned long long) W.mw_items ? n_long : (unsigned long long) W.mw_items);
    if (n == 0) return;
    const int wave = (int) ((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int n_waves = (int) ((gridDim.x * blockDim.x) >> 6);
    int per_item = n_waves / n;
    per_item = per_item < 1 ? 1 : (per_item > 64 ? 64 : per_item);
    if (wave >= n * per_item) return;
    const int64_t kcap = trial_cap(cfg);
    int j = wave % n, exhausted = 0, staged_item = -1;
    bool staged = false;
    LdsTexels<COOP_TILE> lds;
    GlintFootprint fp;
    GlobalTexels tex;
    uint64_t qkey = 0;
    V2 uv_sol = v2(0.f, 0.f);
    int64_t kstart = 0;
    while (exhausted < n) {
        if (staged_item != j) {
            if (staged_item >= 0) window_done();
            const OverflowRec ov = W.long_list.items[j];
            const TrialRec tr = W.trials.items[ov.q];
            footprint_texels(S, tr.f.shape, fp);
            footprint_from_rec(tr.f, fp);
            footprint_proposal(W.prop, tr.f.q, tr.f.shape, fp);
            staged = stage_window(fp, S.shapes[tr.f.shape].tiles, tile, lane, lds);
            tex.slope = fp.slope; tex.res = fp.res; tex.stride = fp.stride;
            qkey = query_key(seed, (uint64_t) tr.f.q);
            uv_sol = v2(ov.sol[0], ov.sol[1]);
            kstart = ov.k;
            staged_item = j;
        }
        unsigned long long c = 0;
        if (lane == 0) c = atomicAdd(&W.mw_next[j], 1ull);
        c = __shfl(c, 0, 64);
        long long found_now = __hip_atomic_load(&W.mw_found[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int64_t lim = kstart + (int64_t) W.mw_trials - 1;
        lim = lim < kcap ? lim : kcap;
        lim = lim < (int64_t) found_now ? lim : (int64_t) found_now;
        const int64_t k0 = kstart + 64 * (int64_t) c;
        if (k0 > lim) {                          // nothing left to claim on this item
            j = j + 1 == n ? 0 : j + 1;
            exhausted++;
            continue;
        }
        exhausted = 0;
        const int64_t k = k0 + lane;
        bool match = false;
        if (k <= kcap && k < kstart + (int64_t) W.mw_trials) {
            GlintWalk w;
            int s = staged ? glint_trial(fp, cfg, qkey, k, w, lds) : glint_trial(fp, cfg, qkey, k, w, tex);
            bool conv = s == WALK_CONVERGED;
            match = conv && norm2(w.uv - uv_sol) < cfg.uniqueness_threshold;
            W.mw_log[(size_t) j * (size_t) W.mw_trials + (size_t) (k - kstart)] =
                (uint16_t) ((unsigned) w.it | (conv ? 0x8000u : 0u));
        }
        unsigned long long ballot = __ballot(match);
        if (ballot && lane == 0) atomicMin(&W.mw_found[j], (long long) (k0 + (int64_t) __ffsll((long long) ballot) - 1));
    }
    if (staged_item >= 0) window_done();
}

// One wavefront per long item: sum the log up to the first match (or the end of the stage's
// range), resolve the item or pass it to the sequential tail.
__global__ void __launch_bounds__(256) k_glint_multiwave_resolve(DScene S, EstimatorWork W) {
    const sms_config &cfg = S.cfg;
    const int lane = threadIdx.x & 63;
    const unsigned long long n_long = *W.long_list.count;
    const int64_t kcap = trial_cap(cfg);
    unsigned long long st_walks = 0, st_conv = 0, st_iters = 0, st_btrials = 0, st_bmax = 0;
    const unsigned long long n_waves = (gridDim.x * blockDim.x) >> 6;
    for (unsigned long long j = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; j < n_long; j += n_waves) {
        const OverflowRec ov = W.long_list.items[j];
        if (j >= (unsigned long long) W.mw_items) {
            if (lane == 0) append_item(W.seq_list, ov);
            continue;
        }
        const long long found = W.mw_found[j];
        const int64_t kstart = ov.k;
        int64_t last = kstart + (int64_t) W.mw_trials - 1;
        last = last < kcap ? last : kcap;
        const bool matched = found != 0x7fffffffffffffffLL;
        const int64_t upto = matched ? (int64_t) found : last;
        unsigned long long w_its = 0, w_conv = 0;
        for (int64_t t = lane; t <= upto - kstart; t += 64) {
            unsigned v = W.mw_log[(size_t) j * (size_t) W.mw_trials + (size_t) t];
            w_its += v & 0x7fffu;
            w_conv += v >> 15;
        }
        if (lane == 0) {
            st_walks += (unsigned long long) (upto - kstart + 1);
            if (matched || last == kcap) {
                W.res[ov.q].inv_prob = matched ? (float) found : 0.f;
                W.res[ov.q].status = 1;
                unsigned long long trials = matched ? (unsigned long long) found : (unsigned long long) kcap + 1ull;
                st_btrials += trials;
                st_bmax = st_bmax > trials ? st_bmax : trials;
            } else {
                OverflowRec nx = ov;
                nx.k = last + 1;
                append_item(W.seq_list, nx);
            }
        }
        st_iters += w_its;
        st_conv += w_conv;
    }
    stat_add(W.stats + 0, st_walks); stat_add(W.stats + 1, st_conv); stat_add(W.stats + 2, st_iters);
    stat_add(W.stats + 4, st_btrials); stat_max(W.stats + 5, st_bmax); stat_add(W.stats + 6, st_iters);
}


// ---------------------------------------------------------------- contributions
__global__ void __launch_bounds__(256) k_glint_finalize(DScene S, GlintQuerySource Q, const uint64_t *__restrict__ seedp,
                                                        float *__restrict__ out, EstimatorWork W) {
    const uint64_t seed = *seedp;
    const sms_config &cfg = S.cfg;
    const unsigned long long u = (unsigned long long) blockIdx.x * blockDim.x + threadIdx.x;
    const SpreadWalk sw = spread_walk(W.trials, u, (unsigned long long) gridDim.x * blockDim.x);
    for (unsigned long long p = sw.pos; p < sw.count; p += sw.step) {
        const unsigned long long slot = (unsigned long long) sw.sub * W.trials.sub_cap + p;
        const GlintRes res = W.res[slot];
        const TrialRec tr = W.trials.items[slot];
        const int64_t q = tr.f.q;
        const sms_glint_query g = glint_query_fetch(S, Q, q);
        SI si = glint_query_si(g);
        Pcg32 rq;
        stream_seed(rq, query_key(seed, (uint64_t) q), 0);
        EI ei = sample_emitter_interaction(S, si.p, S.emitters_all, S.n_emitters_all, rq);
        V3 sensor_p = ld3(g.sensor_p);
        float val = glint_contribution(S, sensor_p, ei, si, v2(tr.uv_sol[0], tr.uv_sol[1]));
        float We = sensor_we(S.cam, sensor_p, si.p);
        // MIS against the BSDF strategy (manifold_glints.cpp:155-179): p_sms is the emitter
        // sampling density, p_bsdf the LEAN pdf of the emitter direction plus 1
        V3 edir = ei.p - si.p;
        float edist = norm(edir);
        edir = edir * (1.f / edist);
        float p_sms = sms_emitter_pdf(S, ei, si.p);
        V3 wo_l = v3(dot(edir, si.sh_s), dot(edir, si.sh_t), dot(edir, si.sh_n));
        float p_bsdf = lean_pdf(S, g.shape, v2(g.uv[0], g.uv[1]), v2(g.duv_dx[0], g.duv_dx[1]),
                                v2(g.duv_dy[0], g.duv_dy[1]), ld3(g.wi), wo_l);
        p_bsdf += 1.f;
        float mis = mis_power(p_sms, p_bsdf);
        if (cfg.bsdf_strategy_only) mis = 0.f;
        write_rgb(out, q, glint_result(val, res.inv_prob, mis * We, ei.weight));
    }
}

// ---------------------------------------------------------------- biased estimator
// Biased glint SMS (manifold_glints.cpp:112-140): one wavefront per query; lane m of a round
// runs walk m of the trial set (stream (q, m + 1), as the unbiased trial k); converged walks
// are de-duplicated in m order by uv distance through an LDS list of the solutions so far and
// the contributions of the new ones are summed in m order.  MIS and sensor importance as the
// unbiased finalize pass.
constexpr int GLINT_BIASED_MAX_SOL = 128;
template <bool BVH>
SMS_DEV void glint_biased_body(const DScene &S, const GlintQuerySource &Q, int64_t n, uint64_t seed,
                               float *__restrict__ out, EstimatorWork W, float2 (*sol_lds)[GLINT_BIASED_MAX_SOL]) {
    const sms_config &cfg = S.cfg;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    unsigned long long st_walks = 0, st_conv = 0, st_iters = 0;
    const int64_t n_waves = ((int64_t) gridDim.x * blockDim.x) >> 6;
    for (int64_t q = ((int64_t) blockIdx.x * blockDim.x + threadIdx.x) >> 6; q < n; q += n_waves) {
        const sms_glint_query g = glint_query_fetch<BVH>(S, Q, q);
        V3 result = v3(0.f, 0.f, 0.f);
        if (g.shape >= 0 && g.shape < S.n_shapes && S.shapes[g.shape].bsdf == SMS_BSDF_NORMALMAP) {
            SI si = glint_query_si(g);
            const uint64_t qkey = query_key(seed, (uint64_t) q);
            Pcg32 rq;
            stream_seed(rq, qkey, 0);
            EI ei = sample_emitter_interaction(S, si.p, S.emitters_all, S.n_emitters_all, rq);
            V3 edir = ei.p - si.p;
            float edist = norm(edir);
            edir = edir * (1.f / edist);
            const bool occluded =
                ray_test_t<BVH>(S, make_ray(si.p, edir, RAY_EPS_F * (1.f + hmax_abs(si.p)), edist * (1.f - RAY_EPS_F)));
            if (!occluded) {
                const V3 sensor_p = ld3(g.sensor_p);
                GlintFootprint fp;
                // sample_glint's side test fails every walk of the set before its solve
                const bool solvable = glint_target(sensor_p, ei.p, si.p, si.sh_s, si.sh_t, si.sh_n, v3(0.f, 0.f, 1.f), fp.target);
                if (solvable) {
                    footprint_texels(S, g.shape, fp);
                    fp.uv = v2(g.uv[0], g.uv[1]);
                    fp.duv_dx = v2(g.duv_dx[0], g.duv_dx[1]);
                    fp.duv_dy = v2(g.duv_dy[0], g.duv_dy[1]);
                    footprint_setup(fp);
                    footprint_proposal(W.prop, q, g.shape, fp);
                }
                float sum = 0.f;
                int n_sol = 0;
                for (int base = 0; base < cfg.max_trials; base += 64) {
                    const int m = base + lane;
                    bool cand = false;
                    V2 uvf = v2(0.f, 0.f);
                    if (m < cfg.max_trials) {
                        st_walks++;
                        if (solvable) {
                            Pcg32 rw;
                            stream_seed(rw, qkey, (uint64_t) m + 1);
                            V2 xi;
                            const bool seeded = glint_seed(fp, rw, xi);
                            GlintWalk w;
                            glint_walk_start(fp, xi, w);
                            int stw = WALK_FAILED;
                            if (seeded)
                                while ((stw = glint_walk_iter(fp, cfg, w)) == WALK_RUNNING) {}
                            st_iters += (unsigned long long) w.it;
                            if (stw == WALK_CONVERGED) {
                                st_conv++;
                                cand = true;
                                uvf = w.uv;
                            }
                        }
                    }
                    // in-order de-duplication against the solution list
                    bool unique = false;
                    unsigned long long cmask = __ballot(cand);
                    while (cmask) {
                        const int j = __ffsll((long long) cmask) - 1;
                        cmask &= cmask - 1ull;
                        const V2 uj = v2(__shfl(uvf.x, j, 64), __shfl(uvf.y, j, 64));
                        bool dup = false;
                        for (int s0 = 0; s0 < n_sol; s0 += 64) {
                            const int s = s0 + lane;
                            bool d = false;
                            if (s < n_sol) {
                                const float2 o = sol_lds[wv][s];
                                d = norm2(uj - v2(o.x, o.y)) < cfg.uniqueness_threshold;
                            }
                            if (__ballot(d)) { dup = true; break; }
                        }
                        if (!dup) {
                            if (lane == 0) sol_lds[wv][n_sol] = make_float2(uj.x, uj.y);
                            n_sol++;
                            if (lane == j) unique = true;
                        }
                        __builtin_amdgcn_wave_barrier();
                    }
                    const float contrib = unique ? glint_contribution(S, sensor_p, ei, si, uvf) : 0.f;
                    unsigned long long umask = __ballot(unique);
                    while (umask) {
                        const int j = __ffsll((long long) umask) - 1;
                        umask &= umask - 1ull;
                        sum += __shfl(contrib, j, 64);
                    }
                }
                if (n_sol > 0) {
                    const float We = sensor_we(S.cam, sensor_p, si.p);
                    const float p_sms = sms_emitter_pdf(S, ei, si.p);
                    const V3 wo_l = v3(dot(edir, si.sh_s), dot(edir, si.sh_t), dot(edir, si.sh_n));
                    float p_bsdf = lean_pdf(S, g.shape, v2(g.uv[0], g.uv[1]), v2(g.duv_dx[0], g.duv_dx[1]),
                                            v2(g.duv_dy[0], g.duv_dy[1]), ld3(g.wi), wo_l);
                    p_bsdf += 1.f;
                    float mis = mis_power(p_sms, p_bsdf);
                    if (cfg.bsdf_strategy_only) mis = 0.f;
                    result = glint_result(sum, 1.f, mis * We, ei.weight);
                }
            }
        }
        if (lane == 0) write_rgb(out, q, result);
    }
    stat_add(W.stats + 0, st_walks); stat_add(W.stats + 1, st_conv); stat_add(W.stats + 2, st_iters);
}

#define SMS_GLINT_BIASED_KERNEL(NAME, BVH)                                                                   \
    __global__ void __launch_bounds__(256) NAME(DScene S, GlintQuerySource Q, int64_t n,                     \
                                                const uint64_t *__restrict__ seedp, float *__restrict__ out, \
                                                EstimatorWork W) {                                           \
        __shared__ float2 sol_lds[4][GLINT_BIASED_MAX_SOL];                                                  \
        glint_biased_body<BVH>(S, Q, n, *seedp, out, W, sol_lds);                                            \
    }
SMS_GLINT_BIASED_KERNEL(k_glint_biased_bvh, true)
SMS_GLINT_BIASED_KERNEL(k_glint_biased_nobvh, false)

// Inputs of the learned proposal for glint queries: shading point, shading normal and the glint
// estimator's light sample (stream (q, 0) over all emitters); zeros for queries without a glinty hit.
__global__ void k_glint_prop_inputs(DScene S, GlintQuerySource Q, int64_t n,
                                    const uint64_t *__restrict__ seedp, float *__restrict__ x) {
    const int64_t q = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const sms_glint_query g = glint_query_fetch(S, Q, q);
    float *r = x + 9 * q;
    if (g.shape < 0) {
        for (int k = 0; k < 9; ++k) r[k] = 0.f;
        return;
    }
    const SI si = glint_query_si(g);
    Pcg32 rq;
    stream_seed(rq, query_key(*seedp, (uint64_t) q), 0);
    const EI ei = sample_emitter_interaction(S, si.p, S.emitters_all, S.n_emitters_all, rq);
    proposal_inputs(S, si.p, si.sh_n, ei.p, r);
}

// ---------------------------------------------------------------- BSDF-sampling strategy
// manifold_glints.cpp:183-220, one lane per query, added to the SMS strategy's result in out:
// an offset uv inside the pixel footprint (PCG32 stream GLINT_BSDF_STREAM of the query), the
// nested smooth conductor's mirror direction in the normal map's level-0 frame there
// (bsdfs/normalmap.cpp:118-136, conductor.cpp:217-275), kept only if it stays on the same side
// of the unperturbed frame, one ray (Interaction::spawn_ray); a hit on the front of an area
// emitter adds mis(p_bsdf, p_sms) * F * L (area.cpp:71-79).  Only an area emitter can receive
// this strategy, so the host launches it for scenes that have one.
constexpr uint64_t GLINT_BSDF_STREAM = 0xFFFFFFFEull;
__global__ void __launch_bounds__(256) k_glint_bsdf(DScene S, GlintQuerySource Q, int64_t n,
                                                    const uint64_t *__restrict__ seedp, float *__restrict__ out) {
    const int64_t q = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const sms_glint_query g = glint_query_fetch(S, Q, q);
    if (g.shape < 0 || g.shape >= S.n_shapes || S.shapes[g.shape].bsdf != SMS_BSDF_NORMALMAP) return;
    Pcg32 r;
    stream_seed(r, query_key(*seedp, (uint64_t) q), GLINT_BSDF_STREAM);
    const float fx = pcg32_float(r), fy = pcg32_float(r);
    const V2 duv_dx = v2(g.duv_dx[0], g.duv_dx[1]), duv_dy = v2(g.duv_dy[0], g.duv_dy[1]);
    const V2 duv = (fx - 0.5f) * duv_dx + (fy - 0.5f) * duv_dy;
    SI so = glint_query_si(g);
    so.uv = so.uv + duv;
    const MV f = mv_from_si(S, so);                       // NormalmapBSDF::frame at the offset uv
    const V3 wi = ld3(g.wi);
    const V3 wi_w = to_world(so.sh_s, so.sh_t, so.sh_n, wi);
    const V3 wi_p = v3(dot(wi_w, f.s), dot(wi_w, f.t), dot(wi_w, f.n));
    if (!(wi_p.z > 0.f)) return;
    const V3 wo_p = v3(-wi_p.x, -wi_p.y, wi_p.z);
    const DShape &sh = S.shapes[g.shape];
    const float F = fresnel_conductor(wi_p.z, sh.eta_r, sh.k);
    const V3 wo_w = to_world(f.s, f.t, f.n, wo_p);
    const V3 wo = v3(dot(wo_w, so.sh_s), dot(wo_w, so.sh_t), dot(wo_w, so.sh_n));
    if (!(wo_p.z * wo.z > 0.f)) return;
    const V3 d = to_world(so.sh_s, so.sh_t, so.sh_n, wo);
    const SI h = ray_intersect(S, spawn_ray(so.p, d));
    if (!h.valid) return;
    const int e = S.shapes[h.shape].emitter;
    if (e < 0 || S.emitters[e].type != SMS_EMITTER_AREA) return;
    float p_bsdf = lean_pdf(S, g.shape, so.uv, duv_dx, duv_dy, wi, wo);
    p_bsdf += 1.f;
    EI eh;
    eh.p = h.p; eh.n = h.sh_n;
    const V3 dh = h.p - so.p;
    eh.d = dh * (1.f / norm(dh));
    eh.weight = v3(0.f, 0.f, 0.f); eh.pdf = 0.f; eh.kind = SMS_EI_AREA; eh.emitter = e;
    const float p_sms = sms_emitter_pdf(S, eh, so.p);
    float mis = mis_power(p_bsdf, p_sms);
    if (S.cfg.sms_strategy_only) mis = 0.f;
    if (!(dot(d * -1.f, h.sh_n) > 0.f)) return;
    const DEmitter &em = S.emitters[e];
    const float w = mis * F;
    out[3 * q] = out[3 * q] + w * em.radiance[0];
    out[3 * q + 1] = out[3 * q + 1] + w * em.radiance[1];
    out[3 * q + 2] = out[3 * q + 2] + w * em.radiance[2];
}

} // namespace smsdev
