// sms_glint_kernels.hip -- unbiased glint SMS estimator (SpecularManifoldGlints::
// specular_manifold_sampling, src/librender/manifold_glints.cpp:28-181) as four passes:
//
//   k_glint_setup     one lane per query: zero the output, sample the emitter, test its
//                     visibility, compute the footprint's target slope; queries that reach the
//                     solve are compacted (one atomic per wavefront) into GlintRec records
//   k_glint_walks     persistent: every lane owns a record and advances its current walk one
//                     Newton iteration per loop trip -- the first walk, then Bernoulli trials
//                     until one lands on the same solution; a lane whose query needs more than
//                     `lane_trials` trials hands it to the overflow list and refills
//   k_glint_coop      tail pass: one wavefront per overflow record; lane j runs trial k0 + j,
//                     the footprint's texels are staged in LDS, the first match comes from a
//                     ballot
//   k_glint_finalize  one lane per record with a solution: contribution (Fresnel x generalized
//                     geometric term), sensor importance, output (f * (val * inv_prob)) * ew
//
// Walk k of query q draws from PCG32 stream (q, k + 1), so the result of every query equals
// the sequential trial loop whatever lane or pass runs which trial.
#include "kernels.h"

namespace smsdev {

SMS_DEV void footprint_from_rec(const GlintRec &r, GlintFootprint &fp) {
    fp.target = v2(r.target[0], r.target[1]);
    fp.uv = v2(r.uv[0], r.uv[1]);
    fp.duv_dx = v2(r.duv_dx[0], r.duv_dx[1]);
    fp.duv_dy = v2(r.duv_dy[0], r.duv_dy[1]);
    footprint_setup(fp);
}

// ---------------------------------------------------------------- pass 1: setup + compaction
__global__ void __launch_bounds__(256) k_glint_setup(DScene S, const sms_glint_query *__restrict__ Q, int64_t n,
                                                     uint64_t seed, float *__restrict__ out, EstimatorWork W) {
    int64_t q = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    bool valid = false;
    unsigned long long st_walks = 0;
    GlintRec rec;
    if (q < n) {
        const sms_glint_query g = Q[q];
        write_rgb(out, q, v3(0.f, 0.f, 0.f));
        if (g.shape >= 0) {
            SI si = glint_query_si(g);
            Pcg32 rq;
            stream_seed(rq, query_key(seed, (uint64_t) q), 0);
            EI ei = sample_emitter_interaction(S, si.p, S.emitters_all, S.n_emitters_all, rq);
            V3 edir = ei.p - si.p;
            float edist = norm(edir);
            edir = edir * (1.f / edist);
            bool occluded = ray_test(S, make_ray(si.p, edir, RAY_EPS_F * (1.f + hmax_abs(si.p)), edist * (1.f - RAY_EPS_F)));
            if (!occluded) {
                V2 target;
                if (glint_target(ld3(g.sensor_p), ei.p, si.p, si.sh_s, si.sh_t, si.sh_n, v3(0.f, 0.f, 1.f), target)) {
                    valid = true;
                    rec.q = (int32_t) q;
                    rec.target[0] = target.x; rec.target[1] = target.y;
                    rec.uv[0] = g.uv[0]; rec.uv[1] = g.uv[1];
                    rec.duv_dx[0] = g.duv_dx[0]; rec.duv_dx[1] = g.duv_dx[1];
                    rec.duv_dy[0] = g.duv_dy[0]; rec.duv_dy[1] = g.duv_dy[1];
                    rec.pad = g.shape;
                } else {
                    st_walks = 1;          // sample_glint fails its side test before the solve
                }
            }
        }
    }
    // wavefront-aggregated append
    unsigned long long ballot = __ballot(valid);
    unsigned long long base = 0;
    int lane = threadIdx.x & 63;
    if (lane == 0 && ballot) base = atomicAdd(W.n_rec, (unsigned long long) __popcll(ballot));
    base = __shfl(base, 0, 64);
    if (valid) {
        unsigned long long rank = __popcll(ballot & ((1ull << lane) - 1ull));
        W.rec[base + rank] = rec;
    }
    wave_add_u64(W.stats + 0, st_walks);
}

// ---------------------------------------------------------------- pass 2: persistent walks
struct WalkLane {
    int32_t r;                  // record index or -1
    int32_t k;                  // 0 = first walk, k >= 1 = Bernoulli trial k
    uint64_t qkey;
    GlintFootprint fp;
    GlintWalk w;
    V2 uv_sol;
};

SMS_DEV void start_walk(WalkLane &L) {
    Pcg32 rw;
    stream_seed(rw, L.qkey, (uint64_t) L.k + 1);
    V2 xi;
    xi.x = pcg32_float(rw);
    xi.y = pcg32_float(rw);
    glint_walk_start(L.fp, xi, L.w);
}

__global__ void __launch_bounds__(256) k_glint_walks(DScene S, uint64_t seed, EstimatorWork W) {
    const sms_config &cfg = S.cfg;
    const unsigned long long n_rec = *W.n_rec;
    unsigned long long st_walks = 0, st_conv = 0, st_iters = 0, st_bcalls = 0, st_btrials = 0, st_bmax = 0;
    WalkLane L;
    L.r = -1;
    while (true) {
        if (L.r < 0) {
            unsigned long long r = atomicAdd(W.cursor, 1ull);
            if (r >= n_rec) break;
            const GlintRec rec = W.rec[r];
            const DShape &gs = S.shapes[rec.pad];
            const DNormalmap nm = S.nmaps[gs.normalmap];
            L.fp.slope = nm.slope; L.fp.res = nm.res; L.fp.stride = nm.stride; L.fp.fres = (float) nm.res;
            L.fp.tiles = gs.tiles;
            footprint_from_rec(rec, L.fp);
            L.r = (int32_t) r;
            L.k = 0;
            L.qkey = query_key(seed, (uint64_t) rec.q);
            start_walk(L);
        }
        int st = glint_walk_iter(L.fp, cfg, L.w);
        if (st == WALK_RUNNING) continue;
        st_walks++;
        st_iters += (unsigned long long) L.w.it;
        bool ok = st == WALK_CONVERGED;
        if (ok) st_conv++;
        if (L.k == 0) {
            if (!ok) {
                GlintRes o; o.uv_sol[0] = o.uv_sol[1] = 0.f; o.inv_prob = 0.f; o.status = 0;
                W.res[L.r] = o;
                L.r = -1;
                continue;
            }
            L.uv_sol = L.w.uv;
            st_bcalls++;
            L.k = 1;
        } else {
            bool done = false;
            float inv_prob = 0.f;
            if (ok && norm2(L.w.uv - L.uv_sol) < cfg.uniqueness_threshold) {
                inv_prob = (float) L.k;          // match at trial k (manifold_glints.cpp:85-104)
                done = true;
            } else {
                L.k++;
                if (cfg.max_trials > 0 && L.k > cfg.max_trials) done = true;     // inv_prob = 0
            }
            if (done) {
                st_btrials += (unsigned long long) L.k;
                st_bmax = st_bmax > (unsigned long long) L.k ? st_bmax : (unsigned long long) L.k;
                GlintRes o; o.uv_sol[0] = L.uv_sol.x; o.uv_sol[1] = L.uv_sol.y; o.inv_prob = inv_prob; o.status = 1;
                W.res[L.r] = o;
                L.r = -1;
                continue;
            }
            if (L.k > W.lane_trials) {
                unsigned long long slot = atomicAdd(W.n_overflow, 1ull);
                OverflowRec ov;
                ov.q = L.r; ov.k = L.k; ov.sol[0] = L.uv_sol.x; ov.sol[1] = L.uv_sol.y; ov.sol[2] = 0.f;
                ov.shapes[0] = ov.shapes[1] = ov.shapes[2] = ov.shapes[3] = -1; ov.pad = 0;
                W.overflow[slot] = ov;
                GlintRes o; o.uv_sol[0] = L.uv_sol.x; o.uv_sol[1] = L.uv_sol.y; o.inv_prob = 0.f; o.status = 2;
                W.res[L.r] = o;
                L.r = -1;
                continue;
            }
        }
        start_walk(L);
    }
    wave_add_u64(W.stats + 0, st_walks); wave_add_u64(W.stats + 1, st_conv); wave_add_u64(W.stats + 2, st_iters);
    wave_add_u64(W.stats + 3, st_bcalls); wave_add_u64(W.stats + 4, st_btrials); wave_max_u64(W.stats + 5, st_bmax);
}

// ---------------------------------------------------------------- pass 3: tail (wave per query)
// The 64 lanes of a wavefront iterate inside one footprint, so its texels (plus the bilinear
// margin) are staged in LDS when they fit; otherwise the walk reads the global slope array.
constexpr int COOP_TILE = 24;                 // texels per side of the staged window
__global__ void __launch_bounds__(256) k_glint_coop(DScene S, uint64_t seed, EstimatorWork W) {
    __shared__ float2 tile[4][COOP_TILE + 1][COOP_TILE + 1];
    const sms_config &cfg = S.cfg;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    unsigned long long st_walks = 0, st_conv = 0, st_iters = 0, st_btrials = 0, st_bmax = 0, st_tail = 0;
    const unsigned long long n_over = *W.n_overflow;
    while (true) {
        unsigned long long item = 0;
        if (lane == 0) item = atomicAdd(W.coop_cursor, 1ull);
        item = __shfl(item, 0, 64);
        if (item >= n_over) break;
        const OverflowRec ov = W.overflow[item];
        const GlintRec rec = W.rec[ov.q];
        const DShape &gs = S.shapes[rec.pad];
        const DNormalmap nm = S.nmaps[gs.normalmap];
        GlintFootprint fp;
        fp.slope = nm.slope; fp.res = nm.res; fp.stride = nm.stride; fp.fres = (float) nm.res; fp.tiles = gs.tiles;
        footprint_from_rec(rec, fp);
        // Stage the footprint's texel window (bounding box of the parallelogram, one texel of
        // margin each side) when it fits; with tiling only footprints inside one tile qualify.
        LdsTexels<COOP_TILE> lds;
        lds.tile = tile[wv];
        lds.g.slope = fp.slope; lds.g.res = fp.res; lds.g.stride = fp.stride;
        bool staged = false;
        if (gs.tiles == 1.f) {
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
                        tile[wv][ty][tx] = fp.slope[(size_t) gy * fp.stride + gx];
                    }
                    lds.x0 = x0; lds.y0 = y0;
                    staged = true;
                }
            }
        }
        // The window belongs to this wavefront only (waves of a block leave the item loop at
        // different times, so no block barrier here): order its LDS writes before the reads.
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        uint64_t qkey = query_key(seed, (uint64_t) rec.q);
        V2 uv_sol = v2(ov.sol[0], ov.sol[1]);
        int64_t kcap = cfg.max_trials > 0 ? (int64_t) cfg.max_trials : (int64_t) 1 << 40;
        int64_t found = -1;
        for (int64_t base = ov.k; base <= kcap && found < 0; base += 64) {
            int64_t k = base + lane;
            bool match = false, ran = false, conv = false;
            int its = 0;
            if (k <= kcap) {
                Pcg32 rw;
                stream_seed(rw, qkey, (uint64_t) k + 1);
                V2 xi; xi.x = pcg32_float(rw); xi.y = pcg32_float(rw);
                GlintWalk w;
                glint_walk_start(fp, xi, w);
                int s;
                if (staged) { while ((s = glint_walk_iter_f(fp, cfg, w, lds)) == WALK_RUNNING) {} }
                else { while ((s = glint_walk_iter(fp, cfg, w)) == WALK_RUNNING) {} }
                ran = true; its = w.it; conv = s == WALK_CONVERGED;
                match = conv && norm2(w.uv - uv_sol) < cfg.uniqueness_threshold;
            }
            unsigned long long ballot = __ballot(match);
            if (ballot) found = base + (int64_t) __ffsll((long long) ballot) - 1;
            // count only the walks the sequential trial loop would have run (k <= first match)
            if (ran && (found < 0 || k <= found)) {
                st_walks++; st_iters += (unsigned long long) its; st_tail += (unsigned long long) its;
                if (conv) st_conv++;
            }
        }
        // every lane is done reading the window before the next item overwrites it
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (lane == 0) {
            W.res[ov.q].inv_prob = found >= 0 ? (float) found : 0.f;
            unsigned long long trials = found >= 0 ? (unsigned long long) found : (unsigned long long) kcap + 1ull;
            st_btrials += trials;
            st_bmax = st_bmax > trials ? st_bmax : trials;
        }
    }
    wave_add_u64(W.stats + 0, st_walks); wave_add_u64(W.stats + 1, st_conv); wave_add_u64(W.stats + 2, st_iters);
    wave_add_u64(W.stats + 4, st_btrials); wave_max_u64(W.stats + 5, st_bmax); wave_add_u64(W.stats + 6, st_tail);
}

// ---------------------------------------------------------------- pass 4: contributions
__global__ void __launch_bounds__(256) k_glint_finalize(DScene S, const sms_glint_query *__restrict__ Q, uint64_t seed,
                                                        float *__restrict__ out, EstimatorWork W) {
    const unsigned long long n_rec = *W.n_rec;
    const sms_config &cfg = S.cfg;
    for (unsigned long long r = (unsigned long long) blockIdx.x * blockDim.x + threadIdx.x; r < n_rec;
         r += (unsigned long long) gridDim.x * blockDim.x) {
        const GlintRes res = W.res[r];
        if (res.status == 0) continue;
        const int64_t q = W.rec[r].q;
        const sms_glint_query g = Q[q];
        SI si = glint_query_si(g);
        Pcg32 rq;
        stream_seed(rq, query_key(seed, (uint64_t) q), 0);
        EI ei = sample_emitter_interaction(S, si.p, S.emitters_all, S.n_emitters_all, rq);
        V3 sensor_p = ld3(g.sensor_p);
        float val = glint_contribution(S, sensor_p, ei, si, v2(res.uv_sol[0], res.uv_sol[1]));
        float We = sensor_we(S.cam, sensor_p, si.p);
        float mis = cfg.bsdf_strategy_only ? 0.f : 1.f;
        write_rgb(out, q, glint_result(val, res.inv_prob, mis * We, ei.weight));
    }
}

} // namespace smsdev
