// dev_ss_est.h -- pieces of the single-scatter estimator shared by its kernels: the query's
// surface interaction, the receiver BSDF, the shadow test, one seeded walk, and the training
// capture of the learned proposal.
#pragma once
#include "kernels.h"

namespace smsdev {

SMS_DEV SI ss_query_si(const sms_ss_query &g) {
    SI si = si_invalid();
    si.valid = true; si.t = 1.f;
    si.p = ld3(g.p); si.n = ld3(g.n); si.sh_n = ld3(g.sh_n); si.sh_s = ld3(g.sh_s); si.sh_t = ld3(g.sh_t);
    si.dp_du = si.sh_s; si.dp_dv = si.sh_t; si.shape = g.shape;
    return si;
}

// Diffuse::eval at the receiver (src/bsdfs/diffuse.cpp:104-120)
SMS_DEV V3 receiver_bsdf_eval(const DScene &S, const SI &si, V3 wi_local, V3 wo_local) {
    const DShape &sh = S.shapes[si.shape];
    if (sh.bsdf != SMS_BSDF_DIFFUSE) return v3(0.f, 0.f, 0.f);
    if (!(wi_local.z > 0.f && wo_local.z > 0.f)) return v3(0.f, 0.f, 0.f);
    return (ld3(sh.refl) * INV_PI_F) * wo_local.z;
}

SMS_DEV bool ss_visible(const DScene &S, const EI &ei, V3 p) {
    // manifold_ss.cpp:88-102: toward a directional emitter the ray leaves the solution
    if (ei.kind == SMS_EI_DIRECTIONAL) return !ray_test(S, make_ray(p, ei.d));
    V3 d = p - ei.p;
    float dist = norm(d);
    d = d * (1.f / dist);
    return !ray_test(S, make_ray(ei.p, d, RAY_EPS_F * (1.f + hmax_abs(ei.p)), dist * (1.f - RAY_EPS_F)));
}

SMS_DEV PathResult ss_walk(const DScene &S, const Proposal &P, int64_t q, uint64_t qkey, int c, int shape, uint64_t k,
                           V3 si_p, const EI &ei, V3 n_offset) {
    Pcg32 rw;
    stream_seed(rw, qkey, ((uint64_t) c << 32) + k + 1);
    V2 s = v2(0.f, 0.f);
    V3 p_start = v3(0.f, 0.f, 0.f);
    if (P.gmm && shape == P.shape) {
        if (!proposal_seed(S, P, P.gmm + q * PROP_RAW, proposal_anchor(P, q), rw, s, p_start)) {
            PathResult res;
            res.ok = false; res.init_valid = false; res.iterations = 0;
            res.si_final = si_invalid(); res.p_init = v3(0.f, 0.f, 0.f);
            return res;
        }
    } else {
        s.x = pcg32_float(rw); s.y = pcg32_float(rw);
    }
    V2 ls = v2(0.f, 0.f);
    if (S.cfg.twostage) { ls.x = pcg32_float(rw); ls.y = pcg32_float(rw); }
    return sample_path_ss(S, shape, si_p, ei, s, n_offset, p_start, ls);
}

// record (network inputs, solution uv, weight) for the proposal's training
SMS_DEV void capture_sample(const DScene &S, const Proposal &P, const SI &si, const EI &ei, V2 uv, float w,
                            const float *anc = nullptr) {
    if (!proposal_target(anc, P.radius, uv, uv)) return;
    const int sub = (int) (blockIdx.x % SPREAD);
    unsigned long long pos = atomicAdd(P.cap_counts + sub * SPREAD_STRIDE, 1ull);
    if (pos >= P.cap_sub) return;
    float *r = P.cap + ((size_t) sub * P.cap_sub + pos) * 12;
    proposal_inputs(S, si.p, si.sh_n, ei.p, r);
    r[9] = uv.x; r[10] = uv.y; r[11] = w;
}

SMS_DEV bool capturing(const Proposal &P, int shape, int mode) { return P.cap && shape == P.shape && P.cap_mode == mode; }
SMS_DEV float mean3(const float *v) { return ((v[0] + v[1]) + v[2]) * (1.f / 3.f); }

} // namespace smsdev
