// dev_glint.h -- glint manifold walk on the GPU (SpecularManifoldGlints,
// src/librender/manifold_glints.cpp): Newton iterations in UV space inside the pixel
// footprint against the bilinearly interpolated level-0 normal-map slope.
//
// The walk is written as a resumable state (GlintWalk) advanced one Newton iteration per
// call, so a persistent kernel can refill lanes whose walk ended while others keep
// iterating.  Per-footprint constants (target slope, parallelogram test terms) live in
// GlintFootprint and are computed once per query.
#pragma once
#include "dev_scene.h"

namespace smsdev {

struct TriTest { V2 a, v0, v1; float dot00, dot01, dot11, inv_denom; };

// the query-constant part of ManifoldGlints::inside_triangle (manifold_glints.h:163-184)
SMS_DEV TriTest tri_test_setup(V2 a, V2 b, V2 c) {
    TriTest t;
    t.a = a; t.v0 = c - a; t.v1 = b - a;
    t.dot00 = dot2(t.v0, t.v0); t.dot01 = dot2(t.v0, t.v1); t.dot11 = dot2(t.v1, t.v1);
    t.inv_denom = 1.f / (t.dot00 * t.dot11 - t.dot01 * t.dot01);
    return t;
}
SMS_DEV bool tri_test(const TriTest &t, V2 p) {
    V2 v2_ = p - t.a;
    float dot02 = dot2(t.v0, v2_), dot12 = dot2(t.v1, v2_);
    float u = (t.dot11 * dot02 - t.dot01 * dot12) * t.inv_denom, v = (t.dot00 * dot12 - t.dot01 * dot02) * t.inv_denom;
    return (u >= 0.f) && (v >= 0.f) && (u + v < 1.f);
}

struct GlintFootprint {
    V2 target, uv, duv_dx, duv_dy;
    TriTest t0, t1;              // (a,b,c) and (a,c,d) of the footprint parallelogram
    const float2 *slope;         // padded level-0 slopes
    int res, stride; float fres, tiles;
    const float *gmm;            // learned proposal: the query's seed mixture over the footprint (or null)
    float defensive;             // its weight of the reference's uniform footprint seed
};

// texel source of a walk: the padded slope array in global memory
struct GlobalTexels {
    const float2 *slope; int res, stride;
    SMS_DEV Texel4 operator()(V2 uv) const { return nm_fetch(slope, res, stride, uv); }
};

// texel source of a wavefront whose walks share one footprint: a (T+1)^2 window of the padded
// slope array staged in LDS at texel (x0, y0); lookups outside it read global memory
template <int T>
struct LdsTexels {
    const float2 (*tile)[T + 1];
    int x0, y0;
    GlobalTexels g;
    SMS_DEV Texel4 operator()(V2 uv) const {
        NmCoord c = nm_coord(g.res, uv);
        int lx = c.x - x0, ly = c.y - y0;
        if (lx < 0 || ly < 0 || lx + 1 > T || ly + 1 > T) return g(uv);
        float2 a = tile[ly][lx], b = tile[ly][lx + 1], d0 = tile[ly + 1][lx], d1 = tile[ly + 1][lx + 1];
        Texel4 t;
        t.v00 = v2(a.x, a.y); t.v10 = v2(b.x, b.y); t.v01 = v2(d0.x, d0.y); t.v11 = v2(d1.x, d1.y);
        t.w = c.w;
        return t;
    }
};

// footprint corners (manifold_glints.cpp:303-307)
SMS_DEV void footprint_setup(GlintFootprint &fp) {
    V2 a = (fp.uv + 0.5f * fp.duv_dx) + 0.5f * fp.duv_dy, b = (fp.uv + 0.5f * fp.duv_dx) - 0.5f * fp.duv_dy,
       c = (fp.uv - 0.5f * fp.duv_dx) - 0.5f * fp.duv_dy, d = (fp.uv - 0.5f * fp.duv_dx) + 0.5f * fp.duv_dy;
    fp.t0 = tri_test_setup(a, b, c);
    fp.t1 = tri_test_setup(a, c, d);
}

// target slope of sample_glint (manifold_glints.cpp:235-247); false if either side fails
SMS_DEV bool glint_target(V3 sensor_p, V3 ei_p, V3 p, V3 s, V3 t, V3 n, V3 n_offset, V2 &target) {
    V3 a = normalize(sensor_p - p), b = normalize(ei_p - p);
    V3 wi = v3(dot(a, s), dot(a, t), dot(a, n)), wo = v3(dot(b, s), dot(b, t), dot(b, n));
    if (wi.z <= 0.f || wo.z <= 0.f) return false;
    V3 h = normalize(wi + wo);
    V2 h_s = v2(-h.x / h.z, -h.y / h.z), o_s = v2(-n_offset.x / n_offset.z, -n_offset.y / n_offset.z);
    target = h_s + o_s;
    return true;
}

struct GlintWalk { V2 uv; float beta; int it; };
enum { WALK_RUNNING = 0, WALK_CONVERGED = 1, WALK_FAILED = 2 };

// seed inside the footprint (manifold_glints.cpp:252-259)
SMS_DEV void glint_walk_start(const GlintFootprint &fp, V2 xi, GlintWalk &w) {
    V2 duv = (xi.x - 0.5f) * fp.duv_dx + (xi.y - 0.5f) * fp.duv_dy;
    w.uv = fp.uv + duv; w.beta = 1.f; w.it = 0;
}

// One iteration of newton_solver + compute_step (manifold_glints.cpp:313-373).
template <class Fetch>
SMS_DEV int glint_walk_iter_f(const GlintFootprint &fp, const sms_config &cfg, GlintWalk &w, const Fetch &fetch) {
    if (w.it >= cfg.max_iterations) return WALK_FAILED;
    V2 tuv = tile_uv(w.uv, fp.tiles);
    Texel4 tx = fetch(tuv);
    V2 slope = nm_slope(tx);
    V2 tu, tv;
    nm_slope_deriv(tx, fp.fres, tu, tv);
    V2 du = tu * fp.tiles, dv = tv * fp.tiles;
    V2 dS = slope - fp.target;
    M2 J = m2(du.x, dv.x, du.y, dv.y);
    if (fabsf(det2(J)) < 1e-6f) return WALK_FAILED;
    V2 step = mulv2(inverse2(J), dS);
    if (norm2(dS) < cfg.solver_threshold) return WALK_CONVERGED;
    V2 uv_prop = w.uv - cfg.step_scale * w.beta * step;
    if (!(tri_test(fp.t0, uv_prop) || tri_test(fp.t1, uv_prop))) {
        w.beta = 0.5f * w.beta;
        w.it++;
        return WALK_RUNNING;
    }
    w.beta = fmin_(1.f, 2.f * w.beta);
    w.uv = uv_prop;
    w.it++;
    return WALK_RUNNING;
}
SMS_DEV int glint_walk_iter(const GlintFootprint &fp, const sms_config &cfg, GlintWalk &w) {
    GlobalTexels f;
    f.slope = fp.slope; f.res = fp.res; f.stride = fp.stride;
    return glint_walk_iter_f(fp, cfg, w, f);
}

// geometric_term glints (manifold_glints.cpp:411-554)
SMS_DEV float geometric_term_glint(const MV &v0, const MV &v1, const MV &v2v) {
    V3 wi = v0.p - v1.p;
    float ili = norm(wi);
    if (ili < 1e-3f) return 0.f;
    ili = 1.f / ili; wi = wi * ili;
    V3 wo = v2v.p - v1.p;
    float ilo = norm(wo);
    if (ilo < 1e-3f) return 0.f;
    ilo = 1.f / ilo; wo = wo * ilo;
    M2 dc1_dx0 = m2(0, 0, 0, 0), dc2_dx1 = m2(0, 0, 0, 0), dc2_dx2 = m2(0, 0, 0, 0);
    if (v2v.fixed_direction) {
        V3 dc2_du1 = ilo * (v1.dp_du - wo * dot(wo, v1.dp_du)), dc2_dv1 = ilo * (v1.dp_dv - wo * dot(wo, v1.dp_dv));
        dc2_dx1 = m2(dot(dc2_du1, v2v.dp_du), dot(dc2_dv1, v2v.dp_du), dot(dc2_du1, v2v.dp_dv), dot(dc2_dv1, v2v.dp_dv));
        V3 dc2_du2 = -ilo * (v2v.dp_du - wo * dot(wo, v2v.dp_du)), dc2_dv2 = -ilo * (v2v.dp_dv - wo * dot(wo, v2v.dp_dv));
        dc2_dx2 = m2(dot(dc2_du2, v2v.dp_du), dot(dc2_dv2, v2v.dp_du), dot(dc2_du2, v2v.dp_dv), dot(dc2_dv2, v2v.dp_dv));
    }
    V3 h = wi + wo;
    float ilh = 1.f / norm(h);
    h = h * ilh;
    ilo *= ilh; ili *= ilh;
    float dot_dpdu_n = dot(v1.dp_du, v1.n), dot_dpdv_n = dot(v1.dp_dv, v1.n);
    V3 s = v1.dp_du - dot_dpdu_n * v1.n, t = v1.dp_dv - dot_dpdv_n * v1.n;
    V3 dh_du, dh_dv;
    if (v2v.fixed_direction) {
        dh_du = ili * (v0.dp_du - wi * dot(wi, v0.dp_du));
        dh_dv = ili * (v0.dp_dv - wi * dot(wi, v0.dp_dv));
        dh_du = dh_du - h * dot(dh_du, h); dh_dv = dh_dv - h * dot(dh_dv, h);
        dc1_dx0 = m2(dot(dh_du, s), dot(dh_dv, s), dot(dh_du, t), dot(dh_dv, t));
    }
    dh_du = (-v1.dp_du * (ili + ilo) + wi * (dot(wi, v1.dp_du) * ili)) + wo * (dot(wo, v1.dp_du) * ilo);
    dh_dv = (-v1.dp_dv * (ili + ilo) + wi * (dot(wi, v1.dp_dv) * ili)) + wo * (dot(wo, v1.dp_dv) * ilo);
    dh_du = dh_du - h * dot(dh_du, h); dh_dv = dh_dv - h * dot(dh_dv, h);
    float dot_h_n = dot(h, v1.n), dot_h_dndu = dot(h, v1.dn_du), dot_h_dndv = dot(h, v1.dn_dv);
    M2 dc1_dx1 = m2((dot(dh_du, s) - dot(v1.dp_du, v1.dn_du) * dot_h_n) - dot_dpdu_n * dot_h_dndu,
                    (dot(dh_dv, s) - dot(v1.dp_du, v1.dn_dv) * dot_h_n) - dot_dpdu_n * dot_h_dndv,
                    (dot(dh_du, t) - dot(v1.dp_dv, v1.dn_du) * dot_h_n) - dot_dpdv_n * dot_h_dndu,
                    (dot(dh_dv, t) - dot(v1.dp_dv, v1.dn_dv) * dot_h_n) - dot_dpdv_n * dot_h_dndv);
    dh_du = ilo * (v2v.dp_du - wo * dot(wo, v2v.dp_du));
    dh_dv = ilo * (v2v.dp_dv - wo * dot(wo, v2v.dp_dv));
    dh_du = dh_du - h * dot(dh_du, h); dh_dv = dh_dv - h * dot(dh_dv, h);
    M2 dc1_dx2 = m2(dot(dh_du, s), dot(dh_dv, s), dot(dh_du, t), dot(dh_dv, t));
    float G;
    if (v2v.fixed_direction) {
        if (fabsf(det2(dc2_dx2)) < 1e-6f) return 0.f;
        M2 Li = inverse2(dc2_dx2);
        M2 tmp = mulm2(Li, dc2_dx1);
        M2 m = subm2(dc1_dx1, mulm2(dc1_dx2, tmp));
        if (fabsf(det2(m)) < 1e-6f) return 0.f;
        Li = inverse2(m);
        M2 sol1 = mulm2(negm2(Li), dc1_dx0);
        M2 sol0 = mulm2(negm2(tmp), sol1);
        G = fabsf(det2(negm2(sol0)));
    } else {
        if (fabsf(det2(dc1_dx1)) < 1e-6f) return 0.f;
        M2 inv = inverse2(dc1_dx1);
        float dx1_dx2 = fabsf(det2(mulm2(inv, dc1_dx2)));
        V3 d = v0.p - v1.p;
        float inv_r2 = 1.f / sqnorm(d);
        d = d * sqrtf(inv_r2);
        float dw0_dx1 = fabsf(dot(d, v1.gn)) * inv_r2;
        G = dw0_dx1 * dx1_dx2;
    }
    return fmin_(1e-5f, G);
}

// evaluate_glint_contribution (manifold_glints.cpp:267-294): Fresnel * G (same in RGB)
SMS_DEV float glint_contribution(const DScene &S, V3 sensor_p, const EI &ei, const SI &si_q, V2 uv) {
    MV vy;
    if (!emitter_interaction_to_vertex(S, ei, si_q.p, vy)) return 0.f;
    MV vx = mv_point(sensor_p);
    V3 dx = normalize(si_q.p - sensor_p);
    vx.n = dx; vx.gn = dx;
    coordinate_system(vx.n, vx.dp_du, vx.dp_dv);
    SI si = si_q;
    si.uv = uv;
    MV vtx = mv_from_si(S, si);
    make_orthonormal(vtx);
    const DShape &sh = S.shapes[si.shape];
    float cos_theta = dot(vtx.n, normalize(ei.p - si.p));
    float F = sh.k == 0.f ? fresnel(fabsf(cos_theta), sh.eta_r) : fresnel_conductor(fabsf(cos_theta), sh.eta_r, sh.k);
    float G = geometric_term_glint(vx, vtx, vy);
    return F * G;
}

// sensor importance x film area (manifold_glints.cpp:146-153; perspective.cpp:224-276)
SMS_DEV float sensor_we(const DCamera &c, V3 sensor_p, V3 p) {
    V3 d = normalize(p - sensor_p);
    V3 dl = v3(dot(d, ld3(c.left)), dot(d, ld3(c.up)), dot(d, ld3(c.dir)));
    float cos_theta = dl.z;
    if (!(cos_theta > 0.f)) return 0.f;
    float inv_cos = 1.f / cos_theta;
    float px = dl.x * inv_cos, py = dl.y * inv_cos;
    if (!(px >= -c.rect_x && px <= c.rect_x && py >= -c.rect_y && py <= c.rect_y)) return 0.f;
    float We = c.normalization * inv_cos * inv_cos * inv_cos;
    return We * (float) (c.w * c.h);
}

} // namespace smsdev
