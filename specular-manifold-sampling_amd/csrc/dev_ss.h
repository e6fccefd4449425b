// dev_ss.h -- single-scatter manifold walk on the GPU: angle-difference / half-vector
// constraints with their 2x2 Jacobians, the Newton solver with reprojection by whole-scene
// ray casts, the seed, and the path contribution (SpecularManifoldSingleScatter,
// src/librender/manifold_ss.cpp).
#pragma once
#include "dev_lean.h"
#include "dev_scene.h"

namespace smsdev {

// ---------------------------------------------------------------- specular transforms (manifold.h:347-444)
SMS_DEV V3 reflect(V3 w, V3 n) { return 2.f * dot(w, n) * n - w; }
SMS_DEV void d_reflect(V3 w, V3 dw_du, V3 dw_dv, V3 n, V3 dn_du, V3 dn_dv, V3 &du, V3 &dv) {
    float dot_w_n = dot(w, n), dot_dwdu_n = dot(dw_du, n), dot_dwdv_n = dot(dw_dv, n),
          dot_w_dndu = dot(w, dn_du), dot_w_dndv = dot(w, dn_dv);
    du = 2.f * ((dot_dwdu_n + dot_w_dndu) * n + dot_w_n * dn_du) - dw_du;
    dv = 2.f * ((dot_dwdv_n + dot_w_dndv) * n + dot_w_n * dn_dv) - dw_dv;
}
SMS_DEV bool refract(V3 w, V3 n_, float eta_, V3 &out) {
    V3 n = n_;
    float eta = 1.f / eta_;
    if (dot(w, n) < 0.f) { eta = 1.f / eta; n = n * -1.f; }
    float dot_w_n = dot(w, n);
    float root_term = 1.f - eta * eta * (1.f - dot_w_n * dot_w_n);
    if (root_term < 0.f) { out = v3(0.f, 0.f, 0.f); return false; }
    out = -eta * (w - dot_w_n * n) - n * sqrtf(root_term);
    return true;
}
SMS_DEV void d_refract(V3 w, V3 dw_du, V3 dw_dv, V3 n_, V3 dn_du_, V3 dn_dv_, float eta_, V3 &du, V3 &dv) {
    V3 n = n_, dn_du = dn_du_, dn_dv = dn_dv_;
    float eta = 1.f / eta_;
    if (dot(w, n) < 0.f) { eta = 1.f / eta; n = n * -1.f; dn_du = dn_du * -1.f; dn_dv = dn_dv * -1.f; }
    float dot_w_n = dot(w, n), dot_dwdu_n = dot(dw_du, n), dot_dwdv_n = dot(dw_dv, n),
          dot_w_dndu = dot(w, dn_du), dot_w_dndv = dot(w, dn_dv);
    float root = sqrtf(1.f - eta * eta * (1.f - dot_w_n * dot_w_n));
    V3 a_u = -eta * (dw_du - ((dot_dwdu_n + dot_w_dndu) * n + dot_w_n * dn_du));
    V3 b1_u = dn_du * root;
    V3 b2_u = n * (1.f / (2.f * root)) * (-eta * eta * (-2.f * dot_w_n * (dot_dwdu_n + dot_w_dndu)));
    V3 b_u = -(b1_u + b2_u);
    V3 a_v = -eta * (dw_dv - ((dot_dwdv_n + dot_w_dndv) * n + dot_w_n * dn_dv));
    V3 b1_v = dn_dv * root;
    V3 b2_v = n * (1.f / (2.f * root)) * (-eta * eta * (-2.f * dot_w_n * (dot_dwdv_n + dot_w_dndv)));
    V3 b_v = -(b1_v + b2_v);
    du = a_u + b_u; dv = a_v + b_v;
}
SMS_DEV void sphcoords(V3 w, float &theta, float &phi) {
    theta = safe_acos(w.z);
    phi = c_atan2f(w.y, w.x);
    if (phi < 0.f) phi += 2.f * PI_F;
}
SMS_DEV void d_sphcoords(V3 w, V3 dw_du, V3 dw_dv, float &dt_du, float &dp_du, float &dt_dv, float &dp_dv) {
    float d_acos = -(1.f / safe_sqrt(1.f - w.z * w.z));
    V2 d_theta = d_acos * v2(dw_du.z, dw_dv.z);
    float yx = w.y / w.x;
    float d_atan = 1.f / (1.f + yx * yx);
    V2 d_phi = (d_atan * v2(w.x * dw_du.y - w.y * dw_du.x, w.x * dw_dv.y - w.y * dw_dv.x)) * (1.f / (w.x * w.x));
    if (w.x == 0.f) d_phi = v2(0.f, 0.f);
    dt_du = d_theta.x; dp_du = d_phi.x; dt_dv = d_theta.y; dp_dv = d_phi.y;
}

struct Step { bool ok; V2 C, dX; };

// compute_step_anglediff (manifold_ss.cpp:494-636)
SMS_DEV Step step_anglediff(V3 v0p, const MV &v1, const EI &v2e, V3 n_offset) {
    Step r; r.ok = false; r.C = v2(__builtin_inff(), __builtin_inff()); r.dX = v2(0.f, 0.f);
    bool directional = v2e.kind == SMS_EI_DIRECTIONAL;
    V3 wi = v0p - v1.p;
    float ili = norm(wi);
    if (ili < 1e-3f) return r;
    ili = 1.f / ili;
    wi = wi * ili;
    V3 dwi_du = -ili * (v1.dp_du - wi * dot(wi, v1.dp_du));
    V3 dwi_dv = -ili * (v1.dp_dv - wi * dot(wi, v1.dp_dv));
    V3 wo = directional ? v2e.d : v2e.p - v1.p;
    float ilo = norm(wo);
    if (ilo < 1e-3f) return r;
    ilo = 1.f / ilo;
    wo = wo * ilo;
    V3 dwo_du, dwo_dv;
    if (directional) { dwo_du = v3(0.f, 0.f, 0.f); dwo_dv = v3(0.f, 0.f, 0.f); }
    else {
        dwo_du = -ilo * (v1.dp_du - wo * dot(wo, v1.dp_du));
        dwo_dv = -ilo * (v1.dp_dv - wo * dot(wo, v1.dp_dv));
    }
    V2 C = v2(0.f, 0.f);
    M2 dC = m2(0.f, 0.f, 0.f, 0.f);
    V3 n = (v1.s * n_offset.x + v1.t * n_offset.y) + v1.n * n_offset.z;
    V3 dn_du = (v1.ds_du * n_offset.x + v1.dt_du * n_offset.y) + v1.dn_du * n_offset.z;
    V3 dn_dv = (v1.ds_dv * n_offset.x + v1.dt_dv * n_offset.y) + v1.dn_dv * n_offset.z;
    bool success = false;
    bool refl = v1.eta == 1.f;
    V3 wio; bool valid_i;
    if (refl) { wio = reflect(wi, n); valid_i = true; } else valid_i = refract(wi, n, v1.eta, wio);
    if (valid_i) {
        V3 dwio_du, dwio_dv;
        if (refl) d_reflect(wi, dwi_du, dwi_dv, n, dn_du, dn_dv, dwio_du, dwio_dv);
        else d_refract(wi, dwi_du, dwi_dv, n, dn_du, dn_dv, v1.eta, dwio_du, dwio_dv);
        float to, po, tio, pio;
        sphcoords(wo, to, po); sphcoords(wio, tio, pio);
        float dt = to - tio, dp = po - pio;
        if (dp < -PI_F) dp += 2.f * PI_F; else if (dp > PI_F) dp -= 2.f * PI_F;
        C = v2(dt, dp);
        float dto_du, dpo_du, dto_dv, dpo_dv, dtio_du, dpio_du, dtio_dv, dpio_dv;
        d_sphcoords(wo, dwo_du, dwo_dv, dto_du, dpo_du, dto_dv, dpo_dv);
        d_sphcoords(wio, dwio_du, dwio_dv, dtio_du, dpio_du, dtio_dv, dpio_dv);
        dC = m2(dto_du - dtio_du, dto_dv - dtio_dv, dpo_du - dpio_du, dpo_dv - dpio_dv);
        success = true;
    }
    V3 woi; bool valid_o;
    if (refl) { woi = reflect(wo, n); valid_o = true; } else valid_o = refract(wo, n, v1.eta, woi);
    if (valid_o && !success) {
        V3 dwoi_du, dwoi_dv;
        if (refl) d_reflect(wo, dwo_du, dwo_dv, n, dn_du, dn_dv, dwoi_du, dwoi_dv);
        else d_refract(wo, dwo_du, dwo_dv, n, dn_du, dn_dv, v1.eta, dwoi_du, dwoi_dv);
        float ti, pi, toi, poi;
        sphcoords(wi, ti, pi); sphcoords(woi, toi, poi);
        float dt = ti - toi, dp = pi - poi;
        if (dp < -PI_F) dp += 2.f * PI_F; else if (dp > PI_F) dp -= 2.f * PI_F;
        C = v2(dt, dp);
        float dti_du, dpi_du, dti_dv, dpi_dv, dtoi_du, dpoi_du, dtoi_dv, dpoi_dv;
        d_sphcoords(wi, dwi_du, dwi_dv, dti_du, dpi_du, dti_dv, dpi_dv);
        d_sphcoords(woi, dwoi_du, dwoi_dv, dtoi_du, dpoi_du, dtoi_dv, dpoi_dv);
        dC = m2(dti_du - dtoi_du, dti_dv - dtoi_dv, dpi_du - dpoi_du, dpi_dv - dpoi_dv);
        success = true;
    }
    float determinant = det2(dC);
    if (fabsf(determinant) < 1e-6f) return r;
    M2 inv = inverse2(dC);
    r.ok = success; r.C = C; r.dX = mulv2(inv, C);
    return r;
}

// compute_step_halfvector (manifold_ss.cpp:407-492)
SMS_DEV Step step_halfvector(V3 v0p, const MV &v1, const EI &v2e, V3 n_offset) {
    Step r; r.ok = false; r.C = v2(__builtin_inff(), __builtin_inff()); r.dX = v2(0.f, 0.f);
    bool directional = v2e.kind == SMS_EI_DIRECTIONAL;
    V3 wo = directional ? v2e.d : v2e.p - v1.p;
    float ilo = norm(wo);
    if (ilo < 1e-3f) return r;
    ilo = 1.f / ilo; wo = wo * ilo;
    V3 wi = v0p - v1.p;
    float ili = norm(wi);
    if (ili < 1e-3f) return r;
    ili = 1.f / ili; wi = wi * ili;
    float eta = v1.eta;
    if (dot(wi, v1.gn) < 0.f) eta = 1.f / eta;
    V3 h = wi + eta * wo;
    if (eta != 1.f) h = h * -1.f;
    float ilh = 1.f / norm(h);
    h = h * ilh;
    ilo *= eta * ilh; ili *= ilh;
    V3 dh_du, dh_dv;
    if (directional) {
        dh_du = ili * (-v1.dp_du + wi * dot(wi, v1.dp_du));
        dh_dv = ili * (-v1.dp_dv + wi * dot(wi, v1.dp_dv));
    } else {
        dh_du = (-v1.dp_du * (ili + ilo) + wi * (dot(wi, v1.dp_du) * ili)) + wo * (dot(wo, v1.dp_du) * ilo);
        dh_dv = (-v1.dp_dv * (ili + ilo) + wi * (dot(wi, v1.dp_dv) * ili)) + wo * (dot(wo, v1.dp_dv) * ilo);
    }
    dh_du = dh_du - h * dot(dh_du, h);
    dh_dv = dh_dv - h * dot(dh_dv, h);
    if (eta != 1.f) { dh_du = dh_du * -1.f; dh_dv = dh_dv * -1.f; }
    M2 dH = m2(dot(v1.ds_du, h) + dot(v1.s, dh_du), dot(v1.ds_dv, h) + dot(v1.s, dh_dv),
               dot(v1.dt_du, h) + dot(v1.t, dh_du), dot(v1.dt_dv, h) + dot(v1.t, dh_dv));
    if (fabsf(det2(dH)) < 1e-6f) return r;
    M2 inv = inverse2(dH);
    V2 H = v2(dot(v1.s, h), dot(v1.t, h));
    V2 dHv = H - v2(n_offset.x, n_offset.y);
    r.ok = true; r.C = dHv; r.dX = mulv2(inv, dHv);
    return r;
}

struct PathResult { bool ok; bool init_valid; SI si_final; V3 p_init; int iterations; };

// SpecularManifold::sample_gaussian (manifold.h:328-345) with warp::square_to_std_normal
// (warp.h:184-190); sigma = (s00, s11, s01)
SMS_DEV V2 sample_gaussian(V2 mu, const float sig[3], V2 s) {
    const float r = sqrtf(-2.f * logf_l(1.f - s.x));
    float sn, cs;
    c_sincosf((2.f * PI_F) * s.y, &sn, &cs);
    const V2 p = v2(cs * r, sn * r);
    const float sx = sqrtf(sig[0]), sy = sqrtf(sig[1]);
    const float rho = sig[2] / (sx * sy);
    const float a = sqrtf(1.f + rho), b = sqrtf(1.f - rho);
    const float d = a * 0.5f + b * 0.5f, o = a * 0.5f + b * -0.5f;     // A = a P + b Q
    V2 q = v2(d * p.x + o * p.y, o * p.x + d * p.y);
    q.x = q.x * sx; q.y = q.y * sy;
    return q + mu;
}

// BSDF::lean at an interaction without uv partials: NormalmapBSDF::lean (bsdfs/normalmap.cpp:
// 292-303) at the coarsest mip level (mipmap_level, :361-364); (0, 0) for other BSDFs
SMS_DEV void lean_coarsest(const DScene &S, const SI &si, V2 &mu, float sig[3]) {
    const DShape &sh = S.shapes[si.shape];
    if (sh.bsdf != SMS_BSDF_NORMALMAP) { mu = v2(0.f, 0.f); sig[0] = sig[1] = sig[2] = 0.f; return; }
    const DNormalmap nm = S.nmaps[sh.normalmap];
    V2 uv = si.uv * sh.tiles;
    uv = v2(uv.x - floorf(uv.x), uv.y - floorf(uv.y));
    lean_moments(nm, uv, nm.levels - 1, mu, sig);
    // without use_slopes, NormalmapBSDF::lean takes the mean from the interpolated RGB as
    // stored, n = rgb (not decoded): mu = (-r / b, -g / b)
    if (nm.encoded) mu = v2(-nm.top_rgb[0] / nm.top_rgb[2], -nm.top_rgb[1] / nm.top_rgb[2]);
}

// newton_solver SS (manifold_ss.cpp:329-405) from vtx; smooth: vertices at smoothing 1
SMS_DEV bool newton_ss(const DScene &S, V3 si_p, MV &vtx, const EI &ei, V3 n_offset, bool smooth, SI &si_current,
                       int &iterations) {
    const sms_config &cfg = S.cfg;
    bool success = false;
    iterations = 0;
    float beta = 1.f;
    si_current = si_invalid();
    while (iterations < cfg.max_iterations) {
        Step st = cfg.halfvector_constraints ? step_halfvector(si_p, vtx, ei, n_offset)
                                             : step_anglediff(si_p, vtx, ei, n_offset);
        if (!st.ok) break;
        if (norm2(st.C) < cfg.solver_threshold) { success = true; break; }
        V3 p_prop = vtx.p - cfg.step_scale * beta * (vtx.dp_du * st.dX.x + vtx.dp_dv * st.dX.y);
        V3 d_prop = normalize(p_prop - si_p);
        si_current = ray_intersect(S, make_ray(si_p, d_prop));
        if (!si_current.valid || vtx.shape != si_current.shape) {
            beta = 0.5f * beta;
            iterations += 1;
            continue;
        }
        beta = fmin_(1.f, 2.f * beta);
        vtx = mv_from_si(S, si_current, smooth);
        iterations++;
    }
    if (!success) return false;
    V3 wx = normalize(si_p - vtx.p);
    V3 wy = ei.kind == SMS_EI_DIRECTIONAL ? ei.d : normalize(ei.p - vtx.p);
    float cos_theta_x = dot(vtx.gn, wx), cos_theta_y = dot(vtx.gn, wy);
    bool refraction = cos_theta_x * cos_theta_y < 0.f, reflection = !refraction;
    return !((vtx.eta == 1.f && !reflection) || (vtx.eta != 1.f && !refraction));
}

// sample_path SS (manifold_ss.cpp:203-279): uniform seed, optional MNEE direction, and with
// config.twostage the LEAN-smoothed pre-solve (:251-270) from an offset slope drawn with
// lean_sample; iterations count both solves
SMS_DEV PathResult sample_path_ss(const DScene &S, int shape, V3 si_p, const EI &ei, V2 sample, V3 n_offset, V3 p_start,
                                  V2 lean_sample) {
    const sms_config &cfg = S.cfg;
    PathResult out;
    out.ok = false; out.init_valid = false; out.si_final = si_invalid(); out.p_init = v3(0.f, 0.f, 0.f);
    out.iterations = 0;
    PS ps = sample_position(S, shape, sample);
    if (p_start.x != 0.f || p_start.y != 0.f || p_start.z != 0.f) ps.p = p_start;
    V3 d_tmp = normalize(ps.p - si_p);
    if (cfg.mnee_init) {
        const DShape &sh = S.shapes[shape];
        if (sh.k != 0.f) d_tmp = normalize(ld3(sh.bbox_center) - si_p);
        else d_tmp = normalize(ei.p - si_p);
    }
    SI si_init = ray_intersect(S, make_ray(si_p, d_tmp));
    if (!si_init.valid || shape != si_init.shape) return out;
    out.init_valid = true; out.p_init = si_init.p;
    MV vtx = mv_from_si(S, si_init, cfg.twostage != 0);
    int smooth_iterations = 0;
    SI si_current;
    int iterations;
    if (cfg.twostage) {
        const V2 mu_offset = v2(-n_offset.x / n_offset.z, -n_offset.y / n_offset.z);
        V2 mu;
        float sig[3];
        lean_coarsest(S, si_init, mu, sig);
        const V2 slope = sample_gaussian(0.5f * (mu + mu_offset), sig, lean_sample);
        const V3 lean_normal = normalize(v3(-slope.x, -slope.y, 1.f));
        const bool ok = newton_ss(S, si_p, vtx, ei, lean_normal, true, si_current, smooth_iterations);
        if (!(ok && si_current.valid)) { out.iterations = smooth_iterations; return out; }
        vtx = mv_from_si(S, si_current);
    }
    const bool ok = newton_ss(S, si_p, vtx, ei, n_offset, false, si_current, iterations);
    out.iterations = smooth_iterations + iterations;
    out.si_final = si_current;
    out.ok = ok && si_current.valid;
    return out;
}

// specular_reflectance SS, smooth branch (manifold_ss.cpp:638-693)
// ---------------------------------------------------------------- rough conductors
// MicrofacetDistribution(Beckmann, alpha, alpha, sample_visible = false) of
// include/mitsuba/render/microfacet.h: sample (:229-290, isotropic), eval (:184-201), smith_g1
// (:331-355, the rational approximation); RoughConductor::eval (src/bsdfs/roughconductor.cpp:
// 283-356, unpolarized).
SMS_DEV V3 beckmann_sample(float alpha, V2 s, float &pdf) {
    float sin_phi, cos_phi;
    c_sincosf((2.f * PI_F) * s.y, &sin_phi, &cos_phi);
    const float alpha_2 = alpha * alpha;
    const float cos_theta = 1.f / sqrtf(fmaf(-alpha_2, logf_l(1.f - s.x), 1.f));
    const float cos_theta_2 = cos_theta * cos_theta;
    const float cos_theta_3 = fmax_(cos_theta_2 * cos_theta, 1e-20f);
    pdf = (1.f - s.x) / (PI_F * alpha * alpha * cos_theta_3);
    const float sin_theta = sqrtf(1.f - cos_theta_2);
    return v3(cos_phi * sin_theta, sin_phi * sin_theta, cos_theta);
}
SMS_DEV float beckmann_eval(float alpha, V3 m) {
    const float alpha_uv = alpha * alpha, cos_theta = m.z, cos_theta_2 = cos_theta * cos_theta;
    const float a = m.x / alpha, b = m.y / alpha;
    const float result = expf_l(-(a * a + b * b) / cos_theta_2) / (PI_F * alpha_uv * (cos_theta_2 * cos_theta_2));
    return result * cos_theta > 1e-20f ? result : 0.f;
}
SMS_DEV float beckmann_g1(float alpha, V3 v, V3 m) {
    const float ax = alpha * v.x, ay = alpha * v.y;
    const float xy_alpha_2 = ax * ax + ay * ay, tan_theta_alpha_2 = xy_alpha_2 / (v.z * v.z);
    const float a = 1.f / sqrtf(tan_theta_alpha_2), a_sqr = a * a;
    float result = a >= 1.6f ? 1.f : (3.535f * a + 2.181f * a_sqr) / ((1.f + 2.276f * a) + 2.577f * a_sqr);
    if (xy_alpha_2 == 0.f) result = 1.f;
    if (dot(v, m) * v.z <= 0.f) result = 0.f;
    return result;
}
SMS_DEV float rough_conductor_eval(const DShape &sh, V3 wi, V3 wo) {
    if (!(wi.z > 0.f && wo.z > 0.f)) return 0.f;
    const V3 H = normalize(wo + wi);
    const float D = beckmann_eval(sh.alpha, H);
    if (D == 0.f) return 0.f;
    const float G = beckmann_g1(sh.alpha, wi, H) * beckmann_g1(sh.alpha, wo, H);
    const float result = D * G / (4.f * wi.z);
    const float F = fresnel_conductor(dot(wi, H), sh.eta_r, sh.k);
    return F * result;
}

// offset normal of caster c of a query (manifold_ss.cpp:61-70): for a rough caster a Beckmann
// sample from the query's stream after the emitter sample, one 2D sample per rough caster in
// caster order; returns its pdf (1 and (0, 0, 1) for smooth casters)
SMS_DEV float caster_offset(const DScene &S, uint64_t qkey, int c, V3 &n_offset) {
    n_offset = v3(0.f, 0.f, 1.f);
    const float alpha = S.shapes[S.casters_ss[c]].alpha;
    if (!(alpha > 0.f)) return 1.f;
    Pcg32 rq;
    stream_seed(rq, qkey, 0);
    if (S.n_emitters_ss > 0) { (void) pcg32_float(rq); (void) pcg32_float(rq); (void) pcg32_float(rq); }
    for (int j = 0; j < c; ++j)
        if (S.shapes[S.casters_ss[j]].alpha > 0.f) { (void) pcg32_float(rq); (void) pcg32_float(rq); }
    V2 s;
    s.x = pcg32_float(rq); s.y = pcg32_float(rq);
    float pdf;
    n_offset = beckmann_sample(alpha, s, pdf);
    return pdf;
}

// specular_reflectance SS (manifold_ss.cpp:638-693): the glossy branch for rough conductors (the
// BSDF in the half-vector domain; frame (s, t, n) and wi the solution interaction's), else the
// smooth one
SMS_DEV float specular_reflectance_ss_n(const DScene &S, int shape, V3 sh_n, V3 wo, V3 sh_s, V3 sh_t, V3 wi_local) {
    const DShape &sh = S.shapes[shape];
    if (sh.alpha > 0.f) {
        const V3 wo_l = v3(dot(wo, sh_s), dot(wo, sh_t), dot(wo, sh_n));
        float val = rough_conductor_eval(sh, wi_local, wo_l);
        const V3 h_l = normalize(wi_local + wo_l);
        return val * (4.f * fabsf(dot(wo_l, h_l)));
    }
    float cos_theta = dot(sh_n, wo);
    if (sh.k != 0.f) return fresnel_conductor(fabsf(cos_theta), sh.eta_r, sh.k);
    float eta = sh.eta_vtx;
    if (cos_theta < 0.f) eta = 1.f / eta;
    float F = fresnel(cos_theta, eta);
    float val = 1.f - F;
    return val * (eta * eta);
}
SMS_DEV float specular_reflectance_ss(const DScene &S, const SI &si, V3 wo) {
    if (!si.valid) return 0.f;
    return specular_reflectance_ss_n(S, si.shape, si.sh_n, wo, si.sh_s, si.sh_t, si.wi);
}

// geometric_term SS (manifold_ss.cpp:695-857)
SMS_DEV float geometric_term_ss(const MV &v0, const MV &v1, const MV &v2v) {
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
    float eta = v1.eta;
    if (dot(wi, v1.gn) < 0.f) eta = 1.f / eta;
    V3 h = wi + eta * wo;
    if (eta != 1.f) h = h * -1.f;
    float ilh = 1.f / norm(h);
    h = h * ilh;
    ilo *= eta * ilh; ili *= ilh;
    float dot_dpdu_n = dot(v1.dp_du, v1.n), dot_dpdv_n = dot(v1.dp_dv, v1.n);
    V3 s = v1.dp_du - dot_dpdu_n * v1.n, t = v1.dp_dv - dot_dpdv_n * v1.n;
    V3 dh_du, dh_dv;
    if (v2v.fixed_direction) {
        dh_du = ili * (v0.dp_du - wi * dot(wi, v0.dp_du));
        dh_dv = ili * (v0.dp_dv - wi * dot(wi, v0.dp_dv));
        dh_du = dh_du - h * dot(dh_du, h); dh_dv = dh_dv - h * dot(dh_dv, h);
        if (eta != 1.f) { dh_du = dh_du * -1.f; dh_dv = dh_dv * -1.f; }
        dc1_dx0 = m2(dot(dh_du, s), dot(dh_dv, s), dot(dh_du, t), dot(dh_dv, t));
    }
    dh_du = (-v1.dp_du * (ili + ilo) + wi * (dot(wi, v1.dp_du) * ili)) + wo * (dot(wo, v1.dp_du) * ilo);
    dh_dv = (-v1.dp_dv * (ili + ilo) + wi * (dot(wi, v1.dp_dv) * ili)) + wo * (dot(wo, v1.dp_dv) * ilo);
    dh_du = dh_du - h * dot(dh_du, h); dh_dv = dh_dv - h * dot(dh_dv, h);
    if (eta != 1.f) { dh_du = dh_du * -1.f; dh_dv = dh_dv * -1.f; }
    float dot_h_n = dot(h, v1.n), dot_h_dndu = dot(h, v1.dn_du), dot_h_dndv = dot(h, v1.dn_dv);
    M2 dc1_dx1 = m2((dot(dh_du, s) - dot(v1.dp_du, v1.dn_du) * dot_h_n) - dot_dpdu_n * dot_h_dndu,
                    (dot(dh_dv, s) - dot(v1.dp_du, v1.dn_dv) * dot_h_n) - dot_dpdu_n * dot_h_dndv,
                    (dot(dh_du, t) - dot(v1.dp_dv, v1.dn_du) * dot_h_n) - dot_dpdv_n * dot_h_dndu,
                    (dot(dh_dv, t) - dot(v1.dp_dv, v1.dn_dv) * dot_h_n) - dot_dpdv_n * dot_h_dndv);
    dh_du = ilo * (v2v.dp_du - wo * dot(wo, v2v.dp_du));
    dh_dv = ilo * (v2v.dp_dv - wo * dot(wo, v2v.dp_dv));
    dh_du = dh_du - h * dot(dh_du, h); dh_dv = dh_dv - h * dot(dh_dv, h);
    if (eta != 1.f) { dh_du = dh_du * -1.f; dh_dv = dh_dv * -1.f; }
    M2 dc1_dx2 = m2(dot(dh_du, s), dot(dh_dv, s), dot(dh_du, t), dot(dh_dv, t));
    if (v2v.fixed_direction) {
        if (fabsf(det2(dc2_dx2)) < 1e-6f) return 0.f;
        M2 Li = inverse2(dc2_dx2);
        M2 tmp = mulm2(Li, dc2_dx1);
        M2 m = subm2(dc1_dx1, mulm2(dc1_dx2, tmp));
        if (fabsf(det2(m)) < 1e-6f) return 0.f;
        Li = inverse2(m);
        M2 sol1 = mulm2(negm2(Li), dc1_dx0);
        M2 sol0 = mulm2(negm2(tmp), sol1);
        float G = fabsf(det2(negm2(sol0)));
        G = fmin_(G, 10.f);
        G /= fabsf(dot(wi, v0.n));
        G *= eta * eta;
        return G;
    }
    if (fabsf(det2(dc1_dx1)) < 1e-6f) return 0.f;
    M2 inv = inverse2(dc1_dx1);
    float dx1_dx2 = fabsf(det2(mulm2(inv, dc1_dx2)));
    dx1_dx2 = fmin_(dx1_dx2, 1.f);
    V3 d = v0.p - v1.p;
    float inv_r2 = 1.f / sqnorm(d);
    d = d * sqrtf(inv_r2);
    float dw0_dx1 = fabsf(dot(d, v1.gn)) * inv_r2;
    return dw0_dx1 * dx1_dx2;
}

// evaluate_path_contribution SS (manifold_ss.cpp:281-327); si = receiver (query) interaction
// vfinal = mv_from_si of the (valid) solution surface interaction, sh_n_final its shading normal
template <bool LDS_STACK = false, int LDS_DEPTH = SMS_BVH_STACK>
SMS_DEV V3 path_contribution_ss_mv(const DScene &S, const SI &si, const EI &ei_, const MV &vfinal, V3 sh_n_final,
                                   V3 wi_final) {
    MV vtx = vfinal;
    make_orthonormal(vtx);
    EI ei = ei_;
    if (ei.kind == SMS_EI_POINT) {
        const DEmitter &em = S.emitters[ei.emitter];
        V3 pos = ld3(em.position);
        V3 d = pos - vfinal.p;
        float dist = norm(d), inv_dist = 1.f / dist;
        d = d * inv_dist;
        float emitter_pdf = ei.pdf;
        ei.p = pos; ei.d = d; ei.n = d; ei.pdf = 1.f;
        ei.weight = (ld3(em.intensity) * (inv_dist * inv_dist)) * dist * dist;
        ei.pdf *= emitter_pdf;
        ei.weight = ei.weight * (1.f / emitter_pdf);
    }
    MV vy;
    if (!emitter_interaction_to_vertex<LDS_STACK, LDS_DEPTH>(S, ei, vtx.p, vy)) return v3(0.f, 0.f, 0.f);
    MV vx = mv_from_si(S, si);
    make_orthonormal(vx);
    float F = specular_reflectance_ss_n(S, vfinal.shape, sh_n_final, normalize(ei.p - vfinal.p), vfinal.s, vfinal.t,
                                        wi_final);
    float G = geometric_term_ss(vx, vtx, vy);
    float thr = F * G;
    return v3(thr * ei.weight.x, thr * ei.weight.y, thr * ei.weight.z);
}

// contribution of a walk's solution (manifold_ss.cpp:281-327)
SMS_DEV V3 path_contribution_ss(const DScene &S, const SI &si, const EI &ei, const SI &si_final) {
    if (!si_final.valid) return v3(0.f, 0.f, 0.f);
    return path_contribution_ss_mv(S, si, ei, mv_from_si(S, si_final), si_final.sh_n, si_final.wi);
}

} // namespace smsdev
