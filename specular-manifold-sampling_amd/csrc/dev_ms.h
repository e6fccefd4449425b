// dev_ms.h -- multi-scatter manifold walk on the GPU (SpecularManifoldMultiScatter,
// src/librender/manifold_ms.cpp): a k-vertex specular chain traced through a seed point,
// per-vertex angle-difference constraints with 2x2 blocks w.r.t. the previous / current /
// next vertex, block-Thomas solve of the block-tridiagonal Newton system, chain reprojection,
// and the generalized geometric term of the whole chain.  MAXB (maximum chain length) is a
// template parameter so the per-lane chain stays in registers for the common k = 2.
#pragma once
#include "dev_ss.h"

namespace smsdev {

struct MSVert { MV v; V2 C; M2 dprev, dcur, dnext, tmp, inv_lambda; V2 dx; };

template <int MAXB>
struct MSPath { int n; MSVert v[MAXB]; int shapes[MAXB]; };


// offset normal (0,0,1) in world space (manifold_ms.cpp:386-388)
SMS_DEV V3 offset_normal_world(const MV &v) { return (v.s * 0.f + v.t * 0.f) + v.n * 1.f; }

SMS_DEV bool invert_exact(M2 a, M2 &inv) {
    if (fabsf(det2(a)) == 0.f) return false;
    inv = inverse2(a);
    return true;
}
SMS_DEV float wrap_pi(float dp) {
    if (dp < -PI_F) dp += 2.f * PI_F; else if (dp > PI_F) dp -= 2.f * PI_F;
    return dp;
}
SMS_DEV bool transform_w(V3 w, V3 n, float eta, V3 &out) {
    if (eta == 1.f) { out = reflect(w, n); return true; }
    return refract(w, n, eta, out);
}
SMS_DEV void d_transform_w(V3 w, V3 du, V3 dv, V3 n, V3 dn_du, V3 dn_dv, float eta, V3 &ou, V3 &ov) {
    if (eta == 1.f) d_reflect(w, du, dv, n, dn_du, dn_dv, ou, ov);
    else d_refract(w, du, dv, n, dn_du, dn_dv, eta, ou, ov);
}

// sample_seed_path (manifold_ms.cpp:280-408), smooth interfaces.  The reference's bounce-0
// target-shape test compares the shape with itself (:332) and never rejects.
template <int MAXB>
SMS_DEV bool ms_seed_path(const DScene &S, int shape, V3 si_p, const EI &ei, V2 sample, bool first_path,
                          const int *seed_shapes, V3 p_start, MSPath<MAXB> &P) {
    P.n = 0;
    V3 x1 = sample_position(S, shape, sample).p;
    if (p_start.x != 0.f || p_start.y != 0.f || p_start.z != 0.f) x1 = p_start;
    V3 wo = normalize(x1 - si_p);
    if (S.cfg.mnee_init) {
        const DShape &sh = S.shapes[shape];
        wo = sh.k != 0.f ? normalize(ld3(sh.bbox_center) - si_p) : normalize(ei.p - si_p);
    }
    Ray ray = make_ray(si_p, wo);
    for (int bounce = 0; bounce < S.cfg.bounces; ++bounce) {
        SI si = ray_intersect(S, ray);
        if (!si.valid) return false;
        int flags = S.shapes[si.shape].flags;
        if (!(flags & SMS_FLAG_CAUSTIC_CASTER_MS) && !(flags & SMS_FLAG_CAUSTIC_BOUNCER)) return false;
        if (!first_path && si.shape != seed_shapes[bounce]) return false;
        MV vertex = mv_from_si(S, si);
        P.v[bounce].v = vertex;
        P.shapes[bounce] = si.shape;
        P.n = bounce + 1;
        if (!S.cfg.mnee_init) {
            V3 m = offset_normal_world(vertex);
            V3 wi = -wo;
            if (vertex.eta == 1.f) wo = reflect(wi, m);
            else if (!refract(wi, m, vertex.eta, wo)) return false;
        }
        ray = spawn_ray(si.p, wo);
    }
    return true;
}

// invert_tridiagonal_step (manifold_ms.cpp:1192-1232): block Thomas elimination
template <int MAXB>
SMS_DEV bool ms_invert_tridiagonal_step(MSPath<MAXB> &P) {
    int n = P.n;
    if (n == 0) return true;
    P.v[0].tmp = P.v[0].dprev;
    if (!invert_exact(P.v[0].dcur, P.v[0].inv_lambda)) return false;
    for (int i = 1; i < n; ++i) {
        P.v[i].tmp = mulm2(P.v[i].dprev, P.v[i - 1].inv_lambda);
        M2 m = subm2(P.v[i].dcur, mulm2(P.v[i].tmp, P.v[i - 1].dnext));
        if (!invert_exact(m, P.v[i].inv_lambda)) return false;
    }
    P.v[0].dx = P.v[0].C;
    for (int i = 1; i < n; ++i) P.v[i].dx = P.v[i].C - mulv2(P.v[i].tmp, P.v[i - 1].dx);
    P.v[n - 1].dx = mulv2(P.v[n - 1].inv_lambda, P.v[n - 1].dx);
    for (int i = n - 2; i >= 0; --i) P.v[i].dx = mulv2(P.v[i].inv_lambda, P.v[i].dx - mulv2(P.v[i].dnext, P.v[i + 1].dx));
    return true;
}

// compute_step_anglediff MS (manifold_ms.cpp:631-857)
template <int MAXB>
SMS_DEV bool ms_step_anglediff(MSPath<MAXB> &P, V3 x0, const EI &ei) {
    const int k = P.n;
    bool success = true;
    const V3 zero = v3(0.f, 0.f, 0.f);
    for (int i = 0; i < k; ++i) {
        MSVert &vi = P.v[i];
        const MV &v = vi.v;
        vi.C = v2(0.f, 0.f);
        vi.dprev = vi.dcur = vi.dnext = m2(0.f, 0.f, 0.f, 0.f);
        V3 x_prev = i == 0 ? x0 : P.v[i - 1].v.p;
        V3 x_next = i == k - 1 ? ei.p : P.v[i + 1].v.p;
        bool fixed = i == k - 1 && ei.kind == SMS_EI_DIRECTIONAL;
        V3 wo = fixed ? ei.d : x_next - v.p;
        float ilo = norm(wo);
        if (ilo < 1e-3f) return false;
        ilo = 1.f / ilo;
        wo = wo * ilo;
        V3 dwo_du_cur = zero, dwo_dv_cur = zero;
        if (!fixed) {
            dwo_du_cur = -ilo * (v.dp_du - wo * dot(wo, v.dp_du));
            dwo_dv_cur = -ilo * (v.dp_dv - wo * dot(wo, v.dp_dv));
        }
        V3 wi = x_prev - v.p;
        float ili = norm(wi);
        if (ili < 1e-3f) return false;
        ili = 1.f / ili;
        wi = wi * ili;
        V3 dwi_du_cur = -ili * (v.dp_du - wi * dot(wi, v.dp_du)), dwi_dv_cur = -ili * (v.dp_dv - wi * dot(wi, v.dp_dv));
        bool success_i = false;
        V3 n = offset_normal_world(v);
        V3 dn_du = (v.ds_du * 0.f + v.dt_du * 0.f) + v.dn_du * 1.f;
        V3 dn_dv = (v.ds_dv * 0.f + v.dt_dv * 0.f) + v.dn_dv * 1.f;
        V3 wio;
        if (transform_w(wi, n, v.eta, wio)) {
            float to, po, tio, pio;
            sphcoords(wo, to, po); sphcoords(wio, tio, pio);
            vi.C = v2(to - tio, wrap_pi(po - pio));
            float dto_du, dpo_du, dto_dv, dpo_dv, dtio_du, dpio_du, dtio_dv, dpio_dv;
            if (i > 0) {
                const MV &vp = P.v[i - 1].v;
                V3 a = ili * (vp.dp_du - wi * dot(wi, vp.dp_du)), b = ili * (vp.dp_dv - wi * dot(wi, vp.dp_dv));
                V3 c, d;
                d_transform_w(wi, a, b, n, zero, zero, v.eta, c, d);
                d_sphcoords(wio, c, d, dtio_du, dpio_du, dtio_dv, dpio_dv);
                vi.dprev = m2(-dtio_du, -dtio_dv, -dpio_du, -dpio_dv);
            }
            V3 a, b;
            d_transform_w(wi, dwi_du_cur, dwi_dv_cur, n, dn_du, dn_dv, v.eta, a, b);
            d_sphcoords(wo, dwo_du_cur, dwo_dv_cur, dto_du, dpo_du, dto_dv, dpo_dv);
            d_sphcoords(wio, a, b, dtio_du, dpio_du, dtio_dv, dpio_dv);
            vi.dcur = m2(dto_du - dtio_du, dto_dv - dtio_dv, dpo_du - dpio_du, dpo_dv - dpio_dv);
            if (i < k - 1) {
                const MV &vn = P.v[i + 1].v;
                V3 c = ilo * (vn.dp_du - wo * dot(wo, vn.dp_du)), d = ilo * (vn.dp_dv - wo * dot(wo, vn.dp_dv));
                d_sphcoords(wo, c, d, dto_du, dpo_du, dto_dv, dpo_dv);
                vi.dnext = m2(dto_du, dto_dv, dpo_du, dpo_dv);
            }
            success_i = true;
        }
        V3 woi;
        if (transform_w(wo, n, v.eta, woi) && !success_i) {
            float ti, pi, toi, poi;
            sphcoords(wi, ti, pi); sphcoords(woi, toi, poi);
            vi.C = v2(ti - toi, wrap_pi(pi - poi));
            float dti_du, dpi_du, dti_dv, dpi_dv, dtoi_du, dpoi_du, dtoi_dv, dpoi_dv;
            if (i > 0) {
                const MV &vp = P.v[i - 1].v;
                V3 a = ili * (vp.dp_du - wi * dot(wi, vp.dp_du)), b = ili * (vp.dp_dv - wi * dot(wi, vp.dp_dv));
                d_sphcoords(wi, a, b, dti_du, dpi_du, dti_dv, dpi_dv);
                vi.dprev = m2(dti_du, dti_dv, dpi_du, dpi_dv);
            }
            V3 a, b;
            d_transform_w(wo, dwo_du_cur, dwo_dv_cur, n, dn_du, dn_dv, v.eta, a, b);
            d_sphcoords(wi, dwi_du_cur, dwi_dv_cur, dti_du, dpi_du, dti_dv, dpi_dv);
            d_sphcoords(woi, a, b, dtoi_du, dpoi_du, dtoi_dv, dpoi_dv);
            vi.dcur = m2(dti_du - dtoi_du, dti_dv - dtoi_dv, dpi_du - dpoi_du, dpi_dv - dpoi_dv);
            if (i < k - 1) {
                const MV &vn = P.v[i + 1].v;
                V3 c = ilo * (vn.dp_du - wo * dot(wo, vn.dp_du)), d = ilo * (vn.dp_dv - wo * dot(wo, vn.dp_dv));
                V3 e, f;
                d_transform_w(wo, c, d, n, zero, zero, v.eta, e, f);
                d_sphcoords(woi, e, f, dtoi_du, dpoi_du, dtoi_dv, dpoi_dv);
                vi.dnext = m2(-dtoi_du, -dtoi_dv, -dpoi_du, -dpoi_dv);
            }
            success_i = true;
        }
        success = success && success_i;
    }
    if (!success || !ms_invert_tridiagonal_step(P)) return false;
    return true;
}

// compute_step_halfvector MS (manifold_ms.cpp:500-629), smooth interfaces: the constraint is
// the tangential part of the generalized half-vector
template <int MAXB>
SMS_DEV bool ms_step_halfvector(MSPath<MAXB> &P, V3 x0, const EI &ei) {
    const int k = P.n;
    for (int i = 0; i < k; ++i) {
        MSVert &vi = P.v[i];
        const MV &v = vi.v;
        vi.C = v2(0.f, 0.f);
        vi.dprev = vi.dcur = vi.dnext = m2(0.f, 0.f, 0.f, 0.f);
        V3 x_prev = i == 0 ? x0 : P.v[i - 1].v.p;
        V3 x_next = i == k - 1 ? ei.p : P.v[i + 1].v.p;
        bool fixed = i == k - 1 && ei.kind == SMS_EI_DIRECTIONAL;
        V3 wo = fixed ? ei.d : x_next - v.p;
        float ilo = norm(wo);
        if (ilo < 1e-3f) return false;
        ilo = 1.f / ilo;
        wo = wo * ilo;
        V3 wi = x_prev - v.p;
        float ili = norm(wi);
        if (ili < 1e-3f) return false;
        ili = 1.f / ili;
        wi = wi * ili;
        float eta = v.eta;
        if (dot(wi, v.gn) < 0.f) eta = 1.f / eta;
        V3 h = wi + eta * wo;
        if (eta != 1.f) h = h * -1.f;
        float ilh = 1.f / norm(h);
        h = h * ilh;
        ilo *= eta * ilh; ili *= ilh;
        V3 dh_du, dh_dv;
        if (i > 0) {
            const MV &vp = P.v[i - 1].v;
            dh_du = ili * (vp.dp_du - wi * dot(wi, vp.dp_du));
            dh_dv = ili * (vp.dp_dv - wi * dot(wi, vp.dp_dv));
            dh_du = dh_du - h * dot(dh_du, h); dh_dv = dh_dv - h * dot(dh_dv, h);
            if (eta != 1.f) { dh_du = dh_du * -1.f; dh_dv = dh_dv * -1.f; }
            vi.dprev = m2(dot(v.s, dh_du), dot(v.s, dh_dv), dot(v.t, dh_du), dot(v.t, dh_dv));
        }
        if (fixed) {
            dh_du = ili * (-v.dp_du + wi * dot(wi, v.dp_du));
            dh_dv = ili * (-v.dp_dv + wi * dot(wi, v.dp_dv));
        } else {
            dh_du = (-v.dp_du * (ili + ilo) + wi * (dot(wi, v.dp_du) * ili)) + wo * (dot(wo, v.dp_du) * ilo);
            dh_dv = (-v.dp_dv * (ili + ilo) + wi * (dot(wi, v.dp_dv) * ili)) + wo * (dot(wo, v.dp_dv) * ilo);
        }
        dh_du = dh_du - h * dot(dh_du, h); dh_dv = dh_dv - h * dot(dh_dv, h);
        if (eta != 1.f) { dh_du = dh_du * -1.f; dh_dv = dh_dv * -1.f; }
        vi.dcur = m2(dot(v.ds_du, h) + dot(v.s, dh_du), dot(v.ds_dv, h) + dot(v.s, dh_dv),
                     dot(v.dt_du, h) + dot(v.t, dh_du), dot(v.dt_dv, h) + dot(v.t, dh_dv));
        if (i < k - 1) {
            const MV &vn = P.v[i + 1].v;
            dh_du = ilo * (vn.dp_du - wo * dot(wo, vn.dp_du));
            dh_dv = ilo * (vn.dp_dv - wo * dot(wo, vn.dp_dv));
            dh_du = dh_du - h * dot(dh_du, h); dh_dv = dh_dv - h * dot(dh_dv, h);
            if (eta != 1.f) { dh_du = dh_du * -1.f; dh_dv = dh_dv * -1.f; }
            vi.dnext = m2(dot(v.s, dh_du), dot(v.s, dh_dv), dot(v.t, dh_du), dot(v.t, dh_dv));
        }
        vi.C = v2(dot(v.s, h) - 0.f, dot(v.t, h) - 0.f);
    }
    return ms_invert_tridiagonal_step(P);
}

// reproject (manifold_ms.cpp:859-918): re-trace the chain through the proposed first vertex
template <int MAXB>
SMS_DEV bool ms_reproject(const DScene &S, V3 si_p, V3 x1, const int *seed_shapes, int n_seed, MSPath<MAXB> &Q) {
    Q.n = 0;
    V3 wo = normalize(x1 - si_p);
    Ray ray = make_ray(si_p, wo);
    for (int bounce = 0; bounce < S.cfg.bounces; ++bounce) {
        SI si = ray_intersect(S, ray);
        if (!si.valid) return false;
        if (si.shape != seed_shapes[bounce]) return false;
        MV vertex = mv_from_si(S, si);
        Q.v[bounce].v = vertex;
        Q.shapes[bounce] = si.shape;
        Q.n = bounce + 1;
        V3 m = offset_normal_world(vertex);
        V3 wi = -wo;
        if (vertex.eta == 1.f) wo = reflect(wi, m);
        else if (!refract(wi, m, vertex.eta, wo)) return false;
        ray = spawn_ray(si.p, wo);
    }
    return Q.n == n_seed;
}

template <int MAXB>
struct MSWalk { bool ok, seed_ok; int iterations; MSPath<MAXB> path; int seed_shapes[MAXB]; };

// sample_path MS (manifold_ms.cpp:203-227) + newton_solver (:410-498)
template <int MAXB>
SMS_DEV void ms_sample_path(const DScene &S, int shape, V3 si_p, const EI &ei, V2 sample, bool first_path,
                            const int *seed_shapes_in, V3 p_start, MSWalk<MAXB> &W) {
    const sms_config &cfg = S.cfg;
    W.ok = false; W.seed_ok = false; W.iterations = 0;
    for (int i = 0; i < MAXB; ++i) W.seed_shapes[i] = seed_shapes_in ? seed_shapes_in[i] : -1;
    if (!ms_seed_path(S, shape, si_p, ei, sample, first_path, W.seed_shapes, p_start, W.path)) return;
    W.seed_ok = true;
    if (first_path) for (int i = 0; i < W.path.n; ++i) W.seed_shapes[i] = W.path.shapes[i];
    const int n_seed = W.path.n;
    bool success = false, needs_step_update = true;
    int iterations = 0;
    float beta = 1.f;
    while (iterations < cfg.max_iterations) {
        bool step_success = true;
        if (needs_step_update)
            step_success = cfg.halfvector_constraints ? ms_step_halfvector(W.path, si_p, ei)
                                                      : ms_step_anglediff(W.path, si_p, ei);
        if (!step_success) break;
        bool converged = true;
        for (int i = 0; i < W.path.n; ++i)
            if (norm2(W.path.v[i].C) > cfg.solver_threshold) { converged = false; break; }
        if (converged) { success = true; break; }
        const MV &v0 = W.path.v[0].v;
        V3 prop0 = v0.p - cfg.step_scale * beta * (v0.dp_du * W.path.v[0].dx.x + v0.dp_dv * W.path.v[0].dx.y);
        MSPath<MAXB> Q;
        if (!ms_reproject(S, si_p, prop0, W.seed_shapes, n_seed, Q)) {
            beta = 0.5f * beta;
            needs_step_update = false;
        } else {
            beta = fmin_(1.f, 2.f * beta);
            W.path = Q;
            needs_step_update = true;
        }
        iterations++;
    }
    W.iterations = iterations;
    if (!success) return;
    const int n = W.path.n;
    for (int i = 0; i < n; ++i) {
        V3 x_prev = i == 0 ? si_p : W.path.v[i - 1].v.p;
        V3 x_next = i == n - 1 ? ei.p : W.path.v[i + 1].v.p;
        V3 x_cur = W.path.v[i].v.p;
        bool fixed = i == n - 1 && ei.kind == SMS_EI_DIRECTIONAL;
        V3 wi = normalize(x_prev - x_cur);
        V3 wo = fixed ? ei.d : normalize(x_next - x_cur);
        float ci = dot(W.path.v[i].v.gn, wi), co = dot(W.path.v[i].v.gn, wo);
        bool refraction = ci * co < 0.f, reflection = !refraction;
        float eta = W.path.v[i].v.eta;
        if ((eta == 1.f && !reflection) || (eta != 1.f && !refraction)) return;
    }
    W.ok = true;
}

// specular_reflectance MS, smooth interfaces (manifold_ms.cpp:920-1023) + light visibility
template <int MAXB>
SMS_DEV float ms_specular_reflectance(const DScene &S, V3 si_p, const EI &ei, const MSPath<MAXB> &P) {
    if (P.n == 0) return 0.f;
    V3 wo = normalize(P.v[0].v.p - si_p);
    Ray ray = make_ray(si_p, wo);
    float val = 1.f;
    SI si = si_invalid();
    for (int k = 0; k < P.n; ++k) {
        si = ray_intersect(S, ray);
        if (!si.valid) return 0.f;
        const DShape &sh = S.shapes[si.shape];
        V3 w;
        if (k < P.n - 1) w = normalize(P.v[k + 1].v.p - si.p);
        else w = ei.kind == SMS_EI_DIRECTIONAL ? ei.d : normalize(ei.p - si.p);
        float cos_theta = dot(si.sh_n, w);
        float f;
        if (sh.k != 0.f) f = fresnel_conductor(fabsf(cos_theta), sh.eta_r, sh.k);
        else {
            float eta = sh.eta_vtx;
            if (cos_theta < 0.f) eta = 1.f / eta;
            f = (1.f - fresnel(cos_theta, eta)) * (eta * eta);
        }
        val = val * f;
        ray = make_ray(si.p, w);
    }
    if (ei.kind != SMS_EI_DIRECTIONAL) {
        V3 d = ei.p - si.p;
        ray = make_ray(si.p, normalize(d), RAY_EPS_F * (1.f + hmax_abs(si.p)), norm(d) * (1.f - RAY_EPS_F));
    }
    if (ray_test(S, ray)) return 0.f;
    return val;
}

// invert_tridiagonal_geo (manifold_ms.cpp:1234-1268)
template <int N>
SMS_DEV float ms_invert_tridiagonal_geo(MSVert (&v)[N], int n) {
    if (n == 0) return 0.f;
    M2 Li;
    if (!invert_exact(v[0].dcur, Li)) return 0.f;
    for (int i = 0; i < n - 2; ++i) {
        v[i].tmp = mulm2(Li, v[i].dnext);
        M2 m = subm2(v[i + 1].dcur, mulm2(v[i + 1].dprev, v[i].tmp));
        if (!invert_exact(m, Li)) return 0.f;
    }
    v[n - 2].inv_lambda = mulm2(negm2(Li), v[n - 2].dnext);
    for (int i = n - 3; i >= 0; --i) v[i].inv_lambda = mulm2(negm2(v[i].tmp), v[i + 1].inv_lambda);
    return fabsf(det2(negm2(v[0].inv_lambda)));
}

// geometric_term MS (manifold_ms.cpp:1025-1190)
template <int MAXB>
SMS_DEV float ms_geometric_term(const MV &vx, const MV &vy, const MSPath<MAXB> &P) {
    MSVert v[MAXB + 2];
    int k = 0;
    if (vy.fixed_direction) {
        v[k++].v = vy;
        for (int i = P.n - 1; i >= 0; --i) v[k++].v = P.v[i].v;
        v[k++].v = vx;
    } else {
        for (int i = 0; i < P.n; ++i) v[k++].v = P.v[i].v;
        v[k++].v = vy;
    }
    for (int i = 0; i < k - 1; ++i) {
        v[i].dprev = v[i].dcur = v[i].dnext = m2(0.f, 0.f, 0.f, 0.f);
        const MV c = v[i].v, nx = v[i + 1].v;
        V3 wo = nx.p - c.p;
        float ilo = norm(wo);
        if (ilo < 1e-3f) return 0.f;
        ilo = 1.f / ilo;
        wo = wo * ilo;
        if (c.fixed_direction) {
            V3 a = -ilo * (c.dp_du - wo * dot(wo, c.dp_du)), b = -ilo * (c.dp_dv - wo * dot(wo, c.dp_dv));
            v[i].dcur = m2(dot(a, c.dp_du), dot(b, c.dp_du), dot(a, c.dp_dv), dot(b, c.dp_dv));
            V3 e = ilo * (nx.dp_du - wo * dot(wo, nx.dp_du)), f = ilo * (nx.dp_dv - wo * dot(wo, nx.dp_dv));
            v[i].dnext = m2(dot(e, c.dp_du), dot(f, c.dp_du), dot(e, c.dp_dv), dot(f, c.dp_dv));
            continue;
        }
        V3 x_prev = i == 0 ? vx.p : v[i - 1].v.p;
        V3 wi = x_prev - c.p;
        float ili = norm(wi);
        if (ili < 1e-3f) return 0.f;
        ili = 1.f / ili;
        wi = wi * ili;
        float eta = c.eta;
        if (dot(wi, c.gn) < 0.f) eta = 1.f / eta;
        V3 h = wi + eta * wo;
        if (eta != 1.f) h = h * -1.f;
        float ilh = 1.f / norm(h);
        h = h * ilh;
        ilo *= eta * ilh; ili *= ilh;
        float dot_dpdu_n = dot(c.dp_du, c.n), dot_dpdv_n = dot(c.dp_dv, c.n);
        V3 s = c.dp_du - dot_dpdu_n * c.n, t = c.dp_dv - dot_dpdv_n * c.n;
        V3 dh_du, dh_dv;
        if (i > 0) {
            const MV pv = v[i - 1].v;
            dh_du = ili * (pv.dp_du - wi * dot(wi, pv.dp_du));
            dh_dv = ili * (pv.dp_dv - wi * dot(wi, pv.dp_dv));
            dh_du = dh_du - h * dot(dh_du, h); dh_dv = dh_dv - h * dot(dh_dv, h);
            if (eta != 1.f) { dh_du = dh_du * -1.f; dh_dv = dh_dv * -1.f; }
            v[i].dprev = m2(dot(s, dh_du), dot(s, dh_dv), dot(t, dh_du), dot(t, dh_dv));
        }
        dh_du = (-c.dp_du * (ili + ilo) + wi * (dot(wi, c.dp_du) * ili)) + wo * (dot(wo, c.dp_du) * ilo);
        dh_dv = (-c.dp_dv * (ili + ilo) + wi * (dot(wi, c.dp_dv) * ili)) + wo * (dot(wo, c.dp_dv) * ilo);
        dh_du = dh_du - h * dot(dh_du, h); dh_dv = dh_dv - h * dot(dh_dv, h);
        if (eta != 1.f) { dh_du = dh_du * -1.f; dh_dv = dh_dv * -1.f; }
        float dot_h_n = dot(h, c.n), dot_h_dndu = dot(h, c.dn_du), dot_h_dndv = dot(h, c.dn_dv);
        v[i].dcur = m2((dot(dh_du, s) - dot(c.dp_du, c.dn_du) * dot_h_n) - dot_dpdu_n * dot_h_dndu,
                       (dot(dh_dv, s) - dot(c.dp_du, c.dn_dv) * dot_h_n) - dot_dpdu_n * dot_h_dndv,
                       (dot(dh_du, t) - dot(c.dp_dv, c.dn_du) * dot_h_n) - dot_dpdv_n * dot_h_dndu,
                       (dot(dh_dv, t) - dot(c.dp_dv, c.dn_dv) * dot_h_n) - dot_dpdv_n * dot_h_dndv);
        dh_du = ilo * (nx.dp_du - wo * dot(wo, nx.dp_du));
        dh_dv = ilo * (nx.dp_dv - wo * dot(wo, nx.dp_dv));
        dh_du = dh_du - h * dot(dh_du, h); dh_dv = dh_dv - h * dot(dh_dv, h);
        if (eta != 1.f) { dh_du = dh_du * -1.f; dh_dv = dh_dv * -1.f; }
        v[i].dnext = m2(dot(s, dh_du), dot(s, dh_dv), dot(t, dh_du), dot(t, dh_dv));
    }
    if (vy.fixed_direction) {
        float G = ms_invert_tridiagonal_geo(v, k);
        V3 d = normalize(v[k - 1].v.p - v[k - 2].v.p);
        return G / fabsf(dot(d, v[k - 1].v.n));
    }
    float dx1_dxend = ms_invert_tridiagonal_geo(v, k);
    V3 d = vx.p - v[0].v.p;
    float inv_r2 = 1.f / sqnorm(d);
    d = d * sqrtf(inv_r2);
    float dw0_dx1 = fabsf(dot(d, v[0].v.gn)) * inv_r2;
    return dw0_dx1 * dx1_dxend;
}

// evaluate_path_contribution MS (manifold_ms.cpp:229-278); P is taken by value because the
// chain vertices are made orthonormal here
template <int MAXB>
SMS_DEV V3 ms_path_contribution(const DScene &S, const SI &si, const EI &ei_, MSPath<MAXB> P) {
    const MV last = P.v[P.n - 1].v;
    EI ei = ei_;
    if (ei.kind == SMS_EI_POINT) {
        const DEmitter &em = S.emitters[ei.emitter];
        V3 pos = ld3(em.position);
        V3 d = pos - last.p;
        float dist = norm(d), inv_dist = 1.f / dist;
        d = d * inv_dist;
        float emitter_pdf = ei.pdf;
        ei.p = pos; ei.d = d; ei.n = d; ei.pdf = 1.f;
        ei.weight = (ld3(em.intensity) * (inv_dist * inv_dist)) * dist * dist;
        ei.pdf *= emitter_pdf;
        ei.weight = ei.weight * (1.f / emitter_pdf);
    }
    MV vy;
    if (!emitter_interaction_to_vertex(S, ei, last.p, vy)) return v3(0.f, 0.f, 0.f);
    for (int k = 0; k < P.n; ++k) make_orthonormal(P.v[k].v);
    MV vx = mv_from_si(S, si);
    make_orthonormal(vx);
    float f = ms_specular_reflectance(S, si.p, ei, P);
    float G = ms_geometric_term(vx, vy, P);
    float thr = f * G;
    return v3(thr * ei.weight.x, thr * ei.weight.y, thr * ei.weight.z);
}

} // namespace smsdev
