// dev_path.h -- the path integrators around the single-scatter SMS term on gfx950:
// path_sms_ss (src/integrators/path_sms_ss.cpp:69-251) and path_filtered_ss
// (src/integrators/path_filtered_ss.cpp:37-194) over diffuse, smooth dielectric and smooth
// conductor BSDFs, one lane per camera sample.
//
// A sample's numbers come from PCG32 stream PATH_STREAM of (seed, sample index): the film
// jitter, then per bounce the Russian-roulette sample (1D), the emitter sample (2D), the BSDF
// sample (1D + 2D) and, where path_sms_ss checks a BSDF-sampled caustic against MNEE, the
// walk's seed sample (2D, and the LEAN sample with twostage), in the reference's order.  The
// lane writes the path's own terms, and, into fixed per-sample slots, the SMS queries of its
// caustic receivers (with the path throughput there) and its MNEE checks; the SMS estimates
// and the MNEE walks run afterwards as batches over the compacted slots.
#pragma once
#include "dev_scene.h"
#include "dev_lean.h"

namespace smsdev {

constexpr uint64_t PATH_STREAM = 0xFFFFFFFDull;

// warp::square_to_uniform_disk_concentric (include/mitsuba/core/warp.h:54-90)
SMS_DEV V2 square_to_uniform_disk_concentric(V2 s) {
    const float x = fmaf(2.f, s.x, -1.f), y = fmaf(2.f, s.y, -1.f);
    const bool is_zero = x == 0.f && y == 0.f, q13 = fabsf(x) < fabsf(y);
    const float r = q13 ? y : x, rp = q13 ? x : y;
    float phi = .25f * PI_F * rp / r;
    if (q13) phi = .5f * PI_F - phi;
    if (is_zero) phi = 0.f;
    float sn, cs;
    c_sincosf(phi, &sn, &cs);
    return v2(r * cs, r * sn);
}

// warp::square_to_cosine_hemisphere (warp.h:332-340)
SMS_DEV V3 square_to_cosine_hemisphere(V2 s) {
    const V2 p = square_to_uniform_disk_concentric(s);
    return v3(p.x, p.y, safe_sqrt(1.f - (p.x * p.x + p.y * p.y)));
}

struct DirSample { V3 p, n, d; float dist, pdf; bool delta; int emitter; };

SMS_DEV DirSample dir_sample_zero() {
    DirSample ds;
    ds.p = ds.n = ds.d = v3(0.f, 0.f, 0.f); ds.dist = 0.f; ds.pdf = 0.f; ds.delta = false; ds.emitter = -1;
    return ds;
}

// Sphere::sample_direction (src/shapes/sphere.cpp:177-252): a uniform direction in the cone the
// sphere subtends seen from outside, else a uniform point on the sphere
SMS_DEV DirSample sphere_sample_direction(const DShape &sh, V3 ref, V2 sample) {
    DirSample ds = dir_sample_zero();
    const V3 c = v3(sh.cx, sh.cy, sh.cz);
    const V3 dc_v = c - ref;
    const float dc_2 = dot(dc_v, dc_v);
    const float radius_adj = sh.radius * (1.f - RAY_EPS_F);
    if (dc_2 > radius_adj * radius_adj) {
        const float inv_dc = 1.f / sqrtf(dc_2), sin_theta_max = sh.radius * inv_dc,
                    sin_theta_max_2 = sin_theta_max * sin_theta_max, inv_sin_theta_max = 1.f / sin_theta_max,
                    cos_theta_max = safe_sqrt(1.f - sin_theta_max_2);
        const float a = fmaf(cos_theta_max - 1.f, sample.x, 1.f);
        const float sin_theta_2 = sin_theta_max_2 > 0.00068523f ? 1.f - a * a : sin_theta_max_2 * sample.x;
        const float cos_theta = safe_sqrt(1.f - sin_theta_2);
        const float cos_alpha = sin_theta_2 * inv_sin_theta_max +
                                cos_theta * safe_sqrt(fmaf(-sin_theta_2, inv_sin_theta_max * inv_sin_theta_max, 1.f));
        const float sin_alpha = safe_sqrt(fmaf(-cos_alpha, cos_alpha, 1.f));
        float sin_phi, cos_phi;
        c_sincosf(sample.y * (2.f * PI_F), &sin_phi, &cos_phi);
        const V3 fn = dc_v * -inv_dc;
        V3 fs, ft;
        coordinate_system(fn, fs, ft);
        const V3 d = to_world(fs, ft, fn, v3(cos_phi * sin_alpha, sin_phi * sin_alpha, cos_alpha));
        ds.p = fmadd3(d, sh.radius, c);
        ds.n = d;
        ds.d = ds.p - ref;
        const float dist2 = dot(ds.d, ds.d);
        ds.dist = sqrtf(dist2);
        ds.d = ds.d / ds.dist;
        ds.pdf = INV_TWOPI_F / (1.f - cos_theta_max);
        if (ds.dist == 0.f) ds.pdf = 0.f;
    } else {
        const V3 d = square_to_uniform_sphere(sample);
        ds.p = fmadd3(d, sh.radius, c);
        ds.n = d;
        ds.d = ds.p - ref;
        const float dist2 = dot(ds.d, ds.d);
        ds.dist = sqrtf(dist2);
        ds.d = ds.d / ds.dist;
        ds.pdf = sh.inv_surface_area * dist2 / fabsf(dot(ds.d, ds.n));
    }
    ds.delta = sh.radius == 0.f;
    return ds;
}

// Shape::sample_direction (src/librender/shape.cpp:259-276) of shapes without their own
// (rectangles, meshes): a point of the shape's area sampling and the solid-angle conversion
SMS_DEV DirSample shape_sample_direction(const DScene &S, int shape, V3 ref, V2 sample) {
    DirSample ds = dir_sample_zero();
    const PS ps = sample_position(S, shape, sample);
    ds.p = ps.p; ds.n = ps.n; ds.pdf = ps.pdf;
    ds.d = ds.p - ref;
    const float dist2 = dot(ds.d, ds.d);
    ds.dist = sqrtf(dist2);
    ds.d = ds.d / ds.dist;
    const float dp = fabsf(dot(ds.d, ds.n));
    ds.pdf = dp != 0.f ? ds.pdf * (dist2 / dp) : 0.f;
    return ds;
}

// Scene::sample_emitter_direction (src/librender/scene.cpp:156-203) over AreaLight
// (area.cpp:104-117), PointLight (point.cpp:74-99) and DirectionalEmitter (directional.cpp:
// 105-128); the visibility ray only when asked (a delta BSDF evaluates to zero whatever it sees)
SMS_DEV DirSample scene_sample_emitter_direction(const DScene &S, V3 ref, V2 sample, bool test_visibility, V3 &spec) {
    DirSample ds = dir_sample_zero();
    spec = v3(0.f, 0.f, 0.f);
    const int n = S.n_emitters;
    if (n == 0) return ds;
    int index = 0;
    float emitter_pdf = 1.f;
    if (n > 1) {
        emitter_pdf = 1.f / (float) n;
        const uint32_t idx = (uint32_t) (sample.x * (float) n);
        index = (int) (idx < (uint32_t) (n - 1) ? idx : (uint32_t) (n - 1));
        sample.x = (sample.x - (float) index * emitter_pdf) * (float) n;
    }
    const DEmitter &em = S.emitters[index];
    if (em.type == SMS_EMITTER_AREA) {
        const DShape &es = S.shapes[em.shape];
        ds = es.type == SMS_SHAPE_SPHERE ? sphere_sample_direction(es, ref, sample)
                                         : shape_sample_direction(S, em.shape, ref, sample);
        if (dot(ds.d, ds.n) < 0.f && ds.pdf != 0.f)
            spec = v3(em.radiance[0] / ds.pdf, em.radiance[1] / ds.pdf, em.radiance[2] / ds.pdf);
    } else if (em.type == SMS_EMITTER_POINT) {
        ds.p = ld3(em.position); ds.pdf = 1.f; ds.delta = true;
        ds.d = ds.p - ref;
        ds.dist = norm(ds.d);
        const float inv_dist = 1.f / ds.dist;
        ds.d = ds.d * inv_dist;
        spec = ld3(em.intensity) * (inv_dist * inv_dist);
    } else {
        const V3 d = ld3(em.direction);
        const float dist = S.bs_dist;
        ds.p = ref - d * dist; ds.n = d; ds.pdf = 1.f; ds.delta = true; ds.d = d * -1.f; ds.dist = dist;
        spec = ld3(em.irradiance);
    }
    ds.emitter = index;
    if (n > 1) {
        ds.pdf = ds.pdf * emitter_pdf;
        spec = spec * (1.f / emitter_pdf);
    }
    if (test_visibility && ds.pdf != 0.f) {
        const Ray r = make_ray(ref, ds.d, RAY_EPS_F * (1.f + hmax_abs(ref)), ds.dist * (1.f - SHADOW_EPS_F));
        if (ray_test(S, r)) spec = v3(0.f, 0.f, 0.f);
    }
    return ds;
}

// Scene::pdf_emitter_direction (scene.cpp:205-219) of a BSDF-sampled hit on an area emitter:
// AreaLight::pdf_direction (area.cpp:119-125), Sphere::pdf_direction (sphere.cpp:254-267) or
// Shape::pdf_direction (shape.cpp:278-289: pdf_position, rectangle.cpp:130-133 and
// mesh.cpp:588-591, times dist^2 / |cos|), with DirectionSample3f(hit, ref) (records.h:168-174)
SMS_DEV float pdf_emitter_hit(const DScene &S, int emitter, const SI &hit, V3 ref) {
    const DEmitter &em = S.emitters[emitter];
    if (em.type != SMS_EMITTER_AREA) return 0.f;
    const DShape &sh = S.shapes[em.shape];
    V3 d = hit.p - ref;
    const float dist = norm(d);
    d = d / dist;
    if (!(dot(d, hit.sh_n) < 0.f)) return 0.f;
    if (sh.type != SMS_SHAPE_SPHERE) {
        const float dp = fabsf(dot(d, hit.sh_n));
        const float pdf_pos = sh.type == SMS_SHAPE_RECTANGLE ? sh.inv_surface_area : sh.area_norm;
        float pdf = dp != 0.f ? pdf_pos * ((dist * dist) / dp) : 0.f;
        if (S.n_emitters > 1) pdf = pdf * (1.f / (float) S.n_emitters);
        return pdf;
    }
    const float sin_alpha = sh.radius * (1.f / norm(v3(sh.cx, sh.cy, sh.cz) - ref));
    const float cos_alpha = safe_sqrt(1.f - sin_alpha * sin_alpha);
    float pdf = sin_alpha < 0.99999994039535522461f ? INV_TWOPI_F / (1.f - cos_alpha)
                                                    : sh.inv_surface_area * (dist * dist) / fabsf(dot(d, hit.sh_n));
    if (S.n_emitters > 1) pdf = pdf * (1.f / (float) S.n_emitters);
    return pdf;
}

// AreaLight::eval (area.cpp:71-79) at a surface interaction on its shape
SMS_DEV V3 emitter_eval_hit(const DScene &S, int emitter, const SI &hit) {
    const DEmitter &em = S.emitters[emitter];
    if (em.type != SMS_EMITTER_AREA || !(hit.wi.z > 0.f)) return v3(0.f, 0.f, 0.f);
    return ld3(em.radiance);
}

struct BsdfSample { V3 wo, weight; float pdf, eta; bool delta, reflection, transmission; };

// BSDF::sample of SmoothDiffuse (diffuse.cpp:77-101), SmoothDielectric (dielectric.cpp:201-313)
// and SmoothConductor (conductor.cpp:217-275), local shading frame
SMS_DEV BsdfSample bsdf_sample(const DShape &sh, V3 wi, float s1, V2 s2) {
    BsdfSample bs;
    bs.wo = v3(0.f, 0.f, 0.f); bs.weight = v3(0.f, 0.f, 0.f); bs.pdf = 0.f; bs.eta = 1.f;
    bs.delta = sh.bsdf != SMS_BSDF_DIFFUSE; bs.reflection = true; bs.transmission = false;
    const float cos_i = wi.z;
    if (sh.bsdf == SMS_BSDF_DIFFUSE) {
        if (!(cos_i > 0.f)) return bs;
        bs.wo = square_to_cosine_hemisphere(s2);
        bs.pdf = INV_PI_F * bs.wo.z;
        if (bs.pdf > 0.f) bs.weight = ld3(sh.refl);
    } else if (sh.bsdf == SMS_BSDF_DIELECTRIC) {
        const float eta = sh.eta_r;
        const bool outside = cos_i >= 0.f;
        const float rcp_eta = 1.f / eta, eta_it = outside ? eta : rcp_eta, eta_ti = outside ? rcp_eta : eta;
        const float cos_t_sqr = fmaf(-fmaf(-cos_i, cos_i, 1.f), eta_ti * eta_ti, 1.f);
        const float cos_t_abs = safe_sqrt(cos_t_sqr);
        const float F = fresnel(cos_i, eta);
        if (s1 <= F) {
            bs.wo = v3(-wi.x, -wi.y, wi.z); bs.pdf = F; bs.weight = v3(1.f, 1.f, 1.f);
        } else {
            const float cos_t = outside ? -cos_t_abs : cos_t_abs;
            bs.wo = v3(-eta_ti * wi.x, -eta_ti * wi.y, cos_t);
            bs.pdf = 1.f - F; bs.eta = eta_it; bs.reflection = false; bs.transmission = true;
            const float f = eta_ti * eta_ti;
            bs.weight = v3(f, f, f);
        }
    } else {
        if (!(cos_i > 0.f)) return bs;
        bs.wo = v3(-wi.x, -wi.y, wi.z); bs.pdf = 1.f;
        const float F = fresnel_conductor(cos_i, sh.eta_r, sh.k);
        bs.weight = v3(F, F, F);
    }
    return bs;
}

// NormalmapBSDF::frame (src/bsdfs/normalmap.cpp:218-239) at smoothing 0: the shading frame
// perturbed by the map, as ManifoldVertex builds it
SMS_DEV void normalmap_frame(const DScene &S, const SI &si, V3 &n, V3 &s, V3 &t) {
    const DShape &sh = S.shapes[si.shape];
    const V3 bn = si.sh_n;
    const V3 bs = normalize(fnmadd3(bn, dot(bn, si.dp_du), si.dp_du));
    const V3 bt = cross(bn, bs);
    const DNormalmap nm = S.nmaps[sh.normalmap];
    V3 nl, du, dv;
    nm_local_normal(nm, tile_uv(si.uv, sh.tiles), nl, du, dv);
    n = normalize(to_world(bs, bt, bn, nl));
    s = normalize(si.dp_du - n * dot(n, si.dp_du));
    t = cross(n, s);
}

// NormalmapBSDF::sample (bsdfs/normalmap.cpp:117-134, without the LEAN fallback): the nested
// smooth dielectric (k = 0) or conductor sampled in the perturbed frame, the direction taken
// back to the shading frame, rejected when the two frames put it on different sides
SMS_DEV BsdfSample bsdf_sample_normalmap(const DScene &S, const SI &si, float s1, V2 s2) {
    const DShape &sh = S.shapes[si.shape];
    DShape nested = sh;
    nested.bsdf = sh.k == 0.f ? SMS_BSDF_DIELECTRIC : SMS_BSDF_CONDUCTOR;
    V3 pn, ps, pt;
    normalmap_frame(S, si, pn, ps, pt);
    const V3 wi_w = to_world(si.sh_s, si.sh_t, si.sh_n, si.wi);
    const V3 wi_p = v3(dot(wi_w, ps), dot(wi_w, pt), dot(wi_w, pn));
    BsdfSample bs = bsdf_sample(nested, wi_p, s1, s2);
    if (bs.weight.x == 0.f && bs.weight.y == 0.f && bs.weight.z == 0.f) return bs;
    const V3 wo_w = to_world(ps, pt, pn, bs.wo);
    const V3 wo = v3(dot(wo_w, si.sh_s), dot(wo_w, si.sh_t), dot(wo_w, si.sh_n));
    if (!(bs.wo.z * wo.z > 0.f)) bs.weight = v3(0.f, 0.f, 0.f);
    bs.wo = wo;
    return bs;
}

// SmoothDiffuse::eval / pdf (diffuse.cpp:103-133); the delta BSDFs evaluate to zero
SMS_DEV V3 bsdf_eval(const DShape &sh, V3 wi, V3 wo) {
    if (sh.bsdf != SMS_BSDF_DIFFUSE || !(wi.z > 0.f && wo.z > 0.f)) return v3(0.f, 0.f, 0.f);
    return v3((sh.refl[0] * INV_PI_F) * wo.z, (sh.refl[1] * INV_PI_F) * wo.z, (sh.refl[2] * INV_PI_F) * wo.z);
}
SMS_DEV float bsdf_pdf(const DShape &sh, V3 wi, V3 wo) {
    if (sh.bsdf != SMS_BSDF_DIFFUSE || !(wi.z > 0.f && wo.z > 0.f)) return 0.f;
    return INV_PI_F * wo.z;
}

SMS_DEV V3 mul3(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }

// film position and primary ray of sample idx (pixel idx / spp, jitter from the path stream)
SMS_DEV Ray path_camera_ray(const DScene &S, int64_t idx, int spp, Pcg32 &r, float &pos_x, float &pos_y) {
    const DCamera &c = S.cam;
    const int64_t pixel = idx / spp;
    const int px = (int) (pixel % c.w), py = (int) (pixel / c.w);
    const float sx = pcg32_float(r), sy = pcg32_float(r);
    pos_x = (float) px + sx; pos_y = (float) py + sy;
    const V2 p = v2(((float) px + sx) / (float) c.w, ((float) py + sy) / (float) c.h);
    const V3 near_p = v3((1.f - 2.f * p.x) * c.tan_half, (1.f - 2.f * p.y) * c.rect_y, 1.f);
    const V3 dloc = normalize(near_p);
    const V3 d = to_world(ld3(c.left), ld3(c.up), ld3(c.dir), dloc);
    return make_ray(ld3(c.o), d, 1e-2f * (1.f / dloc.z), 1e4f * (1.f / dloc.z));
}

// per-sample slot arrays of a pass batch
struct PathSlots {
    int depth_slots;                 // SMS / MNEE slots per sample
    sms_ss_query *q;                 // [i * depth_slots + k]
    float *q_thr;                    // x 3
    int *nq;                         // SMS queries per sample
    sms_ss_walk_in *m;               // MNEE walks
    float *m_val;                    // x 3: throughput * emitter radiance
    float *m_dir;                    // x 3: direction receiver -> caster vertex of the path
    float *m_ref;                    // x 3: receiver position
    int *nm;                         // MNEE checks per sample
    sms_ms_walk_in *mm;              // multi-scatter MNEE checks (values in m_val, counts in nm)
    float *mm_pos;                   // x 3 SMS_MS_MAX_BOUNCES: the chain the path tracer found
    int *mm_npos;
    float *L;                        // x 3: the path's own terms
    float *pos;                      // x 2: film position
};

SMS_DEV void sms_query_from_si(const SI &si, sms_ss_query &q) {
    q.shape = si.shape;
    st3(q.p, si.p); st3(q.n, si.n); st3(q.sh_n, si.sh_n); st3(q.sh_s, si.sh_s); st3(q.sh_t, si.sh_t);
    st3(q.wi, si.wi);
}

// The multi-scatter integrators' record of the specular chain a path is on: MNEEHelper
// (path_sms_ms.cpp:54-143: receiver, first caster, the chain's positions) and PathRecord
// (path_filtered_ms.cpp:237-300: the chain's scattering types); states 0 Default, 1 HitReceiver,
// 2 HitCaster, 3 HitEmitter; depth_receiver = the path depth of the chain's receiver.
struct ChainRecord {
    int state, bounce, shape0, npos, depth_receiver;
    V3 endpoint, pos[SMS_MS_MAX_BOUNCES];
};
SMS_DEV void chain_reset(ChainRecord &c) { c.state = 0; c.bounce = -1; c.shape0 = -1; c.npos = 0; }

// caustic_eta of BSDF::ior (path_filtered_ss.cpp:85-87, path_filtered_ms.cpp:268-270); a
// normal-mapped BSDF gives its nested BSDF's (bsdfs/normalmap.cpp:342-344)
SMS_DEV float caustic_eta_of(const DShape &sh) {
    if (sh.bsdf == SMS_BSDF_DIELECTRIC) return sh.eta_r;
    if ((sh.bsdf == SMS_BSDF_CONDUCTOR || sh.bsdf == SMS_BSDF_NORMALMAP) && sh.k == 0.f) return sh.eta_r;
    return 1.f;
}

// MNEEHelper::state_transition (path_sms_ms.cpp:85-138) to `next` at path depth `depth`
SMS_DEV void mnee_transition(const DScene &S, ChainRecord &c, const SI &next, int depth) {
    const int bounces = S.cfg.bounces;
    const int flags = next.valid ? S.shapes[next.shape].flags : 0;
    if (next.valid && (flags & SMS_FLAG_CAUSTIC_RECEIVER)) {
        chain_reset(c); c.state = 1; c.endpoint = next.p; c.depth_receiver = depth;
        return;
    }
    const int e = next.valid ? S.shapes[next.shape].emitter : -1;
    if (c.state == 1) {
        if (next.valid && (flags & SMS_FLAG_CAUSTIC_CASTER_MS)) {
            c.shape0 = next.shape; c.pos[0] = next.p; c.npos = 1; c.bounce = 1; c.state = 2;
        } else chain_reset(c);
    } else if (c.state == 2) {
        if (c.bounce < bounces && next.valid && (flags & (SMS_FLAG_CAUSTIC_CASTER_MS | SMS_FLAG_CAUSTIC_BOUNCER))) {
            c.pos[c.npos < SMS_MS_MAX_BOUNCES ? c.npos : SMS_MS_MAX_BOUNCES - 1] = next.p;
            c.npos++; c.bounce++;
        } else if (c.bounce == bounces && e >= 0 && S.emitters[e].ms) {
            c.state = 3;
        } else chain_reset(c);
    } else chain_reset(c);
}

// PathRecord::state_transition (path_filtered_ms.cpp:248-299) from `cur` (BSDF-sampled as a
// reflection / transmission) to `next` at path depth `depth`
SMS_DEV void record_transition(const DScene &S, ChainRecord &c, const SI &cur, const SI &next, bool refl, bool trans,
                               int depth) {
    const int bounces = S.cfg.bounces;
    const int flags = next.valid ? S.shapes[next.shape].flags : 0;
    if (next.valid && (flags & SMS_FLAG_CAUSTIC_RECEIVER)) {
        chain_reset(c); c.state = 1; c.depth_receiver = depth;
        return;
    }
    const int e = next.valid ? S.shapes[next.shape].emitter : -1;
    if (c.state == 1) {
        if (next.valid && (flags & SMS_FLAG_CAUSTIC_CASTER_MS)) { c.bounce = 1; c.state = 2; }
        else chain_reset(c);
    } else if (c.state == 2) {
        if (c.bounce < bounces && next.valid && (flags & (SMS_FLAG_CAUSTIC_CASTER_MS | SMS_FLAG_CAUSTIC_BOUNCER))) {
            const float eta = caustic_eta_of(S.shapes[cur.shape]);
            if ((eta == 1.f && trans) || (eta != 1.f && refl)) chain_reset(c);
            else c.bounce++;
        } else if (c.bounce == bounces && e >= 0 && S.emitters[e].ms) {
            c.state = 3;
        } else chain_reset(c);
    } else chain_reset(c);
}

// One sample of path_sms_ss / path_filtered_ss / path_sms_ms / path_filtered_ms
// (SamplingIntegrator::sample); j = index of the sample in the batch (slot row), idx = index in
// the pass (stream)
SMS_DEV void path_sample(const DScene &S, const sms_path_config &pc, uint64_t seed, int64_t idx, int64_t j, int spp,
                         const PathSlots &P) {
    const bool ms = pc.integrator == SMS_PATH_SMS_MS || pc.integrator == SMS_PATH_FILTERED_MS;
    const bool sms = pc.integrator == SMS_PATH_SMS_SS || pc.integrator == SMS_PATH_SMS_MS, only = pc.caustics_only != 0;
    const int depth_cap = pc.max_depth < 0 ? SMS_PATH_MAX_DEPTH : pc.max_depth;
    const int chain = ms ? S.cfg.bounces : 1;           // specular vertices between receiver and light
    Pcg32 r;
    stream_seed(r, query_key(seed, (uint64_t) idx), PATH_STREAM);
    float pos_x, pos_y;
    Ray ray = path_camera_ray(S, idx, spp, r, pos_x, pos_y);
    V3 L = v3(0.f, 0.f, 0.f), thr = v3(1.f, 1.f, 1.f);
    float eta = 1.f;
    int nq = 0, nm = 0;
    SI si = ray_intersect(S, ray);
    SI si_prev = si_invalid();
    int first_diffuse = -1;
    bool specular_camera_path = true;
    ChainRecord rec;
    chain_reset(rec);
    rec.depth_receiver = -1; rec.endpoint = v3(0.f, 0.f, 0.f);
    if (ms) {
        if (sms) mnee_transition(S, rec, si, 1);
        else record_transition(S, rec, si, si, false, false, 1);
    }
    if (si.valid && !only) {
        const int e = S.shapes[si.shape].emitter;
        if (e >= 0) L = L + emitter_eval_hit(S, e, si);
    }
    for (int depth = 1;; ++depth) {
        if (!si.valid) break;
        if (depth > pc.rr_depth) {
            float m = thr.x < thr.y ? thr.y : thr.x;          // std::max / std::min order
            m = m < thr.z ? thr.z : m;
            float q = m * (eta * eta);
            q = .95f < q ? .95f : q;
            if (pcg32_float(r) > q) break;
            thr = thr * (1.f / q);
        }
        if (depth >= depth_cap) break;
        if (only && first_diffuse >= 0 && depth > first_diffuse + chain) break;     // nothing left to count
        const DShape &sh = S.shapes[si.shape];
        const bool smooth = sh.bsdf == SMS_BSDF_DIFFUSE;
        if (first_diffuse < 0 && smooth) first_diffuse = depth;
        const bool receiver = (sh.flags & SMS_FLAG_CAUSTIC_RECEIVER) != 0;
        const bool from_receiver = si_prev.valid && (S.shapes[si_prev.shape].flags & SMS_FLAG_CAUSTIC_RECEIVER);
        const bool glint_case = !ms && depth == 1 && !pc.disable_reflected_emitters;
        bool on_caster;
        float caustic_eta = 1.f;
        if (ms) {
            on_caster = (sh.flags & (SMS_FLAG_CAUSTIC_CASTER_MS | SMS_FLAG_CAUSTIC_BOUNCER)) != 0;
            caustic_eta = caustic_eta_of(sh);
        } else if (sms) {
            on_caster = (sh.flags & SMS_FLAG_CAUSTIC_CASTER_SS) != 0;
        } else {
            on_caster = (sh.flags & (SMS_FLAG_CAUSTIC_CASTER_SS | SMS_FLAG_CAUSTIC_CASTER_MS)) != 0;
            caustic_eta = caustic_eta_of(sh);
        }
        // SMS at a caustic receiver (path_sms_ss.cpp:118-124, path_sms_ms.cpp:228-234)
        if (sms && receiver && !on_caster && (pc.max_depth < 0 || depth + chain < pc.max_depth) &&
            (!only || depth == first_diffuse) && nq < P.depth_slots) {
            const int64_t k = j * P.depth_slots + nq;
            sms_ss_query q;
            float *z = (float *) &q;
            for (int a = 0; a < (int) (sizeof(q) / 4); ++a) z[a] = 0.f;
            sms_query_from_si(si, q);
            P.q[k] = q;
            st3(P.q_thr + 3 * k, thr);
            ++nq;
        }
        // emitter sampling (path_sms_ss.cpp:142-161, path_filtered_ss.cpp:107-142,
        // path_sms_ms.cpp:240-270, path_filtered_ms.cpp:97-140)
        bool nee, valid_connection = false;
        if (sms) {
            nee = (smooth && !on_caster) || glint_case;
        } else {
            valid_connection = ms ? (rec.state == 2 && rec.bounce == S.cfg.bounces)
                                  : on_caster && (from_receiver || glint_case);
            nee = smooth && (!on_caster || valid_connection);
        }
        if (nee) {
            V2 s2;
            s2.x = pcg32_float(r); s2.y = pcg32_float(r);
            V3 ew;
            const DirSample ds = scene_sample_emitter_direction(S, si.p, s2, smooth, ew);
            if (ds.pdf != 0.f) {
                const V3 wo = v3(dot(ds.d, si.sh_s), dot(ds.d, si.sh_t), dot(ds.d, si.sh_n));
                const V3 val = bsdf_eval(sh, si.wi, wo);
                const float bpdf = bsdf_pdf(sh, si.wi, wo);
                float mis = ds.delta ? 1.f : mis_power(ds.pdf, bpdf);
                if (!sms && valid_connection && (ms || depth > 1)) {
                    const bool refraction = wo.z * si.wi.z < 0.f, reflection = !refraction;
                    if ((caustic_eta == 1.f && !reflection) || (caustic_eta != 1.f && !refraction)) mis = 0.f;
                }
                if (!only) L = L + mul3(mul3(thr * mis, val), ew);
            }
        }
        // BSDF sampling
        const float s1 = pcg32_float(r);
        V2 s2;
        s2.x = pcg32_float(r); s2.y = pcg32_float(r);
        const BsdfSample bs = sh.bsdf == SMS_BSDF_NORMALMAP ? bsdf_sample_normalmap(S, si, s1, s2)
                                                            : bsdf_sample(sh, si.wi, s1, s2);
        thr = mul3(thr, bs.weight);
        eta = eta * bs.eta;
        if (!bs.delta) specular_camera_path = false;
        if (thr.x == 0.f && thr.y == 0.f && thr.z == 0.f) break;
        ray = spawn_ray(si.p, to_world(si.sh_s, si.sh_t, si.sh_n, bs.wo));
        const SI hit = ray_intersect(S, ray);
        if (ms) {
            if (sms) mnee_transition(S, rec, hit, depth + 1);
            else record_transition(S, rec, si, hit, bs.reflection, bs.transmission, depth + 1);
        }
        const int e = hit.valid ? S.shapes[hit.shape].emitter : -1;
        if (e >= 0) {
            if (ms && sms) {
                if (!on_caster || specular_camera_path) {
                    if (!only) {
                        const V3 ev = emitter_eval_hit(S, e, hit);
                        const float epdf = bs.delta ? 0.f : pdf_emitter_hit(S, e, hit, si.p);
                        const float mis = mis_power(bs.pdf, epdf);
                        L = L + mul3(thr * mis, ev);
                    }
                } else if (S.cfg.mnee_init && !pc.biased_mnee && rec.state == 3 && rec.bounce == S.cfg.bounces) {
                    // MNEE check (path_sms_ms.cpp:290-329): the walk from the chain's receiver
                    // through its first caster, compared vertex by vertex with the chain found
                    sms_ms_walk_in w;
                    float *z = (float *) &w;
                    for (int a = 0; a < (int) (sizeof(w) / 4); ++a) z[a] = 0.f;
                    w.sample[0] = pcg32_float(r); w.sample[1] = pcg32_float(r);
                    if ((!only || rec.depth_receiver == first_diffuse) && P.mm && nm < P.depth_slots) {
                        const int64_t k = j * P.depth_slots + nm;
                        st3(w.si_p, rec.endpoint);
                        w.shape = rec.shape0;
                        const V3 ed = normalize(hit.p - rec.endpoint);
                        st3(w.ei.p, hit.p); st3(w.ei.n, hit.n); st3(w.ei.d, ed);
                        w.ei.kind = SMS_EI_AREA; w.ei.emitter = e;
                        w.first_path = 1;
                        P.mm[k] = w;
                        st3(P.m_val + 3 * k, mul3(thr, emitter_eval_hit(S, e, hit)));
                        const int np = rec.npos < SMS_MS_MAX_BOUNCES ? rec.npos : SMS_MS_MAX_BOUNCES;
                        for (int v = 0; v < np; ++v) st3(P.mm_pos + 3 * (SMS_MS_MAX_BOUNCES * k + v), rec.pos[v]);
                        P.mm_npos[k] = np;
                        ++nm;
                    }
                }
            } else if (ms) {
                const bool light_hit_valid = rec.state == 3 && rec.bounce == S.cfg.bounces;
                if (!on_caster || light_hit_valid || specular_camera_path) {
                    const bool take = !only || (light_hit_valid && rec.depth_receiver == first_diffuse);
                    if (take) {
                        const V3 ev = emitter_eval_hit(S, e, hit);
                        const float epdf = bs.delta ? 0.f : pdf_emitter_hit(S, e, hit, si.p);
                        const float mis = mis_power(bs.pdf, epdf);
                        L = L + mul3(thr * mis, ev);
                    }
                }
            } else if (sms) {
                if (!on_caster || glint_case) {
                    if (!only) {
                        const V3 ev = emitter_eval_hit(S, e, hit);
                        const float epdf = bs.delta ? 0.f : pdf_emitter_hit(S, e, hit, si.p);
                        const float mis = mis_power(bs.pdf, epdf);
                        L = L + mul3(thr * mis, ev);
                    }
                } else if (S.cfg.mnee_init && !pc.biased_mnee && from_receiver && S.emitters[e].ss) {
                    // MNEE completion (path_sms_ss.cpp:196-241): a walk from the receiver
                    sms_ss_walk_in w;
                    float *z = (float *) &w;
                    for (int a = 0; a < (int) (sizeof(w) / 4); ++a) z[a] = 0.f;
                    w.sample[0] = pcg32_float(r); w.sample[1] = pcg32_float(r);
                    if (S.cfg.twostage) { w.lean_sample[0] = pcg32_float(r); w.lean_sample[1] = pcg32_float(r); }
                    if ((!only || depth - 1 == first_diffuse) && P.m && nm < P.depth_slots) {
                        const int64_t k = j * P.depth_slots + nm;
                        st3(w.si_p, si_prev.p);
                        w.shape = si.shape;
                        const V3 ed = normalize(hit.p - si_prev.p);
                        st3(w.ei.p, hit.p); st3(w.ei.n, hit.n); st3(w.ei.d, ed);
                        w.ei.kind = SMS_EI_AREA; w.ei.emitter = e;
                        w.n_offset[2] = 1.f;
                        P.m[k] = w;
                        st3(P.m_val + 3 * k, mul3(thr, emitter_eval_hit(S, e, hit)));
                        st3(P.m_dir + 3 * k, normalize(si.p - si_prev.p));
                        st3(P.m_ref + 3 * k, si_prev.p);
                        ++nm;
                    }
                }
            } else {
                const bool r_t_valid = (caustic_eta == 1.f && bs.reflection) || (caustic_eta != 1.f && bs.transmission);
                const bool vc = on_caster && ((from_receiver && r_t_valid) || glint_case);
                if (!on_caster || vc) {
                    const bool take = !only || (on_caster && from_receiver && r_t_valid && depth - 1 == first_diffuse);
                    if (take) {
                        const V3 ev = emitter_eval_hit(S, e, hit);
                        const float epdf = bs.delta ? 0.f : pdf_emitter_hit(S, e, hit, si.p);
                        const float mis = mis_power(bs.pdf, epdf);
                        L = L + mul3(thr * mis, ev);
                    }
                }
            }
        }
        si_prev = si;
        si = hit;
    }
    st3(P.L + 3 * j, L);
    P.pos[2 * j] = pos_x; P.pos[2 * j + 1] = pos_y;
    P.nq[j] = nq;
    if (P.nm) P.nm[j] = nm;
}

// Discretized reconstruction filter of the film (rfilter.h:62-66, the table built on the host
// by init_discretization, src/libcore/rfilter.cpp:9-20); box: radius 0
struct FilmFilter { float values[32]; float radius, scale; int n; int box; };

SMS_DEV float filter_disc(const FilmFilter &F, float x) {
    const int index = (int) fabsf(x * F.scale);
    return F.values[index < 31 ? index : 31];
}

// the weight of a sample at film position (Px, Py) at pixel (x, y) of a w x h film with a
// border of 2 pixels (ImageBlock::put, imageblock.cpp:110-163); 0 when (x, y) is not a tap
SMS_DEV float film_weight(const FilmFilter &F, int w, int h, float Px, float Py, int x, int y) {
    const float bx = Px + 1.5f, by = Py + 1.5f;          // pos - (offset - border + .5)
    if (F.box) {
        const int lx = (int) ceilf(bx - .5f) - 2, ly = (int) ceilf(by - .5f) - 2;
        return (lx == x && ly == y) ? 1.f : 0.f;
    }
    const int lcx = (int) ceilf(bx - F.radius), lcy = (int) ceilf(by - F.radius);
    const int lox = lcx > 0 ? lcx : 0, loy = lcy > 0 ? lcy : 0;
    const int hfx = (int) floorf(bx + F.radius), hfy = (int) floorf(by + F.radius);
    const int hix = hfx < w + 3 ? hfx : w + 3, hiy = hfy < h + 3 ? hfy : h + 3;
    const int xr = x + 2 - lox, yr = y + 2 - loy;
    if (xr < 0 || yr < 0 || xr >= F.n || yr >= F.n || x + 2 > hix || y + 2 > hiy) return 0.f;
    const float wx = filter_disc(F, ((float) lox - bx) + (float) xr);
    const float wy = filter_disc(F, ((float) loy - by) + (float) yr);
    return wy * wx;
}

} // namespace smsdev
