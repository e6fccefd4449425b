// dev_lean.h -- LEAN-map density of a glint footprint (SpecularManifoldGlints::lean_pdf,
// src/librender/manifold_glints.cpp:556-579; include/mitsuba/render/lean.h) and the power
// heuristic that weighs the glint SMS strategy against BSDF sampling
// (manifold_glints.cpp:155-179).  Evaluated once per glint solution.
//
// exp, erf and log2 are FP64 series rounded to float, standing in for the correctly rounded
// float functions: exp by range reduction with a split ln 2 and a degree-13 Taylor
// polynomial, erf by its Maclaurin series below 2 and the Laplace continued fraction of erfc
// above, log2 by the exponent split and the atanh series of ln((1+s)/(1-s)).
#pragma once
#include "dev_scene.h"

namespace smsdev {

SMS_DEV double ldexp2(double p, int n) {
    int n1 = n / 2, n2 = n - n1;
    return (p * __longlong_as_double((long long) (n1 + 1023) << 52)) * __longlong_as_double((long long) (n2 + 1023) << 52);
}
SMS_DEV double exp_d(double x) {
    if (x != x) return x;
    if (x > 709.0) return __builtin_inf();
    if (x < -745.0) return 0.0;
    double k = floor(x * 1.4426950408889634 + 0.5);
    double r = (x - k * 6.93147180369123816490e-01) - k * 1.90821492927058770002e-10;
    double p = 1.0 / 6227020800.0;
    p = p * r + 1.0 / 479001600.0; p = p * r + 1.0 / 39916800.0; p = p * r + 1.0 / 3628800.0;
    p = p * r + 1.0 / 362880.0; p = p * r + 1.0 / 40320.0; p = p * r + 1.0 / 5040.0; p = p * r + 1.0 / 720.0;
    p = p * r + 1.0 / 120.0; p = p * r + 1.0 / 24.0; p = p * r + 1.0 / 6.0; p = p * r + 0.5; p = p * r + 1.0;
    p = p * r + 1.0;
    return ldexp2(p, (int) k);
}
SMS_DEV double erf_d(double x) {
    double ax = fabs(x), e;
    if (ax < 2.0) {
        double x2 = ax * ax, term = ax, s = ax;
        for (int n = 1; n <= 36; ++n) {
            term = term * (-x2 / (double) n);
            s = s + term / (double) (2 * n + 1);
        }
        e = s * 1.12837916709551257390;
    } else if (ax < 6.0) {
        double f = ax;
        for (int k = 40; k >= 1; --k) f = ax + (0.5 * (double) k) / f;
        e = 1.0 - exp_d(-ax * ax) * 0.56418958354775628695 / f;
    } else {
        e = 1.0;
    }
    return x < 0.0 ? -e : e;
}
SMS_DEV double log2_d(double x) {
    unsigned long long b = (unsigned long long) __double_as_longlong(x);
    int e = (int) ((b >> 52) & 0x7ff) - 1022;
    double m = __longlong_as_double((long long) ((b & 0x000fffffffffffffULL) | 0x3fe0000000000000ULL));
    if (m < 0.70710678118654752440) { m = m + m; e -= 1; }
    double s = (m - 1.0) / (m + 1.0), s2 = s * s, t = 0.0;
    for (int k = 39; k >= 1; k -= 2) t = t * s2 + 1.0 / (double) k;
    return (double) e + (2.0 * s * t) * 1.44269504088896340736;
}
SMS_DEV float expf_l(float x) { return (float) exp_d((double) x); }
SMS_DEV float erff_l(float x) { return (float) erf_d((double) x); }
SMS_DEV float log2f_l(float x) { return (float) log2_d((double) x); }
SMS_DEV float logf_l(float x) { return (float) (log2_d((double) x) * 0.69314718055994530942); }

constexpr float INV_SQRT_PI_F = 0.56418958354775628695f;
constexpr float INV_TWOPI_F = 0.15915494309189533577f;

struct LeanParams { V3 n; float ax, ay, rho, sqrt_one_minus_rho_sqr, tx_n, ty_n; };

// NormalmapBSDF::mipmap_level (bsdfs/normalmap.cpp:361-374); a negative ceil() becomes a huge
// size_t in the reference, which min() turns into the coarsest level
SMS_DEV int lean_level(float tiles, V2 duv_dx, V2 duv_dy, int levels) {
    V2 a = duv_dx * tiles, b = duv_dy * tiles;
    float width = 2.f * fmax_(fmax_(fabsf(a.x), fabsf(a.y)), fmax_(fabsf(b.x), fabsf(b.y)));
    float l = ceilf((float) (levels - 1) + log2f_l(fmax_(width, 1e-8f)));
    if (l < 0.f) return levels - 1;
    int li = (int) l;
    return li < levels - 1 ? li : levels - 1;
}

// eval_normal (use_slopes) and eval_lean_sigma at one level (normalmap.h:61-92,198-228)
SMS_DEV void lean_moments(const DNormalmap &nm, V2 uv, int level, V2 &mu, float sig[3]) {
    const int res = nm.res >> level;
    const float fres = (float) res, duv = 1.f / fres;
    V2 pos = v2((uv.x - 0.5f * duv) * fres, (uv.y - 0.5f * duv) * fres);
    pos.x = fmin_(fmax_(pos.x, 0.f), fres - 1.f); pos.y = fmin_(fmax_(pos.y, 0.f), fres - 1.f);
    uint32_t xmin = (uint32_t) floorf(pos.x), ymin = (uint32_t) floorf(pos.y);
    uint32_t xmax = xmin + 1, ymax = ymin + 1, lim = (uint32_t) (res - 1);
    xmin = xmin < lim ? xmin : lim; ymin = ymin < lim ? ymin : lim;
    xmax = xmax < lim ? xmax : lim; ymax = ymax < lim ? ymax : lim;
    V2 w1 = v2(pos.x - (float) xmin, pos.y - (float) ymin), w0 = v2(1.f - w1.x, 1.f - w1.y);
    const float *L = nm.lean + nm.lean_off[level];
    const float *v00 = L + 5 * ((size_t) ymin * res + xmin), *v10 = L + 5 * ((size_t) ymin * res + xmax),
                *v01 = L + 5 * ((size_t) ymax * res + xmin), *v11 = L + 5 * ((size_t) ymax * res + xmax);
    float v[5];
    for (int c = 0; c < 5; ++c) {
        float a = fmaf(w0.x, v00[c], w1.x * v10[c]), b = fmaf(w0.x, v01[c], w1.x * v11[c]);
        v[c] = fmaf(w0.y, a, w1.y * b);
    }
    V3 n = v3(-v[0], -v[1], 1.f);
    mu = v2(-n.x / n.z, -n.y / n.z);
    sig[0] = fmax_(0.f, v[2] - v[0] * v[0]);
    sig[1] = fmax_(0.f, v[3] - v[1] * v[1]);
    sig[2] = fmax_(0.f, v[4] - v[0] * v[1]);
}

// LEAN::p_from_lean_and_base (lean.h:44-77)
SMS_DEV LeanParams lean_params(V2 mu, const float sig[3], float alpha_u, float alpha_v) {
    float lean_e1 = mu.x, lean_e2 = mu.y;
    float lean_ax = sqrtf(2.f * sig[0]), lean_ay = sqrtf(2.f * sig[1]);
    float rho = 2.f * sig[2] / (lean_ax * lean_ay);
    float lean_e3 = 0.5f * lean_ax * lean_ax + lean_e1 * lean_e1;
    float lean_e4 = 0.5f * lean_ay * lean_ay + lean_e2 * lean_e2;
    float lean_e5 = 0.5f * rho * lean_ax * lean_ay + lean_e1 * lean_e2;
    float base_e3 = 0.5f * alpha_u * alpha_u, base_e4 = 0.5f * alpha_v * alpha_v;
    float e1 = lean_e1, e2 = lean_e2, e3 = lean_e3 + base_e3, e4 = lean_e4 + base_e4, e5 = lean_e5;
    LeanParams p;
    p.tx_n = e1; p.ty_n = e2;
    float tmp3 = fmax_(0.f, e3 - e1 * e1), tmp4 = fmax_(0.f, e4 - e2 * e2);
    p.ax = sqrtf(2.f * tmp3);
    p.ay = sqrtf(2.f * tmp4);
    p.rho = 2.f * (e5 - e1 * e2) / (p.ax * p.ay);
    p.sqrt_one_minus_rho_sqr = sqrtf(1.f - p.rho * p.rho);
    p.n = normalize(v3(-p.tx_n, -p.ty_n, 1.f));
    return p;
}

// LEAN::sigma_std, sigma, g1, p22, ndf, vndf (lean.h:79-212)
SMS_DEV float lean_sigma_std(float cos_theta) {
    if (cos_theta == 1.f) return 1.f;
    float sin_theta = sqrtf(1.f - cos_theta * cos_theta);
    float nu = cos_theta / sin_theta;
    float tmp = expf_l(-nu * nu) * INV_SQRT_PI_F;
    return 0.5f * (cos_theta * (1.f + erff_l(nu)) + sin_theta * tmp);
}
SMS_DEV float lean_sigma(V3 k, const LeanParams &p) {
    float a = k.x * p.ax + k.y * p.ay * p.rho;
    float b = k.y * p.ay * p.sqrt_one_minus_rho_sqr;
    float c = k.z - k.x * p.tx_n - k.y * p.ty_n;
    float nrm = sqrtf(a * a + b * b + c * c);
    V3 kp = v3(a, b, c) * (1.f / nrm);
    return nrm * lean_sigma_std(kp.z);
}
SMS_DEV float lean_g1(V3 k, const LeanParams &p) {
    if (dot(k, p.n) > 0.f) return k.z / lean_sigma(k, p);
    return 0.f;
}
SMS_DEV float lean_p22(float sx, float sy, const LeanParams &p) {
    sx -= p.tx_n;
    sy -= p.ty_n;
    float nrm = p.ax * p.ay * p.sqrt_one_minus_rho_sqr;
    float x = sx / p.ax;
    float tmp1 = p.ax * sy - p.rho * p.ay * sx;
    float tmp2 = p.ax * p.ay * p.sqrt_one_minus_rho_sqr;
    float y = tmp1 / tmp2;
    float r = x * x + y * y;
    return (INV_PI_F * expf_l(-r)) / nrm;
}
SMS_DEV float lean_ndf(V3 h, const LeanParams &p) {
    if (h.z > 5.9604644775390625e-08f) {
        float tmp1 = h.z * h.z, tmp2 = tmp1 * tmp1;
        return lean_p22(-h.x / h.z, -h.y / h.z, p) / tmp2;
    }
    return 0.f;
}
SMS_DEV float lean_vndf(V3 h, V3 wi, const LeanParams &p) {
    float dp = dot(wi, h);
    if (dp > 0.f) return dp * lean_ndf(h, p) / lean_sigma(wi, p);
    return 0.f;
}

// lean_pdf of a footprint (uv, duv_dx, duv_dy on `shape`) for wi, wo in its shading frame, with
// the smooth conductor's roughness 0 (conductor.cpp:288-290)
SMS_DEV float lean_pdf(const DScene &S, int shape, V2 uv_q, V2 duv_dx, V2 duv_dy, V3 wi, V3 wo) {
    if (wi.z < 0.f || wo.z < 0.f) return 0.f;
    const DShape &sh = S.shapes[shape];
    const DNormalmap nm = S.nmaps[sh.normalmap];
    int level = lean_level(sh.tiles, duv_dx, duv_dy, nm.levels);
    V2 uv = uv_q * sh.tiles;
    uv = v2(uv.x - floorf(uv.x), uv.y - floorf(uv.y));
    V2 mu;
    float sig[3];
    lean_moments(nm, uv, level, mu, sig);
    LeanParams p = lean_params(mu, sig, 0.f, 0.f);
    V3 H = normalize(wo + wi);
    float G = lean_g1(wo, p) * lean_g1(wi, p);
    if (!(G > 0.f)) return 0.f;
    return lean_vndf(H, wi, p) / (4.f * dot(wi, H));
}

// power heuristic (manifold_glints.h:155-159)
SMS_DEV float mis_power(float a, float b) {
    a *= a;
    b *= b;
    return a > 0.f ? a / (a + b) : 0.f;
}

// Scene::pdf_emitter_direction of the SMS strategy's emitter sample (scene.cpp:206-219,
// area.cpp:119-125, sphere.cpp:254-267); 1 for delta emitters (manifold_glints.cpp:158-165)
SMS_DEV float sms_emitter_pdf(const DScene &S, const EI &ei, V3 p) {
    if (ei.kind != SMS_EI_AREA) return 1.f;
    const DEmitter &em = S.emitters[ei.emitter];
    const DShape &sh = S.shapes[em.shape];
    V3 c = v3(sh.cx, sh.cy, sh.cz);
    float sin_alpha = sh.radius * (1.f / norm(c - p));
    float cos_alpha = safe_sqrt(1.f - sin_alpha * sin_alpha);
    float pdf;
    if (sin_alpha < 0.99999994039535522461f) pdf = INV_TWOPI_F / (1.f - cos_alpha);
    else pdf = sh.inv_surface_area * (norm(ei.p - p) * norm(ei.p - p)) / fabsf(dot(ei.d, ei.n));
    if (!(dot(ei.d, ei.n) < 0.f)) pdf = 0.f;
    if (S.n_emitters > 1) pdf = pdf * (1.f / (float) S.n_emitters);
    return pdf;
}

} // namespace smsdev
