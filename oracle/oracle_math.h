/*
 * oracle_math.h -- TEST INFRASTRUCTURE ONLY.  Scalar FP32 helpers of the CPU oracle.
 *
 * This header is part of the CPU restatement of the reference's SMS solver path that the
 * test-suite and the benchmark's CPU-baseline leg use as a checker.  Nothing in the product
 * includes it.
 *
 * Arithmetic contract (shared by construction with the GPU kernels, written independently):
 *   dot(a,b)       = (a.x*b.x + a.y*b.y) + a.z*b.z
 *   norm(a)        = sqrtf(dot(a,a));   normalize(a) = a * (1/sqrtf(dot(a,a)))
 *   rcp(x)         = 1/x
 *   cross(a,b)     = (a.y*b.z - a.z*b.y, a.z*b.x - a.x*b.z, a.x*b.y - a.y*b.x)
 *   fmadd/fnmadd/fmsub map to fmaf wherever the reference writes them; every other
 *   expression is evaluated as written, without contraction (-ffp-contract=off).
 * Transcendentals follow S. L. Moshier's Cephes single-precision routines (sinf/cosf,
 * atanf, asinf/acosf), re-implemented here from the published method; exp, erf and log2 of
 * the LEAN pdf are FP64 series rounded to float (see below).
 */
#pragma once
#include <math.h>
#include <stdint.h>
#include <string.h>

namespace orc {

static const float PI_F      = 3.14159265358979323846f;
static const float PIO2_F    = 1.57079632679489661923f;
static const float PIO4_F    = 0.78539816339744830962f;
static const float INV_PI_F  = 0.31830988618379067154f;
static const float INV_TWOPI_F = 0.15915494309189533577f;
static const float INF_F     = INFINITY;
/* math::Epsilon<float> = FLT_EPSILON/2; RayEpsilon = 1500*Epsilon; ShadowEpsilon = 10*RayEpsilon
   (include/mitsuba/core/math.h:31-38) */
static const float EPS_F        = 5.9604644775390625e-08f;
static const float RAY_EPS_F    = 1500.f * 5.9604644775390625e-08f;
static const float SHADOW_EPS_F = 10.f * (1500.f * 5.9604644775390625e-08f);

struct V3 { float x, y, z; };
struct V2 { float x, y; };

static inline V3 v3(float x, float y, float z) { V3 r = {x, y, z}; return r; }
static inline V2 v2(float x, float y) { V2 r = {x, y}; return r; }
static inline V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
static inline V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
static inline V3 operator*(float s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
static inline V3 operator/(V3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
static inline V2 operator+(V2 a, V2 b) { return v2(a.x + b.x, a.y + b.y); }
static inline V2 operator-(V2 a, V2 b) { return v2(a.x - b.x, a.y - b.y); }
static inline V2 operator*(V2 a, float s) { return v2(a.x * s, a.y * s); }
static inline V2 operator*(float s, V2 a) { return v2(s * a.x, s * a.y); }
static inline float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline float dot2(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }
static inline float sqnorm(V3 a) { return dot(a, a); }
static inline float norm(V3 a) { return sqrtf(dot(a, a)); }
static inline float norm2(V2 a) { return sqrtf(dot2(a, a)); }
static inline float rcp(float x) { return 1.f / x; }
static inline V3 normalize(V3 a) { return a * (1.f / sqrtf(dot(a, a))); }
static inline V3 cross(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline V3 fmadd3(V3 a, float b, V3 c) { return v3(fmaf(a.x, b, c.x), fmaf(a.y, b, c.y), fmaf(a.z, b, c.z)); }
/* fnmadd(a, b, c) = -a*b + c */
static inline V3 fnmadd3(V3 a, float b, V3 c) { return v3(fmaf(-a.x, b, c.x), fmaf(-a.y, b, c.y), fmaf(-a.z, b, c.z)); }
static inline float hmax_abs(V3 a) { float m = fabsf(a.x); if (fabsf(a.y) > m) m = fabsf(a.y); if (fabsf(a.z) > m) m = fabsf(a.z); return m; }
static inline float safe_sqrt(float x) { return sqrtf(x > 0.f ? x : (float) 0.f); }
static inline float mulsign(float a, float b) { return signbit(b) ? -a : a; }

/* 2x2 matrix, m[row][col] */
struct M2 { float m00, m01, m10, m11; };
static inline M2 m2(float a, float b, float c, float d) { M2 r = {a, b, c, d}; return r; }
static inline float det2(M2 a) { return a.m00 * a.m11 - a.m01 * a.m10; }
static inline M2 inverse2(M2 a) {
    float inv_det = 1.f / det2(a);
    return m2(a.m11 * inv_det, -a.m01 * inv_det, -a.m10 * inv_det, a.m00 * inv_det);
}
static inline V2 mulv2(M2 a, V2 v) { return v2(a.m00 * v.x + a.m01 * v.y, a.m10 * v.x + a.m11 * v.y); }
static inline M2 mulm2(M2 a, M2 b) {
    return m2(a.m00 * b.m00 + a.m01 * b.m10, a.m00 * b.m01 + a.m01 * b.m11,
              a.m10 * b.m00 + a.m11 * b.m10, a.m10 * b.m01 + a.m11 * b.m11);
}
static inline M2 subm2(M2 a, M2 b) { return m2(a.m00 - b.m00, a.m01 - b.m01, a.m10 - b.m10, a.m11 - b.m11); }
static inline M2 negm2(M2 a) { return m2(-a.m00, -a.m01, -a.m10, -a.m11); }

/* ---------------------------------------------------------------- Cephes-style float math */
static inline float c_sinf_poly(float x, float z) {
    return ((-1.9515295891E-4f * z + 8.3321608736E-3f) * z - 1.6666654611E-1f) * z * x + x;
}
static inline float c_cosf_poly(float z) {
    return ((2.443315711809948E-005f * z - 1.388731625493765E-003f) * z + 4.166664568298827E-002f) * z * z
           - 0.5f * z + 1.0f;
}
/* sin and cos of one argument (Cephes sinf.c / cosf.c range reduction by pi/4). */
static inline void c_sincosf(float xx, float *s_out, float *c_out) {
    const float FOPI = 1.27323954473516f;
    const float DP1 = 0.78515625f, DP2 = 2.4187564849853515625e-4f, DP3 = 3.77489497744594108e-8f;
    float x = fabsf(xx);
    int sign_s = xx < 0.f ? -1 : 1;
    int sign_c = 1;
    int j = (int) (FOPI * x);
    float y = (float) j;
    if (j & 1) { j += 1; y += 1.0f; }
    j &= 7;
    if (j > 3) { sign_s = -sign_s; sign_c = -sign_c; j -= 4; }
    if (j > 1) sign_c = -sign_c;
    x = ((x - y * DP1) - y * DP2) - y * DP3;
    float z = x * x;
    float ps = c_sinf_poly(x, z), pc = c_cosf_poly(z);
    float s, c;
    if (j == 1 || j == 2) { s = pc; c = ps; } else { s = ps; c = pc; }
    *s_out = sign_s < 0 ? -s : s;
    *c_out = sign_c < 0 ? -c : c;
}
static inline float c_atanf(float xx) {
    float x = fabsf(xx), y;
    if (x > 2.414213562373095f) { y = PIO2_F; x = -(1.0f / x); }
    else if (x > 0.4142135623730950f) { y = PIO4_F; x = (x - 1.0f) / (x + 1.0f); }
    else y = 0.0f;
    float z = x * x;
    y += (((8.05374449538e-2f * z - 1.38776856032E-1f) * z + 1.99777106478E-1f) * z - 3.33329491539E-1f) * z * x + x;
    return xx < 0.f ? -y : y;
}
static inline float c_atan2f(float y, float x) {
    if (x == 0.f) {
        if (y > 0.f) return PIO2_F;
        if (y < 0.f) return -PIO2_F;
        return 0.f;
    }
    float z = c_atanf(y / x);
    if (x < 0.f) {
        if (y < 0.f || (y == 0.f && signbit(y))) z = z - PI_F;
        else z = z + PI_F;
    }
    return z;
}
static inline float c_asinf(float xx) {
    float a = fabsf(xx), x, z;
    if (a > 1.0f) return 0.f;
    if (a < 1.0e-4f) return xx;
    int flag;
    if (a > 0.5f) { z = 0.5f * (1.0f - a); x = sqrtf(z); flag = 1; }
    else { x = a; z = x * x; flag = 0; }
    z = ((((4.2163199048E-2f * z + 2.4181311049E-2f) * z + 4.5470025998E-2f) * z + 7.4953002686E-2f) * z
         + 1.6666752422E-1f) * z * x + x;
    if (flag) { z = z + z; z = PIO2_F - z; }
    return xx < 0.f ? -z : z;
}
static inline float c_acosf(float x) {
    if (x < -1.0f) return PI_F;
    if (x > 1.0f) return 0.f;
    if (x < -0.5f) return PI_F - 2.0f * c_asinf(sqrtf(0.5f * (1.0f + x)));
    if (x > 0.5f) return 2.0f * c_asinf(sqrtf(0.5f * (1.0f - x)));
    return PIO2_F - c_asinf(x);
}
static inline float safe_acos(float x) { return c_acosf(x < -1.f ? (float) -1.f : (x > 1.f ? (float) 1.f : x)); }

/* ---------------------------------------------------------------- PCG32 (O'Neill, pcg-random.org) */
struct Pcg32 { uint64_t state, inc; };
static const uint64_t PCG32_MULT = 0x5851f42d4c957f2dULL;
static inline uint32_t pcg32_next(Pcg32 *r) {
    uint64_t old = r->state;
    r->state = old * PCG32_MULT + r->inc;
    uint32_t xorshifted = (uint32_t) (((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t) (old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((~rot + 1u) & 31));
}
static inline void pcg32_seed(Pcg32 *r, uint64_t initstate, uint64_t initseq) {
    r->state = 0u;
    r->inc = (initseq << 1u) | 1u;
    pcg32_next(r);
    r->state += initstate;
    pcg32_next(r);
}
static inline float pcg32_float(Pcg32 *r) {
    union { uint32_t u; float f; } x;
    x.u = (pcg32_next(r) >> 9) | 0x3f800000u;
    return x.f - 1.0f;
}
static inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
/* Stream for (seed, query, walk): walk 0 holds the query-level samples, walk k>=1 the k-th walk. */
static inline void stream_seed(Pcg32 *r, uint64_t seed, uint64_t query, uint64_t walk) {
    pcg32_seed(r, mix64(seed ^ mix64(query + 0x9E3779B97F4A7C15ULL)), walk);
}

/* ---------------------------------------------------------------- exp / erf / log2
   For the LEAN pdf (the glint MIS weight).  Evaluated in FP64 with fixed-length series and
   rounded to float, so they stand in for the correctly rounded float functions; the GPU code
   runs the same operations.  exp: range reduction by ln 2 (split constant) and a degree-13
   Taylor polynomial; erf: Maclaurin series below 2, Laplace continued fraction of erfc
   above; log2: exponent split and the atanh series of ln((1+s)/(1-s)). */
static inline double d_from_bits(uint64_t b) { double d; memcpy(&d, &b, 8); return d; }
static inline uint64_t d_bits(double d) { uint64_t b; memcpy(&b, &d, 8); return b; }
static inline double ldexp2(double p, int n) {
    int n1 = n / 2, n2 = n - n1;
    return (p * d_from_bits((uint64_t) (n1 + 1023) << 52)) * d_from_bits((uint64_t) (n2 + 1023) << 52);
}
static inline double exp_d(double x) {
    if (x != x) return x;
    if (x > 709.0) return INFINITY;
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
static inline double erf_d(double x) {
    double ax = fabs(x), e;
    if (ax < 2.0) {
        double x2 = ax * ax, term = ax, s = ax;
        for (int n = 1; n <= 36; ++n) {
            term = term * (-x2 / (double) n);
            s = s + term / (double) (2 * n + 1);
        }
        e = s * 1.12837916709551257390;          /* 2 / sqrt(pi) */
    } else if (ax < 6.0) {
        double f = ax;                            /* erfc = exp(-x^2)/sqrt(pi) / (x + (1/2)/(x + 1/(x + (3/2)/(x + ...)))) */
        for (int k = 40; k >= 1; --k) f = ax + (0.5 * (double) k) / f;
        e = 1.0 - exp_d(-ax * ax) * 0.56418958354775628695 / f;
    } else {
        e = 1.0;
    }
    return x < 0.0 ? -e : e;
}
static inline double log2_d(double x) {          /* x positive, normal */
    uint64_t b = d_bits(x);
    int e = (int) ((b >> 52) & 0x7ff) - 1022;
    double m = d_from_bits((b & 0x000fffffffffffffULL) | 0x3fe0000000000000ULL);   /* [0.5, 1) */
    if (m < 0.70710678118654752440) { m = m + m; e -= 1; }
    double s = (m - 1.0) / (m + 1.0), s2 = s * s, t = 0.0;
    for (int k = 39; k >= 1; k -= 2) t = t * s2 + 1.0 / (double) k;
    return (double) e + (2.0 * s * t) * 1.44269504088896340736;
}
static inline float expf_l(float x) { return (float) exp_d((double) x); }
static inline float erff_l(float x) { return (float) erf_d((double) x); }
static inline float log2f_l(float x) { return (float) log2_d((double) x); }

} // namespace orc
