// dev_math.h -- device-side FP32 vector math, Cephes-style transcendentals and PCG32 for the
// gfx950 SMS kernels.
//
// Arithmetic contract (identical op order on the CPU oracle so that a walk's trajectory is
// bit-reproducible from its inputs; the library is compiled with -ffp-contract=off):
//   dot(a,b) = (a.x*b.x + a.y*b.y) + a.z*b.z ; norm = sqrtf(dot) ; normalize = a * (1/sqrtf(dot))
//   rcp(x) = 1/x (IEEE division); fmaf only where the reference writes fmadd/fnmadd/fmsub.
// Transcendentals are S. L. Moshier's Cephes single-precision routines (sinf/cosf, atanf,
// asinf/acosf) re-implemented from the published method; IEEE sqrt and division are
// correctly rounded on gfx950 with the default HIP flags.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SMS_DEV __device__ __forceinline__

namespace smsdev {

constexpr float PI_F = 3.14159265358979323846f;
constexpr float PIO2_F = 1.57079632679489661923f;
constexpr float PIO4_F = 0.78539816339744830962f;
constexpr float INV_PI_F = 0.31830988618379067154f;
constexpr float INV_TWO_PI_F = 0.15915494309189533577f;
constexpr float EPS_F = 5.9604644775390625e-08f;            // math::Epsilon<float>
constexpr float RAY_EPS_F = 1500.f * 5.9604644775390625e-08f;
constexpr float SHADOW_EPS_F = 10.f * (1500.f * 5.9604644775390625e-08f);

struct V3 { float x, y, z; };
struct V2 { float x, y; };

SMS_DEV V3 v3(float x, float y, float z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
SMS_DEV V2 v2(float x, float y) { V2 r; r.x = x; r.y = y; return r; }
SMS_DEV V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
SMS_DEV V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
SMS_DEV V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
SMS_DEV V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
SMS_DEV V3 operator*(float s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
SMS_DEV V3 operator/(V3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
SMS_DEV V2 operator+(V2 a, V2 b) { return v2(a.x + b.x, a.y + b.y); }
SMS_DEV V2 operator-(V2 a, V2 b) { return v2(a.x - b.x, a.y - b.y); }
SMS_DEV V2 operator*(V2 a, float s) { return v2(a.x * s, a.y * s); }
SMS_DEV V2 operator*(float s, V2 a) { return v2(s * a.x, s * a.y); }
SMS_DEV float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
SMS_DEV float dot2(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }
SMS_DEV float sqnorm(V3 a) { return dot(a, a); }
SMS_DEV float norm(V3 a) { return sqrtf(dot(a, a)); }
SMS_DEV float norm2(V2 a) { return sqrtf(dot2(a, a)); }
SMS_DEV V3 normalize(V3 a) { return a * (1.f / sqrtf(dot(a, a))); }
SMS_DEV V3 cross(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
SMS_DEV V3 fmadd3(V3 a, float b, V3 c) { return v3(fmaf(a.x, b, c.x), fmaf(a.y, b, c.y), fmaf(a.z, b, c.z)); }
SMS_DEV V3 fnmadd3(V3 a, float b, V3 c) { return v3(fmaf(-a.x, b, c.x), fmaf(-a.y, b, c.y), fmaf(-a.z, b, c.z)); }
SMS_DEV float fmin_(float a, float b) { return b < a ? b : a; }   // std::min semantics
SMS_DEV float fmax_(float a, float b) { return a < b ? b : a; }   // std::max semantics
SMS_DEV float hmax_abs(V3 a) {
    float m = fabsf(a.x);
    if (fabsf(a.y) > m) m = fabsf(a.y);
    if (fabsf(a.z) > m) m = fabsf(a.z);
    return m;
}
SMS_DEV float safe_sqrt(float x) { return sqrtf(x > 0.f ? x : 0.f); }
SMS_DEV bool sign_bit(float x) { return (__float_as_uint(x) >> 31) != 0u; }
SMS_DEV float mulsign(float a, float b) { return sign_bit(b) ? -a : a; }
SMS_DEV V3 ld3(const float *p) { return v3(p[0], p[1], p[2]); }
SMS_DEV void st3(float *p, V3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }

struct M2 { float m00, m01, m10, m11; };
SMS_DEV M2 m2(float a, float b, float c, float d) { M2 r; r.m00 = a; r.m01 = b; r.m10 = c; r.m11 = d; return r; }
SMS_DEV float det2(M2 a) { return a.m00 * a.m11 - a.m01 * a.m10; }
SMS_DEV M2 inverse2(M2 a) {
    float inv_det = 1.f / det2(a);
    return m2(a.m11 * inv_det, -a.m01 * inv_det, -a.m10 * inv_det, a.m00 * inv_det);
}
SMS_DEV V2 mulv2(M2 a, V2 v) { return v2(a.m00 * v.x + a.m01 * v.y, a.m10 * v.x + a.m11 * v.y); }
SMS_DEV M2 mulm2(M2 a, M2 b) {
    return m2(a.m00 * b.m00 + a.m01 * b.m10, a.m00 * b.m01 + a.m01 * b.m11,
              a.m10 * b.m00 + a.m11 * b.m10, a.m10 * b.m01 + a.m11 * b.m11);
}
SMS_DEV M2 subm2(M2 a, M2 b) { return m2(a.m00 - b.m00, a.m01 - b.m01, a.m10 - b.m10, a.m11 - b.m11); }
SMS_DEV M2 negm2(M2 a) { return m2(-a.m00, -a.m01, -a.m10, -a.m11); }

// ---------------------------------------------------------------- Cephes-style float math
SMS_DEV void c_sincosf(float xx, float *s_out, float *c_out) {
    const float FOPI = 1.27323954473516f;
    const float DP1 = 0.78515625f, DP2 = 2.4187564849853515625e-4f, DP3 = 3.77489497744594108e-8f;
    float x = fabsf(xx);
    int sign_s = xx < 0.f ? -1 : 1, sign_c = 1;
    int j = (int) (FOPI * x);
    float y = (float) j;
    if (j & 1) { j += 1; y += 1.0f; }
    j &= 7;
    if (j > 3) { sign_s = -sign_s; sign_c = -sign_c; j -= 4; }
    if (j > 1) sign_c = -sign_c;
    x = ((x - y * DP1) - y * DP2) - y * DP3;
    float z = x * x;
    float ps = ((-1.9515295891E-4f * z + 8.3321608736E-3f) * z - 1.6666654611E-1f) * z * x + x;
    float pc = ((2.443315711809948E-005f * z - 1.388731625493765E-003f) * z + 4.166664568298827E-002f) * z * z
               - 0.5f * z + 1.0f;
    float s, c;
    if (j == 1 || j == 2) { s = pc; c = ps; } else { s = ps; c = pc; }
    *s_out = sign_s < 0 ? -s : s;
    *c_out = sign_c < 0 ? -c : c;
}
SMS_DEV float c_atanf(float xx) {
    float x = fabsf(xx), y;
    if (x > 2.414213562373095f) { y = PIO2_F; x = -(1.0f / x); }
    else if (x > 0.4142135623730950f) { y = PIO4_F; x = (x - 1.0f) / (x + 1.0f); }
    else y = 0.0f;
    float z = x * x;
    y += (((8.05374449538e-2f * z - 1.38776856032E-1f) * z + 1.99777106478E-1f) * z - 3.33329491539E-1f) * z * x + x;
    return xx < 0.f ? -y : y;
}
SMS_DEV float c_atan2f(float y, float x) {
    if (x == 0.f) {
        if (y > 0.f) return PIO2_F;
        if (y < 0.f) return -PIO2_F;
        return 0.f;
    }
    float z = c_atanf(y / x);
    if (x < 0.f) {
        if (y < 0.f || (y == 0.f && sign_bit(y))) z = z - PI_F;
        else z = z + PI_F;
    }
    return z;
}
SMS_DEV float c_asinf(float xx) {
    float a = fabsf(xx), x, z;
    if (a > 1.0f) return 0.f;
    if (a < 1.0e-4f) return xx;
    bool flag;
    if (a > 0.5f) { z = 0.5f * (1.0f - a); x = sqrtf(z); flag = true; }
    else { x = a; z = x * x; flag = false; }
    z = ((((4.2163199048E-2f * z + 2.4181311049E-2f) * z + 4.5470025998E-2f) * z + 7.4953002686E-2f) * z
         + 1.6666752422E-1f) * z * x + x;
    if (flag) { z = z + z; z = PIO2_F - z; }
    return xx < 0.f ? -z : z;
}
SMS_DEV float c_acosf(float x) {
    if (x < -1.0f) return PI_F;
    if (x > 1.0f) return 0.f;
    if (x < -0.5f) return PI_F - 2.0f * c_asinf(sqrtf(0.5f * (1.0f + x)));
    if (x > 0.5f) return 2.0f * c_asinf(sqrtf(0.5f * (1.0f - x)));
    return PIO2_F - c_asinf(x);
}
SMS_DEV float safe_acos(float x) { return c_acosf(x < -1.f ? -1.f : (x > 1.f ? 1.f : x)); }

// ---------------------------------------------------------------- PCG32 (O'Neill, pcg-random.org)
struct Pcg32 { uint64_t state, inc; };
constexpr uint64_t PCG32_MULT = 0x5851f42d4c957f2dULL;
SMS_DEV uint32_t pcg32_next(Pcg32 &r) {
    uint64_t old = r.state;
    r.state = old * PCG32_MULT + r.inc;
    uint32_t xorshifted = (uint32_t) (((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t) (old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((~rot + 1u) & 31u));
}
SMS_DEV void pcg32_seed(Pcg32 &r, uint64_t initstate, uint64_t initseq) {
    r.state = 0u;
    r.inc = (initseq << 1u) | 1u;
    pcg32_next(r);
    r.state += initstate;
    pcg32_next(r);
}
SMS_DEV float pcg32_float(Pcg32 &r) { return __uint_as_float((pcg32_next(r) >> 9) | 0x3f800000u) - 1.0f; }
SMS_DEV uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
// Per-(seed, query, walk) stream: walk 0 = query-level samples, walk k >= 1 = k-th walk.
SMS_DEV uint64_t query_key(uint64_t seed, uint64_t query) { return mix64(seed ^ mix64(query + 0x9E3779B97F4A7C15ULL)); }
SMS_DEV void stream_seed(Pcg32 &r, uint64_t qkey, uint64_t walk) { pcg32_seed(r, qkey, walk); }

} // namespace smsdev
