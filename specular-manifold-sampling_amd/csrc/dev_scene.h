// dev_scene.h -- device scene for the gfx950 SMS kernels: analytic shapes, a triangle BVH,
// normal maps, emitters and the surface-interaction / ManifoldVertex construction every
// Newton step of a walk performs.
//
// Layout in HBM (built once by sms_ctx_create):
//   shapes     DShape[n]           ~200 B each, read through the scalar cache (uniform index)
//   bvh        DNode[n_nodes]      32 B: (bmin.xyz, left|first) (bmax.xyz, count); children of
//                                  an inner node are adjacent (left, left+1)
//   tris       float4[3*n_tris]    leaf order: (p0, shape) (e1 = p1-p0, prim) (e2 = p2-p0, -)
//   vnrm/vuv   per vertex          only read when a hit is turned into a surface interaction
//   slopes     float2[res*res]     decoded level-0 normal-map slopes
#pragma once
#include "dev_math.h"
#include "../../include/sms.h"

namespace smsdev {

struct DShape {
    int type, bsdf, flags, emitter;
    float cx, cy, cz, radius, inv_radius;
    float M[12], Minv[12];
    float frs[3], frt[3], frn[3];
    float eta_r, k, eta_vtx;
    float refl[3];
    int normalmap; float tiles;
    int fbase, nf, has_n, has_uv;
    float area_sum, area_norm, inv_surface_area;
    float bbox_center[3];
    int cdf_offset;
    float alpha;                  // rough conductor: isotropic Beckmann roughness (0: smooth)
};

struct DEmitter { int type, shape; float radiance[3], position[3], intensity[3]; int ss, ms; float direction[3], irradiance[3]; };

// Level-0 slopes stored with one replicated pad column and row ((res+1)^2 float2, row stride
// res+1): texels (x, y), (x+1, y) of a bilinear footprint are adjacent for every x <= res-1,
// and the pad reproduces the edge clamp of Normalmap::eval_normal (normalmap.h:67-72), so a
// fetch is two 16-byte loads.
// LEAN mip pyramid: level l holds (res >> l)^2 texels of (E[x], E[y], E[x^2], E[y^2], E[xy]),
// unpadded, starting at float offset lean_off[l] of `lean`.
constexpr int SMS_MAX_LEAN_LEVELS = 16;
struct DNormalmap {
    int res; int stride; const float2 *slope;
    const float *lean; int levels; long long lean_off[SMS_MAX_LEAN_LEVELS];
    // encoded-normal maps: the stored RGB in the padded layout of `slope` (w unused) and the RGB
    // of the coarsest level of the box-filtered RGB pyramid
    int encoded; const float4 *rgbn; float top_rgb[3];
};

struct DCamera {
    float o[3], left[3], up[3], dir[3];
    float tan_half, rect_y, rect_x, normalization;
    float dx[3], dy[3];
    int w, h;
};

struct DNode { float4 a, b; };

struct DScene {
    const DShape *shapes; int n_shapes; int n_analytic;
    const int *analytic;                 // indices of sphere/rectangle shapes
    const DEmitter *emitters; int n_emitters;
    const int *emitters_ss; int n_emitters_ss;
    const int *emitters_all; int n_emitters_all;
    const int *casters_ss; int n_casters_ss;
    const int *emitters_ms; int n_emitters_ms;
    const int *casters_ms; int n_casters_ms;
    const DNormalmap *nmaps;
    const float4 *vpos;                  // mesh vertex positions (xyz, 0)
    const float4 *vnrm;                  // mesh vertex normals
    const float2 *vuv;                   // mesh vertex uvs
    const uint4 *faces;                  // global triangle: (i0, i1, i2, shape) (global vertex ids)
    const DNode *nodes; int n_nodes;
    const float4 *tris;                  // leaf-ordered triangle records (3 float4 each)
    const float *cdf, *pmf;              // mesh area distributions
    DCamera cam; int has_cam;
    float box_c[3], box_inv;             // scene bounding box centre, 1 / largest half extent
    float bs_dist;                       // directional emitters: 2 x the scene's bounding-sphere radius
    sms_config cfg;
};

struct Ray { V3 o, d, d_rcp; float mint, maxt; };
SMS_DEV Ray make_ray(V3 o, V3 d, float mint, float maxt) {
    Ray r; r.o = o; r.d = d; r.d_rcp = v3(1.f / d.x, 1.f / d.y, 1.f / d.z); r.mint = mint; r.maxt = maxt;
    return r;
}
// Interaction::spawn_ray (include/mitsuba/render/interaction.h:58-61)
SMS_DEV Ray spawn_ray(V3 p, V3 d) { return make_ray(p, d, (1.f + hmax_abs(p)) * RAY_EPS_F, __builtin_inff()); }
SMS_DEV Ray make_ray(V3 o, V3 d) { return make_ray(o, d, RAY_EPS_F, __builtin_inff()); }

SMS_DEV V3 xform_p(const float *M, V3 p) {
    return v3(((M[0] * p.x + M[1] * p.y) + M[2] * p.z) + M[3],
              ((M[4] * p.x + M[5] * p.y) + M[6] * p.z) + M[7],
              ((M[8] * p.x + M[9] * p.y) + M[10] * p.z) + M[11]);
}
SMS_DEV V3 xform_v(const float *M, V3 v) {
    return v3((M[0] * v.x + M[1] * v.y) + M[2] * v.z,
              (M[4] * v.x + M[5] * v.y) + M[6] * v.z,
              (M[8] * v.x + M[9] * v.y) + M[10] * v.z);
}

// coordinate_system (include/mitsuba/core/vector.h:116-136)
SMS_DEV void coordinate_system(V3 n, V3 &s, V3 &t) {
    float sign = sign_bit(n.z) ? -1.f : 1.f;
    float a = -(1.f / (sign + n.z));
    float b = n.x * n.y * a;
    s = v3(mulsign(n.x * n.x * a, n.z) + 1.f, mulsign(b, n.z), -mulsign(n.x, n.z));
    t = v3(b, sign + n.y * n.y * a, -n.y);
}

struct SI {
    bool valid; float t;
    V3 p, n, sh_n, sh_s, sh_t, dp_du, dp_dv;
    V3 wi;                        // direction toward the ray origin in the shading frame (si.wi)
    V2 uv;
    int shape, prim;
};
SMS_DEV SI si_invalid() {
    SI s;
    s.valid = false; s.t = __builtin_inff(); s.shape = -1; s.prim = -1;
    s.p = s.n = s.sh_n = s.sh_s = s.sh_t = s.dp_du = s.dp_dv = s.wi = v3(0.f, 0.f, 0.f);
    s.uv = v2(0.f, 0.f);
    return s;
}

// ---------------------------------------------------------------- primitive tests
// Sphere::ray_intersect in FP64 (src/shapes/sphere.cpp:276-303) + math::solve_quadratic
SMS_DEV bool sphere_intersect(const DShape &sh, const Ray &ray, float &t_out) {
    double mint = (double) ray.mint, maxt = (double) ray.maxt;
    double ox = (double) ray.o.x - (double) sh.cx, oy = (double) ray.o.y - (double) sh.cy,
           oz = (double) ray.o.z - (double) sh.cz;
    double dx = ray.d.x, dy = ray.d.y, dz = ray.d.z;
    double A = (dx * dx + dy * dy) + dz * dz;
    double B = 2.0 * ((ox * dx + oy * dy) + oz * dz);
    double C = ((ox * ox + oy * oy) + oz * oz) - (double) sh.radius * (double) sh.radius;
    bool linear = A == 0.0;
    bool active = !linear || B > 0.0;
    double x0 = -C / B, x1 = x0;
    double discrim = fma(B, B, -(4.0 * A * C));
    active = active && (linear || discrim >= 0.0);
    if (!active) return false;
    if (!linear) {
        double sq = sqrt(discrim);
        double temp = -0.5 * (B + copysign(sq, B));
        double x0p = temp / A, x1p = C / temp;
        x0 = x1p < x0p ? x1p : x0p;
        x1 = x0p < x1p ? x1p : x0p;
    }
    bool out_bounds = !(x0 <= maxt && x1 >= mint);
    bool in_bounds = x0 < mint && x1 > maxt;
    if (out_bounds || in_bounds) return false;
    t_out = (float) (x0 < mint ? x1 : x0);
    return true;
}

// Rectangle::ray_intersect (src/shapes/rectangle.cpp:141-163)
SMS_DEV bool rect_intersect(const DShape &sh, const Ray &ray, float &t_out, float &lx, float &ly) {
    V3 o = xform_p(sh.Minv, ray.o), d = xform_v(sh.Minv, ray.d);
    float t = -o.z * (1.f / d.z);
    V3 local = fmadd3(d, t, o);
    if (!(t >= ray.mint && t <= ray.maxt && fabsf(local.x) <= 1.f && fabsf(local.y) <= 1.f)) return false;
    t_out = t; lx = local.x; ly = local.y;
    return true;
}

// Mesh::ray_intersect_triangle (include/mitsuba/render/mesh.h:202-230); e1, e2 precomputed
// as p1-p0, p2-p0 in FP32 (same bits as computing them here).
SMS_DEV bool tri_intersect(V3 p0, V3 e1, V3 e2, const Ray &ray, float maxt, float &t_out, float &u_out, float &v_out) {
    V3 pvec = cross(ray.d, e2);
    float inv_det = 1.f / dot(e1, pvec);
    V3 tvec = ray.o - p0;
    float u = dot(tvec, pvec) * inv_det;
    if (!(u >= 0.f && u <= 1.f)) return false;
    V3 qvec = cross(tvec, e1);
    float v = dot(ray.d, qvec) * inv_det;
    if (!(v >= 0.f && u + v <= 1.f)) return false;
    float t = dot(e2, qvec) * inv_det;
    if (!(t >= ray.mint && t <= maxt)) return false;
    t_out = t; u_out = u; v_out = v;
    return true;
}

SMS_DEV bool box_hit(float4 a, float4 b, const Ray &r, float tmax) {
    float t0 = r.mint, t1 = tmax;
    float ta, tb, lo, hi;
    ta = (a.x - r.o.x) * r.d_rcp.x; tb = (b.x - r.o.x) * r.d_rcp.x;
    if (ta != ta) ta = -__builtin_inff();
    if (tb != tb) tb = __builtin_inff();
    lo = fmin_(ta, tb); hi = fmax_(ta, tb); t0 = fmax_(t0, lo); t1 = fmin_(t1, hi);
    ta = (a.y - r.o.y) * r.d_rcp.y; tb = (b.y - r.o.y) * r.d_rcp.y;
    if (ta != ta) ta = -__builtin_inff();
    if (tb != tb) tb = __builtin_inff();
    lo = fmin_(ta, tb); hi = fmax_(ta, tb); t0 = fmax_(t0, lo); t1 = fmin_(t1, hi);
    ta = (a.z - r.o.z) * r.d_rcp.z; tb = (b.z - r.o.z) * r.d_rcp.z;
    if (ta != ta) ta = -__builtin_inff();
    if (tb != tb) tb = __builtin_inff();
    lo = fmin_(ta, tb); hi = fmax_(ta, tb); t0 = fmax_(t0, lo); t1 = fmin_(t1, hi);
    return t0 <= t1 * (1.f + 4.f * EPS_F);
}

SMS_DEV bool better(float t, int shape, int prim, float bt, int bshape, int bprim) {
    if (t < bt) return true;
    if (t > bt) return false;
    if (shape != bshape) return shape < bshape;
    return prim < bprim;
}

struct Hit { bool hit; float t; int shape, prim; float b1, b2, lx, ly; };

// Entry distance of a ray into a (padded) box, or +inf if the conservative slab test misses it;
// the test is box_hit's, so a box is entered exactly when box_hit accepts it.
SMS_DEV float box_enter(float4 a, float4 b, const Ray &r, float tmax) {
    float t0 = r.mint, t1 = tmax;
    float ta, tb, lo, hi;
    ta = (a.x - r.o.x) * r.d_rcp.x; tb = (b.x - r.o.x) * r.d_rcp.x;
    if (ta != ta) ta = -__builtin_inff();
    if (tb != tb) tb = __builtin_inff();
    lo = fmin_(ta, tb); hi = fmax_(ta, tb); t0 = fmax_(t0, lo); t1 = fmin_(t1, hi);
    ta = (a.y - r.o.y) * r.d_rcp.y; tb = (b.y - r.o.y) * r.d_rcp.y;
    if (ta != ta) ta = -__builtin_inff();
    if (tb != tb) tb = __builtin_inff();
    lo = fmin_(ta, tb); hi = fmax_(ta, tb); t0 = fmax_(t0, lo); t1 = fmin_(t1, hi);
    ta = (a.z - r.o.z) * r.d_rcp.z; tb = (b.z - r.o.z) * r.d_rcp.z;
    if (ta != ta) ta = -__builtin_inff();
    if (tb != tb) tb = __builtin_inff();
    lo = fmin_(ta, tb); hi = fmax_(ta, tb); t0 = fmax_(t0, lo); t1 = fmin_(t1, hi);
    return t0 <= t1 * (1.f + 4.f * EPS_F) ? t0 : __builtin_inff();
}

// Scene::ray_intersect / ray_test (scene_native.inl:17-55): closest hit over every shape,
// ties broken by (t, shape, prim) so the result does not depend on traversal order.
// BVH = false drops the triangle-BVH traversal (and its per-lane stack) for scenes made of
// analytic shapes only.  LDS_STACK = true keeps the traversal stack in LDS (LDS_DEPTH entries
// per lane, interleaved by lane; the kernel's blocks must have 256 threads and the scene's BVH
// must be at most LDS_DEPTH deep) instead of per-lane scratch memory.  The host rejects BVHs
// deeper than SMS_BVH_STACK.
//
// Traversal: the two children of a node sit side by side in one 64-byte aligned pair; both
// boxes are tested against the current closest hit, the nearer entered box is visited next and
// the other pushed; a popped node is re-tested against the closest hit found since.  A box is
// skipped only when its entry lies beyond the closest hit, so every triangle that could be the
// closest hit or tie with it is tested and the result is that of an exhaustive search.
constexpr int SMS_BVH_STACK = 64;
template <bool BVH, bool LDS_STACK = false, int LDS_DEPTH = SMS_BVH_STACK>
SMS_DEV Hit scene_hit_t(const DScene &S, const Ray &ray, bool any_hit) {
    Hit h; h.hit = false; h.t = __builtin_inff(); h.shape = 1 << 30; h.prim = 1 << 30;
    h.b1 = h.b2 = h.lx = h.ly = 0.f;
    for (int k = 0; k < S.n_analytic; ++k) {
        int i = S.analytic[k];
        const DShape &sh = S.shapes[i];
        float t = 0.f, lx = 0.f, ly = 0.f;
        bool ok;
        Ray r = ray; r.maxt = fmin_(ray.maxt, h.t);
        if (sh.type == SMS_SHAPE_SPHERE) ok = sphere_intersect(sh, r, t);
        else ok = rect_intersect(sh, r, t, lx, ly);
        if (ok && better(t, i, 0, h.t, h.shape, h.prim)) {
            h.hit = true; h.t = t; h.shape = i; h.prim = 0; h.lx = lx; h.ly = ly;
            if (any_hit) return h;
        }
    }
    // A ray with a non-finite origin or direction (a Newton step that produced NaN) hits no
    // triangle -- tri_intersect rejects it -- but passes every box test, so it would walk the
    // whole tree; skip the traversal (the result is the same).
    const bool finite = isfinite(ray.o.x) && isfinite(ray.o.y) && isfinite(ray.o.z) && isfinite(ray.d.x) &&
                        isfinite(ray.d.y) && isfinite(ray.d.z);
    if (BVH && S.n_nodes > 0 && finite) {
        int local_stack[LDS_STACK ? 1 : SMS_BVH_STACK];
        int *stack = local_stack;
        int stride = 1;
        if (LDS_STACK) {
            __shared__ int lds_stack[LDS_DEPTH * 256];
            stack = lds_stack + threadIdx.x;
            stride = 256;
        }
        int sp = 0;
        DNode nd = S.nodes[0];
        bool have = box_enter(nd.a, nd.b, ray, fmin_(ray.maxt, h.t)) < __builtin_inff();
        while (true) {
            if (have) {
                const int count = __float_as_int(nd.b.w);
                const int link = __float_as_int(nd.a.w);
                if (count > 0) {
                    for (int k = link; k < link + count; ++k) {
                        float4 r0 = S.tris[3 * k], r1 = S.tris[3 * k + 1], r2 = S.tris[3 * k + 2];
                        float t, u, v;
                        if (tri_intersect(v3(r0.x, r0.y, r0.z), v3(r1.x, r1.y, r1.z), v3(r2.x, r2.y, r2.z), ray,
                                          fmin_(ray.maxt, h.t), t, u, v)) {
                            int shp = __float_as_int(r0.w), prim = __float_as_int(r1.w);
                            if (better(t, shp, prim, h.t, h.shape, h.prim)) {
                                h.hit = true; h.t = t; h.shape = shp; h.prim = prim; h.b1 = u; h.b2 = v;
                                if (any_hit) return h;
                            }
                        }
                    }
                } else {
                    const DNode c0 = S.nodes[link], c1 = S.nodes[link + 1];
                    const float tmax = fmin_(ray.maxt, h.t);
                    const float e0 = box_enter(c0.a, c0.b, ray, tmax), e1 = box_enter(c1.a, c1.b, ray, tmax);
                    const bool h0 = e0 < __builtin_inff(), h1 = e1 < __builtin_inff();
                    if (h0 && h1) {
                        const bool first0 = e0 <= e1;
                        stack[stride * sp++] = first0 ? link + 1 : link;
                        nd = first0 ? c0 : c1;
                        continue;
                    }
                    if (h0 || h1) {
                        nd = h0 ? c0 : c1;
                        continue;
                    }
                }
            }
            if (sp == 0) break;
            nd = S.nodes[stack[stride * --sp]];
            have = box_enter(nd.a, nd.b, ray, fmin_(ray.maxt, h.t)) < __builtin_inff();
        }
    }
    return h;
}

// Angle between the unit vector v and +z from the half chord to the pole, which stays accurate
// near both poles (unit_angle_z of Mitsuba 2's core vector.h).
SMS_DEV float unit_angle_z(V3 v) {
    const float dz = v.z - mulsign(1.f, v.z);
    const float temp = 2.f * c_asinf(.5f * sqrtf((v.x * v.x + v.y * v.y) + dz * dz));
    return v.z >= 0.f ? temp : PI_F - temp;
}

// sphere.cpp:343-348: (phi / 2 pi, theta / pi) of the object-space point
SMS_DEV V2 sphere_uv(V3 local) {
    const float theta = unit_angle_z(local);
    float phi = c_atan2f(local.y, local.x);
    if (phi < 0.f) phi = phi + 2.f * PI_F;
    return v2(phi * INV_TWO_PI_F, theta * INV_PI_F);
}

// Shape::fill_surface_interaction + kd-tree shading frame (kdtree.h:2335-2364)
SMS_DEV SI make_si(const DScene &S, const Ray &ray, const Hit &h) {
    if (!h.hit) return si_invalid();
    SI si = si_invalid();
    si.valid = true; si.t = h.t; si.shape = h.shape; si.prim = h.prim;
    const DShape &sh = S.shapes[h.shape];
    if (sh.type == SMS_SHAPE_SPHERE) {
        // sphere.cpp:329-374; uv is read only through a normal map, so only normal-mapped
        // spheres compute it
        V3 c = v3(sh.cx, sh.cy, sh.cz);
        V3 n = normalize(fmadd3(ray.d, si.t, ray.o) - c);
        si.sh_n = n;
        si.p = fmadd3(n, sh.radius, c);
        V3 local = (si.p - c) * sh.inv_radius;
        float rd_2 = local.x * local.x + local.y * local.y;
        if (sh.normalmap >= 0) si.uv = sphere_uv(local);
        si.dp_du = v3(-local.y, local.x, 0.f);
        float rd = sqrtf(rd_2), inv_rd = 1.f / rd;
        float cos_phi = local.x * inv_rd, sin_phi = local.y * inv_rd;
        si.dp_dv = v3(local.z * cos_phi, local.z * sin_phi, -rd);
        if (rd == 0.f) si.dp_dv = v3(1.f, 0.f, 0.f);
        si.dp_du = (si.dp_du * sh.radius) * (2.f * PI_F);
        si.dp_dv = (si.dp_dv * sh.radius) * PI_F;
        si.n = si.sh_n;
    } else if (sh.type == SMS_SHAPE_RECTANGLE) {
        // rectangle.cpp:175-202
        si.n = ld3(sh.frn); si.sh_n = si.n;
        si.dp_du = ld3(sh.frs); si.dp_dv = ld3(sh.frt);
        si.p = fmadd3(ray.d, si.t, ray.o);
        si.uv = v2(fmaf(h.lx, .5f, .5f), fmaf(h.ly, .5f, .5f));
    } else {
        // mesh.cpp:620-691
        float b1 = h.b1, b2 = h.b2, b0 = 1.f - b1 - b2;
        uint4 f = S.faces[sh.fbase + h.prim];
        float4 q0 = S.vpos[f.x], q1 = S.vpos[f.y], q2 = S.vpos[f.z];
        V3 p0 = v3(q0.x, q0.y, q0.z), p1 = v3(q1.x, q1.y, q1.z), p2 = v3(q2.x, q2.y, q2.z);
        V3 dp0 = p1 - p0, dp1 = p2 - p0;
        si.p = (p0 * b0 + p1 * b1) + p2 * b2;
        V3 n = normalize(cross(dp0, dp1));
        si.n = n;
        V3 dpdu, dpdv;
        coordinate_system(n, dpdu, dpdv);
        V2 uv = v2(b1, b2);
        if (sh.has_uv) {
            float2 t0 = S.vuv[f.x], t1 = S.vuv[f.y], t2 = S.vuv[f.z];
            V2 uv0 = v2(t0.x, t0.y), uv1 = v2(t1.x, t1.y), uv2 = v2(t2.x, t2.y);
            uv = (uv0 * b0 + uv1 * b1) + uv2 * b2;
            V2 duv0 = uv1 - uv0, duv1 = uv2 - uv0;
            float det = fmaf(duv0.x, duv1.y, -(duv0.y * duv1.x)), inv_det = 1.f / det;
            if (det != 0.f) {
                V3 a = duv0.y * dp1;
                dpdu = v3(fmaf(duv1.y, dp0.x, -a.x), fmaf(duv1.y, dp0.y, -a.y), fmaf(duv1.y, dp0.z, -a.z)) * inv_det;
                V3 c = duv0.x * dp1;
                dpdv = v3(fmaf(-duv1.x, dp0.x, c.x), fmaf(-duv1.x, dp0.y, c.y), fmaf(-duv1.x, dp0.z, c.z)) * inv_det;
            }
        }
        si.uv = uv;
        if (sh.has_n) {
            float4 a0 = S.vnrm[f.x], a1 = S.vnrm[f.y], a2 = S.vnrm[f.z];
            V3 n0 = v3(a0.x, a0.y, a0.z), n1 = v3(a1.x, a1.y, a1.z), n2 = v3(a2.x, a2.y, a2.z);
            n = normalize((n0 * b0 + n1 * b1) + n2 * b2);
        }
        si.sh_n = n;
        si.dp_du = dpdu; si.dp_dv = dpdv;
    }
    si.sh_s = normalize(fnmadd3(si.sh_n, dot(si.sh_n, si.dp_du), si.dp_du));
    si.sh_t = cross(si.sh_n, si.sh_s);
    const V3 mw = -ray.d;
    si.wi = v3(dot(mw, si.sh_s), dot(mw, si.sh_t), dot(mw, si.sh_n));
    return si;
}

SMS_DEV Hit scene_hit(const DScene &S, const Ray &ray, bool any_hit) { return scene_hit_t<true>(S, ray, any_hit); }
SMS_DEV SI ray_intersect(const DScene &S, const Ray &ray) { return make_si(S, ray, scene_hit(S, ray, false)); }
SMS_DEV bool ray_test(const DScene &S, const Ray &ray) { return scene_hit(S, ray, true).hit; }
template <bool BVH> SMS_DEV bool ray_test_t(const DScene &S, const Ray &ray) { return scene_hit_t<BVH>(S, ray, true).hit; }

// ---------------------------------------------------------------- normal derivatives
SMS_DEV void normal_derivative(const DScene &S, const SI &si, V3 &dn_du, V3 &dn_dv) {
    const DShape &sh = S.shapes[si.shape];
    if (sh.type == SMS_SHAPE_SPHERE) {
        dn_du = si.dp_du * sh.inv_radius; dn_dv = si.dp_dv * sh.inv_radius;     // sphere.cpp:376-382
    } else if (sh.type == SMS_SHAPE_RECTANGLE || !sh.has_n) {
        dn_du = v3(0.f, 0.f, 0.f); dn_dv = v3(0.f, 0.f, 0.f);
    } else {
        // mesh.cpp:694-760
        uint4 f = S.faces[sh.fbase + si.prim];
        float4 q0 = S.vpos[f.x], q1 = S.vpos[f.y], q2 = S.vpos[f.z];
        float4 a0 = S.vnrm[f.x], a1 = S.vnrm[f.y], a2 = S.vnrm[f.z];
        V3 p0 = v3(q0.x, q0.y, q0.z), p1 = v3(q1.x, q1.y, q1.z), p2 = v3(q2.x, q2.y, q2.z);
        V3 n0 = v3(a0.x, a0.y, a0.z), n1 = v3(a1.x, a1.y, a1.z), n2 = v3(a2.x, a2.y, a2.z);
        V3 rel = si.p - p0, du = p1 - p0, dv = p2 - p0;
        float b1 = dot(du, rel), b2 = dot(dv, rel), a11 = dot(du, du), a12 = dot(du, dv), a22 = dot(dv, dv),
              inv_det = 1.f / (a11 * a22 - a12 * a12);
        float u = fmaf(a22, b1, -(a12 * b2)) * inv_det, v = fmaf(-a12, b1, a11 * b2) * inv_det, w = 1.f - u - v;
        V3 N = (u * n1 + v * n2) + w * n0;
        float il = 1.f / sqrtf(sqnorm(N));
        N = N * il;
        V3 dndu = (n1 - n0) * il, dndv = (n2 - n0) * il;
        dndu = fnmadd3(N, dot(N, dndu), dndu);
        dndv = fnmadd3(N, dot(N, dndv), dndv);
        if (sh.has_uv) {
            float2 t0 = S.vuv[f.x], t1 = S.vuv[f.y], t2 = S.vuv[f.z];
            V2 duv1 = v2(t1.x - t0.x, t1.y - t0.y), duv2 = v2(t2.x - t0.x, t2.y - t0.y);
            inv_det = 1.f / (duv1.x * duv2.y - duv1.y * duv2.x);
            V3 a = (duv2.y * dndu - duv1.y * dndv) * inv_det, b = (-duv2.x * dndu + duv1.x * dndv) * inv_det;
            dndu = a; dndv = b;
        }
        dn_du = dndu; dn_dv = dndv;
    }
}

// ---------------------------------------------------------------- normal map (level 0)
struct Texel4 { V2 v00, v10, v01, v11, w; };
struct NmCoord { int x, y; V2 w; };
// Normalmap::eval_normal / eval_normal_derivatives index computation (normalmap.h:62-80):
// the lower-left texel (clamped to res-1) and the bilinear weights; the upper texel is
// x+1 / y+1 in the padded layout.
SMS_DEV NmCoord nm_coord(int res, V2 uv) {
    float fres = (float) res;
    float duv = 1.f / fres;
    V2 pos = v2((uv.x - 0.5f * duv) * fres, (uv.y - 0.5f * duv) * fres);
    float hi = (float) (res - 1);
    pos.x = fmin_(fmax_(pos.x, 0.f), hi); pos.y = fmin_(fmax_(pos.y, 0.f), hi);
    uint32_t xmin = (uint32_t) floorf(pos.x), ymin = (uint32_t) floorf(pos.y);
    uint32_t lim = (uint32_t) (res - 1);
    xmin = xmin < lim ? xmin : lim; ymin = ymin < lim ? ymin : lim;
    NmCoord c;
    c.x = (int) xmin; c.y = (int) ymin;
    c.w = v2(pos.x - (float) xmin, pos.y - (float) ymin);
    return c;
}
struct alignas(8) F4 { float x, y, z, w; };
SMS_DEV Texel4 nm_fetch(const float2 *__restrict__ slope, int res, int stride, V2 uv) {
    NmCoord c = nm_coord(res, uv);
    const float2 *row = slope + (size_t) c.y * stride + c.x;
    F4 a = *(const F4 *) row, b = *(const F4 *) (row + stride);
    Texel4 t;
    t.v00 = v2(a.x, a.y); t.v10 = v2(a.z, a.w); t.v01 = v2(b.x, b.y); t.v11 = v2(b.z, b.w);
    t.w = c.w;
    return t;
}
SMS_DEV V2 nm_slope(const Texel4 &t) {
    V2 w1 = t.w, w0 = v2(1.f - w1.x, 1.f - w1.y);
    V2 a = v2(fmaf(w0.x, t.v00.x, w1.x * t.v10.x), fmaf(w0.x, t.v00.y, w1.x * t.v10.y));
    V2 b = v2(fmaf(w0.x, t.v01.x, w1.x * t.v11.x), fmaf(w0.x, t.v01.y, w1.x * t.v11.y));
    return v2(fmaf(w0.y, a.x, w1.y * b.x), fmaf(w0.y, a.y, w1.y * b.y));
}
SMS_DEV void nm_slope_deriv(const Texel4 &t, float res, V2 &tu, V2 &tv) {
    V2 tmp = (t.v01 + t.v10) - t.v11;
    tu = ((t.v10 + t.v00 * (t.w.y - 1.f)) - tmp * t.w.y) * res;
    tv = ((t.v01 + t.v00 * (t.w.x - 1.f)) - tmp * t.w.x) * res;
}
// The map's local normal (not normalised) at level 0 and its uv derivatives (normalmap.h:62-191
// with NormalmapBSDF::frame / frame_derivative, bsdfs/normalmap.cpp:226-258): a slope map gives
// n = (-s_x, -s_y, 1), an encoded map n = 2 rgb - 1 of the bilinearly interpolated RGB
SMS_DEV void nm_local_normal(const DNormalmap &nm, V2 uv, V3 &nl, V3 &du, V3 &dv) {
    if (!nm.encoded) {
        const Texel4 tx = nm_fetch(nm.slope, nm.res, nm.stride, uv);
        const V2 sl = nm_slope(tx);
        V2 tu, tv;
        nm_slope_deriv(tx, (float) nm.res, tu, tv);
        nl = v3(-sl.x, -sl.y, 1.f);
        du = v3(-tu.x, -tu.y, 0.f); dv = v3(-tv.x, -tv.y, 0.f);
        return;
    }
    const NmCoord c = nm_coord(nm.res, uv);
    const float4 *row = nm.rgbn + (size_t) c.y * nm.stride + c.x;
    const float4 a0 = row[0], a1 = row[1], b0 = row[nm.stride], b1 = row[nm.stride + 1];
    const V3 v00 = v3(a0.x, a0.y, a0.z), v10 = v3(a1.x, a1.y, a1.z), v01 = v3(b0.x, b0.y, b0.z),
             v11 = v3(b1.x, b1.y, b1.z);
    const V2 w1 = c.w, w0 = v2(1.f - w1.x, 1.f - w1.y);
    const V3 r0 = v3(fmaf(w0.x, v00.x, w1.x * v10.x), fmaf(w0.x, v00.y, w1.x * v10.y), fmaf(w0.x, v00.z, w1.x * v10.z));
    const V3 r1 = v3(fmaf(w0.x, v01.x, w1.x * v11.x), fmaf(w0.x, v01.y, w1.x * v11.y), fmaf(w0.x, v01.z, w1.x * v11.z));
    const V3 rgb = v3(fmaf(w0.y, r0.x, w1.y * r1.x), fmaf(w0.y, r0.y, w1.y * r1.y), fmaf(w0.y, r0.z, w1.y * r1.z));
    nl = v3(2.f * rgb.x - 1.f, 2.f * rgb.y - 1.f, 2.f * rgb.z - 1.f);
    const float fres = (float) nm.res;
    const V3 tmp = (v01 + v10) - v11;
    du = (((v10 + v00 * (w1.y - 1.f)) - tmp * w1.y) * fres) * 2.f;
    dv = (((v01 + v00 * (w1.x - 1.f)) - tmp * w1.x) * fres) * 2.f;
}

SMS_DEV V2 tile_uv(V2 uv, float tiles) {
    uv = uv * tiles;
    return v2(uv.x - floorf(uv.x), uv.y - floorf(uv.y));
}

// ---------------------------------------------------------------- ManifoldVertex
struct MV {
    V3 p, dp_du, dp_dv, n, gn, dn_du, dn_dv, s, t, ds_du, ds_dv, dt_du, dt_dv;
    float eta; int shape; bool fixed_direction;
};
SMS_DEV MV mv_point(V3 p) {
    MV v;
    v.p = p; v.dp_du = v.dp_dv = v.n = v.gn = v.dn_du = v.dn_dv = v.s = v.t = v3(0.f, 0.f, 0.f);
    v.ds_du = v.ds_dv = v.dt_du = v.dt_dv = v3(0.f, 0.f, 0.f);
    v.eta = 1.f; v.shape = -1; v.fixed_direction = false;
    return v;
}
SMS_DEV V3 to_world(V3 s, V3 t, V3 n, V3 v) { return (s * v.x + t * v.y) + n * v.z; }

// compute_shading_frame_derivative (include/mitsuba/core/frame.h:224-244)
SMS_DEV void shading_frame_derivative(V3 n, V3 dp_du, V3 dn_du, V3 dn_dv, V3 &ds_du, V3 &ds_dv, V3 &dt_du, V3 &dt_dv) {
    V3 s = dp_du - n * dot(n, dp_du);
    float inv_len_s = 1.f / norm(s);
    s = s * inv_len_s;
    ds_du = inv_len_s * (-dn_du * dot(n, dp_du) - n * dot(dn_du, dp_du));
    ds_dv = inv_len_s * (-dn_dv * dot(n, dp_du) - n * dot(dn_dv, dp_du));
    ds_du = ds_du - s * dot(ds_du, s);
    ds_dv = ds_dv - s * dot(ds_dv, s);
    dt_du = cross(dn_du, s) + cross(n, ds_du);
    dt_dv = cross(dn_dv, s) + cross(n, ds_dv);
}

// ManifoldVertex(si, smoothing = 0) (include/mitsuba/render/manifold.h:89-113) with the BSDF
// frame / frame_derivative of bsdf.cpp:17-31 or of the normal-map BSDF (bsdfs/normalmap.cpp:218-290)
// smooth: ManifoldVertex(si, smoothing = 1) -- a normal map at its coarsest mip level (one
// texel: the mean slope, zero derivatives; normalmap.h:236-268); otherwise smoothing 0
SMS_DEV MV mv_from_si(const DScene &S, const SI &si, bool smooth = false) {
    MV v = mv_point(si.p);
    v.dp_du = si.dp_du; v.dp_dv = si.dp_dv; v.gn = si.n; v.shape = si.shape;
    const DShape &sh = S.shapes[si.shape];
    v.eta = sh.eta_vtx;
    V3 bn = si.sh_n;
    V3 bs = normalize(fnmadd3(bn, dot(bn, si.dp_du), si.dp_du));
    V3 bt = cross(bn, bs);
    V3 sdn_du, sdn_dv;
    normal_derivative(S, si, sdn_du, sdn_dv);
    if (sh.bsdf != SMS_BSDF_NORMALMAP) {
        v.n = bn; v.s = bs; v.t = bt;
        shading_frame_derivative(bn, si.dp_du, sdn_du, sdn_dv, v.ds_du, v.ds_dv, v.dt_du, v.dt_dv);
        v.dn_du = sdn_du; v.dn_dv = sdn_dv;
    } else {
        const DNormalmap nm = S.nmaps[sh.normalmap];
        V3 nl, dl_du, dl_dv;
        if (smooth) {
            // the coarsest level is one texel: bilinear weights (0, 0), derivatives zero (their
            // signs as the slope and RGB branches of eval_normal_derivatives form them)
            if (nm.encoded) {
                nl = v3(2.f * nm.top_rgb[0] - 1.f, 2.f * nm.top_rgb[1] - 1.f, 2.f * nm.top_rgb[2] - 1.f);
                dl_du = v3(0.f, 0.f, 0.f);
            } else {
                const float *top = nm.lean + nm.lean_off[nm.levels - 1];
                nl = v3(-top[0], -top[1], 1.f);
                dl_du = v3(-0.f, -0.f, 0.f);
            }
            dl_dv = dl_du;
        } else {
            nm_local_normal(nm, tile_uv(si.uv, sh.tiles), nl, dl_du, dl_dv);
        }
        V3 dn_du = dl_du * sh.tiles, dn_dv = dl_dv * sh.tiles;
        V3 rn = normalize(to_world(bs, bt, bn, nl));
        v.n = rn;
        v.s = normalize(si.dp_du - rn * dot(rn, si.dp_du));
        v.t = cross(rn, v.s);
        V3 b_ds_du, b_ds_dv, b_dt_du, b_dt_dv;
        shading_frame_derivative(bn, si.dp_du, sdn_du, sdn_dv, b_ds_du, b_ds_dv, b_dt_du, b_dt_dv);
        V3 world_n = to_world(bs, bt, bn, nl);
        float inv_length_n = 1.f / norm(world_n);
        world_n = world_n * inv_length_n;
        V3 dnu = inv_length_n * (to_world(bs, bt, bn, dn_du) + to_world(b_ds_du, b_dt_du, sdn_du, nl));
        V3 dnv = inv_length_n * (to_world(bs, bt, bn, dn_dv) + to_world(b_ds_dv, b_dt_dv, sdn_dv, nl));
        dnu = dnu - world_n * dot(dnu, world_n);
        dnv = dnv - world_n * dot(dnv, world_n);
        V3 s = si.dp_du - world_n * dot(world_n, si.dp_du);
        float inv_length_s = 1.f / norm(s);
        s = s * inv_length_s;
        V3 dsu = inv_length_s * (-dnu * dot(world_n, si.dp_du) - world_n * dot(dnu, si.dp_du));
        V3 dsv = inv_length_s * (-dnv * dot(world_n, si.dp_du) - world_n * dot(dnv, si.dp_du));
        dsu = dsu - s * dot(dsu, s);
        dsv = dsv - s * dot(dsv, s);
        v.dn_du = dnu; v.dn_dv = dnv; v.ds_du = dsu; v.ds_dv = dsv;
        v.dt_du = cross(dnu, s) + cross(world_n, dsu);
        v.dt_dv = cross(dnv, s) + cross(world_n, dsv);
    }
    if (dot(v.n, v.gn) < 0.f) v.gn = v.gn * -1.f;
    return v;
}

// ManifoldVertex::make_orthonormal (manifold.h:115-126)
SMS_DEV void make_orthonormal(MV &v) {
    float inv_norm = 1.f / norm(v.dp_du);
    v.dp_du = v.dp_du * inv_norm;
    v.dn_du = v.dn_du * inv_norm;
    float dp = dot(v.dp_du, v.dp_dv);
    V3 dp_dv_tmp = v.dp_dv - dp * v.dp_du;
    V3 dn_dv_tmp = v.dn_dv - dp * v.dn_du;
    inv_norm = 1.f / norm(dp_dv_tmp);
    v.dp_dv = dp_dv_tmp * inv_norm;
    v.dn_dv = dn_dv_tmp * inv_norm;
}

// ---------------------------------------------------------------- emitters
struct EI { V3 p, n, d, weight; float pdf; int kind, emitter; };

SMS_DEV V3 square_to_uniform_sphere(V2 s) {
    float z = fmaf(-2.f, s.y, 1.f), r = safe_sqrt(fmaf(-z, z, 1.f));
    float sn, cs;
    c_sincosf((2.f * PI_F) * s.x, &sn, &cs);
    return v3(r * cs, r * sn, z);
}

struct PS { V3 p, n; float pdf; };
// Shape::sample_position: sphere.cpp:153-171, rectangle.cpp:114-126, mesh.cpp:542-585
SMS_DEV PS sample_position(const DScene &S, int shape, V2 sample) {
    const DShape &sh = S.shapes[shape];
    PS ps;
    if (sh.type == SMS_SHAPE_SPHERE) {
        V3 local = square_to_uniform_sphere(sample);
        ps.p = fmadd3(local, sh.radius, v3(sh.cx, sh.cy, sh.cz)); ps.n = local; ps.pdf = sh.inv_surface_area;
    } else if (sh.type == SMS_SHAPE_RECTANGLE) {
        ps.p = xform_p(sh.M, v3(sample.x * 2.f - 1.f, sample.y * 2.f - 1.f, 0.f));
        ps.n = ld3(sh.frn); ps.pdf = sh.inv_surface_area;
    } else {
        const float *cdf = S.cdf + sh.cdf_offset, *pmf_a = S.pmf + sh.cdf_offset;
        float value = sample.y;
        float scaled = value * sh.area_sum;
        int lo = 0, hi = sh.nf - 1;
        while (lo < hi) { int mid = (lo + hi) >> 1; if (cdf[mid] < scaled) lo = mid + 1; else hi = mid; }
        int idx = lo;
        float pmf = pmf_a[idx] * sh.area_norm;
        float c = idx > 0 ? cdf[idx - 1] * sh.area_norm : 0.f;
        float y = (value - c) / pmf;
        uint4 f = S.faces[sh.fbase + idx];
        float4 q0 = S.vpos[f.x], q1 = S.vpos[f.y], q2 = S.vpos[f.z];
        V3 p0 = v3(q0.x, q0.y, q0.z), p1 = v3(q1.x, q1.y, q1.z), p2 = v3(q2.x, q2.y, q2.z);
        V3 e0 = p1 - p0, e1 = p2 - p0;
        float t = safe_sqrt(1.f - sample.x);
        V2 b = v2(1.f - t, t * y);
        ps.p = (p0 + e0 * b.x) + e1 * b.y;
        if (sh.has_n) {
            float4 a0 = S.vnrm[f.x], a1 = S.vnrm[f.y], a2 = S.vnrm[f.z];
            V3 n0 = v3(a0.x, a0.y, a0.z), n1 = v3(a1.x, a1.y, a1.z), n2 = v3(a2.x, a2.y, a2.z);
            ps.n = normalize((n0 * (1.f - b.x - b.y) + n1 * b.x) + n2 * b.y);
        } else ps.n = normalize(cross(e0, e1));
        ps.pdf = sh.area_norm;
    }
    return ps;
}

// SpecularManifold::sample_emitter_interaction (manifold.h:203-260)
SMS_DEV EI sample_emitter_interaction(const DScene &S, V3 si_p, const int *emitters, int n_em, Pcg32 &rng) {
    EI ei;
    ei.p = ei.n = ei.d = ei.weight = v3(0.f, 0.f, 0.f); ei.pdf = 0.f; ei.kind = SMS_EI_AREA; ei.emitter = -1;
    if (n_em == 0) return ei;
    float emitter_sample = pcg32_float(rng);
    float emitter_pdf = 1.f / (float) n_em;
    uint32_t index = (uint32_t) (emitter_sample * (float) n_em);
    if (index > (uint32_t) n_em - 1u) index = (uint32_t) n_em - 1u;
    const DEmitter &em = S.emitters[emitters[index]];
    ei.emitter = emitters[index];
    V3 spec = v3(0.f, 0.f, 0.f);
    V2 s2; s2.x = pcg32_float(rng); s2.y = pcg32_float(rng);
    if (em.type == SMS_EMITTER_AREA) {
        ei.kind = SMS_EI_AREA;
        PS ps = sample_position(S, em.shape, s2);
        if (ps.pdf > 0.f) {
            spec = ld3(em.radiance) / ps.pdf;
            ei.p = ps.p; ei.n = ps.n; ei.d = normalize(ps.p - si_p); ei.pdf = ps.pdf;
        }
    } else if (em.type == SMS_EMITTER_DIRECTIONAL) {
        // DirectionalEmitter::sample_direction (src/emitters/directional.cpp:105-128)
        ei.kind = SMS_EI_DIRECTIONAL;
        V3 d = ld3(em.direction);
        ei.p = si_p - d * S.bs_dist; ei.d = -d; ei.n = -d; ei.pdf = 1.f;
        spec = ld3(em.irradiance);
    } else {
        ei.kind = SMS_EI_POINT;
        V3 pos = ld3(em.position);
        V3 d = pos - si_p;
        float dist = norm(d), inv_dist = 1.f / dist;
        d = d * inv_dist;
        ei.p = pos; ei.d = d; ei.n = d; ei.pdf = 1.f;
        spec = ld3(em.intensity) * (inv_dist * inv_dist);
        spec = spec * dist * dist;
    }
    ei.pdf *= emitter_pdf;
    ei.weight = spec * (1.f / emitter_pdf);
    return ei;
}

// SpecularManifold::emitter_interaction_to_vertex (manifold.h:263-302); LDS_STACK as scene_hit_t
template <bool LDS_STACK = false, int LDS_DEPTH = SMS_BVH_STACK>
SMS_DEV bool emitter_interaction_to_vertex(const DScene &S, const EI &ei, V3 p, MV &vy) {
    if (ei.kind == SMS_EI_AREA) {
        V3 d_tmp = normalize(ei.p - p);
        Ray r = make_ray(p + SHADOW_EPS_F * d_tmp, d_tmp);
        SI si_y = make_si(S, r, scene_hit_t<true, LDS_STACK, LDS_DEPTH>(S, r, false));
        if (!si_y.valid) { vy = mv_point(v3(0.f, 0.f, 0.f)); return false; }
        vy = mv_from_si(S, si_y);
        make_orthonormal(vy);
        return true;
    } else if (ei.kind == SMS_EI_DIRECTIONAL) {
        vy = mv_point(ei.p);
        V3 d = normalize(ei.p - p);
        vy.p = p + d; vy.n = -d;
        coordinate_system(vy.n, vy.dp_du, vy.dp_dv);
        vy.fixed_direction = true;
        return true;
    } else {
        vy = mv_point(ei.p);
        V3 d = normalize(p - ei.p);
        vy.n = d; vy.gn = d;
        coordinate_system(d, vy.dp_du, vy.dp_dv);
        return true;
    }
}

// ---------------------------------------------------------------- Fresnel (fresnel.h:34-122)
SMS_DEV float fresnel(float cos_theta_i, float eta) {
    bool outside = cos_theta_i >= 0.f;
    float rcp_eta = 1.f / eta, eta_it = outside ? eta : rcp_eta, eta_ti = outside ? rcp_eta : eta;
    float cos_theta_t_sqr = fmaf(-fmaf(-cos_theta_i, cos_theta_i, 1.f), eta_ti * eta_ti, 1.f);
    float cos_theta_i_abs = fabsf(cos_theta_i), cos_theta_t_abs = safe_sqrt(cos_theta_t_sqr);
    bool index_matched = eta == 1.f, special = index_matched || cos_theta_i_abs == 0.f;
    float r_sc = index_matched ? 0.f : 1.f;
    float a_s = fmaf(-eta_it, cos_theta_t_abs, cos_theta_i_abs) / fmaf(eta_it, cos_theta_t_abs, cos_theta_i_abs);
    float a_p = fmaf(-eta_it, cos_theta_i_abs, cos_theta_t_abs) / fmaf(eta_it, cos_theta_i_abs, cos_theta_t_abs);
    float r = .5f * (a_s * a_s + a_p * a_p);
    return special ? r_sc : r;
}
SMS_DEV float fresnel_conductor(float cos_theta_i, float eta_r, float eta_i) {
    float c2 = cos_theta_i * cos_theta_i, s2 = 1.f - c2, s4 = s2 * s2;
    float temp_1 = eta_r * eta_r - eta_i * eta_i - s2;
    float a_2_pb_2 = safe_sqrt(temp_1 * temp_1 + 4.f * eta_i * eta_i * eta_r * eta_r);
    float a = safe_sqrt(.5f * (a_2_pb_2 + temp_1));
    float term_1 = a_2_pb_2 + c2, term_2 = 2.f * cos_theta_i * a;
    float r_s = (term_1 - term_2) / (term_1 + term_2);
    float term_3 = a_2_pb_2 * c2 + s4, term_4 = term_2 * s2;
    float r_p = r_s * (term_3 - term_4) / (term_3 + term_4);
    return .5f * (r_s + r_p);
}

} // namespace smsdev
