/*
 * sms_oracle.cpp -- TEST INFRASTRUCTURE ONLY: scalar CPU restatement of the reference's
 * Specular Manifold Sampling solver path.  Only tests/, __graft_entry__.smoke() and the
 * CPU-baseline leg of bench.py load it, as the checker.  The product never links it.
 *
 * Each routine cites the reference file:line it restates (paths relative to the reference
 * repository).  Parity status of every part is recorded in the test-suite; parts that no
 * closed form, brute force or invariant pins say so where they are tested.
 *
 * Scene model: analytic spheres and rectangles are tested one by one; triangle meshes go
 * through a median-split BVH of this file (boxes padded so that traversal is conservative
 * and the closest hit equals brute force; ties broken by (t, shape, prim)).
 */
#include "oracle_math.h"
#include "../include/sms.h"

#include <vector>
#include <limits>
#include <algorithm>
#include <thread>
#include <atomic>
#include <stdlib.h>
#include <stdio.h>

using namespace orc;

namespace {

/* ---------------------------------------------------------------- planted-error hooks
   Factors the tests set to plant a known mistake into one term of the oracle and check that the
   pin of that term fails.  Every factor is 1 unless a test sets it; x * 1.0f == x exactly, so
   the unplanted oracle is bit-identical to one without the hooks. */
enum {
    PLANT_SS_G = 0,         /* first row of dC1/dx2 in the single-scatter geometric term      */
    PLANT_SS_TRANSMIT,      /* specular_reflectance of a single-scatter dielectric             */
    PLANT_SS_AD_JAC,        /* dC/du_0 of the single-scatter angle-difference Jacobian          */
    PLANT_SS_HV_JAC,        /* dH/du_0 of the single-scatter half-vector Jacobian               */
    PLANT_MS_AD_JAC,        /* dC_i/dx_i (0,0) of the multi-scatter angle-difference blocks     */
    PLANT_MS_HV_JAC,        /* dC_i/dx_i (0,0) of the multi-scatter half-vector blocks          */
    PLANT_MNEE_TARGET,      /* scale of the light point the MNEE seed ray aims at              */
    PLANT_MS_G,             /* the multi-scatter generalized geometric term                     */
    PLANT_SS_G_DIR,         /* the single-scatter geometric term toward a directional emitter  */
    PLANT_SS_G_CLAMP,       /* the bound (1) of |det dx1/dx2| in the single-scatter term        */
    PLANT_N = 16
};
static float g_plant[PLANT_N] = {1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f};

/* ---------------------------------------------------------------- scene */
struct Shape {
    int type, bsdf, flags, emitter;
    V3 center; float radius, inv_radius;
    float M[12], Minv[12];          // rectangle to_world / to_object (row-major 3x4)
    V3 fr_s, fr_t, fr_n;            // rectangle m_frame (dp_du, dp_dv, normal)
    float eta_r, k, eta_vtx;        // ior() = (eta_r, k) and the ManifoldVertex eta encoding
    float alpha;                    // rough conductor: isotropic Beckmann roughness (0: smooth)
    V3 bbox_center;
    V3 refl;
    int normalmap; float tiles;
    // mesh
    int vbase, fbase, nv, nf; bool has_n, has_uv;
    std::vector<float> cdf, pmf; float area_sum, area_norm;
    float inv_surface_area;
};

struct Emitter { int type, shape; V3 radiance, position, intensity, direction, irradiance; bool ss, ms; };

struct Normalmap {
    int res;
    std::vector<float> slope;       // level-0 slopes (res*res*2), normalmap.cpp:46-66
    /* LEAN mip pyramid: level l has (res >> l)^2 texels of (E[x], E[y], E[x^2], E[y^2], E[xy])
       (normalmap.cpp:36-87) */
    std::vector<std::vector<float>> lean;
    /* maps without use_slopes (sms_normalmap_desc.encoded_normals): the stored RGB (res*res*3)
       and the RGB of the coarsest level of its box-filter pyramid */
    bool encoded = false;
    std::vector<float> rgb;
    float top_rgb[3] = {0.f, 0.f, 0.f};
};

struct Camera {
    V3 o, left, up, dir;            // look-at frame (columns of camera-to-world)
    float tan_half, aspect;
    V3 dx, dy;                      // near-plane differentials (perspective.cpp:136-140)
    float rect_x, rect_y, normalization;
    int w, h;
};

struct Node { float bmin[3], bmax[3]; int left, right, first, count; };

struct Scene {
    std::vector<Shape> shapes;
    std::vector<Emitter> emitters;
    std::vector<Normalmap> nmaps;
    std::vector<float> pos, nrm, uvs;    // all mesh vertices
    std::vector<uint32_t> faces;         // all mesh faces (global vertex indices)
    std::vector<int> tri_shape, tri_prim;
    std::vector<int> tri_order;
    std::vector<Node> nodes;
    std::vector<int> casters_ss, emitters_ss, emitters_all;
    Camera cam;
    bool has_cam;
    float box_c[3], box_inv;             // bounding box centre, 1 / largest half extent
    float bs_dist;                       // directional emitters: 2 x the scene's bounding-sphere radius
    sms_config cfg;
};

struct Ray { V3 o, d, d_rcp; float mint, maxt; };
static inline Ray make_ray(V3 o, V3 d, float mint = RAY_EPS_F, float maxt = INF_F) {
    Ray r; r.o = o; r.d = d; r.d_rcp = v3(1.f / d.x, 1.f / d.y, 1.f / d.z); r.mint = mint; r.maxt = maxt;
    return r;
}
static inline V3 ray_at(const Ray &r, float t) { return fmadd3(r.d, t, r.o); }
/* Interaction::spawn_ray (include/mitsuba/render/interaction.h:58-61) */
static inline Ray spawn_ray(V3 p, V3 d) { return make_ray(p, d, (1.f + hmax_abs(p)) * RAY_EPS_F, INF_F); }

struct SI {
    bool valid; float t;
    V3 p, n, sh_n, sh_s, sh_t, dp_du, dp_dv, wi;
    V2 uv, duv_dx, duv_dy;
    int shape, prim;
    float b1, b2;                    // triangle barycentrics (mesh cache)
    float lx, ly;                    // rectangle local coords (rectangle cache)
};
static SI si_invalid() { SI s; memset(&s, 0, sizeof(s)); s.valid = false; s.t = INF_F; s.shape = -1; s.prim = -1; return s; }

static inline V3 xform_p(const float *M, V3 p) {
    return v3(((M[0] * p.x + M[1] * p.y) + M[2] * p.z) + M[3],
              ((M[4] * p.x + M[5] * p.y) + M[6] * p.z) + M[7],
              ((M[8] * p.x + M[9] * p.y) + M[10] * p.z) + M[11]);
}
static inline V3 xform_v(const float *M, V3 v) {
    return v3((M[0] * v.x + M[1] * v.y) + M[2] * v.z,
              (M[4] * v.x + M[5] * v.y) + M[6] * v.z,
              (M[8] * v.x + M[9] * v.y) + M[10] * v.z);
}

/* coordinate_system, include/mitsuba/core/vector.h:116-136 (Duff et al. 2017) */
static inline void coordinate_system(V3 n, V3 *s, V3 *t) {
    float sign = signbit(n.z) ? -1.f : 1.f;
    float a = -(1.f / (sign + n.z));
    float b = n.x * n.y * a;
    *s = v3(mulsign(n.x * n.x * a, n.z) + 1.f, mulsign(b, n.z), -mulsign(n.x, n.z));
    *t = v3(b, sign + n.y * n.y * a, -n.y);
}

/* hmean of an RGB spectrum whose three channels hold the same value */
static inline float hmean_same(float e) { return ((e + e) + e) / 3.f; }

/* ---------------------------------------------------------------- scene setup */
static void invert_affine(const float *M, float *Minv) {
    double a[3][3], inv[3][3];
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) a[i][j] = M[i * 4 + j];
    double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0])
               + a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
    double id = 1.0 / det;
    inv[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) * id;
    inv[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) * id;
    inv[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) * id;
    inv[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) * id;
    inv[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) * id;
    inv[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) * id;
    inv[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) * id;
    inv[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) * id;
    inv[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) * id;
    double t[3] = {M[3], M[7], M[11]};
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) Minv[i * 4 + j] = (float) inv[i][j];
        Minv[i * 4 + 3] = (float) (-(inv[i][0] * t[0] + inv[i][1] * t[1] + inv[i][2] * t[2]));
    }
}

struct BuildItem { float c[3]; float bmin[3], bmax[3]; int tri; };

static void bvh_build(Scene &S) {
    int nt = (int) S.tri_shape.size();
    S.nodes.clear(); S.tri_order.clear();
    if (nt == 0) return;
    std::vector<BuildItem> items(nt);
    for (int i = 0; i < nt; ++i) {
        BuildItem &it = items[i]; it.tri = i;
        for (int a = 0; a < 3; ++a) { it.bmin[a] = INF_F; it.bmax[a] = -INF_F; }
        for (int k = 0; k < 3; ++k) {
            const float *p = &S.pos[3 * S.faces[3 * i + k]];
            for (int a = 0; a < 3; ++a) { it.bmin[a] = std::min(it.bmin[a], p[a]); it.bmax[a] = std::max(it.bmax[a], p[a]); }
        }
        for (int a = 0; a < 3; ++a) it.c[a] = 0.5f * (it.bmin[a] + it.bmax[a]);
    }
    struct Task { int node, lo, hi; };
    std::vector<Task> stack;
    S.nodes.push_back(Node());
    stack.push_back({0, 0, nt});
    while (!stack.empty()) {
        Task tk = stack.back(); stack.pop_back();
        Node nd; nd.left = nd.right = -1; nd.first = tk.lo; nd.count = tk.hi - tk.lo;
        float cmin[3], cmax[3];
        for (int a = 0; a < 3; ++a) { nd.bmin[a] = INF_F; nd.bmax[a] = -INF_F; cmin[a] = INF_F; cmax[a] = -INF_F; }
        for (int i = tk.lo; i < tk.hi; ++i)
            for (int a = 0; a < 3; ++a) {
                nd.bmin[a] = std::min(nd.bmin[a], items[i].bmin[a]); nd.bmax[a] = std::max(nd.bmax[a], items[i].bmax[a]);
                cmin[a] = std::min(cmin[a], items[i].c[a]); cmax[a] = std::max(cmax[a], items[i].c[a]);
            }
        // pad so that the slab test can never reject a triangle the triangle test accepts
        for (int a = 0; a < 3; ++a) {
            float ext = nd.bmax[a] - nd.bmin[a];
            float pad = 1e-5f * ext + 1e-5f * std::max(fabsf(nd.bmin[a]), fabsf(nd.bmax[a])) + 1e-7f;
            nd.bmin[a] -= pad; nd.bmax[a] += pad;
        }
        if (nd.count <= 4) { S.nodes[tk.node] = nd; continue; }
        int axis = 0; float best = cmax[0] - cmin[0];
        for (int a = 1; a < 3; ++a) if (cmax[a] - cmin[a] > best) { best = cmax[a] - cmin[a]; axis = a; }
        int mid = (tk.lo + tk.hi) / 2;
        std::nth_element(items.begin() + tk.lo, items.begin() + mid, items.begin() + tk.hi,
                         [axis](const BuildItem &x, const BuildItem &y) { return x.c[axis] < y.c[axis]; });
        nd.left = (int) S.nodes.size(); S.nodes.push_back(Node());
        nd.right = (int) S.nodes.size(); S.nodes.push_back(Node());
        nd.count = 0;
        S.nodes[tk.node] = nd;
        stack.push_back({nd.left, tk.lo, mid});
        stack.push_back({nd.right, mid, tk.hi});
    }
    S.tri_order.resize(nt);
    for (int i = 0; i < nt; ++i) S.tri_order[i] = items[i].tri;
}

static void camera_setup(Camera &c, const sms_camera_desc &d) {
    double o[3] = {d.origin[0], d.origin[1], d.origin[2]};
    double tg[3] = {d.target[0], d.target[1], d.target[2]};
    double up[3] = {d.up[0], d.up[1], d.up[2]};
    double dir[3] = {tg[0] - o[0], tg[1] - o[1], tg[2] - o[2]};
    double l = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    for (int i = 0; i < 3; ++i) dir[i] /= l;
    double left[3] = {up[1] * dir[2] - up[2] * dir[1], up[2] * dir[0] - up[0] * dir[2], up[0] * dir[1] - up[1] * dir[0]};
    l = sqrt(left[0] * left[0] + left[1] * left[1] + left[2] * left[2]);
    for (int i = 0; i < 3; ++i) left[i] /= l;
    double nup[3] = {dir[1] * left[2] - dir[2] * left[1], dir[2] * left[0] - dir[0] * left[2], dir[0] * left[1] - dir[1] * left[0]};
    c.o = v3((float) o[0], (float) o[1], (float) o[2]);
    c.left = v3((float) left[0], (float) left[1], (float) left[2]);
    c.up = v3((float) nup[0], (float) nup[1], (float) nup[2]);
    c.dir = v3((float) dir[0], (float) dir[1], (float) dir[2]);
    double th = tan(0.5 * d.fov_deg * M_PI / 180.0);
    double aspect = (double) d.width / (double) d.height;
    c.tan_half = (float) th; c.aspect = (float) aspect;
    c.dx = v3((float) (-2.0 * th / d.width), 0.f, 0.f);
    c.dy = v3(0.f, (float) (-2.0 * th / aspect / d.height), 0.f);
    c.rect_x = (float) th; c.rect_y = (float) (th / aspect);
    c.normalization = (float) (1.0 / ((2.0 * th) * (2.0 * th / aspect)));
    c.w = d.width; c.h = d.height;
}

/* LEAN map (normalmap.cpp:36-87): level 0 = (s_x, s_y, s_x^2 + b, s_y^2 + b, s_x s_y) with the
   base Beckmann variance b = 0.5 * 1e-4^2; each coarser level is the box-filter resample to half
   resolution (Bitmap::resample, bitmap.cpp:234-312, rfilter.h:107-203): two taps of weight 0.5
   per axis, first along x, then along y. */
static void build_lean(Normalmap &nm) {
    const int res = nm.res;
    const float base_alpha = 0.0001f, base_sigma_2 = 0.5f * base_alpha * base_alpha;
    std::vector<float> l0((size_t) res * res * 5);
    for (size_t s = 0; s < (size_t) res * res; ++s) {
        float sx = nm.slope[2 * s], sy = nm.slope[2 * s + 1];
        float *d = &l0[5 * s];
        d[0] = sx; d[1] = sy; d[2] = sx * sx + base_sigma_2; d[3] = sy * sy + base_sigma_2; d[4] = sx * sy;
    }
    nm.lean.clear();
    nm.lean.push_back(std::move(l0));
    for (int r = res; r > 1; r /= 2) {
        const std::vector<float> &src = nm.lean.back();
        const int h = r / 2;
        std::vector<float> tmp((size_t) h * r * 5), dst((size_t) h * h * 5);
        for (int y = 0; y < r; ++y)
            for (int x = 0; x < h; ++x)
                for (int c = 0; c < 5; ++c)
                    tmp[((size_t) y * h + x) * 5 + c] = src[((size_t) y * r + 2 * x) * 5 + c] * 0.5f +
                                                        src[((size_t) y * r + 2 * x + 1) * 5 + c] * 0.5f;
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < h; ++x)
                for (int c = 0; c < 5; ++c)
                    dst[((size_t) y * h + x) * 5 + c] = tmp[((size_t) (2 * y) * h + x) * 5 + c] * 0.5f +
                                                        tmp[((size_t) (2 * y + 1) * h + x) * 5 + c] * 0.5f;
        nm.lean.push_back(std::move(dst));
    }
}

static Scene *scene_create(const sms_scene_desc *desc, const sms_config *cfg) {
    Scene *S = new Scene();
    S->cfg = *cfg;
    for (int i = 0; i < desc->n_normalmaps; ++i) {
        const sms_normalmap_desc &nd = desc->normalmaps[i];
        Normalmap nm; nm.res = nd.res;
        size_t N = (size_t) nd.res * nd.res;
        nm.slope.resize(2 * N);
        /* normalmap.cpp:46-66: rgb = byte/255, n = normalize(2 rgb - 1), slope = -n.xy/n.z */
        for (size_t s = 0; s < N; ++s) {
            /* an OpenEXR map (rgb32f) is used as stored (normalmap.cpp:20-22) */
            float r = nd.rgb ? (float) nd.rgb[3 * s + 0] / 255.f : (float) nd.rgb32f[3 * s + 0];
            float g = nd.rgb ? (float) nd.rgb[3 * s + 1] / 255.f : (float) nd.rgb32f[3 * s + 1];
            float b = nd.rgb ? (float) nd.rgb[3 * s + 2] / 255.f : (float) nd.rgb32f[3 * s + 2];
            V3 n = normalize(v3(2.f * r - 1.f, 2.f * g - 1.f, 2.f * b - 1.f));
            nm.slope[2 * s + 0] = -n.x / n.z;
            nm.slope[2 * s + 1] = -n.y / n.z;
        }
        build_lean(nm);
        if (nd.encoded_normals) {
            nm.encoded = true;
            nm.rgb.resize(3 * N);
            for (size_t s = 0; s < 3 * N; ++s) nm.rgb[s] = nd.rgb ? (float) nd.rgb[s] / 255.f : (float) nd.rgb32f[s];
            /* the normals mipmap (normalmap.cpp:70-87): the same two-tap resample as the LEAN map */
            std::vector<float> cur = nm.rgb;
            for (int r = nd.res; r > 1; r /= 2) {
                const int h = r / 2;
                std::vector<float> tmp((size_t) h * r * 3), dst((size_t) h * h * 3);
                for (int y = 0; y < r; ++y)
                    for (int x = 0; x < h; ++x)
                        for (int c = 0; c < 3; ++c)
                            tmp[((size_t) y * h + x) * 3 + c] = cur[((size_t) y * r + 2 * x) * 3 + c] * 0.5f +
                                                                cur[((size_t) y * r + 2 * x + 1) * 3 + c] * 0.5f;
                for (int y = 0; y < h; ++y)
                    for (int x = 0; x < h; ++x)
                        for (int c = 0; c < 3; ++c)
                            dst[((size_t) y * h + x) * 3 + c] = tmp[((size_t) (2 * y) * h + x) * 3 + c] * 0.5f +
                                                                tmp[((size_t) (2 * y + 1) * h + x) * 3 + c] * 0.5f;
                cur = std::move(dst);
            }
            nm.top_rgb[0] = cur[0]; nm.top_rgb[1] = cur[1]; nm.top_rgb[2] = cur[2];
        }
        S->nmaps.push_back(nm);
    }
    for (int i = 0; i < desc->n_emitters; ++i) {
        const sms_emitter_desc &e = desc->emitters[i];
        Emitter em; em.type = e.type; em.shape = e.shape;
        em.radiance = v3(e.radiance[0], e.radiance[1], e.radiance[2]);
        em.position = v3(e.position[0], e.position[1], e.position[2]);
        em.intensity = v3(e.intensity[0], e.intensity[1], e.intensity[2]);
        em.direction = v3(e.direction[0], e.direction[1], e.direction[2]);
        em.irradiance = v3(e.irradiance[0], e.irradiance[1], e.irradiance[2]);
        em.ss = e.caustic_ss != 0; em.ms = e.caustic_ms != 0;
        S->emitters.push_back(em);
        S->emitters_all.push_back(i);
        if (em.ss) S->emitters_ss.push_back(i);
    }
    for (int i = 0; i < desc->n_shapes; ++i) {
        const sms_shape_desc &d = desc->shapes[i];
        Shape sh; memset(&sh.center, 0, sizeof(V3));
        sh.type = d.type; sh.bsdf = d.bsdf; sh.flags = d.flags; sh.emitter = d.emitter;
        sh.center = v3(d.center[0], d.center[1], d.center[2]);
        sh.radius = d.radius; sh.inv_radius = 1.f / d.radius;
        memcpy(sh.M, d.to_world, sizeof(sh.M));
        /* BSDF::ior(): diffuse (0,1) (bsdf.cpp:59-62), dielectric (eta,0), conductor (eta,k) */
        if (d.bsdf == SMS_BSDF_DIFFUSE) { sh.eta_r = 0.f; sh.k = 1.f; }
        else if (d.bsdf == SMS_BSDF_DIELECTRIC) { sh.eta_r = d.eta; sh.k = 0.f; }
        else { sh.eta_r = d.eta; sh.k = d.k; }
        sh.refl = v3(d.reflectance[0], d.reflectance[1], d.reflectance[2]);
        sh.normalmap = d.normalmap; sh.tiles = d.tiles;
        sh.vbase = sh.fbase = sh.nv = sh.nf = 0; sh.has_n = sh.has_uv = false;
        sh.area_sum = sh.area_norm = 0.f;
        /* ManifoldVertex eta: dielectric real IOR, conductors (k != 0) and others 1
           (manifold.h:93-95; bsdf.cpp:59-62 default ior (0,1)) */
        sh.eta_vtx = sh.k == 0.f ? hmean_same(sh.eta_r) : (float) 1.f;
        sh.alpha = d.bsdf == SMS_BSDF_CONDUCTOR ? d.roughness : (float) 0.f;
        sh.bbox_center = sh.center;
        if (d.type == SMS_SHAPE_SPHERE) {
            sh.inv_surface_area = 1.f / (4.f * PI_F * sh.radius * sh.radius);
        } else if (d.type == SMS_SHAPE_RECTANGLE) {
            invert_affine(sh.M, sh.Minv);
            /* rectangle.cpp:86-90: dp_du = M (2,0,0), dp_dv = M (0,2,0), n = normalize(M^-T (0,0,1)) */
            sh.fr_s = v3(2.f * sh.M[0], 2.f * sh.M[4], 2.f * sh.M[8]);
            sh.fr_t = v3(2.f * sh.M[1], 2.f * sh.M[5], 2.f * sh.M[9]);
            double nx = sh.Minv[8], ny = sh.Minv[9], nz = sh.Minv[10];
            double l = sqrt(nx * nx + ny * ny + nz * nz);
            sh.fr_n = v3((float) (nx / l), (float) (ny / l), (float) (nz / l));
            sh.inv_surface_area = 1.f / norm(cross(sh.fr_s, sh.fr_t));
            sh.bbox_center = v3(sh.M[3], sh.M[7], sh.M[11]);
        } else if (d.type == SMS_SHAPE_MESH) {
            sh.vbase = (int) (S->pos.size() / 3); sh.fbase = (int) (S->faces.size() / 3);
            sh.nv = d.n_vertices; sh.nf = d.n_faces;
            sh.has_n = d.normals != nullptr; sh.has_uv = d.texcoords != nullptr;
            for (int v = 0; v < d.n_vertices; ++v) {
                for (int a = 0; a < 3; ++a) S->pos.push_back(d.positions[3 * v + a]);
                for (int a = 0; a < 3; ++a) S->nrm.push_back(d.normals ? d.normals[3 * v + a] : (float) 0.f);
                for (int a = 0; a < 2; ++a) S->uvs.push_back(d.texcoords ? d.texcoords[2 * v + a] : (float) 0.f);
            }
            /* area distribution: mesh.cpp area_distr_ensure + DiscreteDistribution::update
               (include/mitsuba/core/distr_1d.h: double running sum, float cdf) */
            sh.pmf.resize(d.n_faces); sh.cdf.resize(d.n_faces);
            double sum = 0.0;
            for (int f = 0; f < d.n_faces; ++f) {
                uint32_t i0 = d.faces[3 * f], i1 = d.faces[3 * f + 1], i2 = d.faces[3 * f + 2];
                for (int k = 0; k < 3; ++k) S->faces.push_back((uint32_t) sh.vbase + d.faces[3 * f + k]);
                S->tri_shape.push_back(i); S->tri_prim.push_back(f);
                V3 p0 = v3(d.positions[3 * i0], d.positions[3 * i0 + 1], d.positions[3 * i0 + 2]);
                V3 p1 = v3(d.positions[3 * i1], d.positions[3 * i1 + 1], d.positions[3 * i1 + 2]);
                V3 p2 = v3(d.positions[3 * i2], d.positions[3 * i2 + 1], d.positions[3 * i2 + 2]);
                float a = .5f * norm(cross(p1 - p0, p2 - p0));
                sh.pmf[f] = a;
                sum += (double) a;
                sh.cdf[f] = (float) sum;
            }
            sh.area_sum = (float) sum; sh.area_norm = (float) (1.0 / sum);
            float lo[3] = {INF_F, INF_F, INF_F}, hi[3] = {-INF_F, -INF_F, -INF_F};
            for (int v = 0; v < d.n_vertices; ++v)
                for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], d.positions[3 * v + a]); hi[a] = std::max(hi[a], d.positions[3 * v + a]); }
            sh.bbox_center = v3(0.5f * (lo[0] + hi[0]), 0.5f * (lo[1] + hi[1]), 0.5f * (lo[2] + hi[2]));
            sh.inv_surface_area = sh.area_norm;
        }
        S->shapes.push_back(sh);
        if (d.flags & SMS_FLAG_CAUSTIC_CASTER_SS) S->casters_ss.push_back(i);
    }
    bvh_build(*S);
    /* scene bounding box over mesh vertices, spheres and rectangle corners (inputs of the
       learned proposal are normalised to it) */
    float lo[3] = {INF_F, INF_F, INF_F}, hi[3] = {-INF_F, -INF_F, -INF_F};
    auto grow = [&](float x, float y, float z) {
        const float v[3] = {x, y, z};
        for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], v[a]); hi[a] = std::max(hi[a], v[a]); }
    };
    for (size_t v = 0; v < S->pos.size() / 3; ++v) grow(S->pos[3 * v], S->pos[3 * v + 1], S->pos[3 * v + 2]);
    for (const Shape &sh : S->shapes) {
        if (sh.type == SMS_SHAPE_SPHERE) {
            grow(sh.center.x - sh.radius, sh.center.y - sh.radius, sh.center.z - sh.radius);
            grow(sh.center.x + sh.radius, sh.center.y + sh.radius, sh.center.z + sh.radius);
        } else if (sh.type == SMS_SHAPE_RECTANGLE) {
            for (int cy = -1; cy <= 1; cy += 2)
                for (int cx = -1; cx <= 1; cx += 2)
                    grow(sh.M[0] * cx + sh.M[1] * cy + sh.M[3], sh.M[4] * cx + sh.M[5] * cy + sh.M[7],
                         sh.M[8] * cx + sh.M[9] * cy + sh.M[11]);
        }
    }
    float half = 0.f;
    for (int a = 0; a < 3; ++a) { S->box_c[a] = 0.5f * (lo[a] + hi[a]); half = std::max(half, 0.5f * (hi[a] - lo[a])); }
    S->box_inv = half > 0.f ? 1.f / half : (float) 1.f;
    /* BoundingBox::bounding_sphere of the scene (radius = |max - centre|), enlarged by
       DirectionalEmitter::set_scene (src/emitters/directional.cpp:63-68); sample_direction places
       the emitter 2 radii away (:112) */
    {
        double r2 = 0.0;
        for (int a = 0; a < 3; ++a) {
            double c = 0.5 * ((double) lo[a] + (double) hi[a]), d = (double) hi[a] - c;
            r2 += d * d;
        }
        double r = sqrt(r2) * (1.0 + (double) RAY_EPS_F);
        if (!(r > (double) RAY_EPS_F)) r = (double) RAY_EPS_F;
        S->bs_dist = (float) (2.0 * r);
    }
    S->has_cam = desc->camera.width > 0;
    if (S->has_cam) camera_setup(S->cam, desc->camera);
    return S;
}

/* ---------------------------------------------------------------- intersection */
/* Sphere::ray_intersect, FP64 (src/shapes/sphere.cpp:276-303) with math::solve_quadratic
   (include/mitsuba/core/math.h:371-411). */
static bool sphere_intersect(const Shape &sh, const Ray &ray, float *t_out) {
    double mint = (double) ray.mint, maxt = (double) ray.maxt;
    double ox = (double) ray.o.x - (double) sh.center.x, oy = (double) ray.o.y - (double) sh.center.y,
           oz = (double) ray.o.z - (double) sh.center.z;
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
        x0 = std::min(x0p, x1p); x1 = std::max(x0p, x1p);
    }
    bool out_bounds = !(x0 <= maxt && x1 >= mint);
    bool in_bounds = x0 < mint && x1 > maxt;
    if (out_bounds || in_bounds) return false;
    *t_out = (float) (x0 < mint ? x1 : x0);
    return true;
}

/* Rectangle::ray_intersect (src/shapes/rectangle.cpp:141-163) */
static bool rect_intersect(const Shape &sh, const Ray &ray, float *t_out, float *lx, float *ly) {
    V3 o = xform_p(sh.Minv, ray.o), d = xform_v(sh.Minv, ray.d);
    float t = -o.z * (1.f / d.z);
    V3 local = fmadd3(d, t, o);
    if (!(t >= ray.mint && t <= ray.maxt && fabsf(local.x) <= 1.f && fabsf(local.y) <= 1.f)) return false;
    *t_out = t; *lx = local.x; *ly = local.y;
    return true;
}

/* Mesh::ray_intersect_triangle (include/mitsuba/render/mesh.h:202-230) */
static bool tri_intersect(const Scene &S, int tri, const Ray &ray, float *t_out, float *u_out, float *v_out) {
    const uint32_t *f = &S.faces[3 * tri];
    V3 p0 = v3(S.pos[3 * f[0]], S.pos[3 * f[0] + 1], S.pos[3 * f[0] + 2]);
    V3 p1 = v3(S.pos[3 * f[1]], S.pos[3 * f[1] + 1], S.pos[3 * f[1] + 2]);
    V3 p2 = v3(S.pos[3 * f[2]], S.pos[3 * f[2] + 1], S.pos[3 * f[2] + 2]);
    V3 e1 = p1 - p0, e2 = p2 - p0;
    V3 pvec = cross(ray.d, e2);
    float inv_det = 1.f / dot(e1, pvec);
    V3 tvec = ray.o - p0;
    float u = dot(tvec, pvec) * inv_det;
    if (!(u >= 0.f && u <= 1.f)) return false;
    V3 qvec = cross(tvec, e1);
    float v = dot(ray.d, qvec) * inv_det;
    if (!(v >= 0.f && u + v <= 1.f)) return false;
    float t = dot(e2, qvec) * inv_det;
    if (!(t >= ray.mint && t <= ray.maxt)) return false;
    *t_out = t; *u_out = u; *v_out = v;
    return true;
}

static inline bool box_hit(const Node &n, const Ray &r, float tmax) {
    float t0 = r.mint, t1 = tmax;
    const float o[3] = {r.o.x, r.o.y, r.o.z}, ir[3] = {r.d_rcp.x, r.d_rcp.y, r.d_rcp.z};
    for (int a = 0; a < 3; ++a) {
        float ta = (n.bmin[a] - o[a]) * ir[a], tb = (n.bmax[a] - o[a]) * ir[a];
        if (ta != ta) ta = -INF_F;           // 0 * inf inside the slab
        if (tb != tb) tb = INF_F;
        float lo = std::min(ta, tb), hi = std::max(ta, tb);
        t0 = std::max(t0, lo); t1 = std::min(t1, hi);
    }
    return t0 <= t1 * (1.f + 4.f * EPS_F);
}

static inline bool better(float t, int shape, int prim, float bt, int bshape, int bprim) {
    if (t < bt) return true;
    if (t > bt) return false;
    if (shape != bshape) return shape < bshape;
    return prim < bprim;
}

struct Hit { bool hit; float t; int shape, prim; float b1, b2, lx, ly; };

/* Scene::ray_intersect closest hit over every shape (scene_native.inl:17-34) */
static Hit scene_hit(const Scene &S, const Ray &ray, bool any_hit) {
    Hit h; h.hit = false; h.t = INF_F; h.shape = 1 << 30; h.prim = 1 << 30; h.b1 = h.b2 = h.lx = h.ly = 0.f;
    for (int i = 0; i < (int) S.shapes.size(); ++i) {
        const Shape &sh = S.shapes[i];
        float t, lx = 0.f, ly = 0.f;
        bool ok = false;
        Ray r = ray; r.maxt = std::min(ray.maxt, h.t);
        if (sh.type == SMS_SHAPE_SPHERE) ok = sphere_intersect(sh, r, &t);
        else if (sh.type == SMS_SHAPE_RECTANGLE) ok = rect_intersect(sh, r, &t, &lx, &ly);
        if (ok && better(t, i, 0, h.t, h.shape, h.prim)) {
            h.hit = true; h.t = t; h.shape = i; h.prim = 0; h.lx = lx; h.ly = ly;
            if (any_hit) return h;
        }
    }
    /* a ray with a non-finite origin or direction hits no triangle (tri_intersect rejects it)
       but passes every box test: skip the traversal, the result is the same */
    const bool finite = std::isfinite(ray.o.x) && std::isfinite(ray.o.y) && std::isfinite(ray.o.z) &&
                        std::isfinite(ray.d.x) && std::isfinite(ray.d.y) && std::isfinite(ray.d.z);
    if (!S.nodes.empty() && finite) {
        int stack[128]; int sp = 0; stack[sp++] = 0;
        while (sp > 0) {
            const Node &n = S.nodes[stack[--sp]];
            if (!box_hit(n, ray, std::min(ray.maxt, h.t))) continue;
            if (n.left < 0) {
                for (int k = n.first; k < n.first + n.count; ++k) {
                    int tri = S.tri_order[k];
                    float t, u, v;
                    Ray r = ray; r.maxt = std::min(ray.maxt, h.t);
                    if (tri_intersect(S, tri, r, &t, &u, &v) &&
                        better(t, S.tri_shape[tri], S.tri_prim[tri], h.t, h.shape, h.prim)) {
                        h.hit = true; h.t = t; h.shape = S.tri_shape[tri]; h.prim = S.tri_prim[tri]; h.b1 = u; h.b2 = v;
                        if (any_hit) return h;
                    }
                }
            } else {
                stack[sp++] = n.left; stack[sp++] = n.right;
            }
        }
    }
    return h;
}

/* Shape::fill_surface_interaction for the three shape kinds + the kd-tree's shading frame
   (include/mitsuba/render/kdtree.h:2335-2364) */
static SI make_si(const Scene &S, const Ray &ray, const Hit &h) {
    if (!h.hit) return si_invalid();
    SI si = si_invalid();
    si.valid = true; si.t = h.t; si.shape = h.shape; si.prim = h.prim;
    const Shape &sh = S.shapes[h.shape];
    if (sh.type == SMS_SHAPE_SPHERE) {
        /* sphere.cpp:329-374 */
        V3 n = normalize(ray_at(ray, si.t) - sh.center);
        si.sh_n = n;
        si.p = fmadd3(n, sh.radius, sh.center);
        V3 local = (si.p - sh.center) * sh.inv_radius;
        float rd_2 = local.x * local.x + local.y * local.y;
        si.uv = v2(0.f, 0.f);
        if (sh.normalmap >= 0) {
            /* sphere.cpp:343-348 (read only through a normal map); theta by unit_angle_z of
               Mitsuba 2's core vector.h: 2 asin(|v - sign(v.z) z| / 2), mirrored for v.z < 0 */
            float dz = local.z - (signbit(local.z) ? -1.f : 1.f);
            float temp = 2.f * c_asinf(.5f * sqrtf((local.x * local.x + local.y * local.y) + dz * dz));
            float theta = local.z >= 0.f ? temp : PI_F - temp;
            float phi = c_atan2f(local.y, local.x);
            if (phi < 0.f) phi = phi + 2.f * PI_F;
            si.uv = v2(phi * INV_TWOPI_F, theta * INV_PI_F);
        }
        si.dp_du = v3(-local.y, local.x, 0.f);
        float rd = sqrtf(rd_2), inv_rd = 1.f / rd;
        float cos_phi = local.x * inv_rd, sin_phi = local.y * inv_rd;
        si.dp_dv = v3(local.z * cos_phi, local.z * sin_phi, -rd);
        if (rd == 0.f) si.dp_dv = v3(1.f, 0.f, 0.f);
        si.dp_du = (si.dp_du * sh.radius) * (2.f * PI_F);
        si.dp_dv = (si.dp_dv * sh.radius) * PI_F;
        si.n = si.sh_n;
    } else if (sh.type == SMS_SHAPE_RECTANGLE) {
        /* rectangle.cpp:175-202 */
        si.n = sh.fr_n; si.sh_n = sh.fr_n;
        si.dp_du = sh.fr_s; si.dp_dv = sh.fr_t;
        si.p = ray_at(ray, si.t);
        si.uv = v2(fmaf(h.lx, .5f, .5f), fmaf(h.ly, .5f, .5f));
        si.lx = h.lx; si.ly = h.ly;
    } else {
        /* mesh.cpp:620-691 */
        float b1 = h.b1, b2 = h.b2, b0 = 1.f - b1 - b2;
        si.b1 = b1; si.b2 = b2;
        int tri = sh.fbase + h.prim;
        const uint32_t *f = &S.faces[3 * tri];
        V3 p0 = v3(S.pos[3 * f[0]], S.pos[3 * f[0] + 1], S.pos[3 * f[0] + 2]);
        V3 p1 = v3(S.pos[3 * f[1]], S.pos[3 * f[1] + 1], S.pos[3 * f[1] + 2]);
        V3 p2 = v3(S.pos[3 * f[2]], S.pos[3 * f[2] + 1], S.pos[3 * f[2] + 2]);
        V3 dp0 = p1 - p0, dp1 = p2 - p0;
        si.p = (p0 * b0 + p1 * b1) + p2 * b2;
        V3 n = normalize(cross(dp0, dp1));
        si.n = n;
        V3 dpdu, dpdv; coordinate_system(n, &dpdu, &dpdv);
        V2 uv = v2(b1, b2);
        if (sh.has_uv) {
            V2 uv0 = v2(S.uvs[2 * f[0]], S.uvs[2 * f[0] + 1]), uv1 = v2(S.uvs[2 * f[1]], S.uvs[2 * f[1] + 1]),
               uv2 = v2(S.uvs[2 * f[2]], S.uvs[2 * f[2] + 1]);
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
            V3 n0 = v3(S.nrm[3 * f[0]], S.nrm[3 * f[0] + 1], S.nrm[3 * f[0] + 2]);
            V3 n1 = v3(S.nrm[3 * f[1]], S.nrm[3 * f[1] + 1], S.nrm[3 * f[1] + 2]);
            V3 n2 = v3(S.nrm[3 * f[2]], S.nrm[3 * f[2] + 1], S.nrm[3 * f[2] + 2]);
            n = normalize((n0 * b0 + n1 * b1) + n2 * b2);
        }
        si.sh_n = n;
        si.dp_du = dpdu; si.dp_dv = dpdv;
    }
    si.sh_s = normalize(fnmadd3(si.sh_n, dot(si.sh_n, si.dp_du), si.dp_du));
    si.sh_t = cross(si.sh_n, si.sh_s);
    V3 mw = -ray.d;
    si.wi = v3(dot(mw, si.sh_s), dot(mw, si.sh_t), dot(mw, si.sh_n));
    si.duv_dx = si.duv_dy = v2(0.f, 0.f);
    return si;
}

static SI ray_intersect(const Scene &S, const Ray &ray) { return make_si(S, ray, scene_hit(S, ray, false)); }
static bool ray_test(const Scene &S, const Ray &ray) { return scene_hit(S, ray, true).hit; }

/* ---------------------------------------------------------------- normal derivatives */
static void normal_derivative(const Scene &S, const SI &si, V3 *dn_du, V3 *dn_dv) {
    const Shape &sh = S.shapes[si.shape];
    if (sh.type == SMS_SHAPE_SPHERE) {
        /* sphere.cpp:376-382 */
        *dn_du = si.dp_du * sh.inv_radius; *dn_dv = si.dp_dv * sh.inv_radius;
    } else if (sh.type == SMS_SHAPE_RECTANGLE) {
        *dn_du = v3(0.f, 0.f, 0.f); *dn_dv = v3(0.f, 0.f, 0.f);
    } else {
        /* mesh.cpp:694-760 */
        if (!sh.has_n) { *dn_du = v3(0.f, 0.f, 0.f); *dn_dv = v3(0.f, 0.f, 0.f); return; }
        int tri = sh.fbase + si.prim;
        const uint32_t *f = &S.faces[3 * tri];
        V3 p0 = v3(S.pos[3 * f[0]], S.pos[3 * f[0] + 1], S.pos[3 * f[0] + 2]);
        V3 p1 = v3(S.pos[3 * f[1]], S.pos[3 * f[1] + 1], S.pos[3 * f[1] + 2]);
        V3 p2 = v3(S.pos[3 * f[2]], S.pos[3 * f[2] + 1], S.pos[3 * f[2] + 2]);
        V3 n0 = v3(S.nrm[3 * f[0]], S.nrm[3 * f[0] + 1], S.nrm[3 * f[0] + 2]);
        V3 n1 = v3(S.nrm[3 * f[1]], S.nrm[3 * f[1] + 1], S.nrm[3 * f[1] + 2]);
        V3 n2 = v3(S.nrm[3 * f[2]], S.nrm[3 * f[2] + 1], S.nrm[3 * f[2] + 2]);
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
            V2 uv0 = v2(S.uvs[2 * f[0]], S.uvs[2 * f[0] + 1]), uv1 = v2(S.uvs[2 * f[1]], S.uvs[2 * f[1] + 1]),
               uv2 = v2(S.uvs[2 * f[2]], S.uvs[2 * f[2] + 1]);
            V2 duv1 = uv1 - uv0, duv2 = uv2 - uv0;
            inv_det = 1.f / (duv1.x * duv2.y - duv1.y * duv2.x);
            V3 a = (duv2.y * dndu - duv1.y * dndv) * inv_det, b = (-duv2.x * dndu + duv1.x * dndv) * inv_det;
            dndu = a; dndv = b;
        }
        *dn_du = dndu; *dn_dv = dndv;
    }
}

/* ---------------------------------------------------------------- normal map (level 0) */
struct Texel4 { V2 v00, v10, v01, v11; V2 w; };
/* Normalmap::eval_normal / eval_normal_derivatives index computation (normalmap.h:62-80,148-163) */
static void nm_index(int res, V2 uv, uint32_t idx[4], V2 *w) {
    float fres = (float) res;
    float duv = 1.f / fres;
    V2 pos = v2((uv.x - 0.5f * duv) * fres, (uv.y - 0.5f * duv) * fres);
    float hi = (float) (res - 1);
    pos.x = std::min(std::max(pos.x, 0.f), hi); pos.y = std::min(std::max(pos.y, 0.f), hi);
    uint32_t xmin = (uint32_t) floorf(pos.x), ymin = (uint32_t) floorf(pos.y);
    uint32_t xmax = xmin + 1, ymax = ymin + 1;
    uint32_t lim = (uint32_t) (res - 1);
    xmin = std::min(xmin, lim); ymin = std::min(ymin, lim); xmax = std::min(xmax, lim); ymax = std::min(ymax, lim);
    idx[0] = ymin * res + xmin; idx[1] = ymin * res + xmax; idx[2] = ymax * res + xmin; idx[3] = ymax * res + xmax;
    *w = v2(pos.x - (float) xmin, pos.y - (float) ymin);
}
static Texel4 nm_fetch(const Normalmap &nm, V2 uv) {
    uint32_t i[4];
    Texel4 t;
    nm_index(nm.res, uv, i, &t.w);
    const float *s = nm.slope.data();
    t.v00 = v2(s[2 * i[0]], s[2 * i[0] + 1]); t.v10 = v2(s[2 * i[1]], s[2 * i[1] + 1]);
    t.v01 = v2(s[2 * i[2]], s[2 * i[2] + 1]); t.v11 = v2(s[2 * i[3]], s[2 * i[3] + 1]);
    return t;
}
/* interpolated slope (v[0], v[1]) of eval_normal with use_slopes (normalmap.h:82-92) */
static V2 nm_slope(const Texel4 &t) {
    V2 w1 = t.w, w0 = v2(1.f - w1.x, 1.f - w1.y);
    V2 a = v2(fmaf(w0.x, t.v00.x, w1.x * t.v10.x), fmaf(w0.x, t.v00.y, w1.x * t.v10.y));
    V2 b = v2(fmaf(w0.x, t.v01.x, w1.x * t.v11.x), fmaf(w0.x, t.v01.y, w1.x * t.v11.y));
    return v2(fmaf(w0.y, a.x, w1.y * b.x), fmaf(w0.y, a.y, w1.y * b.y));
}
/* tmp_u, tmp_v of eval_normal_derivatives with use_slopes (normalmap.h:170-176) */
static void nm_slope_deriv(const Texel4 &t, float res, V2 *tu, V2 *tv) {
    V2 tmp = (t.v01 + t.v10) - t.v11;
    *tu = ((t.v10 + t.v00 * (t.w.y - 1.f)) - tmp * t.w.y) * res;
    *tv = ((t.v01 + t.v00 * (t.w.x - 1.f)) - tmp * t.w.x) * res;
}
/* the map's local normal (not normalised) at level 0 and its uv derivatives: with use_slopes
   n = (-s_x, -s_y, 1) (normalmap.h:82-92,166-179); without, n = 2 rgb - 1 of the bilinearly
   interpolated RGB and dn = 2 drgb (normalmap.h:93-103,180-190 with NormalmapBSDF::frame and
   frame_derivative, bsdfs/normalmap.cpp:226-258) */
static void nm_local_normal(const Normalmap &nm, V2 uv, V3 *nl, V3 *du, V3 *dv) {
    if (!nm.encoded) {
        Texel4 tx = nm_fetch(nm, uv);
        V2 sl = nm_slope(tx), tu, tv;
        nm_slope_deriv(tx, (float) nm.res, &tu, &tv);
        *nl = v3(-sl.x, -sl.y, 1.f);
        *du = v3(-tu.x, -tu.y, 0.f); *dv = v3(-tv.x, -tv.y, 0.f);
        return;
    }
    uint32_t i[4];
    V2 w;
    nm_index(nm.res, uv, i, &w);
    const float *c = nm.rgb.data();
    V3 v00 = v3(c[3 * i[0]], c[3 * i[0] + 1], c[3 * i[0] + 2]), v10 = v3(c[3 * i[1]], c[3 * i[1] + 1], c[3 * i[1] + 2]);
    V3 v01 = v3(c[3 * i[2]], c[3 * i[2] + 1], c[3 * i[2] + 2]), v11 = v3(c[3 * i[3]], c[3 * i[3] + 1], c[3 * i[3] + 2]);
    float w0x = 1.f - w.x, w0y = 1.f - w.y;
    V3 a = v3(fmaf(w0x, v00.x, w.x * v10.x), fmaf(w0x, v00.y, w.x * v10.y), fmaf(w0x, v00.z, w.x * v10.z));
    V3 b = v3(fmaf(w0x, v01.x, w.x * v11.x), fmaf(w0x, v01.y, w.x * v11.y), fmaf(w0x, v01.z, w.x * v11.z));
    V3 rgb = v3(fmaf(w0y, a.x, w.y * b.x), fmaf(w0y, a.y, w.y * b.y), fmaf(w0y, a.z, w.y * b.z));
    *nl = v3(2.f * rgb.x - 1.f, 2.f * rgb.y - 1.f, 2.f * rgb.z - 1.f);
    V3 tmp = (v01 + v10) - v11;
    *du = (((v10 + v00 * (w.y - 1.f)) - tmp * w.y) * (float) nm.res) * 2.f;
    *dv = (((v01 + v00 * (w.x - 1.f)) - tmp * w.x) * (float) nm.res) * 2.f;
}
static inline V2 tile_uv(V2 uv, float tiles) {
    uv = uv * tiles;
    return v2(uv.x - floorf(uv.x), uv.y - floorf(uv.y));
}

/* ---------------------------------------------------------------- ManifoldVertex */
struct MV {
    V3 p, dp_du, dp_dv, n, gn, dn_du, dn_dv, s, t, ds_du, ds_dv, dt_du, dt_dv;
    float eta; V2 uv; int shape; bool fixed_direction;
};
static MV mv_point(V3 p) {
    MV v; memset(&v, 0, sizeof(v)); v.p = p; v.eta = 1.f; v.shape = -1; v.fixed_direction = false; return v;
}

/* compute_shading_frame_derivative (include/mitsuba/core/frame.h:224-244) */
static void shading_frame_derivative(V3 n, V3 dp_du, V3 dn_du, V3 dn_dv, MV &v) {
    V3 s = dp_du - n * dot(n, dp_du);
    float inv_len_s = 1.f / norm(s);
    s = s * inv_len_s;
    v.ds_du = inv_len_s * (-dn_du * dot(n, dp_du) - n * dot(dn_du, dp_du));
    v.ds_dv = inv_len_s * (-dn_dv * dot(n, dp_du) - n * dot(dn_dv, dp_du));
    v.ds_du = v.ds_du - s * dot(v.ds_du, s);
    v.ds_dv = v.ds_dv - s * dot(v.ds_dv, s);
    v.dt_du = cross(dn_du, s) + cross(n, v.ds_du);
    v.dt_dv = cross(dn_dv, s) + cross(n, v.ds_dv);
    v.dn_du = dn_du; v.dn_dv = dn_dv;
}

static inline V3 to_world(V3 s, V3 t, V3 n, V3 v) { return (s * v.x + t * v.y) + n * v.z; }

/* ManifoldVertex(si, smoothing) (include/mitsuba/render/manifold.h:89-113) for smoothing 0 and
   1.  Smoothing 1 evaluates a normal map at its coarsest mip level (normalmap.h:236-268 with
   level L - 1): one texel, so the normal is the map's mean slope and its derivatives vanish. */
static MV mv_from_si(const Scene &S, const SI &si, bool smooth = false) {
    MV v = mv_point(si.p);
    v.dp_du = si.dp_du; v.dp_dv = si.dp_dv; v.gn = si.n; v.uv = si.uv; v.shape = si.shape;
    const Shape &sh = S.shapes[si.shape];
    v.eta = sh.eta_vtx;
    V3 bn = si.sh_n;
    V3 bs = normalize(fnmadd3(bn, dot(bn, si.dp_du), si.dp_du));
    V3 bt = cross(bn, bs);
    V3 sdn_du, sdn_dv;
    normal_derivative(S, si, &sdn_du, &sdn_dv);
    if (sh.bsdf != SMS_BSDF_NORMALMAP) {
        /* BSDF::frame / frame_derivative (src/librender/bsdf.cpp:17-31) */
        v.n = bn; v.s = bs; v.t = bt;
        shading_frame_derivative(bn, si.dp_du, sdn_du, sdn_dv, v);
    } else {
        /* NormalmapBSDF::frame / frame_derivative at smoothing 0 (src/bsdfs/normalmap.cpp:218-290) */
        const Normalmap &nm = S.nmaps[sh.normalmap];
        V3 nl, dl_du, dl_dv;
        if (smooth) {
            /* one texel: bilinear weights (0, 0), the derivatives of eval_normal_derivatives zero,
               negated in its slope branch (normalmap.h:175-179) */
            if (nm.encoded) {
                nl = v3(2.f * nm.top_rgb[0] - 1.f, 2.f * nm.top_rgb[1] - 1.f, 2.f * nm.top_rgb[2] - 1.f);
                dl_du = v3(0.f, 0.f, 0.f);
            } else {
                const std::vector<float> &top = nm.lean.back();
                nl = v3(-top[0], -top[1], 1.f);
                dl_du = v3(-0.f, -0.f, 0.f);
            }
            dl_dv = dl_du;
        } else {
            nm_local_normal(nm, tile_uv(si.uv, sh.tiles), &nl, &dl_du, &dl_dv);
        }
        V3 dn_du = dl_du * sh.tiles, dn_dv = dl_dv * sh.tiles;
        V3 rn = normalize(to_world(bs, bt, bn, nl));
        v.n = rn;
        v.s = normalize(si.dp_du - rn * dot(rn, si.dp_du));
        v.t = cross(rn, v.s);
        MV db; shading_frame_derivative(bn, si.dp_du, sdn_du, sdn_dv, db);
        V3 world_n = to_world(bs, bt, bn, nl);
        float inv_length_n = 1.f / norm(world_n);
        world_n = world_n * inv_length_n;
        V3 dnu = inv_length_n * (to_world(bs, bt, bn, dn_du) + to_world(db.ds_du, db.dt_du, db.dn_du, nl));
        V3 dnv = inv_length_n * (to_world(bs, bt, bn, dn_dv) + to_world(db.ds_dv, db.dt_dv, db.dn_dv, nl));
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

/* ManifoldVertex::make_orthonormal (manifold.h:115-126) */
static void make_orthonormal(MV &v) {
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

/* ---------------------------------------------------------------- specular transforms */
/* SpecularManifold::reflect / d_reflect / refract / d_refract (manifold.h:347-417) */
static inline V3 reflect(V3 w, V3 n) { return 2.f * dot(w, n) * n - w; }
static inline void d_reflect(V3 w, V3 dw_du, V3 dw_dv, V3 n, V3 dn_du, V3 dn_dv, V3 *du, V3 *dv) {
    float dot_w_n = dot(w, n), dot_dwdu_n = dot(dw_du, n), dot_dwdv_n = dot(dw_dv, n),
          dot_w_dndu = dot(w, dn_du), dot_w_dndv = dot(w, dn_dv);
    *du = 2.f * ((dot_dwdu_n + dot_w_dndu) * n + dot_w_n * dn_du) - dw_du;
    *dv = 2.f * ((dot_dwdv_n + dot_w_dndv) * n + dot_w_n * dn_dv) - dw_dv;
}
static inline bool refract(V3 w, V3 n_, float eta_, V3 *out) {
    V3 n = n_;
    float eta = 1.f / eta_;
    if (dot(w, n) < 0.f) { eta = 1.f / eta; n = n * -1.f; }
    float dot_w_n = dot(w, n);
    float root_term = 1.f - eta * eta * (1.f - dot_w_n * dot_w_n);
    if (root_term < 0.f) { *out = v3(0.f, 0.f, 0.f); return false; }
    *out = -eta * (w - dot_w_n * n) - n * sqrtf(root_term);
    return true;
}
static inline void d_refract(V3 w, V3 dw_du, V3 dw_dv, V3 n_, V3 dn_du_, V3 dn_dv_, float eta_, V3 *du, V3 *dv) {
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
    *du = a_u + b_u; *dv = a_v + b_v;
}
/* sphcoords / d_sphcoords (manifold.h:419-444) */
static inline void sphcoords(V3 w, float *theta, float *phi) {
    *theta = safe_acos(w.z);
    *phi = c_atan2f(w.y, w.x);
    if (*phi < 0.f) *phi += 2.f * PI_F;
}
static inline void d_sphcoords(V3 w, V3 dw_du, V3 dw_dv, float *dt_du, float *dp_du, float *dt_dv, float *dp_dv) {
    float d_acos = -(1.f / safe_sqrt(1.f - w.z * w.z));
    V2 d_theta = d_acos * v2(dw_du.z, dw_dv.z);
    float yx = w.y / w.x;
    float d_atan = 1.f / (1.f + yx * yx);
    V2 d_phi = (d_atan * v2(w.x * dw_du.y - w.y * dw_du.x, w.x * dw_dv.y - w.y * dw_dv.x)) * (1.f / (w.x * w.x));
    if (w.x == 0.f) d_phi = v2(0.f, 0.f);
    *dt_du = d_theta.x; *dp_du = d_phi.x; *dt_dv = d_theta.y; *dp_dv = d_phi.y;
}

/* ---------------------------------------------------------------- emitter interaction */
struct EI { V3 p, n, d, weight; float pdf; int kind, emitter; };
static inline bool ei_is_directional(const EI &e) { return e.kind == SMS_EI_DIRECTIONAL; }
static inline bool ei_is_delta(const EI &e) { return e.kind == SMS_EI_POINT || e.kind == SMS_EI_DIRECTIONAL; }

static EI ei_from_c(const sms_emitter_interaction &c) {
    EI e; e.p = v3(c.p[0], c.p[1], c.p[2]); e.n = v3(c.n[0], c.n[1], c.n[2]); e.d = v3(c.d[0], c.d[1], c.d[2]);
    e.weight = v3(c.weight[0], c.weight[1], c.weight[2]); e.pdf = c.pdf; e.kind = c.kind; e.emitter = c.emitter;
    return e;
}

/* warp::square_to_uniform_sphere (include/mitsuba/core/warp.h:262-267) */
static inline V3 square_to_uniform_sphere(V2 s) {
    float z = fmaf(-2.f, s.y, 1.f), r = safe_sqrt(fmaf(-z, z, 1.f));
    float sn, cs; c_sincosf((2.f * PI_F) * s.x, &sn, &cs);
    return v3(r * cs, r * sn, z);
}

struct PS { V3 p, n; float pdf; };
/* Shape::sample_position for sphere (sphere.cpp:153-171), rectangle (rectangle.cpp:114-126),
   mesh (mesh.cpp:542-585) */
static PS sample_position(const Scene &S, int shape, V2 sample) {
    const Shape &sh = S.shapes[shape];
    PS ps;
    if (sh.type == SMS_SHAPE_SPHERE) {
        V3 local = square_to_uniform_sphere(sample);
        ps.p = fmadd3(local, sh.radius, sh.center); ps.n = local; ps.pdf = sh.inv_surface_area;
    } else if (sh.type == SMS_SHAPE_RECTANGLE) {
        ps.p = xform_p(sh.M, v3(sample.x * 2.f - 1.f, sample.y * 2.f - 1.f, 0.f));
        ps.n = sh.fr_n; ps.pdf = sh.inv_surface_area;
    } else {
        float value = sample.y;
        float scaled = value * sh.area_sum;
        /* enoki::binary_search over [0, nf-1]: first index whose cdf is not < value */
        int lo = 0, hi = sh.nf - 1;
        while (lo < hi) { int mid = (lo + hi) / 2; if (sh.cdf[mid] < scaled) lo = mid + 1; else hi = mid; }
        int idx = lo;
        float pmf = sh.pmf[idx] * sh.area_norm;
        float cdf = idx > 0 ? sh.cdf[idx - 1] * sh.area_norm : (float) 0.f;
        float y = (value - cdf) / pmf;
        const uint32_t *f = &S.faces[3 * (sh.fbase + idx)];
        V3 p0 = v3(S.pos[3 * f[0]], S.pos[3 * f[0] + 1], S.pos[3 * f[0] + 2]);
        V3 p1 = v3(S.pos[3 * f[1]], S.pos[3 * f[1] + 1], S.pos[3 * f[1] + 2]);
        V3 p2 = v3(S.pos[3 * f[2]], S.pos[3 * f[2] + 1], S.pos[3 * f[2] + 2]);
        V3 e0 = p1 - p0, e1 = p2 - p0;
        float t = safe_sqrt(1.f - sample.x);
        V2 b = v2(1.f - t, t * y);
        ps.p = (p0 + e0 * b.x) + e1 * b.y;
        if (sh.has_n) {
            V3 n0 = v3(S.nrm[3 * f[0]], S.nrm[3 * f[0] + 1], S.nrm[3 * f[0] + 2]);
            V3 n1 = v3(S.nrm[3 * f[1]], S.nrm[3 * f[1] + 1], S.nrm[3 * f[1] + 2]);
            V3 n2 = v3(S.nrm[3 * f[2]], S.nrm[3 * f[2] + 1], S.nrm[3 * f[2] + 2]);
            ps.n = normalize((n0 * (1.f - b.x - b.y) + n1 * b.x) + n2 * b.y);
        } else ps.n = normalize(cross(e0, e1));
        ps.pdf = sh.area_norm;
    }
    return ps;
}

/* SpecularManifold::sample_emitter_interaction (manifold.h:203-260) */
static EI sample_emitter_interaction(const Scene &S, V3 si_p, const std::vector<int> &emitters, Pcg32 *rng) {
    EI ei; memset(&ei, 0, sizeof(ei)); ei.emitter = -1; ei.kind = SMS_EI_AREA;
    if (emitters.empty()) return ei;
    float emitter_sample = pcg32_float(rng);
    float emitter_pdf = 1.f / (float) emitters.size();
    uint32_t index = std::min((uint32_t) (emitter_sample * (float) emitters.size()), (uint32_t) emitters.size() - 1);
    const Emitter &em = S.emitters[emitters[index]];
    ei.emitter = emitters[index];
    V3 spec = v3(0.f, 0.f, 0.f);
    V2 s2; s2.x = pcg32_float(rng); s2.y = pcg32_float(rng);
    if (em.type == SMS_EMITTER_AREA) {
        ei.kind = SMS_EI_AREA;
        PS ps = sample_position(S, em.shape, s2);
        if (ps.pdf > 0.f) {
            spec = em.radiance / ps.pdf;        // area.cpp:71-78 with cos(wi=(0,0,1)) > 0
            ei.p = ps.p; ei.n = ps.n; ei.d = normalize(ps.p - si_p); ei.pdf = ps.pdf;
        }
    } else if (em.type == SMS_EMITTER_DIRECTIONAL) {
        /* DirectionalEmitter::sample_direction (src/emitters/directional.cpp:105-128) */
        ei.kind = SMS_EI_DIRECTIONAL;
        V3 d = em.direction;
        ei.p = si_p - d * S.bs_dist; ei.d = -d; ei.n = -d; ei.pdf = 1.f;
        spec = em.irradiance;
    } else {
        ei.kind = SMS_EI_POINT;
        /* PointLight::sample_direction (src/emitters/point.cpp:74-98) */
        V3 d = em.position - si_p;
        float dist = norm(d), inv_dist = 1.f / dist;
        d = d * inv_dist;
        ei.p = em.position; ei.d = d; ei.n = d; ei.pdf = 1.f;
        spec = em.intensity * (inv_dist * inv_dist);
        spec = spec * dist * dist;
    }
    ei.pdf *= emitter_pdf;
    ei.weight = spec * (1.f / emitter_pdf);
    return ei;
}

/* SpecularManifold::emitter_interaction_to_vertex (manifold.h:263-302) */
static bool emitter_interaction_to_vertex(const Scene &S, const EI &ei, V3 p, MV *vy) {
    if (ei.kind == SMS_EI_AREA) {
        V3 d_tmp = normalize(ei.p - p);
        Ray r = make_ray(p + SHADOW_EPS_F * d_tmp, d_tmp);
        SI si_y = ray_intersect(S, r);
        if (!si_y.valid) { *vy = mv_point(v3(0.f, 0.f, 0.f)); return false; }
        *vy = mv_from_si(S, si_y);
        make_orthonormal(*vy);
        return true;
    } else if (ei.kind == SMS_EI_DIRECTIONAL) {
        *vy = mv_point(ei.p);
        V3 d = normalize(ei.p - p);
        vy->p = p + d; vy->n = -d;
        coordinate_system(vy->n, &vy->dp_du, &vy->dp_dv);
        vy->fixed_direction = true;
        return true;
    } else {
        *vy = mv_point(ei.p);
        V3 d = normalize(p - ei.p);
        vy->n = d; vy->gn = d;
        coordinate_system(d, &vy->dp_du, &vy->dp_dv);
        return true;
    }
}

/* ---------------------------------------------------------------- Fresnel */
/* fresnel() (include/mitsuba/render/fresnel.h:34-75) */
static float fresnel(float cos_theta_i, float eta) {
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
/* fresnel_conductor (fresnel.h:96-122) */
static float fresnel_conductor(float cos_theta_i, float eta_r, float eta_i) {
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

/* ================================================================ single scatter */
struct Step { bool ok; V2 C, dX; };

/* compute_step_anglediff (src/librender/manifold_ss.cpp:494-636) */
static Step step_anglediff(V3 v0p, const MV &v1, const EI &v2e, V3 n_offset) {
    Step r; r.ok = false; r.C = v2(INF_F, INF_F); r.dX = v2(0.f, 0.f);
    V3 wi = v0p - v1.p;
    float ili = norm(wi);
    if (ili < 1e-3f) return r;
    ili = 1.f / ili;
    wi = wi * ili;
    V3 dwi_du = -ili * (v1.dp_du - wi * dot(wi, v1.dp_du));
    V3 dwi_dv = -ili * (v1.dp_dv - wi * dot(wi, v1.dp_dv));
    V3 wo = ei_is_directional(v2e) ? v2e.d : v2e.p - v1.p;
    float ilo = norm(wo);
    if (ilo < 1e-3f) return r;
    ilo = 1.f / ilo;
    wo = wo * ilo;
    V3 dwo_du, dwo_dv;
    if (ei_is_directional(v2e)) { dwo_du = v3(0.f, 0.f, 0.f); dwo_dv = v3(0.f, 0.f, 0.f); }
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
    V3 wio; bool valid_i;
    if (v1.eta == 1.f) { wio = reflect(wi, n); valid_i = true; } else valid_i = refract(wi, n, v1.eta, &wio);
    if (valid_i) {
        V3 dwio_du, dwio_dv;
        if (v1.eta == 1.f) d_reflect(wi, dwi_du, dwi_dv, n, dn_du, dn_dv, &dwio_du, &dwio_dv);
        else d_refract(wi, dwi_du, dwi_dv, n, dn_du, dn_dv, v1.eta, &dwio_du, &dwio_dv);
        float to, po, tio, pio;
        sphcoords(wo, &to, &po); sphcoords(wio, &tio, &pio);
        float dt = to - tio, dp = po - pio;
        if (dp < -PI_F) dp += 2.f * PI_F; else if (dp > PI_F) dp -= 2.f * PI_F;
        C = v2(dt, dp);
        float dto_du, dpo_du, dto_dv, dpo_dv, dtio_du, dpio_du, dtio_dv, dpio_dv;
        d_sphcoords(wo, dwo_du, dwo_dv, &dto_du, &dpo_du, &dto_dv, &dpo_dv);
        d_sphcoords(wio, dwio_du, dwio_dv, &dtio_du, &dpio_du, &dtio_dv, &dpio_dv);
        dC = m2(dto_du - dtio_du, dto_dv - dtio_dv, dpo_du - dpio_du, dpo_dv - dpio_dv);
        success = true;
    }
    V3 woi; bool valid_o;
    if (v1.eta == 1.f) { woi = reflect(wo, n); valid_o = true; } else valid_o = refract(wo, n, v1.eta, &woi);
    if (valid_o && !success) {
        V3 dwoi_du, dwoi_dv;
        if (v1.eta == 1.f) d_reflect(wo, dwo_du, dwo_dv, n, dn_du, dn_dv, &dwoi_du, &dwoi_dv);
        else d_refract(wo, dwo_du, dwo_dv, n, dn_du, dn_dv, v1.eta, &dwoi_du, &dwoi_dv);
        float ti, pi, toi, poi;
        sphcoords(wi, &ti, &pi); sphcoords(woi, &toi, &poi);
        float dt = ti - toi, dp = pi - poi;
        if (dp < -PI_F) dp += 2.f * PI_F; else if (dp > PI_F) dp -= 2.f * PI_F;
        C = v2(dt, dp);
        float dti_du, dpi_du, dti_dv, dpi_dv, dtoi_du, dpoi_du, dtoi_dv, dpoi_dv;
        d_sphcoords(wi, dwi_du, dwi_dv, &dti_du, &dpi_du, &dti_dv, &dpi_dv);
        d_sphcoords(woi, dwoi_du, dwoi_dv, &dtoi_du, &dpoi_du, &dtoi_dv, &dpoi_dv);
        dC = m2(dti_du - dtoi_du, dti_dv - dtoi_dv, dpi_du - dpoi_du, dpi_dv - dpoi_dv);
        success = true;
    }
    dC.m00 = dC.m00 * g_plant[PLANT_SS_AD_JAC];
    float determinant = det2(dC);
    if (fabsf(determinant) < 1e-6f) return r;
    M2 inv = inverse2(dC);
    r.ok = success; r.C = C; r.dX = mulv2(inv, C);
    return r;
}

/* compute_step_halfvector (manifold_ss.cpp:407-492) */
static Step step_halfvector(V3 v0p, const MV &v1, const EI &v2e, V3 n_offset) {
    Step r; r.ok = false; r.C = v2(INF_F, INF_F); r.dX = v2(0.f, 0.f);
    V3 wo = ei_is_directional(v2e) ? v2e.d : v2e.p - v1.p;
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
    if (ei_is_directional(v2e)) {
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
    dH.m00 = dH.m00 * g_plant[PLANT_SS_HV_JAC];
    float determinant = det2(dH);
    if (fabsf(determinant) < 1e-6f) return r;
    M2 inv = inverse2(dH);
    V2 H = v2(dot(v1.s, h), dot(v1.t, h));
    V2 dHv = H - v2(n_offset.x, n_offset.y);
    r.ok = true; r.C = dHv; r.dX = mulv2(inv, dHv);
    return r;
}

struct NewtonOut { bool ok; SI si; int iterations; };

/* newton_solver SS (manifold_ss.cpp:329-405) */
static NewtonOut newton_ss(const Scene &S, V3 si_p, const MV &vtx_init, const EI &ei, V3 n_offset, bool smooth = false) {
    const sms_config &cfg = S.cfg;
    MV vtx = vtx_init;
    bool success = false;
    int iterations = 0;
    float beta = 1.f;
    SI si_current = si_invalid();
    while (iterations < cfg.max_iterations) {
        Step st = cfg.halfvector_constraints ? step_halfvector(si_p, vtx, ei, n_offset)
                                             : step_anglediff(si_p, vtx, ei, n_offset);
        if (!st.ok) break;
        if (norm2(st.C) < cfg.solver_threshold) { success = true; break; }
        V3 p_prop = vtx.p - cfg.step_scale * beta * (vtx.dp_du * st.dX.x + vtx.dp_dv * st.dX.y);
        V3 d_prop = normalize(p_prop - si_p);
        Ray ray = make_ray(si_p, d_prop);
        si_current = ray_intersect(S, ray);
        if (!si_current.valid || vtx.shape != si_current.shape) {
            beta = 0.5f * beta;
            iterations += 1;
            continue;
        }
        beta = std::min(1.f, 2.f * beta);
        vtx = mv_from_si(S, si_current, smooth);
        iterations++;
    }
    NewtonOut out; out.iterations = iterations; out.si = si_current;
    if (!success) { out.ok = false; return out; }
    V3 wx = normalize(si_p - vtx.p);
    V3 wy = ei_is_directional(ei) ? ei.d : normalize(ei.p - vtx.p);
    float cos_theta_x = dot(vtx.gn, wx), cos_theta_y = dot(vtx.gn, wy);
    bool refraction = cos_theta_x * cos_theta_y < 0.f, reflection = !refraction;
    if ((vtx.eta == 1.f && !reflection) || (vtx.eta != 1.f && !refraction)) { out.ok = false; return out; }
    out.ok = true;
    return out;
}

struct PathOut { bool ok; SI si_final, si_init; int iterations; };

static inline float logf_l(float x);
static void lean_moments(const Normalmap &nm, V2 uv, int level, V2 *mu, float sig[3]);

/* SpecularManifold::sample_gaussian (manifold.h:328-345) with warp::square_to_std_normal
   (warp.h:184-190); sigma = (s00, s11, s01) */
static V2 sample_gaussian(V2 mu, const float sig[3], V2 s) {
    const float r = sqrtf(-2.f * logf_l(1.f - s.x));
    float sn, cs;
    c_sincosf((2.f * PI_F) * s.y, &sn, &cs);
    V2 p = v2(cs * r, sn * r);
    const float sx = sqrtf(sig[0]), sy = sqrtf(sig[1]);
    const float rho = sig[2] / (sx * sy);
    const float a = sqrtf(1.f + rho), b = sqrtf(1.f - rho);
    const float d = a * 0.5f + b * 0.5f, o = a * 0.5f + b * -0.5f;     /* A = a P + b Q */
    V2 q = v2(d * p.x + o * p.y, o * p.x + d * p.y);
    q.x = q.x * sx; q.y = q.y * sy;
    return q + mu;
}

/* BSDF::lean at a surface interaction without uv partials: NormalmapBSDF::lean
   (bsdfs/normalmap.cpp:292-303) at the coarsest mip level (mipmap_level, :361-364); (0, 0)
   for other BSDFs (bsdf.cpp:34-38) */
static void lean_coarsest(const Scene &S, const SI &si, V2 *mu, float sig[3]) {
    const Shape &sh = S.shapes[si.shape];
    if (sh.bsdf != SMS_BSDF_NORMALMAP) { *mu = v2(0.f, 0.f); sig[0] = sig[1] = sig[2] = 0.f; return; }
    const Normalmap &nm = S.nmaps[sh.normalmap];
    V2 uv = si.uv * sh.tiles;
    uv = v2(uv.x - floorf(uv.x), uv.y - floorf(uv.y));
    lean_moments(nm, uv, (int) nm.lean.size() - 1, mu, sig);
    /* without use_slopes the mean comes from eval_normal's RGB as stored, not decoded */
    if (nm.encoded) *mu = v2(-nm.top_rgb[0] / nm.top_rgb[2], -nm.top_rgb[1] / nm.top_rgb[2]);
}

/* sample_path SS (manifold_ss.cpp:203-279): uniform seed, optional MNEE direction, and with
   config.twostage the LEAN-smoothed pre-solve (:251-270) from an offset slope drawn with
   lean_sample; iterations count both solves */
static PathOut sample_path_ss(const Scene &S, int shape, V3 si_p, const EI &ei, V2 sample, V3 n_offset, V3 p_start,
                              V2 lean_sample) {
    PathOut out; out.ok = false; out.si_final = si_invalid(); out.si_init = si_invalid(); out.iterations = 0;
    PS ps = sample_position(S, shape, sample);
    if (p_start.x != 0.f || p_start.y != 0.f || p_start.z != 0.f) ps.p = p_start;
    V3 d_tmp = normalize(ps.p - si_p);
    if (S.cfg.mnee_init) {
        const Shape &sh = S.shapes[shape];
        bool reflection = sh.k != 0.f;                 // any(imag(ior) != 0)
        if (reflection) d_tmp = normalize(sh.bbox_center - si_p);
        else d_tmp = normalize(ei.p * g_plant[PLANT_MNEE_TARGET] - si_p);
    }
    Ray ray = make_ray(si_p, d_tmp);
    SI si_init = ray_intersect(S, ray);
    if (!si_init.valid || shape != si_init.shape) return out;
    out.si_init = si_init;
    const bool twostage = S.cfg.twostage != 0;
    MV vtx_init = mv_from_si(S, si_init, twostage);
    int smooth_iterations = 0;
    if (twostage) {
        V2 mu_offset = v2(-n_offset.x / n_offset.z, -n_offset.y / n_offset.z);
        V2 mu; float sig[3];
        lean_coarsest(S, si_init, &mu, sig);
        V2 slope = sample_gaussian(0.5f * (mu + mu_offset), sig, lean_sample);
        V3 lean_normal = normalize(v3(-slope.x, -slope.y, 1.f));
        NewtonOut sm = newton_ss(S, si_p, vtx_init, ei, lean_normal, true);
        smooth_iterations = sm.iterations;
        if (!(sm.ok && sm.si.valid)) { out.iterations = smooth_iterations; return out; }
        vtx_init = mv_from_si(S, sm.si);
    }
    NewtonOut nr = newton_ss(S, si_p, vtx_init, ei, n_offset);
    out.iterations = smooth_iterations + nr.iterations;
    out.si_final = nr.si;
    out.ok = nr.ok && nr.si.valid;
    return out;
}

/* ---------------------------------------------------------------- rough conductors
   MicrofacetDistribution(Beckmann, alpha, alpha, sample_visible = false) of
   include/mitsuba/render/microfacet.h: sample (:229-290, isotropic), eval (:184-201), smith_g1
   (:331-355, the rational approximation); RoughConductor::eval (src/bsdfs/roughconductor.cpp:
   283-356, unpolarized). */
static V3 beckmann_sample(float alpha, V2 s, float *pdf) {
    float sin_phi, cos_phi;
    c_sincosf((2.f * PI_F) * s.y, &sin_phi, &cos_phi);
    const float alpha_2 = alpha * alpha;
    const float cos_theta = 1.f / sqrtf(fmaf(-alpha_2, logf_l(1.f - s.x), 1.f));
    const float cos_theta_2 = cos_theta * cos_theta;
    const float cos_theta_3 = std::max(cos_theta_2 * cos_theta, 1e-20f);
    *pdf = (1.f - s.x) / (PI_F * alpha * alpha * cos_theta_3);
    const float sin_theta = sqrtf(1.f - cos_theta_2);
    return v3(cos_phi * sin_theta, sin_phi * sin_theta, cos_theta);
}
static float beckmann_eval(float alpha, V3 m) {
    const float alpha_uv = alpha * alpha, cos_theta = m.z, cos_theta_2 = cos_theta * cos_theta;
    const float a = m.x / alpha, b = m.y / alpha;
    const float result = expf_l(-(a * a + b * b) / cos_theta_2) / (PI_F * alpha_uv * (cos_theta_2 * cos_theta_2));
    return result * cos_theta > 1e-20f ? result : (float) 0.f;
}
static float beckmann_g1(float alpha, V3 v, V3 m) {
    const float ax = alpha * v.x, ay = alpha * v.y;
    const float xy_alpha_2 = ax * ax + ay * ay, tan_theta_alpha_2 = xy_alpha_2 / (v.z * v.z);
    const float a = 1.f / sqrtf(tan_theta_alpha_2), a_sqr = a * a;
    float result = a >= 1.6f ? (float) 1.f : (3.535f * a + 2.181f * a_sqr) / ((1.f + 2.276f * a) + 2.577f * a_sqr);
    if (xy_alpha_2 == 0.f) result = 1.f;
    if (dot(v, m) * v.z <= 0.f) result = 0.f;
    return result;
}
static float rough_conductor_eval(const Shape &sh, V3 wi, V3 wo) {
    if (!(wi.z > 0.f && wo.z > 0.f)) return 0.f;
    const V3 H = normalize(wo + wi);
    const float D = beckmann_eval(sh.alpha, H);
    if (D == 0.f) return 0.f;
    const float G = beckmann_g1(sh.alpha, wi, H) * beckmann_g1(sh.alpha, wo, H);
    const float result = D * G / (4.f * wi.z);
    const float F = fresnel_conductor(dot(wi, H), sh.eta_r, sh.k);
    return F * result;
}

/* specular_reflectance SS (manifold_ss.cpp:638-693): the glossy branch for rough conductors
   (the BSDF in the half-vector domain), else the smooth one */
static V3 specular_reflectance_ss(const Scene &S, const SI &si, V3 wo) {
    if (!si.valid) return v3(0.f, 0.f, 0.f);
    const Shape &sh = S.shapes[si.shape];
    if (sh.alpha > 0.f) {
        const V3 wo_l = v3(dot(wo, si.sh_s), dot(wo, si.sh_t), dot(wo, si.sh_n));
        float val = rough_conductor_eval(sh, si.wi, wo_l);
        const V3 h_l = normalize(si.wi + wo_l);
        val = val * (4.f * fabsf(dot(wo_l, h_l)));
        return v3(val, val, val);
    }
    bool reflection = sh.k != 0.f;      // any(imag(ior) != 0); the all(imag == 0) sub-case is unreachable
    V3 n = si.sh_n;       // BSDF::frame(si).n == sh_frame.n for non-normal-mapped shapes
    float cos_theta = dot(n, wo);
    float val;
    if (reflection) {
        val = fresnel_conductor(fabsf(cos_theta), sh.eta_r, sh.k);
    } else {
        float eta = hmean_same(sh.eta_r);
        if (cos_theta < 0.f) eta = 1.f / eta;
        float F = fresnel(cos_theta, eta);
        val = 1.f - F;
        val = val * (eta * eta);
        val = val * g_plant[PLANT_SS_TRANSMIT];
    }
    return v3(val, val, val);
}

/* geometric_term SS (manifold_ss.cpp:695-857) */
static float geometric_term_ss(const MV &v0, const MV &v1, const MV &v2v) {
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
    M2 dc1_dx2 = m2(dot(dh_du, s) * g_plant[PLANT_SS_G], dot(dh_dv, s) * g_plant[PLANT_SS_G], dot(dh_du, t),
                    dot(dh_dv, t));
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
        G = std::min(G, 10.f);
        G /= fabsf(dot(wi, v0.n));
        G *= eta * eta;
        return G * g_plant[PLANT_SS_G_DIR];
    }
    if (fabsf(det2(dc1_dx1)) < 1e-6f) return 0.f;
    M2 inv = inverse2(dc1_dx1);
    float dx1_dx2 = fabsf(det2(mulm2(inv, dc1_dx2)));
    dx1_dx2 = std::min(dx1_dx2, g_plant[PLANT_SS_G_CLAMP]);
    V3 d = v0.p - v1.p;
    float inv_r2 = 1.f / sqnorm(d);
    d = d * sqrtf(inv_r2);
    float dw0_dx1 = fabsf(dot(d, v1.gn)) * inv_r2;
    return dw0_dx1 * dx1_dx2;
}

/* evaluate_path_contribution SS (manifold_ss.cpp:281-327) */
static V3 path_contribution_ss(const Scene &S, const SI &si, const EI &ei_, const SI &si_final) {
    MV vtx = mv_from_si(S, si_final);
    make_orthonormal(vtx);
    EI ei = ei_;
    if (ei.kind == SMS_EI_POINT) {
        const Emitter &em = S.emitters[ei.emitter];
        V3 d = em.position - si_final.p;
        float dist = norm(d), inv_dist = 1.f / dist;
        d = d * inv_dist;
        float emitter_pdf = ei.pdf;
        ei.p = em.position; ei.d = d; ei.n = d; ei.pdf = 1.f;
        ei.weight = (em.intensity * (inv_dist * inv_dist)) * dist * dist;
        ei.pdf *= emitter_pdf;
        ei.weight = ei.weight * (1.f / emitter_pdf);
    }
    MV vy;
    if (!emitter_interaction_to_vertex(S, ei, vtx.p, &vy)) return v3(0.f, 0.f, 0.f);
    MV vx = mv_from_si(S, si);
    make_orthonormal(vx);
    V3 thr = specular_reflectance_ss(S, si_final, normalize(ei.p - si_final.p));
    float G = geometric_term_ss(vx, vtx, vy);
    thr = thr * G;
    return v3(thr.x * ei.weight.x, thr.y * ei.weight.y, thr.z * ei.weight.z);
}

/* Diffuse::eval (src/bsdfs/diffuse.cpp:104-120) at the receiver */
static V3 receiver_bsdf_eval(const Scene &S, const SI &si, V3 wo_local) {
    const Shape &sh = S.shapes[si.shape];
    if (sh.bsdf != SMS_BSDF_DIFFUSE) return v3(0.f, 0.f, 0.f);
    float ci = si.wi.z, co = wo_local.z;
    if (!(ci > 0.f && co > 0.f)) return v3(0.f, 0.f, 0.f);
    return (sh.refl * INV_PI_F) * co;
}

static SI si_from_query(const sms_ss_query &q) {
    SI si = si_invalid();
    si.valid = true; si.t = 1.f;
    si.p = v3(q.p[0], q.p[1], q.p[2]); si.n = v3(q.n[0], q.n[1], q.n[2]);
    si.sh_n = v3(q.sh_n[0], q.sh_n[1], q.sh_n[2]); si.sh_s = v3(q.sh_s[0], q.sh_s[1], q.sh_s[2]);
    si.sh_t = v3(q.sh_t[0], q.sh_t[1], q.sh_t[2]); si.wi = v3(q.wi[0], q.wi[1], q.wi[2]);
    si.dp_du = si.sh_s; si.dp_dv = si.sh_t;
    si.shape = q.shape;
    return si;
}

struct Stats { uint64_t walks, conv, iters, bcalls, btrials, bmax; };

/* ================================================================ learned seed proposal
   No counterpart in the reference, whose seeds are uniform over the caster's area
   (manifold_ss.cpp:228-235).  Restates the library's defensive-mixture seed so that estimates
   with a proposal are checked walk for walk: the raw network outputs of a query are the
   input; everything after them (mixture preparation, seed draw, UV lookup, the walk) is
   restated here. */
static const int PROP_COMP = 8, PROP_RAW = 48;

struct UVGrid { int g = 0, shape = -1; std::vector<int> start, faces; };

/* G x G cells over the UV square (G the power of two with 2 G^2 >= faces, at most 1024);
   cell (x, y) lists in ascending order the faces whose UV bounding box spans it, spans taken
   as floor(lo G) .. floor(hi G) clamped to the grid */
static void build_uv_grid(const Scene &S, int shape, UVGrid &G) {
    const Shape &sh = S.shapes[shape];
    int g = 1;
    while (2 * (int64_t) g * g < sh.nf && g < 1024) g *= 2;
    auto cell_of = [&](float t) { int c = (int) floorf(t * (float) g); return std::min(std::max(c, 0), g - 1); };
    std::vector<std::vector<int>> cells((size_t) g * g);
    for (int f = 0; f < sh.nf; ++f) {
        const uint32_t *fc = &S.faces[3 * (size_t) (sh.fbase + f)];
        float ux[3], uy[3];
        for (int k = 0; k < 3; ++k) { ux[k] = S.uvs[2 * (size_t) fc[k]]; uy[k] = S.uvs[2 * (size_t) fc[k] + 1]; }
        const float lx = std::min(std::min(ux[0], ux[1]), ux[2]), hx = std::max(std::max(ux[0], ux[1]), ux[2]);
        const float ly = std::min(std::min(uy[0], uy[1]), uy[2]), hy = std::max(std::max(uy[0], uy[1]), uy[2]);
        for (int y = cell_of(ly); y <= cell_of(hy); ++y)
            for (int x = cell_of(lx); x <= cell_of(hx); ++x) cells[(size_t) y * g + x].push_back(sh.fbase + f);
    }
    G.g = g; G.shape = shape;
    G.start.assign(1, 0); G.faces.clear();
    for (auto &c : cells) { G.faces.insert(G.faces.end(), c.begin(), c.end()); G.start.push_back((int) G.faces.size()); }
}

/* surface point of uv: first listed face of uv's cell whose UV triangle holds it (UV-space
   barycentrics), positions interpolated with the same weights */
static bool uv_to_point(const Scene &S, const UVGrid &G, V2 uv, V3 *p) {
    int gx = (int) (uv.x * (float) G.g), gy = (int) (uv.y * (float) G.g);
    gx = gx < G.g - 1 ? gx : G.g - 1;
    gy = gy < G.g - 1 ? gy : G.g - 1;
    const int cell = gy * G.g + gx;
    for (int i = G.start[cell]; i < G.start[cell + 1]; ++i) {
        const uint32_t *fc = &S.faces[3 * (size_t) G.faces[i]];
        V2 t0 = v2(S.uvs[2 * fc[0]], S.uvs[2 * fc[0] + 1]), t1 = v2(S.uvs[2 * fc[1]], S.uvs[2 * fc[1] + 1]),
           t2 = v2(S.uvs[2 * fc[2]], S.uvs[2 * fc[2] + 1]);
        V2 e1 = t1 - t0, e2 = t2 - t0, d = uv - t0;
        float det = e1.x * e2.y - e1.y * e2.x;
        float b1 = (d.x * e2.y - d.y * e2.x) / det;
        float b2 = (e1.x * d.y - e1.y * d.x) / det;
        float b0 = (1.f - b1) - b2;
        if (b0 >= 0.f && b1 >= 0.f && b2 >= 0.f) {
            V3 q0 = v3(S.pos[3 * fc[0]], S.pos[3 * fc[0] + 1], S.pos[3 * fc[0] + 2]);
            V3 q1 = v3(S.pos[3 * fc[1]], S.pos[3 * fc[1] + 1], S.pos[3 * fc[1] + 2]);
            V3 q2 = v3(S.pos[3 * fc[2]], S.pos[3 * fc[2] + 1], S.pos[3 * fc[2] + 2]);
            *p = (q0 * b0 + q1 * b1) + q2 * b2;
            return true;
        }
    }
    return false;
}

static inline float omin(float a, float b) { return b < a ? b : a; }
static inline float omax(float a, float b) { return a < b ? b : a; }
static inline float sigmoid_l(float x) { return 1.f / (1.f + expf_l(-x)); }
static inline float tanh_l(float x) {
    float t = expf_l(-2.f * fabsf(x));
    float r = (1.f - t) / (1.f + t);
    return x < 0.f ? -r : r;
}
static inline float logf_l(float x) { return (float) (log2_d((double) x) * 0.69314718055994530942); }

/* raw outputs (per component: weight logit, mean u, mean v, log sigma u, log sigma v,
   correlation) -> cdf of the softmax weights, sigmoid means, sigmas exp(clamp(., -7, 0)),
   correlation 0.95 tanh */
static void gmm_prepare(const float *o, float *g) {
    float mx = o[0];
    for (int c = 1; c < PROP_COMP; ++c) mx = omax(mx, o[6 * c]);
    float e[PROP_COMP], sum = 0.f;
    for (int c = 0; c < PROP_COMP; ++c) { e[c] = expf_l(o[6 * c] - mx); sum = sum + e[c]; }
    float cdf = 0.f;
    for (int c = 0; c < PROP_COMP; ++c) {
        cdf = cdf + e[c] / sum;
        g[c] = cdf;
        g[8 + c] = sigmoid_l(o[6 * c + 1]);
        g[16 + c] = sigmoid_l(o[6 * c + 2]);
        g[24 + c] = expf_l(omin(omax(o[6 * c + 3], -7.f), 0.f));
        g[32 + c] = expf_l(omin(omax(o[6 * c + 4], -7.f), 0.f));
        g[40 + c] = 0.95f * tanh_l(o[6 * c + 5]);
    }
}

static void proposal_inputs(const Scene &S, V3 p, V3 n, V3 light, float *x) {
    x[0] = (p.x - S.box_c[0]) * S.box_inv; x[1] = (p.y - S.box_c[1]) * S.box_inv; x[2] = (p.z - S.box_c[2]) * S.box_inv;
    x[3] = n.x; x[4] = n.y; x[5] = n.z;
    x[6] = (light.x - S.box_c[0]) * S.box_inv; x[7] = (light.y - S.box_c[1]) * S.box_inv;
    x[8] = (light.z - S.box_c[2]) * S.box_inv;
}

/* seed of one walk from its stream: t0 < defensive -> the reference's uniform seed from the
   next two draws (kind 0, *out = the 2D sample); else component by t1 against the cdf and a
   Box-Muller draw from (t2, t3) through the Cholesky factor (kind 1, *out = uv) */
static int proposal_seed(const float *g, float defensive, Pcg32 *rw, V2 *out) {
    const float t0 = pcg32_float(rw);
    if (t0 < defensive) { out->x = pcg32_float(rw); out->y = pcg32_float(rw); return 0; }
    const float t1 = pcg32_float(rw), t2 = pcg32_float(rw), t3 = pcg32_float(rw);
    int c = PROP_COMP - 1;
    for (int j = 0; j < PROP_COMP; ++j) if (t1 < g[j]) { c = j; break; }
    const float r = sqrtf(-2.f * logf_l(1.f - t2));
    float sn, cs;
    c_sincosf((2.f * PI_F) * t3, &sn, &cs);
    const float z1 = r * cs, z2 = r * sn;
    const float rho = g[40 + c], q = sqrtf(1.f - rho * rho);
    *out = v2(g[8 + c] + g[24 + c] * z1, g[16 + c] + g[32 + c] * (rho * z1 + q * z2));
    return 1;
}

/* Anchor of a single-scatter query on mesh `shape`: the uv of the first surface hit by the ray
   from the shading point toward the light sample, when that surface is `shape` (the segment
   that manifold next-event estimation starts its walk from, manifold_ss.cpp:247-250), else
   NaN.  Anchored mixtures place the mixture's [0,1]^2 in the uv box a + 2 r (xi - 1/2). */
static V2 proposal_anchor(const Scene &S, int shape, V3 si_p, V3 light_p) {
    const SI h = ray_intersect(S, make_ray(si_p, normalize(light_p - si_p)));
    if (h.valid && h.shape == shape) return h.uv;
    return v2(std::numeric_limits<float>::quiet_NaN(), std::numeric_limits<float>::quiet_NaN());
}

/* raw: network outputs, row (qid - raw_base) for query qid; radius > 0: anchored mixtures */
struct Prop { const float *raw; int shape; float defensive; UVGrid grid; int64_t raw_base = 0; float radius = 0.f; };

/* one walk of a query whose mixture is g: uniform seed, or the mixture's uv mapped to the
   surface and used as p_start; uv outside [0,1)^2 or in no face -> failed walk, 0 iterations */
static PathOut proposal_walk(const Scene &S, const Prop &P, const float *g, int shape, V3 si_p, const EI &ei,
                             V3 n_offset, Pcg32 *rw, const V2 *anchor = nullptr) {
    V2 s;
    V3 p_start = v3(0.f, 0.f, 0.f);
    /* an anchored query without an anchor seeds uniformly: the defensive draw with weight 1 */
    const bool no_anchor = anchor && !(anchor->x == anchor->x);
    float defensive = P.defensive;
    if (no_anchor) defensive = 1.f;
    if (proposal_seed(g, defensive, rw, &s) == 1) {
        PathOut po;
        po.ok = false; po.iterations = 0; po.si_final = si_invalid(); po.si_init = si_invalid();
        if (anchor) {
            const float side = 2.f * P.radius;
            s = v2(anchor->x + side * (s.x - 0.5f), anchor->y + side * (s.y - 0.5f));
        }
        if (!(s.x >= 0.f && s.x < 1.f && s.y >= 0.f && s.y < 1.f)) return po;
        if (!uv_to_point(S, P.grid, s, &p_start)) return po;
        s = v2(0.f, 0.f);
    }
    V2 ls = v2(0.f, 0.f);
    if (S.cfg.twostage) { ls.x = pcg32_float(rw); ls.y = pcg32_float(rw); }
    return sample_path_ss(S, shape, si_p, ei, s, n_offset, p_start, ls);
}

/* specular_manifold_sampling SS (manifold_ss.cpp:28-201), per-(query, walk) PCG32 streams */
static V3 estimate_ss(const Scene &S, const sms_ss_query &q, uint64_t seed, uint64_t qid, Stats &st,
                      const Prop *P = nullptr) {
    const sms_config &cfg = S.cfg;
    if (q.shape < 0) return v3(0.f, 0.f, 0.f);     /* no receiver: zero estimate, no walks */
    SI si = si_from_query(q);
    Pcg32 rq; stream_seed(&rq, seed, qid, 0);
    EI ei = sample_emitter_interaction(S, si.p, S.emitters_ss, &rq);
    V3 result = v3(0.f, 0.f, 0.f);
    if (S.casters_ss.empty()) return result;
    for (size_t c = 0; c < S.casters_ss.size(); ++c) {
        int shape = S.casters_ss[c];
        uint64_t wbase = ((uint64_t) c << 32);
        /* offset normal of a rough caster (manifold_ss.cpp:61-70), from the query's stream after
           the emitter sample, one 2D sample per rough caster in order */
        V3 n_offset = v3(0.f, 0.f, 1.f);
        float p_offset = 1.f;
        if (S.shapes[shape].alpha > 0.f) {
            V2 so; so.x = pcg32_float(&rq); so.y = pcg32_float(&rq);
            n_offset = beckmann_sample(S.shapes[shape].alpha, so, &p_offset);
        }
        V3 value = v3(0.f, 0.f, 0.f);
        float gq[PROP_RAW];
        const bool use_prop = P && shape == P->shape;
        if (use_prop) gmm_prepare(P->raw + ((int64_t) qid - P->raw_base) * PROP_RAW, gq);
        const V2 anchor = use_prop && P->radius > 0.f ? proposal_anchor(S, shape, si.p, ei.p) : v2(0.f, 0.f);
        auto run_walk = [&](uint64_t k) {
            Pcg32 rw; stream_seed(&rw, seed, qid, wbase + k + 1);
            PathOut po;
            if (use_prop) {
                po = proposal_walk(S, *P, gq, shape, si.p, ei, n_offset, &rw, P->radius > 0.f ? &anchor : nullptr);
            } else {
                V2 s; s.x = pcg32_float(&rw); s.y = pcg32_float(&rw);
                V2 ls = v2(0.f, 0.f);
                if (cfg.twostage) { ls.x = pcg32_float(&rw); ls.y = pcg32_float(&rw); }
                po = sample_path_ss(S, shape, si.p, ei, s, n_offset, v3(0.f, 0.f, 0.f), ls);
            }
            st.walks++; st.iters += po.iterations; if (po.ok) st.conv++;
            return po;
        };
        auto visible = [&](const SI &sf) {
            /* manifold_ss.cpp:88-102: toward a directional emitter the ray leaves the solution */
            if (ei_is_directional(ei)) return !ray_test(S, make_ray(sf.p, ei.d));
            V3 d = sf.p - ei.p;
            float dist = norm(d);
            d = d * (1.f / dist);
            Ray rv = make_ray(ei.p, d, RAY_EPS_F * (1.f + hmax_abs(ei.p)), dist * (1.f - RAY_EPS_F));
            return !ray_test(S, rv);
        };
        if (!cfg.biased) {
            PathOut po = run_walk(0);
            if (!po.ok) continue;
            V3 direction = normalize(po.si_final.p - si.p);
            if (!visible(po.si_final)) continue;
            V3 specular_val = path_contribution_ss(S, si, ei, po.si_final);
            V3 dl = v3(dot(direction, si.sh_s), dot(direction, si.sh_t), dot(direction, si.sh_n));
            V3 bsdf_val = receiver_bsdf_eval(S, si, dl);
            float inv_prob = 1.f;
            int iterations = 1;
            uint64_t k = 1;
            st.bcalls++;
            while (true) {
                PathOut pt = run_walk(k++);
                V3 dtr = normalize(pt.si_final.p - si.p);
                if (pt.ok && fabsf(dot(direction, dtr) - 1.f) < cfg.uniqueness_threshold) break;
                inv_prob += 1.f;
                iterations++;
                if (cfg.max_trials > 0 && iterations > cfg.max_trials) { inv_prob = 0.f; break; }
            }
            st.btrials += iterations;
            if ((uint64_t) iterations > st.bmax) st.bmax = iterations;
            value = v3(bsdf_val.x * specular_val.x * inv_prob, bsdf_val.y * specular_val.y * inv_prob,
                       bsdf_val.z * specular_val.z * inv_prob);
        } else {
            std::vector<V3> solutions;
            for (int m = 0; m < cfg.max_trials; ++m) {
                PathOut po = run_walk((uint64_t) m);
                if (!po.ok) continue;
                V3 direction = normalize(po.si_final.p - si.p);
                if (!visible(po.si_final)) continue;
                bool dup = false;
                for (size_t k = 0; k < solutions.size(); ++k)
                    if (fabsf(dot(direction, solutions[k]) - 1.f) < cfg.uniqueness_threshold) { dup = true; break; }
                if (dup) continue;
                solutions.push_back(direction);
                V3 dl = v3(dot(direction, si.sh_s), dot(direction, si.sh_t), dot(direction, si.sh_n));
                V3 bsdf_val = receiver_bsdf_eval(S, si, dl);
                V3 specular_val = path_contribution_ss(S, si, ei, po.si_final);
                value = value + v3(bsdf_val.x * specular_val.x, bsdf_val.y * specular_val.y, bsdf_val.z * specular_val.z);
            }
        }
        result = result + value / p_offset;
    }
    return result;
}

/* ================================================================ glints */
static inline bool inside_triangle(V2 p, V2 a, V2 b, V2 c) {
    V2 v0 = c - a, v1 = b - a, v2_ = p - a;
    float dot00 = dot2(v0, v0), dot01 = dot2(v0, v1), dot02 = dot2(v0, v2_), dot11 = dot2(v1, v1), dot12 = dot2(v1, v2_);
    float inv_denom = 1.f / (dot00 * dot11 - dot01 * dot01);
    float u = (dot11 * dot02 - dot01 * dot12) * inv_denom, v = (dot00 * dot12 - dot01 * dot02) * inv_denom;
    return (u >= 0.f) && (v >= 0.f) && (u + v < 1.f);
}

struct GlintSI { V3 p, gn, n, s, t, dp_du, dp_dv, wi; V2 uv, duv_dx, duv_dy; int shape; };

/* newton_solver + compute_step glints (manifold_glints.cpp:296-373) */
static bool newton_glint(const Scene &S, const GlintSI &si, V2 target, V2 uv_init, V2 *uv_out, int *iters) {
    const sms_config &cfg = S.cfg;
    const Shape &sh = S.shapes[si.shape];
    const Normalmap &nm = S.nmaps[sh.normalmap];
    V2 uv = uv_init;
    V2 fp_a = (si.uv + 0.5f * si.duv_dx) + 0.5f * si.duv_dy, fp_b = (si.uv + 0.5f * si.duv_dx) - 0.5f * si.duv_dy,
       fp_c = (si.uv - 0.5f * si.duv_dx) - 0.5f * si.duv_dy, fp_d = (si.uv - 0.5f * si.duv_dx) + 0.5f * si.duv_dy;
    bool success = false;
    int iterations = 0;
    float beta = 1.f;
    while (iterations < cfg.max_iterations) {
        V2 tuv = tile_uv(uv, sh.tiles);
        Texel4 tx = nm_fetch(nm, tuv);
        V2 slope = nm_slope(tx);
        V2 tu, tv; nm_slope_deriv(tx, (float) nm.res, &tu, &tv);
        V2 du = tu * sh.tiles, dv = tv * sh.tiles;
        V2 dS = slope - target;
        M2 J = m2(du.x, dv.x, du.y, dv.y);
        if (fabsf(det2(J)) < 1e-6f) break;
        V2 step = mulv2(inverse2(J), dS);
        if (norm2(dS) < cfg.solver_threshold) { success = true; break; }
        V2 uv_prop = uv - cfg.step_scale * beta * step;
        if (!(inside_triangle(uv_prop, fp_a, fp_b, fp_c) || inside_triangle(uv_prop, fp_a, fp_c, fp_d))) {
            beta = 0.5f * beta;
            iterations++;
            continue;
        }
        beta = std::min(1.f, 2.f * beta);
        uv = uv_prop;
        iterations++;
    }
    *uv_out = uv; *iters = iterations;
    return success;
}

static inline V3 to_local(const GlintSI &si, V3 v) { return v3(dot(v, si.s), dot(v, si.t), dot(v, si.n)); }

/* sample_glint (manifold_glints.cpp:225-265) */
static bool sample_glint(const Scene &S, V3 sensor_p, V3 ei_p, const GlintSI &si, V2 sample, V3 n_offset,
                         V2 xi_start, V2 *uv_final, V2 *uv_init, int *iters) {
    *iters = 0;
    V3 wi = to_local(si, normalize(sensor_p - si.p)), wo = to_local(si, normalize(ei_p - si.p));
    if (wi.z <= 0.f || wo.z <= 0.f) { *uv_final = v2(0.f, 0.f); *uv_init = v2(0.f, 0.f); return false; }
    V3 h = normalize(wi + wo);
    V2 h_s = v2(-h.x / h.z, -h.y / h.z), o_s = v2(-n_offset.x / n_offset.z, -n_offset.y / n_offset.z);
    V2 target = h_s + o_s;
    V2 fp = sample;
    if (xi_start.x != 1.f || xi_start.y != 1.f) fp = xi_start;
    V2 duv = (fp.x - 0.5f) * si.duv_dx + (fp.y - 0.5f) * si.duv_dy;
    *uv_init = si.uv + duv;
    return newton_glint(S, si, target, *uv_init, uv_final, iters);
}

/* specular_reflectance glints, smooth branch (manifold_glints.cpp:375-409) */
static float specular_reflectance_glint(const Shape &sh, V3 n, V3 wo) {
    float cos_theta = dot(n, wo);
    if (sh.k == 0.f) return fresnel(fabsf(cos_theta), sh.eta_r);
    return fresnel_conductor(fabsf(cos_theta), sh.eta_r, sh.k);
}

/* geometric_term glints (manifold_glints.cpp:411-554) */
static float geometric_term_glint(const MV &v0, const MV &v1, const MV &v2v) {
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
    float G = 0.f;
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
    G = std::min(1e-5f, G);
    return G;
}

/* the three vertices of a glint path: sensor point, the footprint's point with the normal map
   at uv, the emitter vertex (manifold_glints.cpp:267-283) */
static bool glint_vertices(const Scene &S, V3 sensor_p, const EI &ei, const GlintSI &gsi, V2 uv, MV *vx, MV *vtx,
                           MV *vy) {
    if (!emitter_interaction_to_vertex(S, ei, gsi.p, vy)) return false;
    *vx = mv_point(sensor_p);
    V3 dx = normalize(gsi.p - sensor_p);
    vx->n = dx; vx->gn = dx;
    coordinate_system(vx->n, &vx->dp_du, &vx->dp_dv);
    SI si = si_invalid();
    si.valid = true; si.p = gsi.p; si.n = gsi.gn; si.sh_n = gsi.n; si.sh_s = gsi.s; si.sh_t = gsi.t;
    si.dp_du = gsi.dp_du; si.dp_dv = gsi.dp_dv; si.uv = uv; si.shape = gsi.shape;
    *vtx = mv_from_si(S, si);
    make_orthonormal(*vtx);
    return true;
}

/* evaluate_glint_contribution (manifold_glints.cpp:267-294) */
static V3 glint_contribution(const Scene &S, V3 sensor_p, const EI &ei, const GlintSI &gsi, V2 uv) {
    MV vx, vtx, vy;
    if (!glint_vertices(S, sensor_p, ei, gsi, uv, &vx, &vtx, &vy)) return v3(0.f, 0.f, 0.f);
    float F = specular_reflectance_glint(S.shapes[gsi.shape], vtx.n, normalize(ei.p - gsi.p));
    float G = geometric_term_glint(vx, vtx, vy);
    float v = F * G;
    return v3(v, v, v);
}

static GlintSI gsi_from_query(const sms_glint_query &q) {
    GlintSI g;
    g.p = v3(q.p[0], q.p[1], q.p[2]); g.gn = v3(q.n[0], q.n[1], q.n[2]); g.n = v3(q.sh_n[0], q.sh_n[1], q.sh_n[2]);
    g.s = v3(q.sh_s[0], q.sh_s[1], q.sh_s[2]); g.t = v3(q.sh_t[0], q.sh_t[1], q.sh_t[2]);
    g.dp_du = v3(q.dp_du[0], q.dp_du[1], q.dp_du[2]); g.dp_dv = v3(q.dp_dv[0], q.dp_dv[1], q.dp_dv[2]);
    g.wi = v3(q.wi[0], q.wi[1], q.wi[2]);
    g.uv = v2(q.uv[0], q.uv[1]); g.duv_dx = v2(q.duv_dx[0], q.duv_dx[1]); g.duv_dy = v2(q.duv_dy[0], q.duv_dy[1]);
    g.shape = q.shape;
    return g;
}

/* sensor importance We * film area (manifold_glints.cpp:146-153; perspective.cpp:224-276) */
static float sensor_we(const Scene &S, V3 sensor_p, V3 p) {
    const Camera &c = S.cam;
    V3 d = normalize(p - sensor_p);
    V3 dl = v3(dot(d, c.left), dot(d, c.up), dot(d, c.dir));
    float cos_theta = dl.z;
    if (!(cos_theta > 0.f)) return 0.f;
    float inv_cos = 1.f / cos_theta;
    float px = dl.x * inv_cos, py = dl.y * inv_cos;
    if (!(px >= -c.rect_x && px <= c.rect_x && py >= -c.rect_y && py <= c.rect_y)) return 0.f;
    float We = c.normalization * inv_cos * inv_cos * inv_cos;
    return We * (float) (c.w * c.h);
}

/* ---------------------------------------------------------------- LEAN pdf (glint MIS) */
struct LeanParams { V3 n; float ax, ay, rho, sqrt_one_minus_rho_sqr, tx_n, ty_n; };

static const float INV_SQRT_PI_F = 0.56418958354775628695f;

/* NormalmapBSDF::mipmap_level (bsdfs/normalmap.cpp:361-374).  A negative ceil() converts to a
   huge size_t in the reference and min() then selects the coarsest level. */
static int lean_level(float tiles, V2 duv_dx, V2 duv_dy, int levels) {
    V2 a = duv_dx * tiles, b = duv_dy * tiles;
    float width = 2.f * std::max(std::max(fabsf(a.x), fabsf(a.y)), std::max(fabsf(b.x), fabsf(b.y)));
    float l = ceilf((float) (levels - 1) + log2f_l(std::max(width, 1e-8f)));
    if (l < 0.f) return levels - 1;
    return std::min((int) l, levels - 1);
}

/* eval_normal (use_slopes) and eval_lean_sigma at one level (normalmap.h:61-92,198-228) */
static void lean_moments(const Normalmap &nm, V2 uv, int level, V2 *mu, float sig[3]) {
    const int res = nm.res >> level;
    const float fres = (float) res, duv = 1.f / fres;
    V2 pos = v2((uv.x - 0.5f * duv) * fres, (uv.y - 0.5f * duv) * fres);
    pos.x = std::min(std::max(pos.x, 0.f), fres - 1.f); pos.y = std::min(std::max(pos.y, 0.f), fres - 1.f);
    uint32_t xmin = (uint32_t) floorf(pos.x), ymin = (uint32_t) floorf(pos.y);
    uint32_t xmax = xmin + 1, ymax = ymin + 1, lim = (uint32_t) (res - 1);
    xmin = std::min(xmin, lim); ymin = std::min(ymin, lim); xmax = std::min(xmax, lim); ymax = std::min(ymax, lim);
    V2 w1 = v2(pos.x - (float) xmin, pos.y - (float) ymin), w0 = v2(1.f - w1.x, 1.f - w1.y);
    const float *L = nm.lean[level].data();
    const float *v00 = L + 5 * ((size_t) ymin * res + xmin), *v10 = L + 5 * ((size_t) ymin * res + xmax),
                *v01 = L + 5 * ((size_t) ymax * res + xmin), *v11 = L + 5 * ((size_t) ymax * res + xmax);
    float v[5];
    for (int c = 0; c < 5; ++c) {
        float a = fmaf(w0.x, v00[c], w1.x * v10[c]), b = fmaf(w0.x, v01[c], w1.x * v11[c]);
        v[c] = fmaf(w0.y, a, w1.y * b);
    }
    V3 n = v3(-v[0], -v[1], 1.f);
    *mu = v2(-n.x / n.z, -n.y / n.z);
    sig[0] = std::max(0.f, v[2] - v[0] * v[0]);     /* sigma(0,0) */
    sig[1] = std::max(0.f, v[3] - v[1] * v[1]);     /* sigma(1,1) */
    sig[2] = std::max(0.f, v[4] - v[0] * v[1]);     /* sigma(0,1) = sigma(1,0) */
}

/* LEAN::p_from_lean_and_base (lean.h:44-77) */
static LeanParams lean_params(V2 mu, const float sig[3], float alpha_u, float alpha_v) {
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
    float tmp3 = std::max(0.f, e3 - e1 * e1), tmp4 = std::max(0.f, e4 - e2 * e2);
    p.ax = sqrtf(2.f * tmp3);
    p.ay = sqrtf(2.f * tmp4);
    p.rho = 2.f * (e5 - e1 * e2) / (p.ax * p.ay);
    p.sqrt_one_minus_rho_sqr = sqrtf(1.f - p.rho * p.rho);
    p.n = normalize(v3(-p.tx_n, -p.ty_n, 1.f));
    return p;
}

/* LEAN::sigma_std, sigma, g1, gaf, p22, ndf, vndf (lean.h:79-212) */
static float lean_sigma_std(float cos_theta) {
    if (cos_theta == 1.f) return 1.f;
    float sin_theta = sqrtf(1.f - cos_theta * cos_theta);
    float nu = cos_theta / sin_theta;
    float tmp = expf_l(-nu * nu) * INV_SQRT_PI_F;
    return 0.5f * (cos_theta * (1.f + erff_l(nu)) + sin_theta * tmp);
}
static float lean_sigma(V3 k, const LeanParams &p) {
    float a = k.x * p.ax + k.y * p.ay * p.rho;
    float b = k.y * p.ay * p.sqrt_one_minus_rho_sqr;
    float c = k.z - k.x * p.tx_n - k.y * p.ty_n;
    float nrm = sqrtf(a * a + b * b + c * c);
    V3 kp = v3(a, b, c) * (1.f / nrm);
    return nrm * lean_sigma_std(kp.z);
}
static float lean_g1(V3 k, const LeanParams &p) {
    if (dot(k, p.n) > 0.f) return k.z / lean_sigma(k, p);
    return 0.f;
}
static float lean_p22(float sx, float sy, const LeanParams &p) {
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
static float lean_ndf(V3 h, const LeanParams &p) {
    if (h.z > 5.9604644775390625e-08f) {          /* math::Epsilon<float> */
        float tmp1 = h.z * h.z, tmp2 = tmp1 * tmp1;
        return lean_p22(-h.x / h.z, -h.y / h.z, p) / tmp2;
    }
    return 0.f;
}
static float lean_vndf(V3 h, V3 wi, const LeanParams &p) {
    float dp = dot(wi, h);
    if (dp > 0.f) return dp * lean_ndf(h, p) / lean_sigma(wi, p);
    return 0.f;
}

/* SpecularManifoldGlints::lean_pdf (manifold_glints.cpp:556-579) with the smooth conductor's
   roughness 0 (conductor.cpp:288-290); wi, wo in the query's shading frame */
static float lean_pdf(const Scene &S, const GlintSI &si, V3 wo) {
    if (si.wi.z < 0.f || wo.z < 0.f) return 0.f;
    const Shape &sh = S.shapes[si.shape];
    const Normalmap &nm = S.nmaps[sh.normalmap];
    int level = lean_level(sh.tiles, si.duv_dx, si.duv_dy, (int) nm.lean.size());
    V2 uv = si.uv * sh.tiles;
    uv = v2(uv.x - floorf(uv.x), uv.y - floorf(uv.y));
    V2 mu; float sig[3];
    lean_moments(nm, uv, level, &mu, sig);
    LeanParams p = lean_params(mu, sig, 0.f, 0.f);
    V3 H = normalize(wo + si.wi);
    float G1i = lean_g1(wo, p), G1o = lean_g1(si.wi, p);
    float G = G1i * G1o;
    if (!(G > 0.f)) return 0.f;
    return lean_vndf(H, si.wi, p) / (4.f * dot(si.wi, H));
}

/* power heuristic (manifold_glints.h:155-159) */
static float mis_power(float a, float b) {
    a *= a;
    b *= b;
    return a > 0.f ? a / (a + b) : (float) 0.f;
}

/* Scene::pdf_emitter_direction for the SMS strategy's emitter sample (scene.cpp:206-219,
   area.cpp:119-125, sphere.cpp:254-267); 1 for delta emitters (manifold_glints.cpp:158-165) */
static float sms_emitter_pdf(const Scene &S, const EI &ei, V3 p) {
    if (ei.kind != SMS_EI_AREA) return 1.f;
    const Emitter &em = S.emitters[ei.emitter];
    const Shape &sh = S.shapes[em.shape];
    float sin_alpha = sh.radius * (1.f / norm(sh.center - p));
    float cos_alpha = safe_sqrt(1.f - sin_alpha * sin_alpha);
    float pdf;
    if (sin_alpha < 0.99999994039535522461f) pdf = INV_TWOPI_F / (1.f - cos_alpha);
    else pdf = sh.inv_surface_area * (norm(ei.p - p) * norm(ei.p - p)) / fabsf(dot(ei.d, ei.n));
    if (!(dot(ei.d, ei.n) < 0.f)) pdf = 0.f;
    if (S.emitters.size() > 1) pdf = pdf * (1.f / (float) S.emitters.size());
    return pdf;
}

/* BSDF-sampling strategy of the glint estimator (manifold_glints.cpp:183-220): an offset uv
   inside the pixel footprint, NormalmapBSDF::sample there (bsdfs/normalmap.cpp:118-136: the
   nested smooth conductor, conductor.cpp:217-275, in the level-0 perturbed frame, the reflected
   direction kept only if it stays on the same side of the unperturbed frame), one ray
   (Interaction::spawn_ray, interaction.h:58-61); a hit on an area emitter adds
   mis(p_bsdf, p_sms) * F * L with p_bsdf the LEAN pdf at the offset uv plus 1 and p_sms the
   emitter-sampling density of the hit (AreaLight::eval, area.cpp:71-79: front side only).
   The footprint offset is drawn from PCG32 stream GLINT_BSDF_STREAM of the query. */
static const uint64_t GLINT_BSDF_STREAM = 0xFFFFFFFEull;
static bool scene_has_area_emitter(const Scene &S) {
    for (const Emitter &e : S.emitters)
        if (e.type == SMS_EMITTER_AREA) return true;
    return false;
}
/* parts (tests, may be null): [0..1] footprint sample, [2..4] ray direction, [5] emitter hit
   (-1: none), [6] F, [7] p_bsdf, [8] p_sms, [9] mis */
static V3 glint_bsdf_strategy(const Scene &S, const GlintSI &si, uint64_t seed, uint64_t qid, float *parts = nullptr) {
    const V3 zero = v3(0.f, 0.f, 0.f);
    Pcg32 r; stream_seed(&r, seed, qid, GLINT_BSDF_STREAM);
    float fx = pcg32_float(&r), fy = pcg32_float(&r);
    if (parts) {
        for (int i = 0; i < 10; ++i) parts[i] = 0.f;
        parts[0] = fx; parts[1] = fy; parts[5] = -1.f;
    }
    V2 duv = (fx - 0.5f) * si.duv_dx + (fy - 0.5f) * si.duv_dy;
    SI so = si_invalid();
    so.valid = true; so.p = si.p; so.n = si.gn; so.sh_n = si.n; so.sh_s = si.s; so.sh_t = si.t;
    so.dp_du = si.dp_du; so.dp_dv = si.dp_dv; so.uv = si.uv + duv; so.shape = si.shape;
    MV f = mv_from_si(S, so);                     /* NormalmapBSDF::frame (n, s, t) at the offset uv */
    V3 wi_w = to_world(si.s, si.t, si.n, si.wi);
    V3 wi_p = v3(dot(wi_w, f.s), dot(wi_w, f.t), dot(wi_w, f.n));
    if (!(wi_p.z > 0.f)) return zero;
    V3 wo_p = v3(-wi_p.x, -wi_p.y, wi_p.z);
    const Shape &sh = S.shapes[si.shape];
    float F = fresnel_conductor(wi_p.z, sh.eta_r, sh.k);
    V3 wo_w = to_world(f.s, f.t, f.n, wo_p);
    V3 wo = v3(dot(wo_w, si.s), dot(wo_w, si.t), dot(wo_w, si.n));
    if (!(wo_p.z * wo.z > 0.f)) return zero;
    V3 d = to_world(si.s, si.t, si.n, wo);
    if (parts) { parts[2] = d.x; parts[3] = d.y; parts[4] = d.z; parts[6] = F; }
    SI h = ray_intersect(S, spawn_ray(si.p, d));
    if (!h.valid) return zero;
    int e = S.shapes[h.shape].emitter;
    if (e < 0 || S.emitters[e].type != SMS_EMITTER_AREA) return zero;
    if (parts) parts[5] = (float) e;
    GlintSI sio = si;
    sio.uv = so.uv;
    float p_bsdf = lean_pdf(S, sio, wo) + 1.f;
    EI eh;
    memset(&eh, 0, sizeof(eh));
    eh.kind = SMS_EI_AREA; eh.emitter = e; eh.p = h.p; eh.n = h.sh_n;
    V3 dh = h.p - si.p;
    eh.d = dh * (1.f / norm(dh));
    float p_sms = sms_emitter_pdf(S, eh, si.p);
    float mis = mis_power(p_bsdf, p_sms);
    if (S.cfg.sms_strategy_only) mis = 0.f;
    if (parts) { parts[7] = p_bsdf; parts[8] = p_sms; parts[9] = mis; }
    if (!(dot(d * -1.f, h.sh_n) > 0.f)) return zero;
    const V3 L = S.emitters[e].radiance;
    float w = mis * F;
    return v3(w * L.x, w * L.y, w * L.z);
}

/* The glint estimator (manifold_glints.cpp:28-223): SMS strategy, then the BSDF strategy,
   which only an area emitter can receive (skipped in scenes without one) */
static V3 estimate_glint_sms(const Scene &S, const sms_glint_query &q, const GlintSI &si, V3 sensor_p, const EI &ei,
                             V3 edir, uint64_t seed, uint64_t qid, Stats &st, const float *gq, float defensive);
/* gq: the query's prepared seed mixture over the footprint (nullptr: the reference's uniform
   footprint seeds) */
static V3 estimate_glint(const Scene &S, const sms_glint_query &q, uint64_t seed, uint64_t qid, Stats &st,
                         const float *gq = nullptr, float defensive = 1.f) {
    V3 result = v3(0.f, 0.f, 0.f);
    if (q.shape < 0) return result;
    GlintSI si = gsi_from_query(q);
    V3 sensor_p = v3(q.sensor_p[0], q.sensor_p[1], q.sensor_p[2]);
    Pcg32 rq; stream_seed(&rq, seed, qid, 0);
    EI ei = sample_emitter_interaction(S, si.p, S.emitters_all, &rq);
    V3 edir = ei.p - si.p;
    float edist = norm(edir);
    edir = edir * (1.f / edist);
    Ray re = make_ray(si.p, edir, RAY_EPS_F * (1.f + hmax_abs(si.p)), edist * (1.f - RAY_EPS_F));
    if (!ray_test(S, re)) result = estimate_glint_sms(S, q, si, sensor_p, ei, edir, seed, qid, st, gq, defensive);
    if (scene_has_area_emitter(S)) result = result + glint_bsdf_strategy(S, si, seed, qid);
    return result;
}

/* SMS strategy of the glint estimator (manifold_glints.cpp:50-181), emitter sample visible */
static V3 estimate_glint_sms(const Scene &S, const sms_glint_query &q, const GlintSI &si, V3 sensor_p, const EI &ei,
                             V3 edir, uint64_t seed, uint64_t qid, Stats &st, const float *gq, float defensive) {
    const sms_config &cfg = S.cfg;
    V3 result = v3(0.f, 0.f, 0.f);
    V3 n_offset = v3(0.f, 0.f, 1.f);
    bool sms_success = false;
    V3 sms_weight = v3(0.f, 0.f, 0.f);
    auto walk = [&](uint64_t k, V2 *uvf) {
        Pcg32 rw; stream_seed(&rw, seed, qid, k + 1);
        V2 s;
        if (gq) {
            /* learned proposal: the footprint sample from the defensive mixture; a mixture sample
               outside the footprint is a failed walk with 0 iterations */
            if (proposal_seed(gq, defensive, &rw, &s) == 1 && !(s.x >= 0.f && s.x < 1.f && s.y >= 0.f && s.y < 1.f)) {
                st.walks++;
                return false;
            }
        } else {
            s.x = pcg32_float(&rw); s.y = pcg32_float(&rw);
        }
        V2 uvi; int it;
        bool ok = sample_glint(S, sensor_p, ei.p, si, s, n_offset, v2(1.f, 1.f), uvf, &uvi, &it);
        st.walks++; st.iters += it; if (ok) st.conv++;
        return ok;
    };
    if (!cfg.biased) {
        V2 uv_final;
        if (walk(0, &uv_final)) {
            sms_success = true;
            V3 specular_val = glint_contribution(S, sensor_p, ei, si, uv_final);
            float inv_prob = 1.f;
            int iterations = 1;
            uint64_t k = 1;
            st.bcalls++;
            while (true) {
                V2 uvt;
                bool ok = walk(k++, &uvt);
                if (ok && norm2(uvt - uv_final) < cfg.uniqueness_threshold) break;
                inv_prob += 1.f;
                iterations++;
                if (cfg.max_trials > 0 && iterations > cfg.max_trials) { inv_prob = 0.f; break; }
            }
            st.btrials += iterations;
            if ((uint64_t) iterations > st.bmax) st.bmax = iterations;
            sms_weight = specular_val * inv_prob;
        }
    } else {
        std::vector<V2> sols;
        for (int m = 0; m < cfg.max_trials; ++m) {
            V2 uvf;
            if (!walk((uint64_t) m, &uvf)) continue;
            bool dup = false;
            for (size_t k = 0; k < sols.size(); ++k) if (norm2(uvf - sols[k]) < cfg.uniqueness_threshold) { dup = true; break; }
            if (dup) continue;
            sols.push_back(uvf);
            sms_weight = sms_weight + glint_contribution(S, sensor_p, ei, si, uvf);
        }
        sms_success = !sols.empty();
    }
    if (sms_success) {
        float We = sensor_we(S, sensor_p, si.p);
        /* MIS against the BSDF strategy (manifold_glints.cpp:155-179): p_sms is the emitter
           sampling density, p_bsdf the LEAN pdf of the emitter direction plus 1 */
        float p_sms = sms_emitter_pdf(S, ei, si.p);
        V3 wo_l = v3(dot(edir, si.s), dot(edir, si.t), dot(edir, si.n));
        float p_bsdf = lean_pdf(S, si, wo_l);
        p_bsdf += 1.f;
        float mis = mis_power(p_sms, p_bsdf);
        if (cfg.bsdf_strategy_only) mis = 0.f;
        float f = mis * We;
        result = result + v3(f * sms_weight.x * ei.weight.x, f * sms_weight.y * ei.weight.y, f * sms_weight.z * ei.weight.z);
    }
    return result;
}

/* Camera ray + compute_partials for one pixel sample of a glint frame
   (integrator.cpp:243-262, perspective.cpp:191-221, interaction.h:216-248). */
static void make_glint_query(const Scene &S, int64_t idx, int spp, uint64_t seed, sms_glint_query *q) {
    memset(q, 0, sizeof(*q));
    q->shape = -1;
    const Camera &c = S.cam;
    int64_t pixel = idx / spp;
    int px = (int) (pixel % c.w), py = (int) (pixel / c.w);
    q->pixel = (int32_t) pixel;
    Pcg32 r; stream_seed(&r, seed, (uint64_t) idx, 0xFFFFFFFFull);
    float sx = pcg32_float(&r), sy = pcg32_float(&r);
    V2 pos = v2(((float) px + sx) / (float) c.w, ((float) py + sy) / (float) c.h);
    V3 near_p = v3((1.f - 2.f * pos.x) * c.tan_half, (1.f - 2.f * pos.y) * c.rect_y, 1.f);
    V3 dloc = normalize(near_p);
    V3 d = to_world(c.left, c.up, c.dir, dloc);
    V3 dxl = normalize(near_p + c.dx), dyl = normalize(near_p + c.dy);
    V3 d_x = to_world(c.left, c.up, c.dir, dxl), d_y = to_world(c.left, c.up, c.dir, dyl);
    float scale = 1.f / sqrtf((float) spp);
    d_x = fmadd3(d_x - d, scale, d); d_y = fmadd3(d_y - d, scale, d);
    Ray ray = make_ray(c.o, d, 1e-2f * (1.f / dloc.z), 1e4f * (1.f / dloc.z));
    SI si = ray_intersect(S, ray);
    for (int a = 0; a < 3; ++a) (&q->sensor_p[0])[a] = (&c.o.x)[a];
    if (!si.valid) return;
    const Shape &sh = S.shapes[si.shape];
    if (!(sh.flags & SMS_FLAG_GLINTY)) return;
    /* compute_partials (interaction.h:216-248); o_x = o_y = o */
    float dd = dot(si.n, si.p);
    float t_x = (dd - dot(si.n, c.o)) / dot(si.n, d_x), t_y = (dd - dot(si.n, c.o)) / dot(si.n, d_y);
    V3 dp_dx = fmadd3(d_x, t_x, c.o) - si.p, dp_dy = fmadd3(d_y, t_y, c.o) - si.p;
    float a00 = dot(si.dp_du, si.dp_du), a01 = dot(si.dp_du, si.dp_dv), a11 = dot(si.dp_dv, si.dp_dv),
          inv_det = 1.f / (a00 * a11 - a01 * a01);
    float b0x = dot(si.dp_du, dp_dx), b1x = dot(si.dp_dv, dp_dx), b0y = dot(si.dp_du, dp_dy), b1y = dot(si.dp_dv, dp_dy);
    if (!isfinite(inv_det)) inv_det = 0.f;
    V2 duv_dx = v2(fmaf(a11, b0x, -(a01 * b1x)) * inv_det, fmaf(a00, b1x, -(a01 * b0x)) * inv_det);
    V2 duv_dy = v2(fmaf(a11, b0y, -(a01 * b1y)) * inv_det, fmaf(a00, b1y, -(a01 * b0y)) * inv_det);
    q->shape = si.shape;
    for (int a = 0; a < 3; ++a) {
        q->p[a] = (&si.p.x)[a]; q->n[a] = (&si.n.x)[a]; q->sh_n[a] = (&si.sh_n.x)[a]; q->sh_s[a] = (&si.sh_s.x)[a];
        q->sh_t[a] = (&si.sh_t.x)[a]; q->dp_du[a] = (&si.dp_du.x)[a]; q->dp_dv[a] = (&si.dp_dv.x)[a]; q->wi[a] = (&si.wi.x)[a];
    }
    q->uv[0] = si.uv.x; q->uv[1] = si.uv.y;
    q->duv_dx[0] = duv_dx.x; q->duv_dx[1] = duv_dx.y; q->duv_dy[0] = duv_dy.x; q->duv_dy[1] = duv_dy.y;
}

/* Camera pass of the single-scatter integrator (path_sms_ss.cpp:94-251), SMS term only: the
   pixel sample's ray (same camera model and jitter stream as the glint queries) follows smooth
   dielectric interfaces -- reflection with probability F, else refraction
   (dielectric.cpp:201-313: weight 1 for reflection, eta_ti^2 for refraction) -- for up to
   SS_CAMERA_DEPTH vertices; the first caustic receiver that is not a caster becomes the query
   (path_sms_ss.cpp:118-124) with the path throughput as its weight.  Paths that leave the
   scene or stop on another surface give shape -1 and weight 0. */
static const int SS_CAMERA_DEPTH = 8;
static void make_ss_query(const Scene &S, int64_t idx, int spp, uint64_t seed, sms_ss_query *q, float *w) {
    memset(q, 0, sizeof(*q));
    q->shape = -1;
    *w = 0.f;
    const Camera &c = S.cam;
    int64_t pixel = idx / spp;
    int px = (int) (pixel % c.w), py = (int) (pixel / c.w);
    Pcg32 r; stream_seed(&r, seed, (uint64_t) idx, 0xFFFFFFFFull);
    float sx = pcg32_float(&r), sy = pcg32_float(&r);
    V2 pos = v2(((float) px + sx) / (float) c.w, ((float) py + sy) / (float) c.h);
    V3 near_p = v3((1.f - 2.f * pos.x) * c.tan_half, (1.f - 2.f * pos.y) * c.rect_y, 1.f);
    V3 dloc = normalize(near_p);
    V3 d = to_world(c.left, c.up, c.dir, dloc);
    Ray ray = make_ray(c.o, d, 1e-2f * (1.f / dloc.z), 1e4f * (1.f / dloc.z));
    float throughput = 1.f;
    for (int depth = 0; depth < SS_CAMERA_DEPTH; ++depth) {
        SI si = ray_intersect(S, ray);
        if (!si.valid) return;
        const Shape &sh = S.shapes[si.shape];
        V3 wi = v3(-dot(ray.d, si.sh_s), -dot(ray.d, si.sh_t), -dot(ray.d, si.sh_n));
        if ((sh.flags & SMS_FLAG_CAUSTIC_RECEIVER) && !(sh.flags & SMS_FLAG_CAUSTIC_CASTER_SS)) {
            q->shape = si.shape;
            for (int a = 0; a < 3; ++a) {
                q->p[a] = (&si.p.x)[a]; q->n[a] = (&si.n.x)[a]; q->sh_n[a] = (&si.sh_n.x)[a];
                q->sh_s[a] = (&si.sh_s.x)[a]; q->sh_t[a] = (&si.sh_t.x)[a]; q->wi[a] = (&wi.x)[a];
            }
            *w = throughput;
            return;
        }
        if (sh.bsdf != SMS_BSDF_DIELECTRIC) return;
        /* fresnel (fresnel.h:34-80) with the transmitted cosine and relative IOR */
        const float cos_i = wi.z, eta = sh.eta_r;
        const bool outside = cos_i >= 0.f;
        const float rcp_eta = 1.f / eta, eta_ti = outside ? rcp_eta : eta;
        const float cos_t_sqr = fmaf(-fmaf(-cos_i, cos_i, 1.f), eta_ti * eta_ti, 1.f);
        const float cos_t_abs = safe_sqrt(cos_t_sqr);
        const float F = fresnel(cos_i, eta);
        V3 wo;
        if (pcg32_float(&r) <= F) {
            wo = v3(-wi.x, -wi.y, wi.z);
        } else {
            const float cos_t = outside ? -cos_t_abs : cos_t_abs;
            wo = v3(-eta_ti * wi.x, -eta_ti * wi.y, cos_t);
            throughput = throughput * (eta_ti * eta_ti);
        }
        V3 dw = to_world(si.sh_s, si.sh_t, si.sh_n, wo);
        ray = spawn_ray(si.p, dw);
    }
}

template <typename F>
static void parallel_for(int64_t n, int threads, F f) {
    if (threads <= 1 || n < 64) { for (int64_t i = 0; i < n; ++i) f(i); return; }
    std::atomic<int64_t> next(0);
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&]() {
            while (true) {
                int64_t b = next.fetch_add(64);
                if (b >= n) break;
                int64_t e = std::min(n, b + 64);
                for (int64_t i = b; i < e; ++i) f(i);
            }
        });
    for (auto &th : pool) th.join();
}

static void stats_out(const Stats &s, sms_stats *o) {
    if (!o) return;
    memset(o, 0, sizeof(*o));
    o->walks = s.walks; o->walks_converged = s.conv; o->newton_iterations = s.iters;
    o->bernoulli_calls = s.bcalls; o->bernoulli_trials = s.btrials; o->bernoulli_max = s.bmax;
}

/* ================================================================ multi scatter */
static const int MS_MAXB = 4;

struct MSVert { MV v; V2 C; M2 dprev, dcur, dnext, tmp, inv_lambda; V2 dx; };
struct MSPath { int n; MSVert v[MS_MAXB]; int shapes[MS_MAXB]; };

static inline V3 offset_normal_world(const MV &v) { return (v.s * 0.f + v.t * 0.f) + v.n * 1.f; }

/* sample_seed_path (manifold_ms.cpp:280-408) with smooth interfaces (offset normal (0,0,1)).
   The reference's bounce-0 target-shape test compares the shape with itself (:332) and so
   never rejects; this restatement keeps that behaviour. */
static bool ms_seed_path(const Scene &S, int shape, V3 si_p, const EI &ei, V2 sample, bool first_path,
                         const int *seed_shapes, V3 p_start, MSPath &P) {
    P.n = 0;
    V3 x0 = si_p;
    V3 x1 = sample_position(S, shape, sample).p;
    if (p_start.x != 0.f || p_start.y != 0.f || p_start.z != 0.f) x1 = p_start;
    V3 wo = normalize(x1 - x0);
    if (S.cfg.mnee_init) {
        const Shape &sh = S.shapes[shape];
        wo = sh.k != 0.f ? normalize(sh.bbox_center - x0) : normalize(ei.p * g_plant[PLANT_MNEE_TARGET] - x0);
    }
    Ray ray = make_ray(x0, wo);
    while (true) {
        int bounce = P.n;
        if (bounce >= S.cfg.bounces) break;
        SI si = ray_intersect(S, ray);
        if (!si.valid) return false;
        int flags = S.shapes[si.shape].flags;
        if (!(flags & SMS_FLAG_CAUSTIC_CASTER_MS) && !(flags & SMS_FLAG_CAUSTIC_BOUNCER)) return false;
        if (!first_path && si.shape != seed_shapes[bounce]) return false;
        MV vertex = mv_from_si(S, si);
        P.v[bounce].v = vertex;
        P.shapes[bounce] = si.shape;
        P.n++;
        if (!S.cfg.mnee_init) {
            V3 m = offset_normal_world(vertex);
            V3 wi = -wo;
            if (vertex.eta == 1.f) wo = reflect(wi, m);
            else if (!refract(wi, m, vertex.eta, &wo)) return false;
        }
        ray = spawn_ray(si.p, wo);
    }
    return true;
}

static inline bool invert_exact(M2 a, M2 *inv) {
    if (fabsf(det2(a)) == 0.f) return false;
    *inv = inverse2(a);
    return true;
}

/* invert_tridiagonal_step (manifold_ms.cpp:1192-1232): block Thomas elimination */
static bool ms_invert_tridiagonal_step(MSPath &P) {
    int n = P.n;
    if (n == 0) return true;
    P.v[0].tmp = P.v[0].dprev;
    if (!invert_exact(P.v[0].dcur, &P.v[0].inv_lambda)) return false;
    for (int i = 1; i < n; ++i) {
        P.v[i].tmp = mulm2(P.v[i].dprev, P.v[i - 1].inv_lambda);
        M2 m = subm2(P.v[i].dcur, mulm2(P.v[i].tmp, P.v[i - 1].dnext));
        if (!invert_exact(m, &P.v[i].inv_lambda)) return false;
    }
    P.v[0].dx = P.v[0].C;
    for (int i = 1; i < n; ++i) P.v[i].dx = P.v[i].C - mulv2(P.v[i].tmp, P.v[i - 1].dx);
    P.v[n - 1].dx = mulv2(P.v[n - 1].inv_lambda, P.v[n - 1].dx);
    for (int i = n - 2; i >= 0; --i) P.v[i].dx = mulv2(P.v[i].inv_lambda, P.v[i].dx - mulv2(P.v[i].dnext, P.v[i + 1].dx));
    return true;
}

static inline float wrap_pi(float dp) {
    if (dp < -PI_F) dp += 2.f * PI_F; else if (dp > PI_F) dp -= 2.f * PI_F;
    return dp;
}
static inline bool transform_w(V3 w, V3 n, float eta, V3 *out) {
    if (eta == 1.f) { *out = reflect(w, n); return true; }
    return refract(w, n, eta, out);
}
static inline void d_transform_w(V3 w, V3 du, V3 dv, V3 n, V3 dn_du, V3 dn_dv, float eta, V3 *ou, V3 *ov) {
    if (eta == 1.f) d_reflect(w, du, dv, n, dn_du, dn_dv, ou, ov);
    else d_refract(w, du, dv, n, dn_du, dn_dv, eta, ou, ov);
}

/* compute_step_anglediff MS (manifold_ms.cpp:631-857) */
static bool ms_step_anglediff(MSPath &P, V3 x0, const EI &ei) {
    int k = P.n;
    bool success = true;
    const V3 zero = v3(0.f, 0.f, 0.f);
    for (int i = 0; i < k; ++i) {
        MSVert &vi = P.v[i];
        const MV &v = vi.v;
        vi.C = v2(0.f, 0.f);
        vi.dprev = vi.dcur = vi.dnext = m2(0.f, 0.f, 0.f, 0.f);
        V3 x_prev = i == 0 ? x0 : P.v[i - 1].v.p;
        V3 x_next = i == k - 1 ? ei.p : P.v[i + 1].v.p;
        V3 x_cur = v.p;
        bool fixed = i == k - 1 && ei_is_directional(ei);
        V3 wo = fixed ? ei.d : x_next - x_cur;
        float ilo = norm(wo);
        if (ilo < 1e-3f) return false;
        ilo = 1.f / ilo;
        wo = wo * ilo;
        V3 dwo_du_cur, dwo_dv_cur;
        if (fixed) { dwo_du_cur = zero; dwo_dv_cur = zero; }
        else {
            dwo_du_cur = -ilo * (v.dp_du - wo * dot(wo, v.dp_du));
            dwo_dv_cur = -ilo * (v.dp_dv - wo * dot(wo, v.dp_dv));
        }
        V3 wi = x_prev - x_cur;
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
        if (transform_w(wi, n, v.eta, &wio)) {
            float to, po, tio, pio;
            sphcoords(wo, &to, &po); sphcoords(wio, &tio, &pio);
            vi.C = v2(to - tio, wrap_pi(po - pio));
            float dto_du, dpo_du, dto_dv, dpo_dv, dtio_du, dpio_du, dtio_dv, dpio_dv;
            if (i > 0) {
                const MV &vp = P.v[i - 1].v;
                V3 dwi_du_prev = ili * (vp.dp_du - wi * dot(wi, vp.dp_du)), dwi_dv_prev = ili * (vp.dp_dv - wi * dot(wi, vp.dp_dv));
                V3 a, b;
                d_transform_w(wi, dwi_du_prev, dwi_dv_prev, n, zero, zero, v.eta, &a, &b);
                d_sphcoords(wio, a, b, &dtio_du, &dpio_du, &dtio_dv, &dpio_dv);
                vi.dprev = m2(-dtio_du, -dtio_dv, -dpio_du, -dpio_dv);
            }
            V3 a, b;
            d_transform_w(wi, dwi_du_cur, dwi_dv_cur, n, dn_du, dn_dv, v.eta, &a, &b);
            d_sphcoords(wo, dwo_du_cur, dwo_dv_cur, &dto_du, &dpo_du, &dto_dv, &dpo_dv);
            d_sphcoords(wio, a, b, &dtio_du, &dpio_du, &dtio_dv, &dpio_dv);
            vi.dcur = m2(dto_du - dtio_du, dto_dv - dtio_dv, dpo_du - dpio_du, dpo_dv - dpio_dv);
            if (i < k - 1) {
                const MV &vn = P.v[i + 1].v;
                V3 dwo_du_next = ilo * (vn.dp_du - wo * dot(wo, vn.dp_du)), dwo_dv_next = ilo * (vn.dp_dv - wo * dot(wo, vn.dp_dv));
                d_sphcoords(wo, dwo_du_next, dwo_dv_next, &dto_du, &dpo_du, &dto_dv, &dpo_dv);
                vi.dnext = m2(dto_du, dto_dv, dpo_du, dpo_dv);
            }
            success_i = true;
        }
        V3 woi;
        if (transform_w(wo, n, v.eta, &woi) && !success_i) {
            float ti, pi, toi, poi;
            sphcoords(wi, &ti, &pi); sphcoords(woi, &toi, &poi);
            vi.C = v2(ti - toi, wrap_pi(pi - poi));
            float dti_du, dpi_du, dti_dv, dpi_dv, dtoi_du, dpoi_du, dtoi_dv, dpoi_dv;
            if (i > 0) {
                const MV &vp = P.v[i - 1].v;
                V3 dwi_du_prev = ili * (vp.dp_du - wi * dot(wi, vp.dp_du)), dwi_dv_prev = ili * (vp.dp_dv - wi * dot(wi, vp.dp_dv));
                d_sphcoords(wi, dwi_du_prev, dwi_dv_prev, &dti_du, &dpi_du, &dti_dv, &dpi_dv);
                vi.dprev = m2(dti_du, dti_dv, dpi_du, dpi_dv);
            }
            V3 a, b;
            d_transform_w(wo, dwo_du_cur, dwo_dv_cur, n, dn_du, dn_dv, v.eta, &a, &b);
            d_sphcoords(wi, dwi_du_cur, dwi_dv_cur, &dti_du, &dpi_du, &dti_dv, &dpi_dv);
            d_sphcoords(woi, a, b, &dtoi_du, &dpoi_du, &dtoi_dv, &dpoi_dv);
            vi.dcur = m2(dti_du - dtoi_du, dti_dv - dtoi_dv, dpi_du - dpoi_du, dpi_dv - dpoi_dv);
            if (i < k - 1) {
                const MV &vn = P.v[i + 1].v;
                V3 dwo_du_next = ilo * (vn.dp_du - wo * dot(wo, vn.dp_du)), dwo_dv_next = ilo * (vn.dp_dv - wo * dot(wo, vn.dp_dv));
                V3 c, d;
                d_transform_w(wo, dwo_du_next, dwo_dv_next, n, zero, zero, v.eta, &c, &d);
                d_sphcoords(woi, c, d, &dtoi_du, &dpoi_du, &dtoi_dv, &dpoi_dv);
                vi.dnext = m2(-dtoi_du, -dtoi_dv, -dpoi_du, -dpoi_dv);
            }
            success_i = true;
        }
        vi.dcur.m00 = vi.dcur.m00 * g_plant[PLANT_MS_AD_JAC];
        success = success && success_i;
    }
    if (!success || !ms_invert_tridiagonal_step(P)) return false;
    return true;
}

/* compute_step_halfvector MS (manifold_ms.cpp:500-629), smooth interfaces: offset normal
   (0,0,1), so the constraint is the tangential part of the generalized half-vector */
static bool ms_step_halfvector(MSPath &P, V3 x0, const EI &ei) {
    int k = P.n;
    for (int i = 0; i < k; ++i) {
        MSVert &vi = P.v[i];
        const MV &v = vi.v;
        vi.C = v2(0.f, 0.f);
        vi.dprev = vi.dcur = vi.dnext = m2(0.f, 0.f, 0.f, 0.f);
        V3 x_prev = i == 0 ? x0 : P.v[i - 1].v.p;
        V3 x_next = i == k - 1 ? ei.p : P.v[i + 1].v.p;
        bool fixed = i == k - 1 && ei_is_directional(ei);
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
        vi.dcur.m00 = vi.dcur.m00 * g_plant[PLANT_MS_HV_JAC];
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

/* reproject (manifold_ms.cpp:859-918): re-trace the chain through the proposed first vertex */
static bool ms_reproject(const Scene &S, V3 si_p, V3 x1, const int *seed_shapes, int n_seed, MSPath &Q) {
    Q.n = 0;
    V3 wo = normalize(x1 - si_p);
    Ray ray = make_ray(si_p, wo);
    while (true) {
        int bounce = Q.n;
        if (bounce >= S.cfg.bounces) break;
        SI si = ray_intersect(S, ray);
        if (!si.valid) return false;
        if (si.shape != seed_shapes[bounce]) return false;
        MV vertex = mv_from_si(S, si);
        Q.v[bounce].v = vertex;
        Q.shapes[bounce] = si.shape;
        Q.n++;
        V3 m = offset_normal_world(vertex);
        V3 wi = -wo;
        if (vertex.eta == 1.f) wo = reflect(wi, m);
        else if (!refract(wi, m, vertex.eta, &wo)) return false;
        ray = spawn_ray(si.p, wo);
    }
    return Q.n == n_seed;
}

struct MSWalk { bool ok, seed_ok; int iterations; MSPath path; int seed_shapes[MS_MAXB]; };

/* sample_path MS (manifold_ms.cpp:203-227) + newton_solver (:410-498).  With first_path the
   seed chain's shapes become the template later reprojections must hit. */
static MSWalk ms_sample_path(const Scene &S, int shape, V3 si_p, const EI &ei, V2 sample, bool first_path,
                             const int *seed_shapes_in, V3 p_start) {
    const sms_config &cfg = S.cfg;
    MSWalk W;
    W.ok = false; W.seed_ok = false; W.iterations = 0;
    for (int i = 0; i < MS_MAXB; ++i) W.seed_shapes[i] = seed_shapes_in ? seed_shapes_in[i] : -1;
    if (!ms_seed_path(S, shape, si_p, ei, sample, first_path, W.seed_shapes, p_start, W.path)) return W;
    W.seed_ok = true;
    if (first_path) for (int i = 0; i < W.path.n; ++i) W.seed_shapes[i] = W.path.shapes[i];
    int n_seed = W.path.n;
    bool success = false, needs_step_update = true;
    int iterations = 0;
    float beta = 1.f;
    MSPath Q;
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
        V3 prop0 = W.path.v[0].v.p - cfg.step_scale * beta * (W.path.v[0].v.dp_du * W.path.v[0].dx.x +
                                                              W.path.v[0].v.dp_dv * W.path.v[0].dx.y);
        if (!ms_reproject(S, si_p, prop0, W.seed_shapes, n_seed, Q)) {
            beta = 0.5f * beta;
            needs_step_update = false;
        } else {
            beta = std::min(1.f, 2.f * beta);
            W.path = Q;
            needs_step_update = true;
        }
        iterations++;
    }
    W.iterations = iterations;
    if (!success) return W;
    int n = W.path.n;
    for (int i = 0; i < n; ++i) {
        V3 x_prev = i == 0 ? si_p : W.path.v[i - 1].v.p;
        V3 x_next = i == n - 1 ? ei.p : W.path.v[i + 1].v.p;
        V3 x_cur = W.path.v[i].v.p;
        bool fixed = i == n - 1 && ei_is_directional(ei);
        V3 wi = normalize(x_prev - x_cur);
        V3 wo = fixed ? ei.d : normalize(x_next - x_cur);
        float ci = dot(W.path.v[i].v.gn, wi), co = dot(W.path.v[i].v.gn, wo);
        bool refraction = ci * co < 0.f, reflection = !refraction;
        float eta = W.path.v[i].v.eta;
        if ((eta == 1.f && !reflection) || (eta != 1.f && !refraction)) return W;
    }
    W.ok = true;
    return W;
}

/* specular_reflectance MS, smooth interfaces (manifold_ms.cpp:920-1023), with the final
   visibility test towards the light */
static V3 ms_specular_reflectance(const Scene &S, V3 si_p, const EI &ei, const MSPath &P) {
    if (P.n == 0) return v3(0.f, 0.f, 0.f);
    V3 wo = normalize(P.v[0].v.p - si_p);
    Ray ray = make_ray(si_p, wo);
    float val = 1.f;
    SI si = si_invalid();
    for (int k = 0; k < P.n; ++k) {
        si = ray_intersect(S, ray);
        if (!si.valid) return v3(0.f, 0.f, 0.f);
        const Shape &sh = S.shapes[si.shape];
        V3 w;
        if (k < P.n - 1) w = normalize(P.v[k + 1].v.p - si.p);
        else w = ei_is_directional(ei) ? ei.d : normalize(ei.p - si.p);
        float cos_theta = dot(si.sh_n, w);
        float f;
        if (sh.k != 0.f) f = fresnel_conductor(fabsf(cos_theta), sh.eta_r, sh.k);
        else {
            float eta = hmean_same(sh.eta_r);
            if (cos_theta < 0.f) eta = 1.f / eta;
            f = (1.f - fresnel(cos_theta, eta)) * (eta * eta);
        }
        val = val * f;
        ray = make_ray(si.p, w);
    }
    if (!ei_is_directional(ei)) {
        V3 d = ei.p - si.p;
        ray = make_ray(si.p, normalize(d), RAY_EPS_F * (1.f + hmax_abs(si.p)), norm(d) * (1.f - RAY_EPS_F));
    }
    if (ray_test(S, ray)) return v3(0.f, 0.f, 0.f);
    return v3(val, val, val);
}

/* invert_tridiagonal_geo (manifold_ms.cpp:1234-1268) */
static float ms_invert_tridiagonal_geo(MSVert *v, int n) {
    if (n == 0) return 0.f;
    M2 Li;
    if (!invert_exact(v[0].dcur, &Li)) return 0.f;
    for (int i = 0; i < n - 2; ++i) {
        v[i].tmp = mulm2(Li, v[i].dnext);
        M2 m = subm2(v[i + 1].dcur, mulm2(v[i + 1].dprev, v[i].tmp));
        if (!invert_exact(m, &Li)) return 0.f;
    }
    v[n - 2].inv_lambda = mulm2(negm2(Li), v[n - 2].dnext);
    for (int i = n - 3; i >= 0; --i) v[i].inv_lambda = mulm2(negm2(v[i].tmp), v[i + 1].inv_lambda);
    return fabsf(det2(negm2(v[0].inv_lambda)));
}

/* geometric_term MS (manifold_ms.cpp:1025-1190) */
static float ms_geometric_term(const MV &vx, const MV &vy, const MSPath &P) {
    MSVert v[MS_MAXB + 2];
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
        const MV &c = v[i].v, &nx = v[i + 1].v;
        V3 wo = nx.p - c.p;
        float ilo = norm(wo);
        if (ilo < 1e-3f) return 0.f;
        ilo = 1.f / ilo;
        wo = wo * ilo;
        if (c.fixed_direction) {
            V3 dc_du_cur = -ilo * (c.dp_du - wo * dot(wo, c.dp_du)), dc_dv_cur = -ilo * (c.dp_dv - wo * dot(wo, c.dp_dv));
            v[i].dcur = m2(dot(dc_du_cur, c.dp_du), dot(dc_dv_cur, c.dp_du), dot(dc_du_cur, c.dp_dv), dot(dc_dv_cur, c.dp_dv));
            V3 dc_du_next = ilo * (nx.dp_du - wo * dot(wo, nx.dp_du)), dc_dv_next = ilo * (nx.dp_dv - wo * dot(wo, nx.dp_dv));
            v[i].dnext = m2(dot(dc_du_next, c.dp_du), dot(dc_dv_next, c.dp_du), dot(dc_du_next, c.dp_dv), dot(dc_dv_next, c.dp_dv));
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
            const MV &pv = v[i - 1].v;
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
    float dx1_dxend = ms_invert_tridiagonal_geo(v, k) * g_plant[PLANT_MS_G];
    V3 d = vx.p - v[0].v.p;
    float inv_r2 = 1.f / sqnorm(d);
    d = d * sqrtf(inv_r2);
    float dw0_dx1 = fabsf(dot(d, v[0].v.gn)) * inv_r2;
    return dw0_dx1 * dx1_dxend;
}

/* evaluate_path_contribution MS (manifold_ms.cpp:229-278) */
static V3 ms_path_contribution(const Scene &S, const SI &si, const EI &ei_, MSPath P) {
    const MV &last = P.v[P.n - 1].v;
    EI ei = ei_;
    if (ei.kind == SMS_EI_POINT) {
        const Emitter &em = S.emitters[ei.emitter];
        V3 d = em.position - last.p;
        float dist = norm(d), inv_dist = 1.f / dist;
        d = d * inv_dist;
        float emitter_pdf = ei.pdf;
        ei.p = em.position; ei.d = d; ei.n = d; ei.pdf = 1.f;
        ei.weight = (em.intensity * (inv_dist * inv_dist)) * dist * dist;
        ei.pdf *= emitter_pdf;
        ei.weight = ei.weight * (1.f / emitter_pdf);
    }
    MV vy;
    if (!emitter_interaction_to_vertex(S, ei, last.p, &vy)) return v3(0.f, 0.f, 0.f);
    for (int k = 0; k < P.n; ++k) make_orthonormal(P.v[k].v);
    MV vx = mv_from_si(S, si);
    make_orthonormal(vx);
    V3 f = ms_specular_reflectance(S, si.p, ei, P);
    float G = ms_geometric_term(vx, vy, P);
    V3 thr = f * G;
    return v3(thr.x * ei.weight.x, thr.y * ei.weight.y, thr.z * ei.weight.z);
}

/* specular_manifold_sampling MS (manifold_ms.cpp:27-201), per-(query, walk) PCG32 streams */
/* specular_manifold_sampling MS (manifold_ms.cpp:28-201).  With a proposal on a caster, the
   unbiased estimator's walks on it seed the chain through the defensive mixture's point
   (uniform branch: the reference's area sample); the biased estimator keeps uniform seeds. */
static V3 estimate_ms(const Scene &S, const sms_ss_query &q, uint64_t seed, uint64_t qid, Stats &st,
                      const Prop *P = nullptr) {
    if (q.shape < 0) return v3(0.f, 0.f, 0.f);     /* no receiver: zero estimate, no walks */
    const sms_config &cfg = S.cfg;
    SI si = si_from_query(q);
    Pcg32 rq; stream_seed(&rq, seed, qid, 0);
    std::vector<int> em_ms;
    for (int i = 0; i < (int) S.emitters.size(); ++i) if (S.emitters[i].ms) em_ms.push_back(i);
    EI ei = sample_emitter_interaction(S, si.p, em_ms, &rq);
    V3 result = v3(0.f, 0.f, 0.f);
    std::vector<int> casters;
    for (int i = 0; i < (int) S.shapes.size(); ++i) if (S.shapes[i].flags & SMS_FLAG_CAUSTIC_CASTER_MS) casters.push_back(i);
    for (size_t c = 0; c < casters.size(); ++c) {
        int shape = casters[c];
        uint64_t wbase = (uint64_t) c << 32;
        V3 value = v3(0.f, 0.f, 0.f);
        float gq[PROP_RAW];
        const bool use_prop = P && shape == P->shape && !cfg.biased;
        if (use_prop) gmm_prepare(P->raw + ((int64_t) qid - P->raw_base) * PROP_RAW, gq);
        auto run = [&](uint64_t k, bool first, const int *shapes) {
            Pcg32 rw; stream_seed(&rw, seed, qid, wbase + k + 1);
            V2 s = v2(0.f, 0.f);
            V3 p_start = v3(0.f, 0.f, 0.f);
            bool seeded = true;
            if (use_prop) {
                if (proposal_seed(gq, P->defensive, &rw, &s) == 1) {
                    seeded = s.x >= 0.f && s.x < 1.f && s.y >= 0.f && s.y < 1.f && uv_to_point(S, P->grid, s, &p_start);
                    s = v2(0.f, 0.f);
                }
            } else {
                s.x = pcg32_float(&rw); s.y = pcg32_float(&rw);
            }
            MSWalk w;
            if (seeded) {
                w = ms_sample_path(S, shape, si.p, ei, s, first, shapes, p_start);
            } else {                                   /* mixture uv outside the caster: failed, 0 iterations */
                w.ok = false; w.seed_ok = false; w.iterations = 0; w.path.n = 0;
                for (int i = 0; i < MS_MAXB; ++i) w.seed_shapes[i] = shapes ? shapes[i] : -1;
            }
            st.walks++; st.iters += w.iterations; if (w.ok) st.conv++;
            return w;
        };
        if (!cfg.biased) {
            MSWalk w0 = run(0, true, nullptr);
            if (!w0.ok) continue;
            V3 direction = normalize(w0.path.v[0].v.p - si.p);
            V3 spec = ms_path_contribution(S, si, ei, w0.path);
            V3 dl = v3(dot(direction, si.sh_s), dot(direction, si.sh_t), dot(direction, si.sh_n));
            V3 bsdf = receiver_bsdf_eval(S, si, dl);
            float inv_prob = 1.f;
            int iterations = 1;
            uint64_t k = 1;
            st.bcalls++;
            while (true) {
                MSWalk t = run(k++, false, w0.seed_shapes);
                V3 dtr = t.path.n > 0 ? normalize(t.path.v[0].v.p - si.p) : v3(0.f, 0.f, 0.f);
                if (t.ok && fabsf(dot(direction, dtr) - 1.f) < cfg.uniqueness_threshold) break;
                inv_prob += 1.f;
                iterations++;
                if (cfg.max_trials > 0 && iterations > cfg.max_trials) { inv_prob = 0.f; break; }
            }
            st.btrials += iterations;
            if ((uint64_t) iterations > st.bmax) st.bmax = iterations;
            value = v3(bsdf.x * spec.x * inv_prob, bsdf.y * spec.y * inv_prob, bsdf.z * spec.z * inv_prob);
        } else {
            std::vector<V3> solutions;
            bool have_seed = false;
            int seed_shapes[MS_MAXB];
            for (int m = 0; m < cfg.max_trials; ++m) {
                MSWalk w = run((uint64_t) m, !have_seed, have_seed ? seed_shapes : nullptr);
                if (!have_seed && w.seed_ok) {
                    have_seed = true;
                    for (int i = 0; i < MS_MAXB; ++i) seed_shapes[i] = w.seed_shapes[i];
                }
                if (!w.ok) continue;
                V3 direction = normalize(w.path.v[0].v.p - si.p);
                bool dup = false;
                for (size_t j = 0; j < solutions.size(); ++j)
                    if (fabsf(dot(direction, solutions[j]) - 1.f) < cfg.uniqueness_threshold) { dup = true; break; }
                if (dup) continue;
                solutions.push_back(direction);
                V3 dl = v3(dot(direction, si.sh_s), dot(direction, si.sh_t), dot(direction, si.sh_n));
                V3 bsdf = receiver_bsdf_eval(S, si, dl);
                V3 spec = ms_path_contribution(S, si, ei, w.path);
                value = value + v3(bsdf.x * spec.x, bsdf.y * spec.y, bsdf.z * spec.z);
            }
        }
        result = result + value / 1.f;
    }
    return result;
}

/* ================================================================ path integrators
   path_sms_ss (src/integrators/path_sms_ss.cpp:69-251) and path_filtered_ss
   (src/integrators/path_filtered_ss.cpp:37-194) over diffuse, smooth dielectric and smooth
   conductor BSDFs.  A sample's numbers come from PCG32 stream PATH_STREAM of (seed, sample
   index): the film jitter, then per bounce the Russian-roulette sample (1D), the emitter
   sample (2D), the BSDF sample (1D + 2D) and, where path_sms_ss checks a BSDF-sampled caustic
   against MNEE, the walk's seed sample (2D, and the LEAN sample with twostage), in the
   reference's order.  The path's own terms are summed first, then its SMS estimates in depth
   order, then its MNEE completions. */
static const uint64_t PATH_STREAM = 0xFFFFFFFDull;

/* warp::square_to_uniform_disk_concentric (include/mitsuba/core/warp.h:54-90) */
static V2 square_to_uniform_disk_concentric(V2 s) {
    float x = fmaf(2.f, s.x, -1.f), y = fmaf(2.f, s.y, -1.f);
    bool is_zero = x == 0.f && y == 0.f, q13 = fabsf(x) < fabsf(y);
    float r = q13 ? y : x, rp = q13 ? x : y;
    float phi = .25f * PI_F * rp / r;
    if (q13) phi = .5f * PI_F - phi;
    if (is_zero) phi = 0.f;
    float sn, cs; c_sincosf(phi, &sn, &cs);
    return v2(r * cs, r * sn);
}

/* warp::square_to_cosine_hemisphere (warp.h:332-340) */
static V3 square_to_cosine_hemisphere(V2 s) {
    V2 p = square_to_uniform_disk_concentric(s);
    return v3(p.x, p.y, safe_sqrt(1.f - (p.x * p.x + p.y * p.y)));
}

struct DirSample { V3 p, n, d; float dist, pdf; bool delta; int emitter; };

static DirSample dir_sample_zero() {
    DirSample ds;
    ds.p = ds.n = ds.d = v3(0.f, 0.f, 0.f); ds.dist = 0.f; ds.pdf = 0.f; ds.delta = false; ds.emitter = -1;
    return ds;
}

/* Sphere::sample_direction (src/shapes/sphere.cpp:177-252): a uniform direction in the cone
   the sphere subtends seen from outside, else a uniform point on the sphere */
static DirSample sphere_sample_direction(const Shape &sh, V3 ref, V2 sample) {
    DirSample ds = dir_sample_zero();
    V3 dc_v = sh.center - ref;
    float dc_2 = dot(dc_v, dc_v);
    float radius_adj = sh.radius * (1.f - RAY_EPS_F);
    if (dc_2 > radius_adj * radius_adj) {
        float inv_dc = 1.f / sqrtf(dc_2), sin_theta_max = sh.radius * inv_dc,
              sin_theta_max_2 = sin_theta_max * sin_theta_max, inv_sin_theta_max = 1.f / sin_theta_max,
              cos_theta_max = safe_sqrt(1.f - sin_theta_max_2);
        float a = fmaf(cos_theta_max - 1.f, sample.x, 1.f);
        float sin_theta_2 = sin_theta_max_2 > 0.00068523f ? 1.f - a * a : sin_theta_max_2 * sample.x;
        float cos_theta = safe_sqrt(1.f - sin_theta_2);
        float cos_alpha = sin_theta_2 * inv_sin_theta_max +
                          cos_theta * safe_sqrt(fmaf(-sin_theta_2, inv_sin_theta_max * inv_sin_theta_max, 1.f));
        float sin_alpha = safe_sqrt(fmaf(-cos_alpha, cos_alpha, 1.f));
        float sin_phi, cos_phi; c_sincosf(sample.y * (2.f * PI_F), &sin_phi, &cos_phi);
        V3 fn = dc_v * -inv_dc, fs, ft;
        coordinate_system(fn, &fs, &ft);
        V3 d = to_world(fs, ft, fn, v3(cos_phi * sin_alpha, sin_phi * sin_alpha, cos_alpha));
        ds.p = fmadd3(d, sh.radius, sh.center);
        ds.n = d;
        ds.d = ds.p - ref;
        float dist2 = dot(ds.d, ds.d);
        ds.dist = sqrtf(dist2);
        ds.d = ds.d / ds.dist;
        ds.pdf = INV_TWOPI_F / (1.f - cos_theta_max);
        if (ds.dist == 0.f) ds.pdf = 0.f;
    } else {
        V3 d = square_to_uniform_sphere(sample);
        ds.p = fmadd3(d, sh.radius, sh.center);
        ds.n = d;
        ds.d = ds.p - ref;
        float dist2 = dot(ds.d, ds.d);
        ds.dist = sqrtf(dist2);
        ds.d = ds.d / ds.dist;
        ds.pdf = sh.inv_surface_area * dist2 / fabsf(dot(ds.d, ds.n));
    }
    ds.delta = sh.radius == 0.f;
    return ds;
}

/* Scene::sample_emitter_direction (src/librender/scene.cpp:156-203) with the visibility test,
   over AreaLight (area.cpp:104-117), PointLight (point.cpp:74-99) and DirectionalEmitter
   (directional.cpp:105-128) */
static DirSample scene_sample_emitter_direction(const Scene &S, V3 ref, V2 sample, bool test_visibility, V3 *spec_out) {
    DirSample ds = dir_sample_zero();
    V3 spec = v3(0.f, 0.f, 0.f);
    const int n = (int) S.emitters.size();
    *spec_out = spec;
    if (n == 0) return ds;
    int index = 0;
    float emitter_pdf = 1.f;
    if (n > 1) {
        emitter_pdf = 1.f / (float) n;
        uint32_t idx = (uint32_t) (sample.x * (float) n);
        index = (int) std::min(idx, (uint32_t) (n - 1));
        sample.x = (sample.x - (float) index * emitter_pdf) * (float) n;
    }
    const Emitter &em = S.emitters[(size_t) index];
    if (em.type == SMS_EMITTER_AREA) {
        const Shape &es = S.shapes[(size_t) em.shape];
        if (es.type == SMS_SHAPE_SPHERE) {
            ds = sphere_sample_direction(es, ref, sample);
        } else {
            /* Shape::sample_direction (shape.cpp:259-276): area sample, solid-angle conversion */
            PS ps = sample_position(S, em.shape, sample);
            ds.p = ps.p; ds.n = ps.n; ds.pdf = ps.pdf;
            ds.d = ds.p - ref;
            float dist2 = dot(ds.d, ds.d);
            ds.dist = sqrtf(dist2);
            ds.d = ds.d / ds.dist;
            float dp = fabsf(dot(ds.d, ds.n));
            if (dp != 0.f) ds.pdf = ds.pdf * (dist2 / dp);
            else ds.pdf = 0.f;
        }
        bool ok = dot(ds.d, ds.n) < 0.f && ds.pdf != 0.f;
        if (ok) spec = v3(em.radiance.x / ds.pdf, em.radiance.y / ds.pdf, em.radiance.z / ds.pdf);
    } else if (em.type == SMS_EMITTER_POINT) {
        ds.p = em.position; ds.pdf = 1.f; ds.delta = true;
        ds.d = ds.p - ref;
        ds.dist = norm(ds.d);
        float inv_dist = 1.f / ds.dist;
        ds.d = ds.d * inv_dist;
        spec = em.intensity * (inv_dist * inv_dist);
    } else {
        V3 d = em.direction;
        float dist = S.bs_dist;
        ds.p = ref - d * dist; ds.n = d; ds.pdf = 1.f; ds.delta = true; ds.d = d * -1.f; ds.dist = dist;
        spec = em.irradiance;
    }
    ds.emitter = index;
    if (n > 1) {
        ds.pdf = ds.pdf * emitter_pdf;
        spec = spec * (1.f / emitter_pdf);
    }
    if (test_visibility && ds.pdf != 0.f) {
        Ray r = make_ray(ref, ds.d, RAY_EPS_F * (1.f + hmax_abs(ref)), ds.dist * (1.f - SHADOW_EPS_F));
        if (ray_test(S, r)) spec = v3(0.f, 0.f, 0.f);
    }
    *spec_out = spec;
    return ds;
}

/* Scene::pdf_emitter_direction (scene.cpp:205-219) of a BSDF-sampled hit on an area emitter:
   AreaLight::pdf_direction (area.cpp:119-125) and Sphere::pdf_direction (sphere.cpp:254-267)
   with DirectionSample3f(si_hit, ref) (records.h:168-174) */
static float pdf_emitter_hit(const Scene &S, int emitter, const SI &hit, V3 ref) {
    const Emitter &em = S.emitters[(size_t) emitter];
    if (em.type != SMS_EMITTER_AREA) return 0.f;
    const Shape &sh = S.shapes[(size_t) em.shape];
    V3 d = hit.p - ref;
    float dist = norm(d);
    d = d / dist;
    if (!(dot(d, hit.sh_n) < 0.f)) return 0.f;
    if (sh.type != SMS_SHAPE_SPHERE) {
        /* Shape::pdf_direction (shape.cpp:278-289): pdf_position (rectangle: 1 / area,
           rectangle.cpp:130-133; mesh: the area distribution's normalisation, mesh.cpp:588-591)
           times dist^2 / |cos| */
        float dp = fabsf(dot(d, hit.sh_n));
        float pdf_pos = sh.type == SMS_SHAPE_RECTANGLE ? sh.inv_surface_area : sh.area_norm;
        float pdf = 0.f;
        if (dp != 0.f) pdf = pdf_pos * ((dist * dist) / dp);
        if (S.emitters.size() > 1) pdf = pdf * (1.f / (float) S.emitters.size());
        return pdf;
    }
    float sin_alpha = sh.radius * (1.f / norm(sh.center - ref));
    float cos_alpha = safe_sqrt(1.f - sin_alpha * sin_alpha);
    float pdf = sin_alpha < 0.99999994039535522461f ? INV_TWOPI_F / (1.f - cos_alpha)
                                                    : sh.inv_surface_area * (dist * dist) / fabsf(dot(d, hit.sh_n));
    if (S.emitters.size() > 1) pdf = pdf * (1.f / (float) S.emitters.size());
    return pdf;
}

/* AreaLight::eval (area.cpp:71-79) at a surface interaction on its shape */
static V3 emitter_eval_hit(const Scene &S, int emitter, const SI &hit) {
    const Emitter &em = S.emitters[(size_t) emitter];
    if (em.type != SMS_EMITTER_AREA || !(hit.wi.z > 0.f)) return v3(0.f, 0.f, 0.f);
    return em.radiance;
}

struct BsdfSample { V3 wo, weight; float pdf, eta; bool delta, reflection, transmission; };

/* BSDF::sample of SmoothDiffuse (diffuse.cpp:77-101), SmoothDielectric (dielectric.cpp:201-313)
   and SmoothConductor (conductor.cpp:217-275) in the local shading frame */
static BsdfSample bsdf_sample(const Shape &sh, V3 wi, float s1, V2 s2) {
    BsdfSample bs;
    bs.wo = v3(0.f, 0.f, 0.f); bs.weight = v3(0.f, 0.f, 0.f); bs.pdf = 0.f; bs.eta = 1.f;
    bs.delta = sh.bsdf != SMS_BSDF_DIFFUSE; bs.reflection = true; bs.transmission = false;
    const float cos_i = wi.z;
    if (sh.bsdf == SMS_BSDF_DIFFUSE) {
        if (!(cos_i > 0.f)) { bs.pdf = 0.f; return bs; }
        bs.wo = square_to_cosine_hemisphere(s2);
        bs.pdf = INV_PI_F * bs.wo.z;
        if (bs.pdf > 0.f) bs.weight = sh.refl;
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

/* NormalmapBSDF::frame (src/bsdfs/normalmap.cpp:218-239) at smoothing 0: the shading frame
   perturbed by the map, as ManifoldVertex builds it */
static void normalmap_frame(const Scene &S, const SI &si, V3 *n, V3 *s, V3 *t) {
    const Shape &sh = S.shapes[si.shape];
    V3 bn = si.sh_n;
    V3 bs = normalize(fnmadd3(bn, dot(bn, si.dp_du), si.dp_du));
    V3 bt = cross(bn, bs);
    const Normalmap &nm = S.nmaps[sh.normalmap];
    V3 nl, du, dv;
    nm_local_normal(nm, tile_uv(si.uv, sh.tiles), &nl, &du, &dv);
    *n = normalize(to_world(bs, bt, bn, nl));
    *s = normalize(si.dp_du - *n * dot(*n, si.dp_du));
    *t = cross(*n, *s);
}

/* NormalmapBSDF::sample (bsdfs/normalmap.cpp:117-134, without the LEAN fallback): the nested
   smooth dielectric (k = 0) or conductor sampled in the perturbed frame, the direction taken
   back to the shading frame, rejected when the two frames put it on different sides */
static BsdfSample bsdf_sample_normalmap(const Scene &S, const SI &si, float s1, V2 s2) {
    const Shape &sh = S.shapes[si.shape];
    Shape nested = sh;
    nested.bsdf = sh.k == 0.f ? SMS_BSDF_DIELECTRIC : SMS_BSDF_CONDUCTOR;
    V3 pn, ps, pt;
    normalmap_frame(S, si, &pn, &ps, &pt);
    V3 wi_w = to_world(si.sh_s, si.sh_t, si.sh_n, si.wi);
    V3 wi_p = v3(dot(wi_w, ps), dot(wi_w, pt), dot(wi_w, pn));
    BsdfSample bs = bsdf_sample(nested, wi_p, s1, s2);
    if (bs.weight.x == 0.f && bs.weight.y == 0.f && bs.weight.z == 0.f) return bs;
    V3 wo_w = to_world(ps, pt, pn, bs.wo);
    V3 wo = v3(dot(wo_w, si.sh_s), dot(wo_w, si.sh_t), dot(wo_w, si.sh_n));
    if (!(bs.wo.z * wo.z > 0.f)) bs.weight = v3(0.f, 0.f, 0.f);
    bs.wo = wo;
    return bs;
}

/* SmoothDiffuse::eval / pdf (diffuse.cpp:103-133); the delta BSDFs evaluate to zero */
static V3 bsdf_eval(const Shape &sh, V3 wi, V3 wo) {
    if (sh.bsdf != SMS_BSDF_DIFFUSE || !(wi.z > 0.f && wo.z > 0.f)) return v3(0.f, 0.f, 0.f);
    return v3((sh.refl.x * INV_PI_F) * wo.z, (sh.refl.y * INV_PI_F) * wo.z, (sh.refl.z * INV_PI_F) * wo.z);
}
static float bsdf_pdf(const Shape &sh, V3 wi, V3 wo) {
    if (sh.bsdf != SMS_BSDF_DIFFUSE || !(wi.z > 0.f && wo.z > 0.f)) return 0.f;
    return INV_PI_F * wo.z;
}

static inline V3 mul3(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }

/* the film position and primary ray of sample idx (pixel idx / spp, jitter from the path stream) */
static Ray path_camera_ray(const Scene &S, int64_t idx, int spp, Pcg32 *r, float *pos) {
    const Camera &c = S.cam;
    int64_t pixel = idx / spp;
    int px = (int) (pixel % c.w), py = (int) (pixel / c.w);
    float sx = pcg32_float(r), sy = pcg32_float(r);
    pos[0] = (float) px + sx; pos[1] = (float) py + sy;
    V2 p = v2(((float) px + sx) / (float) c.w, ((float) py + sy) / (float) c.h);
    V3 near_p = v3((1.f - 2.f * p.x) * c.tan_half, (1.f - 2.f * p.y) * c.rect_y, 1.f);
    V3 dloc = normalize(near_p);
    V3 d = to_world(c.left, c.up, c.dir, dloc);
    return make_ray(c.o, d, 1e-2f * (1.f / dloc.z), 1e4f * (1.f / dloc.z));
}

struct PathSms { sms_ss_query q; V3 thr; };
struct PathMnee { sms_ss_walk_in w; V3 value, dir_pt, ref_p; };
/* a multi-scatter MNEE check: the walk from the receiver and the chain the path tracer found */
struct PathMneeMs { sms_ms_walk_in w; V3 value; V3 pos[MS_MAXB]; int npos; };
struct PathSample {
    V3 L; float pos[2]; std::vector<PathSms> sms; std::vector<PathMnee> mnee; std::vector<PathMneeMs> mnee_ms;
};

/* The multi-scatter integrators' record of the specular chain the path is on: MNEEHelper
   (path_sms_ms.cpp:54-143: receiver, first caster and the chain's positions) and PathRecord
   (path_filtered_ms.cpp:237-300: the chain's scattering types); states Default, HitReceiver,
   HitCaster, HitEmitter.  depth_receiver: the path depth of the chain's receiver. */
struct ChainRecord {
    int state = 0, bounce = -1, shape0 = -1, npos = 0, depth_receiver = -1;
    V3 endpoint = v3(0.f, 0.f, 0.f), pos[MS_MAXB];
    void reset() { state = 0; bounce = -1; shape0 = -1; npos = 0; }
};

/* caustic_eta of BSDF::ior (path_filtered_ss.cpp:85-87, path_filtered_ms.cpp:268-270): a
   dielectric's (eta, 0); conductors (eta, k) and the default (0, 1) give 1 unless k = 0 */
static float caustic_eta_of(const Shape &sh) {
    if (sh.bsdf == SMS_BSDF_DIELECTRIC) return sh.eta_r;
    /* a normal-mapped BSDF gives its nested BSDF's (bsdfs/normalmap.cpp:342-344) */
    if ((sh.bsdf == SMS_BSDF_CONDUCTOR || sh.bsdf == SMS_BSDF_NORMALMAP) && sh.k == 0.f) return sh.eta_r;
    return 1.f;
}

/* MNEEHelper::state_transition (path_sms_ms.cpp:85-138) to the interaction `next` at `depth` */
static void mnee_transition(const Scene &S, ChainRecord &c, const SI &next, int depth) {
    const int bounces = S.cfg.bounces;
    const int flags = next.valid ? S.shapes[(size_t) next.shape].flags : 0;
    if (next.valid && (flags & SMS_FLAG_CAUSTIC_RECEIVER)) {
        c.reset(); c.state = 1; c.endpoint = next.p; c.depth_receiver = depth;
        return;
    }
    const int e = next.valid ? S.shapes[(size_t) next.shape].emitter : -1;
    if (c.state == 1) {
        if (next.valid && (flags & SMS_FLAG_CAUSTIC_CASTER_MS)) {
            c.shape0 = next.shape; c.pos[0] = next.p; c.npos = 1; c.bounce = 1; c.state = 2;
        } else c.reset();
    } else if (c.state == 2) {
        if (c.bounce < bounces && next.valid && (flags & (SMS_FLAG_CAUSTIC_CASTER_MS | SMS_FLAG_CAUSTIC_BOUNCER))) {
            c.pos[c.npos++] = next.p; c.bounce++;
        } else if (c.bounce == bounces && e >= 0 && S.emitters[(size_t) e].ms) {
            c.state = 3;
        } else c.reset();
    } else c.reset();
}

/* PathRecord::state_transition (path_filtered_ms.cpp:248-299) from `cur` (BSDF-sampled with
   reflection / transmission flags refl / trans) to `next` at `depth` */
static void record_transition(const Scene &S, ChainRecord &c, const SI &cur, const SI &next, bool refl, bool trans,
                              int depth) {
    const int bounces = S.cfg.bounces;
    const int flags = next.valid ? S.shapes[(size_t) next.shape].flags : 0;
    if (next.valid && (flags & SMS_FLAG_CAUSTIC_RECEIVER)) {
        c.reset(); c.state = 1; c.depth_receiver = depth;
        return;
    }
    const int e = next.valid ? S.shapes[(size_t) next.shape].emitter : -1;
    if (c.state == 1) {
        if (next.valid && (flags & SMS_FLAG_CAUSTIC_CASTER_MS)) { c.bounce = 1; c.state = 2; }
        else c.reset();
    } else if (c.state == 2) {
        if (c.bounce < bounces && next.valid && (flags & (SMS_FLAG_CAUSTIC_CASTER_MS | SMS_FLAG_CAUSTIC_BOUNCER))) {
            const float eta = caustic_eta_of(S.shapes[(size_t) cur.shape]);
            if ((eta == 1.f && trans) || (eta != 1.f && refl)) c.reset();
            else c.bounce++;
        } else if (c.bounce == bounces && e >= 0 && S.emitters[(size_t) e].ms) {
            c.state = 3;
        } else c.reset();
    } else c.reset();
}

static void sms_query_from_si(const SI &si, sms_ss_query *q) {
    memset(q, 0, sizeof(*q));
    q->shape = si.shape;
    for (int a = 0; a < 3; ++a) {
        q->p[a] = (&si.p.x)[a]; q->n[a] = (&si.n.x)[a]; q->sh_n[a] = (&si.sh_n.x)[a];
        q->sh_s[a] = (&si.sh_s.x)[a]; q->sh_t[a] = (&si.sh_t.x)[a]; q->wi[a] = (&si.wi.x)[a];
    }
}

/* one sample of path_sms_ss / path_filtered_ss / path_sms_ms / path_filtered_ms
   (SamplingIntegrator::sample) */
static void path_sample(const Scene &S, const sms_path_config &pc, uint64_t seed, int64_t idx, int spp, PathSample &o) {
    const bool ms = pc.integrator == SMS_PATH_SMS_MS || pc.integrator == SMS_PATH_FILTERED_MS;
    const bool sms = pc.integrator == SMS_PATH_SMS_SS || pc.integrator == SMS_PATH_SMS_MS, only = pc.caustics_only != 0;
    const int depth_cap = pc.max_depth < 0 ? SMS_PATH_MAX_DEPTH : pc.max_depth;
    const int chain = ms ? S.cfg.bounces : 1;           /* specular vertices between receiver and light */
    Pcg32 r; stream_seed(&r, seed, (uint64_t) idx, PATH_STREAM);
    o.sms.clear(); o.mnee.clear(); o.mnee_ms.clear();
    Ray ray = path_camera_ray(S, idx, spp, &r, o.pos);
    V3 L = v3(0.f, 0.f, 0.f), thr = v3(1.f, 1.f, 1.f);
    float eta = 1.f;
    SI si = ray_intersect(S, ray), si_prev = si_invalid();
    int first_diffuse = -1;
    bool specular_camera_path = true;
    ChainRecord rec;
    if (ms) {
        if (sms) mnee_transition(S, rec, si, 1);
        else record_transition(S, rec, si, si, false, false, 1);
    }
    if (si.valid && !only) {
        int e = S.shapes[(size_t) si.shape].emitter;
        if (e >= 0) L = L + emitter_eval_hit(S, e, si);
    }
    for (int depth = 1;; ++depth) {
        if (!si.valid) break;
        if (depth > pc.rr_depth) {
            float m = thr.x < thr.y ? thr.y : thr.x;          /* std::max / std::min order */
            m = m < thr.z ? thr.z : m;
            float q = m * (eta * eta);
            q = .95f < q ? (float) .95f : q;
            if (pcg32_float(&r) > q) break;
            thr = thr * (1.f / q);
        }
        if (depth >= depth_cap) break;
        if (only && first_diffuse >= 0 && depth > first_diffuse + chain) break;     /* nothing left to count */
        const Shape &sh = S.shapes[(size_t) si.shape];
        const bool smooth = sh.bsdf == SMS_BSDF_DIFFUSE;
        if (first_diffuse < 0 && smooth) first_diffuse = depth;
        const bool receiver = (sh.flags & SMS_FLAG_CAUSTIC_RECEIVER) != 0;
        const bool from_receiver = si_prev.valid && (S.shapes[(size_t) si_prev.shape].flags & SMS_FLAG_CAUSTIC_RECEIVER);
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
        /* SMS at a caustic receiver (path_sms_ss.cpp:118-124, path_sms_ms.cpp:228-234) */
        if (sms && receiver && !on_caster && (pc.max_depth < 0 || depth + chain < pc.max_depth) &&
            (!only || depth == first_diffuse)) {
            PathSms ps;
            sms_query_from_si(si, &ps.q);
            ps.thr = thr;
            o.sms.push_back(ps);
        }
        /* emitter sampling (path_sms_ss.cpp:142-161, path_filtered_ss.cpp:107-142,
           path_sms_ms.cpp:240-270, path_filtered_ms.cpp:97-140) */
        bool nee;
        bool valid_connection = false;
        if (sms) nee = (smooth && !on_caster) || glint_case;
        else {
            valid_connection = ms ? (rec.state == 2 && rec.bounce == S.cfg.bounces)
                                  : on_caster && (from_receiver || glint_case);
            nee = smooth && (!on_caster || valid_connection);
        }
        if (nee) {
            V2 s2; s2.x = pcg32_float(&r); s2.y = pcg32_float(&r);
            V3 ew;
            DirSample ds;
            V3 wo = v3(0.f, 0.f, 0.f), val = v3(0.f, 0.f, 0.f);
            /* a delta BSDF evaluates to zero whatever the visibility: no shadow ray */
            ds = scene_sample_emitter_direction(S, si.p, s2, smooth, &ew);
            if (ds.pdf != 0.f) {
                wo = v3(dot(ds.d, si.sh_s), dot(ds.d, si.sh_t), dot(ds.d, si.sh_n));
                val = bsdf_eval(sh, si.wi, wo);
                float bpdf = bsdf_pdf(sh, si.wi, wo);
                float mis = ds.delta ? (float) 1.f : mis_power(ds.pdf, bpdf);
                if (!sms && valid_connection && (ms || depth > 1)) {
                    bool refraction = wo.z * si.wi.z < 0.f, reflection = !refraction;
                    if ((caustic_eta == 1.f && !reflection) || (caustic_eta != 1.f && !refraction)) mis = 0.f;
                }
                if (!only) L = L + mul3(mul3(thr * mis, val), ew);
            }
        }
        /* BSDF sampling */
        float s1 = pcg32_float(&r);
        V2 s2; s2.x = pcg32_float(&r); s2.y = pcg32_float(&r);
        BsdfSample bs = sh.bsdf == SMS_BSDF_NORMALMAP ? bsdf_sample_normalmap(S, si, s1, s2)
                                                      : bsdf_sample(sh, si.wi, s1, s2);
        thr = mul3(thr, bs.weight);
        eta = eta * bs.eta;
        if (!bs.delta) specular_camera_path = false;
        if (thr.x == 0.f && thr.y == 0.f && thr.z == 0.f) break;
        ray = spawn_ray(si.p, to_world(si.sh_s, si.sh_t, si.sh_n, bs.wo));
        SI hit = ray_intersect(S, ray);
        if (ms) {
            if (sms) mnee_transition(S, rec, hit, depth + 1);
            else record_transition(S, rec, si, hit, bs.reflection, bs.transmission, depth + 1);
        }
        int e = hit.valid ? S.shapes[(size_t) hit.shape].emitter : -1;
        if (e >= 0) {
            if (ms && sms) {
                if (!on_caster || specular_camera_path) {
                    if (!only) {
                        V3 ev = emitter_eval_hit(S, e, hit);
                        float epdf = bs.delta ? (float) 0.f : pdf_emitter_hit(S, e, hit, si.p);
                        float mis = mis_power(bs.pdf, epdf);
                        L = L + mul3(thr * mis, ev);
                    }
                } else if (S.cfg.mnee_init && !pc.biased_mnee && rec.state == 3 && rec.bounce == S.cfg.bounces) {
                    /* MNEE check (path_sms_ms.cpp:290-329): the walk from the chain's receiver through
                       its first caster, compared vertex by vertex with the chain found */
                    PathMneeMs m;
                    memset(&m.w, 0, sizeof(m.w));
                    m.w.sample[0] = pcg32_float(&r); m.w.sample[1] = pcg32_float(&r);
                    if (!only || rec.depth_receiver == first_diffuse) {
                        for (int a = 0; a < 3; ++a) m.w.si_p[a] = (&rec.endpoint.x)[a];
                        m.w.shape = rec.shape0;
                        V3 ed = normalize(hit.p - rec.endpoint);
                        for (int a = 0; a < 3; ++a) {
                            m.w.ei.p[a] = (&hit.p.x)[a]; m.w.ei.n[a] = (&hit.n.x)[a]; m.w.ei.d[a] = (&ed.x)[a];
                        }
                        m.w.ei.kind = SMS_EI_AREA; m.w.ei.emitter = e;
                        m.w.first_path = 1;
                        m.value = mul3(thr, emitter_eval_hit(S, e, hit));
                        m.npos = rec.npos;
                        for (int k = 0; k < rec.npos; ++k) m.pos[k] = rec.pos[k];
                        o.mnee_ms.push_back(m);
                    }
                }
            } else if (ms) {
                const bool light_hit_valid = rec.state == 3 && rec.bounce == S.cfg.bounces;
                if (!on_caster || light_hit_valid || specular_camera_path) {
                    bool take = !only || (light_hit_valid && rec.depth_receiver == first_diffuse);
                    if (take) {
                        V3 ev = emitter_eval_hit(S, e, hit);
                        float epdf = bs.delta ? (float) 0.f : pdf_emitter_hit(S, e, hit, si.p);
                        float mis = mis_power(bs.pdf, epdf);
                        L = L + mul3(thr * mis, ev);
                    }
                }
            } else if (sms) {
                if (!on_caster || glint_case) {
                    if (!only) {
                        V3 ev = emitter_eval_hit(S, e, hit);
                        float epdf = bs.delta ? (float) 0.f : pdf_emitter_hit(S, e, hit, si.p);
                        float mis = mis_power(bs.pdf, epdf);
                        L = L + mul3(thr * mis, ev);
                    }
                } else if (S.cfg.mnee_init && !pc.biased_mnee && from_receiver && S.emitters[(size_t) e].ss) {
                    /* MNEE completion (path_sms_ss.cpp:196-241): the walk from the receiver */
                    PathMnee m;
                    memset(&m.w, 0, sizeof(m.w));
                    m.w.sample[0] = pcg32_float(&r); m.w.sample[1] = pcg32_float(&r);
                    if (S.cfg.twostage) { m.w.lean_sample[0] = pcg32_float(&r); m.w.lean_sample[1] = pcg32_float(&r); }
                    if (!only || depth - 1 == first_diffuse) {
                        for (int a = 0; a < 3; ++a) m.w.si_p[a] = (&si_prev.p.x)[a];
                        m.w.shape = si.shape;
                        V3 ed = normalize(hit.p - si_prev.p);
                        for (int a = 0; a < 3; ++a) {
                            m.w.ei.p[a] = (&hit.p.x)[a]; m.w.ei.n[a] = (&hit.n.x)[a]; m.w.ei.d[a] = (&ed.x)[a];
                        }
                        m.w.ei.kind = SMS_EI_AREA; m.w.ei.emitter = e;
                        m.w.n_offset[2] = 1.f;
                        m.value = mul3(thr, emitter_eval_hit(S, e, hit));
                        m.dir_pt = normalize(si.p - si_prev.p);
                        m.ref_p = si_prev.p;
                        o.mnee.push_back(m);
                    }
                }
            } else {
                bool r_t_valid = (caustic_eta == 1.f && bs.reflection) || (caustic_eta != 1.f && bs.transmission);
                bool vc = on_caster && ((from_receiver && r_t_valid) || glint_case);
                if (!on_caster || vc) {
                    bool take = !only || (on_caster && from_receiver && r_t_valid && depth - 1 == first_diffuse);
                    if (take) {
                        V3 ev = emitter_eval_hit(S, e, hit);
                        float epdf = bs.delta ? (float) 0.f : pdf_emitter_hit(S, e, hit, si.p);
                        float mis = mis_power(bs.pdf, epdf);
                        L = L + mul3(thr * mis, ev);
                    }
                }
            }
        }
        si_prev = si;
        si = hit;
    }
    o.L = L;
}

/* estimator seed of batch c of a pass (SMS_PATH_CHUNK samples each) */
static uint64_t path_chunk_seed(uint64_t seed, uint64_t c) { return mix64(seed ^ mix64(c + 0x632BE59BD9B4E019ULL)); }

/* ImageBlock::put (src/librender/imageblock.cpp:80-165) of one sample into a w x h film with a
   border of 2 pixels (radius > 1: the Gaussian of `tab`, rfilter.h:62-66) or the box branch
   (tab = nullptr), gathered per target pixel: sample j of the pass adds weight * rgb and
   weight to pixel (x, y) if (x, y) is one of its taps */
struct FilterTab { float values[32]; float radius, scale; int n; };
static void filter_table(float stddev, FilterTab *t) {
    /* GaussianFilter (src/rfilters/gaussian.cpp:34-48) and init_discretization (src/libcore/rfilter.cpp:9-20) */
    float radius = 4.f * stddev, alpha = -1.f / (2.f * stddev * stddev), bias = expf(alpha * (radius * radius));
    for (int i = 0; i < 31; ++i) {
        float x = (radius * (float) i) / 31.f;
        t->values[i] = std::max(0.f, expf(alpha * (x * x)) - bias);
    }
    t->values[31] = 0.f;
    t->radius = radius; t->scale = 31.f / radius;
    t->n = (int) ceilf((radius - 2.f * RAY_EPS_F) * 2.f);
}
static inline float filter_disc(const FilterTab &t, float x) {
    int index = (int) fabsf(x * t.scale);
    return t.values[index < 31 ? index : 31];
}
/* the weight of sample position P (film pixels) at image pixel (x, y): 0 when not a tap */
static float splat_weight(const FilterTab *t, int w, int h, float Px, float Py, int x, int y) {
    const float bx = Px + 1.5f, by = Py + 1.5f;          /* pos - (offset - border + .5) */
    if (!t) {
        int lx = (int) ceilf(bx - .5f) - 2, ly = (int) ceilf(by - .5f) - 2;
        return (lx == x && ly == y) ? 1.f : 0.f;
    }
    int lox = std::max((int) ceilf(bx - t->radius), 0), loy = std::max((int) ceilf(by - t->radius), 0);
    int hix = std::min((int) floorf(bx + t->radius), w + 3), hiy = std::min((int) floorf(by + t->radius), h + 3);
    int xr = x + 2 - lox, yr = y + 2 - loy;
    if (xr < 0 || yr < 0 || xr >= t->n || yr >= t->n || x + 2 > hix || y + 2 > hiy) return 0.f;
    float wx = filter_disc(*t, ((float) lox - bx) + (float) xr);
    float wy = filter_disc(*t, ((float) loy - by) + (float) yr);
    return wy * wx;
}
static bool sample_valid(const float *v) {
    for (int k = 0; k < 3; ++k) if (!(isfinite(v[k]) && v[k] >= -1e-5f)) return false;
    return true;
}

} // namespace

/* ================================================================ C entry points */
extern "C" {

void *oracle_create(const sms_scene_desc *desc, const sms_config *cfg) { return scene_create(desc, cfg); }
void oracle_destroy(void *h) { delete (Scene *) h; }
void oracle_set_config(void *h, const sms_config *cfg) { ((Scene *) h)->cfg = *cfg; }

int oracle_ss_walk_batch(void *h, const sms_ss_walk_in *in, int64_t n, sms_ss_walk_out *out, int threads) {
    const Scene &S = *(Scene *) h;
    parallel_for(n, threads, [&](int64_t i) {
        const sms_ss_walk_in &w = in[i];
        EI ei = ei_from_c(w.ei);
        PathOut po = sample_path_ss(S, w.shape, v3(w.si_p[0], w.si_p[1], w.si_p[2]), ei, v2(w.sample[0], w.sample[1]),
                                    v3(w.n_offset[0], w.n_offset[1], w.n_offset[2]),
                                    v3(w.p_start[0], w.p_start[1], w.p_start[2]), v2(w.lean_sample[0], w.lean_sample[1]));
        sms_ss_walk_out &o = out[i];
        memset(&o, 0, sizeof(o));
        o.success = po.ok; o.iterations = po.iterations;
        o.final_shape = po.si_final.valid ? po.si_final.shape : -1;
        o.init_valid = po.si_init.valid;
        for (int a = 0; a < 3; ++a) {
            o.p[a] = (&po.si_final.p.x)[a]; o.n[a] = (&po.si_final.sh_n.x)[a]; o.p_init[a] = (&po.si_init.p.x)[a];
        }
    });
    return 0;
}

int oracle_glint_walk_batch(void *h, const sms_glint_walk_in *in, int64_t n, sms_glint_walk_out *out, int threads) {
    const Scene &S = *(Scene *) h;
    parallel_for(n, threads, [&](int64_t i) {
        const sms_glint_walk_in &w = in[i];
        GlintSI g;
        g.p = v3(w.si_p[0], w.si_p[1], w.si_p[2]);
        g.s = v3(w.frame_s[0], w.frame_s[1], w.frame_s[2]); g.t = v3(w.frame_t[0], w.frame_t[1], w.frame_t[2]);
        g.n = v3(w.frame_n[0], w.frame_n[1], w.frame_n[2]); g.gn = g.n;
        g.uv = v2(w.uv[0], w.uv[1]); g.duv_dx = v2(w.duv_dx[0], w.duv_dx[1]); g.duv_dy = v2(w.duv_dy[0], w.duv_dy[1]);
        g.shape = w.shape; g.dp_du = g.s; g.dp_dv = g.t; g.wi = v3(0.f, 0.f, 1.f);
        V2 uvf, uvi; int it;
        bool ok = sample_glint(S, v3(w.sensor_p[0], w.sensor_p[1], w.sensor_p[2]), v3(w.ei_p[0], w.ei_p[1], w.ei_p[2]), g,
                               v2(w.sample[0], w.sample[1]), v3(w.n_offset[0], w.n_offset[1], w.n_offset[2]),
                               v2(w.xi_start[0], w.xi_start[1]), &uvf, &uvi, &it);
        sms_glint_walk_out &o = out[i];
        o.success = ok; o.iterations = it; o.uv[0] = uvf.x; o.uv[1] = uvf.y; o.uv_init[0] = uvi.x; o.uv_init[1] = uvi.y;
    });
    return 0;
}

int oracle_ss_estimate_batch(void *h, const sms_ss_query *q, int64_t n, uint64_t seed, float *out_rgb,
                             sms_stats *stats, int threads) {
    const Scene &S = *(Scene *) h;
    std::vector<Stats> per((size_t) n);
    parallel_for(n, threads, [&](int64_t i) {
        Stats st; memset(&st, 0, sizeof(st));
        V3 r = estimate_ss(S, q[i], seed, (uint64_t) i, st);
        out_rgb[3 * i] = r.x; out_rgb[3 * i + 1] = r.y; out_rgb[3 * i + 2] = r.z;
        per[(size_t) i] = st;
    });
    Stats tot; memset(&tot, 0, sizeof(tot));
    for (auto &s : per) { tot.walks += s.walks; tot.conv += s.conv; tot.iters += s.iters; tot.bcalls += s.bcalls;
        tot.btrials += s.btrials; if (s.bmax > tot.bmax) tot.bmax = s.bmax; }
    stats_out(tot, stats);
    return 0;
}

int oracle_ss_make_queries(void *h, int spp, uint64_t seed, sms_ss_query *q, float *w, int threads) {
    const Scene &S = *(Scene *) h;
    if (!S.has_cam) return -1;
    int64_t n = (int64_t) S.cam.w * S.cam.h * spp;
    parallel_for(n, threads, [&](int64_t i) { make_ss_query(S, i, spp, seed, q + i, w + i); });
    return 0;
}

/* path integrators (path_sms_ss / path_filtered_ss): per-sample radiance and film positions of a
   pass of n = w*h*spp samples, the SMS estimates in batches of SMS_PATH_CHUNK samples */
int oracle_path_samples(void *h, const sms_path_config *pc, int spp, uint64_t seed, float *rgb, float *pos,
                        sms_stats *stats, int threads) {
    const Scene &S = *(Scene *) h;
    if (!S.has_cam || spp <= 0) return -1;
    const int64_t n = (int64_t) S.cam.w * S.cam.h * spp;
    Stats tot; memset(&tot, 0, sizeof(tot));
    for (int64_t c0 = 0; c0 < n; c0 += SMS_PATH_CHUNK) {
        const int64_t c1 = std::min<int64_t>(n, c0 + SMS_PATH_CHUNK);
        std::vector<PathSample> ps((size_t) (c1 - c0));
        parallel_for(c1 - c0, threads, [&](int64_t i) { path_sample(S, *pc, seed, c0 + i, spp, ps[(size_t) i]); });
        std::vector<sms_ss_query> Q;
        std::vector<const PathMnee *> M;
        std::vector<const PathMneeMs *> MM;
        for (size_t i = 0; i < ps.size(); ++i) {
            for (const PathSms &x : ps[i].sms) Q.push_back(x.q);
            for (const PathMnee &m : ps[i].mnee) M.push_back(&m);
            for (const PathMneeMs &m : ps[i].mnee_ms) MM.push_back(&m);
        }
        const uint64_t cs = path_chunk_seed(seed, (uint64_t) (c0 / SMS_PATH_CHUNK));
        const bool ms = pc->integrator == SMS_PATH_SMS_MS;
        std::vector<V3> est(Q.size());
        std::vector<Stats> per(Q.size());
        parallel_for((int64_t) Q.size(), threads, [&](int64_t k) {
            memset(&per[(size_t) k], 0, sizeof(Stats));
            est[(size_t) k] = ms ? estimate_ms(S, Q[(size_t) k], cs, (uint64_t) k, per[(size_t) k])
                                 : estimate_ss(S, Q[(size_t) k], cs, (uint64_t) k, per[(size_t) k]);
        });
        for (const Stats &x : per) {
            tot.walks += x.walks; tot.conv += x.conv; tot.iters += x.iters; tot.bcalls += x.bcalls;
            tot.btrials += x.btrials; if (x.bmax > tot.bmax) tot.bmax = x.bmax;
        }
        std::vector<char> add(M.size(), 0);
        parallel_for((int64_t) M.size(), threads, [&](int64_t k) {
            const PathMnee &m = *M[(size_t) k];
            const sms_ss_walk_in &w = m.w;
            PathOut po = sample_path_ss(S, w.shape, v3(w.si_p[0], w.si_p[1], w.si_p[2]), ei_from_c(w.ei),
                                        v2(w.sample[0], w.sample[1]), v3(w.n_offset[0], w.n_offset[1], w.n_offset[2]),
                                        v3(w.p_start[0], w.p_start[1], w.p_start[2]),
                                        v2(w.lean_sample[0], w.lean_sample[1]));
            bool ok = po.ok;
            if (ok) {
                V3 dm = normalize(po.si_final.p - m.ref_p);
                if (fabsf(dot(m.dir_pt, dm) - 1.f) >= S.cfg.uniqueness_threshold) ok = false;
            }
            add[(size_t) k] = !ok;
        });
        std::vector<char> add_ms(MM.size(), 0);
        parallel_for((int64_t) MM.size(), threads, [&](int64_t k) {
            const PathMneeMs &m = *MM[(size_t) k];
            const sms_ms_walk_in &w = m.w;
            MSWalk wk = ms_sample_path(S, w.shape, v3(w.si_p[0], w.si_p[1], w.si_p[2]), ei_from_c(w.ei),
                                       v2(w.sample[0], w.sample[1]), true, nullptr, v3(0.f, 0.f, 0.f));
            bool ok = wk.ok;
            for (int v = 0; ok && v < wk.path.n; ++v)
                if (!(v < m.npos) || !(norm(m.pos[v] - wk.path.v[v].v.p) < 1e-5f)) ok = false;
            add_ms[(size_t) k] = !ok;
        });
        size_t kq = 0, km = 0, kmm = 0;
        for (size_t i = 0; i < ps.size(); ++i) {
            V3 L = ps[i].L;
            for (const PathSms &x : ps[i].sms) L = L + mul3(x.thr, est[kq++]);
            for (const PathMnee &m : ps[i].mnee) { if (add[km]) L = L + m.value; ++km; }
            for (const PathMneeMs &m : ps[i].mnee_ms) { if (add_ms[kmm]) L = L + m.value; ++kmm; }
            const int64_t j = c0 + (int64_t) i;
            rgb[3 * j] = L.x; rgb[3 * j + 1] = L.y; rgb[3 * j + 2] = L.z;
            pos[2 * j] = ps[i].pos[0]; pos[2 * j + 1] = ps[i].pos[1];
        }
    }
    stats_out(tot, stats);
    return 0;
}

/* film reconstruction of a pass (sms_film_splat): image += filter-weighted sums, weight += weights */
int oracle_film_splat(void *h, const float *rgb, const float *pos, int spp, float stddev, float *image, float *weight,
                      int threads) {
    const Scene &S = *(Scene *) h;
    if (!S.has_cam || spp <= 0) return -1;
    const int w = S.cam.w, hh = S.cam.h;
    FilterTab tab;
    const FilterTab *t = nullptr;
    if (stddev > 0.f && 4.f * stddev > 1.f) { filter_table(stddev, &tab); t = &tab; }   /* radius <= 1: box branch */
    parallel_for((int64_t) w * hh, threads, [&](int64_t p) {
        const int x = (int) (p % w), y = (int) (p / w);
        float acc[3] = {0.f, 0.f, 0.f}, ws = 0.f;
        for (int ny = y - 2; ny <= y + 2; ++ny)
            for (int nx = x - 2; nx <= x + 2; ++nx) {
                if (nx < 0 || ny < 0 || nx >= w || ny >= hh) continue;
                const int64_t base = ((int64_t) ny * w + nx) * spp;
                for (int k = 0; k < spp; ++k) {
                    const float *v = rgb + 3 * (base + k);
                    if (!sample_valid(v)) continue;
                    const float wt = splat_weight(t, w, hh, pos[2 * (base + k)], pos[2 * (base + k) + 1], x, y);
                    if (wt == 0.f) continue;
                    for (int a = 0; a < 3; ++a) acc[a] = acc[a] + v[a] * wt;
                    ws = ws + wt;
                }
            }
        for (int a = 0; a < 3; ++a) image[3 * p + a] = image[3 * p + a] + acc[a];
        weight[p] = weight[p] + ws;
    });
    return 0;
}

/* estimator with the learned proposal on caster `shape`: raw = n x 48 raw network outputs;
   radius > 0: anchored mixtures */
int oracle_ss_estimate_batch_prop_anchored(void *h, const sms_ss_query *q, int64_t n, uint64_t seed, const float *raw,
                                           int shape, float defensive, float radius, float *out_rgb, sms_stats *stats,
                                           int threads) {
    const Scene &S = *(Scene *) h;
    if (shape < 0 || shape >= (int) S.shapes.size() || S.shapes[shape].type != SMS_SHAPE_MESH) return -1;
    Prop P;
    P.raw = raw; P.shape = shape; P.defensive = defensive; P.radius = radius;
    build_uv_grid(S, shape, P.grid);
    std::vector<Stats> per((size_t) n);
    parallel_for(n, threads, [&](int64_t i) {
        Stats st; memset(&st, 0, sizeof(st));
        V3 r = estimate_ss(S, q[i], seed, (uint64_t) i, st, &P);
        out_rgb[3 * i] = r.x; out_rgb[3 * i + 1] = r.y; out_rgb[3 * i + 2] = r.z;
        per[(size_t) i] = st;
    });
    Stats tot; memset(&tot, 0, sizeof(tot));
    for (auto &s : per) { tot.walks += s.walks; tot.conv += s.conv; tot.iters += s.iters; tot.bcalls += s.bcalls;
        tot.btrials += s.btrials; if (s.bmax > tot.bmax) tot.bmax = s.bmax; }
    stats_out(tot, stats);
    return 0;
}

int oracle_ss_estimate_batch_prop(void *h, const sms_ss_query *q, int64_t n, uint64_t seed, const float *raw, int shape,
                                  float defensive, float *out_rgb, sms_stats *stats, int threads) {
    return oracle_ss_estimate_batch_prop_anchored(h, q, n, seed, raw, shape, defensive, 0.f, out_rgb, stats, threads);
}

/* anchors (n x 2, NaN: none) of the queries of an estimate on mesh `shape` */
void oracle_prop_anchors(void *h, const sms_ss_query *q, int64_t n, uint64_t seed, int shape, float *anchor) {
    const Scene &S = *(Scene *) h;
    for (int64_t i = 0; i < n; ++i) {
        V2 a = v2(std::numeric_limits<float>::quiet_NaN(), std::numeric_limits<float>::quiet_NaN());
        if (q[i].shape >= 0) {
            SI si = si_from_query(q[i]);
            Pcg32 rq; stream_seed(&rq, seed, (uint64_t) i, 0);
            EI ei = sample_emitter_interaction(S, si.p, S.emitters_ss, &rq);
            a = proposal_anchor(S, shape, si.p, ei.p);
        }
        anchor[2 * i] = a.x; anchor[2 * i + 1] = a.y;
    }
}

/* network inputs of the queries of an estimate (same light sample as the estimator) */
void oracle_prop_inputs(void *h, const sms_ss_query *q, int64_t n, uint64_t seed, float *x) {
    const Scene &S = *(Scene *) h;
    for (int64_t i = 0; i < n; ++i) {
        SI si = si_from_query(q[i]);
        Pcg32 rq; stream_seed(&rq, seed, (uint64_t) i, 0);
        EI ei = sample_emitter_interaction(S, si.p, S.emitters_ss, &rq);
        proposal_inputs(S, si.p, si.sh_n, ei.p, x + 9 * i);
    }
}

/* the same for the multi-scatter estimator's light samples */
void oracle_prop_inputs_ms(void *h, const sms_ss_query *q, int64_t n, uint64_t seed, float *x) {
    const Scene &S = *(Scene *) h;
    std::vector<int> em_ms;
    for (int i = 0; i < (int) S.emitters.size(); ++i) if (S.emitters[i].ms) em_ms.push_back(i);
    for (int64_t i = 0; i < n; ++i) {
        SI si = si_from_query(q[i]);
        Pcg32 rq; stream_seed(&rq, seed, (uint64_t) i, 0);
        EI ei = sample_emitter_interaction(S, si.p, em_ms, &rq);
        proposal_inputs(S, si.p, si.sh_n, ei.p, x + 9 * i);
    }
}

/* multi-scatter estimates with the proposal of raw network outputs raw (n x 48) on `shape` */
int oracle_ms_estimate_batch_prop(void *h, const sms_ss_query *q, int64_t n, uint64_t seed, const float *raw, int shape,
                                  float defensive, float *out_rgb, sms_stats *stats, int threads) {
    const Scene &S = *(Scene *) h;
    if (S.cfg.bounces < 1 || S.cfg.bounces > MS_MAXB) return -5;
    if (shape < 0 || shape >= (int) S.shapes.size() || S.shapes[shape].type != SMS_SHAPE_MESH) return -1;
    Prop P;
    P.raw = raw; P.shape = shape; P.defensive = defensive;
    build_uv_grid(S, shape, P.grid);
    std::vector<Stats> per((size_t) n);
    parallel_for(n, threads, [&](int64_t i) {
        Stats st; memset(&st, 0, sizeof(st));
        V3 r = estimate_ms(S, q[i], seed, (uint64_t) i, st, &P);
        out_rgb[3 * i] = r.x; out_rgb[3 * i + 1] = r.y; out_rgb[3 * i + 2] = r.z;
        per[(size_t) i] = st;
    });
    Stats tot; memset(&tot, 0, sizeof(tot));
    for (auto &s : per) { tot.walks += s.walks; tot.conv += s.conv; tot.iters += s.iters; tot.bcalls += s.bcalls;
        tot.btrials += s.btrials; if (s.bmax > tot.bmax) tot.bmax = s.bmax; }
    stats_out(tot, stats);
    return 0;
}

/* generalized geometric term of a glint path (tests): the vertex at query q's point with the
   normal map at uv, the sensor at x0, a point light at x2 */
float oracle_glint_geometric(void *h, const sms_glint_query *q, const float *uv, const float *x0, const float *x2) {
    const Scene &S = *(Scene *) h;
    GlintSI gsi = gsi_from_query(*q);
    EI ei;
    memset(&ei, 0, sizeof(ei));
    ei.kind = SMS_EI_POINT;
    ei.p = v3(x2[0], x2[1], x2[2]);
    MV vx, vtx, vy;
    if (!glint_vertices(S, v3(x0[0], x0[1], x0[2]), ei, gsi, v2(uv[0], uv[1]), &vx, &vtx, &vy)) return -1.f;
    return geometric_term_glint(vx, vtx, vy);
}

/* the glint BSDF-sampling strategy of query q (index qid) alone, with its parts (tests) */
void oracle_glint_bsdf_parts(void *h, const sms_glint_query *q, uint64_t seed, uint64_t qid, float *rgb, float *parts) {
    const Scene &S = *(Scene *) h;
    V3 v = glint_bsdf_strategy(S, gsi_from_query(*q), seed, qid, parts);
    rgb[0] = v.x; rgb[1] = v.y; rgb[2] = v.z;
}

/* SpecularManifold::sample_gaussian of the two-stage solver (tests): mu[2], sig = (s00, s11, s01) */
void oracle_sample_gaussian(const float *mu, const float *sig, const float *s, float *out) {
    V2 r = sample_gaussian(v2(mu[0], mu[1]), sig, v2(s[0], s[1]));
    out[0] = r.x; out[1] = r.y;
}

void oracle_scene_box(void *h, float *c, float *inv) {
    const Scene &S = *(Scene *) h;
    for (int a = 0; a < 3; ++a) c[a] = S.box_c[a];
    *inv = S.box_inv;
}

/* estimates of selected queries of a batch: q[i] is query ids[i] of the batch (its random
   streams), raw row i its network outputs; sequential */
int oracle_ss_estimate_ids_prop(void *h, const sms_ss_query *q, const int64_t *ids, int64_t n, uint64_t seed,
                                const float *raw, int shape, float defensive, float *out_rgb, sms_stats *stats) {
    const Scene &S = *(Scene *) h;
    if (shape < 0 || shape >= (int) S.shapes.size() || S.shapes[shape].type != SMS_SHAPE_MESH) return -1;
    Prop P;
    P.shape = shape; P.defensive = defensive;
    build_uv_grid(S, shape, P.grid);
    Stats st; memset(&st, 0, sizeof(st));
    for (int64_t i = 0; i < n; ++i) {
        P.raw = raw + i * PROP_RAW;
        P.raw_base = ids[i];
        V3 r = estimate_ss(S, q[i], seed, (uint64_t) ids[i], st, &P);
        out_rgb[3 * i] = r.x; out_rgb[3 * i + 1] = r.y; out_rgb[3 * i + 2] = r.z;
    }
    if (stats) {
        memset(stats, 0, sizeof(*stats));
        stats->walks = st.walks; stats->walks_converged = st.conv; stats->newton_iterations = st.iters;
        stats->bernoulli_calls = st.bcalls; stats->bernoulli_trials = st.btrials; stats->bernoulli_max = st.bmax;
    }
    return 0;
}

void oracle_gmm_prepare(const float *raw, int64_t n, float *g) {
    for (int64_t i = 0; i < n; ++i) gmm_prepare(raw + PROP_RAW * i, g + PROP_RAW * i);
}

/* seed draw of walk `walk` of query `qid`: out = (kind, u, v) per record */
void oracle_prop_seed(const float *g, float defensive, uint64_t seed, uint64_t qid, uint64_t walk0, int64_t n,
                      float *out) {
    for (int64_t i = 0; i < n; ++i) {
        Pcg32 rw; stream_seed(&rw, seed, qid, walk0 + (uint64_t) i);
        V2 s;
        int kind = proposal_seed(g, defensive, &rw, &s);
        out[3 * i] = (float) kind; out[3 * i + 1] = s.x; out[3 * i + 2] = s.y;
    }
}

/* uv -> surface point through the proposal's grid of mesh `shape`; ok[i] = 0 when no face
   holds uv; brute = 1 scans every face of the mesh in order instead of the grid */
int oracle_uv_to_point(void *h, int shape, const float *uv, int64_t n, int brute, float *p, int32_t *ok) {
    const Scene &S = *(Scene *) h;
    if (shape < 0 || shape >= (int) S.shapes.size() || S.shapes[shape].type != SMS_SHAPE_MESH) return -1;
    UVGrid G;
    if (brute) {
        const Shape &sh = S.shapes[shape];
        G.g = 1; G.start = {0, sh.nf};
        for (int f = 0; f < sh.nf; ++f) G.faces.push_back(sh.fbase + f);
    } else {
        build_uv_grid(S, shape, G);
    }
    for (int64_t i = 0; i < n; ++i) {
        V3 x = v3(0.f, 0.f, 0.f);
        ok[i] = uv_to_point(S, G, v2(uv[2 * i], uv[2 * i + 1]), &x) ? 1 : 0;
        p[3 * i] = x.x; p[3 * i + 1] = x.y; p[3 * i + 2] = x.z;
    }
    return 0;
}

int oracle_glint_estimate_batch(void *h, const sms_glint_query *q, int64_t n, uint64_t seed, float *out_rgb,
                                sms_stats *stats, int threads) {
    const Scene &S = *(Scene *) h;
    std::vector<Stats> per((size_t) n);
    parallel_for(n, threads, [&](int64_t i) {
        Stats st; memset(&st, 0, sizeof(st));
        V3 r = estimate_glint(S, q[i], seed, (uint64_t) i, st);
        out_rgb[3 * i] = r.x; out_rgb[3 * i + 1] = r.y; out_rgb[3 * i + 2] = r.z;
        per[(size_t) i] = st;
    });
    Stats tot; memset(&tot, 0, sizeof(tot));
    for (auto &s : per) { tot.walks += s.walks; tot.conv += s.conv; tot.iters += s.iters; tot.bcalls += s.bcalls;
        tot.btrials += s.btrials; if (s.bmax > tot.bmax) tot.bmax = s.bmax; }
    stats_out(tot, stats);
    return 0;
}

/* The glint estimator with the learned proposal: raw network outputs (n x 48) give each query's
   defensive mixture over its pixel footprint (defensive weight `defensive`). */
int oracle_glint_estimate_batch_prop(void *h, const sms_glint_query *q, int64_t n, uint64_t seed, const float *raw,
                                     float defensive, float *out_rgb, sms_stats *stats, int threads) {
    const Scene &S = *(Scene *) h;
    std::vector<Stats> per((size_t) n);
    parallel_for(n, threads, [&](int64_t i) {
        Stats st; memset(&st, 0, sizeof(st));
        float gq[PROP_RAW];
        gmm_prepare(raw + i * PROP_RAW, gq);
        V3 r = estimate_glint(S, q[i], seed, (uint64_t) i, st, gq, defensive);
        out_rgb[3 * i] = r.x; out_rgb[3 * i + 1] = r.y; out_rgb[3 * i + 2] = r.z;
        per[(size_t) i] = st;
    });
    Stats tot; memset(&tot, 0, sizeof(tot));
    for (auto &s : per) { tot.walks += s.walks; tot.conv += s.conv; tot.iters += s.iters; tot.bcalls += s.bcalls;
        tot.btrials += s.btrials; if (s.bmax > tot.bmax) tot.bmax = s.bmax; }
    stats_out(tot, stats);
    return 0;
}

/* Network inputs of glint queries: shading point, shading normal and the glint estimator's light
   sample (its emitters are all emitters). */
void oracle_glint_prop_inputs(void *h, const sms_glint_query *q, int64_t n, uint64_t seed, float *x) {
    const Scene &S = *(Scene *) h;
    for (int64_t i = 0; i < n; ++i) {
        if (q[i].shape < 0) { for (int k = 0; k < 9; ++k) x[9 * i + k] = 0.f; continue; }
        GlintSI si = gsi_from_query(q[i]);
        Pcg32 rq; stream_seed(&rq, seed, (uint64_t) i, 0);
        EI ei = sample_emitter_interaction(S, si.p, S.emitters_all, &rq);
        proposal_inputs(S, si.p, si.n, ei.p, x + 9 * i);
    }
}

int oracle_ms_walk_batch(void *h, const sms_ms_walk_in *in, int64_t n, sms_ms_walk_out *out, int threads) {
    const Scene &S = *(Scene *) h;
    if (S.cfg.bounces < 1 || S.cfg.bounces > MS_MAXB) return -5;
    parallel_for(n, threads, [&](int64_t i) {
        const sms_ms_walk_in &w = in[i];
        EI ei = ei_from_c(w.ei);
        MSWalk r = ms_sample_path(S, w.shape, v3(w.si_p[0], w.si_p[1], w.si_p[2]), ei, v2(w.sample[0], w.sample[1]),
                                  w.first_path != 0, w.seed_shapes, v3(w.p_start[0], w.p_start[1], w.p_start[2]));
        sms_ms_walk_out &o = out[i];
        memset(&o, 0, sizeof(o));
        o.success = r.ok; o.iterations = r.iterations; o.seed_valid = r.seed_ok; o.n_vertices = r.path.n;
        for (int k = 0; k < MS_MAXB; ++k) {
            o.shapes[k] = r.seed_ok ? r.seed_shapes[k] : -1;
            if (k < r.path.n) { o.p[k][0] = r.path.v[k].v.p.x; o.p[k][1] = r.path.v[k].v.p.y; o.p[k][2] = r.path.v[k].v.p.z; }
        }
    });
    return 0;
}

int oracle_ms_estimate_batch(void *h, const sms_ss_query *q, int64_t n, uint64_t seed, float *out_rgb,
                             sms_stats *stats, int threads) {
    const Scene &S = *(Scene *) h;
    if (S.cfg.bounces < 1 || S.cfg.bounces > MS_MAXB) return -5;
    std::vector<Stats> per((size_t) n);
    parallel_for(n, threads, [&](int64_t i) {
        Stats st; memset(&st, 0, sizeof(st));
        V3 r = estimate_ms(S, q[i], seed, (uint64_t) i, st);
        out_rgb[3 * i] = r.x; out_rgb[3 * i + 1] = r.y; out_rgb[3 * i + 2] = r.z;
        per[(size_t) i] = st;
    });
    Stats tot; memset(&tot, 0, sizeof(tot));
    for (auto &s : per) { tot.walks += s.walks; tot.conv += s.conv; tot.iters += s.iters; tot.bcalls += s.bcalls;
        tot.btrials += s.btrials; if (s.bmax > tot.bmax) tot.bmax = s.bmax; }
    stats_out(tot, stats);
    return 0;
}

int oracle_glint_make_queries(void *h, int spp, uint64_t seed, sms_glint_query *q, int threads) {
    const Scene &S = *(Scene *) h;
    if (!S.has_cam) return -1;
    int64_t n = (int64_t) S.cam.w * S.cam.h * spp;
    parallel_for(n, threads, [&](int64_t i) { make_glint_query(S, i, spp, seed, &q[i]); });
    return 0;
}

/* Scene ray cast exposed for the BVH-vs-brute-force test. */
int oracle_ray_intersect(void *h, const float *o, const float *d, float *t, int *shape, int *prim, float *p) {
    const Scene &S = *(Scene *) h;
    Ray r = make_ray(v3(o[0], o[1], o[2]), v3(d[0], d[1], d[2]));
    SI si = ray_intersect(S, r);
    *t = si.t; *shape = si.valid ? si.shape : -1; *prim = si.prim;
    p[0] = si.p.x; p[1] = si.p.y; p[2] = si.p.z;
    return si.valid;
}

/* Constraint value and Jacobian of the angle-difference / half-vector step at a point on a
   specular shape (for finite-difference tests).  Returns step.ok. */
int oracle_ss_step_at(void *h, int halfvector, const float *x, const float *o, const float *d,
                      const sms_emitter_interaction *e, float *C, float *dX, float *vtx_p) {
    const Scene &S = *(Scene *) h;
    SI si = ray_intersect(S, make_ray(v3(o[0], o[1], o[2]), v3(d[0], d[1], d[2])));
    if (!si.valid) return -1;
    MV v = mv_from_si(S, si);
    EI ei = ei_from_c(*e);
    Step st = halfvector ? step_halfvector(v3(x[0], x[1], x[2]), v, ei, v3(0.f, 0.f, 1.f))
                         : step_anglediff(v3(x[0], x[1], x[2]), v, ei, v3(0.f, 0.f, 1.f));
    C[0] = st.C.x; C[1] = st.C.y; dX[0] = st.dX.x; dX[1] = st.dX.y;
    vtx_p[0] = v.p.x; vtx_p[1] = v.p.y; vtx_p[2] = v.p.z;
    return st.ok;
}

/* Block-tridiagonal kernels of the multi-scatter solver on caller-given blocks (row-major 2x2
   per vertex: dprev, dcur, dnext; C two floats per vertex), exposed for checks against dense
   linear algebra.  step: dx = A^{-1} C (returns 0 when a pivot block is singular);
   geo: |det| of the vertex-0 block of -A'^{-1} B (manifold_ms.cpp:1234-1268). */
int oracle_ms_tridiag_step(int n, const float *dprev, const float *dcur, const float *dnext, const float *C, float *dx) {
    if (n < 1 || n > MS_MAXB) return -1;
    MSPath P;
    P.n = n;
    for (int i = 0; i < n; ++i) {
        P.v[i].dprev = m2(dprev[4 * i], dprev[4 * i + 1], dprev[4 * i + 2], dprev[4 * i + 3]);
        P.v[i].dcur = m2(dcur[4 * i], dcur[4 * i + 1], dcur[4 * i + 2], dcur[4 * i + 3]);
        P.v[i].dnext = m2(dnext[4 * i], dnext[4 * i + 1], dnext[4 * i + 2], dnext[4 * i + 3]);
        P.v[i].C = v2(C[2 * i], C[2 * i + 1]);
    }
    if (!ms_invert_tridiagonal_step(P)) return 0;
    for (int i = 0; i < n; ++i) { dx[2 * i] = P.v[i].dx.x; dx[2 * i + 1] = P.v[i].dx.y; }
    return 1;
}
float oracle_ms_tridiag_geo(int n, const float *dprev, const float *dcur, const float *dnext) {
    if (n < 2 || n > MS_MAXB + 2) return -1.f;
    MSVert v[MS_MAXB + 2];
    for (int i = 0; i < n; ++i) {
        v[i].dprev = m2(dprev[4 * i], dprev[4 * i + 1], dprev[4 * i + 2], dprev[4 * i + 3]);
        v[i].dcur = m2(dcur[4 * i], dcur[4 * i + 1], dcur[4 * i + 2], dcur[4 * i + 3]);
        v[i].dnext = m2(dnext[4 * i], dnext[4 * i + 1], dnext[4 * i + 2], dnext[4 * i + 3]);
    }
    return ms_invert_tridiagonal_geo(v, n);
}

/* LEAN pieces exposed for the pdf-normalisation and mip-pyramid checks. */
float oracle_lean_pdf_params(const float *mu, const float *sig, const float *wi, const float *wo) {
    LeanParams p = lean_params(v2(mu[0], mu[1]), sig, 0.f, 0.f);
    V3 i = v3(wi[0], wi[1], wi[2]), o = v3(wo[0], wo[1], wo[2]);
    if (i.z < 0.f || o.z < 0.f) return 0.f;
    V3 H = normalize(o + i);
    float G = lean_g1(o, p) * lean_g1(i, p);
    if (!(G > 0.f)) return 0.f;
    return lean_vndf(H, i, p) / (4.f * dot(i, H));
}
int oracle_lean_levels(void *h, int nmap) { return (int) ((Scene *) h)->nmaps[nmap].lean.size(); }
int oracle_lean_level(float tiles, const float *duv_dx, const float *duv_dy, int levels) {
    return lean_level(tiles, v2(duv_dx[0], duv_dx[1]), v2(duv_dy[0], duv_dy[1]), levels);
}
void oracle_lean_texel(void *h, int nmap, int level, int x, int y, float *out5) {
    const Normalmap &nm = ((Scene *) h)->nmaps[nmap];
    const int res = nm.res >> level;
    for (int c = 0; c < 5; ++c) out5[c] = nm.lean[level][((size_t) y * res + x) * 5 + c];
}
/* Set a planted-error factor (tests only; see the PLANT_* sites).  Returns 0, or -1 for an
   unknown site. */
int oracle_plant(int site, float factor) {
    if (site < 0 || site >= PLANT_N) return -1;
    g_plant[site] = factor;
    return 0;
}

/* The pieces of evaluate_path_contribution SS (manifold_ss.cpp:281-327) for a receiver query, an
   emitter interaction and the specular vertex hit by the ray from the query point toward p1:
   out[0] = geometric term, out[1] = specular_reflectance (first channel), out[2] = emitter
   weight (first channel), out[3..5] = the contribution, out[6..8] = the vertex used.  Returns 0
   when the ray does not reach a caster. */
int oracle_ss_contribution(void *h, const sms_ss_query *q, const sms_emitter_interaction *e, const float *p1,
                           float *out) {
    const Scene &S = *(Scene *) h;
    SI si = si_from_query(*q);
    V3 x = si.p;
    SI sf = ray_intersect(S, make_ray(x, normalize(v3(p1[0], p1[1], p1[2]) - x)));
    if (!sf.valid || !(S.shapes[sf.shape].flags & SMS_FLAG_CAUSTIC_CASTER_SS)) return 0;
    EI ei = ei_from_c(*e);
    V3 c = path_contribution_ss(S, si, ei, sf);
    /* the geometric term and the emitter weight as path_contribution_ss forms them */
    MV vtx = mv_from_si(S, sf);
    make_orthonormal(vtx);
    EI e2 = ei;
    if (e2.kind == SMS_EI_POINT) {
        const Emitter &em = S.emitters[e2.emitter];
        V3 d = em.position - sf.p;
        float dist = norm(d), inv_dist = 1.f / dist;
        d = d * inv_dist;
        float emitter_pdf = e2.pdf;
        e2.p = em.position; e2.d = d; e2.n = d; e2.pdf = 1.f;
        e2.weight = (em.intensity * (inv_dist * inv_dist)) * dist * dist;
        e2.pdf *= emitter_pdf;
        e2.weight = e2.weight * (1.f / emitter_pdf);
    }
    MV vy;
    float G = 0.f;
    if (emitter_interaction_to_vertex(S, e2, vtx.p, &vy)) {
        MV vx = mv_from_si(S, si);
        make_orthonormal(vx);
        G = geometric_term_ss(vx, vtx, vy);
    }
    out[0] = G;
    out[1] = specular_reflectance_ss(S, sf, normalize(e2.p - sf.p)).x;
    out[2] = e2.weight.x;
    out[3] = c.x; out[4] = c.y; out[5] = c.z;
    out[6] = sf.p.x; out[7] = sf.p.y; out[8] = sf.p.z;
    return 1;
}

/* Constraint values and Newton step of the multi-scatter solver (angle-difference or half-vector
   constraints, manifold_ms.cpp:500-857, block-tridiagonal solve :1192-1232) for the chain that a
   ray from x0 through p[0] and then through p[1], ... hits (k vertices), with emitter
   interaction e.  Writes C (k x 2), dx (k x 2) and the vertices used (k x 3).  Returns 1 when the
   step exists, 0 when the constraints or the solve fail, -1 when the chain is not hit. */
int oracle_ms_step_at(void *h, int halfvector, const float *x0, int k, const float *p,
                      const sms_emitter_interaction *e, float *C, float *dx, float *vp) {
    const Scene &S = *(Scene *) h;
    if (k < 1 || k > MS_MAXB) return -1;
    MSPath P;
    memset(&P, 0, sizeof(P));
    V3 o = v3(x0[0], x0[1], x0[2]);
    Ray ray = make_ray(o, normalize(v3(p[0], p[1], p[2]) - o));
    for (int i = 0; i < k; ++i) {
        SI si = ray_intersect(S, ray);
        if (!si.valid) return -1;
        P.v[i].v = mv_from_si(S, si);
        P.shapes[i] = si.shape;
        P.n++;
        if (i + 1 < k) ray = spawn_ray(si.p, normalize(v3(p[3 * (i + 1)], p[3 * (i + 1) + 1], p[3 * (i + 1) + 2]) - si.p));
    }
    EI ei = ei_from_c(*e);
    bool ok = halfvector ? ms_step_halfvector(P, o, ei) : ms_step_anglediff(P, o, ei);
    for (int i = 0; i < k; ++i) {
        C[2 * i] = P.v[i].C.x; C[2 * i + 1] = P.v[i].C.y;
        dx[2 * i] = P.v[i].dx.x; dx[2 * i + 1] = P.v[i].dx.y;
        vp[3 * i] = P.v[i].v.p.x; vp[3 * i + 1] = P.v[i].v.p.y; vp[3 * i + 2] = P.v[i].v.p.z;
    }
    return ok ? 1 : 0;
}

double oracle_exp_d(double x) { return exp_d(x); }
double oracle_erf_d(double x) { return erf_d(x); }
double oracle_log2_d(double x) { return log2_d(x); }

/* Math routines exposed for pinning against libm. */
float oracle_acosf(float x) { return safe_acos(x); }
float oracle_atan2f(float y, float x) { return c_atan2f(y, x); }
void oracle_sincosf(float x, float *s, float *c) { c_sincosf(x, s, c); }
float oracle_fresnel(float c, float eta) { return fresnel(c, eta); }
float oracle_fresnel_conductor(float c, float er, float ei) { return fresnel_conductor(c, er, ei); }
void oracle_square_to_uniform_sphere(float u, float v, float *out) {
    V3 r = square_to_uniform_sphere(v2(u, v)); out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
uint32_t oracle_pcg32_first(uint64_t initstate, uint64_t initseq, uint32_t *out, int n) {
    Pcg32 r; pcg32_seed(&r, initstate, initseq);
    for (int i = 0; i < n; ++i) out[i] = pcg32_next(&r);
    return 0;
}
int oracle_sample_emitter_interaction(void *h, const float *p, uint64_t seed, uint64_t qid, sms_emitter_interaction *out) {
    const Scene &S = *(Scene *) h;
    Pcg32 rq; stream_seed(&rq, seed, qid, 0);
    EI ei = sample_emitter_interaction(S, v3(p[0], p[1], p[2]), S.emitters_ss.empty() ? S.emitters_all : S.emitters_ss, &rq);
    memset(out, 0, sizeof(*out));
    for (int a = 0; a < 3; ++a) { out->p[a] = (&ei.p.x)[a]; out->n[a] = (&ei.n.x)[a]; out->d[a] = (&ei.d.x)[a]; out->weight[a] = (&ei.weight.x)[a]; }
    out->pdf = ei.pdf; out->kind = ei.kind; out->emitter = ei.emitter;
    return 0;
}

} // extern "C"
