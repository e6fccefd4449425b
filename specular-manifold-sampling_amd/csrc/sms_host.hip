// sms_host.hip -- C ABI (include/sms.h) of the gfx950 SMS library: context, scene upload,
// BVH construction, normal-map decoding and kernel launches.
//
// One context lives on the HIP device that is current when sms_ctx_create is called and
// owns one stream (default priority).  Estimator calls run two persistent passes on that
// stream; their work-queue buffers grow on demand and are reused across calls.
#include "host_ctx.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

using namespace smsdev;
using namespace smshost;


namespace {

template <typename T>
sms_status upload(sms_ctx *ctx, const std::vector<T> &v, const T **dst) {
    if (v.empty()) { *dst = nullptr; return SMS_OK; }
    void *p = nullptr;
    HIP_TRY(ctx, hipMalloc(&p, v.size() * sizeof(T)));
    ctx->scene_allocs.push_back(p);
    HIP_TRY(ctx, hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    *dst = (const T *) p;
    return SMS_OK;
}

void invert_affine(const float *M, float *Minv) {
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

float hmean_same(float e) { return ((e + e) + e) / 3.f; }

float i2f(int x) { float f; memcpy(&f, &x, 4); return f; }

// ---------------------------------------------------------------- binned-SAH BVH (host build)
struct BItem { float bmin[3], bmax[3], c[3]; int tri; };

struct BvhOut { std::vector<DNode> nodes; std::vector<int> order; int max_depth = 0; };

void pad_box(float *bmin, float *bmax) {
    for (int a = 0; a < 3; ++a) {
        float ext = bmax[a] - bmin[a];
        float pad = 1e-5f * ext + 1e-5f * std::max(std::fabs(bmin[a]), std::fabs(bmax[a])) + 1e-7f;
        bmin[a] -= pad; bmax[a] += pad;
    }
}

BvhOut build_bvh(std::vector<BItem> &items) {
    BvhOut out;
    int n = (int) items.size();
    if (n == 0) return out;
    const int BINS = 16, MAX_LEAF = 8;
    struct Task { int node, lo, hi, depth; };
    std::vector<Task> stack;
    // node 0 is the root and node 1 unused, so every sibling pair starts at an even index: the
    // traversal reads both children of a node as one 64-byte aligned pair
    out.nodes.push_back(DNode());
    out.nodes.push_back(DNode());
    stack.push_back({0, 0, n, 0});
    while (!stack.empty()) {
        Task tk = stack.back(); stack.pop_back();
        out.max_depth = std::max(out.max_depth, tk.depth);
        float bmin[3] = {INFINITY, INFINITY, INFINITY}, bmax[3] = {-INFINITY, -INFINITY, -INFINITY};
        float cmin[3] = {INFINITY, INFINITY, INFINITY}, cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int i = tk.lo; i < tk.hi; ++i)
            for (int a = 0; a < 3; ++a) {
                bmin[a] = std::min(bmin[a], items[i].bmin[a]); bmax[a] = std::max(bmax[a], items[i].bmax[a]);
                cmin[a] = std::min(cmin[a], items[i].c[a]); cmax[a] = std::max(cmax[a], items[i].c[a]);
            }
        float pb0[3] = {bmin[0], bmin[1], bmin[2]}, pb1[3] = {bmax[0], bmax[1], bmax[2]};
        pad_box(pb0, pb1);
        int count = tk.hi - tk.lo;
        DNode nd;
        nd.a = make_float4(pb0[0], pb0[1], pb0[2], 0.f);
        nd.b = make_float4(pb1[0], pb1[1], pb1[2], 0.f);
        auto make_leaf = [&]() {
            nd.a.w = i2f(tk.lo);
            nd.b.w = i2f(count);
            out.nodes[tk.node] = nd;
        };
        if (count <= 2) { make_leaf(); continue; }
        // SAH over BINS bins per axis
        auto area = [](const float *lo, const float *hi) {
            float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
            return dx * dy + dy * dz + dz * dx;
        };
        int best_axis = -1, best_split = -1;
        float best_cost = INFINITY;
        for (int a = 0; a < 3; ++a) {
            float ext = cmax[a] - cmin[a];
            if (!(ext > 0.f)) continue;
            int cnt[BINS] = {0};
            float lo[BINS][3], hi[BINS][3];
            for (int b = 0; b < BINS; ++b) for (int k = 0; k < 3; ++k) { lo[b][k] = INFINITY; hi[b][k] = -INFINITY; }
            float scale = BINS / ext;
            for (int i = tk.lo; i < tk.hi; ++i) {
                int b = std::min(BINS - 1, (int) ((items[i].c[a] - cmin[a]) * scale));
                cnt[b]++;
                for (int k = 0; k < 3; ++k) { lo[b][k] = std::min(lo[b][k], items[i].bmin[k]); hi[b][k] = std::max(hi[b][k], items[i].bmax[k]); }
            }
            float rl[BINS][3], rh[BINS][3]; int rc[BINS];
            float al[3] = {INFINITY, INFINITY, INFINITY}, ah[3] = {-INFINITY, -INFINITY, -INFINITY}; int ac = 0;
            for (int b = BINS - 1; b > 0; --b) {
                ac += cnt[b];
                for (int k = 0; k < 3; ++k) { al[k] = std::min(al[k], lo[b][k]); ah[k] = std::max(ah[k], hi[b][k]); }
                rc[b] = ac; for (int k = 0; k < 3; ++k) { rl[b][k] = al[k]; rh[b][k] = ah[k]; }
            }
            float ll[3] = {INFINITY, INFINITY, INFINITY}, lh[3] = {-INFINITY, -INFINITY, -INFINITY}; int lc = 0;
            for (int b = 0; b < BINS - 1; ++b) {
                lc += cnt[b];
                for (int k = 0; k < 3; ++k) { ll[k] = std::min(ll[k], lo[b][k]); lh[k] = std::max(lh[k], hi[b][k]); }
                if (lc == 0 || rc[b + 1] == 0) continue;
                float cost = lc * area(ll, lh) + rc[b + 1] * area(rl[b + 1], rh[b + 1]);
                if (cost < best_cost) { best_cost = cost; best_axis = a; best_split = b; }
            }
        }
        float leaf_cost = count * area(bmin, bmax);
        if (best_axis < 0 || (count <= MAX_LEAF && best_cost >= leaf_cost)) {
            if (best_axis < 0 && count > MAX_LEAF) {
                // degenerate centroids: split in the middle of the item range
                int mid = (tk.lo + tk.hi) / 2;
                nd.a.w = i2f((int) out.nodes.size());
                nd.b.w = i2f(0);
                out.nodes[tk.node] = nd;
                int l = (int) out.nodes.size();
                out.nodes.push_back(DNode()); out.nodes.push_back(DNode());
                stack.push_back({l, tk.lo, mid, tk.depth + 1}); stack.push_back({l + 1, mid, tk.hi, tk.depth + 1});
                continue;
            }
            make_leaf();
            continue;
        }
        float ext = cmax[best_axis] - cmin[best_axis], scale = BINS / ext;
        auto mid_it = std::partition(items.begin() + tk.lo, items.begin() + tk.hi, [&](const BItem &it) {
            return std::min(BINS - 1, (int) ((it.c[best_axis] - cmin[best_axis]) * scale)) <= best_split;
        });
        int mid = (int) (mid_it - items.begin());
        if (mid == tk.lo || mid == tk.hi) mid = (tk.lo + tk.hi) / 2;
        int l = (int) out.nodes.size();
        nd.a.w = i2f(l);
        nd.b.w = i2f(0);
        out.nodes[tk.node] = nd;
        out.nodes.push_back(DNode()); out.nodes.push_back(DNode());
        stack.push_back({l, tk.lo, mid, tk.depth + 1}); stack.push_back({l + 1, mid, tk.hi, tk.depth + 1});
    }
    out.order.resize(n);
    for (int i = 0; i < n; ++i) out.order[i] = items[i].tri;
    return out;
}

void camera_setup(DCamera &c, const sms_camera_desc &d) {
    double o[3] = {d.origin[0], d.origin[1], d.origin[2]};
    double tg[3] = {d.target[0], d.target[1], d.target[2]};
    double up[3] = {d.up[0], d.up[1], d.up[2]};
    double dir[3] = {tg[0] - o[0], tg[1] - o[1], tg[2] - o[2]};
    double l = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    for (int i = 0; i < 3; ++i) dir[i] /= l;
    double left[3] = {up[1] * dir[2] - up[2] * dir[1], up[2] * dir[0] - up[0] * dir[2], up[0] * dir[1] - up[1] * dir[0]};
    l = std::sqrt(left[0] * left[0] + left[1] * left[1] + left[2] * left[2]);
    for (int i = 0; i < 3; ++i) left[i] /= l;
    double nup[3] = {dir[1] * left[2] - dir[2] * left[1], dir[2] * left[0] - dir[0] * left[2], dir[0] * left[1] - dir[1] * left[0]};
    for (int i = 0; i < 3; ++i) {
        c.o[i] = (float) o[i]; c.left[i] = (float) left[i]; c.up[i] = (float) nup[i]; c.dir[i] = (float) dir[i];
    }
    double th = std::tan(0.5 * d.fov_deg * M_PI / 180.0);
    double aspect = (double) d.width / (double) d.height;
    c.tan_half = (float) th;
    c.dx[0] = (float) (-2.0 * th / d.width); c.dx[1] = 0.f; c.dx[2] = 0.f;
    c.dy[0] = 0.f; c.dy[1] = (float) (-2.0 * th / aspect / d.height); c.dy[2] = 0.f;
    c.rect_x = (float) th; c.rect_y = (float) (th / aspect);
    c.normalization = (float) (1.0 / ((2.0 * th) * (2.0 * th / aspect)));
    c.w = d.width; c.h = d.height;
}

// Host-side form of a scene: everything sms_ctx_create uploads, built without a device.
struct HostNormalmap { DNormalmap m; std::vector<float2> slope; std::vector<float> lean; std::vector<float4> rgbn; };
struct HostScene {
    std::vector<HostNormalmap> nmaps;
    std::vector<DEmitter> emitters;
    std::vector<int> em_ss, em_ms, em_all;
    bool area_emitters = false;
    std::vector<DShape> shapes;
    std::vector<int> analytic, casters, casters_ms;
    std::vector<float4> vpos, vnrm;
    std::vector<float2> vuv;
    std::vector<uint4> faces;
    std::vector<float> cdf, pmf;
    std::vector<DNode> nodes;
    int bvh_depth = 0;
    std::vector<float4> tris;
    float box_c[3] = {0.f, 0.f, 0.f}, box_inv = 1.f, bs_dist = 0.f;
    bool has_cam = false;
    DCamera cam;

    // device bytes of the scene arrays (what upload_scene allocates)
    uint64_t device_bytes() const {
        uint64_t b = 0;
        for (const HostNormalmap &m : nmaps)
            b += m.slope.size() * sizeof(float2) + m.lean.size() * sizeof(float) + m.rgbn.size() * sizeof(float4) +
                 sizeof(DNormalmap);
        b += emitters.size() * sizeof(DEmitter) + (em_ss.size() + em_ms.size() + em_all.size()) * sizeof(int);
        b += shapes.size() * sizeof(DShape) + (analytic.size() + casters.size() + casters_ms.size()) * sizeof(int);
        b += (vpos.size() + vnrm.size()) * sizeof(float4) + vuv.size() * sizeof(float2) + faces.size() * sizeof(uint4);
        b += (cdf.size() + pmf.size()) * sizeof(float) + nodes.size() * sizeof(DNode) + tris.size() * sizeof(float4);
        return b;
    }
};

// Validates a scene description and builds its host form (normal-map slopes and LEAN pyramid,
// shape records, mesh arrays, the BVH); no device call.  On failure *err names the problem.
sms_status prepare_scene(const sms_scene_desc *desc, HostScene &H, std::string &err) {
    auto bad = [&](sms_status code, const char *msg) { err = msg; return code; };
    if (desc->n_shapes <= 0 || !desc->shapes) return bad(SMS_E_INVALID, "scene has no shapes");
    if (desc->n_emitters < 0 || (desc->n_emitters > 0 && !desc->emitters) || desc->n_normalmaps < 0 ||
        (desc->n_normalmaps > 0 && !desc->normalmaps))
        return bad(SMS_E_INVALID, "bad emitter or normal-map arrays");
    // normal maps: level-0 slopes (src/librender/normalmap.cpp:46-66)
    for (int i = 0; i < desc->n_normalmaps; ++i) {
        const sms_normalmap_desc &nd = desc->normalmaps[i];
        if (nd.res <= 0 || (nd.res & (nd.res - 1)) != 0 || !(nd.rgb || nd.rgb32f))
            return bad(SMS_E_INVALID, "normal map must be square power-of-two");
        if (nd.encoded_normals != 0 && nd.encoded_normals != 1) return bad(SMS_E_INVALID, "encoded_normals must be 0 or 1");
        const size_t res = (size_t) nd.res, stride = res + 1;
        // padded layout: column res and row res replicate column/row res-1 (edge clamp)
        std::vector<float2> sl(stride * stride);
        std::vector<float4> rgbn(nd.encoded_normals ? stride * stride : 0);
        for (size_t y = 0; y < stride; ++y) {
            for (size_t x = 0; x < stride; ++x) {
                size_t s = std::min(y, res - 1) * res + std::min(x, res - 1);
                float r, g, b;
                if (nd.rgb) {
                    r = (float) nd.rgb[3 * s] / 255.f; g = (float) nd.rgb[3 * s + 1] / 255.f; b = (float) nd.rgb[3 * s + 2] / 255.f;
                } else {
                    r = nd.rgb32f[3 * s]; g = nd.rgb32f[3 * s + 1]; b = nd.rgb32f[3 * s + 2];
                }
                if (nd.encoded_normals) rgbn[y * stride + x] = make_float4(r, g, b, 0.f);
                float vx = 2.f * r - 1.f, vy = 2.f * g - 1.f, vz = 2.f * b - 1.f;
                float inv = 1.f / std::sqrt((vx * vx + vy * vy) + vz * vz);
                float nx = vx * inv, ny = vy * inv, nz = vz * inv;
                sl[y * stride + x] = make_float2(-nx / nz, -ny / nz);
            }
        }
        DNormalmap m;
        memset(&m, 0, sizeof(m));
        m.res = nd.res; m.stride = (int) stride;
        // LEAN map (normalmap.cpp:36-87): level 0 = (s_x, s_y, s_x^2 + b, s_y^2 + b, s_x s_y) with
        // the base Beckmann variance b = 0.5 * 1e-4^2; coarser levels are the two-tap box resample
        // to half resolution (weights 0.5), along x first, then along y (bitmap.cpp:234-312)
        const float base_alpha = 0.0001f, base_sigma_2 = 0.5f * base_alpha * base_alpha;
        std::vector<float> lean((size_t) res * res * 5);
        for (size_t y = 0; y < res; ++y)
            for (size_t x = 0; x < res; ++x) {
                float2 v = sl[y * stride + x];
                float *d = &lean[(y * res + x) * 5];
                d[0] = v.x; d[1] = v.y; d[2] = v.x * v.x + base_sigma_2; d[3] = v.y * v.y + base_sigma_2;
                d[4] = v.x * v.y;
            }
        m.levels = 1;
        size_t off = 0;
        for (size_t r = res; r > 1; r /= 2) {
            if (m.levels >= SMS_MAX_LEAN_LEVELS) return bad(SMS_E_INVALID, "normal map too large");
            const size_t h = r / 2;
            std::vector<float> tmp(h * r * 5);
            for (size_t y = 0; y < r; ++y)
                for (size_t x = 0; x < h; ++x)
                    for (int c = 0; c < 5; ++c)
                        tmp[(y * h + x) * 5 + c] = lean[off + (y * r + 2 * x) * 5 + c] * 0.5f +
                                                   lean[off + (y * r + 2 * x + 1) * 5 + c] * 0.5f;
            const size_t next = off + r * r * 5;
            lean.resize(next + h * h * 5);
            for (size_t y = 0; y < h; ++y)
                for (size_t x = 0; x < h; ++x)
                    for (int c = 0; c < 5; ++c)
                        lean[next + (y * h + x) * 5 + c] = tmp[((2 * y) * h + x) * 5 + c] * 0.5f +
                                                           tmp[((2 * y + 1) * h + x) * 5 + c] * 0.5f;
            m.lean_off[m.levels] = (long long) next;
            m.levels++;
            off = next;
        }
        if (nd.encoded_normals) {
            // the RGB pyramid's coarsest texel, by the same two-tap resample (normalmap.cpp:70-87)
            m.encoded = 1;
            std::vector<float> cur(res * res * 3), tmp;
            for (size_t y = 0; y < res; ++y)
                for (size_t x = 0; x < res; ++x) {
                    const float4 v = rgbn[y * stride + x];
                    cur[(y * res + x) * 3] = v.x; cur[(y * res + x) * 3 + 1] = v.y; cur[(y * res + x) * 3 + 2] = v.z;
                }
            for (size_t r = res; r > 1; r /= 2) {
                const size_t h = r / 2;
                tmp.assign(h * r * 3, 0.f);
                for (size_t y = 0; y < r; ++y)
                    for (size_t x = 0; x < h; ++x)
                        for (int c = 0; c < 3; ++c)
                            tmp[(y * h + x) * 3 + c] = cur[(y * r + 2 * x) * 3 + c] * 0.5f + cur[(y * r + 2 * x + 1) * 3 + c] * 0.5f;
                cur.assign(h * h * 3, 0.f);
                for (size_t y = 0; y < h; ++y)
                    for (size_t x = 0; x < h; ++x)
                        for (int c = 0; c < 3; ++c)
                            cur[(y * h + x) * 3 + c] = tmp[((2 * y) * h + x) * 3 + c] * 0.5f + tmp[((2 * y + 1) * h + x) * 3 + c] * 0.5f;
            }
            m.top_rgb[0] = cur[0]; m.top_rgb[1] = cur[1]; m.top_rgb[2] = cur[2];
        }
        H.nmaps.push_back(HostNormalmap{m, std::move(sl), std::move(lean), std::move(rgbn)});
    }
    std::vector<DEmitter> &emitters = H.emitters;
    std::vector<int> &em_ss = H.em_ss, &em_ms = H.em_ms, &em_all = H.em_all;
    for (int i = 0; i < desc->n_emitters; ++i) {
        const sms_emitter_desc &e = desc->emitters[i];
        DEmitter d;
        memset(&d, 0, sizeof(d));
        d.type = e.type; d.shape = e.shape;
        for (int a = 0; a < 3; ++a) {
            d.radiance[a] = e.radiance[a]; d.position[a] = e.position[a]; d.intensity[a] = e.intensity[a];
            d.direction[a] = e.direction[a]; d.irradiance[a] = e.irradiance[a];
        }
        if (e.type != SMS_EMITTER_AREA && e.type != SMS_EMITTER_POINT && e.type != SMS_EMITTER_DIRECTIONAL)
            return bad(SMS_E_INVALID, "unknown emitter type");
        if (e.type == SMS_EMITTER_DIRECTIONAL &&
            !(e.direction[0] * e.direction[0] + e.direction[1] * e.direction[1] + e.direction[2] * e.direction[2] > 0.f))
            return bad(SMS_E_INVALID, "directional emitter without a direction");
        d.ss = e.caustic_ss; d.ms = e.caustic_ms;
        if (e.type == SMS_EMITTER_AREA && (e.shape < 0 || e.shape >= desc->n_shapes))
            return bad(SMS_E_INVALID, "area emitter without a shape");
        emitters.push_back(d);
        if (e.type == SMS_EMITTER_AREA) H.area_emitters = true;
        em_all.push_back(i);
        if (e.caustic_ss) em_ss.push_back(i);
        if (e.caustic_ms) em_ms.push_back(i);
    }
    std::vector<DShape> &shapes = H.shapes;
    std::vector<int> &analytic = H.analytic, &casters = H.casters, &casters_ms = H.casters_ms;
    std::vector<float4> &vpos = H.vpos, &vnrm = H.vnrm;
    std::vector<float2> &vuv = H.vuv;
    std::vector<uint4> &faces = H.faces;
    std::vector<float> &cdf = H.cdf, &pmf = H.pmf;
    for (int i = 0; i < desc->n_shapes; ++i) {
        const sms_shape_desc &d = desc->shapes[i];
        DShape sh;
        memset(&sh, 0, sizeof(sh));
        sh.type = d.type; sh.bsdf = d.bsdf; sh.flags = d.flags; sh.emitter = d.emitter;
        sh.cx = d.center[0]; sh.cy = d.center[1]; sh.cz = d.center[2];
        sh.radius = d.radius; sh.inv_radius = 1.f / d.radius;
        memcpy(sh.M, d.to_world, sizeof(sh.M));
        if (d.bsdf == SMS_BSDF_DIFFUSE) { sh.eta_r = 0.f; sh.k = 1.f; }
        else if (d.bsdf == SMS_BSDF_DIELECTRIC) { sh.eta_r = d.eta; sh.k = 0.f; }
        else { sh.eta_r = d.eta; sh.k = d.k; }
        sh.eta_vtx = sh.k == 0.f ? hmean_same(sh.eta_r) : 1.f;
        if (!(d.roughness >= 0.f) || !std::isfinite(d.roughness)) return bad(SMS_E_INVALID, "roughness must be finite and >= 0");
        if (d.roughness > 0.f && d.bsdf != SMS_BSDF_CONDUCTOR)
            return bad(SMS_E_UNSUPPORTED, "rough BSDFs other than the rough conductor are not built");
        sh.alpha = d.roughness;
        for (int a = 0; a < 3; ++a) sh.refl[a] = d.reflectance[a];
        sh.normalmap = d.normalmap; sh.tiles = d.tiles;
        if ((d.flags & SMS_FLAG_GLINTY) && d.bsdf != SMS_BSDF_NORMALMAP)
            return bad(SMS_E_INVALID, "glinty shapes need the normal-mapped BSDF");
        if (d.bsdf == SMS_BSDF_NORMALMAP && (d.normalmap < 0 || d.normalmap >= desc->n_normalmaps))
            return bad(SMS_E_INVALID, "normal-mapped shape without a normal map");
        if ((d.flags & SMS_FLAG_GLINTY) && desc->normalmaps[d.normalmap].encoded_normals)
            return bad(SMS_E_UNSUPPORTED, "glinty shapes need a slope normal map (encoded_normals 0)");
        sh.bbox_center[0] = sh.cx; sh.bbox_center[1] = sh.cy; sh.bbox_center[2] = sh.cz;
        if (d.type == SMS_SHAPE_SPHERE) {
            sh.inv_surface_area = 1.f / (4.f * PI_F * sh.radius * sh.radius);
            analytic.push_back(i);
        } else if (d.type == SMS_SHAPE_RECTANGLE) {
            invert_affine(sh.M, sh.Minv);
            sh.frs[0] = 2.f * sh.M[0]; sh.frs[1] = 2.f * sh.M[4]; sh.frs[2] = 2.f * sh.M[8];
            sh.frt[0] = 2.f * sh.M[1]; sh.frt[1] = 2.f * sh.M[5]; sh.frt[2] = 2.f * sh.M[9];
            double nx = sh.Minv[8], ny = sh.Minv[9], nz = sh.Minv[10];
            double l = std::sqrt(nx * nx + ny * ny + nz * nz);
            sh.frn[0] = (float) (nx / l); sh.frn[1] = (float) (ny / l); sh.frn[2] = (float) (nz / l);
            float cs[3] = {sh.frs[1] * sh.frt[2] - sh.frs[2] * sh.frt[1], sh.frs[2] * sh.frt[0] - sh.frs[0] * sh.frt[2],
                           sh.frs[0] * sh.frt[1] - sh.frs[1] * sh.frt[0]};
            sh.inv_surface_area = 1.f / std::sqrt((cs[0] * cs[0] + cs[1] * cs[1]) + cs[2] * cs[2]);
            sh.bbox_center[0] = sh.M[3]; sh.bbox_center[1] = sh.M[7]; sh.bbox_center[2] = sh.M[11];
            analytic.push_back(i);
        } else if (d.type == SMS_SHAPE_MESH) {
            if (d.n_faces <= 0 || d.n_vertices <= 0 || !d.positions || !d.faces)
                return bad(SMS_E_INVALID, "mesh without faces/positions");
            int vbase = (int) vpos.size();
            sh.fbase = (int) faces.size(); sh.nf = d.n_faces;
            sh.has_n = d.normals != nullptr; sh.has_uv = d.texcoords != nullptr;
            float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (int v = 0; v < d.n_vertices; ++v) {
                vpos.push_back(make_float4(d.positions[3 * v], d.positions[3 * v + 1], d.positions[3 * v + 2], 0.f));
                vnrm.push_back(d.normals ? make_float4(d.normals[3 * v], d.normals[3 * v + 1], d.normals[3 * v + 2], 0.f)
                                         : make_float4(0.f, 0.f, 0.f, 0.f));
                vuv.push_back(d.texcoords ? make_float2(d.texcoords[2 * v], d.texcoords[2 * v + 1]) : make_float2(0.f, 0.f));
                for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], d.positions[3 * v + a]); hi[a] = std::max(hi[a], d.positions[3 * v + a]); }
            }
            for (int a = 0; a < 3; ++a) sh.bbox_center[a] = 0.5f * (lo[a] + hi[a]);
            sh.cdf_offset = (int) cdf.size();
            double sum = 0.0;
            for (int f = 0; f < d.n_faces; ++f) {
                uint32_t i0 = d.faces[3 * f], i1 = d.faces[3 * f + 1], i2 = d.faces[3 * f + 2];
                if (i0 >= (uint32_t) d.n_vertices || i1 >= (uint32_t) d.n_vertices || i2 >= (uint32_t) d.n_vertices)
                    return bad(SMS_E_INVALID, "mesh face index out of range");
                faces.push_back(make_uint4(vbase + i0, vbase + i1, vbase + i2, (uint32_t) i));
                const float *p0 = &d.positions[3 * i0], *p1 = &d.positions[3 * i1], *p2 = &d.positions[3 * i2];
                float e0[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]}, e1[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
                float cr[3] = {e0[1] * e1[2] - e0[2] * e1[1], e0[2] * e1[0] - e0[0] * e1[2], e0[0] * e1[1] - e0[1] * e1[0]};
                float a = .5f * std::sqrt((cr[0] * cr[0] + cr[1] * cr[1]) + cr[2] * cr[2]);
                pmf.push_back(a);
                sum += (double) a;
                cdf.push_back((float) sum);
            }
            sh.area_sum = (float) sum; sh.area_norm = (float) (1.0 / sum);
            sh.inv_surface_area = sh.area_norm;
        } else {
            return bad(SMS_E_INVALID, "unknown shape type");
        }
        if (d.flags & SMS_FLAG_CAUSTIC_CASTER_SS) casters.push_back(i);
        if (d.flags & SMS_FLAG_CAUSTIC_CASTER_MS) casters_ms.push_back(i);
        shapes.push_back(sh);
    }
    // BVH over all mesh triangles
    std::vector<BItem> items(faces.size());
    for (size_t t = 0; t < faces.size(); ++t) {
        BItem &it = items[t];
        it.tri = (int) t;
        uint32_t ids[3] = {faces[t].x, faces[t].y, faces[t].z};
        for (int a = 0; a < 3; ++a) { it.bmin[a] = INFINITY; it.bmax[a] = -INFINITY; }
        for (int k = 0; k < 3; ++k) {
            float p[3] = {vpos[ids[k]].x, vpos[ids[k]].y, vpos[ids[k]].z};
            for (int a = 0; a < 3; ++a) { it.bmin[a] = std::min(it.bmin[a], p[a]); it.bmax[a] = std::max(it.bmax[a], p[a]); }
        }
        for (int a = 0; a < 3; ++a) it.c[a] = 0.5f * (it.bmin[a] + it.bmax[a]);
    }
    BvhOut bvh = build_bvh(items);
    if (bvh.max_depth > SMS_BVH_STACK)
        return bad(SMS_E_UNSUPPORTED, "BVH deeper than the traversal stack");
    H.nodes = std::move(bvh.nodes);
    H.bvh_depth = bvh.max_depth;
    std::vector<float4> &tris = H.tris;
    tris.resize(3 * bvh.order.size());
    std::vector<int> face_prim_base(desc->n_shapes, 0);
    for (int i = 0; i < desc->n_shapes; ++i) face_prim_base[i] = shapes[i].fbase;
    for (size_t k = 0; k < bvh.order.size(); ++k) {
        const uint4 f = faces[bvh.order[k]];
        float4 p0 = vpos[f.x], p1 = vpos[f.y], p2 = vpos[f.z];
        int shape = (int) f.w, prim = bvh.order[k] - face_prim_base[shape];
        tris[3 * k] = make_float4(p0.x, p0.y, p0.z, i2f(shape));
        tris[3 * k + 1] = make_float4(p1.x - p0.x, p1.y - p0.y, p1.z - p0.z, i2f(prim));
        tris[3 * k + 2] = make_float4(p2.x - p0.x, p2.y - p0.y, p2.z - p0.z, 0.f);
    }
    // scene bounding box (positions of the learned proposal's inputs are normalised to it)
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    auto grow = [&](float x, float y, float z) {
        const float v[3] = {x, y, z};
        for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], v[a]); hi[a] = std::max(hi[a], v[a]); }
    };
    for (const float4 &v : vpos) grow(v.x, v.y, v.z);
    for (const DShape &sh : shapes) {
        if (sh.type == SMS_SHAPE_SPHERE) {
            grow(sh.cx - sh.radius, sh.cy - sh.radius, sh.cz - sh.radius);
            grow(sh.cx + sh.radius, sh.cy + sh.radius, sh.cz + sh.radius);
        } else if (sh.type == SMS_SHAPE_RECTANGLE) {
            for (int cy = -1; cy <= 1; cy += 2)
                for (int cx = -1; cx <= 1; cx += 2)
                    grow(sh.M[0] * cx + sh.M[1] * cy + sh.M[3], sh.M[4] * cx + sh.M[5] * cy + sh.M[7],
                         sh.M[8] * cx + sh.M[9] * cy + sh.M[11]);
        }
    }
    float half = 0.f;
    for (int a = 0; a < 3; ++a) { H.box_c[a] = 0.5f * (lo[a] + hi[a]); half = std::max(half, 0.5f * (hi[a] - lo[a])); }
    H.box_inv = half > 0.f ? 1.f / half : 1.f;
    // BoundingBox::bounding_sphere of the scene (radius = |max - centre|), enlarged by
    // DirectionalEmitter::set_scene (src/emitters/directional.cpp:63-68); sample_direction places
    // the emitter 2 radii away (:112)
    {
        double r2 = 0.0;
        for (int a = 0; a < 3; ++a) {
            double c = 0.5 * ((double) lo[a] + (double) hi[a]), d = (double) hi[a] - c;
            r2 += d * d;
        }
        double r = std::sqrt(r2) * (1.0 + (double) RAY_EPS_F);
        if (!(r > (double) RAY_EPS_F)) r = (double) RAY_EPS_F;
        H.bs_dist = (float) (2.0 * r);
    }
    H.has_cam = desc->camera.width > 0 && desc->camera.height > 0;
    if (H.has_cam) camera_setup(H.cam, desc->camera);
    return SMS_OK;
}

// Copies a prepared scene to the context's device.
sms_status upload_scene(sms_ctx *ctx, HostScene &H, const sms_config *cfg) {
    DScene &S = ctx->scene;
    memset(&S, 0, sizeof(S));
    S.cfg = *cfg;
    sms_status st;
    std::vector<DNormalmap> nmaps;
    for (HostNormalmap &hm : H.nmaps) {
        DNormalmap m = hm.m;
        if ((st = upload(ctx, hm.slope, &m.slope))) return st;
        if ((st = upload(ctx, hm.lean, &m.lean))) return st;
        if (!hm.rgbn.empty() && (st = upload(ctx, hm.rgbn, &m.rgbn))) return st;
        nmaps.push_back(m);
    }
    ctx->area_emitters = H.area_emitters;
    ctx->bvh_depth = H.bvh_depth;
    for (int a = 0; a < 3; ++a) S.box_c[a] = H.box_c[a];
    S.box_inv = H.box_inv;
    S.bs_dist = H.bs_dist;
    ctx->h_vuv = H.vuv; ctx->h_faces = H.faces; ctx->h_shapes = H.shapes;
    if ((st = upload(ctx, H.shapes, &S.shapes))) return st;
    S.n_shapes = (int) H.shapes.size();
    if ((st = upload(ctx, H.analytic, &S.analytic))) return st;
    S.n_analytic = (int) H.analytic.size();
    if ((st = upload(ctx, H.emitters, &S.emitters))) return st;
    S.n_emitters = (int) H.emitters.size();
    if ((st = upload(ctx, H.em_ss, &S.emitters_ss))) return st;
    S.n_emitters_ss = (int) H.em_ss.size();
    if ((st = upload(ctx, H.em_all, &S.emitters_all))) return st;
    S.n_emitters_all = (int) H.em_all.size();
    if ((st = upload(ctx, H.casters, &S.casters_ss))) return st;
    S.n_casters_ss = (int) H.casters.size();
    if ((st = upload(ctx, H.em_ms, &S.emitters_ms))) return st;
    S.n_emitters_ms = (int) H.em_ms.size();
    if ((st = upload(ctx, H.casters_ms, &S.casters_ms))) return st;
    S.n_casters_ms = (int) H.casters_ms.size();
    if ((st = upload(ctx, nmaps, &S.nmaps))) return st;
    if ((st = upload(ctx, H.vpos, &S.vpos))) return st;
    if ((st = upload(ctx, H.vnrm, &S.vnrm))) return st;
    if ((st = upload(ctx, H.vuv, &S.vuv))) return st;
    if ((st = upload(ctx, H.faces, &S.faces))) return st;
    if ((st = upload(ctx, H.nodes, &S.nodes))) return st;
    S.n_nodes = (int) H.nodes.size();
    if ((st = upload(ctx, H.tris, &S.tris))) return st;
    if ((st = upload(ctx, H.cdf, &S.cdf))) return st;
    if ((st = upload(ctx, H.pmf, &S.pmf))) return st;
    S.has_cam = H.has_cam;
    if (S.has_cam) S.cam = H.cam;
    return SMS_OK;
}

sms_status check_config(sms_ctx *ctx, const sms_config *cfg) {
    if (!cfg) return fail(ctx, SMS_E_INVALID, "null config");
    if (cfg->max_iterations < 0) return fail(ctx, SMS_E_INVALID, "max_iterations < 0");
    return SMS_OK;
}

// Multi-scatter is built for smooth specular chains of 1..SMS_MS_MAX_BOUNCES vertices.
sms_status check_ms(sms_ctx *ctx) {
    const sms_config &cfg = ctx->scene.cfg;
    if (cfg.bounces < 1 || cfg.bounces > SMS_MS_MAX_BOUNCES)
        return fail(ctx, SMS_E_UNSUPPORTED, "multi-scatter chains of 1..4 vertices are built");
    return SMS_OK;
}

__global__ void k_zero_u64(unsigned long long *p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}


} // namespace

// Frame start: the estimator and camera seed slots take `seed` and the work counters start at
// zero -- a kernel rather than a memset, so that a captured frame graph holds kernel nodes only.
__global__ void k_frame_begin(unsigned long long *counters, int n, uint64_t *seeds, uint64_t seed) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) counters[i] = 0ull;
    if (i == 0) {
        seeds[SEED_EST] = seed;
        seeds[SEED_CAM] = seed;
    }
}

__global__ void k_set_seed(uint64_t *slot, uint64_t seed) {
    if (threadIdx.x == 0) *slot = seed;
}

// Add a frame's counters to total[0..8]: the spread copies of the stats summed (bernoulli_max
// as a maximum) into total[0..7] in sms_stats order, the overflow count into total[8].  One
// wavefront: lane s reads copy s of every counter, then butterfly reductions.
static_assert(SPREAD == 64, "k_stats_fold maps one lane to each spread copy");
__global__ void k_stats_fold(const unsigned long long *counters, unsigned long long *total) {
    const int s = threadIdx.x;
    unsigned long long v[9];
    for (int j = 0; j < 8; ++j) v[j] = counters[STATS_OFF + s * SPREAD_STRIDE + j];
    v[8] = counters[OVER_OFF + s * SPREAD_STRIDE] + (s == 0 ? counters[1] : 0ull);
    for (int off = 32; off > 0; off >>= 1)
        for (int j = 0; j < 9; ++j) {
            const unsigned long long o = __shfl_xor(v[j], off, 64);
            v[j] = j == 5 ? (o > v[j] ? o : v[j]) : v[j] + o;
        }
    if (s == 0)
        for (int j = 0; j < 9; ++j) total[j] = j == 5 ? (v[j] > total[j] ? v[j] : total[j]) : total[j] + v[j];
}

sms_status estimator_stats(sms_ctx *ctx, sms_stats *stats);

sms_status frame_begin(sms_ctx *ctx, uint64_t seed) {
    hipLaunchKernelGGL(k_frame_begin, dim3((N_COUNTERS + 255) / 256), dim3(256), 0, ctx->stream, ctx->counters,
                       N_COUNTERS, ctx->seeds, seed);
    HIP_TRY(ctx, hipGetLastError());
    return SMS_OK;
}

namespace {

sms_status ensure_work(sms_ctx *ctx, size_t n_overflow, size_t n_part) {
    if (!ctx->counters) HIP_TRY(ctx, hipMalloc((void **) &ctx->counters, N_COUNTERS * sizeof(unsigned long long)));
    if (!ctx->seeds) {
        HIP_TRY(ctx, hipMalloc((void **) &ctx->seeds, SEED_SLOTS * sizeof(uint64_t)));
        HIP_TRY(ctx, hipMemset(ctx->seeds, 0, SEED_SLOTS * sizeof(uint64_t)));
    }
    if (n_overflow > ctx->overflow_cap) {
        ctx->gen++;
        if (ctx->overflow) HIP_TRY(ctx, hipFree(ctx->overflow));
        ctx->overflow = nullptr;
        size_t cap = std::max(n_overflow, (size_t) 1 << 16);
        HIP_TRY(ctx, hipMalloc((void **) &ctx->overflow, cap * sizeof(OverflowRec)));
        ctx->overflow_cap = cap;
    }
    if (n_part > ctx->part_cap) {
        ctx->gen++;
        if (ctx->part) HIP_TRY(ctx, hipFree(ctx->part));
        if (ctx->invp) HIP_TRY(ctx, hipFree(ctx->invp));
        ctx->part = ctx->invp = nullptr;
        HIP_TRY(ctx, hipMalloc((void **) &ctx->part, 3 * n_part * sizeof(float)));
        HIP_TRY(ctx, hipMalloc((void **) &ctx->invp, n_part * sizeof(float)));
        ctx->part_cap = n_part;
    }
    return SMS_OK;
}

// Spread overflow list with sub-lists of sub_cap items (kept for the call that asks).
sms_status ensure_over(sms_ctx *ctx, size_t sub_cap) {
    if (SPREAD * sub_cap > ctx->gover_cap) {
        ctx->gen++;
        if (ctx->gover) HIP_TRY(ctx, hipFree(ctx->gover));
        ctx->gover = nullptr;
        HIP_TRY(ctx, hipMalloc((void **) &ctx->gover, SPREAD * sub_cap * sizeof(OverflowRec)));
        ctx->gover_cap = SPREAD * sub_cap;
    }
    if (ctx->over_sub != sub_cap) ctx->gen++;
    ctx->over_sub = sub_cap;
    return SMS_OK;
}

// Per-sub-list capacity for a lane-per-query pass over n queries with nc items each, one
// block per 256 queries, block b appending to sub-list b % SPREAD.
size_t query_sub_cap(int64_t n, int nc) {
    const size_t items = (size_t) std::max<int64_t>(n, 1) * (size_t) std::max(nc, 1);
    return (items + SPREAD - 1) / SPREAD + 256 * (size_t) std::max(nc, 1);
}

// Multi-wavefront stage of the glint tail: long chains it holds and trials per chain.
constexpr int MW_ITEMS = 1024;
constexpr int MW_TRIALS = 32768;

sms_status ensure_glint_work(sms_ctx *ctx, size_t n) {
    if (n > ctx->trial_cap) {
        ctx->gen++;
        void **bufs[4] = {(void **) &ctx->trial, (void **) &ctx->res, (void **) &ctx->long_items,
                          (void **) &ctx->seq_items};
        for (void **b : bufs) {
            if (*b) HIP_TRY(ctx, hipFree(*b));
            *b = nullptr;
        }
        // spread lists: SPREAD sub-lists; a sub-list takes the items of every SPREAD-th block
        // (at most 256 each) or of every SPREAD-th trial slot
        const size_t sub_cap = (n + SPREAD - 1) / SPREAD + 256;
        HIP_TRY(ctx, hipMalloc((void **) &ctx->trial, SPREAD * sub_cap * sizeof(TrialRec)));
        HIP_TRY(ctx, hipMalloc((void **) &ctx->res, SPREAD * sub_cap * sizeof(GlintRes)));
        HIP_TRY(ctx, hipMalloc((void **) &ctx->long_items, n * sizeof(OverflowRec)));
        HIP_TRY(ctx, hipMalloc((void **) &ctx->seq_items, n * sizeof(OverflowRec)));
        ctx->trial_cap = n;
        ctx->sub_cap = sub_cap;
    }
    {
        sms_status st = ensure_over(ctx, ctx->sub_cap);
        if (st) return st;
    }
    if (!ctx->mw_log) {
        ctx->gen++;
        HIP_TRY(ctx, hipMalloc((void **) &ctx->mw_found, MW_ITEMS * sizeof(long long)));
        HIP_TRY(ctx, hipMalloc((void **) &ctx->mw_next, MW_ITEMS * sizeof(unsigned long long)));
        HIP_TRY(ctx, hipMalloc((void **) &ctx->mw_log, (size_t) MW_ITEMS * MW_TRIALS * sizeof(uint16_t)));
    }
    // the per-trial log stores iteration counts in 15 bits
    const int mw_items = ctx->scene.cfg.max_iterations < 0x8000 ? ctx->long_chains : 0;
    if (mw_items != ctx->mw_items || ctx->long_trials != ctx->mw_trials) ctx->gen++;
    ctx->mw_items = mw_items;
    ctx->mw_trials = ctx->long_trials;
    return SMS_OK;
}

sms_schedule schedule_of(const sms_ctx *ctx) {
    sms_schedule s;
    memset(&s, 0, sizeof(s));
    s.lane_trials = ctx->lane_trials; s.blocks_per_cu = ctx->blocks_per_cu;
    s.trial_group = ctx->trial_group; s.tail_rounds = ctx->tail_rounds;
    s.long_chains = ctx->long_chains; s.long_trials = ctx->long_trials;
    s.ss_first_pass = ctx->ss_first_pass;
    return s;
}

void apply_schedule(sms_ctx *ctx, const sms_schedule &s) {
    ctx->gen++;
    ctx->lane_trials = s.lane_trials; ctx->blocks_per_cu = s.blocks_per_cu;
    ctx->trial_group = s.trial_group; ctx->tail_rounds = s.tail_rounds;
    ctx->long_chains = s.long_chains; ctx->long_trials = s.long_trials;
    ctx->ss_first_pass = s.ss_first_pass;
}

EstimatorWork work_of(sms_ctx *ctx) {
    EstimatorWork W;
    memset(&W, 0, sizeof(W));
    unsigned long long *c = ctx->counters;
    W.cursor = c + 0; W.n_overflow = c + 1; W.coop_cursor = c + 2;
    W.stats = c + STATS_OFF; W.overflow = ctx->overflow; W.lane_trials = ctx->lane_trials;
    W.res = ctx->res;
    W.trials.items = ctx->trial; W.trials.counts = c + TRIALS_OFF; W.trials.sub_cap = ctx->sub_cap;
    W.over.items = ctx->gover; W.over.counts = c + OVER_OFF; W.over.sub_cap = ctx->over_sub;
    W.group = ctx->trial_group; W.tail_rounds = ctx->tail_rounds;
    W.long_list.items = ctx->long_items; W.long_list.count = c + 4; W.long_list.cursor = c + 5;
    W.seq_list.items = ctx->seq_items; W.seq_list.count = c + 6; W.seq_list.cursor = c + 7;
    W.mw_found = ctx->mw_found; W.mw_next = ctx->mw_next; W.mw_log = ctx->mw_log;
    W.mw_items = ctx->mw_items; W.mw_trials = ctx->mw_trials;
    W.prop = proposal_of(ctx);
    return W;
}

sms_status read_stats(sms_ctx *ctx, sms_stats *stats) {
    std::vector<unsigned long long> h(N_COUNTERS);
    HIP_TRY(ctx, hipMemcpyAsync(h.data(), ctx->counters, N_COUNTERS * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (stats) {
        unsigned long long v[8] = {0, 0, 0, 0, 0, 0, 0, 0}, over = 0;
        for (int s = 0; s < SPREAD; ++s) {
            const unsigned long long *cp = &h[STATS_OFF + s * SPREAD_STRIDE];
            for (int j = 0; j < 8; ++j) v[j] = j == 5 ? std::max(v[j], cp[j]) : v[j] + cp[j];
            over += h[OVER_OFF + s * SPREAD_STRIDE];
        }
        memset(stats, 0, sizeof(*stats));
        stats->walks = v[0]; stats->walks_converged = v[1]; stats->newton_iterations = v[2];
        stats->bernoulli_calls = v[3]; stats->bernoulli_trials = v[4]; stats->bernoulli_max = v[5];
        stats->overflow_queries = h[1] + over;
        stats->tail_iterations = v[6];
    }
    return SMS_OK;
}

} // namespace

// statistics of the last estimator call (waits for the stream)
sms_status estimator_stats(sms_ctx *ctx, sms_stats *stats) { return read_stats(ctx, stats); }

namespace {

// Persistent kernels take their work from device counters, so blocks beyond the resident
// limit simply start later and find the queue drained; the default is the occupancy limit,
// at least 8 blocks per CU.
int grid_for(sms_ctx *ctx, int occ) {
    int per_cu = ctx->blocks_per_cu > 0 ? ctx->blocks_per_cu : std::max(8, occ);
    return per_cu * ctx->n_cu;
}

template <typename K>
sms_status occupancy(sms_ctx *ctx, K kernel, int *out) {
    int nb = 0;
    HIP_TRY(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, 256, 0));
    *out = std::max(1, nb);
    return SMS_OK;
}

} // namespace

namespace {

// Device part of sms_ctx_create: stream, events, scene upload, occupancy queries, work buffers.
sms_status create_on_device(sms_ctx *ctx, HostScene &H, const sms_config *cfg) {
    sms_status st;
    HIP_TRY(ctx, hipGetDevice(&ctx->device));
    if ((st = noalloc_init_device(ctx))) return st;
    hipDeviceProp_t prop;
    HIP_TRY(ctx, hipGetDeviceProperties(&prop, ctx->device));
    ctx->n_cu = prop.multiProcessorCount;
    snprintf(ctx->dev_name, sizeof(ctx->dev_name), "%s (%s)", prop.name, prop.gcnArchName);
    HIP_TRY(ctx, hipStreamCreate(&ctx->stream));
    HIP_TRY(ctx, hipEventCreate(&ctx->ev0));
    HIP_TRY(ctx, hipEventCreate(&ctx->ev1));
    HIP_TRY(ctx, hipEventCreate(&ctx->ev2));
    HIP_TRY(ctx, hipEventCreate(&ctx->ev3));
    st = upload_scene(ctx, H, cfg);
    if (!st) st = occupancy(ctx, k_glint_trials_g8, &ctx->occ_glint);
    if (!st) st = occupancy(ctx, k_glint_coop_rounds, &ctx->occ_coop);
    if (!st) st = occupancy(ctx, k_glint_multiwave, &ctx->occ_mw);
    if (!st) st = occupancy(ctx, k_ss_estimate, &ctx->occ_ss);
    if (!st) st = occupancy(ctx, k_ss_walker, &ctx->occ_walker);
    if (!st) st = occupancy(ctx, k_ss_coop, &ctx->occ_sscoop);
    if (!st) st = occupancy(ctx, k_ss_biased, &ctx->occ_biased);
    if (!st) st = occupancy(ctx, k_glint_biased_bvh, &ctx->occ_glint_biased);
    if (!st) st = occupancy(ctx, k_ms_estimate_2, &ctx->occ_ms[0]);
    if (!st) st = occupancy(ctx, k_ms_estimate_4, &ctx->occ_ms[1]);
    if (!st) st = occupancy(ctx, k_ms_persist_2, &ctx->occ_mspersist[0]);
    if (!st) st = occupancy(ctx, k_ms_persist_4, &ctx->occ_mspersist[1]);
    if (!st) st = occupancy(ctx, k_ms_coop_2, &ctx->occ_mscoop[0]);
    if (!st) st = occupancy(ctx, k_ms_coop_4, &ctx->occ_mscoop[1]);
    if (!st) st = occupancy(ctx, k_ms_biased_2, &ctx->occ_msbiased[0]);
    if (!st) st = occupancy(ctx, k_ms_biased_4, &ctx->occ_msbiased[1]);
    if (!st) st = ensure_work(ctx, 0, 0);
    return st;
}

} // namespace

// ==================================================================== C ABI
extern "C" {

void sms_config_default(sms_config *cfg) {
    if (!cfg) return;
    cfg->biased = 0; cfg->twostage = 0; cfg->halfvector_constraints = 0; cfg->mnee_init = 0;
    cfg->step_scale = 1.f; cfg->max_iterations = 20; cfg->solver_threshold = 1e-5f;
    cfg->uniqueness_threshold = 1e-4f; cfg->max_trials = -1; cfg->bounces = 1;
    cfg->bsdf_strategy_only = 0; cfg->sms_strategy_only = 0;
}

sms_status sms_ctx_create(const sms_scene_desc *scene, const sms_config *cfg, sms_ctx **out) {
    if (!out || !scene || !cfg) return SMS_E_INVALID;
    *out = nullptr;
    sms_ctx *ctx = nullptr;
    try {
        sms_status st = check_config(nullptr, cfg);
        if (st) return st;
        // the scene's host form first: a bad description or a host allocation failure is reported
        // before any device is touched
        HostScene H;
        std::string err;
        st = prepare_scene(scene, H, err);
        if (st) {
            fprintf(stderr, "sms_ctx_create: %s\n", err.c_str());
            return st;
        }
        int n_dev = 0;
        if (hipGetDeviceCount(&n_dev) != hipSuccess || n_dev == 0) return SMS_E_NO_DEVICE;
        ctx = new sms_ctx();
        st = create_on_device(ctx, H, cfg);
        if (st) {
            fprintf(stderr, "sms_ctx_create: %s\n", ctx->err.c_str());
            sms_ctx_destroy(ctx);
            return st;
        }
        *out = ctx;
        return SMS_OK;
    } catch (const std::bad_alloc &) {
        if (ctx) sms_ctx_destroy(ctx);
        return SMS_E_OOM;
    } catch (...) {
        if (ctx) sms_ctx_destroy(ctx);
        return SMS_E_INVALID;
    }
}

sms_status sms_scene_footprint(const sms_scene_desc *scene, uint64_t *device_bytes, char *err, int32_t err_len) {
    if (!scene || !device_bytes) return SMS_E_INVALID;
    try {
        HostScene H;
        std::string msg;
        sms_status st = prepare_scene(scene, H, msg);
        if (err && err_len > 0) snprintf(err, (size_t) err_len, "%s", msg.c_str());
        *device_bytes = st ? 0 : H.device_bytes();
        return st;
    } catch (const std::bad_alloc &) {
        if (err && err_len > 0) snprintf(err, (size_t) err_len, "host allocation failed");
        return SMS_E_OOM;
    } catch (...) {
        if (err && err_len > 0) snprintf(err, (size_t) err_len, "unexpected host exception");
        return SMS_E_INVALID;
    }
}


sms_status sms_ctx_destroy(sms_ctx *ctx) {
    return guarded(ctx, [&]() -> sms_status {
        if (!ctx) return SMS_E_INVALID;
        if (ctx->stream) (void) hipStreamSynchronize(ctx->stream);
        for (void *p : ctx->scene_allocs) (void) hipFree(p);
        if (ctx->counters) (void) hipFree(ctx->counters);
        if (ctx->overflow) (void) hipFree(ctx->overflow);
        if (ctx->part) (void) hipFree(ctx->part);
        if (ctx->invp) (void) hipFree(ctx->invp);
        if (ctx->trial) (void) hipFree(ctx->trial);
        if (ctx->long_items) (void) hipFree(ctx->long_items);
        if (ctx->seq_items) (void) hipFree(ctx->seq_items);
        if (ctx->mw_found) (void) hipFree(ctx->mw_found);
        if (ctx->mw_next) (void) hipFree(ctx->mw_next);
        if (ctx->mw_log) (void) hipFree(ctx->mw_log);
        if (ctx->gover) (void) hipFree(ctx->gover);
        if (ctx->res) (void) hipFree(ctx->res);
        if (ctx->seeds) (void) hipFree(ctx->seeds);
        proposal_release(ctx);
        path_release(ctx);
        if (ctx->ev3) (void) hipEventDestroy(ctx->ev3);
        if (ctx->ev0) (void) hipEventDestroy(ctx->ev0);
        if (ctx->ev1) (void) hipEventDestroy(ctx->ev1);
        if (ctx->ev2) (void) hipEventDestroy(ctx->ev2);
        if (ctx->stream) (void) hipStreamDestroy(ctx->stream);
        delete ctx;
        return SMS_OK;
    });
}

sms_status sms_ctx_set_config(sms_ctx *ctx, const sms_config *cfg) {
    return guarded(ctx, [&]() -> sms_status {
        if (!ctx) return SMS_E_INVALID;
        sms_status st = check_config(ctx, cfg);
        if (st) return st;
        ctx->scene.cfg = *cfg;
        ctx->gen++;
        return SMS_OK;
    });
}

const char *sms_ctx_last_error(const sms_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

sms_status sms_ctx_set_schedule(sms_ctx *ctx, const sms_schedule *s) {
    return guarded(ctx, [&]() -> sms_status {
        if (!ctx || !s) return SMS_E_INVALID;
        const int g = s->trial_group;
        if (s->lane_trials < 1 || s->blocks_per_cu < 0 || s->tail_rounds < 0 ||
            !(g == 0 || g == 1 || g == 2 || g == 4 || g == 8 || g == 16) ||
            s->long_chains < 0 || s->long_chains > MW_ITEMS || s->long_trials < 64 || s->long_trials > MW_TRIALS ||
            s->long_trials % 64 != 0 || s->ss_first_pass < 0 || s->ss_first_pass > 1)
            return fail(ctx, SMS_E_INVALID, "bad schedule");
        sms_schedule v = *s;
        if (v.trial_group == 0) v.trial_group = 8;
        apply_schedule(ctx, v);
        return SMS_OK;
    });
}

sms_status sms_ctx_get_schedule(const sms_ctx *ctx, sms_schedule *s) {
    return guarded(const_cast<sms_ctx *>(ctx), [&]() -> sms_status {
        if (!ctx || !s) return SMS_E_INVALID;
        *s = schedule_of(ctx);
        return SMS_OK;
    });
}

sms_status sms_ctx_device_info(const sms_ctx *ctx, char *name, int32_t name_len, int32_t *n_cu) {
    return guarded(const_cast<sms_ctx *>(ctx), [&]() -> sms_status {
        if (!ctx) return SMS_E_INVALID;
        if (name && name_len > 0) snprintf(name, (size_t) name_len, "%s", ctx->dev_name);
        if (n_cu) *n_cu = ctx->n_cu;
        return SMS_OK;
    });
}

sms_status sms_ctx_sync(sms_ctx *ctx) {
    return guarded(ctx, [&]() -> sms_status {
        if (!ctx) return SMS_E_INVALID;
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        return SMS_OK;
    });
}

sms_status sms_ctx_last_kernel_ms(const sms_ctx *ctx, float *ms) {
    return guarded(const_cast<sms_ctx *>(ctx), [&]() -> sms_status {
        if (!ctx || !ms) return SMS_E_INVALID;
        ms[0] = ctx->last_ms;
        ms[1] = ctx->last_ms2;
        return SMS_OK;
    });
}

sms_status sms_ss_walk_batch(sms_ctx *ctx, const sms_ss_walk_in *in, int64_t n, sms_ss_walk_out *out) {
    return guarded(ctx, [&]() -> sms_status {
        if (!ctx || n < 0 || (n > 0 && (!in || !out))) return fail(ctx, SMS_E_INVALID, "bad walk batch arguments");
        if (n == 0) return SMS_OK;
        int blocks = (int) ((n + 255) / 256);
        hipLaunchKernelGGL(k_ss_walk, dim3(blocks), dim3(256), 0, ctx->stream, ctx->scene, in, n, out);
        HIP_TRY(ctx, hipGetLastError());
        return SMS_OK;
    });
}

sms_status sms_glint_walk_batch(sms_ctx *ctx, const sms_glint_walk_in *in, int64_t n, sms_glint_walk_out *out) {
    return guarded(ctx, [&]() -> sms_status {
        if (!ctx || n < 0 || (n > 0 && (!in || !out))) return fail(ctx, SMS_E_INVALID, "bad walk batch arguments");
        if (n == 0) return SMS_OK;
        if (!ctx->scene.nmaps) return fail(ctx, SMS_E_INVALID, "scene has no normal map");
        int blocks = (int) ((n + 255) / 256);
        hipLaunchKernelGGL(k_glint_walk, dim3(blocks), dim3(256), 0, ctx->stream, ctx->scene, in, n, out);
        HIP_TRY(ctx, hipGetLastError());
        return SMS_OK;
    });
}

sms_status sms_ss_make_queries(sms_ctx *ctx, int32_t spp, uint64_t seed, sms_ss_query *q, float *weight) {
    return guarded(ctx, [&]() -> sms_status {
        if (!ctx || spp <= 0 || !q || !weight) return fail(ctx, SMS_E_INVALID, "bad query arguments");
        if (!ctx->scene.has_cam) return fail(ctx, SMS_E_INVALID, "scene has no camera");
        hipLaunchKernelGGL(k_set_seed, dim3(1), dim3(64), 0, ctx->stream, ctx->seeds + SEED_CAM, seed);
        HIP_TRY(ctx, hipGetLastError());
        return ss_queries_enqueue(ctx, spp, q, weight);
    });
}

sms_status sms_film_accumulate(sms_ctx *ctx, const float *rgb, const float *weight, int64_t pixels, int32_t spp,
                               float *image, float *image_sq, float *image_n) {
    return guarded(ctx, [&]() -> sms_status {
        if (!ctx || pixels < 0 || spp <= 0 || (pixels > 0 && (!rgb || !weight || !image)))
            return fail(ctx, SMS_E_INVALID, "bad film arguments");
        return film_enqueue(ctx, rgb, weight, pixels, spp, image, image_sq, image_n);
    });
}

sms_status sms_glint_make_queries(sms_ctx *ctx, int32_t spp, uint64_t seed, sms_glint_query *q) {
    return guarded(ctx, [&]() -> sms_status {
        if (!ctx || spp <= 0 || !q) return fail(ctx, SMS_E_INVALID, "bad query arguments");
        if (!ctx->scene.has_cam) return fail(ctx, SMS_E_INVALID, "scene has no camera");
        int64_t n = (int64_t) ctx->scene.cam.w * ctx->scene.cam.h * spp;
        int blocks = (int) ((n + 255) / 256);
        hipLaunchKernelGGL(k_set_seed, dim3(1), dim3(64), 0, ctx->stream, ctx->seeds + SEED_CAM, seed);
        hipLaunchKernelGGL(k_glint_queries, dim3(blocks), dim3(256), 0, ctx->stream, ctx->scene, (int) spp,
                           (const uint64_t *) ctx->seeds + SEED_CAM, q, n);
        HIP_TRY(ctx, hipGetLastError());
        return SMS_OK;
    });
}

} // extern "C"

sms_status glint_prepare(sms_ctx *ctx, const sms_glint_query *q, int64_t n, const float *out_rgb, bool from_camera) {
    if (!ctx || n < 0 || (n > 0 && ((!q && !from_camera) || !out_rgb)))
        return fail(ctx, SMS_E_INVALID, "bad estimator arguments");
    if (from_camera && !ctx->scene.has_cam) return fail(ctx, SMS_E_INVALID, "scene has no camera");
    if (ctx->scene.cfg.biased && ctx->scene.cfg.max_trials > 128)
        return fail(ctx, SMS_E_UNSUPPORTED, "biased trial sets above 128 are not built");
    if (!ctx->scene.nmaps) return fail(ctx, SMS_E_INVALID, "scene has no normal map");
    sms_status st = proposal_alloc(ctx, n);
    if (st) return st;
    if (ctx->scene.cfg.biased) return ensure_work(ctx, 0, 0);
    st = ensure_work(ctx, (size_t) std::max<int64_t>(n, 1), 0);
    if (!st) st = ensure_glint_work(ctx, (size_t) std::max<int64_t>(n, 1));
    return st;
}

// The glint pass sequence on the context stream (no host synchronisation; the frame's seed and
// zeroed counters come first); events ev0 / ev1 / ev2 bracket the first-walk pass and the rest.
sms_status glint_enqueue(sms_ctx *ctx, const GlintQuerySource &q, int64_t n, float *out_rgb) {
    const uint64_t *seed = ctx->seeds + SEED_EST;
    EV_RECORD(ctx, ctx->ev0);
    {
        sms_status st = proposal_enqueue_glint(ctx, q, n);
        if (st) return st;
    }
    if (n > 0 && ctx->scene.cfg.biased) {
        EstimatorWork W = work_of(ctx);
        const dim3 grid(grid_for(ctx, ctx->occ_glint_biased));
        if (ctx->scene.n_nodes > 0)
            hipLaunchKernelGGL(k_glint_biased_bvh, grid, dim3(256), 0, ctx->stream, ctx->scene, q, n, seed, out_rgb, W);
        else
            hipLaunchKernelGGL(k_glint_biased_nobvh, grid, dim3(256), 0, ctx->stream, ctx->scene, q, n, seed, out_rgb, W);
        HIP_TRY(ctx, hipGetLastError());
        EV_RECORD(ctx, ctx->ev1);
        if (ctx->area_emitters) {
            hipLaunchKernelGGL(k_glint_bsdf, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, ctx->stream, ctx->scene, q,
                               n, seed, out_rgb);
            HIP_TRY(ctx, hipGetLastError());
        }
    } else if (n > 0) {
        EstimatorWork W = work_of(ctx);
        const unsigned blocks = (unsigned) ((n + 255) / 256);
        if (ctx->scene.n_nodes > 0)
            hipLaunchKernelGGL(k_glint_first_bvh, dim3(blocks), dim3(256), 0, ctx->stream, ctx->scene, q, n, seed, out_rgb, W);
        else
            hipLaunchKernelGGL(k_glint_first_nobvh, dim3(blocks), dim3(256), 0, ctx->stream, ctx->scene, q, n, seed, out_rgb, W);
        HIP_TRY(ctx, hipGetLastError());
        EV_RECORD(ctx, ctx->ev1);
        const dim3 gt(grid_for(ctx, ctx->occ_glint)), gc(grid_for(ctx, ctx->occ_coop)), gm(grid_for(ctx, ctx->occ_mw));
        switch (ctx->trial_group) {
        case 1: hipLaunchKernelGGL(k_glint_trials_g1, gt, dim3(256), 0, ctx->stream, ctx->scene, seed, W); break;
        case 2: hipLaunchKernelGGL(k_glint_trials_g2, gt, dim3(256), 0, ctx->stream, ctx->scene, seed, W); break;
        case 4: hipLaunchKernelGGL(k_glint_trials_g4, gt, dim3(256), 0, ctx->stream, ctx->scene, seed, W); break;
        case 16: hipLaunchKernelGGL(k_glint_trials_g16, gt, dim3(256), 0, ctx->stream, ctx->scene, seed, W); break;
        default: hipLaunchKernelGGL(k_glint_trials_g8, gt, dim3(256), 0, ctx->stream, ctx->scene, seed, W); break;
        }
        HIP_TRY(ctx, hipGetLastError());
        hipLaunchKernelGGL(k_glint_coop_rounds, gc, dim3(256), 0, ctx->stream, ctx->scene, seed, W);
        HIP_TRY(ctx, hipGetLastError());
        hipLaunchKernelGGL(k_glint_multiwave, gm, dim3(256), 0, ctx->stream, ctx->scene, seed, W);
        HIP_TRY(ctx, hipGetLastError());
        hipLaunchKernelGGL(k_glint_multiwave_resolve, dim3(ctx->n_cu), dim3(256), 0, ctx->stream, ctx->scene, W);
        HIP_TRY(ctx, hipGetLastError());
        hipLaunchKernelGGL(k_glint_coop_seq, gc, dim3(256), 0, ctx->stream, ctx->scene, seed, W);
        HIP_TRY(ctx, hipGetLastError());
        hipLaunchKernelGGL(k_glint_finalize, dim3(4 * ctx->n_cu), dim3(256), 0, ctx->stream, ctx->scene, q, seed,
                           out_rgb, W);
        HIP_TRY(ctx, hipGetLastError());
        if (ctx->area_emitters) {
            hipLaunchKernelGGL(k_glint_bsdf, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, ctx->stream, ctx->scene, q,
                               n, seed, out_rgb);
            HIP_TRY(ctx, hipGetLastError());
        }
    } else {
        EV_RECORD(ctx, ctx->ev1);
    }
    EV_RECORD(ctx, ctx->ev2);
    return SMS_OK;
}

sms_status ss_prepare(sms_ctx *ctx, const sms_ss_query *q, int64_t n, const float *out_rgb) {
    if (!ctx || n < 0 || (n > 0 && (!q || !out_rgb))) return fail(ctx, SMS_E_INVALID, "bad estimator arguments");
    const sms_config &cfg = ctx->scene.cfg;
    const int nc = ctx->scene.n_casters_ss;
    if (cfg.biased && cfg.max_trials > 128) return fail(ctx, SMS_E_UNSUPPORTED, "biased trial sets above 128 are not built");
    sms_status st = cfg.biased ? SMS_OK : ensure_over(ctx, query_sub_cap(n, nc));
    if (!st) st = ensure_work(ctx, 0, cfg.biased ? 0 : (size_t) std::max<int64_t>(n * nc, 1));
    if (!st) st = proposal_alloc(ctx, n);
    return st;
}

// single-scatter estimator: proposal mixtures (if a network is set), then the biased pass or
// the unbiased first pass (persistent walker or lane per query), cooperative tail, finalize
sms_status ss_enqueue(sms_ctx *ctx, const sms_ss_query *q, int64_t n, float *out_rgb) {
    const sms_config &cfg = ctx->scene.cfg;
    const int nc = ctx->scene.n_casters_ss;
    const uint64_t *seed = ctx->seeds + SEED_EST;
    EV_RECORD(ctx, ctx->ev0);
    if (n > 0) {
        sms_status st = proposal_enqueue(ctx, q, n, 0);
        if (st) return st;
        EstimatorWork W = work_of(ctx);
        if (cfg.biased) {
            hipLaunchKernelGGL(k_ss_biased, dim3(grid_for(ctx, ctx->occ_biased)), dim3(256), 0, ctx->stream,
                               ctx->scene, q, n, seed, out_rgb, W);
            HIP_TRY(ctx, hipGetLastError());
            EV_RECORD(ctx, ctx->ev1);
        } else {
            if (ctx->ss_first_pass == 0 && !cfg.twostage && ctx->bvh_depth <= WALKER_STACK) {
                // persistent: the resident blocks take queries from the device queue
                // (the two-stage solver, and scenes whose BVH is deeper than the walker's LDS
                // stack, run the lane-per-query pass)
                const int per_cu = ctx->blocks_per_cu > 0 ? ctx->blocks_per_cu : ctx->occ_walker;
                hipLaunchKernelGGL(k_ss_walker, dim3((unsigned) (per_cu * ctx->n_cu)), dim3(256), 0, ctx->stream,
                                   ctx->scene, q, n, seed, ctx->part, ctx->invp, W);
            } else {
                hipLaunchKernelGGL(k_ss_estimate, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, ctx->stream,
                                   ctx->scene, q, n, seed, ctx->part, ctx->invp, W);
            }
            HIP_TRY(ctx, hipGetLastError());
            EV_RECORD(ctx, ctx->ev1);
            hipLaunchKernelGGL(k_ss_coop, dim3(grid_for(ctx, ctx->occ_sscoop)), dim3(256), 0, ctx->stream,
                               ctx->scene, q, seed, ctx->part, ctx->invp, W);
            HIP_TRY(ctx, hipGetLastError());
            hipLaunchKernelGGL(k_ss_finalize, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, ctx->stream,
                               ctx->scene, seed, ctx->part, ctx->invp, nc, n, 1, out_rgb);
            HIP_TRY(ctx, hipGetLastError());
        }
    } else {
        EV_RECORD(ctx, ctx->ev1);
    }
    EV_RECORD(ctx, ctx->ev2);
    return SMS_OK;
}

sms_status ms_prepare(sms_ctx *ctx, const sms_ss_query *q, int64_t n, const float *out_rgb) {
    if (!ctx || n < 0 || (n > 0 && (!q || !out_rgb))) return fail(ctx, SMS_E_INVALID, "bad estimator arguments");
    sms_status st = check_ms(ctx);
    if (st) return st;
    const sms_config &cfg = ctx->scene.cfg;
    const int nc = ctx->scene.n_casters_ms;
    if (cfg.biased && cfg.max_trials > 128) return fail(ctx, SMS_E_UNSUPPORTED, "biased trial sets above 128 are not built");
    st = cfg.biased ? SMS_OK : ensure_over(ctx, query_sub_cap(n, nc));
    if (!st) st = ensure_work(ctx, 0, cfg.biased ? 0 : (size_t) std::max<int64_t>(n * nc, 1));
    if (!st && !cfg.biased) st = proposal_alloc(ctx, n);
    return st;
}

sms_status ms_enqueue(sms_ctx *ctx, const sms_ss_query *q, int64_t n, float *out_rgb) {
    const sms_config &cfg = ctx->scene.cfg;
    const int nc = ctx->scene.n_casters_ms;
    const int b = cfg.bounces <= 2 ? 0 : 1;
    const uint64_t *seed = ctx->seeds + SEED_EST;
    EV_RECORD(ctx, ctx->ev0);
    if (n > 0) {
        if (!cfg.biased) {
            sms_status st = proposal_enqueue(ctx, q, n, 1);
            if (st) return st;
        }
        EstimatorWork W = work_of(ctx);
        if (cfg.biased) {
            dim3 grid(grid_for(ctx, ctx->occ_msbiased[b]));
            if (b == 0) hipLaunchKernelGGL(k_ms_biased_2, grid, dim3(256), 0, ctx->stream, ctx->scene, q, n, seed, out_rgb, W);
            else hipLaunchKernelGGL(k_ms_biased_4, grid, dim3(256), 0, ctx->stream, ctx->scene, q, n, seed, out_rgb, W);
            HIP_TRY(ctx, hipGetLastError());
            EV_RECORD(ctx, ctx->ev1);
        } else {
            dim3 g1((unsigned) ((n + 255) / 256)), g2(grid_for(ctx, ctx->occ_mscoop[b]));
            if (ctx->ss_first_pass == 0) {
                // persistent: resident lanes take (query, caster) items from the device cursor
                const dim3 gp(grid_for(ctx, ctx->occ_mspersist[b]));
                if (b == 0) hipLaunchKernelGGL(k_ms_persist_2, gp, dim3(256), 0, ctx->stream, ctx->scene, q, n, seed, ctx->part, ctx->invp, W);
                else hipLaunchKernelGGL(k_ms_persist_4, gp, dim3(256), 0, ctx->stream, ctx->scene, q, n, seed, ctx->part, ctx->invp, W);
            } else {
                if (b == 0) hipLaunchKernelGGL(k_ms_estimate_2, g1, dim3(256), 0, ctx->stream, ctx->scene, q, n, seed, ctx->part, ctx->invp, W);
                else hipLaunchKernelGGL(k_ms_estimate_4, g1, dim3(256), 0, ctx->stream, ctx->scene, q, n, seed, ctx->part, ctx->invp, W);
            }
            HIP_TRY(ctx, hipGetLastError());
            EV_RECORD(ctx, ctx->ev1);
            const float *part = ctx->part;
            if (b == 0) hipLaunchKernelGGL(k_ms_coop_2, g2, dim3(256), 0, ctx->stream, ctx->scene, q, seed, part, ctx->invp, W);
            else hipLaunchKernelGGL(k_ms_coop_4, g2, dim3(256), 0, ctx->stream, ctx->scene, q, seed, part, ctx->invp, W);
            HIP_TRY(ctx, hipGetLastError());
            hipLaunchKernelGGL(k_ss_finalize, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, ctx->stream,
                               ctx->scene, seed, ctx->part, ctx->invp, nc, n, 0, out_rgb);
            HIP_TRY(ctx, hipGetLastError());
        }
    } else {
        EV_RECORD(ctx, ctx->ev1);
    }
    EV_RECORD(ctx, ctx->ev2);
    return SMS_OK;
}

// camera pass of the single-scatter integrator (seed slot SEED_CAM)
sms_status ss_queries_enqueue(sms_ctx *ctx, int32_t spp, sms_ss_query *q, float *weight) {
    const int64_t n = (int64_t) ctx->scene.cam.w * ctx->scene.cam.h * spp;
    hipLaunchKernelGGL(k_ss_queries, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, ctx->stream, ctx->scene, (int) spp,
                       (const uint64_t *) ctx->seeds + SEED_CAM, q, weight, n);
    HIP_TRY(ctx, hipGetLastError());
    return SMS_OK;
}

sms_status film_enqueue(sms_ctx *ctx, const float *rgb, const float *weight, int64_t pixels, int32_t spp, float *image,
                        float *image_sq, float *image_n) {
    if (pixels == 0) return SMS_OK;
    hipLaunchKernelGGL(k_film_accumulate, dim3((unsigned) ((pixels + 255) / 256)), dim3(256), 0, ctx->stream, rgb, weight,
                       pixels, (int) spp, image, image_sq, image_n);
    HIP_TRY(ctx, hipGetLastError());
    return SMS_OK;
}

namespace {

sms_status estimator_done(sms_ctx *ctx, sms_stats *stats) {
    sms_status st = read_stats(ctx, stats);
    if (st) return st;
    HIP_TRY(ctx, hipEventElapsedTime(&ctx->last_ms, ctx->ev0, ctx->ev1));
    HIP_TRY(ctx, hipEventElapsedTime(&ctx->last_ms2, ctx->ev1, ctx->ev2));
    return SMS_OK;
}

} // namespace

extern "C" {

sms_status sms_glint_estimate_batch(sms_ctx *ctx, const sms_glint_query *q, int64_t n, uint64_t seed, float *out_rgb,
                                    sms_stats *stats) {
    return guarded(ctx, [&]() -> sms_status {
        sms_status st = glint_prepare(ctx, q, n, out_rgb, false);
        if (!st) st = frame_begin(ctx, seed);
        if (!st) st = glint_enqueue(ctx, glint_queries_of(q), n, out_rgb);
        return st ? st : estimator_done(ctx, stats);
    });
}

sms_status sms_glint_estimate_frame(sms_ctx *ctx, int32_t spp, uint64_t seed, float *out_rgb, sms_stats *stats) {
    return guarded(ctx, [&]() -> sms_status {
        if (!ctx || spp <= 0) return fail(ctx, SMS_E_INVALID, "sms_glint_estimate_frame: bad arguments");
        const int64_t n = (int64_t) ctx->scene.cam.w * ctx->scene.cam.h * spp;
        sms_status st = glint_prepare(ctx, nullptr, n, out_rgb, true);
        if (!st) st = frame_begin(ctx, seed);
        if (!st) st = glint_enqueue(ctx, glint_camera_queries(ctx, spp), n, out_rgb);
        return st ? st : estimator_done(ctx, stats);
    });
}

sms_status sms_glint_autotune(sms_ctx *ctx, const sms_glint_query *q, int64_t n, uint64_t seed, float *out_rgb,
                              sms_schedule *best) {
    return guarded(ctx, [&]() -> sms_status {
        if (!best) return fail(ctx, SMS_E_INVALID, "null schedule");
        sms_status st = glint_prepare(ctx, q, n, out_rgb, false);
        if (st) return st;
        const sms_schedule keep = schedule_of(ctx);
        sms_schedule win = keep;
        float win_ms = INFINITY;
        const int groups[4] = {2, 4, 8, 16}, mults[2] = {1, 2}, rounds[2] = {1, 2}, bpcs[2] = {0, 2 * ctx->occ_glint};
        for (int g : groups) for (int m : mults) for (int r : rounds) for (int b : bpcs) {
            ctx->trial_group = g; ctx->lane_trials = g * m; ctx->tail_rounds = r; ctx->blocks_per_cu = b;
            // the first run warms caches and allocations; the best of the next three is kept, as a
            // frame of a few hundred microseconds varies by several percent from run to run
            float ms = INFINITY;
            for (int rep = 0; rep < 4; ++rep) {
                st = frame_begin(ctx, seed + (uint64_t) rep);
                if (!st) st = glint_enqueue(ctx, glint_queries_of(q), n, out_rgb);
                if (st) break;
                HIP_TRY(ctx, hipEventSynchronize(ctx->ev2));
                float t = 0.f;
                HIP_TRY(ctx, hipEventElapsedTime(&t, ctx->ev0, ctx->ev2));
                if (rep > 0) ms = std::min(ms, t);
            }
            if (st) break;
            if (ms < win_ms) {
                win_ms = ms;
                win = schedule_of(ctx);
            }
        }
        apply_schedule(ctx, st ? keep : win);
        if (st) return st;
        *best = win;
        return SMS_OK;
    });
}

sms_status sms_ss_autotune(sms_ctx *ctx, const sms_ss_query *q, int64_t n, uint64_t seed, float *out_rgb,
                           sms_schedule *best) {
    return guarded(ctx, [&]() -> sms_status {
        if (!ctx) return SMS_E_INVALID;
        if (!best) return fail(ctx, SMS_E_INVALID, "null schedule");
        if (ctx->scene.cfg.biased) return fail(ctx, SMS_E_INVALID, "the autotuner times the unbiased estimator");
        const sms_schedule keep = schedule_of(ctx);
        sms_schedule win = keep;
        float win_ms = INFINITY;
        sms_status st = SMS_OK;
        struct Cand { int first_pass, lane_trials; };
        const Cand cands[8] = {{0, 8}, {0, 16}, {0, 32}, {0, 64}, {0, 128}, {1, 4}, {1, 8}, {1, 16}};
        for (const Cand &cd : cands) {
            ctx->ss_first_pass = cd.first_pass; ctx->lane_trials = cd.lane_trials; ctx->blocks_per_cu = 0;
            float ms = 0.f;
            for (int rep = 0; rep < 2; ++rep) {           // the first run warms caches and allocations
                st = sms_ss_estimate_batch(ctx, q, n, seed + (uint64_t) rep, out_rgb, nullptr);
                if (st) break;
                ms = ctx->last_ms + ctx->last_ms2;
            }
            if (st) break;
            if (ms < win_ms) {
                win_ms = ms;
                win = schedule_of(ctx);
            }
        }
        apply_schedule(ctx, st ? keep : win);
        if (st) return st;
        *best = win;
        return SMS_OK;
    });
}

sms_status sms_ms_autotune(sms_ctx *ctx, const sms_ss_query *q, int64_t n, uint64_t seed, float *out_rgb,
                           sms_schedule *best) {
    return guarded(ctx, [&]() -> sms_status {
        if (!ctx) return SMS_E_INVALID;
        if (!best) return fail(ctx, SMS_E_INVALID, "null schedule");
        if (ctx->scene.cfg.biased) return fail(ctx, SMS_E_INVALID, "the autotuner times the unbiased estimator");
        const sms_schedule keep = schedule_of(ctx);
        sms_schedule win = keep;
        float win_ms = INFINITY;
        sms_status st = SMS_OK;
        // the persistent pass refills a lane whose chain runs long, so longer lane runs pay off there
        struct Cand { int first_pass, lane_trials; };
        const Cand cands[11] = {{0, 4}, {0, 16}, {0, 64}, {0, 256}, {0, 1024},
                                {1, 1}, {1, 2}, {1, 4}, {1, 8}, {1, 16}, {1, 32}};
        for (const Cand &cd : cands) {
            ctx->ss_first_pass = cd.first_pass;
            ctx->lane_trials = cd.lane_trials;
            ctx->blocks_per_cu = 0;
            float ms = 0.f;
            for (int rep = 0; rep < 2; ++rep) {           // the first run warms caches and allocations
                st = sms_ms_estimate_batch(ctx, q, n, seed + (uint64_t) rep, out_rgb, nullptr);
                if (st) break;
                ms = ctx->last_ms + ctx->last_ms2;
            }
            if (st) break;
            if (ms < win_ms) {
                win_ms = ms;
                win = schedule_of(ctx);
            }
        }
        apply_schedule(ctx, st ? keep : win);
        if (st) return st;
        *best = win;
        return SMS_OK;
    });
}

sms_status sms_ss_estimate_batch(sms_ctx *ctx, const sms_ss_query *q, int64_t n, uint64_t seed, float *out_rgb,
                                 sms_stats *stats) {
    return guarded(ctx, [&]() -> sms_status {
        sms_status st = ss_prepare(ctx, q, n, out_rgb);
        if (!st) st = frame_begin(ctx, seed);
        if (!st) st = ss_enqueue(ctx, q, n, out_rgb);
        return st ? st : estimator_done(ctx, stats);
    });
}

sms_status sms_ms_walk_batch(sms_ctx *ctx, const sms_ms_walk_in *in, int64_t n, sms_ms_walk_out *out) {
    return guarded(ctx, [&]() -> sms_status {
        if (!ctx || n < 0 || (n > 0 && (!in || !out))) return fail(ctx, SMS_E_INVALID, "bad walk batch arguments");
        sms_status st = check_ms(ctx);
        if (st) return st;
        if (n == 0) return SMS_OK;
        int blocks = (int) ((n + 255) / 256);
        if (ctx->scene.cfg.bounces <= 2)
            hipLaunchKernelGGL(k_ms_walk_2, dim3(blocks), dim3(256), 0, ctx->stream, ctx->scene, in, n, out);
        else
            hipLaunchKernelGGL(k_ms_walk_4, dim3(blocks), dim3(256), 0, ctx->stream, ctx->scene, in, n, out);
        HIP_TRY(ctx, hipGetLastError());
        return SMS_OK;
    });
}

sms_status sms_ms_estimate_batch(sms_ctx *ctx, const sms_ss_query *q, int64_t n, uint64_t seed, float *out_rgb,
                                 sms_stats *stats) {
    return guarded(ctx, [&]() -> sms_status {
        sms_status st = ms_prepare(ctx, q, n, out_rgb);
        if (!st) st = frame_begin(ctx, seed);
        if (!st) st = ms_enqueue(ctx, q, n, out_rgb);
        return st ? st : estimator_done(ctx, stats);
    });
}

// ---------------------------------------------------------------- device memory helpers
sms_status sms_device_alloc(sms_ctx *ctx, size_t bytes, void **ptr) {
    return guarded(ctx, [&]() -> sms_status {
        if (!ctx || !ptr) return SMS_E_INVALID;
        HIP_TRY(ctx, hipMalloc(ptr, bytes == 0 ? 1 : bytes));
        return SMS_OK;
    });
}
sms_status sms_device_free(sms_ctx *ctx, void *ptr) {
    return guarded(ctx, [&]() -> sms_status {
        if (!ctx) return SMS_E_INVALID;
        if (ptr) HIP_TRY(ctx, hipFree(ptr));
        return SMS_OK;
    });
}
sms_status sms_copy_h2d(sms_ctx *ctx, void *dst, const void *src, size_t bytes) {
    return guarded(ctx, [&]() -> sms_status {
        if (!ctx) return SMS_E_INVALID;
        if (bytes) HIP_TRY(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        return SMS_OK;
    });
}
sms_status sms_copy_d2h(sms_ctx *ctx, void *dst, const void *src, size_t bytes) {
    return guarded(ctx, [&]() -> sms_status {
        if (!ctx) return SMS_E_INVALID;
        if (bytes) HIP_TRY(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        return SMS_OK;
    });
}

// ---------------------------------------------------------------- host-buffer variants
#define HOST_VARIANT_BODY(INT, OUTT, CALL)                                                      \
    if (!ctx || n < 0) return SMS_E_INVALID;                                                     \
    if (n == 0) return SMS_OK;                                                                   \
    void *d_in = nullptr, *d_out = nullptr;                                                      \
    sms_status st = sms_device_alloc(ctx, (size_t) n * sizeof(INT), &d_in);                      \
    if (!st) st = sms_device_alloc(ctx, (size_t) n * (OUTT), &d_out);                            \
    if (!st) st = sms_copy_h2d(ctx, d_in, in, (size_t) n * sizeof(INT));                         \
    if (!st) st = CALL;                                                                          \
    if (!st) st = sms_copy_d2h(ctx, out, d_out, (size_t) n * (OUTT));                            \
    if (d_in) sms_device_free(ctx, d_in);                                                        \
    if (d_out) sms_device_free(ctx, d_out);                                                      \
    return st;

sms_status sms_ss_walk_batch_host(sms_ctx *ctx, const sms_ss_walk_in *in, int64_t n, sms_ss_walk_out *out) {
    return guarded(ctx, [&]() -> sms_status {
        HOST_VARIANT_BODY(sms_ss_walk_in, sizeof(sms_ss_walk_out),
                          sms_ss_walk_batch(ctx, (const sms_ss_walk_in *) d_in, n, (sms_ss_walk_out *) d_out))
    });
}
sms_status sms_glint_walk_batch_host(sms_ctx *ctx, const sms_glint_walk_in *in, int64_t n, sms_glint_walk_out *out) {
    return guarded(ctx, [&]() -> sms_status {
        HOST_VARIANT_BODY(sms_glint_walk_in, sizeof(sms_glint_walk_out),
                          sms_glint_walk_batch(ctx, (const sms_glint_walk_in *) d_in, n, (sms_glint_walk_out *) d_out))
    });
}
sms_status sms_ss_estimate_batch_host(sms_ctx *ctx, const sms_ss_query *in, int64_t n, uint64_t seed, float *out,
                                      sms_stats *stats) {
    return guarded(ctx, [&]() -> sms_status {
        HOST_VARIANT_BODY(sms_ss_query, 3 * sizeof(float),
                          sms_ss_estimate_batch(ctx, (const sms_ss_query *) d_in, n, seed, (float *) d_out, stats))
    });
}
sms_status sms_glint_estimate_batch_host(sms_ctx *ctx, const sms_glint_query *in, int64_t n, uint64_t seed, float *out,
                                         sms_stats *stats) {
    return guarded(ctx, [&]() -> sms_status {
        HOST_VARIANT_BODY(sms_glint_query, 3 * sizeof(float),
                          sms_glint_estimate_batch(ctx, (const sms_glint_query *) d_in, n, seed, (float *) d_out, stats))
    });
}

sms_status sms_ms_walk_batch_host(sms_ctx *ctx, const sms_ms_walk_in *in, int64_t n, sms_ms_walk_out *out) {
    return guarded(ctx, [&]() -> sms_status {
        HOST_VARIANT_BODY(sms_ms_walk_in, sizeof(sms_ms_walk_out),
                          sms_ms_walk_batch(ctx, (const sms_ms_walk_in *) d_in, n, (sms_ms_walk_out *) d_out))
    });
}
sms_status sms_ms_estimate_batch_host(sms_ctx *ctx, const sms_ss_query *in, int64_t n, uint64_t seed, float *out,
                                      sms_stats *stats) {
    return guarded(ctx, [&]() -> sms_status {
        HOST_VARIANT_BODY(sms_ss_query, 3 * sizeof(float),
                          sms_ms_estimate_batch(ctx, (const sms_ss_query *) d_in, n, seed, (float *) d_out, stats))
    });
}

} // extern "C"
