// dev_proposal.h -- learned seed proposal on the walk side: network inputs of a query, the
// mixture prepared from the network's raw outputs, the defensive-mixture seed of a walk, and
// the UV -> surface-point lookup of the proposal's mesh caster.
//
// Everything a seed depends on is computed with operations that round identically on the host
// (exp through the FP64 series of dev_lean.h, Cephes sincos, IEEE sqrt / division), so a walk
// seeded from the same raw network outputs is bit-identical on every device and on the CPU.
#pragma once
#include "dev_lean.h"
#include "dev_ss.h"

namespace smsdev {

constexpr int PROP_COMP = 8;
constexpr int PROP_RAW = 48;

// Uniform grid over a mesh's UV square [0,1)^2: cell (x, y) lists, in face order, the faces
// whose UV bounding box overlaps it (CSR: start[G*G + 1], faces = global face indices).
struct DUVGrid { int g; const int *start; const int *faces; };

struct Proposal {
    const float *gmm;                 // per query: cdf[8], mu_u[8], mu_v[8], sx[8], sy[8], rho[8]
    int shape;                        // mesh caster or glinty shape the proposal seeds (-1: none)
    float defensive;                  // weight of the reference's uniform seed
    // anchored mixtures (single scatter): per query the uv (2 floats) where the segment from the
    // shading point to the light sample first meets the proposal's caster (NaN: none), and the
    // half side of the uv box around it that the mixture's [0,1]^2 coordinates span
    // (anchor null or radius 0: the mixture lives in the caster's uv square itself)
    const float *anchor;
    float radius;
    DUVGrid grid;
    // training capture: records (inputs[9], uv[2], weight)
    float *cap;                       // SPREAD sub-lists of cap_sub records of 12 floats
    unsigned long long *cap_counts;   // SPREAD counters, SPREAD_STRIDE apart
    unsigned long long cap_sub;
    int cap_mode;                     // SMS_CAPTURE_SOLUTIONS or SMS_CAPTURE_ESTIMATES
};

SMS_DEV float sigmoid_l(float x) { return 1.f / (1.f + expf_l(-x)); }
SMS_DEV float tanh_l(float x) {
    float t = expf_l(-2.f * fabsf(x));
    float r = (1.f - t) / (1.f + t);
    return x < 0.f ? -r : r;
}

// raw network outputs -> mixture parameters used by the seeds
SMS_DEV void gmm_prepare(const float *o, float *g) {
    float mx = o[0];
    for (int c = 1; c < PROP_COMP; ++c) mx = fmax_(mx, o[6 * c]);
    float e[PROP_COMP], sum = 0.f;
    for (int c = 0; c < PROP_COMP; ++c) { e[c] = expf_l(o[6 * c] - mx); sum = sum + e[c]; }
    float cdf = 0.f;
    for (int c = 0; c < PROP_COMP; ++c) {
        cdf = cdf + e[c] / sum;
        g[c] = cdf;
        g[8 + c] = sigmoid_l(o[6 * c + 1]);
        g[16 + c] = sigmoid_l(o[6 * c + 2]);
        g[24 + c] = expf_l(fmin_(fmax_(o[6 * c + 3], -7.f), 0.f));
        g[32 + c] = expf_l(fmin_(fmax_(o[6 * c + 4], -7.f), 0.f));
        g[40 + c] = 0.95f * tanh_l(o[6 * c + 5]);
    }
}

// network inputs of a query: shading point and light sample normalised to the scene box,
// shading normal as is
SMS_DEV void proposal_inputs(const DScene &S, V3 p, V3 n, V3 light, float *x) {
    x[0] = (p.x - S.box_c[0]) * S.box_inv; x[1] = (p.y - S.box_c[1]) * S.box_inv; x[2] = (p.z - S.box_c[2]) * S.box_inv;
    x[3] = n.x; x[4] = n.y; x[5] = n.z;
    x[6] = (light.x - S.box_c[0]) * S.box_inv; x[7] = (light.y - S.box_c[1]) * S.box_inv;
    x[8] = (light.z - S.box_c[2]) * S.box_inv;
}

// point of the mesh whose uv is `uv` (first face in grid order that contains it)
SMS_DEV bool uv_to_point(const DScene &S, const DUVGrid &G, V2 uv, V3 &p) {
    int gx = (int) (uv.x * (float) G.g), gy = (int) (uv.y * (float) G.g);
    gx = gx < G.g - 1 ? gx : G.g - 1;
    gy = gy < G.g - 1 ? gy : G.g - 1;
    const int cell = gy * G.g + gx;
    for (int i = G.start[cell]; i < G.start[cell + 1]; ++i) {
        const uint4 f = S.faces[G.faces[i]];
        const float2 t0 = S.vuv[f.x], t1 = S.vuv[f.y], t2 = S.vuv[f.z];
        V2 e1 = v2(t1.x - t0.x, t1.y - t0.y), e2 = v2(t2.x - t0.x, t2.y - t0.y), d = v2(uv.x - t0.x, uv.y - t0.y);
        float det = e1.x * e2.y - e1.y * e2.x;
        float b1 = (d.x * e2.y - d.y * e2.x) / det;
        float b2 = (e1.x * d.y - e1.y * d.x) / det;
        float b0 = (1.f - b1) - b2;
        if (b0 >= 0.f && b1 >= 0.f && b2 >= 0.f) {
            const float4 q0 = S.vpos[f.x], q1 = S.vpos[f.y], q2 = S.vpos[f.z];
            p = (v3(q0.x, q0.y, q0.z) * b0 + v3(q1.x, q1.y, q1.z) * b1) + v3(q2.x, q2.y, q2.z) * b2;
            return true;
        }
    }
    return false;
}

// Seed of one single-scatter walk from the defensive mixture.  Draws from the walk's stream:
// t0 chooses the component (t0 < defensive: the reference's uniform seed, whose 2D sample is
// the next two draws, returned in `sample`); otherwise t1 picks the Gaussian and (t2, t3) give
// it a Box-Muller sample, whose surface point is returned in `p_start`.  False: the mixture's
// uv is outside [0,1)^2 or in no face (a failed walk with 0 iterations).  The caller runs one
// sample_path_ss for both kinds, so the lanes of a wavefront share one Newton solver.
// The draw of a defensive-mixture seed: 0 = the uniform component (its 2D sample in `out`),
// 1 = a Gaussian of the mixture (its point in `out`, not range-checked).
SMS_DEV int proposal_draw(const float *g, float defensive, Pcg32 &rw, V2 &out) {
    const float t0 = pcg32_float(rw);
    if (t0 < defensive) {
        out.x = pcg32_float(rw); out.y = pcg32_float(rw);
        return 0;
    }
    const float t1 = pcg32_float(rw), t2 = pcg32_float(rw), t3 = pcg32_float(rw);
    int c = PROP_COMP - 1;
    for (int j = 0; j < PROP_COMP; ++j) if (t1 < g[j]) { c = j; break; }
    const float r = sqrtf(-2.f * logf_l(1.f - t2));
    float sn, cs;
    c_sincosf((2.f * PI_F) * t3, &sn, &cs);
    const float z1 = r * cs, z2 = r * sn;
    const float rho = g[40 + c], q = sqrtf(1.f - rho * rho);
    out = v2(g[8 + c] + g[24 + c] * z1, g[16 + c] + g[32 + c] * (rho * z1 + q * z2));
    return 1;
}

// The anchor of query q (null: the mixture is absolute).
SMS_DEV const float *proposal_anchor(const Proposal &P, int64_t q) {
    return P.anchor && P.radius > 0.f ? P.anchor + 2 * q : nullptr;
}

// mixture coordinates -> uv: absolute, or the anchor's box a + 2 r (xi - 1/2)
SMS_DEV V2 proposal_uv(const float *anc, float radius, V2 xi) {
    if (!anc) return xi;
    const float side = 2.f * radius;
    return v2(anc[0] + side * (xi.x - 0.5f), anc[1] + side * (xi.y - 0.5f));
}

// Seed of a walk on the proposal's mesh: `anc` = the query's anchor (null: absolute mixture).
// A query whose anchor is missing (NaN) takes the reference's uniform seed on every walk.
SMS_DEV bool proposal_seed(const DScene &S, const Proposal &P, const float *g, const float *anc, Pcg32 &rw, V2 &sample,
                           V3 &p_start) {
    const float t0 = pcg32_float(rw);
    if (t0 < P.defensive || (anc && !(anc[0] == anc[0]))) {
        sample.x = pcg32_float(rw); sample.y = pcg32_float(rw);
        return true;
    }
    const float t1 = pcg32_float(rw), t2 = pcg32_float(rw), t3 = pcg32_float(rw);
    int c = PROP_COMP - 1;
    for (int j = 0; j < PROP_COMP; ++j) if (t1 < g[j]) { c = j; break; }
    const float r = sqrtf(-2.f * logf_l(1.f - t2));
    float sn, cs;
    c_sincosf((2.f * PI_F) * t3, &sn, &cs);
    const float z1 = r * cs, z2 = r * sn;
    const float rho = g[40 + c], q = sqrtf(1.f - rho * rho);
    const V2 xi = v2(g[8 + c] + g[24 + c] * z1, g[16 + c] + g[32 + c] * (rho * z1 + q * z2));
    const V2 uv = proposal_uv(anc, P.radius, xi);
    if (!(uv.x >= 0.f && uv.x < 1.f && uv.y >= 0.f && uv.y < 1.f)) return false;
    return uv_to_point(S, P.grid, uv, p_start);
}

// Training target of a solution at `uv`: the uv itself, or its anchor-box coordinates
// (uv - a) / (2 r) + 1/2; false when the query has no anchor.
SMS_DEV bool proposal_target(const float *anc, float radius, V2 uv, V2 &t) {
    if (!anc) { t = uv; return true; }
    if (!(anc[0] == anc[0])) return false;
    const float side = 2.f * radius;
    t = v2((uv.x - anc[0]) / side + 0.5f, (uv.y - anc[1]) / side + 0.5f);
    return true;
}

} // namespace smsdev
