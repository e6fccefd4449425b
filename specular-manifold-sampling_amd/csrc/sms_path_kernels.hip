// sms_path_kernels.hip -- kernels of the path integrators (path_sms_ss / path_filtered_ss /
// path_sms_ms / path_filtered_ms) and of the film reconstruction.
//
// A pass batch runs: k_path_trace (one lane per camera sample: the whole path, its own terms,
// the SMS queries of its caustic receivers and its MNEE checks into per-sample slots), an
// exclusive scan of the slot counts, k_path_compact (the slots into dense (sample, depth)-ordered
// lists), the single- or multi-scatter estimator over the dense queries and k_ss_walk /
// k_ms_walk over the dense MNEE walks, then k_path_gather (per sample, in depth order: the path terms, + throughput x SMS
// estimate, + the MNEE completions whose walk did not re-find the path).
#include "kernels.h"
#include "dev_path.h"

namespace smsdev {

__global__ void __launch_bounds__(256) k_path_trace(DScene S, sms_path_config pc, uint64_t seed, int64_t begin,
                                                    int64_t n, int spp, PathSlots P) {
    const int64_t j = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    path_sample(S, pc, seed, begin + j, j, spp, P);
}

// dense[off[j] + k] = slots[j * D + k] for k < count[j]; off = exclusive scan of count
template <typename T>
__device__ void compact_rows(const T *__restrict__ slots, const int *__restrict__ count, const int *__restrict__ off,
                             int D, int64_t j, T *__restrict__ dense) {
    const int c = count[j];
    for (int k = 0; k < c; ++k) dense[off[j] + k] = slots[j * D + k];
}

__global__ void k_path_compact(PathSlots P, int64_t n, const int *__restrict__ q_off, sms_ss_query *__restrict__ dq,
                               const int *__restrict__ m_off, sms_ss_walk_in *__restrict__ dm,
                               sms_ms_walk_in *__restrict__ dmm) {
    const int64_t j = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    if (dq) compact_rows(P.q, P.nq, q_off, P.depth_slots, j, dq);
    if (dm) compact_rows(P.m, P.nm, m_off, P.depth_slots, j, dm);
    if (dmm) compact_rows(P.mm, P.nm, m_off, P.depth_slots, j, dmm);
}

// per sample: L (already in P.L) + sum_k thr_k * est_k + the MNEE completions; written in place.
// A single-scatter MNEE walk re-finds the path when its solution direction matches within the
// uniqueness threshold (path_sms_ss.cpp:213-221); a multi-scatter one when every vertex of its
// chain lies within 1e-5 of the path's (path_sms_ms.cpp:302-311).
__global__ void k_path_gather(PathSlots P, int64_t n, const int *__restrict__ q_off, const float *__restrict__ est,
                              const int *__restrict__ m_off, const sms_ss_walk_out *__restrict__ mout,
                              const sms_ms_walk_out *__restrict__ mmout, float uniqueness_threshold) {
    const int64_t j = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    V3 L = ld3(P.L + 3 * j);
    const int D = P.depth_slots;
    const int nq = P.nq[j];
    for (int k = 0; k < nq; ++k) {
        const int64_t d = q_off[j] + k;
        L = L + mul3(ld3(P.q_thr + 3 * (j * D + k)), ld3(est + 3 * d));
    }
    if (P.nm && mout) {
        const int nm = P.nm[j];
        for (int k = 0; k < nm; ++k) {
            const int64_t s = j * D + k;
            const sms_ss_walk_out &o = mout[m_off[j] + k];
            bool ok = o.success != 0;
            if (ok) {
                const V3 dm = normalize(ld3(o.p) - ld3(P.m_ref + 3 * s));
                if (fabsf(dot(ld3(P.m_dir + 3 * s), dm) - 1.f) >= uniqueness_threshold) ok = false;
            }
            if (!ok) L = L + ld3(P.m_val + 3 * s);
        }
    } else if (P.nm && mmout) {
        const int nm = P.nm[j];
        for (int k = 0; k < nm; ++k) {
            const int64_t s = j * D + k;
            const sms_ms_walk_out &o = mmout[m_off[j] + k];
            bool ok = o.success != 0;
            const int npos = P.mm_npos[s];
            for (int v = 0; ok && v < o.n_vertices && v < SMS_MS_MAX_BOUNCES; ++v)
                if (!(v < npos) || !(norm(ld3(P.mm_pos + 3 * (SMS_MS_MAX_BOUNCES * s + v)) - ld3(o.p[v])) < 1e-5f))
                    ok = false;
            if (!ok) L = L + ld3(P.m_val + 3 * s);
        }
    }
    st3(P.L + 3 * j, L);
}

// ImageBlock::put (src/librender/imageblock.cpp:80-165) gathered per target pixel: the film has
// a border of 2 pixels; with radius > 1 a sample's taps are the n x n pixels from
// lo = ceil(pos - radius) (pos = film position + 1.5, the border offset) with separable weights
// from the discretized filter (rfilter.h:62-66), else (box) the one pixel ceil(pos - .5).
// Pixel (x, y) sums, over the samples of its 5 x 5 neighbourhood in row order, the valid
// samples' weight * rgb and weights.
__global__ void k_film_splat(FilmFilter F, int w, int h, int spp, const float *__restrict__ rgb,
                             const float *__restrict__ pos, float *__restrict__ image, float *__restrict__ weight) {
    const int64_t p = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (int64_t) w * h) return;
    const int x = (int) (p % w), y = (int) (p / w);
    float acc0 = 0.f, acc1 = 0.f, acc2 = 0.f, ws = 0.f;
    for (int ny = y - 2; ny <= y + 2; ++ny) {
        if (ny < 0 || ny >= h) continue;
        for (int nx = x - 2; nx <= x + 2; ++nx) {
            if (nx < 0 || nx >= w) continue;
            const int64_t base = ((int64_t) ny * w + nx) * spp;
            for (int k = 0; k < spp; ++k) {
                const float *v = rgb + 3 * (base + k);
                const float v0 = v[0], v1 = v[1], v2 = v[2];
                if (!(isfinite(v0) && v0 >= -1e-5f && isfinite(v1) && v1 >= -1e-5f && isfinite(v2) && v2 >= -1e-5f))
                    continue;
                const float wt = film_weight(F, w, h, pos[2 * (base + k)], pos[2 * (base + k) + 1], x, y);
                if (wt == 0.f) continue;
                acc0 = acc0 + v0 * wt; acc1 = acc1 + v1 * wt; acc2 = acc2 + v2 * wt;
                ws = ws + wt;
            }
        }
    }
    image[3 * p] += acc0; image[3 * p + 1] += acc1; image[3 * p + 2] += acc2;
    weight[p] += ws;
}

} // namespace smsdev
