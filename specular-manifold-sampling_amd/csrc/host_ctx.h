// host_ctx.h -- host-side state of an SMS context, shared by the library's host sources.
#pragma once
#include "kernels.h"

#include <hip/hip_runtime.h>

#include <exception>
#include <new>
#include <string>
#include <vector>

using namespace smsdev;

struct sms_mlp;

struct sms_ctx {
    int device = 0;
    int n_cu = 0;
    char dev_name[256] = {0};
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;
    float last_ms = 0.f, last_ms2 = 0.f;
    std::string err;
    DScene scene;
    std::vector<void *> scene_allocs;
    // estimator work buffers
    unsigned long long *counters = nullptr;     // [0] cursor [1] n_overflow [2] coop cursor [8..15] stats
    OverflowRec *overflow = nullptr; size_t overflow_cap = 0;
    float *part = nullptr; float *invp = nullptr; size_t part_cap = 0;
    TrialRec *trial = nullptr; GlintRes *res = nullptr; size_t trial_cap = 0, sub_cap = 0;
    OverflowRec *gover = nullptr; size_t gover_cap = 0, over_sub = 0;   // spread overflow list
    int lane_trials = 8;
    int blocks_per_cu = 0;
    int trial_group = 8;
    int tail_rounds = 2;
    int long_chains = 1024, long_trials = 32768;
    int ss_first_pass = 0;                      // 0: persistent walker, 1: lane per query
    // glint tail lists and the multi-wavefront stage (allocated with the glint work buffers)
    OverflowRec *long_items = nullptr; OverflowRec *seq_items = nullptr;
    long long *mw_found = nullptr; unsigned long long *mw_next = nullptr; uint16_t *mw_log = nullptr;
    int mw_items = 0, mw_trials = 0;
    int occ_walker = 0;
    int occ_glint = 0, occ_ss = 0, occ_coop = 0, occ_biased = 0, occ_sscoop = 0, occ_mw = 0;   // occ_glint: trial pass
    int occ_glint_biased = 0;
    bool area_emitters = false;                 // glint BSDF-sampling strategy can contribute
    int bvh_depth = 0;                          // depth of the triangle BVH (0: no meshes)
    int occ_ms[2] = {0, 0}, occ_mscoop[2] = {0, 0}, occ_msbiased[2] = {0, 0};   // [0]: chains <= 2, [1]: <= 4
    int occ_mspersist[2] = {0, 0};
    // host copies the learned proposal's UV grid is built from
    std::vector<float2> h_vuv;
    std::vector<uint4> h_faces;
    std::vector<DShape> h_shapes;
    // learned seed proposal (single scatter)
    sms_mlp *prop_mlp = nullptr;
    int prop_shape = -1;
    float prop_defensive = 1.f;
    int grid_g = 0, grid_shape = -1;
    int *grid_start = nullptr, *grid_faces = nullptr;
    float *prop_x = nullptr, *prop_raw = nullptr, *prop_g = nullptr; size_t prop_cap = 0;
    // anchored single-scatter mixtures: half side of the anchor's uv box (0: absolute mixtures)
    // and the per-query anchors (prop_cap x 2 floats)
    float prop_radius = 0.f;
    float *prop_anchor = nullptr;
    // capture of (inputs, solution uv) of converged first walks
    float *cap = nullptr; unsigned long long *cap_counts = nullptr; size_t cap_sub = 0; int cap_mode = 0;
    // seeds the kernels read from device memory (slots SEED_*), so that a captured frame graph
    // takes a new seed per launch by updating its first node only
    uint64_t *seeds = nullptr;
    // bumped whenever buffers or settings that a captured graph holds by value change
    uint64_t gen = 1;
    bool capturing = false;                     // launch sequence being captured: no event records
    // device address of the allocation-free estimator's scratch array on this context's device
    void *scratch = nullptr;
    // path integrators' per-batch buffers (sms_path_host.hip)
    struct PathBuffers *path = nullptr;
};

// seed slots: estimator, camera pass, the proposal's input / output queries
constexpr int SEED_EST = 0, SEED_CAM = 1, SEED_API = 2, SEED_SLOTS = 4;

// counters buffer: queue counters 0..31, then the spread stats copies, then the spread
// counters of the glint trial and overflow lists
constexpr int STATS_OFF = 32;
constexpr int TRIALS_OFF = STATS_OFF + SPREAD * SPREAD_STRIDE;
constexpr int OVER_OFF = TRIALS_OFF + SPREAD * SPREAD_STRIDE;
constexpr int N_COUNTERS = OVER_OFF + SPREAD * SPREAD_STRIDE;
// folded frame statistics (k_stats_fold): sms_stats fields 0..7 in order, then overflow queries
constexpr int N_FOLDED = 9;


namespace smshost {

inline sms_status fail(sms_ctx *ctx, sms_status code, const std::string &msg) {
    if (ctx) ctx->err = msg;
    return code;
}

inline sms_status fail_noexcept(sms_ctx *ctx, sms_status code, const char *msg) noexcept {
    if (ctx) {
        try { ctx->err = msg; } catch (...) {}
    }
    return code;
}

// Runs the body of a C entry point with every C++ exception turned into a status code, so that
// none crosses the C ABI: std::bad_alloc -> SMS_E_OOM (host allocation), anything else ->
// SMS_E_INVALID with the exception's message as the context's last error.
template <typename F>
inline sms_status guarded(sms_ctx *ctx, F &&body) noexcept {
    try {
        return body();
    } catch (const std::bad_alloc &) {
        return fail_noexcept(ctx, SMS_E_OOM, "host allocation failed");
    } catch (const std::exception &e) {
        return fail_noexcept(ctx, SMS_E_INVALID, e.what());
    } catch (...) {
        return fail_noexcept(ctx, SMS_E_INVALID, "unexpected host exception");
    }
}

#define HIP_TRY(ctx, call)                                                                      \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return fail(ctx, e_ == hipErrorOutOfMemory ? SMS_E_OOM : SMS_E_HIP,                 \
                        std::string(#call) + ": " + hipGetErrorString(e_));                     \
    } while (0)

} // namespace smshost

// record an event unless the launch sequence is being captured into a graph
#define EV_RECORD(ctx, ev)                                                                      \
    do {                                                                                        \
        if (!(ctx)->capturing) HIP_TRY(ctx, hipEventRecord(ev, (ctx)->stream));                 \
    } while (0)

// allocation-free estimator (sms_noalloc_host.hip): per-device scratch address lookup
sms_status noalloc_init_device(sms_ctx *ctx);

// learned seed proposal (sms_mlp_host.hip)
Proposal proposal_of(sms_ctx *ctx);
sms_status proposal_alloc(sms_ctx *ctx, int64_t n);
sms_status proposal_enqueue(sms_ctx *ctx, const sms_ss_query *q, int64_t n, int ms);
sms_status proposal_enqueue_glint(sms_ctx *ctx, const GlintQuerySource &q, int64_t n);
void proposal_release(sms_ctx *ctx);

// launch sequences shared by the estimator calls and the frame graphs (sms_host.hip).  The
// *_prepare functions validate and allocate (never while capturing); the *_enqueue functions
// only launch kernels, reading the seed of slot SEED_EST (SEED_CAM for the camera passes).
__global__ void k_frame_begin(unsigned long long *counters, int n, uint64_t *seeds, uint64_t seed);
__global__ void k_set_seed(uint64_t *slot, uint64_t seed);
__global__ void k_stats_fold(const unsigned long long *counters, unsigned long long *total);
sms_status frame_begin(sms_ctx *ctx, uint64_t seed);
// glint queries from an array, or generated from the camera where they are read (the camera
// seed is slot SEED_CAM, which frame_begin sets with the estimator's)
inline GlintQuerySource glint_queries_of(const sms_glint_query *q) { return GlintQuerySource{q, nullptr, 0}; }
inline GlintQuerySource glint_camera_queries(sms_ctx *ctx, int32_t spp) {
    return GlintQuerySource{nullptr, ctx->seeds + SEED_CAM, (int) spp};
}
sms_status glint_prepare(sms_ctx *ctx, const sms_glint_query *q, int64_t n, const float *out_rgb, bool from_camera);
sms_status glint_enqueue(sms_ctx *ctx, const GlintQuerySource &q, int64_t n, float *out_rgb);
sms_status ss_prepare(sms_ctx *ctx, const sms_ss_query *q, int64_t n, const float *out_rgb);
sms_status ss_enqueue(sms_ctx *ctx, const sms_ss_query *q, int64_t n, float *out_rgb);
sms_status ms_prepare(sms_ctx *ctx, const sms_ss_query *q, int64_t n, const float *out_rgb);
sms_status ms_enqueue(sms_ctx *ctx, const sms_ss_query *q, int64_t n, float *out_rgb);
sms_status ss_queries_enqueue(sms_ctx *ctx, int32_t spp, sms_ss_query *q, float *weight);
sms_status film_enqueue(sms_ctx *ctx, const float *rgb, const float *weight, int64_t pixels, int32_t spp, float *image,
                        float *image_sq, float *image_n);
// statistics of the last estimator call (waits for the stream)
sms_status estimator_stats(sms_ctx *ctx, sms_stats *stats);
// path integrators' buffers (sms_path_host.hip)
void path_release(sms_ctx *ctx);
