/*
 * sms.hpp -- C++17 interface of the gfx950 SMS library, shaped like the reference's plugin and
 * class interfaces and built only on the C ABI of sms.h (header only; link libsms_amd.so).
 *
 *   sms::Properties                      the subset of mitsuba::Properties the integrators read
 *                                        (include/mitsuba/core/properties.h: bool_, int_, float_)
 *   sms::ss_config / ms_config /         the property mapping of path_sms_ss.cpp:47-61,
 *   glints_config / path_config          path_sms_ms.cpp:149-164, path_sms_glints.cpp:42-54,
 *                                        path_filtered_ss.cpp:33-35 onto sms_config / sms_path_config
 *   sms::Context                         owner of an sms_ctx (scene upload, BVH, normal maps)
 *   sms::SpecularManifoldSingleScatter   manifold_ss.h:19-145: specular_manifold_sampling and
 *   sms::SpecularManifoldMultiScatter    sample_path, one call per batch of shading points
 *   sms::SpecularManifoldGlints          (manifold_ms.h:187-306, manifold_glints.h:28-149)
 *   sms::SingleScatterSMSPathIntegrator  path_sms_ss / path_filtered_ss: render() of spp passes
 *   sms::SingleScatterFilteredPathIntegrator  into the Gaussian film, developed image returned
 *   sms::MultiScatterSMSIntegrator       path_sms_ms / path_sms_glints: the SMS term of the camera
 *   sms::GlintsSMSIntegrator             pass (multi-scatter caustics at the first receiver,
 *                                        glints at the first glinty hit) into a box film
 *
 * Errors of the library become sms::Error on the host side; nothing here crosses the C ABI.
 * The reference evaluates one shading point per call; here every call takes a batch (host
 * vectors, copied over PCIe) or device buffers (sms_device_alloc) of them.
 */
#ifndef SMS_HPP
#define SMS_HPP

#include "sms.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace sms {

class Error : public std::runtime_error {
public:
    Error(sms_status status, const std::string &msg) : std::runtime_error(msg), status(status) {}
    sms_status status;
};

inline void check(sms_status st, const sms_ctx *ctx, const char *what) {
    if (st == SMS_OK) return;
    const char *e = ctx ? sms_ctx_last_error(ctx) : nullptr;
    throw Error(st, std::string(what) + ": status " + std::to_string(st) + (e && *e ? std::string(": ") + e : ""));
}

/* ---------------------------------------------------------------- properties */
class Properties {
public:
    enum class Type { Bool, Int, Float };

    Properties &set_bool(const std::string &name, bool v) { return put(name, Type::Bool, v ? 1.0 : 0.0); }
    Properties &set_int(const std::string &name, int64_t v) { return put(name, Type::Int, (double) v); }
    Properties &set_float(const std::string &name, double v) { return put(name, Type::Float, v); }

    bool has_property(const std::string &name) const { return m_values.count(name) != 0; }

    /* Properties::bool_ / int_ / float_: the stored value or the default; a value of another
       type is an error (an integer is accepted where a float is read, as the reference does) */
    bool bool_(const std::string &name, bool def) const {
        const Entry *e = find(name, Type::Bool, false);
        return e ? e->value != 0.0 : def;
    }
    int64_t int_(const std::string &name, int64_t def) const {
        const Entry *e = find(name, Type::Int, false);
        return e ? (int64_t) e->value : def;
    }
    float float_(const std::string &name, float def) const {
        const Entry *e = find(name, Type::Float, true);
        return e ? (float) e->value : def;
    }
    /* properties no integrator or solver read (the reference warns about them) */
    std::vector<std::string> unqueried() const {
        std::vector<std::string> out;
        for (const auto &kv : m_values)
            if (!m_queried.count(kv.first)) out.push_back(kv.first);
        return out;
    }

private:
    struct Entry { Type type; double value; };
    Properties &put(const std::string &name, Type t, double v) {
        m_values[name] = Entry{t, v};
        return *this;
    }
    const Entry *find(const std::string &name, Type t, bool int_ok) const {
        auto it = m_values.find(name);
        if (it == m_values.end()) return nullptr;
        m_queried.insert(name);
        if (it->second.type != t && !(int_ok && it->second.type == Type::Int))
            throw Error(SMS_E_INVALID, "property \"" + name + "\" has another type");
        return &it->second;
    }
    std::map<std::string, Entry> m_values;
    mutable std::set<std::string> m_queried;
};

/* ---------------------------------------------------------------- property mapping */
inline sms_config default_config() {
    sms_config c;
    sms_config_default(&c);
    return c;
}

/* the SMSConfig keys shared by the SMS integrators (path_sms_ss.cpp:49-57, path_sms_ms.cpp:151-159) */
inline sms_config ss_config(const Properties &p) {
    sms_config c = default_config();
    c.biased = p.bool_("biased", false);
    c.twostage = p.bool_("twostage", false);
    c.halfvector_constraints = p.bool_("halfvector_constraints", false);
    c.mnee_init = p.bool_("mnee_init", false);
    c.step_scale = p.float_("step_scale", 1.f);
    c.max_iterations = (int32_t) p.int_("max_iterations", 20);
    c.solver_threshold = p.float_("solver_threshold", 1e-5f);
    c.uniqueness_threshold = p.float_("uniqueness_threshold", 1e-4f);
    c.max_trials = (int32_t) p.int_("max_trials", -1);
    return c;
}

/* path_sms_ms.cpp:149-164: the shared keys and "bounces" (2) */
inline sms_config ms_config(const Properties &p) {
    sms_config c = ss_config(p);
    c.bounces = (int32_t) p.int_("bounces", 2);
    return c;
}

/* path_sms_glints.cpp:42-54; "glints_vectorized" selects the reference's vectorized glint
   solver, which the batched GPU estimator always is: read and accepted */
inline sms_config glints_config(const Properties &p) {
    sms_config c = default_config();
    c.biased = p.bool_("biased", false);
    c.step_scale = p.float_("step_scale", 1.f);
    c.max_iterations = (int32_t) p.int_("max_iterations", 20);
    c.solver_threshold = p.float_("solver_threshold", 1e-5f);
    c.uniqueness_threshold = p.float_("uniqueness_threshold", 1e-4f);
    c.max_trials = (int32_t) p.int_("max_trials", -1);
    c.bsdf_strategy_only = p.bool_("bsdf_strategy_only", false);
    c.sms_strategy_only = p.bool_("sms_strategy_only", false);
    (void) p.bool_("glints_vectorized", false);
    return c;
}

/* the path integrators' own keys (MonteCarloIntegrator max_depth / rr_depth,
   path_sms_ss.cpp:59-60, path_filtered_ss.cpp:34) and the Gaussian rfilter's stddev */
inline sms_path_config path_config(const Properties &p, int32_t integrator, float filter_stddev = 0.5f) {
    sms_path_config c;
    sms_path_config_default(&c);
    c.integrator = integrator;
    c.max_depth = (int32_t) p.int_("max_depth", -1);
    c.rr_depth = (int32_t) p.int_("rr_depth", 5);
    if (integrator == SMS_PATH_SMS_SS || integrator == SMS_PATH_FILTERED_SS)
        c.disable_reflected_emitters = p.bool_("disable_reflected_emitters", false);
    if (integrator == SMS_PATH_SMS_SS || integrator == SMS_PATH_SMS_MS) c.biased_mnee = p.bool_("biased_mnee", false);
    c.filter_stddev = filter_stddev;
    return c;
}

/* ---------------------------------------------------------------- context and buffers */
class Context {
public:
    Context(const sms_scene_desc &scene, const sms_config &cfg) {
        check(sms_ctx_create(&scene, &cfg, &m_ctx), nullptr, "sms_ctx_create");
        m_cfg = cfg;
        m_width = scene.camera.width;
        m_height = scene.camera.height;
    }
    ~Context() {
        if (m_ctx) (void) sms_ctx_destroy(m_ctx);
    }
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    Context(Context &&o) noexcept : m_ctx(o.m_ctx), m_cfg(o.m_cfg), m_width(o.m_width), m_height(o.m_height) {
        o.m_ctx = nullptr;
    }

    sms_ctx *get() const { return m_ctx; }
    /* film size of the scene's camera */
    int width() const { return m_width; }
    int height() const { return m_height; }
    const sms_config &config() const { return m_cfg; }
    void set_config(const sms_config &cfg) {
        check(sms_ctx_set_config(m_ctx, &cfg), m_ctx, "sms_ctx_set_config");
        m_cfg = cfg;
    }
    void sync() { check(sms_ctx_sync(m_ctx), m_ctx, "sms_ctx_sync"); }

private:
    sms_ctx *m_ctx = nullptr;
    sms_config m_cfg;
    int m_width = 0, m_height = 0;
};

/* device array of T (sms_device_alloc), freed with its owner */
template <typename T>
class DeviceArray {
public:
    DeviceArray(Context &ctx, size_t n) : m_ctx(&ctx), m_n(n) {
        check(sms_device_alloc(ctx.get(), std::max<size_t>(n, 1) * sizeof(T), (void **) &m_p), ctx.get(),
              "sms_device_alloc");
    }
    ~DeviceArray() {
        if (m_p) (void) sms_device_free(m_ctx->get(), m_p);
    }
    DeviceArray(const DeviceArray &) = delete;
    DeviceArray &operator=(const DeviceArray &) = delete;
    T *data() const { return m_p; }
    size_t size() const { return m_n; }
    void upload(const T *src, size_t n) {
        check(sms_copy_h2d(m_ctx->get(), m_p, src, n * sizeof(T)), m_ctx->get(), "sms_copy_h2d");
    }
    void download(T *dst, size_t n) const {
        check(sms_copy_d2h(m_ctx->get(), dst, m_p, n * sizeof(T)), m_ctx->get(), "sms_copy_d2h");
    }
    void zero() {
        std::vector<T> z(m_n);
        std::memset((void *) z.data(), 0, z.size() * sizeof(T));
        upload(z.data(), m_n);
    }

private:
    Context *m_ctx;
    size_t m_n;
    T *m_p = nullptr;
};

/* ---------------------------------------------------------------- solvers */
/* SpecularManifoldSingleScatter (include/mitsuba/render/manifold_ss.h:19-145) */
class SpecularManifoldSingleScatter {
public:
    SpecularManifoldSingleScatter(Context &ctx, const sms_config &cfg) : m_ctx(ctx) { m_ctx.set_config(cfg); }
    /* specular_manifold_sampling(si, sampler) for every receiver query: RGB per query */
    std::vector<float> specular_manifold_sampling(const std::vector<sms_ss_query> &q, uint64_t seed,
                                                  sms_stats *stats = nullptr) const {
        std::vector<float> rgb(3 * q.size());
        check(sms_ss_estimate_batch_host(m_ctx.get(), q.data(), (int64_t) q.size(), seed, rgb.data(), stats),
              m_ctx.get(), "sms_ss_estimate_batch_host");
        return rgb;
    }
    /* the same on device-resident queries and output */
    void specular_manifold_sampling(const sms_ss_query *q, int64_t n, uint64_t seed, float *rgb,
                                    sms_stats *stats = nullptr) const {
        check(sms_ss_estimate_batch(m_ctx.get(), q, n, seed, rgb, stats), m_ctx.get(), "sms_ss_estimate_batch");
    }
    /* sample_path(shape, si, ei, sampler, n_offset, p_start) per walk record */
    std::vector<sms_ss_walk_out> sample_path(const std::vector<sms_ss_walk_in> &w) const {
        std::vector<sms_ss_walk_out> out(w.size());
        check(sms_ss_walk_batch_host(m_ctx.get(), w.data(), (int64_t) w.size(), out.data()), m_ctx.get(),
              "sms_ss_walk_batch_host");
        return out;
    }

private:
    Context &m_ctx;
};

/* SpecularManifoldMultiScatter (include/mitsuba/render/manifold_ms.h:187-306) */
class SpecularManifoldMultiScatter {
public:
    SpecularManifoldMultiScatter(Context &ctx, const sms_config &cfg) : m_ctx(ctx) { m_ctx.set_config(cfg); }
    std::vector<float> specular_manifold_sampling(const std::vector<sms_ss_query> &q, uint64_t seed,
                                                  sms_stats *stats = nullptr) const {
        std::vector<float> rgb(3 * q.size());
        check(sms_ms_estimate_batch_host(m_ctx.get(), q.data(), (int64_t) q.size(), seed, rgb.data(), stats),
              m_ctx.get(), "sms_ms_estimate_batch_host");
        return rgb;
    }
    std::vector<sms_ms_walk_out> sample_path(const std::vector<sms_ms_walk_in> &w) const {
        std::vector<sms_ms_walk_out> out(w.size());
        check(sms_ms_walk_batch_host(m_ctx.get(), w.data(), (int64_t) w.size(), out.data()), m_ctx.get(),
              "sms_ms_walk_batch_host");
        return out;
    }

private:
    Context &m_ctx;
};

/* SpecularManifoldGlints (include/mitsuba/render/manifold_glints.h:28-149) */
class SpecularManifoldGlints {
public:
    SpecularManifoldGlints(Context &ctx, const sms_config &cfg) : m_ctx(ctx) { m_ctx.set_config(cfg); }
    std::vector<float> specular_manifold_sampling(const std::vector<sms_glint_query> &q, uint64_t seed,
                                                  sms_stats *stats = nullptr) const {
        std::vector<float> rgb(3 * q.size());
        check(sms_glint_estimate_batch_host(m_ctx.get(), q.data(), (int64_t) q.size(), seed, rgb.data(), stats),
              m_ctx.get(), "sms_glint_estimate_batch_host");
        return rgb;
    }
    /* sample_glint(sensor_p, ei, si, sampler, n_offset, xi_start) per walk record */
    std::vector<sms_glint_walk_out> sample_glint(const std::vector<sms_glint_walk_in> &w) const {
        std::vector<sms_glint_walk_out> out(w.size());
        check(sms_glint_walk_batch_host(m_ctx.get(), w.data(), (int64_t) w.size(), out.data()), m_ctx.get(),
              "sms_glint_walk_batch_host");
        return out;
    }

private:
    Context &m_ctx;
};

/* ---------------------------------------------------------------- integrators */
struct Image {
    int width = 0, height = 0;
    std::vector<float> rgb;            /* width * height * 3, developed */
    sms_stats stats;                   /* summed over the passes (bernoulli_max: maximum) */
};

inline void add_stats(sms_stats &t, const sms_stats &s) {
    t.walks += s.walks; t.walks_converged += s.walks_converged; t.newton_iterations += s.newton_iterations;
    t.bernoulli_calls += s.bernoulli_calls; t.bernoulli_trials += s.bernoulli_trials;
    t.bernoulli_max = t.bernoulli_max > s.bernoulli_max ? t.bernoulli_max : s.bernoulli_max;
    t.overflow_queries += s.overflow_queries; t.tail_iterations += s.tail_iterations;
}

inline std::vector<float> develop(const std::vector<float> &sum, const std::vector<float> &w) {
    std::vector<float> out(sum.size());
    for (size_t i = 0; i < w.size(); ++i)
        for (int a = 0; a < 3; ++a) out[3 * i + a] = w[i] > 0.f ? sum[3 * i + a] / w[i] : 0.f;
    return out;
}

/* path_sms_ss / path_filtered_ss (src/integrators/path_sms_ss.cpp:39-278,
   path_filtered_ss.cpp:27-219): the plugin's properties; render() runs spp samples per pixel in
   passes of pass_spp (pass p with seed seed + p) into the Gaussian film and develops it */
class PathIntegrator {
public:
    PathIntegrator(const Properties &props, int32_t integrator, float filter_stddev)
        : m_cfg(ss_config(props)), m_pc(path_config(props, integrator, filter_stddev)) {}
    PathIntegrator(const sms_config &cfg, const Properties &props, int32_t integrator, float filter_stddev)
        : m_cfg(cfg), m_pc(path_config(props, integrator, filter_stddev)) {}
    virtual ~PathIntegrator() = default;
    const sms_config &config() const { return m_cfg; }
    const sms_path_config &path() const { return m_pc; }
    /* keep only the single-scatter caustic paths through the first non-specular vertex */
    void set_caustics_only(bool v) { m_pc.caustics_only = v ? 1 : 0; }

    Image render(Context &ctx, int spp, uint64_t seed, int pass_spp = 16) const {
        if (spp <= 0 || pass_spp <= 0) throw Error(SMS_E_INVALID, "render: spp and pass_spp must be positive");
        ctx.set_config(m_cfg);
        const int64_t px = (int64_t) ctx.width() * ctx.height();
        const int per = pass_spp < spp ? pass_spp : spp;
        DeviceArray<float> rgb(ctx, (size_t) (px * per * 3)), pos(ctx, (size_t) (px * per * 2));
        DeviceArray<float> img(ctx, (size_t) (px * 3)), wt(ctx, (size_t) px);
        img.zero();
        wt.zero();
        Image out;
        out.width = ctx.width(); out.height = ctx.height();
        std::memset(&out.stats, 0, sizeof(out.stats));
        int done = 0;
        for (uint64_t p = 0; done < spp; ++p) {
            const int k = spp - done < per ? spp - done : per;
            sms_stats st;
            check(sms_path_samples(ctx.get(), &m_pc, k, seed + p, rgb.data(), pos.data(), &st), ctx.get(),
                  "sms_path_samples");
            check(sms_film_splat(ctx.get(), rgb.data(), pos.data(), k, m_pc.filter_stddev, img.data(), wt.data()),
                  ctx.get(), "sms_film_splat");
            add_stats(out.stats, st);
            done += k;
        }
        std::vector<float> sum((size_t) px * 3), ws((size_t) px);
        img.download(sum.data(), sum.size());
        wt.download(ws.data(), ws.size());
        out.rgb = develop(sum, ws);
        return out;
    }

protected:
    sms_config m_cfg;
    sms_path_config m_pc;
};

/* SingleScatterSMSPathIntegrator (path_sms_ss.cpp:39-61) */
class SingleScatterSMSPathIntegrator : public PathIntegrator {
public:
    explicit SingleScatterSMSPathIntegrator(const Properties &props, float filter_stddev = 0.5f)
        : PathIntegrator(props, SMS_PATH_SMS_SS, filter_stddev) {}
};

/* SingleScatterFilteredPathIntegrator (path_filtered_ss.cpp:27-35) */
class SingleScatterFilteredPathIntegrator : public PathIntegrator {
public:
    explicit SingleScatterFilteredPathIntegrator(const Properties &props, float filter_stddev = 0.5f)
        : PathIntegrator(props, SMS_PATH_FILTERED_SS, filter_stddev) {}
};

/* MultiScatterSMSPathIntegrator (path_sms_ms.cpp:39-164): the shared SMS keys and "bounces" (2),
   "biased_mnee" */
class MultiScatterSMSPathIntegrator : public PathIntegrator {
public:
    explicit MultiScatterSMSPathIntegrator(const Properties &props, float filter_stddev = 0.5f)
        : PathIntegrator(ms_config(props), props, SMS_PATH_SMS_MS, filter_stddev) {}
};

/* MultiScatterFilteredPathIntegrator (path_filtered_ms.cpp:27-36): "bounces" (2), the length of
   the chains it keeps */
class MultiScatterFilteredPathIntegrator : public PathIntegrator {
public:
    explicit MultiScatterFilteredPathIntegrator(const Properties &props, float filter_stddev = 0.5f)
        : PathIntegrator(filtered_ms_config(props), props, SMS_PATH_FILTERED_MS, filter_stddev) {}

private:
    static sms_config filtered_ms_config(const Properties &p) {
        sms_config c = default_config();
        c.bounces = (int32_t) p.int_("bounces", 2);
        return c;
    }
};

/* ---------------------------------------------------------------- scene files
   A Mitsuba 2 scene file read by sms_scene_load: its description (for Context), the
   integrator's type and properties, the sampler's sample count and the film's filter; and the
   path integrator it names, created from those properties as the reference's plugin manager
   would (src/libcore/plugin.cpp:27-59). */
class SceneFile {
public:
    explicit SceneFile(const std::string &path, const std::map<std::string, std::string> &defines = {},
                       bool substitute = false) {
        std::vector<std::string> defs;
        for (const auto &kv : defines) defs.push_back(kv.first + "=" + kv.second);
        std::vector<const char *> ptrs;
        for (const auto &d : defs) ptrs.push_back(d.c_str());
        char err[1024] = {0};
        const sms_status st = sms_scene_load(path.c_str(), ptrs.empty() ? nullptr : ptrs.data(), (int32_t) ptrs.size(),
                                             substitute ? SMS_LOAD_SUBSTITUTE : 0, &m_file, err, (int32_t) sizeof(err));
        if (st) throw Error(st, std::string("sms_scene_load: ") + err);
        sms_scene_info info;
        check(sms_scene_file_info(m_file, &info), nullptr, "sms_scene_file_info");
        m_integrator = info.integrator;
        m_spp = info.spp;
        m_stddev = info.filter_gaussian ? info.filter_stddev : 0.f;
        for (int32_t i = 0; i < info.n_properties; ++i) {
            sms_property p;
            check(sms_scene_file_property(m_file, i, &p), nullptr, "sms_scene_file_property");
            if (p.type == SMS_PROP_BOOL) m_props.set_bool(p.name, p.value != 0.0);
            else if (p.type == SMS_PROP_INT) m_props.set_int(p.name, (int64_t) p.value);
            else if (p.type == SMS_PROP_FLOAT) m_props.set_float(p.name, p.value);
        }
        for (int32_t i = 0; i < info.n_notes; ++i) m_notes.push_back(sms_scene_file_note(m_file, i));
    }
    ~SceneFile() {
        if (m_file) sms_scene_file_free(m_file);
    }
    SceneFile(const SceneFile &) = delete;
    SceneFile &operator=(const SceneFile &) = delete;

    const sms_scene_desc &desc() const { return *sms_scene_file_desc(m_file); }
    const std::string &integrator() const { return m_integrator; }
    const Properties &properties() const { return m_props; }
    int spp() const { return m_spp; }
    float filter_stddev() const { return m_stddev; }
    const std::vector<std::string> &notes() const { return m_notes; }

    /* the path integrator the file names (path_sms_ss, path_filtered_ss, path_sms_ms,
       path_filtered_ms); others are SMS_E_UNSUPPORTED */
    std::unique_ptr<PathIntegrator> make_path_integrator() const {
        if (m_integrator == "path_sms_ss")
            return std::unique_ptr<PathIntegrator>(new SingleScatterSMSPathIntegrator(m_props, m_stddev));
        if (m_integrator == "path_filtered_ss")
            return std::unique_ptr<PathIntegrator>(new SingleScatterFilteredPathIntegrator(m_props, m_stddev));
        if (m_integrator == "path_sms_ms")
            return std::unique_ptr<PathIntegrator>(new MultiScatterSMSPathIntegrator(m_props, m_stddev));
        if (m_integrator == "path_filtered_ms")
            return std::unique_ptr<PathIntegrator>(new MultiScatterFilteredPathIntegrator(m_props, m_stddev));
        throw Error(SMS_E_UNSUPPORTED, "scene file: integrator \"" + m_integrator + "\" is not a built path integrator");
    }

private:
    sms_scene_file *m_file = nullptr;
    std::string m_integrator;
    Properties m_props;
    int m_spp = 4;
    float m_stddev = 0.5f;
    std::vector<std::string> m_notes;
};

/* The SMS term of a camera pass into the box film (sms_film_accumulate): per pixel the mean of
   its spp samples' weight x estimate, developed by the kept-sample count */
class CameraPassIntegrator {
public:
    const sms_config &config() const { return m_cfg; }

protected:
    explicit CameraPassIntegrator(const sms_config &cfg) : m_cfg(cfg) {}
    template <typename Pass>
    Image render_passes(Context &ctx, int spp, uint64_t seed, int pass_spp, Pass &&pass) const {
        if (spp <= 0 || pass_spp <= 0) throw Error(SMS_E_INVALID, "render: spp and pass_spp must be positive");
        ctx.set_config(m_cfg);
        const int64_t px = (int64_t) ctx.width() * ctx.height();
        DeviceArray<float> img(ctx, (size_t) (px * 3)), cnt(ctx, (size_t) px);
        img.zero();
        cnt.zero();
        Image out;
        out.width = ctx.width(); out.height = ctx.height();
        std::memset(&out.stats, 0, sizeof(out.stats));
        int done = 0;
        for (uint64_t p = 0; done < spp; ++p) {
            const int k = spp - done < pass_spp ? spp - done : pass_spp;
            sms_stats st;
            pass(k, seed + p, img.data(), cnt.data(), &st);
            add_stats(out.stats, st);
            done += k;
        }
        std::vector<float> sum((size_t) px * 3), n((size_t) px);
        img.download(sum.data(), sum.size());
        cnt.download(n.data(), n.size());
        out.rgb = develop(sum, n);
        return out;
    }
    sms_config m_cfg;
};

/* MultiScatterSMSPathIntegrator's SMS term (path_sms_ms.cpp:149-164, :228-234): multi-scatter
   caustics at the first caustic receiver of the camera pass */
class MultiScatterSMSIntegrator : public CameraPassIntegrator {
public:
    explicit MultiScatterSMSIntegrator(const Properties &props) : CameraPassIntegrator(ms_config(props)) {}
    Image render(Context &ctx, int spp, uint64_t seed, int pass_spp = 8) const {
        const int64_t px = (int64_t) ctx.width() * ctx.height();
        DeviceArray<sms_ss_query> q(ctx, (size_t) (px * pass_spp));
        DeviceArray<float> w(ctx, (size_t) (px * pass_spp)), rgb(ctx, (size_t) (px * pass_spp * 3));
        return render_passes(ctx, spp, seed, pass_spp, [&](int k, uint64_t s, float *img, float *cnt, sms_stats *st) {
            check(sms_ss_make_queries(ctx.get(), k, s, q.data(), w.data()), ctx.get(), "sms_ss_make_queries");
            check(sms_ms_estimate_batch(ctx.get(), q.data(), px * k, s, rgb.data(), st), ctx.get(),
                  "sms_ms_estimate_batch");
            check(sms_film_accumulate(ctx.get(), rgb.data(), w.data(), px, k, img, nullptr, cnt), ctx.get(),
                  "sms_film_accumulate");
        });
    }
};

/* GlintsSMSPathIntegrator's glint term (path_sms_glints.cpp:42-54, :119-127): the glint
   estimator at the camera ray's first glinty hit */
class GlintsSMSIntegrator : public CameraPassIntegrator {
public:
    explicit GlintsSMSIntegrator(const Properties &props) : CameraPassIntegrator(glints_config(props)) {}
    Image render(Context &ctx, int spp, uint64_t seed, int pass_spp = 4) const {
        const int64_t px = (int64_t) ctx.width() * ctx.height();
        DeviceArray<sms_glint_query> q(ctx, (size_t) (px * pass_spp));
        DeviceArray<float> rgb(ctx, (size_t) (px * pass_spp * 3)), ones(ctx, (size_t) (px * pass_spp));
        std::vector<float> one((size_t) (px * pass_spp), 1.f);
        ones.upload(one.data(), one.size());
        return render_passes(ctx, spp, seed, pass_spp, [&](int k, uint64_t s, float *img, float *cnt, sms_stats *st) {
            check(sms_glint_make_queries(ctx.get(), k, s, q.data()), ctx.get(), "sms_glint_make_queries");
            check(sms_glint_estimate_batch(ctx.get(), q.data(), px * k, s, rgb.data(), st), ctx.get(),
                  "sms_glint_estimate_batch");
            check(sms_film_accumulate(ctx.get(), rgb.data(), ones.data(), px, k, img, nullptr, cnt), ctx.get(),
                  "sms_film_accumulate");
        });
    }
};

}  // namespace sms

#endif /* SMS_HPP */
