// sms_hpp_check.cpp -- a program written the way a C++ integrator uses include/sms.hpp, driven
// by tests/test_sms_hpp.py:
//   sms_hpp_check config            the integrators' property mapping for the reference's scene
//                                   files, one "key value" line per mapped field
//   sms_hpp_check render <out>      (GPU) the glass dome through SingleScatterSMSPathIntegrator,
//                                   SingleScatterFilteredPathIntegrator and
//                                   SpecularManifoldSingleScatter, raw floats written to <out>
//   sms_hpp_check scene <xml> [k=v]...       the integrator a scene file names, its mapped
//                                   configuration and the file's notes
//   sms_hpp_check scene-render <xml> <out> <spp> [k=v]...   (GPU) that integrator's render of the
//                                   file (seed 7, passes of 8 spp), raw floats written to <out>
#include "sms.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace {

void print_config(const char *name, const sms_config &c) {
    printf("%s.biased %d\n%s.twostage %d\n%s.halfvector_constraints %d\n%s.mnee_init %d\n", name, c.biased, name,
           c.twostage, name, c.halfvector_constraints, name, c.mnee_init);
    printf("%s.step_scale %.9g\n%s.max_iterations %d\n%s.solver_threshold %.9g\n%s.uniqueness_threshold %.9g\n", name,
           c.step_scale, name, c.max_iterations, name, c.solver_threshold, name, c.uniqueness_threshold);
    printf("%s.max_trials %d\n%s.bounces %d\n%s.bsdf_strategy_only %d\n%s.sms_strategy_only %d\n", name, c.max_trials,
           name, c.bounces, name, c.bsdf_strategy_only, name, c.sms_strategy_only);
}

void print_path(const char *name, const sms_path_config &c) {
    printf("%s.integrator %d\n%s.max_depth %d\n%s.rr_depth %d\n%s.disable_reflected_emitters %d\n", name,
           c.integrator, name, c.max_depth, name, c.rr_depth, name, c.disable_reflected_emitters);
    printf("%s.biased_mnee %d\n%s.caustics_only %d\n%s.filter_stddev %.9g\n", name, c.biased_mnee, name,
           c.caustics_only, name, c.filter_stddev);
}

int config() {
    // results/Figure_14_15_MainComparison/sphere_sms.xml:26-41 with its defaults
    sms::Properties sphere;
    sphere.set_int("max_depth", 6).set_bool("biased", false).set_bool("twostage", false)
        .set_bool("halfvector_constraints", false).set_bool("mnee_init", false).set_int("max_iterations", 20)
        .set_float("solver_threshold", 0.00001).set_float("uniqueness_threshold", 0.00001)
        .set_int("max_trials", 10000000).set_bool("biased_mnee", false);
    sms::SingleScatterSMSPathIntegrator ss(sphere);
    print_config("sphere", ss.config());
    print_path("sphere", ss.path());
    printf("sphere.unqueried %zu\n", sphere.unqueried().size());
    // an MNEE comparison as path_sms_ss.cpp:30-36 describes it
    sms::Properties mnee;
    mnee.set_bool("mnee_init", true).set_bool("biased", true).set_int("max_trials", 1)
        .set_bool("halfvector_constraints", true).set_bool("biased_mnee", true).set_int("max_depth", 7);
    sms::SingleScatterSMSPathIntegrator mi(mnee);
    print_config("mnee", mi.config());
    print_path("mnee", mi.path());
    // path_filtered_ss with its one key
    sms::Properties pt;
    pt.set_int("max_depth", 7).set_bool("disable_reflected_emitters", true);
    sms::SingleScatterFilteredPathIntegrator fi(pt);
    print_path("filtered", fi.path());
    // path_sms_ms (Figure_18_DoubleRefraction) and path_sms_glints (Figure_19_GlintsComparison)
    sms::Properties ms;
    ms.set_int("bounces", 2).set_bool("biased", true).set_int("max_trials", 32).set_float("step_scale", 0.5);
    print_config("ms", sms::ms_config(ms));
    sms::Properties gl;
    gl.set_float("solver_threshold", 0.00004).set_float("uniqueness_threshold", 0.0002).set_int("max_trials", 100000)
        .set_bool("glints_vectorized", true).set_bool("sms_strategy_only", true);
    print_config("glints", sms::glints_config(gl));
    printf("glints.unqueried %zu\n", gl.unqueried().size());
    // a property of the wrong type is an error; an integer where a float is read is accepted
    sms::Properties bad;
    bad.set_float("max_trials", 3.5).set_int("step_scale", 2);
    try {
        sms::ss_config(bad);
        printf("type_error 0\n");
    } catch (const sms::Error &e) {
        printf("type_error %d\n", e.status);
    }
    sms::Properties ok;
    ok.set_int("step_scale", 2);
    printf("int_as_float %.9g\n", sms::ss_config(ok).step_scale);
    return 0;
}

// the glass dome (a sphere of radius 1, IOR 1.5, centred on a diffuse plane) under a sphere
// light of radius 0.5 and radiance 4 at (0, 5, 0), camera at (0, 3, 6) looking at (0, 0.5, 0)
struct Dome {
    sms_shape_desc shapes[3];
    sms_emitter_desc emitter;
    sms_scene_desc desc;
    Dome(int w, int h) {
        memset(shapes, 0, sizeof(shapes));
        memset(&emitter, 0, sizeof(emitter));
        memset(&desc, 0, sizeof(desc));
        for (sms_shape_desc &s : shapes) {
            s.emitter = -1; s.radius = 1.f; s.eta = 1.5f; s.normalmap = -1; s.tiles = 1.f;
            s.reflectance[0] = s.reflectance[1] = s.reflectance[2] = 0.5f;
        }
        shapes[0].type = SMS_SHAPE_SPHERE; shapes[0].bsdf = SMS_BSDF_DIELECTRIC;
        shapes[0].flags = SMS_FLAG_CAUSTIC_CASTER_SS;
        const float plane[12] = {10.f, 0.f, 0.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, -10.f, 0.f, 0.f};
        shapes[1].type = SMS_SHAPE_RECTANGLE; shapes[1].bsdf = SMS_BSDF_DIFFUSE;
        shapes[1].flags = SMS_FLAG_CAUSTIC_RECEIVER;
        memcpy(shapes[1].to_world, plane, sizeof(plane));
        shapes[2].type = SMS_SHAPE_SPHERE; shapes[2].bsdf = SMS_BSDF_DIFFUSE; shapes[2].emitter = 0;
        shapes[2].center[1] = 5.f; shapes[2].radius = 0.5f;
        shapes[2].reflectance[0] = shapes[2].reflectance[1] = shapes[2].reflectance[2] = 0.f;
        emitter.type = SMS_EMITTER_AREA; emitter.shape = 2;
        emitter.radiance[0] = emitter.radiance[1] = emitter.radiance[2] = 4.f;
        emitter.caustic_ss = 1; emitter.caustic_ms = 1;
        desc.n_shapes = 3; desc.shapes = shapes; desc.n_emitters = 1; desc.emitters = &emitter;
        desc.camera = sms_camera_desc{{0.f, 3.f, 6.f}, {0.f, 0.5f, 0.f}, {0.f, 1.f, 0.f}, 45.f, w, h};
    }
};

int render(const char *path) {
    Dome dome(24, 16);
    sms::Properties props;
    props.set_int("max_depth", 6).set_int("max_trials", 100000);
    sms::SingleScatterSMSPathIntegrator integrator(props);
    sms::Context ctx(dome.desc, integrator.config());
    std::vector<float> out;
    sms::Image a = integrator.render(ctx, 40, 11, 16);        // passes of 16, 16 and 8 spp
    out.insert(out.end(), a.rgb.begin(), a.rgb.end());
    sms::SingleScatterSMSPathIntegrator caustic(props);
    caustic.set_caustics_only(true);
    sms::Image b = caustic.render(ctx, 16, 21, 16);
    out.insert(out.end(), b.rgb.begin(), b.rgb.end());
    sms::SingleScatterFilteredPathIntegrator filtered(props);
    sms::Image c = filtered.render(ctx, 16, 31, 8);
    out.insert(out.end(), c.rgb.begin(), c.rgb.end());
    // the solver class on the camera pass's receivers
    const int64_t n = 24 * 16 * 2;
    sms::DeviceArray<sms_ss_query> q(ctx, (size_t) n);
    sms::DeviceArray<float> w(ctx, (size_t) n);
    sms::check(sms_ss_make_queries(ctx.get(), 2, 41, q.data(), w.data()), ctx.get(), "sms_ss_make_queries");
    std::vector<sms_ss_query> hq((size_t) n);
    q.download(hq.data(), hq.size());
    sms::SpecularManifoldSingleScatter solver(ctx, integrator.config());
    sms_stats st;
    std::vector<float> e = solver.specular_manifold_sampling(hq, 41, &st);
    out.insert(out.end(), e.begin(), e.end());
    const float counts[3] = {(float) a.stats.walks, (float) b.stats.walks, (float) st.walks};
    out.insert(out.end(), counts, counts + 3);
    FILE *f = fopen(path, "wb");
    if (!f) return 2;
    fwrite(out.data(), sizeof(float), out.size(), f);
    fclose(f);
    printf("wrote %zu floats\n", out.size());
    return 0;
}

std::map<std::string, std::string> defines(int argc, char **argv, int first) {
    std::map<std::string, std::string> d;
    for (int i = first; i < argc; ++i) {
        const char *eq = strchr(argv[i], '=');
        if (eq) d[std::string(argv[i], eq - argv[i])] = eq + 1;
    }
    return d;
}

int scene(const char *xml, const std::map<std::string, std::string> &defs) {
    sms::SceneFile file(xml, defs);
    std::unique_ptr<sms::PathIntegrator> integrator = file.make_path_integrator();
    printf("integrator %s\nspp %d\nshapes %d\nemitters %d\n", file.integrator().c_str(), file.spp(),
           file.desc().n_shapes, file.desc().n_emitters);
    print_config("file", integrator->config());
    print_path("file", integrator->path());
    printf("unqueried %zu\nnotes %zu\n", file.properties().unqueried().size(), file.notes().size());
    return 0;
}

int scene_render(const char *xml, const char *path, int spp, const std::map<std::string, std::string> &defs) {
    sms::SceneFile file(xml, defs);
    std::unique_ptr<sms::PathIntegrator> integrator = file.make_path_integrator();
    sms::Context ctx(file.desc(), integrator->config());
    sms::Image img = integrator->render(ctx, spp, 7, 8);
    FILE *f = fopen(path, "wb");
    if (!f) return 2;
    fwrite(img.rgb.data(), sizeof(float), img.rgb.size(), f);
    fclose(f);
    printf("wrote %zu floats, %llu walks\n", img.rgb.size(), (unsigned long long) img.stats.walks);
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    try {
        if (argc >= 2 && !strcmp(argv[1], "config")) return config();
        if (argc >= 3 && !strcmp(argv[1], "render")) return render(argv[2]);
        if (argc >= 3 && !strcmp(argv[1], "scene")) return scene(argv[2], defines(argc, argv, 3));
        if (argc >= 5 && !strcmp(argv[1], "scene-render"))
            return scene_render(argv[2], argv[3], atoi(argv[4]), defines(argc, argv, 5));
    } catch (const sms::Error &e) {
        fprintf(stderr, "sms error: %s\n", e.what());
        return 1;
    }
    fprintf(stderr, "usage: sms_hpp_check config | render <out> | scene <xml> [k=v].. | scene-render <xml> <out> <spp> [k=v]..\n");
    return 2;
}
