/*
 * sms.h -- C ABI of the MI355X (gfx950) Specular Manifold Sampling solver library.
 *
 * Plain C types only: pointers, sizes and POD structs.  No C++ exception crosses this
 * boundary; every entry point returns an sms_status.  Per-walk failure is data (a flag in
 * the output record), never an error code.
 *
 * The library owns a context (sms_ctx) holding device copies of the scene (analytic shapes,
 * triangle meshes with their BVH, normal maps), the solver configuration and its streams.
 * One context is used by one host thread at a time.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   SMSConfig                                   include/mitsuba/render/manifold.h:11-28
 *   SpecularManifoldSingleScatter::sample_path  src/librender/manifold_ss.cpp:203-279
 *   SpecularManifoldSingleScatter::specular_manifold_sampling
 *                                               src/librender/manifold_ss.cpp:28-201
 *   SpecularManifoldGlints::sample_glint        src/librender/manifold_glints.cpp:225-265
 *   SpecularManifoldGlints::specular_manifold_sampling
 *                                               src/librender/manifold_glints.cpp:28-223
 *   SpecularManifoldMultiScatter::sample_path   src/librender/manifold_ms.cpp:203-227
 *   SpecularManifoldMultiScatter::specular_manifold_sampling
 *                                               src/librender/manifold_ms.cpp:27-201
 *   Scene::ray_intersect (whole-scene closest hit used by every reprojection)
 *                                               src/librender/scene_native.inl:17-34
 */
#ifndef SMS_H
#define SMS_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status codes */
typedef int32_t sms_status;
#define SMS_OK              0
#define SMS_E_INVALID      -1   /* bad argument / unsupported scene description      */
#define SMS_E_HIP          -2   /* a HIP runtime call failed                         */
#define SMS_E_OOM          -3   /* device allocation failed                          */
#define SMS_E_QUEUE_FULL   -4   /* a device work queue overflowed its capacity       */
#define SMS_E_UNSUPPORTED  -5   /* requested mode not built                          */
#define SMS_E_NO_DEVICE    -6   /* no HIP device visible                             */

/* ---------------------------------------------------------------- configuration */
/* Field-for-field mirror of SMSConfig (include/mitsuba/render/manifold.h:11-28). */
typedef struct sms_config {
    int32_t biased;                  /* biased (trial set of size max_trials) vs unbiased SMS  */
    int32_t twostage;                /* single scatter: LEAN-smoothed pre-solve on normal maps  */
    int32_t halfvector_constraints;  /* half-vector constraint instead of angle difference     */
    int32_t mnee_init;               /* deterministic MNEE seed instead of a random one        */
    float   step_scale;              /* Newton step scale                                      */
    int32_t max_iterations;          /* Newton iteration cap                                   */
    float   solver_threshold;        /* |C| stopping criterion                                 */
    float   uniqueness_threshold;    /* duplicate-solution test                                */
    int32_t max_trials;              /* biased: trial-set size M; unbiased: Bernoulli cap      */
    int32_t bounces;                 /* multi-scatter chain length                             */
    int32_t bsdf_strategy_only;      /* glints: disable MIS, BSDF strategy only                */
    int32_t sms_strategy_only;       /* glints: disable MIS, SMS strategy only                 */
} sms_config;

/* Defaults of SMSConfig() plus the integrator property defaults
   (src/integrators/path_sms_ss.cpp:49-57). */
void sms_config_default(sms_config *cfg);

/* ---------------------------------------------------------------- scene description */
enum sms_shape_type { SMS_SHAPE_SPHERE = 0, SMS_SHAPE_RECTANGLE = 1, SMS_SHAPE_MESH = 2 };

enum sms_bsdf_type {
    SMS_BSDF_DIFFUSE    = 0,   /* src/bsdfs/diffuse.cpp                                   */
    SMS_BSDF_DIELECTRIC = 1,   /* smooth dielectric, ior() = (eta, 0)  dielectric.cpp:332  */
    SMS_BSDF_CONDUCTOR  = 2,   /* smooth conductor, ior() = (eta, k)   conductor.cpp:292   */
    SMS_BSDF_NORMALMAP  = 3    /* normal-mapped smooth conductor  src/bsdfs/normalmap.cpp  */
};

/* Shape flags (include/mitsuba/render/shape.h:330-343). */
#define SMS_FLAG_CAUSTIC_CASTER_SS   0x01
#define SMS_FLAG_CAUSTIC_CASTER_MS   0x02
#define SMS_FLAG_CAUSTIC_RECEIVER    0x04
#define SMS_FLAG_CAUSTIC_BOUNCER     0x08
#define SMS_FLAG_GLINTY              0x10

typedef struct sms_shape_desc {
    int32_t type;              /* sms_shape_type                                          */
    int32_t bsdf;              /* sms_bsdf_type                                           */
    int32_t flags;             /* SMS_FLAG_*                                              */
    int32_t emitter;           /* index into sms_scene_desc.emitters, or -1              */
    /* sphere */
    float center[3];
    float radius;
    /* rectangle: row-major 3x4 affine to_world (local [-1,1]^2 x {0}) */
    float to_world[12];
    /* bsdf parameters */
    float eta;                 /* dielectric: int_ior/ext_ior; conductor: real part       */
    float k;                   /* conductor: imaginary part                                */
    float reflectance[3];      /* diffuse albedo                                          */
    int32_t normalmap;         /* index into sms_scene_desc.normalmaps (NORMALMAP bsdf)   */
    float tiles;               /* normal-map tiling factor (bsdfs/normalmap.cpp:43)       */
    /* triangle mesh (SMS_SHAPE_MESH); host pointers, copied at sms_ctx_create */
    int32_t n_vertices;
    int32_t n_faces;
    const float    *positions;  /* n_vertices x 3                                         */
    const float    *normals;    /* n_vertices x 3, or NULL                                */
    const float    *texcoords;  /* n_vertices x 2, or NULL                                */
    const uint32_t *faces;      /* n_faces x 3                                            */
    /* CONDUCTOR only: isotropic Beckmann roughness alpha of a rough conductor
       (src/bsdfs/roughconductor.cpp); 0 = smooth.  A rough single-scatter caster gets an offset
       normal drawn per estimate from the Beckmann distribution (manifold_ss.cpp:61-70) */
    float roughness;
} sms_shape_desc;

enum sms_emitter_type { SMS_EMITTER_AREA = 0, SMS_EMITTER_POINT = 1, SMS_EMITTER_DIRECTIONAL = 2 };

typedef struct sms_emitter_desc {
    int32_t type;              /* sms_emitter_type                                         */
    int32_t shape;             /* AREA: index of the emitting shape (a sphere)             */
    float   radiance[3];       /* AREA: radiance (src/emitters/area.cpp:71-78)             */
    float   position[3];       /* POINT: position                                          */
    float   intensity[3];      /* POINT: intensity (src/emitters/point.cpp:74-105)         */
    int32_t caustic_ss;        /* emitter is an SMS single-scatter target                  */
    int32_t caustic_ms;        /* emitter is an SMS multi-scatter target                   */
    float   direction[3];      /* DIRECTIONAL: direction the light travels (unit length)   */
    float   irradiance[3];     /* DIRECTIONAL: irradiance (src/emitters/directional.cpp:105-128) */
} sms_emitter_desc;

typedef struct sms_normalmap_desc {
    int32_t res;               /* square power-of-two resolution                           */
    const uint8_t *rgb;        /* res*res*3 bytes, row-major, byte/255 (no linearisation)  */
    const float *rgb32f;       /* or, when rgb is NULL: res*res*3 floats, row-major, as an
                                  OpenEXR file stores them (src/librender/normalmap.cpp:20-22) */
    int32_t encoded_normals;   /* 0: frames interpolate the texels' slopes (NormalmapBSDF's
                                  use_slopes); 1: they interpolate the stored RGB and decode
                                  n = 2 rgb - 1 (use_slopes false, the reference's default;
                                  bsdfs/normalmap.cpp:226-258).  Glinty shapes need 0. */
} sms_normalmap_desc;

typedef struct sms_camera_desc {
    float origin[3];
    float target[3];
    float up[3];
    float fov_deg;             /* horizontal field of view (perspective.cpp)               */
    int32_t width, height;
} sms_camera_desc;

typedef struct sms_scene_desc {
    int32_t n_shapes;
    const sms_shape_desc *shapes;
    int32_t n_emitters;
    const sms_emitter_desc *emitters;
    int32_t n_normalmaps;
    const sms_normalmap_desc *normalmaps;
    sms_camera_desc camera;
} sms_scene_desc;

/* ---------------------------------------------------------------- scene and asset files
   Host-side readers (no device work) of the reference's scene inputs into the arrays above.
     sms_obj_load      Wavefront OBJ (src/shapes/obj.cpp:95-345): positions (and normals, with
                       the inverse transpose) through `to_world` (4x4 row-major, NULL: identity),
                       corners de-duplicated by their (v, vt, vn) triple in order of first use,
                       polygons fanned, texture v flipped under SMS_OBJ_FLIP_TEXCOORDS (the
                       reference's default), missing normals recomputed angle-weighted
                       (src/librender/mesh.cpp:420-470) unless SMS_OBJ_FACE_NORMALS.
     sms_image_load    OpenEXR scanline files (NONE, RLE, ZIPS, ZIP; HALF, FLOAT, UINT channels;
                       R, G, B or Y) and PNG (1-16 bit grey, grey + alpha, RGB, RGBA, palette;
                       not interlaced) as float RGB -- EXR values as stored, PNG code values /
                       (2^depth - 1), not linearised (src/librender/normalmap.cpp:14-22).
     sms_scene_load    A Mitsuba 2 scene file (src/libcore/xml.cpp): <default> parameters and
                       `defines` ("name=value", overriding them), $name substitution, nested and
                       referenced (id / <ref>) plugins, transforms composed as the reference does
                       (each op multiplied on the left); shapes sphere / rectangle / obj with
                       the caustic_* and glinty flags; bsdfs diffuse, dielectric, conductor,
                       roughconductor (beckmann), normalmap (over a dielectric or conductor, with
                       use_slopes; the file read by sms_image_load), twosided; emitters area (in a shape),
                       point, directional with caustic_emitter_single / _multi; a perspective
                       sensor (fov and fov_axis, to_world), hdrfilm size, gaussian or box
                       rfilter, independent sampler; the integrator's type and properties are
                       returned as they are (the SMS and path integrators read them).  Anything
                       else is SMS_E_UNSUPPORTED, or with SMS_LOAD_SUBSTITUTE replaced --
                       textures by their image's mean colour (or 0.5 grey), other bsdfs by a
                       diffuse bsdf of their diffuse colour, named conductor materials by a
                       perfect mirror -- or dropped (other shapes and emitters), each recorded
                       as a note.  Ignored properties are noted too.
   Errors are reported by status and, when err is not NULL, a message in err[err_len]. */
typedef struct sms_mesh_data {
    int64_t n_vertices, n_faces;
    float *positions;          /* n_vertices x 3                                           */
    float *normals;            /* n_vertices x 3, or NULL                                  */
    float *texcoords;          /* n_vertices x 2, or NULL                                  */
    uint32_t *faces;           /* n_faces x 3                                              */
} sms_mesh_data;
#define SMS_OBJ_FLIP_TEXCOORDS 0x1
#define SMS_OBJ_FACE_NORMALS   0x2
sms_status sms_obj_load(const char *path, const float *to_world, int32_t flags, sms_mesh_data **out, char *err,
                        int32_t err_len);
void sms_mesh_data_free(sms_mesh_data *m);

typedef struct sms_image {
    int32_t width, height, channels;   /* channels = 3 (RGB)                               */
    float *pixels;                     /* row-major, top row first                          */
} sms_image;
sms_status sms_image_load(const char *path, sms_image **out, char *err, int32_t err_len);
void sms_image_free(sms_image *img);

typedef struct sms_scene_file sms_scene_file;
#define SMS_LOAD_SUBSTITUTE 0x1
#define SMS_PROP_BOOL   0
#define SMS_PROP_INT    1
#define SMS_PROP_FLOAT  2
#define SMS_PROP_STRING 3
typedef struct sms_property {
    char name[64];
    int32_t type;              /* SMS_PROP_*                                                */
    double value;              /* BOOL (0 / 1), INT, FLOAT                                  */
    char text[128];            /* STRING                                                    */
} sms_property;
typedef struct sms_scene_info {
    char integrator[32];       /* the integrator plugin's type ("path" when none is given)  */
    int32_t n_properties;      /* its properties (sms_scene_file_property)                  */
    int32_t spp;               /* the sampler's sample_count (independent.cpp default 4)    */
    int32_t filter_gaussian;   /* 1: gaussian rfilter of filter_stddev, 0: box              */
    float filter_stddev;
    int32_t n_notes;           /* substitutions and ignored entries (sms_scene_file_note)   */
} sms_scene_info;
sms_status sms_scene_load(const char *path, const char *const *defines, int32_t n_defines, int32_t flags,
                          sms_scene_file **out, char *err, int32_t err_len);
/* The scene as a description for sms_ctx_create, valid until sms_scene_file_free. */
const sms_scene_desc *sms_scene_file_desc(const sms_scene_file *f);
sms_status sms_scene_file_info(const sms_scene_file *f, sms_scene_info *info);
sms_status sms_scene_file_property(const sms_scene_file *f, int32_t i, sms_property *p);
const char *sms_scene_file_note(const sms_scene_file *f, int32_t i);
void sms_scene_file_free(sms_scene_file *f);

/* ---------------------------------------------------------------- emitter interaction */
/* EmitterInteraction (include/mitsuba/render/manifold.h:145-190), flattened. */
enum sms_ei_kind { SMS_EI_AREA = 0, SMS_EI_POINT = 1, SMS_EI_DIRECTIONAL = 2 };

typedef struct sms_emitter_interaction {
    float p[3];
    float n[3];
    float d[3];
    float weight[3];
    float pdf;
    int32_t kind;              /* sms_ei_kind                                              */
    int32_t emitter;           /* emitter index                                            */
} sms_emitter_interaction;

/* ---------------------------------------------------------------- walk records (parity API) */
/* One single-scatter walk = SpecularManifoldSingleScatter::sample_path
   (src/librender/manifold_ss.cpp:203-279): seed on `shape`, seed-visibility ray cast,
   Newton solve with reprojection.  A zero p_start means "no override" (:224-226). */
typedef struct sms_ss_walk_in {
    float si_p[3];             /* shading point x                                          */
    int32_t shape;             /* target specular shape                                    */
    sms_emitter_interaction ei;
    float n_offset[3];         /* (0,0,1) for smooth casters                               */
    float sample[2];           /* 2D sample for Shape::sample_position                     */
    float p_start[3];          /* seed override                                            */
    float lean_sample[2];      /* twostage: 2D sample of the LEAN offset slope             */
} sms_ss_walk_in;

typedef struct sms_ss_walk_out {
    int32_t success;           /* Mask returned by sample_path                             */
    int32_t iterations;        /* Newton iterations consumed (0 if the seed ray failed)    */
    int32_t final_shape;       /* shape of si_final, -1 if invalid                         */
    int32_t init_valid;        /* seed ray hit the target shape                            */
    float p[3];                /* si_final.p                                               */
    float n[3];                /* si_final.sh_frame.n                                      */
    float p_init[3];           /* si_init.p                                                */
} sms_ss_walk_out;

/* One multi-scatter walk = SpecularManifoldMultiScatter::sample_path
   (src/librender/manifold_ms.cpp:203-227): seed point on `shape`, a ray traced through it
   scattering (reflect/refract) at every caster/bouncer for `bounces` vertices, then the
   block-tridiagonal Newton solve with chain reprojection.  first_path = 0 requires the
   chain to hit seed_shapes[i] at vertex i (manifold_ms.cpp:346-351,884-887). */
#define SMS_MS_MAX_BOUNCES 4
typedef struct sms_ms_walk_in {
    float si_p[3];
    int32_t shape;
    sms_emitter_interaction ei;
    float sample[2];
    float p_start[3];
    int32_t first_path;
    int32_t seed_shapes[SMS_MS_MAX_BOUNCES];
} sms_ms_walk_in;

typedef struct sms_ms_walk_out {
    int32_t success;
    int32_t iterations;
    int32_t seed_valid;        /* sample_seed_path succeeded                               */
    int32_t n_vertices;
    float p[SMS_MS_MAX_BOUNCES][3];      /* solution chain (valid when success)            */
    int32_t shapes[SMS_MS_MAX_BOUNCES];  /* shapes of the seed chain                       */
} sms_ms_walk_out;

/* One glint walk = SpecularManifoldGlints::sample_glint (manifold_glints.cpp:225-265).
   xi_start == (1,1) means "no override" (:253-255). */
typedef struct sms_glint_walk_in {
    float sensor_p[3];
    float ei_p[3];
    float si_p[3];
    float frame_s[3], frame_t[3], frame_n[3];   /* si.sh_frame                          */
    float uv[2];
    float duv_dx[2], duv_dy[2];
    int32_t shape;             /* normal-mapped shape                                      */
    float n_offset[3];
    float sample[2];
    float xi_start[2];
} sms_glint_walk_in;

typedef struct sms_glint_walk_out {
    int32_t success;
    int32_t iterations;
    float uv[2];
    float uv_init[2];
} sms_glint_walk_out;

/* ---------------------------------------------------------------- estimator queries */
/* A shading point handed to the single-scatter estimator (the SurfaceInteraction
   passed to specular_manifold_sampling, manifold_ss.cpp:28-30). */
typedef struct sms_ss_query {
    float p[3];
    float n[3];                /* geometric normal                                         */
    float sh_n[3], sh_s[3], sh_t[3];
    float wi[3];               /* local incident direction (si.wi)                         */
    int32_t shape;             /* receiver shape                                           */
    int32_t pad;
} sms_ss_query;

/* A glinty first hit handed to the glint estimator (manifold_glints.cpp:28-31). */
typedef struct sms_glint_query {
    float sensor_p[3];
    float p[3];
    float n[3];
    float sh_n[3], sh_s[3], sh_t[3];
    float dp_du[3], dp_dv[3];
    float wi[3];
    float uv[2];
    float duv_dx[2], duv_dy[2];
    int32_t shape;
    int32_t pixel;             /* film pixel index (for We * film area)                    */
} sms_glint_query;

/* Counters returned per call (the reference keeps static atomics,
   manifold_ss.cpp:14-18; here they are per call). */
typedef struct sms_stats {
    uint64_t walks;             /* sample_path / sample_glint calls                        */
    uint64_t walks_converged;   /* calls that returned success                             */
    uint64_t newton_iterations; /* Newton iterations over all walks                        */
    uint64_t bernoulli_calls;   /* unbiased: estimates that started a trial loop           */
    uint64_t bernoulli_trials;  /* unbiased: trials over all loops (incl. the match)        */
    uint64_t bernoulli_max;     /* unbiased: longest trial loop                            */
    uint64_t overflow_queries;  /* queries finished by the cooperative tail pass           */
    uint64_t tail_iterations;   /* Newton iterations run by the tail pass (subset of the   */
                                /* newton_iterations total)                                */
} sms_stats;

/* ---------------------------------------------------------------- context */
typedef struct sms_ctx sms_ctx;

/* Copy the scene to the current HIP device, build the BVH, decode normal maps.  The scene is
   validated and built on the host before any device call, so a bad description returns
   SMS_E_INVALID (and a host allocation failure SMS_E_OOM) also where no device is visible. */
sms_status sms_ctx_create(const sms_scene_desc *scene, const sms_config *cfg, sms_ctx **out);
/* Host only, no device needed: validates a scene description as sms_ctx_create does and
   returns in *device_bytes the device memory its scene arrays would take (normal-map slopes and
   LEAN pyramid, shape records, mesh vertices and faces, BVH nodes, triangle records, area
   CDFs), for capacity planning.  err (if not NULL) receives the failure reason.  No
   counterpart in the reference (Scene construction, src/librender/scene.cpp:18-112, has no
   dry run). */
sms_status sms_scene_footprint(const sms_scene_desc *scene, uint64_t *device_bytes, char *err, int32_t err_len);
sms_status sms_ctx_destroy(sms_ctx *ctx);
sms_status sms_ctx_set_config(sms_ctx *ctx, const sms_config *cfg);
/* Scheduling knobs of the persistent estimator kernels (chosen by the autotuner).
   lane_trials: Bernoulli trials one lane (glints: one lane group) runs for a query before the
   query is handed on to the wavefront-wide tail passes; blocks_per_cu: persistent blocks per
   compute unit (0 = the occupancy limit); trial_group (glints): lanes that run one query's
   trials side by side in the first trial pass (1, 2, 4, 8 or 16); tail_rounds (glints):
   64-trial rounds of the one-wavefront pass before a chain moves to the multi-wavefront stage
   (0 = straight to it); long_chains / long_trials (glints): chains the multi-wavefront stage
   holds (at most 1024) and trials it runs per chain (a multiple of 64, at most 32768); the
   rest continue one wavefront per chain; ss_first_pass (unbiased single and multi scatter): 0 =
   persistent -- single scatter: the walker, whose lanes take their next walk as soon as one
   ends (the two-stage solver uses 1); multi scatter: resident lanes that take their next
   (query, caster) item from a device queue as soon as a chain ends --, 1 = one lane per query
   for its first walk and up to lane_trials trials.  No counterpart in the reference (its trial
   loops are serial). */
typedef struct sms_schedule {
    int32_t lane_trials;
    int32_t blocks_per_cu;
    int32_t trial_group;
    int32_t tail_rounds;
    int32_t long_chains;
    int32_t long_trials;
    int32_t ss_first_pass;
    int32_t pad;
} sms_schedule;
sms_status sms_ctx_set_schedule(sms_ctx *ctx, const sms_schedule *s);
sms_status sms_ctx_get_schedule(const sms_ctx *ctx, sms_schedule *s);

/* Last error message for this context (static storage owned by the context). */
const char *sms_ctx_last_error(const sms_ctx *ctx);
/* Device name and compute-unit count of the device the context lives on. */
sms_status sms_ctx_device_info(const sms_ctx *ctx, char *name, int32_t name_len, int32_t *n_cu);

/* ---------------------------------------------------------------- device-pointer entry points
   All pointers below are DEVICE pointers (already resident in HBM); the calls are
   asynchronous on the context stream unless noted.  sms_ctx_sync waits for completion. */
sms_status sms_ctx_sync(sms_ctx *ctx);

/* Parity entry points: one walk per record. */
sms_status sms_ss_walk_batch(sms_ctx *ctx, const sms_ss_walk_in *in, int64_t n,
                             sms_ss_walk_out *out);
sms_status sms_glint_walk_batch(sms_ctx *ctx, const sms_glint_walk_in *in, int64_t n,
                                sms_glint_walk_out *out);
sms_status sms_ms_walk_batch(sms_ctx *ctx, const sms_ms_walk_in *in, int64_t n,
                             sms_ms_walk_out *out);

/* Throughput entry points: the whole estimator per query, run by a persistent kernel that
   refills idle lanes from a device work queue.  Random numbers are drawn from per-(query,
   walk) PCG32 streams derived from `seed`, so results do not depend on scheduling.
   out_rgb: n x 3 floats.  stats: host pointer, filled after completion (may be NULL). */
sms_status sms_ss_estimate_batch(sms_ctx *ctx, const sms_ss_query *q, int64_t n, uint64_t seed,
                                 float *out_rgb, sms_stats *stats);
/* sms_ss_estimate_batch without any allocation: no counterpart in the reference; for callers
   that cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
   that forbid hipMalloc).  Same results and statistics as sms_ss_estimate_batch, bit for bit.  The
   call's scratch (work counters, seed slots, per-(query, caster) partial results, the overflow
   list of the Bernoulli tail) is a fixed device array of the library, about 20 MB on every device
   the library runs on, held until the process ends: at most 262144 (query, caster) pairs per call
   (n x number of single-scatter casters), larger calls return SMS_E_INVALID.  Every call on a
   device shares that device's array, so calls that overlap on one device (other contexts' or
   other streams') race; calls on one context's stream follow each other.  Runs without a learned
   proposal or training capture (SMS_E_UNSUPPORTED otherwise).  Synchronises like
   sms_ss_estimate_batch (stats are read back). */
sms_status sms_ss_estimate_batch_noalloc(sms_ctx *ctx, const sms_ss_query *q, int64_t n, uint64_t seed,
                                         float *out_rgb, sms_stats *stats);
/* Glint estimator (SpecularManifoldGlints::specular_manifold_sampling, manifold_glints.cpp:
   28-223): the SMS strategy -- unbiased Bernoulli trials, or with config.biased the distinct
   solutions of a trial set of max_trials <= 128 walks -- and, in scenes with an area emitter,
   the BSDF-sampling strategy at a footprint offset (:183-220), MIS-weighted as the reference. */
sms_status sms_glint_estimate_batch(sms_ctx *ctx, const sms_glint_query *q, int64_t n,
                                    uint64_t seed, float *out_rgb, sms_stats *stats);
/* The glint estimate of a whole camera frame without a query buffer: the same results as
   sms_glint_make_queries(ctx, spp, seed, q) followed by sms_glint_estimate_batch(ctx, q,
   w*h*spp, seed, out_rgb, stats), each query generated by the kernel that reads it
   (path_sms_glints.cpp:119-127 per pixel sample).  out_rgb: device, w*h*spp x 3. */
sms_status sms_glint_estimate_frame(sms_ctx *ctx, int32_t spp, uint64_t seed, float *out_rgb, sms_stats *stats);
/* Autotuner (no counterpart in the reference): times sms_glint_estimate_batch on the given
   device-resident queries for every candidate schedule (trial_group x lane_trials x
   tail_rounds x blocks_per_cu), keeps the fastest in the context and returns it in *best.
   out_rgb is device scratch of 3 * n floats; results do not depend on the schedule. */
sms_status sms_glint_autotune(sms_ctx *ctx, const sms_glint_query *q, int64_t n, uint64_t seed,
                              float *out_rgb, sms_schedule *best);
/* Autotuner of the unbiased single-scatter estimator (no counterpart in the reference): times
   sms_ss_estimate_batch on the given device-resident queries for every candidate first pass
   and lane_trials, keeps the fastest in the context and returns it in *best.  out_rgb is
   device scratch of 3 * n floats; results do not depend on the schedule. */
/* Autotuner of the unbiased multi-scatter estimator (no counterpart in the reference): times
   sms_ms_estimate_batch for both first passes (ss_first_pass: persistent with lane_trials 4 ..
   1024, lane per query with 1 .. 32; lane_trials = trials a lane runs before a chain moves to
   the one-wavefront pass), keeps the fastest and returns it in *best. */
sms_status sms_ms_autotune(sms_ctx *ctx, const sms_ss_query *q, int64_t n, uint64_t seed, float *out_rgb,
                           sms_schedule *best);
sms_status sms_ss_autotune(sms_ctx *ctx, const sms_ss_query *q, int64_t n, uint64_t seed, float *out_rgb,
                           sms_schedule *best);
/* Multi-scatter estimator (SpecularManifoldMultiScatter::specular_manifold_sampling,
   manifold_ms.cpp:27-201) over receiver queries; config.bounces vertices per chain. */
sms_status sms_ms_estimate_batch(sms_ctx *ctx, const sms_ss_query *q, int64_t n, uint64_t seed,
                                 float *out_rgb, sms_stats *stats);

/* Camera-ray generation for glint frames: for every pixel sample of a width x height x spp
   frame, trace the primary ray differential, intersect the glinty shape and write the
   glint query (shape = -1 when the first hit is not glinty).  q: device, n = w*h*spp. */
sms_status sms_glint_make_queries(sms_ctx *ctx, int32_t spp, uint64_t seed, sms_glint_query *q);

/* Camera pass of the single-scatter integrator (path_sms_ss.cpp:94-251), SMS term only: every
   pixel sample of a width x height x spp frame follows smooth dielectric interfaces
   (reflection with probability F, else refraction with weight eta_ti^2, dielectric.cpp:201-313)
   until it reaches a caustic receiver that is not a caster (path_sms_ss.cpp:118-124), which
   becomes its query; weight receives the path throughput.  Samples that reach no receiver get
   shape -1 (the estimators return 0 for them, without walks) and weight 0.
   q: device, n = w*h*spp queries; weight: device, n floats. */
sms_status sms_ss_make_queries(sms_ctx *ctx, int32_t spp, uint64_t seed, sms_ss_query *q, float *weight);
/* Box-filter film (device memory): for each of `pixels` pixels, adds the sum over its spp
   consecutive samples of weight * rgb to image (pixels x 3) and, if image_sq is not NULL, the
   sum of squares to image_sq.  A sample with a non-finite or negative (below -1e-5) channel is
   dropped, as the reference's ImageBlock::put drops it (src/librender/imageblock.cpp:85-107);
   image_n (pixels floats, may be NULL) receives the number of samples kept, by which the
   caller divides. */
sms_status sms_film_accumulate(sms_ctx *ctx, const float *rgb, const float *weight, int64_t pixels, int32_t spp,
                               float *image, float *image_sq, float *image_n);

/* Kernel timing of the most recent estimator call, measured with HIP events on the context
   stream the kernels ran on: ms[0] = first (persistent lane-per-query) pass, ms[1] = the
   tail / finalize passes.  ms must hold 2 floats. */
sms_status sms_ctx_last_kernel_ms(const sms_ctx *ctx, float *ms);

/* ---------------------------------------------------------------- host-buffer convenience
   Same as above but with host buffers; copies in/out over PCIe and synchronises. */
sms_status sms_ss_walk_batch_host(sms_ctx *ctx, const sms_ss_walk_in *in, int64_t n,
                                  sms_ss_walk_out *out);
sms_status sms_glint_walk_batch_host(sms_ctx *ctx, const sms_glint_walk_in *in, int64_t n,
                                     sms_glint_walk_out *out);
sms_status sms_ms_walk_batch_host(sms_ctx *ctx, const sms_ms_walk_in *in, int64_t n,
                                  sms_ms_walk_out *out);
sms_status sms_ms_estimate_batch_host(sms_ctx *ctx, const sms_ss_query *q, int64_t n,
                                      uint64_t seed, float *out_rgb, sms_stats *stats);
sms_status sms_ss_estimate_batch_host(sms_ctx *ctx, const sms_ss_query *q, int64_t n,
                                      uint64_t seed, float *out_rgb, sms_stats *stats);
sms_status sms_glint_estimate_batch_host(sms_ctx *ctx, const sms_glint_query *q, int64_t n,
                                         uint64_t seed, float *out_rgb, sms_stats *stats);

/* ---------------------------------------------------------------- learned seed proposal
   No counterpart in the reference, whose seeds are uniform over the caster's area
   (manifold_ss.cpp:228-235).  A small network maps (shading point, shading normal, light
   sample), positions normalised to the scene's bounding box, through 8-frequency Fourier
   features and 4 ReLU layers of width 64 (BF16 weights and activations, FP32 accumulation on
   the matrix cores) to an 8-component Gaussian mixture over the UV of one mesh caster
   (48 raw outputs: per component weight logit, mean u, mean v (sigmoid), log sigma u, log sigma
   v (clamped to [-7, 0]), correlation (0.95 tanh)).  With a proposal set, the seed of every
   walk on that caster -- the first walk and every Bernoulli trial -- is drawn from
   defensive * (uniform area seed of the reference) + (1 - defensive) * mixture, so the
   inverse-PDF estimate stays unbiased; a mixture seed whose uv falls outside the mesh's UV
   domain is a failed walk with 0 iterations.  Training minimises the mixture's (weighted)
   negative log-likelihood of captured solutions' uv (Adam). */
#define SMS_MLP_INPUTS 9
#define SMS_MLP_OUTPUTS 48
#define SMS_MLP_PARAMS 25904
typedef struct sms_mlp sms_mlp;
/* He-initialised network on ctx's device (the output layer scaled by 0.1). */
sms_status sms_mlp_create(sms_ctx *ctx, uint64_t seed, sms_mlp **out);
sms_status sms_mlp_destroy(sms_mlp *mlp);
/* FP32 master parameters (SMS_MLP_PARAMS floats, host memory): per layer the weights
   output-major W[n][k] (layer 1: k < 160, the 153 features zero-padded), then the bias. */
sms_status sms_mlp_get_params(sms_mlp *mlp, float *params);
sms_status sms_mlp_set_params(sms_mlp *mlp, const float *params);
/* x: n x 9 inputs, out: n x 48 raw outputs (device memory). */
sms_status sms_mlp_infer(sms_mlp *mlp, const float *x, int64_t n, float *out);
/* Weighted mean negative log-likelihood sum_j w_j NLL_j / sum_j w_j of uv (n x 2, device)
   under the mixtures of x (n x 9, device), w (n, device; NULL: all 1, non-negative, not all 0)
   and its gradient; the gradient is kept for sms_mlp_adam and, if grad is not NULL, copied to
   grad (SMS_MLP_PARAMS host floats).  *loss (if not NULL) receives the loss. */
sms_status sms_mlp_grad(sms_mlp *mlp, const float *x, const float *uv, const float *w, int64_t n, float *grad,
                        float *loss);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) with the kept gradient, which is cleared. */
sms_status sms_mlp_adam(sms_mlp *mlp, float lr);
/* n_steps training steps, each the gradient of sms_mlp_grad over the batch window
   [offsets[i], offsets[i] + batch) of the device sample arrays x (records x 9), uv (records x
   2), w (records; NULL: all 1) followed by an Adam step with learning rate lr; offsets is a host
   array whose windows must lie inside the arrays.  The steps are queued on the context's
   stream without host synchronisation; losses (if not NULL, n_steps host floats) receives each
   step's weighted mean NLL (then the call waits for the steps).  Gradients are summed in a
   fixed order: the same inputs give bit-identical parameters.  The reference has no trained
   component; this is the online training loop of the learned proposal. */
sms_status sms_mlp_train(sms_mlp *mlp, const float *x, const float *uv, const float *w, int32_t batch,
                         const int64_t *offsets, int32_t n_steps, float lr, float *losses);
/* Use `mlp` as the seed proposal of caster mesh `shape` (with uvs): in sms_ss_estimate_batch if
   it is a single-scatter caster, in the unbiased sms_ms_estimate_batch if it is a multi-scatter
   caster (the seed point of the chain on it; the biased multi-scatter estimator keeps uniform
   seeds); or of a glinty normal-mapped `shape`, in sms_glint_estimate_batch (the footprint
   sample, see sms_glint_proposal_inputs); defensive in (0, 1].  mlp NULL: uniform seeds, `shape` still names the caster whose
   solutions the training capture records; shape -1 clears the proposal. */
sms_status sms_ctx_set_proposal(sms_ctx *ctx, sms_mlp *mlp, int32_t shape, float defensive);
/* Anchored single-scatter mixtures (no counterpart in the reference, whose seeds are uniform):
   with radius r > 0 and a single-scatter caster mesh as the proposal's shape, each query's
   mixture lives in the uv box of half side r around its anchor a -- the uv where the ray from
   the shading point toward the estimator's light sample first meets a surface, if that surface
   is the proposal's mesh (the segment manifold next-event estimation starts from,
   manifold_ss.cpp:247-250) --, uv = a + 2 r (xi - 1/2) for a mixture sample xi, and the
   training capture records xi = (uv - a) / (2 r) + 1/2.  A query without an anchor seeds every
   walk uniformly and records nothing.  r = 0 (the default): mixtures over the mesh's uv square.
   Multi-scatter and glint proposals are not affected. */
sms_status sms_ctx_set_proposal_anchor(sms_ctx *ctx, float radius);
/* Anchors (device, n x 2 floats, NaN when none) of the queries of an estimate with this seed on
   the proposal's shape, as the anchored mixtures use them. */
sms_status sms_ss_proposal_anchors(sms_ctx *ctx, const sms_ss_query *q, int64_t n, uint64_t seed, float *anchor);
/* Training capture of later unbiased sms_ss_estimate_batch / sms_ms_estimate_batch calls on the
   proposal shape, as records (network inputs, solution uv -- for a multi-scatter chain the uv
   of its vertex on that shape --, weight):
     SMS_CAPTURE_SOLUTIONS  every converged walk -- first walks and Bernoulli trials, whose seeds
                            follow the same distribution -- with weight 1;
     SMS_CAPTURE_ESTIMATES  every (query, caster) whose estimate has a positive RGB mean: the
                            first walk's solution with that mean as weight (bsdf * contribution
                            * inverse probability), so a weighted fit follows the solutions'
                            contributions instead of their basins.
   64 lists of ceil(capacity / 64) records, filled by the estimator's blocks in turn, so up to
   64 ceil(capacity / 64) in total; a full list drops further records (capacity 0 disables
   and frees). */
#define SMS_CAPTURE_SOLUTIONS 0
#define SMS_CAPTURE_ESTIMATES 1
sms_status sms_ctx_set_training_capture(sms_ctx *ctx, int64_t capacity, int32_t mode);
/* Number of recorded samples; when x / uv are not NULL, copies them to device buffers of
   count x 9 and count x 2 floats, and the weights to w (count floats) if w is not NULL, in an
   order that depends only on the records (sorted by a hash of their contents), not on the
   order the estimator's lanes recorded them. */
sms_status sms_ctx_training_samples(sms_ctx *ctx, int64_t *count, float *x, float *uv, float *w);
sms_status sms_ctx_clear_training(sms_ctx *ctx);
/* Network inputs (device, n x 9) of the queries of an estimate with this seed: shading point,
   shading normal and the estimator's light sample, positions mapped by (p - c) / h with c the
   centre and h the largest half extent of the scene's bounding box. */
sms_status sms_ss_proposal_inputs(sms_ctx *ctx, const sms_ss_query *q, int64_t n, uint64_t seed, float *x);
/* Raw network outputs (device, n x 48) the proposal gives the queries of an estimate with this
   seed -- the exact values sms_ss_estimate_batch turns into its seed mixtures. */
sms_status sms_ss_proposal_outputs(sms_ctx *ctx, const sms_ss_query *q, int64_t n, uint64_t seed, float *out);
/* The same for the glint estimator (sms_glint_estimate_batch): with the proposal set on a glinty
   normal-mapped shape, the mixture is over the pixel footprint -- the reference's footprint
   sample xi in [0,1)^2, uv = uv0 + (xi.x - 1/2) duv_dx + (xi.y - 1/2) duv_dy
   (manifold_glints.cpp:252-259) -- and it seeds the first walk and every Bernoulli trial (and
   every member of a biased trial set); a mixture sample outside [0,1)^2 is a failed walk with 0
   iterations.  Inputs: shading point, shading normal and the glint estimator's light sample.
   The training capture (SMS_CAPTURE_SOLUTIONS) records converged first walks as footprint
   samples xi. */
sms_status sms_glint_proposal_inputs(sms_ctx *ctx, const sms_glint_query *q, int64_t n, uint64_t seed, float *x);
sms_status sms_glint_proposal_outputs(sms_ctx *ctx, const sms_glint_query *q, int64_t n, uint64_t seed, float *out);
/* The same for the multi-scatter estimator (its light samples come from the multi-scatter
   emitters). */
sms_status sms_ms_proposal_inputs(sms_ctx *ctx, const sms_ss_query *q, int64_t n, uint64_t seed, float *x);
sms_status sms_ms_proposal_outputs(sms_ctx *ctx, const sms_ss_query *q, int64_t n, uint64_t seed, float *out);

/* ---------------------------------------------------------------- frame graphs
   No counterpart in the reference (which renders through per-thread scalar calls).  A graph
   holds one estimator call's launch sequence on the context stream -- frame start (seed and
   work counters), the estimator's kernels, a statistics fold -- captured once, then replayed by
   sms_graph_launch with a new seed and no per-kernel host work; launches are asynchronous and
   may follow each other without host synchronisation.  Each launch gives the results of the
   matching sms_*_estimate_batch call with that seed, bit for bit.  The graph keeps the
   context's buffers, configuration, schedule, proposal and training capture as they were at
   creation: any later change of them (including a larger estimate on the same context, which
   grows the work buffers) makes sms_graph_launch return SMS_E_INVALID until the graph is
   recreated.  Queries, outputs and (for the pass graph) film buffers are device memory that
   must outlive the graph; the proposal network's weights are read at each launch, so
   training between launches is seen by the next one. */
typedef struct sms_graph sms_graph;
/* Graph of sms_glint_estimate_batch(ctx, q, n, seed, out_rgb, ...). */
sms_status sms_glint_graph_create(sms_ctx *ctx, const sms_glint_query *q, int64_t n, float *out_rgb, sms_graph **out);
/* Graph of sms_ss_estimate_batch(ctx, q, n, seed, out_rgb, ...). */
sms_status sms_ss_graph_create(sms_ctx *ctx, const sms_ss_query *q, int64_t n, float *out_rgb, sms_graph **out);
/* Graph of sms_ms_estimate_batch(ctx, q, n, seed, out_rgb, ...). */
sms_status sms_ms_graph_create(sms_ctx *ctx, const sms_ss_query *q, int64_t n, float *out_rgb, sms_graph **out);
/* Graph of one single-scatter frame pass: sms_ss_make_queries(ctx, spp, seed, q, weight), then
   sms_ss_estimate_batch(ctx, q, w*h*spp, seed, out_rgb), then sms_film_accumulate(ctx, out_rgb,
   weight, w*h, spp, image, image_sq, image_n) (image_sq and image_n may be NULL). */
sms_status sms_ss_pass_graph_create(sms_ctx *ctx, int32_t spp, sms_ss_query *q, float *weight, float *out_rgb,
                                    float *image, float *image_sq, float *image_n, sms_graph **out);
/* Graph of sms_glint_estimate_frame(ctx, spp, seed, out_rgb): the whole frame, queries included. */
sms_status sms_glint_frame_graph_create(sms_ctx *ctx, int32_t spp, float *out_rgb, sms_graph **out);
/* Enqueue one replay with `seed` on the context stream. */
sms_status sms_graph_launch(sms_graph *g, uint64_t seed);
/* Waits for the stream; *stats (if not NULL) = the statistics summed over the launches since
   creation or the last reset (bernoulli_max: their maximum); reset != 0 starts a new sum. */
sms_status sms_graph_stats(sms_graph *g, sms_stats *stats, int32_t reset);
sms_status sms_graph_destroy(sms_graph *g);

/* ---------------------------------------------------------------- path integrators
   The path tracers around the SMS terms, over the scene's diffuse, smooth dielectric and smooth
   conductor BSDFs, normal-mapped ones (the nested BSDF sampled in the perturbed frame,
   src/bsdfs/normalmap.cpp:117-134, without the LEAN fallback; scenes with rough shapes are
   refused with SMS_E_UNSUPPORTED) and area (on spheres, rectangles and meshes), point and
   directional emitters:
     SMS_PATH_SMS_SS       path_sms_ss (src/integrators/path_sms_ss.cpp:69-251): emitter hits,
                           emitter sampling on smooth BSDFs, BSDF sampling, and at every caustic
                           receiver that is not a caster the SMS estimate of
                           sms_ss_estimate_batch; light paths through a caster found by BSDF
                           sampling are dropped -- except, with config.mnee_init and not
                           biased_mnee, those an MNEE walk from the receiver does not re-find
                           (:196-241)
     SMS_PATH_FILTERED_SS  path_filtered_ss (src/integrators/path_filtered_ss.cpp:37-194): the
                           same light paths by BSDF / emitter sampling only
     SMS_PATH_SMS_MS       path_sms_ms (src/integrators/path_sms_ms.cpp:175-380): chains of
                           config.bounces vertices on multi-scatter casters and bouncers, the SMS
                           estimate of sms_ms_estimate_batch at every receiver; no emitter sampling
                           on casters; emitter hits through them only on purely specular camera
                           paths -- and, with config.mnee_init and not biased_mnee, those an MNEE
                           walk from the chain's receiver does not re-find vertex for vertex
                           (MNEEHelper, :54-143; :290-329)
     SMS_PATH_FILTERED_MS  path_filtered_ms (src/integrators/path_filtered_ms.cpp:57-186): the
                           same chains by BSDF / emitter sampling, tracked by PathRecord (:237-300)
   Samples are PCG32 streams of (seed, sample index) (stream 0xFFFFFFFD: film jitter, then per
   bounce the Russian-roulette, emitter and BSDF samples in the reference's order); the SMS
   estimates of a pass run in batches of SMS_PATH_CHUNK samples, each batch's queries in
   (sample, depth) order with its own estimator seed.  max_depth -1 (the reference's unbounded
   default) ends paths at SMS_PATH_MAX_DEPTH. */
#define SMS_PATH_SMS_SS       0
#define SMS_PATH_FILTERED_SS  1
#define SMS_PATH_SMS_MS       2
#define SMS_PATH_FILTERED_MS  3
#define SMS_PATH_MAX_DEPTH    16
#define SMS_PATH_CHUNK        (1 << 20)
typedef struct sms_path_config {
    int32_t integrator;                 /* SMS_PATH_*                                           */
    int32_t max_depth;                  /* "max_depth" (MonteCarloIntegrator, -1: unbounded)    */
    int32_t rr_depth;                   /* "rr_depth" (5)                                        */
    int32_t disable_reflected_emitters; /* "disable_reflected_emitters" (path_sms_ss.cpp:59)     */
    int32_t biased_mnee;                /* "biased_mnee" (path_sms_ss.cpp:60)                    */
    /* No counterpart in the reference: 1 keeps only the single-scatter caustic light paths
       whose receiver is the path's first non-specular vertex (SMS mode: the SMS term there and
       MNEE completions through it; filtered mode: BSDF-sampled emitter hits through one caster
       after it) -- the term the bench's caustic images measure; the path's samples and every
       other choice are those of the full integrator. */
    int32_t caustics_only;
    float   filter_stddev;              /* gaussian rfilter "stddev" (src/rfilters/gaussian.cpp) */
    int32_t pad;
} sms_path_config;

/* Defaults of the reference's integrators and film: max_depth -1, rr_depth 5, stddev 0.5. */
sms_status sms_path_config_default(sms_path_config *pc);
/* One pass of `spp` samples per pixel over the camera's w x h film: per-sample radiance
   (device, n = w*h*spp in pixel-major order, x 3 floats) and film positions (device, n x 2
   floats, in pixels: pixel + jitter), and the summed statistics of the pass's SMS estimates
   (path_sms_ss; zeros for path_filtered_ss). */
sms_status sms_path_samples(sms_ctx *ctx, const sms_path_config *pc, int32_t spp, uint64_t seed, float *rgb,
                            float *pos, sms_stats *stats);
/* Film reconstruction of one pass (ImageBlock::put, src/librender/imageblock.cpp:80-165, with
   the Gaussian filter of `stddev`, radius 4 stddev, discretized as rfilter.h:62-66): adds,
   per pixel, the filter-weighted sums of the pass's samples (rgb, pos as written by
   sms_path_samples) to image (device, w*h*3) and the summed weights to weight (device, w*h);
   the developed pixel is image / weight.  Samples with a non-finite or negative (below -1e-5)
   channel are dropped whole.  stddev 0: box filter (each sample into its own pixel, weight 1). */
sms_status sms_film_splat(sms_ctx *ctx, const float *rgb, const float *pos, int32_t spp, float stddev, float *image,
                          float *weight);

/* ---------------------------------------------------------------- device memory helpers */
sms_status sms_device_alloc(sms_ctx *ctx, size_t bytes, void **ptr);
sms_status sms_device_free(sms_ctx *ctx, void *ptr);
sms_status sms_copy_h2d(sms_ctx *ctx, void *dst, const void *src, size_t bytes);
sms_status sms_copy_d2h(sms_ctx *ctx, void *dst, const void *src, size_t bytes);

#ifdef __cplusplus
}
#endif

#endif /* SMS_H */
