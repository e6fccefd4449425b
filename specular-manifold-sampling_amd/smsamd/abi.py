"""ctypes / numpy mirror of ``include/sms.h`` (the C ABI of the gfx950 SMS library).

Record types used in batches (walk inputs/outputs, estimator queries) are numpy structured
dtypes whose layout equals the C structs (all members are 4-byte scalars, so there is no
padding); scene-description structs that carry pointers are ctypes Structures.
"""
import ctypes as C
import os

import numpy as np

SMS_OK = 0
STATUS_NAMES = {0: "SMS_OK", -1: "SMS_E_INVALID", -2: "SMS_E_HIP", -3: "SMS_E_OOM",
                -4: "SMS_E_QUEUE_FULL", -5: "SMS_E_UNSUPPORTED", -6: "SMS_E_NO_DEVICE"}

SHAPE_SPHERE, SHAPE_RECTANGLE, SHAPE_MESH = 0, 1, 2
BSDF_DIFFUSE, BSDF_DIELECTRIC, BSDF_CONDUCTOR, BSDF_NORMALMAP = 0, 1, 2, 3
FLAG_CAUSTIC_CASTER_SS, FLAG_CAUSTIC_CASTER_MS = 0x01, 0x02
FLAG_CAUSTIC_RECEIVER, FLAG_CAUSTIC_BOUNCER, FLAG_GLINTY = 0x04, 0x08, 0x10
EMITTER_AREA, EMITTER_POINT, EMITTER_DIRECTIONAL = 0, 1, 2
EI_AREA, EI_POINT, EI_DIRECTIONAL = 0, 1, 2


class SmsConfig(C.Structure):
    """sms_config: field-for-field SMSConfig (include/mitsuba/render/manifold.h:11-28)."""
    _fields_ = [("biased", C.c_int32), ("twostage", C.c_int32),
                ("halfvector_constraints", C.c_int32), ("mnee_init", C.c_int32),
                ("step_scale", C.c_float), ("max_iterations", C.c_int32),
                ("solver_threshold", C.c_float), ("uniqueness_threshold", C.c_float),
                ("max_trials", C.c_int32), ("bounces", C.c_int32),
                ("bsdf_strategy_only", C.c_int32), ("sms_strategy_only", C.c_int32)]

    @classmethod
    def default(cls, **kw):
        """SMSConfig() plus the integrator property defaults (path_sms_ss.cpp:49-57)."""
        c = cls(0, 0, 0, 0, 1.0, 20, 1e-5, 1e-4, -1, 1, 0, 0)
        for k, v in kw.items():
            setattr(c, k, v)
        return c


class SmsShapeDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("bsdf", C.c_int32), ("flags", C.c_int32), ("emitter", C.c_int32),
                ("center", C.c_float * 3), ("radius", C.c_float), ("to_world", C.c_float * 12),
                ("eta", C.c_float), ("k", C.c_float), ("reflectance", C.c_float * 3),
                ("normalmap", C.c_int32), ("tiles", C.c_float),
                ("n_vertices", C.c_int32), ("n_faces", C.c_int32),
                ("positions", C.POINTER(C.c_float)), ("normals", C.POINTER(C.c_float)),
                ("texcoords", C.POINTER(C.c_float)), ("faces", C.POINTER(C.c_uint32)),
                ("roughness", C.c_float)]


class SmsEmitterDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("shape", C.c_int32), ("radiance", C.c_float * 3),
                ("position", C.c_float * 3), ("intensity", C.c_float * 3),
                ("caustic_ss", C.c_int32), ("caustic_ms", C.c_int32),
                ("direction", C.c_float * 3), ("irradiance", C.c_float * 3)]


class SmsNormalmapDesc(C.Structure):
    _fields_ = [("res", C.c_int32), ("rgb", C.POINTER(C.c_uint8)), ("rgb32f", C.POINTER(C.c_float)),
                ("encoded_normals", C.c_int32)]


class SmsCameraDesc(C.Structure):
    _fields_ = [("origin", C.c_float * 3), ("target", C.c_float * 3), ("up", C.c_float * 3),
                ("fov_deg", C.c_float), ("width", C.c_int32), ("height", C.c_int32)]


class SmsSceneDesc(C.Structure):
    _fields_ = [("n_shapes", C.c_int32), ("shapes", C.POINTER(SmsShapeDesc)),
                ("n_emitters", C.c_int32), ("emitters", C.POINTER(SmsEmitterDesc)),
                ("n_normalmaps", C.c_int32), ("normalmaps", C.POINTER(SmsNormalmapDesc)),
                ("camera", SmsCameraDesc)]


class SmsStats(C.Structure):
    _fields_ = [("walks", C.c_uint64), ("walks_converged", C.c_uint64),
                ("newton_iterations", C.c_uint64), ("bernoulli_calls", C.c_uint64),
                ("bernoulli_trials", C.c_uint64), ("bernoulli_max", C.c_uint64),
                ("overflow_queries", C.c_uint64), ("tail_iterations", C.c_uint64)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class SmsSchedule(C.Structure):
    _fields_ = [("lane_trials", C.c_int32), ("blocks_per_cu", C.c_int32), ("trial_group", C.c_int32),
                ("tail_rounds", C.c_int32), ("long_chains", C.c_int32), ("long_trials", C.c_int32),
                ("ss_first_pass", C.c_int32), ("pad", C.c_int32)]

    KEYS = ("lane_trials", "blocks_per_cu", "trial_group", "tail_rounds", "long_chains", "long_trials",
            "ss_first_pass")

    def as_dict(self):
        return {k: int(getattr(self, k)) for k in self.KEYS}

    @classmethod
    def from_dict(cls, d):
        s = cls()
        for k in cls.KEYS:
            setattr(s, k, int(d[k]))
        return s


class SmsPathConfig(C.Structure):
    """sms_path_config: the path integrators' properties (path_sms_ss / path_filtered_ss)."""
    _fields_ = [("integrator", C.c_int32), ("max_depth", C.c_int32), ("rr_depth", C.c_int32),
                ("disable_reflected_emitters", C.c_int32), ("biased_mnee", C.c_int32),
                ("caustics_only", C.c_int32), ("filter_stddev", C.c_float), ("pad", C.c_int32)]

    @classmethod
    def default(cls, **kw):
        c = cls(integrator=PATH_SMS_SS, max_depth=-1, rr_depth=5, filter_stddev=0.5)
        for k, v in kw.items():
            if not hasattr(c, k):
                raise ValueError(k)
            setattr(c, k, v)
        return c


PATH_SMS_SS, PATH_FILTERED_SS, PATH_SMS_MS, PATH_FILTERED_MS = 0, 1, 2, 3
PATH_MAX_DEPTH = 16
PATH_CHUNK = 1 << 20

F3 = (np.float32, 3)
F2 = (np.float32, 2)
EI_DTYPE = np.dtype([("p", F3), ("n", F3), ("d", F3), ("weight", F3), ("pdf", np.float32),
                     ("kind", np.int32), ("emitter", np.int32)])
SS_WALK_IN = np.dtype([("si_p", F3), ("shape", np.int32), ("ei", EI_DTYPE), ("n_offset", F3),
                       ("sample", F2), ("p_start", F3), ("lean_sample", F2)])
SS_WALK_OUT = np.dtype([("success", np.int32), ("iterations", np.int32), ("final_shape", np.int32),
                        ("init_valid", np.int32), ("p", F3), ("n", F3), ("p_init", F3)])
GLINT_WALK_IN = np.dtype([("sensor_p", F3), ("ei_p", F3), ("si_p", F3), ("frame_s", F3),
                          ("frame_t", F3), ("frame_n", F3), ("uv", F2), ("duv_dx", F2),
                          ("duv_dy", F2), ("shape", np.int32), ("n_offset", F3), ("sample", F2),
                          ("xi_start", F2)])
GLINT_WALK_OUT = np.dtype([("success", np.int32), ("iterations", np.int32), ("uv", F2),
                           ("uv_init", F2)])
MS_MAX_BOUNCES = 4
MS_WALK_IN = np.dtype([("si_p", F3), ("shape", np.int32), ("ei", EI_DTYPE), ("sample", F2), ("p_start", F3),
                       ("first_path", np.int32), ("seed_shapes", (np.int32, MS_MAX_BOUNCES))])
MS_WALK_OUT = np.dtype([("success", np.int32), ("iterations", np.int32), ("seed_valid", np.int32),
                        ("n_vertices", np.int32), ("p", (np.float32, (MS_MAX_BOUNCES, 3))),
                        ("shapes", (np.int32, MS_MAX_BOUNCES))])
SS_QUERY = np.dtype([("p", F3), ("n", F3), ("sh_n", F3), ("sh_s", F3), ("sh_t", F3), ("wi", F3),
                     ("shape", np.int32), ("pad", np.int32)])
GLINT_QUERY = np.dtype([("sensor_p", F3), ("p", F3), ("n", F3), ("sh_n", F3), ("sh_s", F3),
                        ("sh_t", F3), ("dp_du", F3), ("dp_dv", F3), ("wi", F3), ("uv", F2),
                        ("duv_dx", F2), ("duv_dy", F2), ("shape", np.int32), ("pixel", np.int32)])

# sizes the C header implies (checked against the compiled library in the tests)
C_SIZES = {"sms_emitter_interaction": 60, "sms_ss_walk_in": 116, "sms_ss_walk_out": 52,
           "sms_glint_walk_in": 128, "sms_glint_walk_out": 24, "sms_ss_query": 80,
           "sms_glint_query": 140, "sms_config": 48, "sms_path_config": 32, "sms_stats": 64, "sms_ms_walk_in": 116,
           "sms_ms_walk_out": 80}
DTYPES = {"sms_emitter_interaction": EI_DTYPE, "sms_ss_walk_in": SS_WALK_IN,
          "sms_ms_walk_in": MS_WALK_IN, "sms_ms_walk_out": MS_WALK_OUT,
          "sms_ss_walk_out": SS_WALK_OUT, "sms_glint_walk_in": GLINT_WALK_IN,
          "sms_glint_walk_out": GLINT_WALK_OUT, "sms_ss_query": SS_QUERY,
          "sms_glint_query": GLINT_QUERY}

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libsms_amd.so")

# every symbol include/sms.h declares
EXPORTS = ["sms_config_default", "sms_ctx_create", "sms_ctx_destroy", "sms_ctx_set_config",
           "sms_ctx_set_schedule", "sms_ctx_get_schedule",
           "sms_ctx_last_error", "sms_ctx_device_info", "sms_ctx_sync", "sms_ss_walk_batch",
           "sms_glint_walk_batch", "sms_ss_estimate_batch", "sms_glint_estimate_batch",
           "sms_glint_make_queries", "sms_ctx_last_kernel_ms", "sms_ss_walk_batch_host",
           "sms_glint_walk_batch_host", "sms_ss_estimate_batch_host",
           "sms_glint_estimate_batch_host", "sms_device_alloc", "sms_device_free",
           "sms_copy_h2d", "sms_copy_d2h", "sms_ms_walk_batch", "sms_ms_estimate_batch",
           "sms_ms_walk_batch_host", "sms_ms_estimate_batch_host", "sms_glint_autotune",
           "sms_mlp_create", "sms_mlp_destroy", "sms_mlp_get_params", "sms_mlp_set_params",
           "sms_mlp_infer", "sms_mlp_grad", "sms_mlp_adam", "sms_mlp_train", "sms_glint_graph_create",
           "sms_ms_proposal_inputs", "sms_ms_proposal_outputs",
           "sms_ss_graph_create", "sms_ms_graph_create", "sms_ss_pass_graph_create", "sms_graph_launch",
           "sms_graph_stats", "sms_graph_destroy", "sms_ctx_set_proposal",
           "sms_ctx_set_training_capture", "sms_ctx_training_samples", "sms_ctx_clear_training",
           "sms_ss_proposal_inputs", "sms_ss_proposal_outputs", "sms_ss_make_queries",
           "sms_film_accumulate", "sms_ss_autotune", "sms_ms_autotune", "sms_scene_footprint",
           "sms_ss_estimate_batch_noalloc", "sms_glint_proposal_inputs", "sms_glint_proposal_outputs",
           "sms_path_config_default", "sms_path_samples", "sms_film_splat", "sms_glint_estimate_frame",
           "sms_glint_frame_graph_create", "sms_ctx_set_proposal_anchor", "sms_ss_proposal_anchors",
           "sms_obj_load", "sms_mesh_data_free", "sms_image_load", "sms_image_free", "sms_scene_load",
           "sms_scene_file_desc", "sms_scene_file_info", "sms_scene_file_property", "sms_scene_file_note",
           "sms_scene_file_free"]

# learned seed proposal (include/sms.h)
MLP_INPUTS = 9
MLP_OUTPUTS = 48
# scene and asset files
OBJ_FLIP_TEXCOORDS, OBJ_FACE_NORMALS = 0x1, 0x2
LOAD_SUBSTITUTE = 0x1
PROP_BOOL, PROP_INT, PROP_FLOAT, PROP_STRING = 0, 1, 2, 3


class SmsMeshData(C.Structure):
    _fields_ = [("n_vertices", C.c_int64), ("n_faces", C.c_int64), ("positions", C.POINTER(C.c_float)),
                ("normals", C.POINTER(C.c_float)), ("texcoords", C.POINTER(C.c_float)),
                ("faces", C.POINTER(C.c_uint32))]


class SmsImage(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("channels", C.c_int32),
                ("pixels", C.POINTER(C.c_float))]


class SmsProperty(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("type", C.c_int32), ("value", C.c_double), ("text", C.c_char * 128)]


class SmsSceneInfo(C.Structure):
    _fields_ = [("integrator", C.c_char * 32), ("n_properties", C.c_int32), ("spp", C.c_int32),
                ("filter_gaussian", C.c_int32), ("filter_stddev", C.c_float), ("n_notes", C.c_int32)]


MLP_PARAMS = 25904
CAPTURE_SOLUTIONS, CAPTURE_ESTIMATES = 0, 1


def load_library(path=LIB_PATH):
    """Load the HIP library.  Raises if it is missing: there is no CPU fallback."""
    if not os.path.exists(path):
        raise RuntimeError(f"libsms_amd.so not built ({path}); run __graft_entry__.build()")
    lib = C.CDLL(path)
    vp = C.c_void_p
    i64 = C.c_int64
    lib.sms_config_default.argtypes = [C.POINTER(SmsConfig)]
    lib.sms_config_default.restype = None
    lib.sms_ctx_create.argtypes = [C.POINTER(SmsSceneDesc), C.POINTER(SmsConfig), C.POINTER(vp)]
    lib.sms_scene_footprint.argtypes = [C.POINTER(SmsSceneDesc), C.POINTER(C.c_uint64), C.c_char_p, C.c_int32]
    lib.sms_ctx_destroy.argtypes = [vp]
    lib.sms_ctx_set_config.argtypes = [vp, C.POINTER(SmsConfig)]
    lib.sms_ctx_set_schedule.argtypes = [vp, C.POINTER(SmsSchedule)]
    lib.sms_ctx_get_schedule.argtypes = [vp, C.POINTER(SmsSchedule)]
    lib.sms_ctx_last_error.argtypes = [vp]
    lib.sms_ctx_last_error.restype = C.c_char_p
    lib.sms_ctx_device_info.argtypes = [vp, C.c_char_p, C.c_int32, C.POINTER(C.c_int32)]
    lib.sms_ctx_sync.argtypes = [vp]
    for name in ("sms_ss_walk_batch", "sms_glint_walk_batch", "sms_ss_walk_batch_host",
                 "sms_glint_walk_batch_host", "sms_ms_walk_batch", "sms_ms_walk_batch_host"):
        getattr(lib, name).argtypes = [vp, vp, i64, vp]
    for name in ("sms_ss_estimate_batch", "sms_glint_estimate_batch",
                 "sms_ss_estimate_batch_host", "sms_glint_estimate_batch_host",
                 "sms_ms_estimate_batch", "sms_ms_estimate_batch_host", "sms_ss_estimate_batch_noalloc"):
        getattr(lib, name).argtypes = [vp, vp, i64, C.c_uint64, vp, C.POINTER(SmsStats)]
    lib.sms_glint_make_queries.argtypes = [vp, C.c_int32, C.c_uint64, vp]
    lib.sms_glint_autotune.argtypes = [vp, vp, i64, C.c_uint64, vp, C.POINTER(SmsSchedule)]
    lib.sms_ss_autotune.argtypes = [vp, vp, i64, C.c_uint64, vp, C.POINTER(SmsSchedule)]
    lib.sms_ms_autotune.argtypes = [vp, vp, i64, C.c_uint64, vp, C.POINTER(SmsSchedule)]
    lib.sms_ctx_last_kernel_ms.argtypes = [vp, vp]
    lib.sms_device_alloc.argtypes = [vp, C.c_size_t, C.POINTER(vp)]
    lib.sms_device_free.argtypes = [vp, vp]
    lib.sms_copy_h2d.argtypes = [vp, vp, vp, C.c_size_t]
    lib.sms_copy_d2h.argtypes = [vp, vp, vp, C.c_size_t]
    lib.sms_mlp_create.argtypes = [vp, C.c_uint64, C.POINTER(vp)]
    lib.sms_mlp_destroy.argtypes = [vp]
    lib.sms_mlp_get_params.argtypes = [vp, vp]
    lib.sms_mlp_set_params.argtypes = [vp, vp]
    lib.sms_mlp_infer.argtypes = [vp, vp, i64, vp]
    lib.sms_mlp_grad.argtypes = [vp, vp, vp, vp, i64, vp, C.POINTER(C.c_float)]
    lib.sms_mlp_adam.argtypes = [vp, C.c_float]
    lib.sms_mlp_train.argtypes = [vp, vp, vp, vp, C.c_int32, vp, C.c_int32, C.c_float, vp]
    for name in ("sms_glint_graph_create", "sms_ss_graph_create", "sms_ms_graph_create"):
        getattr(lib, name).argtypes = [vp, vp, i64, vp, C.POINTER(vp)]
    lib.sms_ss_pass_graph_create.argtypes = [vp, C.c_int32, vp, vp, vp, vp, vp, vp, C.POINTER(vp)]
    lib.sms_graph_launch.argtypes = [vp, C.c_uint64]
    lib.sms_graph_stats.argtypes = [vp, C.POINTER(SmsStats), C.c_int32]
    lib.sms_graph_destroy.argtypes = [vp]
    lib.sms_ctx_set_proposal.argtypes = [vp, vp, C.c_int32, C.c_float]
    lib.sms_ctx_set_proposal_anchor.argtypes = [vp, C.c_float]
    lib.sms_ss_proposal_anchors.argtypes = [vp, vp, i64, C.c_uint64, vp]
    lib.sms_ctx_set_training_capture.argtypes = [vp, i64, C.c_int32]
    lib.sms_ctx_training_samples.argtypes = [vp, C.POINTER(i64), vp, vp, vp]
    lib.sms_ctx_clear_training.argtypes = [vp]
    lib.sms_ss_proposal_inputs.argtypes = [vp, vp, i64, C.c_uint64, vp]
    lib.sms_ss_proposal_outputs.argtypes = [vp, vp, i64, C.c_uint64, vp]
    lib.sms_ms_proposal_inputs.argtypes = [vp, vp, i64, C.c_uint64, vp]
    lib.sms_ms_proposal_outputs.argtypes = [vp, vp, i64, C.c_uint64, vp]
    lib.sms_glint_proposal_inputs.argtypes = [vp, vp, i64, C.c_uint64, vp]
    lib.sms_glint_proposal_outputs.argtypes = [vp, vp, i64, C.c_uint64, vp]
    lib.sms_ss_make_queries.argtypes = [vp, C.c_int32, C.c_uint64, vp, vp]
    lib.sms_film_accumulate.argtypes = [vp, vp, vp, i64, C.c_int32, vp, vp, vp]
    lib.sms_path_config_default.argtypes = [C.POINTER(SmsPathConfig)]
    lib.sms_glint_estimate_frame.argtypes = [vp, C.c_int32, C.c_uint64, vp, C.POINTER(SmsStats)]
    lib.sms_glint_frame_graph_create.argtypes = [vp, C.c_int32, vp, C.POINTER(vp)]
    lib.sms_path_samples.argtypes = [vp, C.POINTER(SmsPathConfig), C.c_int32, C.c_uint64, vp, vp, C.POINTER(SmsStats)]
    lib.sms_film_splat.argtypes = [vp, vp, vp, C.c_int32, C.c_float, vp, vp]
    lib.sms_obj_load.argtypes = [C.c_char_p, vp, C.c_int32, C.POINTER(C.POINTER(SmsMeshData)), C.c_char_p, C.c_int32]
    lib.sms_mesh_data_free.argtypes = [C.POINTER(SmsMeshData)]
    lib.sms_mesh_data_free.restype = None
    lib.sms_image_load.argtypes = [C.c_char_p, C.POINTER(C.POINTER(SmsImage)), C.c_char_p, C.c_int32]
    lib.sms_image_free.argtypes = [C.POINTER(SmsImage)]
    lib.sms_image_free.restype = None
    lib.sms_scene_load.argtypes = [C.c_char_p, C.POINTER(C.c_char_p), C.c_int32, C.c_int32, C.POINTER(vp),
                                   C.c_char_p, C.c_int32]
    lib.sms_scene_file_desc.argtypes = [vp]
    lib.sms_scene_file_desc.restype = C.POINTER(SmsSceneDesc)
    lib.sms_scene_file_info.argtypes = [vp, C.POINTER(SmsSceneInfo)]
    lib.sms_scene_file_property.argtypes = [vp, C.c_int32, C.POINTER(SmsProperty)]
    lib.sms_scene_file_note.argtypes = [vp, C.c_int32]
    lib.sms_scene_file_note.restype = C.c_char_p
    lib.sms_scene_file_free.argtypes = [vp]
    lib.sms_scene_file_free.restype = None
    for name in EXPORTS:
        fn = getattr(lib, name)
        if fn.restype is C.c_int and name not in ("sms_ctx_last_error", "sms_config_default"):
            fn.restype = C.c_int32
    return lib
