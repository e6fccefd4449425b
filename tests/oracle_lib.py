"""ctypes loader of the CPU oracle (oracle/liboracle.so) -- the checker of the parity tests.

Only the test-suite, __graft_entry__.smoke() and bench.py's CPU-baseline leg load it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from smsamd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")
# the op-counting build of the same source (counting float/double types, oracle/opcount.h)
OPCOUNT_SO = os.path.join(ORACLE_DIR, "liboracle_opcount.so")


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def load(path=ORACLE_SO):
    if not os.path.exists(path):
        subprocess.run(["make", "-s", "-C", ORACLE_DIR, os.path.basename(path)], check=True)
    lib = C.CDLL(path)
    vp, i64 = C.c_void_p, C.c_int64
    lib.oracle_create.argtypes = [C.POINTER(abi.SmsSceneDesc), C.POINTER(abi.SmsConfig)]
    lib.oracle_create.restype = vp
    lib.oracle_destroy.argtypes = [vp]
    lib.oracle_set_config.argtypes = [vp, C.POINTER(abi.SmsConfig)]
    lib.oracle_ss_walk_batch.argtypes = [vp, vp, i64, vp, C.c_int]
    lib.oracle_glint_walk_batch.argtypes = [vp, vp, i64, vp, C.c_int]
    lib.oracle_ss_estimate_batch.argtypes = [vp, vp, i64, C.c_uint64, vp, C.POINTER(abi.SmsStats), C.c_int]
    lib.oracle_glint_estimate_batch.argtypes = [vp, vp, i64, C.c_uint64, vp, C.POINTER(abi.SmsStats), C.c_int]
    lib.oracle_glint_make_queries.argtypes = [vp, C.c_int, C.c_uint64, vp, C.c_int]
    lib.oracle_ms_walk_batch.argtypes = [vp, vp, i64, vp, C.c_int]
    lib.oracle_ms_estimate_batch.argtypes = [vp, vp, i64, C.c_uint64, vp, C.POINTER(abi.SmsStats), C.c_int]
    lib.oracle_ray_intersect.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    lib.oracle_ss_step_at.argtypes = [vp, C.c_int, vp, vp, vp, vp, vp, vp, vp]
    lib.oracle_ms_tridiag_step.argtypes = [C.c_int, vp, vp, vp, vp, vp]
    lib.oracle_ms_tridiag_geo.argtypes = [C.c_int, vp, vp, vp]
    lib.oracle_ms_tridiag_geo.restype = C.c_float
    lib.oracle_lean_pdf_params.argtypes = [vp, vp, vp, vp]
    lib.oracle_lean_pdf_params.restype = C.c_float
    lib.oracle_lean_levels.argtypes = [vp, C.c_int]
    lib.oracle_lean_level.argtypes = [C.c_float, vp, vp, C.c_int]
    lib.oracle_lean_texel.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, vp]
    for fn in ("oracle_exp_d", "oracle_erf_d", "oracle_log2_d"):
        getattr(lib, fn).argtypes = [C.c_double]
        getattr(lib, fn).restype = C.c_double
    lib.oracle_acosf.argtypes = [C.c_float]
    lib.oracle_acosf.restype = C.c_float
    lib.oracle_atan2f.argtypes = [C.c_float, C.c_float]
    lib.oracle_atan2f.restype = C.c_float
    lib.oracle_sincosf.argtypes = [C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    lib.oracle_fresnel.argtypes = [C.c_float, C.c_float]
    lib.oracle_fresnel.restype = C.c_float
    lib.oracle_fresnel_conductor.argtypes = [C.c_float, C.c_float, C.c_float]
    lib.oracle_fresnel_conductor.restype = C.c_float
    lib.oracle_square_to_uniform_sphere.argtypes = [C.c_float, C.c_float, vp]
    lib.oracle_pcg32_first.argtypes = [C.c_uint64, C.c_uint64, vp, C.c_int]
    lib.oracle_sample_emitter_interaction.argtypes = [vp, vp, C.c_uint64, C.c_uint64, vp]
    lib.oracle_ss_estimate_batch_prop.argtypes = [vp, vp, i64, C.c_uint64, vp, C.c_int, C.c_float, vp,
                                                  C.POINTER(abi.SmsStats), C.c_int]
    lib.oracle_ss_estimate_batch_prop_anchored.argtypes = [vp, vp, i64, C.c_uint64, vp, C.c_int, C.c_float, C.c_float,
                                                          vp, C.POINTER(abi.SmsStats), C.c_int]
    lib.oracle_prop_anchors.argtypes = [vp, vp, i64, C.c_uint64, C.c_int, vp]
    lib.oracle_prop_inputs.argtypes = [vp, vp, i64, C.c_uint64, vp]
    lib.oracle_prop_inputs_ms.argtypes = [vp, vp, i64, C.c_uint64, vp]
    lib.oracle_ms_estimate_batch_prop.argtypes = [vp, vp, i64, C.c_uint64, vp, C.c_int, C.c_float, vp,
                                                  C.POINTER(abi.SmsStats), C.c_int]
    lib.oracle_ss_make_queries.argtypes = [vp, C.c_int, C.c_uint64, vp, vp, C.c_int]
    lib.oracle_scene_box.argtypes = [vp, vp, vp]
    lib.oracle_gmm_prepare.argtypes = [vp, i64, vp]
    lib.oracle_prop_seed.argtypes = [vp, C.c_float, C.c_uint64, C.c_uint64, C.c_uint64, i64, vp]
    lib.oracle_uv_to_point.argtypes = [vp, C.c_int, vp, i64, C.c_int, vp, vp]
    lib.oracle_glint_estimate_batch_prop.argtypes = [vp, vp, i64, C.c_uint64, vp, C.c_float, vp,
                                                     C.POINTER(abi.SmsStats), C.c_int]
    lib.oracle_glint_prop_inputs.argtypes = [vp, vp, i64, C.c_uint64, vp]
    lib.oracle_path_samples.argtypes = [vp, C.POINTER(abi.SmsPathConfig), C.c_int, C.c_uint64, vp, vp,
                                        C.POINTER(abi.SmsStats), C.c_int]
    lib.oracle_film_splat.argtypes = [vp, vp, vp, C.c_int, C.c_float, vp, vp, C.c_int]
    return lib


class Oracle:
    """One oracle scene (mirrors sms_ctx)."""

    def __init__(self, scene, cfg, threads=None, lib_path=ORACLE_SO):
        self.lib = load(lib_path)
        self.scene = scene
        self.desc = scene.desc()
        self.cfg = cfg
        self.h = self.lib.oracle_create(C.byref(self.desc), C.byref(cfg))
        self.threads = threads or min(8, os.cpu_count() or 1)

    def close(self):
        if self.h:
            self.lib.oracle_destroy(self.h)
            self.h = None

    def set_config(self, cfg):
        self.cfg = cfg
        self.lib.oracle_set_config(self.h, C.byref(cfg))

    def ss_walks(self, win):
        out = np.zeros(len(win), abi.SS_WALK_OUT)
        self.lib.oracle_ss_walk_batch(self.h, win.ctypes.data, len(win), out.ctypes.data, self.threads)
        return out

    def glint_walks(self, win):
        out = np.zeros(len(win), abi.GLINT_WALK_OUT)
        self.lib.oracle_glint_walk_batch(self.h, win.ctypes.data, len(win), out.ctypes.data, self.threads)
        return out

    def ss_estimate(self, q, seed):
        out = np.zeros((len(q), 3), np.float32)
        st = abi.SmsStats()
        self.lib.oracle_ss_estimate_batch(self.h, q.ctypes.data, len(q), seed, out.ctypes.data, C.byref(st), self.threads)
        return out, st.as_dict()

    def glint_estimate(self, q, seed):
        out = np.zeros((len(q), 3), np.float32)
        st = abi.SmsStats()
        self.lib.oracle_glint_estimate_batch(self.h, q.ctypes.data, len(q), seed, out.ctypes.data, C.byref(st), self.threads)
        return out, st.as_dict()

    def ms_walks(self, win):
        out = np.zeros(len(win), abi.MS_WALK_OUT)
        rc = self.lib.oracle_ms_walk_batch(self.h, win.ctypes.data, len(win), out.ctypes.data, self.threads)
        assert rc == 0, rc
        return out

    def ms_estimate(self, q, seed):
        out = np.zeros((len(q), 3), np.float32)
        st = abi.SmsStats()
        rc = self.lib.oracle_ms_estimate_batch(self.h, q.ctypes.data, len(q), seed, out.ctypes.data, C.byref(st),
                                               self.threads)
        assert rc == 0, rc
        return out, st.as_dict()

    def glint_queries(self, spp, seed):
        cam = self.scene.camera
        q = np.zeros(cam["width"] * cam["height"] * spp, abi.GLINT_QUERY)
        self.lib.oracle_glint_make_queries(self.h, spp, seed, q.ctypes.data, self.threads)
        return q

    def ss_queries(self, spp, seed):
        cam = self.scene.camera
        n = cam["width"] * cam["height"] * spp
        q = np.zeros(n, abi.SS_QUERY)
        w = np.zeros(n, np.float32)
        rc = self.lib.oracle_ss_make_queries(self.h, spp, seed, q.ctypes.data, w.ctypes.data, self.threads)
        assert rc == 0, rc
        return q, w

    # ---------------------------------------------------------------- learned seed proposal
    def ss_estimate_prop(self, q, seed, raw, shape, defensive, radius=0.0):
        """Estimator with the defensive-mixture seed of caster ``shape``; raw: (n, 48) outputs;
        radius > 0: anchored mixtures."""
        raw = np.ascontiguousarray(raw, np.float32)
        assert raw.shape == (len(q), 48)
        out = np.zeros((len(q), 3), np.float32)
        st = abi.SmsStats()
        rc = self.lib.oracle_ss_estimate_batch_prop_anchored(self.h, q.ctypes.data, len(q), seed, raw.ctypes.data,
                                                             shape, defensive, radius, out.ctypes.data, C.byref(st),
                                                             self.threads)
        assert rc == 0, rc
        return out, st.as_dict()

    def prop_anchors(self, q, seed, shape):
        """Anchors (n, 2; NaN: none) of the queries of an estimate on mesh ``shape``."""
        a = np.zeros((len(q), 2), np.float32)
        self.lib.oracle_prop_anchors(self.h, q.ctypes.data, len(q), seed, shape, a.ctypes.data)
        return a

    def prop_inputs(self, q, seed):
        x = np.zeros((len(q), 9), np.float32)
        self.lib.oracle_prop_inputs(self.h, q.ctypes.data, len(q), seed, x.ctypes.data)
        return x

    def ms_prop_inputs(self, q, seed):
        x = np.zeros((len(q), 9), np.float32)
        self.lib.oracle_prop_inputs_ms(self.h, q.ctypes.data, len(q), seed, x.ctypes.data)
        return x

    def ms_estimate_prop(self, q, seed, raw, shape, defensive):
        """Multi-scatter estimator with the defensive-mixture seed on caster ``shape``."""
        raw = np.ascontiguousarray(raw, np.float32)
        assert raw.shape == (len(q), 48)
        out = np.zeros((len(q), 3), np.float32)
        st = abi.SmsStats()
        rc = self.lib.oracle_ms_estimate_batch_prop(self.h, q.ctypes.data, len(q), seed, raw.ctypes.data, shape,
                                                    defensive, out.ctypes.data, C.byref(st), self.threads)
        assert rc == 0, rc
        return out, st.as_dict()

    def glint_estimate_prop(self, q, seed, raw, defensive):
        """Glint estimator with each query's defensive mixture over its footprint; raw: (n, 48)."""
        raw = np.ascontiguousarray(raw, np.float32)
        assert raw.shape == (len(q), 48)
        out = np.zeros((len(q), 3), np.float32)
        st = abi.SmsStats()
        rc = self.lib.oracle_glint_estimate_batch_prop(self.h, q.ctypes.data, len(q), seed, raw.ctypes.data, defensive,
                                                       out.ctypes.data, C.byref(st), self.threads)
        assert rc == 0, rc
        return out, st.as_dict()

    def glint_prop_inputs(self, q, seed):
        x = np.zeros((len(q), 9), np.float32)
        self.lib.oracle_glint_prop_inputs(self.h, q.ctypes.data, len(q), seed, x.ctypes.data)
        return x

    def path_samples(self, pc, spp, seed):
        """(rgb (n, 3), film positions (n, 2), stats) of one pass of the path integrator pc."""
        cam = self.scene.camera
        n = cam["width"] * cam["height"] * spp
        rgb = np.zeros((n, 3), np.float32)
        pos = np.zeros((n, 2), np.float32)
        st = abi.SmsStats()
        rc = self.lib.oracle_path_samples(self.h, C.byref(pc), spp, seed, rgb.ctypes.data, pos.ctypes.data,
                                          C.byref(st), self.threads)
        assert rc == 0, rc
        return rgb, pos, st.as_dict()

    def film_splat(self, rgb, pos, spp, stddev, image=None, weight=None):
        cam = self.scene.camera
        px = cam["width"] * cam["height"]
        image = np.zeros((px, 3), np.float32) if image is None else image
        weight = np.zeros(px, np.float32) if weight is None else weight
        rgb = np.ascontiguousarray(rgb, np.float32)
        pos = np.ascontiguousarray(pos, np.float32)
        rc = self.lib.oracle_film_splat(self.h, rgb.ctypes.data, pos.ctypes.data, spp, stddev, image.ctypes.data,
                                        weight.ctypes.data, self.threads)
        assert rc == 0, rc
        return image, weight

    def scene_box(self):
        c = np.zeros(3, np.float32)
        inv = C.c_float(0)
        self.lib.oracle_scene_box(self.h, c.ctypes.data, C.byref(inv))
        return c, float(inv.value)

    def uv_to_point(self, shape, uv, brute=False):
        uv = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
        p = np.zeros((len(uv), 3), np.float32)
        ok = np.zeros(len(uv), np.int32)
        rc = self.lib.oracle_uv_to_point(self.h, shape, uv.ctypes.data, len(uv), int(brute), p.ctypes.data,
                                         ok.ctypes.data)
        assert rc == 0, rc
        return p, ok.astype(bool)


def gmm_prepare(raw):
    lib = load()
    raw = np.ascontiguousarray(raw, np.float32).reshape(-1, 48)
    g = np.zeros_like(raw)
    lib.oracle_gmm_prepare(raw.ctypes.data, len(raw), g.ctypes.data)
    return g


def prop_seeds(g48, defensive, seed, qid, walk0, n):
    """(kind, u, v) of walks walk0 .. walk0 + n - 1 of query qid drawn from prepared mixture g48."""
    lib = load()
    g = np.ascontiguousarray(g48, np.float32).reshape(48)
    out = np.zeros((n, 3), np.float32)
    lib.oracle_prop_seed(g.ctypes.data, defensive, seed, qid, walk0, n, out.ctypes.data)
    return out


def upgrade_ss_walk_in(a):
    """Fixtures written before sms_ss_walk_in gained lean_sample: copy into the current layout
    (lean_sample 0); other arrays pass through."""
    names = a.dtype.names
    if not names or "sample" not in names or "p_start" not in names or "lean_sample" in names or "first_path" in names:
        return a
    out = np.zeros(a.shape, abi.SS_WALK_IN)
    for k in names:
        out[k] = a[k]
    return out
