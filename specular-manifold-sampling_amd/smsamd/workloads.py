"""Seeded inputs for the walk and estimator entry points (numpy, float32).

These are the synthetic batches the parity tests and the benchmark feed to both the GPU
library and the CPU oracle: receiver points on the ground plane of the sphere-caustic
scene, and pixel footprints on the glint plane.
"""
import numpy as np

from . import abi


def _normalize(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def uniform_sphere(u, v):
    z = 1.0 - 2.0 * v
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    phi = 2.0 * np.pi * u
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], -1)


def ss_walk_inputs(scene, n, seed, caster=0, light_shape=2, extent=3.0, p_start=None):
    """Single-scatter walks of the sphere-caustic scene: receiver x uniform in a disc of
    radius ``extent`` on y=0, emitter sample uniform on the light sphere, seed sample uniform."""
    rng = np.random.default_rng(seed)
    w = np.zeros(n, abi.SS_WALK_IN)
    r = extent * np.sqrt(rng.random(n))
    a = 2 * np.pi * rng.random(n)
    w["si_p"][:, 0] = r * np.cos(a)
    w["si_p"][:, 2] = r * np.sin(a)
    ls = scene.shapes[light_shape]
    c = np.asarray(ls["center"], np.float64)
    ld = uniform_sphere(rng.random(n), rng.random(n))
    w["ei"]["p"] = (c + ls["radius"] * ld).astype(np.float32)
    w["ei"]["n"] = ld.astype(np.float32)
    w["ei"]["d"] = _normalize(w["ei"]["p"].astype(np.float64) - w["si_p"]).astype(np.float32)
    w["ei"]["weight"] = 1.0
    w["ei"]["pdf"] = 1.0
    w["ei"]["kind"] = abi.EI_AREA
    w["ei"]["emitter"] = 0
    w["shape"] = caster
    w["n_offset"] = (0.0, 0.0, 1.0)
    w["sample"] = rng.random((n, 2)).astype(np.float32)
    if p_start is not None:
        w["p_start"] = p_start
    return w


def plane_queries(scene, n, seed, extent=3.0, shape=1):
    """Estimator queries on the diffuse ground plane (y=0) seen from the scene camera."""
    rng = np.random.default_rng(seed)
    q = np.zeros(n, abi.SS_QUERY)
    r = extent * np.sqrt(rng.random(n))
    a = 2 * np.pi * rng.random(n)
    p = np.zeros((n, 3))
    p[:, 0] = r * np.cos(a)
    p[:, 2] = r * np.sin(a)
    q["p"] = p
    q["n"] = (0.0, 1.0, 0.0)
    q["sh_n"] = (0.0, 1.0, 0.0)
    q["sh_s"] = (1.0, 0.0, 0.0)
    q["sh_t"] = (0.0, 0.0, -1.0)
    cam = np.asarray(scene.camera["origin"], np.float64)
    wi = _normalize(cam - p)
    q["wi"] = np.stack([wi[:, 0], -wi[:, 2], wi[:, 1]], -1)
    q["shape"] = shape
    return q


def glint_walk_inputs(scene, n, seed, footprint_texels=4.0, nm_res=None):
    """Glint walks on the 1x1 plane: random footprints (parallelograms of about
    ``footprint_texels`` texels), camera origin as sensor, the point light as emitter."""
    rng = np.random.default_rng(seed)
    res = nm_res or scene.normalmaps[0][0]
    w = np.zeros(n, abi.GLINT_WALK_IN)
    uv = 0.05 + 0.9 * rng.random((n, 2))
    w["uv"] = uv
    w["si_p"][:, 0] = uv[:, 0] - 0.5
    w["si_p"][:, 2] = 0.5 - uv[:, 1]
    w["frame_s"] = (1.0, 0.0, 0.0)
    w["frame_t"] = (0.0, 0.0, -1.0)
    w["frame_n"] = (0.0, 1.0, 0.0)
    s = footprint_texels / res
    ang = rng.random(n) * np.pi
    sk = 0.3 * (rng.random(n) - 0.5)
    w["duv_dx"] = np.stack([np.cos(ang), np.sin(ang)], -1) * s
    w["duv_dy"] = np.stack([-np.sin(ang + sk), np.cos(ang + sk)], -1) * s * (0.5 + rng.random(n))[:, None]
    w["sensor_p"] = scene.camera["origin"]
    w["ei_p"] = scene.emitters[0]["position"]
    w["shape"] = 0
    w["n_offset"] = (0.0, 0.0, 1.0)
    w["sample"] = rng.random((n, 2)).astype(np.float32)
    w["xi_start"] = (1.0, 1.0)
    return w


def ms_walk_inputs(scene, n, seed, caster=0, light_shape=3, extent=0.3, p_start=None, first_path=True,
                   seed_shapes=(-1, -1, -1, -1)):
    """Multi-scatter walks of the slab scene: receiver x uniform in a disc of radius ``extent``
    on the floor (y=0), emitter sample uniform on the light sphere, seed sample uniform."""
    rng = np.random.default_rng(seed)
    w = np.zeros(n, abi.MS_WALK_IN)
    r = extent * np.sqrt(rng.random(n))
    a = 2 * np.pi * rng.random(n)
    w["si_p"][:, 0] = r * np.cos(a)
    w["si_p"][:, 2] = r * np.sin(a)
    ls = scene.shapes[light_shape]
    c = np.asarray(ls["center"], np.float64)
    ld = uniform_sphere(rng.random(n), rng.random(n))
    w["ei"]["p"] = (c + ls["radius"] * ld).astype(np.float32)
    w["ei"]["n"] = ld.astype(np.float32)
    w["ei"]["d"] = _normalize(w["ei"]["p"].astype(np.float64) - w["si_p"]).astype(np.float32)
    w["ei"]["weight"] = 1.0
    w["ei"]["pdf"] = 1.0
    w["ei"]["kind"] = abi.EI_AREA
    w["ei"]["emitter"] = 0
    w["shape"] = caster
    w["sample"] = rng.random((n, 2)).astype(np.float32)
    w["first_path"] = 1 if first_path else 0
    w["seed_shapes"] = seed_shapes
    if p_start is not None:
        w["p_start"] = p_start
    return w


def floor_queries_at(points, camera_origin, shape):
    """Estimator queries at given floor points (y=0, normal +y), viewed from ``camera_origin``."""
    p = np.asarray(points, np.float64).reshape(-1, 3)
    q = np.zeros(len(p), abi.SS_QUERY)
    q["p"] = p
    q["n"] = (0.0, 1.0, 0.0)
    q["sh_n"] = (0.0, 1.0, 0.0)
    q["sh_s"] = (1.0, 0.0, 0.0)
    q["sh_t"] = (0.0, 0.0, -1.0)
    wi = _normalize(np.asarray(camera_origin, np.float64) - p)
    q["wi"] = np.stack([wi[:, 0], -wi[:, 2], wi[:, 1]], -1)
    q["shape"] = shape
    return q
