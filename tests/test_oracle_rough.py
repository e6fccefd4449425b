"""Pins of rough single-scatter SMS (manifold_ss.cpp:61-70 offset normals, :649-667 the glossy
branch of specular_reflectance) on a rough conductor mirror (isotropic Beckmann, roughness
alpha), by mathematics that does not reuse the oracle's code:

* brute force: the unbiased estimate at a floor point equals the FP64 Monte-Carlo integral over
  mirror points p and light points y of f_d cos_x |x-p|^-2 f_r(p) cos_p_in cos_p_out cos_y |y-p|^-2 L,
  f_r the Beckmann microfacet conductor BRDF D G F / (4 cos_i cos_o) with the complex-index
  Fresnel term -- the offset-normal sampling, the half-vector Jacobian 4 |wo.h| and the division
  by the offset-normal pdf have to combine into exactly this integral;
A 10 % error planted in the contribution fails the brute force.
"""
import math

import numpy as np
import pytest

import oracle_lib
from smsamd import abi, scenes, workloads
from test_oracle_pins import planted

ALPHA = 0.3
ETA_C = 0.2 + 3.0j
LIGHT_C, LIGHT_R, LIGHT_L = np.array([0.2, 2.0, 0.3]), 0.15, 40.0


def rough_mirror(alpha=ALPHA):
    sc = scenes.Scene()
    M = np.array([[0, 0, -1, 1.2], [0, 2.0, 0, 1.0], [2.0, 0, 0, 0.0]], np.float32)   # plane x = 1.2, normal -x
    sc.add_rectangle(M, abi.BSDF_CONDUCTOR, flags=abi.FLAG_CAUSTIC_CASTER_SS, eta=ETA_C.real, k=ETA_C.imag,
                     roughness=alpha)
    sc.add_rectangle(scenes.plane_to_world((0, 0, 0), 10.0), abi.BSDF_DIFFUSE, flags=abi.FLAG_CAUSTIC_RECEIVER)
    light = sc.add_sphere(tuple(LIGHT_C), LIGHT_R, abi.BSDF_DIFFUSE, reflectance=(0, 0, 0))
    sc.add_area_light(light, (LIGHT_L, LIGHT_L, LIGHT_L))
    sc.set_camera((-2.0, 2.5, 0.5), (0.8, 0.0, 0.0), (0, 1, 0), 45.0, 64, 64)
    return sc


def _fresnel_conductor(cos_i, eta):
    sin2 = 1.0 - cos_i * cos_i
    cos_t = np.sqrt(1.0 - sin2 / (eta * eta))
    rs = (cos_i - eta * cos_t) / (cos_i + eta * cos_t)
    rp = (eta * cos_i - cos_t) / (eta * cos_i + cos_t)
    return 0.5 * (np.abs(rs) ** 2 + np.abs(rp) ** 2)


def _g1(alpha, v_z, v_t2, vm):
    """Beckmann Smith masking by the rational approximation (v_t2: squared tangential length)."""
    a = v_z / (alpha * np.sqrt(np.maximum(v_t2, 1e-300)))
    g = np.where(a >= 1.6, 1.0, (3.535 * a + 2.181 * a * a) / (1.0 + 2.276 * a + 2.577 * a * a))
    g = np.where(v_t2 == 0.0, 1.0, g)
    return np.where(vm * v_z <= 0.0, 0.0, g)


def _brute_force(x, alpha=ALPHA, n=3_000_000, batches=4, seed=0):
    rng = np.random.default_rng(seed)
    nrm = np.array([-1.0, 0.0, 0.0])
    vals = []
    for _ in range(batches):
        zz, yy = rng.uniform(-2, 2, n), rng.uniform(-1, 3, n)
        p = np.stack([np.full(n, 1.2), yy, zz], -1)
        u, v = rng.random(n), rng.random(n)
        zc = 1 - 2 * v
        rc = np.sqrt(np.maximum(0, 1 - zc * zc))
        ln = np.stack([rc * np.cos(2 * np.pi * u), rc * np.sin(2 * np.pi * u), zc], -1)
        y = LIGHT_C + LIGHT_R * ln
        dx, dy = x - p, y - p
        rx2, ry2 = (dx * dx).sum(1), (dy * dy).sum(1)
        wi, wo = dx / np.sqrt(rx2)[:, None], dy / np.sqrt(ry2)[:, None]
        cos_i, cos_o = wi @ nrm, wo @ nrm
        cos_x = -wi[:, 1]                        # floor normal +y, direction x -> p
        cos_y = -(wo * ln).sum(1)                # light normal, direction y -> p
        h = wi + wo
        h /= np.linalg.norm(h, axis=1)[:, None]
        ch = h @ nrm
        ht2 = np.maximum(1.0 - ch * ch, 0.0)
        D = np.exp(-ht2 / (ch * ch) / (alpha * alpha)) / (np.pi * alpha * alpha * ch ** 4)
        D = np.where(D * ch > 1e-20, D, 0.0)
        G = _g1(alpha, cos_i, 1 - cos_i ** 2, (wi * h).sum(1)) * _g1(alpha, cos_o, 1 - cos_o ** 2, (wo * h).sum(1))
        F = _fresnel_conductor((wi * h).sum(1), ETA_C)
        fr = D * G * F / (4.0 * cos_i * cos_o)
        ok = (cos_i > 0) & (cos_o > 0) & (cos_x > 0) & (cos_y > 0)
        I = np.where(ok, 0.5 / np.pi * cos_x * cos_i / rx2 * fr * cos_o * cos_y / ry2 * LIGHT_L, 0.0)
        vals.append(I * 16.0 * 4 * np.pi * LIGHT_R ** 2)
    vals = np.concatenate(vals)
    return vals.mean(), vals.std() / math.sqrt(len(vals))


POINTS = [(0.3, 0.0, 0.2), (-0.2, 0.0, -0.4)]


def _estimate(x, n=60000, seed=5):
    sc = rough_mirror()
    o = oracle_lib.Oracle(sc, abi.SmsConfig.default(max_trials=1000000))
    q = workloads.floor_queries_at(np.tile(np.asarray(x, np.float64), (n, 1)), sc.camera["origin"], shape=1)
    est, st = o.ss_estimate(q, seed)
    o.close()
    m = est[:, 0].astype(np.float64)
    return m.mean(), m.std() / math.sqrt(n), st


@pytest.mark.parametrize("x", POINTS)
def test_rough_mirror_estimate_matches_brute_force(x):
    mean, se, st = _estimate(x)
    ref, ref_se = _brute_force(np.asarray(x, np.float64))
    print(f"{x}: rough SMS {mean:.5g} +- {se:.2g}, brute force {ref:.5g} +- {ref_se:.2g}, "
          f"walks {st['walks']} converged {st['walks_converged']}")
    assert ref > 0 and st["walks_converged"] > 0
    assert abs(mean - ref) < 4 * math.hypot(se, ref_se) + 1e-3 * ref
    assert 4 * math.hypot(se, ref_se) < 0.04 * ref


def test_rough_mirror_brute_force_detects_planted_error():
    with planted("ss_g"):
        mean, se, _ = _estimate(POINTS[0], n=30000)
    ref, ref_se = _brute_force(np.asarray(POINTS[0], np.float64), batches=2)
    assert abs(mean - ref) > 4 * math.hypot(se, ref_se) + 1e-3 * ref
