"""Pins of the oracle's multi-scatter solver by mathematics that does not reuse its code:

* dense linear algebra: the block-Thomas Newton solve (dx = A^-1 C) and the generalized
  geometric term's block elimination (|det| of the vertex-0 block of -A^-1 B) equal FP64
  dense solves of the same block-tridiagonal systems;
* closed form: through a flat glass slab the converged two-vertex chain is the Snell path
  found by a one-dimensional FP64 root search (angle-difference and half-vector
  constraints alike);
* brute force: the unbiased multi-scatter estimate at a floor point under a flat slab equals
  the FP64 Monte-Carlo integral of light radiance x slab transmittance x cosine over the
  directions whose refracted ray reaches the light sphere.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_lib
from smsamd import abi, scenes, workloads

ETA = 1.5
SLAB_BOTTOM, SLAB_TOP, SLAB_HALF = 0.4, 0.6, 0.5


def _blocks(rng, n):
    dprev = rng.normal(0, 0.3, (n, 4)).astype(np.float32)
    dnext = rng.normal(0, 0.3, (n, 4)).astype(np.float32)
    dcur = rng.normal(0, 0.3, (n, 4)).astype(np.float32)
    dcur[:, 0] += 2.0
    dcur[:, 3] += 2.0
    return dprev, dcur, dnext


def _dense(dprev, dcur, dnext, rows):
    A = np.zeros((2 * rows, 2 * rows))
    for i in range(rows):
        A[2 * i:2 * i + 2, 2 * i:2 * i + 2] = dcur[i].reshape(2, 2)
        if i > 0:
            A[2 * i:2 * i + 2, 2 * i - 2:2 * i] = dprev[i].reshape(2, 2)
        if i < rows - 1:
            A[2 * i:2 * i + 2, 2 * i + 2:2 * i + 4] = dnext[i].reshape(2, 2)
    return A


@pytest.mark.parametrize("n", [1, 2, 3, 4])
def test_tridiagonal_step_matches_dense_solve(n):
    lib = oracle_lib.load()
    rng = np.random.default_rng(100 + n)
    for _ in range(20):
        dprev, dcur, dnext = _blocks(rng, n)
        Cv = rng.normal(0, 1, (n, 2)).astype(np.float32)
        dx = np.zeros((n, 2), np.float32)
        ok = lib.oracle_ms_tridiag_step(n, dprev.ctypes.data, dcur.ctypes.data, dnext.ctypes.data,
                                        Cv.ctypes.data, dx.ctypes.data)
        assert ok == 1
        ref = np.linalg.solve(_dense(dprev, dcur, dnext, n), Cv.astype(np.float64).ravel())
        np.testing.assert_allclose(dx.ravel(), ref, rtol=1e-4, atol=1e-5)


def test_tridiagonal_step_reports_singular_pivot():
    lib = oracle_lib.load()
    dprev = np.zeros((2, 4), np.float32)
    dnext = np.zeros((2, 4), np.float32)
    dcur = np.array([[1, 0, 0, 1], [1, 2, 2, 4]], np.float32)   # second block singular
    Cv = np.ones((2, 2), np.float32)
    dx = np.zeros((2, 2), np.float32)
    assert lib.oracle_ms_tridiag_step(2, dprev.ctypes.data, dcur.ctypes.data, dnext.ctypes.data,
                                      Cv.ctypes.data, dx.ctypes.data) == 0


@pytest.mark.parametrize("n", [2, 3, 4, 5, 6])
def test_tridiagonal_geometric_term_matches_dense(n):
    lib = oracle_lib.load()
    rng = np.random.default_rng(200 + n)
    for _ in range(20):
        dprev, dcur, dnext = _blocks(rng, n)
        g = lib.oracle_ms_tridiag_geo(n, dprev.ctypes.data, dcur.ctypes.data, dnext.ctypes.data)
        rows = n - 1
        A = _dense(dprev, dcur, dnext, rows)
        B = np.zeros((2 * rows, 2))
        B[2 * rows - 2:, :] = dnext[rows - 1].reshape(2, 2)
        M = -np.linalg.solve(A, B)
        ref = abs(np.linalg.det(M[0:2, :]))
        assert g == pytest.approx(ref, rel=2e-4, abs=1e-7)


def _slab_chain(x0, L):
    """FP64 Snell path floor -> bottom face -> top face -> L through the flat slab."""
    h = np.array([L[0] - x0[0], L[2] - x0[2]])
    D = float(np.linalg.norm(h))
    u = h / D if D > 0 else np.zeros(2)
    h_air = SLAB_BOTTOM + (L[1] - SLAB_TOP)
    t_glass = SLAB_TOP - SLAB_BOTTOM

    def f(a):
        g = math.asin(math.sin(a) / ETA)
        return h_air * math.tan(a) + t_glass * math.tan(g) - D

    lo, hi = 0.0, math.pi / 2 - 1e-9
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if f(mid) > 0:
            hi = mid
        else:
            lo = mid
    a = 0.5 * (lo + hi)
    g = math.asin(math.sin(a) / ETA)
    s1 = SLAB_BOTTOM * math.tan(a)
    v1 = np.array([x0[0] + u[0] * s1, SLAB_BOTTOM, x0[2] + u[1] * s1])
    s2 = t_glass * math.tan(g)
    v2 = np.array([v1[0] + u[0] * s2, SLAB_TOP, v1[2] + u[1] * s2])
    return v1, v2


@pytest.fixture(scope="module")
def flat_slab():
    return scenes.slab(width=32, height=32, n=8, amplitude=0.0)


@pytest.mark.parametrize("halfvector", [0, 1])
def test_flat_slab_walks_match_snell_chain(flat_slab, halfvector):
    cfg = abi.SmsConfig.default(bounces=2, halfvector_constraints=halfvector)
    o = oracle_lib.Oracle(flat_slab, cfg)
    w = workloads.ms_walk_inputs(flat_slab, 400, seed=11, extent=0.25)
    out = o.ms_walks(w)
    o.close()
    ok = out["success"] == 1
    assert ok.mean() > 0.5
    assert np.all(out["seed_valid"][ok] == 1)
    assert np.all(out["n_vertices"][ok] == 2)
    assert np.all(out["shapes"][ok][:, :2] == [1, 0])
    for i in np.nonzero(ok)[0]:
        v1, v2 = _slab_chain(w["si_p"][i].astype(np.float64), w["ei"]["p"][i].astype(np.float64))
        np.testing.assert_allclose(out["p"][i, 0], v1, atol=2e-4)
        np.testing.assert_allclose(out["p"][i, 1], v2, atol=2e-4)


def _fresnel(cos_i, eta):
    """Unpolarized dielectric Fresnel reflectance, incident from the eta=1 side (FP64)."""
    sin_t2 = (1.0 - cos_i * cos_i) / (eta * eta)
    cos_t = np.sqrt(np.maximum(0.0, 1.0 - sin_t2))
    rs = (cos_i - eta * cos_t) / (cos_i + eta * cos_t)
    rp = (eta * cos_i - cos_t) / (eta * cos_i + cos_t)
    return 0.5 * (rs * rs + rp * rp)


def _brute_force_irradiance(x0, c, r, radiance, n=4_000_000, seed=0):
    """Integral over directions at x0 of radiance * slab transmittance * cos(theta), FP64."""
    rng = np.random.default_rng(seed)
    axis = (c - x0) / np.linalg.norm(c - x0)
    half = 0.12
    cos_max = math.cos(half)
    z = 1.0 - rng.random(n) * (1.0 - cos_max)
    phi = 2 * np.pi * rng.random(n)
    s = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    t1 = np.cross(axis, [1.0, 0.0, 0.0])
    t1 /= np.linalg.norm(t1)
    t2 = np.cross(axis, t1)
    d = s[:, None] * (np.cos(phi)[:, None] * t1 + np.sin(phi)[:, None] * t2) + z[:, None] * axis
    p1 = x0 + d * (SLAB_BOTTOM / d[:, 1])[:, None]
    cos_a = d[:, 1]
    dh = d[:, [0, 2]] / ETA
    gy = np.sqrt(1.0 - (dh * dh).sum(1))
    p2 = p1.copy()
    p2[:, [0, 2]] += dh * ((SLAB_TOP - SLAB_BOTTOM) / gy)[:, None]
    p2[:, 1] = SLAB_TOP
    inside = (np.abs(p1[:, [0, 2]]) < SLAB_HALF).all(1) & (np.abs(p2[:, [0, 2]]) < SLAB_HALF).all(1)
    oc = p2 - c
    b = (oc * d).sum(1)
    disc = b * b - ((oc * oc).sum(1) - r * r)
    hit = inside & (disc > 0) & (-b - np.sqrt(np.maximum(disc, 0)) > 0)
    T = (1.0 - _fresnel(cos_a, ETA)) * (1.0 - _fresnel(gy, 1.0 / ETA))
    vals = np.where(hit, radiance * T * cos_a, 0.0)
    omega = 2 * np.pi * (1.0 - cos_max)
    return vals.mean() * omega, vals.std() * omega / math.sqrt(n)


@pytest.mark.parametrize("point", [(0.0, 0.0, 0.0), (0.12, 0.0, -0.05)])
def test_flat_slab_estimate_matches_brute_force(flat_slab, point):
    cfg = abi.SmsConfig.default(bounces=2)
    o = oracle_lib.Oracle(flat_slab, cfg)
    n = 80000
    x0 = np.asarray(point, np.float64)
    q = workloads.floor_queries_at(np.tile(x0, (n, 1)), flat_slab.camera["origin"], shape=2)
    est, st = o.ms_estimate(q, seed=3)
    o.close()
    light = flat_slab.shapes[3]
    radiance = 2000.0
    irr, irr_err = _brute_force_irradiance(x0, np.asarray(light["center"], np.float64), light["radius"], radiance)
    ref = 0.5 / math.pi * irr
    m = est[:, 0].astype(np.float64)
    se = m.std() / math.sqrt(n)
    assert st["walks_converged"] > 0
    assert abs(m.mean() - ref) < 4 * se + 4 * 0.5 / math.pi * irr_err + 1e-3 * ref, (m.mean(), ref, se)
    np.testing.assert_array_equal(est[:, 0], est[:, 1])


def test_ms_mixture_estimator_is_unbiased():
    """Two-vertex slab chains seeded through a concentrated defensive mixture on the top face:
    same mean as the reference's uniform seeds within the standard errors, more converged walks."""
    sc = scenes.slab(width=32, height=32, n=64)
    cfg = abi.SmsConfig.default(bounces=2, max_trials=100000)
    orc = oracle_lib.Oracle(sc, cfg)
    q = workloads.plane_queries(sc, 24, 5, extent=0.3, shape=2)
    light = np.asarray(sc.shapes[3]["center"], np.float64)
    p = q["p"].astype(np.float64)
    hit = p + (0.6 / light[1]) * (light - p)               # toward the light, at the top face
    u, v = np.clip(hit[:, 0] + 0.5, 0.02, 0.98), np.clip(hit[:, 2] + 0.5, 0.02, 0.98)
    raw = np.zeros((len(q), 48), np.float32)
    for c in range(8):
        raw[:, 6 * c + 1] = np.log(u / (1 - u)) + 0.03 * (c - 3.5)
        raw[:, 6 * c + 2] = np.log(v / (1 - v)) - 0.03 * (c - 3.5)
        raw[:, 6 * c + 3] = raw[:, 6 * c + 4] = np.log(0.08)
    reps = 600
    qq, rr = np.concatenate([q] * reps), np.concatenate([raw] * reps)
    uni, su = orc.ms_estimate(qq, 3)
    mix, sm = orc.ms_estimate_prop(qq, 3, rr, 0, 0.2)
    lu, lm = uni[:, 0].astype(np.float64), mix[:, 0].astype(np.float64)
    se = np.sqrt(lu.var() / len(lu) + lm.var() / len(lm))
    print(f"uniform {lu.mean():.5f} mixture {lm.mean():.5f} se {se:.5f} conv {su['walks_converged'] / su['walks']:.3f} "
          f"{sm['walks_converged'] / sm['walks']:.3f}")
    assert lu.mean() > 0
    assert abs(lu.mean() - lm.mean()) < 4 * se
    assert sm["walks_converged"] / sm["walks"] > su["walks_converged"] / su["walks"]
    orc.close()
