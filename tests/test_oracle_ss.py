"""Pins of the oracle's single-scatter solver (a4-a11 of the hot path) by mathematics that does
not reuse its code:

* brute force: on the glass dome every converged walk must land on a refraction solution
  found by an FP64 root search on the great circle through receiver, light and centre;
* closed form: on a plane mirror the walk must converge to the image-point construction;
* finite differences of both constraints' Newton steps: tests/test_oracle_pins.py;
* statistics: the unbiased (Bernoulli) estimator and the biased (M = 64) estimator agree in
  expectation when the trial set finds every solution.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_lib
from smsamd import abi, scenes, workloads

ETA = 1.5


def _refract(w, n, eta_):
    eta = 1.0 / eta_
    if np.dot(w, n) < 0:
        eta = 1.0 / eta
        n = -n
    c = np.dot(w, n)
    root = 1.0 - eta * eta * (1.0 - c * c)
    if root < 0:
        return None
    return -eta * (w - c * n) - n * math.sqrt(root)


def _dome_roots(x, y, c=np.zeros(3), r=1.0):
    """All refraction solutions p on the sphere in the plane of (x, y, c), FP64."""
    e1 = (x - c) / np.linalg.norm(x - c)
    t = (y - c) - np.dot(y - c, e1) * e1
    e2 = t / np.linalg.norm(t)
    nrm = np.cross(e1, e2)

    def f(th):
        p = c + r * (math.cos(th) * e1 + math.sin(th) * e2)
        n = (p - c) / r
        wi = (x - p) / np.linalg.norm(x - p)
        wo = (y - p) / np.linalg.norm(y - p)
        wt = _refract(wi, n, ETA)
        if wt is None:
            return None
        return float(np.dot(np.cross(wt, wo), nrm))

    ths = np.linspace(0, 2 * math.pi, 20001)
    vals = [f(t) for t in ths]
    roots = []
    for i in range(len(ths) - 1):
        a, b = vals[i], vals[i + 1]
        if a is None or b is None or a * b > 0:
            continue
        lo, hi, flo = ths[i], ths[i + 1], a
        for _ in range(60):
            mid = 0.5 * (lo + hi)
            fm = f(mid)
            if fm is None:
                break
            if flo * fm <= 0:
                hi = mid
            else:
                lo, flo = mid, fm
        th = 0.5 * (lo + hi)
        p = c + r * (math.cos(th) * e1 + math.sin(th) * e2)
        n = (p - c) / r
        wi = (x - p) / np.linalg.norm(x - p)
        wo = (y - p) / np.linalg.norm(y - p)
        wt = _refract(wi, n, ETA)
        if wt is not None and np.dot(wt, wo) > 0.999:
            roots.append(p)
    return roots


@pytest.fixture(scope="module")
def dome():
    sc = scenes.sphere_dome()
    o = oracle_lib.Oracle(sc, abi.SmsConfig.default())
    yield sc, o
    o.close()


def test_dome_walks_land_on_brute_force_roots(dome):
    sc, o = dome
    rng = np.random.default_rng(5)
    checked = found_all = cases = 0
    for case in range(16):
        x = np.array([rng.uniform(-0.6, 0.6), 0.0, rng.uniform(-0.6, 0.6)])
        y = np.array([0.0, 5.0, 0.0]) + 0.1 * workloads.uniform_sphere(rng.random(1), rng.random(1))[0]
        roots = _dome_roots(x, y)
        if not roots:
            continue
        cases += 1
        w = np.zeros(400, abi.SS_WALK_IN)
        w["si_p"] = x
        w["shape"] = 0
        w["ei"]["p"] = y
        w["ei"]["kind"] = abi.EI_AREA
        w["n_offset"] = (0, 0, 1)
        w["sample"] = rng.random((400, 2))
        out = o.ss_walks(w)
        ok = out["success"] == 1
        assert ok.mean() > 0.3
        hit = set()
        for p in out["p"][ok].astype(np.float64):
            d = [np.linalg.norm(p - q) / np.linalg.norm(q) for q in roots]
            k = int(np.argmin(d))
            assert d[k] < 1e-4, (case, p, roots)
            hit.add(k)
            checked += 1
        found_all += len(hit) == len(roots)
    assert cases >= 10
    assert checked > 1000
    assert found_all >= cases - 2


def test_plane_mirror_closed_form():
    """Conductor mirror x = 1.2 (facing -x); receiver on the floor, point light: the solution is
    where the segment receiver -> mirrored light crosses the mirror plane."""
    sc = scenes.Scene()
    M = np.array([[0, 0, -1, 1.2], [0, 2.0, 0, 1.0], [2.0, 0, 0, 0.0]], np.float32)   # local z -> -x
    sc.add_rectangle(M, abi.BSDF_CONDUCTOR, flags=abi.FLAG_CAUSTIC_CASTER_SS, eta=0.0, k=1.0)
    sc.add_rectangle(scenes.plane_to_world((0, 0, 0), 10.0), abi.BSDF_DIFFUSE, flags=abi.FLAG_CAUSTIC_RECEIVER)
    sc.add_point_light((0.2, 2.0, 0.3), (1, 1, 1))
    o = oracle_lib.Oracle(sc, abi.SmsConfig.default())
    rng = np.random.default_rng(2)
    n = 300
    w = np.zeros(n, abi.SS_WALK_IN)
    xs = np.stack([rng.uniform(-1, 0.8, n), np.zeros(n), rng.uniform(-1, 1, n)], -1)
    w["si_p"] = xs
    w["shape"] = 0
    light = np.array([0.2, 2.0, 0.3])
    w["ei"]["p"] = light
    w["ei"]["kind"] = abi.EI_POINT
    w["n_offset"] = (0, 0, 1)
    w["sample"] = rng.random((n, 2))
    out = o.ss_walks(w)
    ok = out["success"] == 1
    assert ok.mean() > 0.5
    img = light.copy()
    img[0] = 2 * 1.2 - img[0]
    for x, p in zip(xs[ok], out["p"][ok].astype(np.float64)):
        t = (1.2 - x[0]) / (img[0] - x[0])
        ref = x + t * (img - x)
        assert np.linalg.norm(p - ref) / np.linalg.norm(ref) < 1e-4
    o.close()


def test_unbiased_and_biased_estimators_agree(dome):
    sc, _ = dome
    q = workloads.plane_queries(sc, 1500, 3, extent=0.8)
    ob = oracle_lib.Oracle(sc, abi.SmsConfig.default(biased=1, max_trials=64))
    rb, _ = ob.ss_estimate(q, 1)
    ou = oracle_lib.Oracle(sc, abi.SmsConfig.default(max_trials=1000000))
    runs = np.stack([ou.ss_estimate(q, s)[0][:, 0] for s in range(1, 9)])
    mean_u = runs.mean()
    se = runs.mean(0).std() / math.sqrt(runs.shape[1])
    assert abs(mean_u - rb[:, 0].mean()) < 4 * se + 1e-4 * abs(mean_u)
    ob.close()
    ou.close()


def test_camera_pass_refracts_by_snell_law():
    """Flat water (no waves): every camera sample that reaches the floor through one
    refraction carries weight (1/1.33)^2, and tracing its receiver's incident direction back to
    the water plane meets the camera ray under Snell's law; weight-1 samples pass beside the
    water square."""
    sc = scenes.pool(width=48, height=48, n=8, n_waves=0)
    orc = oracle_lib.Oracle(sc, abi.SmsConfig.default())
    q, w = orc.ss_queries(4, 3)
    floor = q["shape"] == 1
    assert floor.mean() > 0.3
    assert np.all(w[~floor] == 0) and np.all(q["shape"][~floor] == -1)
    o = np.asarray(sc.camera["origin"], np.float64)
    # the pool has no walls: samples that pass beside the water square reach the floor directly
    direct = np.isclose(w, 1.0)
    assert np.all(np.isclose(w[floor], 1.0) | np.isclose(w[floor], (1 / 1.33) ** 2, rtol=1e-6))
    pd = q["p"][floor & direct].astype(np.float64)
    s_ = (0.0 - o[1]) / (pd[:, 1] - o[1])
    x0 = o + s_[:, None] * (pd - o)                      # where the straight ray crosses y = 0
    assert (np.abs(x0[:, [0, 2]]).max(1) > 1.0 - 1e-5).all()
    sel = floor & ~direct
    assert sel.mean() > 0.2
    p = q["p"][sel].astype(np.float64)
    np.testing.assert_allclose(p[:, 1], -1.0, atol=1e-6)
    # floor frame: sh_s, sh_t, sh_n -> world direction of wi (toward the water)
    wi = q["wi"][sel].astype(np.float64)
    d = (q["sh_s"][sel] * wi[:, :1] + q["sh_t"][sel] * wi[:, 1:2] + q["sh_n"][sel] * wi[:, 2:3]).astype(np.float64)
    t = (0.0 - p[:, 1]) / d[:, 1]
    h = p + t[:, None] * d
    a = (h - o) / np.linalg.norm(h - o, axis=1, keepdims=True)          # incident, in air
    b = -d / np.linalg.norm(d, axis=1, keepdims=True)                    # transmitted, in water
    n = np.array([0.0, 1.0, 0.0])
    sin_a = np.linalg.norm(np.cross(a, n), axis=1)
    sin_b = np.linalg.norm(np.cross(b, n), axis=1)
    np.testing.assert_allclose(sin_a, 1.33 * sin_b, atol=2e-5)
    # coplanar: the refracted ray stays in the plane of incidence
    assert np.abs(np.einsum("ij,ij->i", np.cross(a, n), b)).max() < 2e-5
    orc.close()


def test_camera_pass_through_glass_sphere():
    """Glass sphere over the floor: samples that pass through the sphere refract twice,
    (1/1.5)^2 * 1.5^2 = 1; direct floor hits have weight 1; internal reflections keep 1."""
    sc = scenes.sphere_caustic(48, 48)
    orc = oracle_lib.Oracle(sc, abi.SmsConfig.default())
    q, w = orc.ss_queries(4, 5)
    hit = q["shape"] >= 0
    assert hit.mean() > 0.5
    np.testing.assert_allclose(w[hit], 1.0, rtol=1e-6)
    orc.close()


def test_nonfinite_newton_step_walk_fixture():
    """A pool walk whose Newton step turns NaN (tests/golden/pool_nan_walk.npy): its
    reprojection rays hit nothing, every iteration is consumed, the walk fails; the rays skip
    the BVH instead of visiting every node."""
    import os
    import time
    w = oracle_lib.upgrade_ss_walk_in(
        np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pool_nan_walk.npy")))
    sc = scenes.pool(512, 512)
    orc = oracle_lib.Oracle(sc, abi.SmsConfig.default(max_trials=1000000, uniqueness_threshold=1e-5), threads=1)
    t = time.perf_counter()
    o = orc.ss_walks(w)
    dt = time.perf_counter() - t
    assert o["success"][0] == 0 and o["iterations"][0] == 20 and o["init_valid"][0] == 1
    assert dt < 0.05, dt
    orc.close()


@pytest.mark.parametrize("mu,sig", [((0.1, -0.2), (0.04, 0.01, -0.012)), ((0.0, 0.3), (0.02, 0.05, 0.018))])
def test_sample_gaussian_moments(mu, sig):
    """SpecularManifold::sample_gaussian (manifold.h:328-345), the two-stage solver's offset
    slope: samples have the requested mean and covariance (a sign slip in the mixing matrix
    flips the correlation)."""
    lib = oracle_lib.load()
    lib.oracle_sample_gaussian.argtypes = [C.c_void_p] * 4
    rng = np.random.default_rng(2)
    n = 40000
    s = rng.random((n, 2)).astype(np.float32)
    m, sg = np.asarray(mu, np.float32), np.asarray(sig, np.float32)
    out = np.zeros((n, 2), np.float32)
    o = np.zeros(2, np.float32)
    for i in range(n):
        lib.oracle_sample_gaussian(m.ctypes.data, sg.ctypes.data, s[i].ctypes.data, o.ctypes.data)
        out[i] = o
    x = out.astype(np.float64)
    cov = np.cov(x.T)
    se = np.sqrt(np.array([sig[0], sig[1]]) / n)
    assert np.all(np.abs(x.mean(0) - mu) < 4 * se)
    assert cov[0, 0] == pytest.approx(sig[0], rel=0.04)
    assert cov[1, 1] == pytest.approx(sig[1], rel=0.04)
    assert cov[0, 1] == pytest.approx(sig[2], abs=0.04 * math.sqrt(sig[0] * sig[1]))


def _wall_queries(sc, n, seed):
    return workloads.plane_queries(sc, n, seed, extent=0.9, shape=0)


def test_twostage_estimator_agrees_with_single_stage():
    """The two-stage solver (manifold_ss.cpp:251-270) only changes how a walk reaches a
    solution of the normal-mapped mirror; the unbiased estimate's expectation is the same with
    and without it."""
    sc = scenes.mirror_wall(nm_res=64, n_scratches=16, seed=4)
    q = _wall_queries(sc, 600, 5)
    means, stats = {}, {}
    for ts in (0, 1):
        o = oracle_lib.Oracle(sc, abi.SmsConfig.default(max_trials=100000, twostage=ts))
        runs = np.stack([o.ss_estimate(q, s)[0][:, 0] for s in range(1, 13)]).astype(np.float64)
        st = o.ss_estimate(q, 1)[1]
        means[ts], stats[ts] = runs, st
        o.close()
    mu0, mu1 = means[0].mean(), means[1].mean()
    d = means[1].mean(0) - means[0].mean(0)            # paired by query: the queries' own spread cancels
    se = d.std() / math.sqrt(d.size)
    print(f"single {mu0:.5g} two-stage {mu1:.5g} se {se:.3g} converged/walk "
          f"{stats[0]['walks_converged'] / stats[0]['walks']:.3f} {stats[1]['walks_converged'] / stats[1]['walks']:.3f}")
    assert (means[0] > 0).mean() > 0.05 and (means[1] > 0).mean() > 0.05
    assert not np.array_equal(means[0], means[1])
    assert abs(mu0 - mu1) < 4 * se + 1e-3 * abs(mu0)
