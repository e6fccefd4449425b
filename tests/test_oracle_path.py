"""Pins of the oracle's path integrators (path_sms_ss.cpp:69-251, path_filtered_ss.cpp:37-194)
and film (imageblock.cpp:80-165, gaussian.cpp).

The SMS estimator and the filtered path tracer share nothing on the caustic paths but the
receiver's BSDF: SMS solves for the specular vertex and divides by the Bernoulli-estimated
probability, the path tracer finds the light by BSDF sampling through the caster.  Their
caustic images must agree in expectation, and a 10 % error planted in the SMS geometric term
must break the agreement.  MNEE (mnee_init, one biased walk) plus the path tracer's completion
of the caustics MNEE misses must agree with both; the biased MNEE variant, which drops them,
must not.
"""
import contextlib
import ctypes as C
import math

import numpy as np
import pytest

import oracle_lib
from smsamd import abi, scenes


def _mean_se(v):
    v = np.asarray(v, np.float64)
    return v.mean(), v.std() / math.sqrt(len(v))


def _caustic(orc, pc, passes, spp=256, seed0=100):
    vals = []
    for s in range(passes):
        rgb, _, _ = orc.path_samples(pc, spp, seed0 + s)
        vals.append(rgb[:, 0])
    return np.concatenate(vals)


@contextlib.contextmanager
def _planted(site, factor):
    lib = oracle_lib.load()
    lib.oracle_plant.argtypes = [C.c_int, C.c_float]
    assert lib.oracle_plant(site, factor) == 0
    try:
        yield
    finally:
        lib.oracle_plant(site, 1.0)


@pytest.fixture(scope="module")
def dome():
    # the glass dome under a larger light: the path tracer's caustic converges quickly
    sc = scenes.sphere_dome(16, 16, light_radius=0.5, light_radiance=4.0)
    orc = oracle_lib.Oracle(sc, abi.SmsConfig.default(max_trials=100000), threads=8)
    yield sc, orc
    orc.close()


def test_sms_caustic_matches_filtered_path_tracer(dome):
    sc, orc = dome
    sms = abi.SmsPathConfig.default(max_depth=6, caustics_only=1)
    flt = abi.SmsPathConfig.default(integrator=abi.PATH_FILTERED_SS, max_depth=6, caustics_only=1)
    vs = _caustic(orc, sms, 16)
    vf = _caustic(orc, flt, 200, seed0=1000)
    ms, ss = _mean_se(vs)
    mf, sf = _mean_se(vf)
    se = math.hypot(ss, sf)
    print(f"SMS {ms:.6g} +- {ss:.2g}  filtered {mf:.6g} +- {sf:.2g}  ({(ms - mf) / se:+.2f} se)")
    assert ms > 0 and mf > 0
    assert abs(ms - mf) < 4 * se
    # per 4x4-pixel block (16 blocks): the caustic's shape agrees as well
    w = sc.camera["width"]
    blocks_s = vs.reshape(16, w, w, 256).transpose(1, 2, 0, 3).reshape(4, 4, 4, 4, -1).transpose(0, 2, 1, 3, 4)
    blocks_f = vf.reshape(200, w, w, 256).transpose(1, 2, 0, 3).reshape(4, 4, 4, 4, -1).transpose(0, 2, 1, 3, 4)
    for by in range(4):
        for bx in range(4):
            a, sa = _mean_se(blocks_s[by, bx].ravel())
            b, sb = _mean_se(blocks_f[by, bx].ravel())
            assert abs(a - b) < 5 * math.hypot(sa, sb) + 1e-12, (by, bx, a, b)
    # a 10 % error in the SMS geometric term is visible against the path tracer
    with _planted(0, 1.1):
        vp = _caustic(orc, sms, 16)
    mp, sp = _mean_se(vp)
    print(f"planted SMS {mp:.6g}: {(mp - mf) / math.hypot(sp, sf):+.2f} se from the path tracer")
    assert abs(mp - mf) > 4 * math.hypot(sp, sf)


def test_full_images_agree(dome):
    """The whole integrators: emitter hits, emitter sampling on the diffuse floor, SMS at every
    caustic receiver of the path against the path tracer's filtered connections."""
    sc, orc = dome
    vs = _caustic(orc, abi.SmsPathConfig.default(max_depth=6), 4)
    vf = _caustic(orc, abi.SmsPathConfig.default(integrator=abi.PATH_FILTERED_SS, max_depth=6), 16, seed0=1000)
    ms, ss = _mean_se(vs)
    mf, sf = _mean_se(vf)
    print(f"full SMS {ms:.6g} +- {ss:.2g}  filtered {mf:.6g} +- {sf:.2g}")
    assert abs(ms - mf) < 4 * math.hypot(ss, sf)
    # the filtered path tracer's caustic term is part of its full image, sample for sample
    # (same samples; the SMS integrator's estimator streams depend on the batch's query list)
    flt = dict(integrator=abi.PATH_FILTERED_SS, max_depth=6)
    rgb_c, _, _ = orc.path_samples(abi.SmsPathConfig.default(caustics_only=1, **flt), 64, 7)
    rgb_f, _, _ = orc.path_samples(abi.SmsPathConfig.default(**flt), 64, 7)
    assert (rgb_f >= rgb_c * (1 - 1e-6)).all() and (rgb_c > 0).any()


def test_caustic_term_matches_camera_pass(dome):
    """The SMS caustic term of the path tracer (SMS at the first non-specular vertex) and the
    camera pass of the SMS estimator (sms_ss_make_queries + sms_ss_estimate_batch) estimate the
    same image with different samples."""
    sc, orc = dome
    vs = _caustic(orc, abi.SmsPathConfig.default(max_depth=6, caustics_only=1), 8)
    vals = []
    for s in range(8):
        q, w = orc.ss_queries(256, 500 + s)
        rgb, _ = orc.ss_estimate(q, 500 + s)
        vals.append(w * rgb[:, 0])
    vc = np.concatenate(vals)
    ms, ss = _mean_se(vs)
    mc, sc_ = _mean_se(vc)
    print(f"path SMS term {ms:.6g} +- {ss:.2g}  camera pass {mc:.6g} +- {sc_:.2g}")
    assert abs(ms - mc) < 4 * math.hypot(ss, sc_)


def test_mnee_completion_is_unbiased():
    """Pool of many solutions per floor point: MNEE (one deterministic walk toward the light)
    finds one; path_sms_ss's unbiased MNEE adds, by BSDF sampling, the caustic paths the MNEE
    walk does not re-find (path_sms_ss.cpp:196-241) and agrees with SMS and the path tracer;
    biased_mnee drops them."""
    sc = scenes.pool(16, 16, n=64, light_radius=0.3, light_radiance=60.0)
    cfg_sms = abi.SmsConfig.default(max_trials=100000, uniqueness_threshold=1e-5, solver_threshold=1e-5)
    cfg_mnee = abi.SmsConfig.default(mnee_init=1, biased=1, max_trials=1, halfvector_constraints=1)
    res = {}
    for name, cfg, kw, passes in [("sms", cfg_sms, {}, 8), ("mnee", cfg_mnee, dict(biased_mnee=0), 8),
                                  ("mnee_biased", cfg_mnee, dict(biased_mnee=1), 4),
                                  ("filtered", cfg_sms, dict(integrator=abi.PATH_FILTERED_SS), 24)]:
        orc = oracle_lib.Oracle(sc, cfg, threads=8)
        res[name] = _mean_se(_caustic(orc, abi.SmsPathConfig.default(max_depth=6, caustics_only=1, **kw), passes))
        orc.close()
    print({k: (round(m, 6), round(s, 6)) for k, (m, s) in res.items()})
    mf, sf = res["filtered"]
    for k in ("sms", "mnee"):
        m, s = res[k]
        assert abs(m - mf) < 4 * math.hypot(s, sf), k
    mb, sb = res["mnee_biased"]
    assert mb < mf - 4 * math.hypot(sb, sf)


# ---------------------------------------------------------------- normal-mapped BSDFs in the path tracer
def _const_normalmap(res, sx=0.1, sy=0.05):
    n = np.array([-sx, -sy, 1.0])
    n /= np.linalg.norm(n)
    return np.tile((0.5 * n + 0.5).astype(np.float32), (res, res, 1))


def _nm_wall():
    sc = scenes.mirror_wall(normalmap=(_const_normalmap(64) * 255).round().astype(np.uint8))
    sc.shapes[2]["radius"] = 0.3
    sc.emitters[0]["radiance"] = (5.0, 5.0, 5.0)
    sc.set_camera((-1.2, 2.2, 0.0), (0.2, 0.0, 0.0), (0, 1, 0), 50.0, 16, 16)
    return sc


def test_normal_mapped_mirror_sms_matches_filtered_path_tracer():
    """A mirror whose normal map tilts every shading normal: the filtered path tracer samples
    the nested conductor in the perturbed frame (bsdfs/normalmap.cpp:117-134), SMS solves with
    the perturbed normals; the caustic terms agree.  An untilted map (the perturbed frame is the
    shading frame) gives the unmapped mirror's caustic, sample for sample."""
    sc = _nm_wall()
    orc = oracle_lib.Oracle(sc, abi.SmsConfig.default(max_trials=1000000), threads=8)
    ms, ss = _mean_se(_caustic(orc, abi.SmsPathConfig.default(max_depth=6, caustics_only=1), 8, seed0=100))
    flt = abi.SmsPathConfig.default(integrator=abi.PATH_FILTERED_SS, max_depth=6, caustics_only=1)
    mf, sf = _mean_se(_caustic(orc, flt, 100, seed0=100))
    print(f"SMS {ms:.6g} +- {ss:.2g}  filtered {mf:.6g} +- {sf:.2g}")
    assert ms > 0 and abs(ms - mf) < 4 * math.hypot(ss, sf)
    orc.close()
    flat = _nm_wall()
    flat.normalmaps[0] = (64, np.full((64, 64, 3), [0.5, 0.5, 1.0], np.float32))
    plain = _nm_wall()
    plain.shapes[1].update(bsdf=abi.BSDF_CONDUCTOR, normalmap=-1)
    a = oracle_lib.Oracle(flat, abi.SmsConfig.default())
    b = oracle_lib.Oracle(plain, abi.SmsConfig.default())
    ra, _, _ = a.path_samples(flt, 64, 5)
    rb, _, _ = b.path_samples(flt, 64, 5)
    assert (rb[:, 0] > 0).any() and np.array_equal(ra, rb)
    a.close()
    b.close()


@pytest.mark.parametrize("encoded", [False, True])
def test_normal_mapped_glass_dome_sms_matches_filtered_path_tracer(encoded):
    """A glass dome whose normal map varies over the sphere's (phi / 2 pi, theta / pi) uv
    (sphere.cpp:343-348): SMS solves with the map's normals and its uv derivatives in the
    geometric term, the filtered path tracer refracts through the same normals; the caustic
    terms agree, for frames from interpolated slopes (use_slopes) and from decoded interpolated
    RGB (the reference's default, normalmap.cpp:226-258).  A sphere whose uv were left constant
    would read one texel's normal in the path tracer but the map's derivative in the geometric
    term (SMS 26 % low at this amplitude)."""
    res, a, k = 256, 0.05, 8
    u = (np.arange(res) + 0.5) / res
    U, V = np.meshgrid(u, u, indexing="xy")
    sx, sy = a * np.sin(2 * np.pi * k * U), a * np.cos(2 * np.pi * k * V)
    n = np.stack([-sx, -sy, np.ones_like(sx)], -1)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    sc = scenes.sphere_dome(16, 16, light_radius=0.5, light_radiance=4.0)
    nm = sc.add_normalmap((0.5 * n + 0.5).astype(np.float32), encoded=encoded)
    sc.shapes[0].update(bsdf=abi.BSDF_NORMALMAP, k=0.0, normalmap=nm)
    orc = oracle_lib.Oracle(sc, abi.SmsConfig.default(max_trials=100000), threads=8)
    ms, ss = _mean_se(_caustic(orc, abi.SmsPathConfig.default(max_depth=6, caustics_only=1), 16, seed0=100))
    flt = abi.SmsPathConfig.default(integrator=abi.PATH_FILTERED_SS, max_depth=6, caustics_only=1)
    mf, sf = _mean_se(_caustic(orc, flt, 400, seed0=100))
    print(f"SMS {ms:.6g} +- {ss:.2g}  filtered {mf:.6g} +- {sf:.2g}")
    assert ms > 0 and abs(ms - mf) < 4 * math.hypot(ss, sf)
    orc.close()


def test_two_stage_on_an_encoded_map_is_unbiased():
    """Two-stage walks on a map without use_slopes start from the map's coarsest level, with the
    LEAN mean of the stored (undecoded) RGB (bsdfs/normalmap.cpp:292-303): only the seeds
    change, so the unbiased SMS caustic with and without two-stage walks agree on a scratched
    mirror."""
    sc = scenes.mirror_wall(nm_res=64, n_scratches=16, seed=4)
    sc.encoded_normalmaps.add(0)
    sc.shapes[2]["radius"] = 0.3
    sc.emitters[0]["radiance"] = (5.0, 5.0, 5.0)
    sc.set_camera((-1.2, 2.2, 0.0), (0.2, 0.0, 0.0), (0, 1, 0), 50.0, 16, 16)
    sms = abi.SmsPathConfig.default(max_depth=6, caustics_only=1)
    est = []
    for twostage in (0, 1):
        orc = oracle_lib.Oracle(sc, abi.SmsConfig.default(max_trials=1000000, twostage=twostage), threads=8)
        est.append(_mean_se(_caustic(orc, sms, 8, seed0=100 + 50 * twostage)))
        orc.close()
    print("single-stage, two-stage:", est)
    (m0, s0), (m1, s1) = est
    assert m0 > 0 and abs(m0 - m1) < 4 * math.hypot(s0, s1)


# ---------------------------------------------------------------- multi-scatter integrators
@pytest.fixture(scope="module")
def slab():
    """The double-refraction slab with gentle faces (two-vertex chains: floor -> bottom face,
    the caster next to the receiver -> top face, a bouncer -> light) under a larger light."""
    sc = scenes.slab(16, 16, n=16, amplitude=0.005, caster="bottom")
    sc.shapes[3]["radius"] = 0.4
    sc.emitters[0]["radiance"] = (30.0, 30.0, 30.0)
    return sc


def test_ms_sms_caustic_matches_filtered_path_tracer(slab):
    """path_sms_ms's caustic term (the multi-scatter estimator at the receiver) and
    path_filtered_ms's (BSDF-sampled chains through the slab that PathRecord accepts) agree; so
    do the whole images (emitter hits through specular camera paths, emitter sampling on the
    floor, the chain's terms)."""
    orc = oracle_lib.Oracle(slab, abi.SmsConfig.default(bounces=2, max_trials=100000), threads=8)
    res = {}
    for name, integ, only, passes in [("sms", abi.PATH_SMS_MS, 1, 8), ("filtered", abi.PATH_FILTERED_MS, 1, 32),
                                      ("sms_full", abi.PATH_SMS_MS, 0, 8), ("filtered_full", abi.PATH_FILTERED_MS, 0, 32)]:
        pc = abi.SmsPathConfig.default(integrator=integ, max_depth=8, caustics_only=only)
        res[name] = _mean_se(_caustic(orc, pc, passes, seed0=100 if integ == abi.PATH_SMS_MS else 1000))
    print({k: (round(m, 6), round(s, 6)) for k, (m, s) in res.items()})
    for a, b in (("sms", "filtered"), ("sms_full", "filtered_full")):
        (ma, sa), (mb, sb) = res[a], res[b]
        assert ma > 0 and abs(ma - mb) < 4 * math.hypot(sa, sb), (a, b)
    # the caustic is part of the whole image, which holds more (floor lit directly)
    assert res["filtered_full"][0] > res["filtered"][0] + 4 * res["filtered_full"][1]
    orc.close()


def test_ms_mnee_completion_is_unbiased(slab):
    """Multi-scatter MNEE (one deterministic walk per receiver, the half-vector constraints)
    plus path_sms_ms's completions (the BSDF-sampled chains the MNEE walk does not re-find vertex
    for vertex) agrees with the filtered path tracer; biased MNEE, which drops them, is lower."""
    cfg = abi.SmsConfig.default(bounces=2, mnee_init=1, biased=1, max_trials=1, halfvector_constraints=1)
    orc = oracle_lib.Oracle(slab, cfg, threads=8)
    res = {}
    for bm in (0, 1):
        pc = abi.SmsPathConfig.default(integrator=abi.PATH_SMS_MS, max_depth=8, caustics_only=1, biased_mnee=bm)
        res[bm] = _mean_se(_caustic(orc, pc, 24, seed0=300))
    flt = _mean_se(_caustic(orc, abi.SmsPathConfig.default(integrator=abi.PATH_FILTERED_MS, max_depth=8,
                                                           caustics_only=1), 48, seed0=2000))
    print(res, flt)
    (mu, su), (mb, sb), (mf, sf) = res[0], res[1], flt
    assert abs(mu - mf) < 4 * math.hypot(su, sf)
    assert mb < mu - 4 * math.hypot(sb, su)
    orc.close()


# ---------------------------------------------------------------- film
@pytest.fixture(scope="module")
def film_scene():
    sc = scenes.sphere_dome(24, 20)
    orc = oracle_lib.Oracle(sc, abi.SmsConfig.default(), threads=8)
    yield sc, orc
    orc.close()


def _positions(w, h, spp, seed):
    rng = np.random.default_rng(seed)
    px = np.repeat(np.arange(w * h), spp)
    return np.stack([(px % w) + rng.random(len(px)), (px // w) + rng.random(len(px))], -1).astype(np.float32)


def test_film_reproduces_constant_and_linear_images(film_scene):
    sc, orc = film_scene
    w, h, spp = sc.camera["width"], sc.camera["height"], 16
    pos = _positions(w, h, spp, 1)
    img, wt = orc.film_splat(np.ones((len(pos), 3), np.float32), pos, spp, 0.5)
    assert (wt > 0).all()
    np.testing.assert_allclose(img / wt[:, None], 1.0, rtol=2e-6)
    # a linear ramp f = x of the film position sampled on a 4 x 4 grid per pixel: the symmetric
    # filter reproduces it at the pixel centres of the interior, where every tap lies inside the
    # film (borders drop the outside taps)
    g = (np.arange(4) + 0.5) / 4
    off = np.stack(np.meshgrid(g, g, indexing="xy"), -1).reshape(-1, 2)
    px = np.arange(w * h)
    pos = (np.stack([px % w, px // w], -1)[:, None, :] + off[None]).reshape(-1, 2).astype(np.float32)
    val = np.repeat(pos[:, :1], 3, 1)
    img, wt = orc.film_splat(val, pos, spp, 0.5)
    dev = (img[:, 0] / wt).reshape(h, w)
    xs = np.arange(w) + 0.5
    inner = dev[2:-2, 2:-2] - xs[None, 2:-2]
    assert np.abs(inner).max() < 1e-4
    # at the border the outside taps are missing: the ramp's value is pulled inward
    assert dev[5, 0] > 0.5 + 1e-3 and dev[5, w - 1] < w - 0.5 - 1e-3


def test_film_filter_shape(film_scene):
    """One sample at a pixel centre: weights at the neighbours follow the truncated Gaussian of
    stddev 0.5 (radius 2) through its 31-step table; the box filter puts the sample in its pixel."""
    sc, orc = film_scene
    w, h = sc.camera["width"], sc.camera["height"]
    spp = 1
    rgb = np.zeros((w * h, 3), np.float32)
    pos = _positions(w, h, spp, 2)
    c = 10 * w + 12
    rgb[c] = 1.0
    pos[c] = (12.5, 10.5)
    img, wt = orc.film_splat(rgb, pos, spp, 0.5)
    g = img[:, 0].reshape(h, w)
    g0 = g[10, 12]
    # separable and symmetric; ratio of a neighbour to the centre ~ exp(-x^2 / (2 s^2)) at x in [30/31, 1]
    assert g[10, 11] == g[10, 13] == g[9, 12] == g[11, 12]
    r1 = g[10, 11] / g0
    lo, hi = math.exp(-2.0) - math.exp(-8.0), math.exp(-2.0 * (30 / 31) ** 2)
    assert lo * 0.999 <= r1 <= hi * 1.001, r1
    assert abs(g[9, 11] / g0 - r1 * r1) < 1e-6
    assert g[10, 14] == 0 and g[10, 10] == 0          # the tap at distance 2 carries weight 0
    img, wt = orc.film_splat(rgb, pos, spp, 0.0)
    gb = img[:, 0].reshape(h, w)
    assert gb[10, 12] == 1.0 and gb.sum() == 1.0
    # invalid samples are dropped whole (value and weight)
    rgb2 = np.ones((w * h, 3), np.float32)
    rgb2[c] = (np.nan, 1, 1)
    img2, wt2 = orc.film_splat(rgb2, pos, spp, 0.5)
    img1, wt1 = orc.film_splat(np.ones((w * h, 3), np.float32), pos, spp, 0.5)
    assert wt2[c] < wt1[c] and np.isfinite(img2).all()


# ------------------------------------------------------------------ area lights on rectangles and meshes
def _polygon_irradiance(x, n, verts, L):
    """Lambert's formula: irradiance at x (normal n) from a uniform polygon of radiance L,
    E = L / 2 * |sum_i theta_i (n . u_i)|, theta_i the angle subtended by edge i and u_i the
    unit normal of the plane through x and the edge."""
    v = [np.asarray(p, np.float64) - x for p in verts]
    s = 0.0
    for i in range(len(v)):
        a, b = v[i], v[(i + 1) % len(v)]
        c = np.cross(a, b)
        th = math.atan2(np.linalg.norm(c), np.dot(a, b))
        s += th * np.dot(n, c / np.linalg.norm(c))
    return L * 0.5 * abs(s)


def _area_light_scene(light, x0, z0):
    """Diffuse floor (y = 0, albedo 0.5) under a light facing down 2 above it: a rectangle of
    half side 0.5 centred at (0.3, 2, -0.2), or the same square as a two-triangle mesh; a
    narrow camera looking straight down at (x0, 0, z0)."""
    sc = scenes.Scene()
    sc.add_rectangle(scenes.plane_to_world((0, 0, 0), 4.0), abi.BSDF_DIFFUSE, reflectance=(0.5, 0.5, 0.5))
    c, a = np.array([0.3, 2.0, -0.2]), 0.5
    corners = [c + [-a, 0, -a], c + [a, 0, -a], c + [a, 0, a], c + [-a, 0, a]]
    if light == "rectangle":
        to_world = np.array([[a, 0, 0, c[0]], [0, 0, -1, c[1]], [0, a, 0, c[2]]], np.float32)
        s = sc.add_rectangle(to_world, abi.BSDF_DIFFUSE, reflectance=(0, 0, 0))
    else:
        P = np.array(corners, np.float32)
        F = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)          # wound so the face normal points down
        s = sc.add_mesh(P, F, bsdf=abi.BSDF_DIFFUSE, reflectance=(0, 0, 0))
    sc.add_area_light(s, (10.0, 10.0, 10.0))
    sc.set_camera((x0, 1.0, z0 + 1e-4), (x0, 0.0, z0), (0, 0, -1), 0.05, 4, 4)
    return sc, corners


@pytest.mark.parametrize("light", ["rectangle", "mesh"])
@pytest.mark.parametrize("x0,z0", [(0.3, -0.2), (-0.9, 0.6)])
def test_area_light_direct_lighting_matches_polygon_formula(light, x0, z0):
    """Direct lighting of a diffuse floor by a rectangle or mesh area light (emitter sampling
    and BSDF sampling combined by MIS) equals albedo / pi x Lambert's polygon irradiance."""
    sc, corners = _area_light_scene(light, x0, z0)
    orc = oracle_lib.Oracle(sc, abi.SmsConfig.default())
    pc = abi.SmsPathConfig.default(integrator=abi.PATH_FILTERED_SS, max_depth=2)
    vals = []
    for s in range(6):
        rgb, _, _ = orc.path_samples(pc, 4096, 1 + s)
        vals.append(rgb[:, 0])
    m, se = _mean_se(np.concatenate(vals))
    ref = 0.5 / math.pi * _polygon_irradiance(np.array([x0, 0.0, z0]), np.array([0, 1.0, 0]), corners, 10.0)
    print(f"{light} at ({x0}, {z0}): {m:.5f} +- {se:.5f}, closed form {ref:.5f}")
    assert abs(m - ref) < 4 * se + 1e-3 * ref
    orc.close()
