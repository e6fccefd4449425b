"""Pins of the oracle's glint solver (a16-a18): every converged walk must sit on a root of
"bilinearly interpolated level-0 slope = target slope" found per texel cell in FP64 from the
decoded RGB8 map (normalmap.cpp:46-66, normalmap.h:62-104); walks on a map whose slopes vary
linearly across a cell converge to the closed-form point."""
import ctypes
import math

import numpy as np
import pytest

import oracle_lib
from smsamd import abi, scenes, workloads


C_FLOAT, C_VP = ctypes.c_float, ctypes.c_void_p


def decode_slopes(rgb):
    n = rgb.astype(np.float64) / 255.0 * 2.0 - 1.0
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    return np.stack([-n[..., 0] / n[..., 2], -n[..., 1] / n[..., 2]], -1)


def target_slope(w):
    def nz(v):
        return v / np.linalg.norm(v)
    p = w["si_p"].astype(np.float64)
    s, t, n = (w[k].astype(np.float64) for k in ("frame_s", "frame_t", "frame_n"))
    a, b = nz(w["sensor_p"] - p), nz(w["ei_p"] - p)
    wi = np.array([a @ s, a @ t, a @ n])
    wo = np.array([b @ s, b @ t, b @ n])
    h = nz(wi + wo)
    return np.array([-h[0] / h[2], -h[1] / h[2]])


def cell_roots(S, T, res, uv_lo, uv_hi):
    """Roots of the bilinear slope interpolant = T in every texel cell touching the box."""
    roots = []
    i0 = max(0, int(math.floor(uv_lo[0] * res - 0.5)) - 1)
    i1 = min(res - 2, int(math.floor(uv_hi[0] * res - 0.5)) + 1)
    j0 = max(0, int(math.floor(uv_lo[1] * res - 0.5)) - 1)
    j1 = min(res - 2, int(math.floor(uv_hi[1] * res - 0.5)) + 1)
    for j in range(j0, j1 + 1):
        for i in range(i0, i1 + 1):
            v00, v10, v01, v11 = S[j, i], S[j, i + 1], S[j + 1, i], S[j + 1, i + 1]
            A, B, D, E = v10 - v00, v01 - v00, v00 - v10 - v01 + v11, v00 - T
            # (E + b B + a (A + b D)) = 0 for both components -> quadratic in b
            c2 = B[0] * D[1] - B[1] * D[0]
            c1 = E[0] * D[1] - E[1] * D[0] + B[0] * A[1] - B[1] * A[0]
            c0 = E[0] * A[1] - E[1] * A[0]
            bs = [-c0 / c1] if abs(c2) < 1e-14 and abs(c1) > 1e-14 else (
                [] if abs(c2) < 1e-14 else [r.real for r in np.roots([c2, c1, c0]) if abs(r.imag) < 1e-9])
            for b in bs:
                den = A + b * D
                k = 0 if abs(den[0]) > abs(den[1]) else 1
                if abs(den[k]) < 1e-14:
                    continue
                a = -(E[k] + b * B[k]) / den[k]
                if -1e-9 <= a <= 1 + 1e-9 and -1e-9 <= b <= 1 + 1e-9:
                    roots.append(np.array([(i + a + 0.5) / res, (j + b + 0.5) / res]))
    return roots


@pytest.fixture(scope="module")
def glint_scene():
    sc = scenes.glint_plane(64, 64, nm_res=128, n_scratches=24, seed=7)
    o = oracle_lib.Oracle(sc, abi.SmsConfig.default(max_trials=100000))
    yield sc, o
    o.close()


def test_glint_walks_land_on_cell_roots(glint_scene):
    sc, o = glint_scene
    res, rgb = sc.normalmaps[0]
    S = decode_slopes(rgb)
    w = workloads.glint_walk_inputs(sc, 6000, 3, footprint_texels=5.0)
    out = o.glint_walks(w)
    ok = np.nonzero(out["success"] == 1)[0]
    assert len(ok) > 60
    for i in ok[:400]:
        T = target_slope(w[i])
        uv = out["uv"][i].astype(np.float64)
        roots = cell_roots(S, T, res, uv - 2.0 / res, uv + 2.0 / res)
        assert roots, (i, uv)
        d = min(np.linalg.norm(uv - r) for r in roots)
        assert d < 1e-4, (i, uv, roots)


def test_glint_linear_cell_closed_form():
    """Slopes affine in uv through the target slope at the plane centre: the bilinear
    interpolant is exact, so Newton converges in a couple of iterations to the cell root."""
    res = 8
    probe = workloads.glint_walk_inputs(scenes.glint_plane(16, 16, nm_res=8, n_scratches=1), 1, 0, nm_res=res)
    probe["uv"] = (0.5, 0.5)
    probe["si_p"] = (0.0, 0.0, 0.0)
    T0 = target_slope(probe[0])
    c = (np.arange(res) + 0.5) / res - 0.5
    sl = np.zeros((res, res, 2))
    sl[..., 0] = T0[0] + 0.9 * c[None, :]
    sl[..., 1] = T0[1] + 0.7 * c[:, None]
    n = np.stack([-sl[..., 0], -sl[..., 1], np.ones((res, res))], -1)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    rgb = np.clip(np.round((n * 0.5 + 0.5) * 255), 0, 255).astype(np.uint8)
    S = decode_slopes(rgb)
    sc = scenes.glint_plane(16, 16, normalmap=rgb)
    o = oracle_lib.Oracle(sc, abi.SmsConfig.default())
    w = workloads.glint_walk_inputs(sc, 500, 8, footprint_texels=3.0, nm_res=res)
    rng = np.random.default_rng(1)
    w["uv"] = 0.5 + 0.02 * (rng.random((500, 2)) - 0.5)
    w["si_p"][:, 0] = w["uv"][:, 0] - 0.5
    w["si_p"][:, 2] = 0.5 - w["uv"][:, 1]
    out = o.glint_walks(w)
    ok = np.nonzero(out["success"] == 1)[0]
    assert len(ok) > 100
    for i in ok:
        T = target_slope(w[i])
        uv = out["uv"][i].astype(np.float64)
        roots = cell_roots(S, T, res, uv - 1.0 / res, uv + 1.0 / res)
        assert min(np.linalg.norm(uv - r) for r in roots) < 1e-4
        assert out["iterations"][i] <= 6
    o.close()


def _bilinear(S, uv, res):
    """Texel-centre bilinear interpolation of the decoded slopes (interior uv)."""
    x, y = uv[0] * res - 0.5, uv[1] * res - 0.5
    i, j = int(math.floor(x)), int(math.floor(y))
    fx, fy = x - i, y - j
    return (S[j, i] * (1 - fx) + S[j, i + 1] * fx) * (1 - fy) + (S[j + 1, i] * (1 - fx) + S[j + 1, i + 1] * fx) * fy


def test_glint_geometric_term_matches_finite_differences():
    """The glint path's generalized geometric term for a point light (manifold_glints.cpp:411-554)
    is |cos| / r0^2 x |det dx1/dx2|: how the specular vertex x1 -- moving on the plane, its normal
    the normal map's at its uv -- follows the light x2 over the light's tangent plane.  An
    independent FP64 solve of "half-vector parallel to the mapped normal", differentiated by
    central differences over light moves, must give the same value.  Both sides use orthonormal
    bases, so the determinant does not depend on which ones."""
    res = 16
    g = (np.arange(res) + 0.5) / res
    U, V = np.meshgrid(g, g, indexing="xy")
    sl = np.stack([0.12 * np.sin(2 * np.pi * U) + 0.05 * V, 0.10 * np.cos(2 * np.pi * V) - 0.04 * U], -1)
    n = np.concatenate([-sl, np.ones((res, res, 1))], -1)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    rgb = np.clip(np.round((n * 0.5 + 0.5) * 255), 0, 255).astype(np.uint8)
    S = decode_slopes(rgb)
    sc = scenes.glint_plane(16, 16, normalmap=rgb)
    orc = oracle_lib.Oracle(sc, abi.SmsConfig.default())
    geo = orc.lib.oracle_glint_geometric
    geo.restype, geo.argtypes = C_FLOAT, [C_VP] * 5
    q0 = orc.glint_queries(1, 3)
    q0 = q0[q0["shape"] >= 0][:1].copy()
    f64 = lambda k: q0[k][0].astype(np.float64)
    p_ref, uv_ref = f64("p"), f64("uv")
    dpu, dpv, bn, gn = f64("dp_du"), f64("dp_dv"), f64("sh_n"), f64("n")
    nz = lambda v: v / np.linalg.norm(v)
    bs = nz(dpu - bn * (bn @ dpu))
    bt = np.cross(bn, bs)

    def x1_of(uv):
        return p_ref + dpu * (uv[0] - uv_ref[0]) + dpv * (uv[1] - uv_ref[1])

    def normal(uv):
        s = _bilinear(S, uv, res)
        return nz(-s[0] * bs - s[1] * bt + bn)

    def residual(uv, x0, x2):
        nn = normal(uv)
        e1 = nz(bs - nn * (nn @ bs))
        e2 = np.cross(nn, e1)
        x1 = x1_of(uv)
        h = nz(nz(x0 - x1) + nz(x2 - x1))
        return np.array([h @ e1, h @ e2])

    def solve(uv, x0, x2):
        """Newton in uv, kept inside the bilinear cell it starts in (None if it leaves)."""
        uv = np.array(uv, np.float64)
        cell = np.floor(uv * res - 0.5)
        for _ in range(60):
            if np.any(np.floor(uv * res - 0.5) != cell):
                return None
            J = np.zeros((2, 2))
            for k in range(2):
                d = np.zeros(2)
                d[k] = 1e-7
                J[:, k] = (residual(uv + d, x0, x2) - residual(uv - d, x0, x2)) / 2e-7
            step = np.linalg.solve(J, residual(uv, x0, x2))
            uv = uv - step
            if np.abs(step).max() < 1e-15:
                break
        return uv if np.abs(residual(uv, x0, x2)).max() < 1e-13 else None

    a1 = nz(dpu)                                                  # orthonormal basis of the plane
    a2 = nz(dpv - a1 * (a1 @ dpv))
    rng = np.random.default_rng(5)
    errs = []
    for trial in range(40):
        cell = rng.integers(2, res - 3, 2)
        uv = (cell + 0.5 + rng.uniform(0.1, 0.9, 2)) / res         # inside one bilinear cell
        x1, nn = x1_of(uv), normal(uv)
        w0 = nz(np.array([rng.uniform(-0.5, 0.5), 1.0, rng.uniform(0.2, 0.8)]))
        w2 = 2 * (w0 @ nn) * nn - w0
        if w2 @ gn < 0.2:
            continue
        x0, x2 = x1 + rng.uniform(150, 300) * w0, x1 + rng.uniform(150, 300) * w2
        d = nz(x1 - x2)
        b1 = nz(np.cross(d, [0.0, 0.0, 1.0]))
        b2 = np.cross(d, b1)
        eps = 0.5
        moved = [solve(uv, x0, x2 + sg * eps * b) for b in (b1, b2) for sg in (1, -1)]
        if any(m is None for m in moved):                          # near a fold of the slope map
            continue
        J = np.array([[(x1_of(moved[2 * k]) - x1_of(moved[2 * k + 1])) @ a / (2 * eps) for k in range(2)]
                      for a in (a1, a2)])
        d0 = x0 - x1
        expect = abs(nz(d0) @ gn) / (d0 @ d0) * abs(np.linalg.det(J))
        if not expect < 1e-5:                                      # the reference clamps G at 1e-5
            continue
        qq = q0.copy()
        qq["p"], qq["uv"] = x1, uv
        uv32, x032, x232 = (np.ascontiguousarray(v, np.float32) for v in (uv, x0, x2))
        G = geo(orc.h, qq.ctypes.data, uv32.ctypes.data, x032.ctypes.data, x232.ctypes.data)
        # FP32 oracle against FP64 differences: about 1e-5 typically, a few 1e-3 near folds
        assert G == pytest.approx(expect, rel=5e-3), (trial, uv, G, expect)
        errs.append(abs(G / expect - 1))
    assert len(errs) >= 15
    assert np.median(errs) < 1e-4
    orc.close()


def test_glint_biased_and_unbiased_estimators_agree():
    """The biased glint estimator (manifold_glints.cpp:112-140) sums the contributions of the
    distinct solutions a trial set of M walks finds; with M large enough to find every solution
    of a footprint its mean equals the unbiased Bernoulli estimator's (:70-111)."""
    sc = scenes.glint_plane(32, 32, nm_res=256, n_scratches=64, seed=3)
    ob = oracle_lib.Oracle(sc, abi.SmsConfig.default(biased=1, max_trials=128))
    q = ob.glint_queries(4, 5)
    rb, sb = ob.glint_estimate(q, 1)
    ou = oracle_lib.Oracle(sc, abi.SmsConfig.default(max_trials=100000))
    runs = np.stack([ou.glint_estimate(q, s)[0][:, 0] for s in range(1, 9)])
    mean_u, mean_b = runs.mean(), rb[:, 0].astype(np.float64).mean()
    se = runs.mean(0).std() / math.sqrt(runs.shape[1])
    print(f"unbiased {mean_u:.6g} biased {mean_b:.6g} se {se:.3g} walks {sb['walks']} converged {sb['walks_converged']}")
    assert (rb[:, 0] > 0).sum() > 50
    assert sb["walks"] > 128 * 50 and sb["bernoulli_calls"] == 0
    assert abs(mean_u - mean_b) < 4 * se + 1e-3 * abs(mean_u)
    ob.close()
    ou.close()


def _smooth_map(res=16):
    g = (np.arange(res) + 0.5) / res
    U, V = np.meshgrid(g, g, indexing="xy")
    sl = np.stack([0.12 * np.sin(2 * np.pi * U) + 0.05 * V, 0.10 * np.cos(2 * np.pi * V) - 0.04 * U], -1)
    n = np.concatenate([-sl, np.ones((res, res, 1))], -1)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    return np.clip(np.round((n * 0.5 + 0.5) * 255), 0, 255).astype(np.uint8)


def _bilinear_clamped(S, uv, res):
    """Normalmap::eval_normal's texel-centre bilinear with edge clamp (normalmap.h:67-80)."""
    x = min(max(uv[0] * res - 0.5, 0.0), res - 1.0)
    y = min(max(uv[1] * res - 0.5, 0.0), res - 1.0)
    i0, j0 = int(math.floor(x)), int(math.floor(y))
    i1, j1 = min(i0 + 1, res - 1), min(j0 + 1, res - 1)
    fx, fy = x - i0, y - j0
    return (S[j0, i0] * (1 - fx) + S[j0, i1] * fx) * (1 - fy) + (S[j1, i0] * (1 - fx) + S[j1, i1] * fx) * fy


AREA_LIGHT = ((0.0, 1.1, -1.05), 0.35, (4.0, 3.0, 2.0))


def test_glint_bsdf_strategy_geometry():
    """BSDF-sampling strategy of the glint estimator (manifold_glints.cpp:183-220) on a mirror
    plane with a smooth normal map under a point light and a spherical area light.  An
    independent FP64 restatement of its geometry -- the footprint offset, the level-0 bilinear
    normal there, the mirror direction and its side tests, the ray against the light sphere --
    must agree with the oracle on every query's direction and on whether it reaches the light;
    where it does, the emitter density is the sphere's solid-angle cone density over the two
    emitters, and the value is the power heuristic of the two densities times F (1 for this
    conductor) times the radiance.  The full estimate is the SMS strategy plus this term."""
    rgb = _smooth_map()
    res = rgb.shape[0]
    S = decode_slopes(rgb)
    sc = scenes.glint_plane(48, 48, normalmap=rgb, area_light=AREA_LIGHT)
    orc = oracle_lib.Oracle(sc, abi.SmsConfig.default(max_trials=100000))
    lib = orc.lib
    lib.oracle_glint_bsdf_parts.argtypes = [C_VP, C_VP, ctypes.c_uint64, ctypes.c_uint64, C_VP, C_VP]
    q = orc.glint_queries(2, 7)
    centre, radius, L = np.array(AREA_LIGHT[0]), AREA_LIGHT[1], np.array(AREA_LIGHT[2])
    nz = lambda v: v / np.linalg.norm(v)
    seed, hits, checked = 3, 0, 0
    values = np.zeros((len(q), 3), np.float32)
    for i in range(len(q)):
        if q["shape"][i] < 0:
            continue
        out, parts = np.zeros(3, np.float32), np.zeros(10, np.float32)
        lib.oracle_glint_bsdf_parts(orc.h, q[i:i + 1].ctypes.data, seed, i, out.ctypes.data, parts.ctypes.data)
        values[i] = out
        f64 = lambda k: q[k][i].astype(np.float64)
        uv = f64("uv") + (parts[0] - 0.5) * f64("duv_dx") + (parts[1] - 0.5) * f64("duv_dy")
        uv = uv - np.floor(uv)
        s_, t_, n_ = f64("sh_s"), f64("sh_t"), f64("sh_n")
        bn, dpu = n_, f64("dp_du")
        bs = nz(dpu - bn * (bn @ dpu))
        bt = np.cross(bn, bs)
        sl = _bilinear_clamped(S, uv, res)
        m = nz(-sl[0] * bs - sl[1] * bt + bn)
        wi = f64("wi")
        wi_w = s_ * wi[0] + t_ * wi[1] + n_ * wi[2]
        if abs(wi_w @ m) < 1e-4:
            continue
        d = 2 * (wi_w @ m) * m - wi_w
        reflected = wi_w @ m > 0 and d @ n_ > 0
        if not reflected:
            assert parts[5] == -1 and not out.any(), i
            continue
        checked += 1
        assert np.allclose(parts[2:5], d, atol=2e-5), (i, parts[2:5], d)
        p = f64("p")
        oc = p - centre
        b = oc @ d
        disc = b * b - (oc @ oc - radius * radius)
        if abs(disc) < 1e-4:                                   # grazing: FP32 may decide either way
            continue
        hit = disc > 0 and -b - math.sqrt(disc) > 0
        assert (parts[5] == 1) == hit, (i, parts[5], disc)
        if not hit:
            assert not out.any()
            continue
        hits += 1
        sin_a = radius / np.linalg.norm(centre - p)
        p_sms = 1.0 / (2 * math.pi * (1 - math.sqrt(1 - sin_a * sin_a))) / 2.0
        assert parts[8] == pytest.approx(p_sms, rel=1e-4)
        assert parts[7] >= 1.0 and parts[6] == pytest.approx(1.0, abs=1e-6)
        pb, ps = float(parts[7]), p_sms
        expect = pb * pb / (pb * pb + ps * ps) * L
        assert np.allclose(out, expect, rtol=1e-4), (i, out, expect)
    print(f"queries {int((q['shape'] >= 0).sum())} reflected {checked} reaching the light {hits}")
    assert checked > 500 and hits > 100
    # the estimator adds the BSDF strategy to the SMS strategy's result
    full, _ = orc.glint_estimate(q, seed)
    orc.set_config(abi.SmsConfig.default(max_trials=100000, sms_strategy_only=1))
    sms_only, _ = orc.glint_estimate(q, seed)
    zero_bsdf = np.zeros_like(values)
    for i in np.nonzero(values.any(1))[0][:50]:
        out, parts = np.zeros(3, np.float32), np.zeros(10, np.float32)
        lib.oracle_glint_bsdf_parts(orc.h, q[i:i + 1].ctypes.data, seed, int(i), out.ctypes.data, parts.ctypes.data)
        zero_bsdf[i] = out                                     # mis = 0 under sms_strategy_only
    assert not zero_bsdf.any()
    assert np.array_equal(full, sms_only + values)
    orc.close()


def test_glint_mixture_estimator_is_unbiased():
    """The learned proposal on glint footprints: seeds drawn from a defensive mixture over the
    footprint (concentrated away from the centre) instead of the uniform footprint sample; the
    Bernoulli estimator's mean is unchanged within standard errors (the same mixture drives the
    first walk and the trials), the walks converge at another rate."""
    sc = scenes.glint_plane(32, 32, nm_res=256, n_scratches=64, seed=3)
    cfg = abi.SmsConfig.default(max_trials=100000)
    orc = oracle_lib.Oracle(sc, cfg)
    q = orc.glint_queries(2, 5)
    q = np.ascontiguousarray(q[q["shape"] >= 0][:400])
    rng = np.random.default_rng(4)
    raw = np.zeros((len(q), 48), np.float32)
    for c in range(8):
        raw[:, 6 * c] = rng.normal(0, 1, len(q))
        raw[:, 6 * c + 1] = rng.normal(0.8, 0.5, len(q))
        raw[:, 6 * c + 2] = rng.normal(-0.5, 0.5, len(q))
        raw[:, 6 * c + 3] = raw[:, 6 * c + 4] = np.log(0.15)
        raw[:, 6 * c + 5] = rng.normal(0, 0.5, len(q))
    reps = 500
    qq, rr = np.concatenate([q] * reps), np.concatenate([raw] * reps)
    uni, su = orc.glint_estimate(qq, 3)
    mix, sm = orc.glint_estimate_prop(qq, 3, rr, 0.2)
    lu, lm = uni[:, 0].astype(np.float64), mix[:, 0].astype(np.float64)
    # paired by query: the footprints' own spread cancels
    du = lu.reshape(reps, -1).mean(0)
    dm = lm.reshape(reps, -1).mean(0)
    se = math.sqrt(lu.var() / len(lu) + lm.var() / len(lm))
    print(f"uniform {lu.mean():.6g} mixture {lm.mean():.6g} se {se:.3g} conv {su['walks_converged'] / su['walks']:.4f} "
          f"{sm['walks_converged'] / sm['walks']:.4f}")
    assert (du > 0).sum() > 20
    assert abs(lu.mean() - lm.mean()) < 4 * se
    assert sm["walks"] != su["walks"]
