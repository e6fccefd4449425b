"""Pins of the oracle's single- and multi-scatter terms that its estimators alone cannot check,
each by mathematics that does not reuse the oracle's code, and each shown to fail on a mistake
planted in the oracle (``oracle_plant``: a factor of 1.1 on one term, on the CPU):

* SS geometric term (manifold_ss.cpp:695-857): FP64 central differences of the refraction
  solution on the glass dome over moves of the light point, |cos| / r^2 x |det dx1/dx2|;
* SS contribution (manifold_ss.cpp:281-327,638-693 with the estimator :28-201): the unbiased
  estimate at a floor point inside the dome equals the FP64 Monte-Carlo irradiance of the light
  seen through the glass, (1 - F) eta^2 L cos over the directions that refract to the light;
* SS Newton steps (manifold_ss.cpp:407-636): J^-1 C of the angle-difference and the half-vector
  constraint against FP64 finite-difference Jacobians of independent numpy restatements;
* MS Newton steps (manifold_ms.cpp:500-857, :1192-1232): the same for both constraints of a
  two-vertex chain through a glass ball, 4 x 4 finite-difference Jacobians;
* MNEE initialisation (manifold_ss.cpp:228-241, manifold_ms.cpp:311-318): the seed vertex is
  the closed-form hit of the straight ray toward the light (dielectric) or toward the caster's
  bounding-box centre (conductor), independent of the random sample.
"""
import contextlib
import ctypes as C
import math

import numpy as np
import pytest

import oracle_lib
from smsamd import abi, scenes, workloads
from test_oracle_ss import _dome_roots, _refract

ETA = 1.5
PLANT = {"ss_g": 0, "ss_transmit": 1, "ss_ad_jac": 2, "ss_hv_jac": 3, "ms_ad_jac": 4, "ms_hv_jac": 5,
         "mnee_target": 6, "ms_g": 7, "ss_g_dir": 8}


def _lib():
    lib = oracle_lib.load()
    lib.oracle_plant.argtypes = [C.c_int, C.c_float]
    lib.oracle_ss_contribution.argtypes = [C.c_void_p] * 5
    lib.oracle_ms_step_at.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_void_p]
    return lib


@contextlib.contextmanager
def planted(site, factor=1.1):
    """Plant factor x (one term of the oracle) for the duration of the block."""
    lib = _lib()
    assert lib.oracle_plant(PLANT[site], factor) == 0
    try:
        yield
    finally:
        lib.oracle_plant(PLANT[site], 1.0)


def _nz(v):
    return v / np.linalg.norm(v)


def _ei(p, n, kind=abi.EI_AREA, emitter=0):
    e = np.zeros(1, abi.EI_DTYPE)
    e["p"], e["n"], e["kind"], e["emitter"] = p, n, kind, emitter
    e["weight"], e["pdf"] = 1.0, 1.0
    return e


# ==================================================================== dome: FP64 refraction roots
def _dome_f(th, x, y, e1, e2, nrm):
    p = math.cos(th) * e1 + math.sin(th) * e2
    wi = _nz(x - p)
    wo = _nz(y - p)
    wt = _refract(wi, p, ETA)
    if wt is None:
        return None
    return float(np.dot(np.cross(wt, wo), nrm))


def _dome_root_near(x, y, p0, delta=0.05):
    """The refraction solution on the unit dome (centre 0) for receiver x and light point y
    nearest to p0: a bracketed FP64 root in the plane of (x, y, 0)."""
    e1 = _nz(x)
    t = y - np.dot(y, e1) * e1
    e2 = _nz(t)
    nrm = np.cross(e1, e2)
    th0 = math.atan2(np.dot(p0, e2), np.dot(p0, e1))
    ths = np.linspace(th0 - delta, th0 + delta, 201)
    vals = [_dome_f(a, x, y, e1, e2, nrm) for a in ths]
    best = None
    for i in range(len(ths) - 1):
        a, b = vals[i], vals[i + 1]
        if a is None or b is None or a * b > 0:
            continue
        if best is None or abs(0.5 * (ths[i] + ths[i + 1]) - th0) < abs(0.5 * (best[0] + best[1]) - th0):
            best = (ths[i], ths[i + 1], a)
    if best is None:
        return None
    lo, hi, flo = best
    for _ in range(100):
        mid = 0.5 * (lo + hi)
        fm = _dome_f(mid, x, y, e1, e2, nrm)
        if flo * fm <= 0:
            hi = mid
        else:
            lo, flo = mid, fm
    th = 0.5 * (lo + hi)
    return math.cos(th) * e1 + math.sin(th) * e2


def _tangent_basis(n):
    n = _nz(n)
    a = np.array([1.0, 0.0, 0.0]) if abs(n[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    b1 = _nz(np.cross(n, a))
    return b1, np.cross(n, b1)


def _ss_g_fd(x, y, light_normal, p1, eps=1e-4):
    """|cos(x0 - x1, n1)| / |x0 - x1|^2 x min(1, |det dx1/dx2|), x2 moving in the plane normal to
    light_normal, by FP64 central differences of the dome's refraction solution."""
    b1, b2 = _tangent_basis(light_normal)
    moved = []
    for b in (b1, b2):
        for sg in (1, -1):
            r = _dome_root_near(x, y + sg * eps * b, p1)
            if r is None:
                return None
            moved.append(r)
    a1, a2 = _tangent_basis(p1)
    J = np.array([[(moved[2 * k] - moved[2 * k + 1]) @ a / (2 * eps) for k in range(2)] for a in (a1, a2)])
    d0 = x - p1
    return abs(_nz(d0) @ _nz(p1)) / (d0 @ d0) * min(1.0, abs(np.linalg.det(J)))


def _dome_point_light(pos=(0.3, 4.0, -0.2)):
    sc = scenes.Scene()
    sc.add_sphere((0.0, 0.0, 0.0), 1.0, abi.BSDF_DIELECTRIC, flags=abi.FLAG_CAUSTIC_CASTER_SS, eta=ETA)
    sc.add_rectangle(scenes.plane_to_world((0, 0, 0), 10.0), abi.BSDF_DIFFUSE, flags=abi.FLAG_CAUSTIC_RECEIVER)
    sc.add_point_light(pos, (5.0, 5.0, 5.0))
    sc.set_camera((0, 3, 6), (0, 0.5, 0), (0, 1, 0), 45.0, 64, 64)
    return sc


def _ss_g_errors(point_light, n_cases=24, seed=4):
    sc = _dome_point_light() if point_light else scenes.sphere_dome(64, 64)
    orc = oracle_lib.Oracle(sc, abi.SmsConfig.default())
    lib = _lib()
    rng = np.random.default_rng(seed)
    errs = []
    out = np.zeros(9, np.float32)
    for _ in range(n_cases):
        x = np.array([rng.uniform(-0.6, 0.6), 0.0, rng.uniform(-0.6, 0.6)])
        if point_light:
            y = np.asarray(sc.emitters[0]["position"], np.float64)
            ln = None
        else:
            c = np.asarray(sc.shapes[2]["center"], np.float64)
            ln = _nz(np.array([0.0, -1.0, 0.0]) + 0.6 * rng.uniform(-1, 1, 3))
            y = c + sc.shapes[2]["radius"] * ln
        roots = [p for p in _dome_roots(x, y) if p[1] > 0.05]
        if not roots:
            continue
        p1 = roots[0]
        q = workloads.floor_queries_at(x[None], sc.camera["origin"], shape=1)
        e = _ei(y, ln if ln is not None else _nz(p1 - y), abi.EI_POINT if point_light else abi.EI_AREA)
        p132 = p1.astype(np.float32)
        if lib.oracle_ss_contribution(orc.h, q.ctypes.data, e.ctypes.data, p132.ctypes.data, out.ctypes.data) != 1:
            continue
        v = out[6:9].astype(np.float64)
        r = _dome_root_near(x, y, v)
        if r is None or np.linalg.norm(r - v) > 1e-4:
            continue
        # the point light's vertex moves in the plane normal to the direction light -> vertex
        expect = _ss_g_fd(x, y, (r - y) if point_light else ln, r)
        if expect is None:
            continue
        errs.append(abs(out[0] / expect - 1.0))
    orc.close()
    return np.array(errs)


@pytest.mark.parametrize("point_light", [False, True])
def test_ss_geometric_term_matches_light_move_differences(point_light):
    errs = _ss_g_errors(point_light)
    print(f"SS G vs FD: {len(errs)} cases, median rel err {np.median(errs):.2e}, max {errs.max():.2e}")
    assert len(errs) >= 12
    assert np.median(errs) < 1e-4
    assert errs.max() < 2e-3


def test_ss_geometric_term_pin_detects_planted_error():
    with planted("ss_g"):
        errs = _ss_g_errors(False, n_cases=12)
    assert len(errs) >= 6 and np.median(errs) > 0.05


# ==================================================================== dome: brute-force irradiance
def _fresnel(cos_i, eta):
    """Unpolarized dielectric Fresnel reflectance, incident from the eta=1 side (FP64)."""
    sin_t2 = (1.0 - cos_i * cos_i) / (eta * eta)
    cos_t = np.sqrt(np.maximum(0.0, 1.0 - sin_t2))
    rs = (cos_i - eta * cos_t) / (cos_i + eta * cos_t)
    rp = (eta * cos_i - cos_t) / (eta * cos_i + cos_t)
    return 0.5 * (rs * rs + rp * rp)


def _dome_irradiance(x, light_c, light_r, radiance, n=1_500_000, half=0.06, seed=0):
    """Irradiance at floor point x inside the unit glass dome from the sphere light seen through
    the glass: integral over directions of radiance (1 - F) eta^2 cos(theta), FP64 Monte Carlo
    in a cone about each refraction image of the light."""
    rng = np.random.default_rng(seed)
    axes = [_nz(p - x) for p in _dome_roots(x, light_c) if p[1] > 0]
    axes = [a for a in axes if a[1] > 0]
    assert axes
    for i in range(len(axes)):
        for j in range(i):
            assert math.acos(min(1.0, axes[i] @ axes[j])) > 2 * half
    total, var, max_ang = 0.0, 0.0, 0.0
    cos_max = math.cos(half)
    omega = 2 * np.pi * (1.0 - cos_max)
    for a in axes:
        z = 1.0 - rng.random(n) * (1.0 - cos_max)
        phi = 2 * np.pi * rng.random(n)
        s = np.sqrt(np.maximum(0.0, 1.0 - z * z))
        t1, t2 = _tangent_basis(a)
        d = s[:, None] * (np.cos(phi)[:, None] * t1 + np.sin(phi)[:, None] * t2) + z[:, None] * a
        b = d @ x
        t = -b + np.sqrt(b * b - (x @ x - 1.0))          # exit point of the unit sphere
        p = x + t[:, None] * d
        cos_i = np.einsum("ij,ij->i", d, p)               # inside, toward the outward normal
        cos_t2 = 1.0 - ETA * ETA * (1.0 - cos_i * cos_i)
        ok = (cos_t2 > 0) & (p[:, 1] > 0) & (d[:, 1] > 0)
        cos_t = np.sqrt(np.maximum(cos_t2, 0.0))
        wt = ETA * d + (cos_t - ETA * cos_i)[:, None] * p
        oc = p - light_c
        bb = np.einsum("ij,ij->i", oc, wt)
        disc = bb * bb - (np.einsum("ij,ij->i", oc, oc) - light_r * light_r)
        hit = ok & (disc > 0) & (-bb - np.sqrt(np.maximum(disc, 0)) > 0)
        vals = np.where(hit, radiance * (1.0 - _fresnel(cos_t, ETA)) * ETA * ETA * d[:, 1], 0.0)
        total += vals.mean() * omega
        var += (vals.std() * omega) ** 2 / n
        if hit.any():
            max_ang = max(max_ang, float(np.arccos(np.clip(z[hit], -1, 1)).max()))
    assert max_ang < 0.8 * half, max_ang          # the cone holds the whole image of the light
    return total, math.sqrt(var)


DOME_POINTS = [(0.35, 0.0, -0.2), (-0.55, 0.0, 0.3)]


def _dome_estimate(point, n=60000, seed=3):
    sc = scenes.sphere_dome(64, 64)
    o = oracle_lib.Oracle(sc, abi.SmsConfig.default(max_trials=1000000))
    x0 = np.asarray(point, np.float64)
    q = workloads.floor_queries_at(np.tile(x0, (n, 1)), sc.camera["origin"], shape=1)
    est, st = o.ss_estimate(q, seed=seed)
    o.close()
    light = sc.shapes[2]
    irr, irr_err = _dome_irradiance(x0, np.asarray(light["center"], np.float64), light["radius"], 100.0)
    ref, ref_err = 0.5 / math.pi * irr, 0.5 / math.pi * irr_err
    m = est[:, 0].astype(np.float64)
    return m.mean(), m.std() / math.sqrt(n), ref, ref_err, st


@pytest.mark.parametrize("point", DOME_POINTS)
def test_ss_dome_estimate_matches_brute_force_irradiance(point):
    mean, se, ref, ref_err, st = _dome_estimate(point)
    print(f"dome {point}: estimate {mean:.5g} +- {se:.2g}, brute force {ref:.5g} +- {ref_err:.2g}, "
          f"walks {st['walks']} converged {st['walks_converged']}")
    assert st["walks_converged"] > 0 and ref > 0
    assert abs(mean - ref) < 4 * math.hypot(se, ref_err) + 1e-3 * ref
    assert 4 * math.hypot(se, ref_err) < 0.03 * ref          # resolves a 10 % error


@pytest.mark.parametrize("site", ["ss_g", "ss_transmit"])
def test_ss_dome_brute_force_detects_planted_error(site):
    with planted(site):
        mean, se, ref, ref_err, _ = _dome_estimate(DOME_POINTS[0], n=30000)
    assert abs(mean - ref) > 4 * math.hypot(se, ref_err) + 1e-3 * ref


# ==================================================================== SS steps: FD Jacobians
def _sph(w):
    t = math.acos(max(-1.0, min(1.0, w[2])))
    ph = math.atan2(w[1], w[0])
    return t, ph + (2 * math.pi if ph < 0 else 0)


def _wrap(dp):
    if dp < -math.pi:
        dp += 2 * math.pi
    elif dp > math.pi:
        dp -= 2 * math.pi
    return dp


def _sphere_point(u, v, c, r):
    """Sphere parameterisation u = phi / 2 pi, v = theta / pi (sphere.cpp:329-374): point,
    outward normal and the shading tangent s = normalize(dp/du)."""
    phi, th = 2 * math.pi * u, math.pi * v
    loc = np.array([math.sin(th) * math.cos(phi), math.sin(th) * math.sin(phi), math.cos(th)])
    s = np.array([-math.sin(phi), math.cos(phi), 0.0])
    return c + r * loc, loc, s


def _anglediff(wi, wo, n, eta):
    """Angle-difference constraint between wo and T(wi), or between wi and T(wo) when T(wi)
    does not exist (manifold_ss.cpp:573-624)."""
    wt = _refract(wi, n, eta)
    if wt is not None:
        a, b = _sph(wo), _sph(wt)
    else:
        wt = _refract(wo, n, eta)
        if wt is None:
            return None
        a, b = _sph(wi), _sph(wt)
    return np.array([a[0] - b[0], _wrap(a[1] - b[1])])


def _halfvector(wi, wo, n, s, eta):
    """Tangential part (h.s, h.t) of the generalized half-vector (manifold_ss.cpp:438-486)."""
    e = eta if wi @ n >= 0 else 1.0 / eta
    h = wi + e * wo
    if e != 1.0:
        h = -h
    h = _nz(h)
    return np.array([h @ s, h @ np.cross(n, s)])


def _frame_s(n, dp_du0):
    """Shading tangent s = normalize(dp_du - n (n . dp_du)) (frame.h:197-200).  The reference
    differentiates it with dp_du held at its value at the current vertex (frame.h:224-244 takes
    dn_du and dn_dv only), so the finite differences hold dp_du0 fixed as well."""
    return _nz(dp_du0 - n * (n @ dp_du0))


def _ss_constraint(u, v, x, y, c, r, halfvector, dp_du0=None):
    p, n, s = _sphere_point(u, v, c, r)
    wi, wo = _nz(x - p), _nz(y - p)
    if halfvector:
        return _halfvector(wi, wo, n, _frame_s(n, s if dp_du0 is None else dp_du0), ETA)
    return _anglediff(wi, wo, n, ETA)


def _ss_step_errors(halfvector, n_cases=60, seed=9):
    sc = scenes.sphere_dome(64, 64)
    o = oracle_lib.Oracle(sc, abi.SmsConfig.default())
    rng = np.random.default_rng(seed)
    c = np.zeros(3)
    errs = []
    for _ in range(n_cases):
        x = np.array([rng.uniform(-0.6, 0.6), 0.0, rng.uniform(-0.6, 0.6)])
        y = np.array([rng.uniform(-1, 1), 5.0, rng.uniform(-1, 1)])
        qv = workloads.uniform_sphere(rng.random(1), rng.uniform(0.0, 0.35, 1))[0]
        p = np.array([qv[0], qv[2], qv[1]])
        d = _nz(p - x)
        e = _ei(y, (0, -1, 0))
        Co, dXo, vp = np.zeros(2, np.float32), np.zeros(2, np.float32), np.zeros(3, np.float32)
        xo, do = x.astype(np.float32), d.astype(np.float32)
        ok = o.lib.oracle_ss_step_at(o.h, halfvector, xo.ctypes.data, xo.ctypes.data, do.ctypes.data, e.ctypes.data,
                                     Co.ctypes.data, dXo.ctypes.data, vp.ctypes.data)
        if ok != 1:
            continue
        loc = _nz(vp.astype(np.float64) - c)
        v0 = math.acos(loc[2]) / math.pi
        u0 = (math.atan2(loc[1], loc[0]) % (2 * math.pi)) / (2 * math.pi)
        xf = xo.astype(np.float64)
        s0 = _sphere_point(u0, v0, c, 1.0)[2]
        f = lambda u, v: _ss_constraint(u, v, xf, y, c, 1.0, halfvector, s0)
        C0 = f(u0, v0)
        h = 1e-6
        ev = [f(u0 + h, v0), f(u0 - h, v0), f(u0, v0 + h), f(u0, v0 - h)]
        if C0 is None or any(a is None for a in ev):
            continue
        J = np.stack([(ev[0] - ev[1]) / (2 * h), (ev[2] - ev[3]) / (2 * h)], 1)
        if abs(np.linalg.det(J)) < 1e-3 or np.linalg.cond(J) > 1e4:
            continue
        dX = np.linalg.solve(J, C0)
        assert np.allclose(Co, C0, atol=2e-5), (Co, C0)
        errs.append(np.linalg.norm(dXo - dX) / max(np.linalg.norm(dX), 1e-2))
    o.close()
    return np.array(errs)


@pytest.mark.parametrize("halfvector", [0, 1])
def test_ss_newton_step_matches_finite_differences(halfvector):
    errs = _ss_step_errors(halfvector)
    print(f"SS step halfvector={halfvector}: {len(errs)} cases, median {np.median(errs):.2e} max {errs.max():.2e}")
    assert len(errs) >= 30
    assert errs.max() < 2e-3


@pytest.mark.parametrize("halfvector,site", [(0, "ss_ad_jac"), (1, "ss_hv_jac")])
def test_ss_newton_step_pin_detects_planted_error(halfvector, site):
    with planted(site):
        errs = _ss_step_errors(halfvector, n_cases=30)
    assert len(errs) >= 10 and np.median(errs) > 0.01


# ==================================================================== MS steps: FD Jacobians
def ball_lens():
    """Glass ball (r 1, IOR 1.5, multi-scatter caster) at (0, 1.5, 0) over the floor, sphere light
    above: chains floor -> ball (enter) -> ball (exit) -> light have two specular vertices."""
    sc = scenes.Scene()
    sc.add_sphere((0.0, 1.5, 0.0), 1.0, abi.BSDF_DIELECTRIC, flags=abi.FLAG_CAUSTIC_CASTER_MS, eta=ETA)
    sc.add_rectangle(scenes.plane_to_world((0, 0, 0), 10.0), abi.BSDF_DIFFUSE, flags=abi.FLAG_CAUSTIC_RECEIVER)
    light = sc.add_sphere((0.0, 5.0, 0.0), 0.1, abi.BSDF_DIFFUSE, reflectance=(0, 0, 0))
    sc.add_area_light(light, (100.0, 100.0, 100.0))
    sc.set_camera((0, 3, 6), (0, 0.5, 0), (0, 1, 0), 45.0, 64, 64)
    return sc


def _ms_constraint(uv, x0, y, c, halfvector, dp_du0):
    """Per-vertex constraints of a two-vertex chain on one sphere (manifold_ms.cpp:500-857);
    dp_du0: the tangents the shading frames are differentiated with (see _frame_s)."""
    pts = [_sphere_point(uv[2 * i], uv[2 * i + 1], c, 1.0) for i in range(2)]
    out = []
    for i in range(2):
        p, n, _ = pts[i]
        s = _frame_s(n, dp_du0[i])
        prev = x0 if i == 0 else pts[0][0]
        nxt = y if i == 1 else pts[1][0]
        wi, wo = _nz(prev - p), _nz(nxt - p)
        ci = _halfvector(wi, wo, n, s, ETA) if halfvector else _anglediff(wi, wo, n, ETA)
        if ci is None:
            return None
        out.append(ci)
    return np.concatenate(out)


def _ms_step_errors(halfvector, n_cases=50, seed=13):
    sc = ball_lens()
    o = oracle_lib.Oracle(sc, abi.SmsConfig.default(bounces=2))
    lib = _lib()
    rng = np.random.default_rng(seed)
    c = np.array([0.0, 1.5, 0.0])
    errs = []
    for _ in range(n_cases):
        x0 = np.array([rng.uniform(-0.5, 0.5), 0.0, rng.uniform(-0.5, 0.5)])
        y = np.array([rng.uniform(-0.3, 0.3), 5.0, rng.uniform(-0.3, 0.3)])
        a = workloads.uniform_sphere(rng.random(1), rng.uniform(0.8, 1.0, 1))[0]      # lower cap (-z)
        p1 = c + np.array([a[0], a[2], a[1]])
        b = workloads.uniform_sphere(rng.random(1), rng.uniform(0.0, 0.2, 1))[0]      # upper cap (+z)
        p2 = c + np.array([b[0], b[2], b[1]])
        pts = np.concatenate([p1, p2]).astype(np.float32)
        e = _ei(y, (0, -1, 0))
        Co, dxo, vp = np.zeros(4, np.float32), np.zeros(4, np.float32), np.zeros(6, np.float32)
        x32 = x0.astype(np.float32)
        ok = lib.oracle_ms_step_at(o.h, halfvector, x32.ctypes.data, 2, pts.ctypes.data, e.ctypes.data,
                                   Co.ctypes.data, dxo.ctypes.data, vp.ctypes.data)
        if ok != 1:
            continue
        uv0 = []
        for i in range(2):
            loc = _nz(vp[3 * i:3 * i + 3].astype(np.float64) - c)
            uv0 += [(math.atan2(loc[1], loc[0]) % (2 * math.pi)) / (2 * math.pi), math.acos(loc[2]) / math.pi]
        uv0 = np.array(uv0)
        xf = x32.astype(np.float64)
        s0 = [_sphere_point(uv0[2 * i], uv0[2 * i + 1], c, 1.0)[2] for i in range(2)]
        C0 = _ms_constraint(uv0, xf, y, c, halfvector, s0)
        if C0 is None:
            continue
        h = 1e-6
        J = np.zeros((4, 4))
        bad = False
        for k in range(4):
            d = np.zeros(4)
            d[k] = h
            cp, cm = _ms_constraint(uv0 + d, xf, y, c, halfvector, s0), _ms_constraint(uv0 - d, xf, y, c, halfvector, s0)
            if cp is None or cm is None:
                bad = True
                break
            J[:, k] = (cp - cm) / (2 * h)
        if bad or np.linalg.cond(J) > 1e4:
            continue
        dx = np.linalg.solve(J, C0)
        assert np.allclose(Co, C0, atol=3e-5), (Co, C0)
        errs.append(np.linalg.norm(dxo - dx) / max(np.linalg.norm(dx), 1e-2))
    o.close()
    return np.array(errs)


@pytest.mark.parametrize("halfvector", [0, 1])
def test_ms_newton_step_matches_finite_differences(halfvector):
    errs = _ms_step_errors(halfvector)
    print(f"MS step halfvector={halfvector}: {len(errs)} cases, median {np.median(errs):.2e} max {errs.max():.2e}")
    assert len(errs) >= 25
    assert errs.max() < 3e-3


@pytest.mark.parametrize("halfvector,site", [(0, "ms_ad_jac"), (1, "ms_hv_jac")])
def test_ms_newton_step_pin_detects_planted_error(halfvector, site):
    with planted(site):
        errs = _ms_step_errors(halfvector, n_cases=30)
    assert len(errs) >= 10 and np.median(errs) > 0.01


# ==================================================================== MNEE initialisation
def _ray_unit_sphere_exit(x, d):
    b = d @ x
    return x + (-b + math.sqrt(b * b - (x @ x - 1.0))) * d


def _mnee_dome_errors(n=300, seed=7):
    sc = scenes.sphere_dome(64, 64)
    o = oracle_lib.Oracle(sc, abi.SmsConfig.default(mnee_init=1))
    w = workloads.ss_walk_inputs(sc, n, seed, extent=0.7)
    out = o.ss_walks(w)
    o.close()
    assert out["init_valid"].all()
    err = []
    for i in range(n):
        x = w["si_p"][i].astype(np.float64)
        ref = _ray_unit_sphere_exit(x, _nz(w["ei"]["p"][i].astype(np.float64) - x))
        err.append(np.linalg.norm(out["p_init"][i] - ref))
    return np.array(err), w, out


def test_mnee_init_seed_is_the_straight_ray_to_the_light():
    err, w, out = _mnee_dome_errors()
    assert err.max() < 2e-6
    # the random seed sample plays no part: other samples give the same walks
    sc = scenes.sphere_dome(64, 64)
    o = oracle_lib.Oracle(sc, abi.SmsConfig.default(mnee_init=1))
    w2 = w.copy()
    w2["sample"] = np.random.default_rng(1).random((len(w2), 2))
    out2 = o.ss_walks(w2)
    o.close()
    for k in ("success", "iterations", "p", "p_init"):
        np.testing.assert_array_equal(out[k], out2[k])
    # converged MNEE walks land on a brute-force refraction root
    ok = np.nonzero(out["success"] == 1)[0]
    assert len(ok) > 0.5 * len(out)
    for i in ok[:40]:
        roots = _dome_roots(w["si_p"][i].astype(np.float64), w["ei"]["p"][i].astype(np.float64))
        d = min(np.linalg.norm(out["p"][i] - r) for r in roots)
        assert d < 1e-4


def test_mnee_init_pin_detects_planted_error():
    with planted("mnee_target"):
        err, _, _ = _mnee_dome_errors(n=60)
    assert np.median(err) > 1e-3            # the pin's tolerance is 2e-6


def test_mnee_init_conductor_aims_at_the_bounding_box_centre():
    """Reflective casters aim the MNEE seed ray at the shape's bounding-box centre
    (manifold_ss.cpp:231-236): on the plane mirror x = 1.2 that is the mirror's centre."""
    sc = scenes.Scene()
    M = np.array([[0, 0, -1, 1.2], [0, 2.0, 0, 1.0], [2.0, 0, 0, 0.0]], np.float32)
    sc.add_rectangle(M, abi.BSDF_CONDUCTOR, flags=abi.FLAG_CAUSTIC_CASTER_SS, eta=0.0, k=1.0)
    sc.add_rectangle(scenes.plane_to_world((0, 0, 0), 10.0), abi.BSDF_DIFFUSE, flags=abi.FLAG_CAUSTIC_RECEIVER)
    sc.add_point_light((0.2, 2.0, 0.3), (1, 1, 1))
    o = oracle_lib.Oracle(sc, abi.SmsConfig.default(mnee_init=1))
    rng = np.random.default_rng(2)
    n = 100
    w = np.zeros(n, abi.SS_WALK_IN)
    w["si_p"] = np.stack([rng.uniform(-1, 0.8, n), np.zeros(n), rng.uniform(-1, 1, n)], -1)
    w["shape"] = 0
    w["ei"]["p"] = (0.2, 2.0, 0.3)
    w["ei"]["kind"] = abi.EI_POINT
    w["n_offset"] = (0, 0, 1)
    w["sample"] = rng.random((n, 2))
    out = o.ss_walks(w)
    o.close()
    assert out["init_valid"].all()
    np.testing.assert_allclose(out["p_init"], np.tile([1.2, 1.0, 0.0], (n, 1)), atol=2e-6)
    assert (out["success"] == 1).mean() > 0.5


def test_mnee_init_multi_scatter_is_deterministic_and_solves_the_slab():
    """MS MNEE seeds trace the straight line toward the light without bending
    (manifold_ms.cpp:311-318, :386-389): the walks do not depend on the sample and converge to the
    flat slab's Snell chain."""
    from test_oracle_ms import _slab_chain
    sc = scenes.slab(width=32, height=32, n=8, amplitude=0.0)
    o = oracle_lib.Oracle(sc, abi.SmsConfig.default(bounces=2, mnee_init=1))
    w = workloads.ms_walk_inputs(sc, 200, seed=11, extent=0.25)
    out = o.ms_walks(w)
    w2 = w.copy()
    w2["sample"] = np.random.default_rng(3).random((len(w2), 2))
    out2 = o.ms_walks(w2)
    o.close()
    for k in ("success", "iterations", "p", "shapes"):
        np.testing.assert_array_equal(out[k], out2[k])
    ok = out["success"] == 1
    assert ok.mean() > 0.8
    for i in np.nonzero(ok)[0]:
        v1, v2 = _slab_chain(w["si_p"][i].astype(np.float64), w["ei"]["p"][i].astype(np.float64))
        np.testing.assert_allclose(out["p"][i, 0], v1, atol=2e-4)
        np.testing.assert_allclose(out["p"][i, 1], v2, atol=2e-4)
