"""Pins of the oracle's glint solver (a16-a18): every converged walk must sit on a root of
"bilinearly interpolated level-0 slope = target slope" found per texel cell in FP64 from the
decoded RGB8 map (normalmap.cpp:46-66, normalmap.h:62-104); walks on a map whose slopes vary
linearly across a cell converge to the closed-form point."""
import math

import numpy as np
import pytest

import oracle_lib
from smsamd import abi, scenes, workloads


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
