"""Pins of the learned-proposal oracle (CPU only).

The network oracle (oracle/mlp_oracle.py) is checked against torch autograd on an
independently written float64 forward (features, layers, mixture log-density through
torch.distributions), against central finite differences, and against scipy's bivariate
normal density.  The seed oracle (oracle/sms_oracle.cpp: mixture preparation, Box-Muller draw,
UV lookup) is checked against numpy closed forms, sampling statistics and a brute-force scan
of every face.  Unbiasedness of the defensive-mixture estimator is checked against uniform
seeding on a small pool.
"""
import math
import os
import sys

import numpy as np
import pytest

import oracle_lib
from smsamd import abi, scenes, workloads

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as M  # noqa: E402


def _batch(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n, 9)).astype(np.float32)
    uv = rng.uniform(0.05, 0.95, (n, 2)).astype(np.float32)
    return x, uv


def _torch_loss(params, x, uv):
    """Independent float64 forward: torch layers and a torch.distributions mixture."""
    import torch
    p = torch.tensor(params, dtype=torch.float64, requires_grad=True)
    xt = torch.tensor(x, dtype=torch.float64)
    feats = [xt]
    for k in range(8):
        a = math.pi * (2 ** k) * xt
        feats += [torch.sin(a), torch.cos(a)]
    h = torch.cat(feats + [torch.zeros(len(x), 7, dtype=torch.float64)], 1)
    o = 0
    for i, (n, k) in enumerate(M.LAYERS):
        W = p[o:o + n * k].reshape(n, k)
        b = p[o + n * k:o + n * k + n]
        o += n * k + n
        h = h @ W.T + b
        if i < 4:
            h = torch.relu(h)
    r = h.reshape(-1, 8, 6)
    logits = r[..., 0]
    mu = torch.sigmoid(r[..., 1:3])
    sig = torch.exp(torch.clamp(r[..., 3:5], -7.0, 0.0))
    rho = 0.95 * torch.tanh(r[..., 5])
    cov = torch.stack([torch.stack([sig[..., 0] ** 2, rho * sig[..., 0] * sig[..., 1]], -1),
                       torch.stack([rho * sig[..., 0] * sig[..., 1], sig[..., 1] ** 2], -1)], -2)
    comp = torch.distributions.MultivariateNormal(mu, covariance_matrix=cov)
    lp = comp.log_prob(torch.tensor(uv, dtype=torch.float64)[:, None, :]) + torch.log_softmax(logits, -1)
    loss = -torch.logsumexp(lp, -1).mean()
    loss.backward()
    return float(loss.detach()), p.grad.numpy()


def test_bf16_rounding_known_answers():
    v = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -8 + 2 ** -20, -2.5, 3.0e38, np.inf], np.float32)
    r = M.bf16(v)
    # ties to even: 1 + 2^-8 sits half-way between 1 and 1 + 2^-7 -> 1; 1 + 3 2^-8 -> 1 + 2^-6
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1 + 2 ** -6 and r[3] == 1 + 2 ** -7
    assert r[4] == -2.5 and np.isinf(r[6])
    assert abs(r[5] / 3.0e38 - 1) < 2 ** -8
    assert np.isnan(M.bf16(np.array([np.nan], np.float32)))[0]


def test_mixture_nll_matches_scipy_density():
    from scipy.stats import multivariate_normal
    rng = np.random.default_rng(4)
    o = rng.normal(0, 1.5, (20, 48))
    uv = rng.uniform(0, 1, (20, 2))
    ours = M.nll(o, uv)
    w, mu, sig, rho = M.head(o)
    for i in range(20):
        dens = 0.0
        for c in range(8):
            s = sig[i, c]
            cov = [[s[0] ** 2, rho[i, c] * s[0] * s[1]], [rho[i, c] * s[0] * s[1], s[1] ** 2]]
            dens += w[i, c] * multivariate_normal(mu[i, c], cov).pdf(uv[i])
        assert abs(ours[i] + math.log(dens)) < 1e-8 * max(1.0, abs(ours[i]))
    loss2, _ = M.nll_grad(o, uv)
    np.testing.assert_allclose(loss2, ours, rtol=1e-10, atol=1e-10)


def test_mlp_gradient_matches_torch_autograd():
    pytest.importorskip("torch")
    p = M.init_like_library(1).astype(np.float64)
    p += np.random.default_rng(2).normal(0, 0.05, p.shape) * (p != 0)
    # biases non-zero so that every path is exercised
    for (n, k), (w0, b0) in zip(M.LAYERS, M.OFFSETS):
        p[b0:b0 + n] = np.random.default_rng(b0).normal(0, 0.1, n)
    x, uv = _batch(96, 5)
    loss, g = M.loss_and_grad(p, x, uv, "f64")
    tl, tg = _torch_loss(p, x, uv)
    assert abs(loss - tl) < 1e-10 * max(1, abs(tl))
    err = np.abs(g - tg).max() / np.abs(tg).max()
    assert err < 1e-9, err
    # padded feature columns get zero gradient
    W1 = g[:M.OFFSETS[0][1]].reshape(64, 160)
    assert not W1[:, 153:].any()


def test_head_gradient_finite_differences():
    rng = np.random.default_rng(8)
    o = rng.normal(0, 1.0, (16, 48))
    o[:, 3::6] = rng.uniform(-5, -0.5, (16, 8))      # inside the clamp
    o[:, 4::6] = rng.uniform(-5, -0.5, (16, 8))
    uv = rng.uniform(0.1, 0.9, (16, 2))
    _, g = M.nll_grad(o, uv)
    h = 1e-6
    for j in range(48):
        e = np.zeros(48)
        e[j] = h
        fd = (M.nll(o + e, uv) - M.nll(o - e, uv)) / (2 * h)
        np.testing.assert_allclose(g[:, j], fd, rtol=2e-5, atol=2e-7)
    # clamped log-sigmas have zero gradient
    o2 = o.copy()
    o2[:, 3] = -9.0
    _, g2 = M.nll_grad(o2, uv)
    assert not g2[:, 3].any()


def test_bf16_mode_tracks_exact_gradient():
    p = M.init_like_library(3)
    x, uv = _batch(512, 6)
    l64, g64 = M.loss_and_grad(p, x, uv, "f64")
    lbf, gbf = M.loss_and_grad(p, x, uv, "bf16")
    assert abs(l64 - lbf) < 2e-2 * abs(l64)
    cos = g64 @ gbf / np.linalg.norm(g64) / np.linalg.norm(gbf)
    assert cos > 0.99, cos


def test_adam_first_step_closed_form():
    """Step 1 from zero moments: m_hat = g, v_hat = g^2, so p -= lr g / (|g| + eps)."""
    rng = np.random.default_rng(9)
    p = rng.normal(0, 1, 1000).astype(np.float32)
    g = rng.normal(0, 1e-2, 1000).astype(np.float32)
    p1, m, v = M.adam(p, g, np.zeros(1000, np.float32), np.zeros(1000, np.float32), 1, 1e-3)
    np.testing.assert_allclose(p1, p - 1e-3 * g / (np.abs(g) + 1e-8), rtol=0, atol=2e-6)
    np.testing.assert_allclose(m, 0.1 * g, rtol=1e-6)
    # (1 - 0.999) rounded to float32 is 0.00099998713
    np.testing.assert_allclose(v, 0.001 * g * g, rtol=3e-5)


# ---------------------------------------------------------------- seed oracle (C)
def test_gmm_prepare_matches_head():
    rng = np.random.default_rng(10)
    raw = rng.normal(0, 2.0, (200, 48)).astype(np.float32)
    raw[:, 3::6] = rng.uniform(-9, 2, (200, 8))
    g = oracle_lib.gmm_prepare(raw)
    w, mu, sig, rho = M.head(raw)
    np.testing.assert_allclose(g[:, :8], np.cumsum(w, 1), rtol=0, atol=3e-6)
    np.testing.assert_allclose(g[:, 8:16], mu[..., 0], rtol=2e-6, atol=1e-7)
    np.testing.assert_allclose(g[:, 16:24], mu[..., 1], rtol=2e-6, atol=1e-7)
    np.testing.assert_allclose(g[:, 24:32], sig[..., 0], rtol=2e-6)
    np.testing.assert_allclose(g[:, 32:40], sig[..., 1], rtol=2e-6)
    np.testing.assert_allclose(g[:, 40:48], rho, rtol=2e-6, atol=1e-7)


def test_seed_draw_statistics():
    """Mixture of two well-separated components with a 0.3 defensive weight: the draws follow
    the defensive split, the component weights, means and covariances."""
    raw = np.zeros(48, np.float32)
    raw[0::6] = -30.0
    raw[0], raw[6] = math.log(0.25), math.log(0.75)
    raw[1], raw[2], raw[3], raw[4], raw[5] = -1.0, 0.5, math.log(0.05), math.log(0.1), 0.4
    raw[7], raw[8], raw[9], raw[10], raw[11] = 1.2, -0.8, math.log(0.08), math.log(0.03), -1.0
    g = oracle_lib.gmm_prepare(raw)[0]
    n = 200000
    s = oracle_lib.prop_seeds(g, 0.3, 17, 5, 1, n)
    uni = s[:, 0] == 0
    sd = math.sqrt(0.3 * 0.7 / n)
    assert abs(uni.mean() - 0.3) < 5 * sd
    assert abs(s[uni, 1].mean() - 0.5) < 0.01 and abs(s[uni, 2].mean() - 0.5) < 0.01
    assert (s[uni, 1:] >= 0).all() and (s[uni, 1:] < 1).all()
    mix = s[~uni, 1:].astype(np.float64)
    w, mu, sig, rho = M.head(raw[None])
    first = np.linalg.norm(mix - mu[0, 0], axis=1) < np.linalg.norm(mix - mu[0, 1], axis=1)
    assert abs(first.mean() - 0.25) < 0.01
    for c, sel in ((0, first), (1, ~first)):
        pts = mix[sel]
        np.testing.assert_allclose(pts.mean(0), mu[0, c], atol=4e-3)
        cov = np.cov(pts.T)
        sx, sy = sig[0, c]
        np.testing.assert_allclose(cov[0, 0], sx * sx, rtol=3e-2)
        np.testing.assert_allclose(cov[1, 1], sy * sy, rtol=3e-2)
        assert abs(cov[0, 1] / (sx * sy) - rho[0, c]) < 2e-2


@pytest.fixture(scope="module")
def small_pool():
    sc = scenes.pool(width=32, height=32, n=32, n_waves=16, seed=3)
    return sc


def test_uv_lookup_grid_equals_brute_force(small_pool):
    orc = oracle_lib.Oracle(small_pool, abi.SmsConfig.default())
    rng = np.random.default_rng(12)
    uv = rng.random((20000, 2)).astype(np.float32)
    # include exact grid-line and vertex coordinates
    uv[:2000] = (rng.integers(0, 33, (2000, 2)) / 32.0).astype(np.float32)
    uv = np.minimum(uv, np.float32(np.nextafter(np.float32(1), np.float32(0))))
    pg, okg = orc.uv_to_point(0, uv)
    pb, okb = orc.uv_to_point(0, uv, brute=True)
    assert okg.all() and np.array_equal(okg, okb)
    assert np.array_equal(pg, pb)
    # the regular grid mesh: the point is the barycentric interpolation over the quad's triangle
    sh = small_pool.shapes[0]
    P = np.asarray(sh["positions"], np.float64).reshape(33, 33, 3)
    u, v = uv[:, 0].astype(np.float64) * 32, uv[:, 1].astype(np.float64) * 32
    i, j = np.minimum(np.floor(u).astype(int), 31), np.minimum(np.floor(v).astype(int), 31)
    fu, fv = u - i, v - j
    lower = fu + fv <= 1.0
    p00, p10, p01, p11 = P[j, i], P[j, i + 1], P[j + 1, i], P[j + 1, i + 1]
    ref = np.where(lower[:, None], p00 + fu[:, None] * (p10 - p00) + fv[:, None] * (p01 - p00),
                   p11 + (1 - fu)[:, None] * (p01 - p11) + (1 - fv)[:, None] * (p10 - p11))
    np.testing.assert_allclose(pg, ref, atol=2e-6)
    orc.close()


def test_scene_box_and_inputs(small_pool):
    orc = oracle_lib.Oracle(small_pool, abi.SmsConfig.default())
    c, inv = orc.scene_box()
    pts = [np.asarray(small_pool.shapes[0]["positions"]).reshape(-1, 3)]
    light = small_pool.shapes[2]
    pts.append(np.asarray(light["center"]) + np.array([[-1, -1, -1], [1, 1, 1]]) * light["radius"])
    pts.append(np.array([[-1, -1, -1], [1, -1, 1]], np.float64))   # floor rectangle corners
    allp = np.concatenate(pts)
    lo, hi = allp.min(0), allp.max(0)
    np.testing.assert_allclose(c, 0.5 * (lo + hi), atol=1e-6)
    np.testing.assert_allclose(inv, 1.0 / (0.5 * (hi - lo)).max(), rtol=1e-6)
    q = workloads.plane_queries(small_pool, 500, 2, extent=0.9, shape=1)
    q["p"][:, 1] = -1.0
    x = orc.prop_inputs(q, 7)
    np.testing.assert_allclose(x[:, :3] / inv + c, q["p"], atol=1e-5)
    np.testing.assert_array_equal(x[:, 3:6], q["sh_n"])
    lp = x[:, 6:9].astype(np.float64) / inv + c
    np.testing.assert_allclose(np.linalg.norm(lp - light["center"], axis=1), light["radius"], atol=1e-5)
    orc.close()


def _concentrated_raw(q, sc, sigma=0.06):
    """Hand-made mixtures centred on the uv above each query point (shifted toward the light)."""
    n = len(q)
    raw = np.full((n, 48), 0.0, np.float32)
    light = np.asarray(sc.shapes[2]["center"])
    p = q["p"].astype(np.float64)
    t = (0.0 - p[:, 1]) / (light[1] - p[:, 1])
    hit = p + t[:, None] * (light - p)
    u = np.clip((hit[:, 0] + 1) / 2, 0.02, 0.98)
    v = np.clip((hit[:, 2] + 1) / 2, 0.02, 0.98)
    for c in range(8):
        raw[:, 6 * c] = 0.0
        raw[:, 6 * c + 1] = np.log(u / (1 - u)) + 0.02 * (c - 3.5)
        raw[:, 6 * c + 2] = np.log(v / (1 - v)) - 0.02 * (c - 3.5)
        raw[:, 6 * c + 3] = math.log(sigma)
        raw[:, 6 * c + 4] = math.log(sigma)
        raw[:, 6 * c + 5] = 0.0
    return raw


def test_mixture_estimator_is_unbiased(small_pool):
    """Same queries, many seeds: the defensive-mixture estimate's mean equals the uniform
    estimate's mean (both unbiased for the same integral) within their standard errors."""
    cfg = abi.SmsConfig.default(max_trials=100000)
    orc = oracle_lib.Oracle(small_pool, cfg)
    q = workloads.plane_queries(small_pool, 24, 5, extent=0.6, shape=1)
    q["p"][:, 1] = -1.0
    raw = _concentrated_raw(q, small_pool)
    reps = 60
    qq = np.concatenate([q] * reps)
    rr = np.concatenate([raw] * reps)
    uni, su = orc.ss_estimate(qq, 3)
    mix, sm = orc.ss_estimate_prop(qq, 3, rr, 0, 0.2)
    lu, lm = uni[:, 0].astype(np.float64), mix[:, 0].astype(np.float64)
    se = math.sqrt(lu.var() / len(lu) + lm.var() / len(lm))
    print(f"uniform {lu.mean():.5f} mixture {lm.mean():.5f} se {se:.5f} trials {su['bernoulli_trials']} "
          f"{sm['bernoulli_trials']}")
    assert lu.mean() > 0
    assert abs(lu.mean() - lm.mean()) < 4 * se
    # the concentrated proposal converges more often per walk
    assert sm["walks_converged"] / sm["walks"] > su["walks_converged"] / su["walks"]
    orc.close()


def test_anchor_closed_form_on_flat_water():
    """Flat water at y = 0 over the floor at y = -1: a query's anchor is the uv where the segment
    from its floor point to its light sample crosses the plane (u = (x + 1) / 2, v = (z + 1) / 2
    on the grid mesh), and NaN where the segment passes beside the water."""
    sc = scenes.pool(width=32, height=32, n=32, n_waves=0, seed=3)
    orc = oracle_lib.Oracle(sc, abi.SmsConfig.default())
    q = workloads.plane_queries(sc, 600, 2, extent=2.2, shape=1)
    q["p"][:, 1] = -1.0
    a = orc.prop_anchors(q, 7, 0)
    c, inv = orc.scene_box()
    lp = orc.prop_inputs(q, 7)[:, 6:9].astype(np.float64) / inv + c
    p = q["p"].astype(np.float64)
    t = (0.0 - p[:, 1]) / (lp[:, 1] - p[:, 1])
    hit = p + t[:, None] * (lp - p)
    inside = (np.abs(hit[:, 0]) < 0.999) & (np.abs(hit[:, 2]) < 0.999)
    outside = (np.abs(hit[:, 0]) > 1.001) | (np.abs(hit[:, 2]) > 1.001)
    assert inside.sum() > 100 and outside.sum() > 50
    assert np.isfinite(a[inside]).all() and np.isnan(a[outside]).all()
    np.testing.assert_allclose(a[inside, 0], (hit[inside, 0] + 1) / 2, atol=2e-5)
    np.testing.assert_allclose(a[inside, 1], (hit[inside, 2] + 1) / 2, atol=2e-5)
    # a query without a receiver has no anchor
    q2 = q[:4].copy()
    q2["shape"] = -1
    assert np.isnan(orc.prop_anchors(q2, 7, 0)).all()
    orc.close()


def test_anchored_mixture_estimator_is_unbiased(small_pool):
    """Anchored mixtures (uv box of half side 0.125 around each query's anchor) centred on the
    anchor: the estimate's mean equals the uniform one within the standard errors, walks
    converge more often than uniform ones, and a mixture far off the solutions (a box of half
    side 0.5 centred a quarter of the uv square away) still leaves the mean unbiased."""
    cfg = abi.SmsConfig.default(max_trials=100000)
    orc = oracle_lib.Oracle(small_pool, cfg)
    q = workloads.plane_queries(small_pool, 24, 5, extent=0.6, shape=1)
    q["p"][:, 1] = -1.0
    raw = np.zeros((len(q), 48), np.float32)
    for c in range(8):
        raw[:, 6 * c + 1] = 0.05 * (c - 3.5)
        raw[:, 6 * c + 2] = -0.05 * (c - 3.5)
        raw[:, 6 * c + 3] = raw[:, 6 * c + 4] = math.log(0.15)
    reps = 60
    qq = np.concatenate([q] * reps)
    rr = np.concatenate([raw] * reps)
    uni, su = orc.ss_estimate(qq, 3)
    mix, sm = orc.ss_estimate_prop(qq, 3, rr, 0, 0.2, radius=0.125)
    off = rr.copy()
    off[:, 1::6] = 2.0                                   # mixture means at xi = 0.88: uv = anchor + 0.38
    far, sf = orc.ss_estimate_prop(qq, 3, off, 0, 0.2, radius=0.5)
    lu, lm, lf = (r[:, 0].astype(np.float64) for r in (uni, mix, far))
    se = math.sqrt(lu.var() / len(lu) + lm.var() / len(lm))
    se_f = math.sqrt(lu.var() / len(lu) + lf.var() / len(lf))
    print(f"uniform {lu.mean():.5f} anchored {lm.mean():.5f} se {se:.5f} far {lf.mean():.5f} se {se_f:.5f}")
    assert lu.mean() > 0
    assert abs(lu.mean() - lm.mean()) < 4 * se
    assert abs(lu.mean() - lf.mean()) < 4 * se_f
    assert sm["walks_converged"] / sm["walks"] > su["walks_converged"] / su["walks"]
    assert sf["walks_converged"] / sf["walks"] < sm["walks_converged"] / sm["walks"]
    orc.close()


def test_weighted_gradient_matches_torch_and_reduces_to_unweighted():
    """Weighted NLL: unit weights give the unweighted gradient; integer weights equal repeating
    samples; the weighted gradient matches torch autograd of the weighted loss."""
    pytest.importorskip("torch")
    p = M.init_like_library(4).astype(np.float64)
    x, uv = _batch(48, 11)
    l1, g1 = M.loss_and_grad(p, x, uv, "f64")
    lu, gu = M.loss_and_grad(p, x, uv, "f64", np.ones(48))
    assert abs(l1 - lu) < 1e-12 and np.abs(g1 - gu).max() < 1e-12
    reps = np.random.default_rng(3).integers(1, 4, 48)
    lw, gw = M.loss_and_grad(p, x, uv, "f64", reps.astype(np.float64))
    lr_, gr = M.loss_and_grad(p, np.repeat(x, reps, 0), np.repeat(uv, reps, 0), "f64")
    assert abs(lw - lr_) < 1e-10 and np.abs(gw - gr).max() < 1e-10 * max(1.0, np.abs(gr).max())
