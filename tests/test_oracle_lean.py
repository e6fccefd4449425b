"""Pins of the oracle's LEAN pdf (the glint MIS weight, manifold_glints.cpp:155-179,556-579)
by mathematics that does not reuse its code:

* the FP64 exp / erf / log2 it is built on equal Python's libm to 1e-15;
* every LEAN mip texel is the mean of its block of level-0 texels (box filter, bitmap.cpp:234-312)
  and the level-0 second moments carry the base variance 0.5 * 1e-4^2 (normalmap.cpp:36-60);
* for zero-mean, uncorrelated slopes the LEAN pdf equals the closed-form Beckmann reflection pdf
  D(h) / (4 cos(theta_i) (1 + Lambda(wi))) with D and Lambda of Walter et al. 2007;
* the pdf integrates to one over outgoing directions (it is a density of wo);
* the mip-level rule, including the reference's coarsest level for footprints below one texel.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_lib
from smsamd import abi, scenes


@pytest.fixture(scope="module")
def lib():
    return oracle_lib.load()


def test_fp64_transcendentals(lib):
    for x in np.concatenate([np.linspace(-40, 40, 801), [-745.0, -700.3, 1e-12, 0.0, 709.0]]):
        assert lib.oracle_exp_d(float(x)) == pytest.approx(math.exp(x), rel=1e-15, abs=1e-300)
    for x in np.concatenate([np.linspace(-7, 7, 1401), [2.4999999, 2.5, 5.999, 6.0, 1e-9]]):
        assert lib.oracle_erf_d(float(x)) == pytest.approx(math.erf(x), rel=2e-15, abs=1e-16)
    for x in np.concatenate([np.geomspace(1e-8, 1e4, 997), [0.5, 1.0, 2.0, 1024.0, 2.0 ** -20]]):
        assert lib.oracle_log2_d(float(x)) == pytest.approx(math.log2(x), rel=1e-15, abs=1e-15)
    for e in range(-30, 31):                      # powers of two are exact (level boundaries)
        assert lib.oracle_log2_d(2.0 ** e) == float(e)


def _slopes(rgb):
    v = 2.0 * (rgb.astype(np.float64) / 255.0) - 1.0
    n = v / np.linalg.norm(v, axis=-1, keepdims=True)
    return -n[..., 0] / n[..., 2], -n[..., 1] / n[..., 2]


def test_lean_pyramid_is_block_mean(lib):
    rgb = scenes.scratch_bump_normalmap(res=64, n_scratches=40, seed=4)
    sc = scenes.glint_plane(16, 16, nm_res=64, n_scratches=40, seed=4, normalmap=rgb)
    o = oracle_lib.Oracle(sc, abi.SmsConfig.default())
    levels = lib.oracle_lean_levels(o.h, 0)
    assert levels == 7
    sx, sy = _slopes(rgb.reshape(64, 64, 3))
    base = 0.5 * 1e-4 * 1e-4
    l0 = np.stack([sx, sy, sx * sx + base, sy * sy + base, sx * sy], -1)
    out = (C.c_float * 5)()
    for level in range(levels):
        r = 64 >> level
        for (x, y) in [(0, 0), (r - 1, r - 1), (r // 2, r // 3)]:
            lib.oracle_lean_texel(o.h, 0, level, x, y, out)
            b = 1 << level
            ref = l0[y * b:(y + 1) * b, x * b:(x + 1) * b].reshape(-1, 5).mean(0)
            np.testing.assert_allclose(np.array(out[:]), ref, rtol=2e-5, atol=2e-7)
    o.close()


def _beckmann_pdf(wi, wo, alpha):
    h = wi + wo
    h = h / np.linalg.norm(h)
    cos_h = h[2]
    tan2 = (1.0 - cos_h * cos_h) / (cos_h * cos_h)
    D = math.exp(-tan2 / (alpha * alpha)) / (math.pi * alpha * alpha * cos_h ** 4)
    tan_i = math.sqrt(1.0 - wi[2] ** 2) / wi[2]
    a = 1.0 / (alpha * tan_i)
    lam = (math.erf(a) - 1.0) / 2.0 + math.exp(-a * a) / (2.0 * a * math.sqrt(math.pi))
    return D / (4.0 * wi[2] * (1.0 + lam))


def _dir(theta, phi):
    return np.array([math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta)])


def _pdf(lib, mu, sig, wi, wo):
    f = lambda v: (C.c_float * len(v))(*[float(t) for t in v])
    return lib.oracle_lean_pdf_params(f(mu), f(sig), f(wi), f(wo))


@pytest.mark.parametrize("alpha", [0.05, 0.2, 0.45])
def test_lean_pdf_equals_beckmann(lib, alpha):
    s2 = 0.5 * alpha * alpha                      # slope variance of a Beckmann surface
    rng = np.random.default_rng(1)
    for _ in range(40):
        wi = _dir(rng.uniform(0.0, 1.2), rng.uniform(0, 2 * math.pi))
        wo = _dir(rng.uniform(0.0, 1.2), rng.uniform(0, 2 * math.pi))
        ref = _beckmann_pdf(wi, wo, alpha)
        got = _pdf(lib, (0, 0), (s2, s2, 0.0), wi, wo)
        assert got == pytest.approx(ref, rel=2e-4, abs=1e-30), (wi, wo)


@pytest.mark.parametrize("mu,sig,theta_i", [((0.0, 0.0), (0.02, 0.03, 0.005), 0.35),
                                            ((0.1, -0.05), (0.03, 0.02, -0.004), 0.6),
                                            ((0.0, 0.0), (0.08, 0.08, 0.0), 0.1)])
def test_lean_pdf_integrates_to_one(lib, mu, sig, theta_i):
    # the covariance clamp of eval_lean_sigma keeps sigma(0,1) >= 0: pass it clamped
    sig = (sig[0], sig[1], max(sig[2], 0.0))
    wi = _dir(theta_i, 0.7)
    nt, npf = 240, 240
    th = (np.arange(nt) + 0.5) * (0.5 * math.pi / nt)
    ph = (np.arange(npf) + 0.5) * (2 * math.pi / npf)
    total = 0.0
    for t in th:
        row = sum(_pdf(lib, mu, sig, wi, _dir(t, p)) for p in ph)
        total += row * math.sin(t)
    total *= (0.5 * math.pi / nt) * (2 * math.pi / npf)
    assert total == pytest.approx(1.0, abs=0.02)


def test_mip_level_rule(lib):
    f2 = lambda a, b: (C.c_float * 2)(a, b)
    lv = lambda t, dx, dy, L: lib.oracle_lean_level(t, f2(*dx), f2(*dy), L)
    for w in [0.004, 0.003, 0.02, 0.1, 0.9]:
        ref = min(math.ceil(11 + math.log2(2 * w)), 11)
        assert lv(1.0, (w, 0.0), (0.0, w * 0.5), 12) == ref, w
    assert lv(1.0, (2.0 ** -12, 0.0), (0.0, 0.0), 12) == 0          # width 2^-11: level 0 exactly
    assert lv(1.0, (1e-5, 0.0), (0.0, 1e-5), 12) == 11              # below one texel: coarsest level
    assert lv(4.0, (0.004, 0.0), (0.0, 0.0), 12) == min(math.ceil(11 + math.log2(2 * 4 * 0.004)), 11)
