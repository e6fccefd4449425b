"""Pins of the oracle's building blocks against known answers and independent math.

Known answers are taken from the reference's own tests (PCG32 from the published pcg32 demo
output; warps from src/libcore/tests/test_warp.py:121-130; Fresnel from
src/librender/tests/test_fresnel.py:7-40); transcendentals are checked against numpy's
float64 libm.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_lib


@pytest.fixture(scope="module")
def lib():
    oracle_lib.build()
    return oracle_lib.load()


def test_pcg32_known_answer(lib):
    # pcg32_srandom_r(&rng, 42u, 54u) -> first six outputs of the reference pcg32 demo
    out = np.zeros(6, np.uint32)
    lib.oracle_pcg32_first(42, 54, out.ctypes.data, 6)
    assert [hex(x) for x in out] == ["0xa15c02b7", "0x7b47f409", "0xba1d3330", "0x83d2f293", "0xbfa4784b", "0xcbed606e"]


def _ulp_err(got, ref):
    got = np.asarray(got, np.float64)
    ulp = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
    return np.max(np.abs(got - ref) / np.maximum(ulp, np.finfo(np.float32).tiny))


def test_acos_against_libm(lib):
    xs = np.linspace(-1, 1, 20001, dtype=np.float32)
    got = np.array([lib.oracle_acosf(float(x)) for x in xs])
    ref = np.arccos(xs.astype(np.float64))
    assert np.max(np.abs(got - ref)) < 5e-7
    assert _ulp_err(got[np.abs(xs) < 0.999], ref[np.abs(xs) < 0.999]) <= 4
    # clamping of safe_acos
    assert lib.oracle_acosf(1.0000001) == 0.0
    assert abs(lib.oracle_acosf(-1.0000001) - math.pi) < 1e-6


def test_atan2_against_libm(lib):
    rng = np.random.default_rng(0)
    ys = rng.normal(size=5000).astype(np.float32)
    xs = rng.normal(size=5000).astype(np.float32)
    got = np.array([lib.oracle_atan2f(float(y), float(x)) for y, x in zip(ys, xs)])
    ref = np.arctan2(ys.astype(np.float64), xs.astype(np.float64))
    assert np.max(np.abs(got - ref)) < 1e-6
    assert _ulp_err(got, ref) <= 4
    for y, x in [(0.0, 1.0), (1.0, 0.0), (-1.0, 0.0), (0.0, -1.0), (-0.0, -1.0)]:
        assert abs(lib.oracle_atan2f(y, x) - math.atan2(y, x)) < 1e-6


def test_sincos_against_libm(lib):
    xs = np.linspace(0, 2 * np.pi, 10001, dtype=np.float32)
    s, c = C.c_float(), C.c_float()
    got_s, got_c = [], []
    for x in xs:
        lib.oracle_sincosf(float(x), C.byref(s), C.byref(c))
        got_s.append(s.value)
        got_c.append(c.value)
    assert np.max(np.abs(np.array(got_s) - np.sin(xs.astype(np.float64)))) < 3e-7
    assert np.max(np.abs(np.array(got_c) - np.cos(xs.astype(np.float64)))) < 3e-7


def test_square_to_uniform_sphere_known_answers(lib):
    out = np.zeros(3, np.float32)
    for uv, ref in [((0, 0), (0, 0, 1)), ((0, 1), (0, 0, -1)), ((0.5, 0.5), (-1, 0, 0))]:
        lib.oracle_square_to_uniform_sphere(uv[0], uv[1], out.ctypes.data)
        np.testing.assert_allclose(out, ref, atol=1e-6)


def test_fresnel_known_answers(lib):
    f = lib.oracle_fresnel
    assert abs(f(1.0, 1.5) - 0.04) < 1e-6
    assert abs(f(-1.0, 1.5) - 0.04) < 1e-6
    assert abs(f(1.0, 1 / 1.5) - 0.04) < 1e-6
    assert abs(f(-1.0, 1 / 1.5) - 0.04) < 1e-6
    assert f(0.0, 1.5) == 1.0
    assert f(0.0, 1 / 1.5) == 1.0
    c45 = math.cos(math.radians(45))
    assert abs(f(c45, 1.5) - 0.5 * (0.09201336304552442 ** 2 + 0.3033370452904235 ** 2)) < 1e-6
    assert abs(f(c45, 1 / 1.5) - 1.0) < 1e-6            # total internal reflection
    c10 = math.cos(math.radians(10))
    assert abs(f(c10, 1 / 1.5) - 0.5 * (0.19046797197779405 ** 2 + 0.20949431963852014 ** 2)) < 1e-6


def _fresnel_conductor_complex(c, eta):
    # unpolarized Fresnel reflectance of a complex IOR from the s/p amplitudes
    s2 = 1 - c * c
    ct = np.sqrt(1 - s2 / (eta * eta) + 0j)
    rs = (c - eta * ct) / (c + eta * ct)
    rp = (eta * c - ct) / (eta * c + ct)
    return 0.5 * (abs(rs) ** 2 + abs(rp) ** 2)


@pytest.mark.parametrize("eta", [complex(0.2, 3.0), complex(1.3, 0.8), complex(0.0, 1.0)])
def test_fresnel_conductor_against_complex_formula(lib, eta):
    for c in np.linspace(0.05, 1.0, 40):
        got = lib.oracle_fresnel_conductor(float(c), eta.real, eta.imag)
        ref = _fresnel_conductor_complex(c, eta)
        assert abs(got - ref) < 2e-5 * max(1.0, ref), (c, got, ref)
