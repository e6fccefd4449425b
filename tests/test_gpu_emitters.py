"""GPU-vs-oracle parity with directional emitters (src/emitters/directional.cpp): single-scatter
walks toward a fixed direction, the single-scatter estimator (persistent walker, lane per query,
biased), the multi-scatter estimator and the glint estimator under a directional light.  Bit-equal
results and counters, as the other parity tests."""
import numpy as np
import pytest

import oracle_lib
from smsamd import abi, scenes, workloads

pytestmark = pytest.mark.gpu

DIR = np.array([0.12, -1.0, 0.07]) / np.linalg.norm([0.12, -1.0, 0.07])
KEYS = ("walks", "walks_converged", "newton_iterations", "bernoulli_calls", "bernoulli_trials", "bernoulli_max")


@pytest.fixture(scope="module")
def gpu():
    from smsamd import gpu as g
    g.lib()
    return g


def _dome_dir():
    sc = scenes.Scene()
    sc.add_sphere((0.0, 0.0, 0.0), 1.0, abi.BSDF_DIELECTRIC, flags=abi.FLAG_CAUSTIC_CASTER_SS, eta=1.5)
    sc.add_rectangle(scenes.plane_to_world((0, 0, 0), 10.0), abi.BSDF_DIFFUSE, flags=abi.FLAG_CAUSTIC_RECEIVER)
    sc.add_directional_light(DIR, (3.0, 3.0, 3.0))
    sc.set_camera((0, 3, 6), (0, 0.5, 0), (0, 1, 0), 45.0, 64, 64)
    return sc


def _compare(rg, sg, ro, so):
    print("gpu", sg, "oracle", so, flush=True)
    assert np.array_equal(rg, ro, equal_nan=True)
    for k in KEYS:
        assert sg[k] == so[k], k


def test_ss_walks_toward_a_fixed_direction(gpu):
    sc = _dome_dir()
    w = workloads.ss_walk_inputs(sc, 6000, 5, light_shape=0, extent=0.9)
    w["ei"]["kind"] = abi.EI_DIRECTIONAL
    w["ei"]["d"] = -DIR
    w["ei"]["p"] = w["si_p"] - 20.0 * DIR
    for hv in (0, 1):
        cfg = abi.SmsConfig.default(halfvector_constraints=hv)
        ctx = gpu.Context(sc, cfg)
        orc = oracle_lib.Oracle(sc, cfg)
        go, oo = ctx.ss_walks(w), orc.ss_walks(w)
        assert oo["success"].mean() > 0.1
        for k in ("success", "iterations", "final_shape", "init_valid", "p", "n", "p_init"):
            assert np.array_equal(go[k], oo[k]), k
        ctx.close()
        orc.close()


@pytest.mark.parametrize("biased,first_pass", [(0, 0), (0, 1), (1, 0)])
def test_ss_estimate_directional(gpu, biased, first_pass):
    sc = _dome_dir()
    cfg = abi.SmsConfig.default(biased=biased, max_trials=16 if biased else 100000)
    ctx = gpu.Context(sc, cfg)
    orc = oracle_lib.Oracle(sc, cfg)
    ctx.set_schedule(ss_first_pass=first_pass)
    q = workloads.plane_queries(sc, 3000, 2, extent=0.9)
    rg, sg = ctx.ss_estimate(q, 7)
    ro, so = orc.ss_estimate(q, 7)
    assert (ro[:, 0] > 0).mean() > 0.2
    _compare(rg, sg, ro, so)
    ctx.close()
    orc.close()


def test_ms_estimate_directional(gpu):
    sc = scenes.slab(width=64, height=64, n=256)
    sc.emitters = []
    sc.shapes[3]["emitter"] = -1
    sc.add_directional_light(DIR, (2.0, 2.0, 2.0))
    cfg = abi.SmsConfig.default(bounces=2, max_trials=100000)
    ctx = gpu.Context(sc, cfg)
    orc = oracle_lib.Oracle(sc, cfg)
    q = workloads.plane_queries(sc, 1500, 5, extent=0.35, shape=2)
    rg, sg = ctx.ms_estimate(q, 9)
    ro, so = orc.ms_estimate(q, 9)
    assert (ro[:, 0] > 0).mean() > 0.02
    _compare(rg, sg, ro, so)
    ctx.close()
    orc.close()


def test_glint_estimate_directional(gpu):
    sc = scenes.glint_plane(64, 48, nm_res=256, n_scratches=64, seed=3, point_light=False)
    sc.add_directional_light((-0.3, -1.0, 0.35), (2.0, 2.0, 2.0))
    cfg = abi.SmsConfig.default(max_trials=100000)
    ctx = gpu.Context(sc, cfg)
    orc = oracle_lib.Oracle(sc, cfg)
    q = orc.glint_queries(2, 5)
    rg, sg = ctx.glint_estimate(q, 9)
    ro, so = orc.glint_estimate(q, 9)
    _compare(rg, sg, ro, so)
    ctx.close()
    orc.close()
