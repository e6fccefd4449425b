"""GPU-vs-oracle parity of rough single-scatter SMS (manifold_ss.cpp:61-70, :649-667) on rough
conductors: walks with Beckmann offset normals (both constraints), and the unbiased (persistent
walker, lane per query) and biased estimators, whose offset normal per (query, caster) comes from
the query's stream and divides the estimate by its pdf.  Bit-equal results and counters."""
import numpy as np
import pytest

import oracle_lib
from smsamd import abi, scenes, workloads
from test_oracle_rough import rough_mirror

pytestmark = pytest.mark.gpu
KEYS = ("walks", "walks_converged", "newton_iterations", "bernoulli_calls", "bernoulli_trials", "bernoulli_max")


@pytest.fixture(scope="module")
def gpu():
    from smsamd import gpu as g
    g.lib()
    return g


def rough_ball():
    """Rough conductor ball (alpha 0.15) over the floor, sphere light: glossy reflective caustics."""
    sc = scenes.Scene()
    sc.add_sphere((0.0, 0.6, 0.0), 0.5, abi.BSDF_CONDUCTOR, flags=abi.FLAG_CAUSTIC_CASTER_SS, eta=0.2, k=3.0,
                  roughness=0.15)
    sc.add_rectangle(scenes.plane_to_world((0, 0, 0), 10.0), abi.BSDF_DIFFUSE, flags=abi.FLAG_CAUSTIC_RECEIVER)
    light = sc.add_sphere((0.9, 1.8, 0.2), 0.1, abi.BSDF_DIFFUSE, reflectance=(0, 0, 0))
    sc.add_area_light(light, (50.0, 50.0, 50.0))
    sc.set_camera((0, 3, 6), (0, 0.5, 0), (0, 1, 0), 45.0, 64, 64)
    return sc


@pytest.mark.parametrize("which", ["mirror", "ball"])
def test_rough_walks(gpu, which):
    sc = rough_mirror() if which == "mirror" else rough_ball()
    w = workloads.ss_walk_inputs(sc, 6000, 4, caster=0, light_shape=2, extent=1.0)
    rng = np.random.default_rng(3)
    th = np.arctan(0.2 * np.sqrt(-np.log(1 - rng.random(len(w)))))
    ph = 2 * np.pi * rng.random(len(w))
    w["n_offset"] = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], -1)
    for hv in (0, 1):
        cfg = abi.SmsConfig.default(halfvector_constraints=hv)
        ctx = gpu.Context(sc, cfg)
        orc = oracle_lib.Oracle(sc, cfg)
        go, oo = ctx.ss_walks(w), orc.ss_walks(w)
        print(which, hv, "converged", oo["success"].mean(), flush=True)
        assert oo["success"].mean() > 0.05
        for k in ("success", "iterations", "final_shape", "init_valid", "p", "n", "p_init"):
            assert np.array_equal(go[k], oo[k]), k
        ctx.close()
        orc.close()


@pytest.mark.parametrize("which,biased,first_pass", [("mirror", 0, 0), ("mirror", 0, 1), ("mirror", 1, 0),
                                                     ("ball", 0, 0), ("ball", 1, 0)])
def test_rough_estimate(gpu, which, biased, first_pass):
    sc = rough_mirror() if which == "mirror" else rough_ball()
    cfg = abi.SmsConfig.default(biased=biased, max_trials=16 if biased else 100000)
    ctx = gpu.Context(sc, cfg)
    orc = oracle_lib.Oracle(sc, cfg)
    ctx.set_schedule(ss_first_pass=first_pass)
    q = workloads.plane_queries(sc, 3000, 2, extent=1.0)
    rg, sg = ctx.ss_estimate(q, 7)
    ro, so = orc.ss_estimate(q, 7)
    print("gpu", sg, "oracle", so, flush=True)
    assert (ro[:, 0] > 0).mean() > 0.05
    assert np.array_equal(rg, ro, equal_nan=True)
    for k in KEYS:
        assert sg[k] == so[k], k
    ctx.close()
    orc.close()


def test_rough_dielectric_is_refused(gpu):
    sc = scenes.Scene()
    sc.add_sphere((0.0, 1.0, 0.0), 0.5, abi.BSDF_DIELECTRIC, flags=abi.FLAG_CAUSTIC_CASTER_SS, eta=1.5, roughness=0.1)
    sc.add_rectangle(scenes.plane_to_world((0, 0, 0), 10.0), abi.BSDF_DIFFUSE, flags=abi.FLAG_CAUSTIC_RECEIVER)
    with pytest.raises(gpu.SmsError):
        gpu.Context(sc, abi.SmsConfig.default())
