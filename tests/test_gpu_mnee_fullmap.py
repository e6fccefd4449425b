"""GPU-vs-oracle parity of the MNEE-initialised solvers (manifold_ss.cpp:228-241,
manifold_ms.cpp:311-318) and of the glint estimator on the full-size 2048^2 scratch/bump map.

Bit-equal walks, estimates and counters, as the other parity tests.
"""
import numpy as np
import pytest

import oracle_lib
from smsamd import abi, scenes, workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    from smsamd import gpu as g
    g.lib()
    return g


def _eq_walks(go, oo, keys):
    for k in keys:
        assert np.array_equal(go[k], oo[k]), k


def _mirror_scene():
    sc = scenes.Scene()
    M = np.array([[0, 0, -1, 1.2], [0, 2.0, 0, 1.0], [2.0, 0, 0, 0.0]], np.float32)
    sc.add_rectangle(M, abi.BSDF_CONDUCTOR, flags=abi.FLAG_CAUSTIC_CASTER_SS, eta=0.0, k=1.0)
    sc.add_rectangle(scenes.plane_to_world((0, 0, 0), 10.0), abi.BSDF_DIFFUSE, flags=abi.FLAG_CAUSTIC_RECEIVER)
    light = sc.add_sphere((0.2, 2.0, 0.3), 0.1, abi.BSDF_DIFFUSE, reflectance=(0, 0, 0))
    sc.add_area_light(light, (20.0, 20.0, 20.0))
    sc.set_camera((0, 3, 6), (0, 0.5, 0), (0, 1, 0), 45.0, 64, 64)
    return sc


@pytest.mark.parametrize("which", ["dome", "sphere", "mirror", "pool"])
def test_ss_mnee_walks(gpu, which):
    if which == "pool":
        sc = scenes.pool(64, 64, n=128)
        w = workloads.ss_walk_inputs(sc, 6000, 3, extent=0.9)
        w["si_p"][:, 1] = -1.0
    elif which == "mirror":
        sc = _mirror_scene()
        w = workloads.ss_walk_inputs(sc, 6000, 3, extent=1.0)
    else:
        sc = scenes.sphere_dome() if which == "dome" else scenes.sphere_caustic()
        w = workloads.ss_walk_inputs(sc, 6000, 3, extent=0.9)
    for hv in (0, 1):
        cfg = abi.SmsConfig.default(mnee_init=1, halfvector_constraints=hv)
        ctx = gpu.Context(sc, cfg)
        orc = oracle_lib.Oracle(sc, cfg)
        go, oo = ctx.ss_walks(w), orc.ss_walks(w)
        print(which, hv, "mnee walks converged", oo["success"].mean(), flush=True)
        assert oo["init_valid"].mean() > 0.3
        _eq_walks(go, oo, ("success", "iterations", "final_shape", "init_valid", "p", "n", "p_init"))
        ctx.close()
        orc.close()


@pytest.mark.parametrize("biased,first_pass", [(0, 0), (0, 1), (1, 0)])
def test_ss_mnee_estimate_dome(gpu, biased, first_pass):
    """MNEE seeds make every walk of a query the same walk: the unbiased estimator's Bernoulli
    loop ends at its first trial, the biased one keeps one solution."""
    sc = scenes.sphere_dome()
    cfg = abi.SmsConfig.default(mnee_init=1, biased=biased, max_trials=8 if biased else 100000)
    ctx = gpu.Context(sc, cfg)
    orc = oracle_lib.Oracle(sc, cfg)
    q = workloads.plane_queries(sc, 3000, 2, extent=0.9)
    ctx.set_schedule(ss_first_pass=first_pass)
    rg, sg = ctx.ss_estimate(q, 7)
    ro, so = orc.ss_estimate(q, 7)
    print("gpu", sg, "oracle", so, flush=True)
    assert (ro[:, 0] > 0).mean() > 0.2
    assert np.array_equal(rg, ro, equal_nan=True)
    for k in ("walks", "walks_converged", "newton_iterations", "bernoulli_calls", "bernoulli_trials", "bernoulli_max"):
        assert sg[k] == so[k], k
    if not biased:
        assert so["bernoulli_trials"] == so["bernoulli_calls"]
    ctx.close()
    orc.close()


@pytest.mark.parametrize("biased", [0, 1])
def test_ms_mnee_walks_and_estimate(gpu, biased):
    sc = scenes.slab(width=64, height=64, n=256)
    cfg = abi.SmsConfig.default(bounces=2, mnee_init=1, biased=biased, max_trials=8 if biased else 100000)
    ctx = gpu.Context(sc, cfg)
    orc = oracle_lib.Oracle(sc, cfg)
    w = workloads.ms_walk_inputs(sc, 8000, 21, extent=0.35)
    go, oo = ctx.ms_walks(w), orc.ms_walks(w)
    print("ms mnee walks converged", oo["success"].mean(), flush=True)
    assert oo["success"].mean() > 0.05
    _eq_walks(go, oo, ("success", "seed_valid", "iterations", "n_vertices", "shapes", "p"))
    q = workloads.plane_queries(sc, 1500, 5, extent=0.35, shape=2)
    rg, sg = ctx.ms_estimate(q, 9)
    ro, so = orc.ms_estimate(q, 9)
    print("gpu", sg, "oracle", so, flush=True)
    assert (ro[:, 0] > 0).mean() > 0.02
    assert np.array_equal(rg, ro, equal_nan=True)
    for k in ("walks", "walks_converged", "newton_iterations", "bernoulli_calls", "bernoulli_trials"):
        assert sg[k] == so[k], k
    ctx.close()
    orc.close()


def test_glint_estimate_full_map_window(gpu):
    """The glint frame's own inputs -- the 2048^2 RGB8 map of 4096 scratches and Gaussian bumps,
    512x512 pixels at 4 spp -- restricted to a 64x64-pixel window: queries, estimates and counters
    of the GPU equal the oracle's."""
    sc = scenes.glint_plane(512, 512, nm_res=2048, n_scratches=4096, seed=1)
    cfg = abi.SmsConfig.default(max_trials=100000)
    ctx = gpu.Context(sc, cfg)
    orc = oracle_lib.Oracle(sc, cfg, threads=16)
    qo = orc.glint_queries(4, 1)
    qg = ctx.glint_queries(4, 1)
    assert np.array_equal(qg.view(np.uint8), qo.view(np.uint8))
    px, py = qo["pixel"] % 512, qo["pixel"] // 512
    sel = (px >= 224) & (px < 288) & (py >= 224) & (py < 288)
    q = np.ascontiguousarray(qo[sel])
    assert len(q) == 64 * 64 * 4
    rg, sg = ctx.glint_estimate(q, 1)
    ro, so = orc.glint_estimate(q, 1)
    print("gpu", sg, "oracle", so, flush=True)
    assert so["walks_converged"] > 100 and (ro[:, 0] > 0).sum() > 50
    assert np.array_equal(rg, ro, equal_nan=True)
    for k in ("walks", "walks_converged", "newton_iterations", "bernoulli_calls", "bernoulli_trials", "bernoulli_max"):
        assert sg[k] == so[k], k
    ctx.close()
    orc.close()
