"""GPU-vs-oracle parity of the HIP kernels, called through the C ABI.

The kernels follow the oracle's arithmetic contract (same op order, no FP contraction,
the same Cephes transcendentals and PCG32 streams), so walks must agree exactly: identical
converge/fail flags and iteration counts, solution points within 1e-4 relative world-space
distance (expected bit-equal).  Estimator outputs use the same per-(query, walk) random
streams and must agree to 1e-5 relative.
"""
import numpy as np
import pytest

import oracle_lib
from smsamd import abi, scenes, workloads

pytestmark = pytest.mark.gpu

REL_TOL = 1e-4


@pytest.fixture(scope="module")
def gpu():
    from smsamd import gpu as g
    g.lib()
    return g


@pytest.fixture(scope="module")
def oracle_mod():
    import oracle_lib
    oracle_lib.build()
    return oracle_lib


def _rel_dist(a, b):
    d = np.linalg.norm(a.astype(np.float64) - b.astype(np.float64), axis=-1)
    return d / np.maximum(np.linalg.norm(b.astype(np.float64), axis=-1), 1e-6)


def _check_ss_walks(go, oo):
    print(f"walks={len(go)} success={oo['success'].mean():.3f} iters={oo['iterations'].mean():.2f}", flush=True)
    assert np.array_equal(go["success"], oo["success"]), "converge/fail flags differ"
    assert np.array_equal(go["init_valid"], oo["init_valid"])
    assert np.array_equal(go["iterations"], oo["iterations"])
    assert np.array_equal(go["final_shape"], oo["final_shape"])
    ok = oo["success"] == 1
    assert _rel_dist(go["p"][ok], oo["p"][ok]).max(initial=0) <= REL_TOL
    assert np.array_equal(go["p"], oo["p"]), "solution vertices not bit-equal"


@pytest.mark.parametrize("scene_fn,extent", [(scenes.sphere_caustic, 1.2), (scenes.sphere_dome, 0.9)])
def test_ss_walks_sphere(gpu, oracle_mod, scene_fn, extent):
    sc = scene_fn()
    cfg = abi.SmsConfig.default()
    ctx = gpu.Context(sc, cfg)
    orc = oracle_mod.Oracle(sc, cfg)
    w = workloads.ss_walk_inputs(sc, 20000, 11, extent=extent)
    _check_ss_walks(ctx.ss_walks(w), orc.ss_walks(w))
    # explicit seeds (p_start override, manifold_ss.cpp:224-226)
    w2 = workloads.ss_walk_inputs(sc, 4000, 12, extent=extent)
    rng = np.random.default_rng(3)
    w2["p_start"] = (np.array([0.0, 1.5 if scene_fn is scenes.sphere_caustic else 0.0, 0.0])
                     + workloads.uniform_sphere(rng.random(4000), rng.random(4000))).astype(np.float32)
    _check_ss_walks(ctx.ss_walks(w2), orc.ss_walks(w2))
    # half-vector constraints
    cfg_h = abi.SmsConfig.default(halfvector_constraints=1)
    ctx.set_config(cfg_h)
    orc.set_config(cfg_h)
    _check_ss_walks(ctx.ss_walks(w), orc.ss_walks(w))
    ctx.close()
    orc.close()


def test_glint_walks(gpu, oracle_mod):
    sc = scenes.glint_plane(64, 64, nm_res=256, n_scratches=64, seed=3)
    cfg = abi.SmsConfig.default(max_trials=100000)
    ctx = gpu.Context(sc, cfg)
    orc = oracle_mod.Oracle(sc, cfg)
    for fp in (2.0, 6.0):
        w = workloads.glint_walk_inputs(sc, 30000, 21, footprint_texels=fp)
        go, oo = ctx.glint_walks(w), orc.glint_walks(w)
        print(f"glint walks fp={fp} success={oo['success'].mean():.4f} iters={oo['iterations'].mean():.2f}", flush=True)
        assert np.array_equal(go["success"], oo["success"])
        assert np.array_equal(go["iterations"], oo["iterations"])
        assert np.array_equal(go["uv"], oo["uv"])
        assert np.array_equal(go["uv_init"], oo["uv_init"])
    ctx.close()
    orc.close()


@pytest.mark.parametrize("max_trials", [1, 40, 128])
def test_glint_biased_estimate(gpu, oracle_mod, max_trials):
    """Biased glint estimator: trial sets of max_trials walks, uv de-duplication in walk order."""
    sc = scenes.glint_plane(64, 48, nm_res=256, n_scratches=64, seed=3)
    cfg = abi.SmsConfig.default(biased=1, max_trials=max_trials)
    ctx = gpu.Context(sc, cfg)
    orc = oracle_mod.Oracle(sc, cfg)
    q = orc.glint_queries(2, 5)
    rg, sg = ctx.glint_estimate(q, 9)
    ro, so = orc.glint_estimate(q, 9)
    print(max_trials, "gpu", sg, "oracle", so, flush=True)
    assert (ro[:, 0] > 0).sum() > 20
    assert np.array_equal(rg, ro, equal_nan=True)
    for k in ("walks", "walks_converged", "newton_iterations", "bernoulli_calls"):
        assert sg[k] == so[k], k
    if max_trials == 128:
        with pytest.raises(gpu.SmsError):
            ctx.set_config(abi.SmsConfig.default(biased=1, max_trials=129))
            ctx.glint_estimate(q, 9)
    ctx.close()
    orc.close()


@pytest.mark.parametrize("biased", [0, 1])
def test_glint_area_light_estimate(gpu, oracle_mod, biased):
    """Point light plus a spherical area light seen in the mirror: the SMS strategy samples
    either emitter, and the BSDF-sampling strategy (manifold_glints.cpp:183-220) adds the area
    light where a footprint's mirror direction reaches it."""
    area = ((0.0, 1.1, -1.05), 0.35, (4.0, 3.0, 2.0))
    sc = scenes.glint_plane(64, 48, nm_res=256, n_scratches=64, seed=3, area_light=area)
    cfg = abi.SmsConfig.default(biased=biased, max_trials=32 if biased else 100000)
    ctx = gpu.Context(sc, cfg)
    orc = oracle_mod.Oracle(sc, cfg)
    q = orc.glint_queries(2, 5)
    rg, sg = ctx.glint_estimate(q, 9)
    ro, so = orc.glint_estimate(q, 9)
    print(biased, "gpu", sg, "oracle", so, flush=True)
    assert np.array_equal(rg, ro, equal_nan=True)
    for k in ("walks", "walks_converged", "newton_iterations", "bernoulli_calls", "bernoulli_trials", "bernoulli_max"):
        assert sg[k] == so[k], k
    # without the BSDF strategy's share the result changes where the mirror sees the light
    cfg2 = abi.SmsConfig.default(biased=biased, max_trials=cfg.max_trials, sms_strategy_only=1)
    ctx.set_config(cfg2)
    orc.set_config(cfg2)
    rg2, _ = ctx.glint_estimate(q, 9)
    ro2, _ = orc.glint_estimate(q, 9)
    assert np.array_equal(rg2, ro2, equal_nan=True)
    assert (rg2[:, 0] != rg[:, 0]).sum() > 20
    ctx.close()
    orc.close()


@pytest.mark.parametrize("biased,encoded", [(0, False), (1, False), (0, True)])
def test_ss_twostage(gpu, oracle_mod, biased, encoded):
    """Two-stage solver (manifold_ss.cpp:251-270) on a normal-mapped mirror wall (frames from
    slopes, or from decoded RGB with the LEAN mean of the stored RGB): walks with explicit LEAN
    samples and whole estimates (the lane-per-query pass) equal the oracle."""
    sc = scenes.mirror_wall(nm_res=64, n_scratches=16, seed=4)
    if encoded:
        sc.encoded_normalmaps.add(0)
    cfg = abi.SmsConfig.default(max_trials=32 if biased else 100000, biased=biased, twostage=1)
    ctx = gpu.Context(sc, cfg)
    orc = oracle_mod.Oracle(sc, cfg)
    w = workloads.ss_walk_inputs(sc, 4000, 7, caster=1, light_shape=2, extent=0.9)
    w["lean_sample"] = np.random.default_rng(8).random((len(w), 2)).astype(np.float32)
    go, oo = ctx.ss_walks(w), orc.ss_walks(w)
    print("two-stage walks converged", oo["success"].mean(), flush=True)
    assert oo["success"].sum() > 200
    for k in ("success", "iterations", "final_shape", "init_valid", "p", "n", "p_init"):
        assert np.array_equal(go[k], oo[k]), k
    q = workloads.plane_queries(sc, 2000, 5, extent=0.9, shape=0)
    rg, sg = ctx.ss_estimate(q, 3)
    ro, so = orc.ss_estimate(q, 3)
    print(biased, "gpu", sg, "oracle", so, flush=True)
    assert so["walks_converged"] > 100 and (ro[:, 0] > 0).sum() > 20
    assert np.array_equal(rg, ro, equal_nan=True)
    for k in ("walks", "walks_converged", "newton_iterations", "bernoulli_calls", "bernoulli_trials", "bernoulli_max"):
        assert sg[k] == so[k], k
    ctx.close()
    orc.close()


def test_glint_queries_and_estimate(gpu, oracle_mod):
    sc = scenes.glint_plane(64, 48, nm_res=256, n_scratches=64, seed=3)
    cfg = abi.SmsConfig.default(max_trials=100000)
    ctx = gpu.Context(sc, cfg)
    orc = oracle_mod.Oracle(sc, cfg)
    qg, qo = ctx.glint_queries(2, 5), orc.glint_queries(2, 5)
    assert np.array_equal(qg.view(np.uint8), qo.view(np.uint8)), "glint queries differ"
    rg, sg = ctx.glint_estimate(qo, 9)
    ro, so = orc.glint_estimate(qo, 9)
    print("gpu", sg, "oracle", so, flush=True)
    assert so["bernoulli_calls"] > 20
    assert np.array_equal(rg, ro, equal_nan=True)
    for k in ("walks", "walks_converged", "newton_iterations", "bernoulli_calls", "bernoulli_trials", "bernoulli_max"):
        assert sg[k] == so[k], k
    # Every schedule splits the trial chains differently between the grouped lane pass, the
    # one-wavefront rounds, the multi-wavefront stage and the sequential tail; the result and
    # the counters must not change.  tail_rounds = 0 sends every chain left after the lane pass
    # to the multi-wavefront stage; long_chains / long_trials small force chains on to the
    # sequential tail.
    assert so["bernoulli_max"] > 8
    schedules = [dict(trial_group=1, lane_trials=1, tail_rounds=0),
                 dict(trial_group=1, lane_trials=1, tail_rounds=0, long_chains=8, long_trials=64),
                 dict(trial_group=2, lane_trials=2, tail_rounds=0, long_chains=1024, long_trials=128),
                 dict(trial_group=2, lane_trials=2, tail_rounds=1, long_chains=1024, long_trials=32768),
                 dict(trial_group=4, lane_trials=8, tail_rounds=2),
                 dict(trial_group=8, lane_trials=8, tail_rounds=1),
                 dict(trial_group=16, lane_trials=16, tail_rounds=3),
                 dict(trial_group=8, lane_trials=8, tail_rounds=2, blocks_per_cu=1)]
    for sch in schedules:
        ctx.set_schedule(**sch)
        rg2, sg2 = ctx.glint_estimate(qo, 9)
        print("schedule", sch, sg2, flush=True)
        assert np.array_equal(rg2, ro, equal_nan=True), sch
        for k in ("walks", "walks_converged", "newton_iterations", "bernoulli_trials", "bernoulli_max"):
            assert sg2[k] == so[k], (k, sch)
        if sch["lane_trials"] < 8:
            assert sg2["overflow_queries"] > 0
    # trial cap: inv_prob = 0 past max_trials (manifold_glints.cpp:96-103)
    for cap, group, lane_trials in [(3, 2, 2), (100, 1, 1), (100, 4, 4)]:
        cfg2 = abi.SmsConfig.default(max_trials=cap)
        ctx.set_config(cfg2)
        orc.set_config(cfg2)
        ctx.set_schedule(trial_group=group, lane_trials=lane_trials, tail_rounds=0, long_chains=4, long_trials=64)
        rg3, sg3 = ctx.glint_estimate(qo, 9)
        ro3, so3 = orc.glint_estimate(qo, 9)
        assert np.array_equal(rg3, ro3, equal_nan=True), cap
        for k in ("walks", "newton_iterations", "bernoulli_trials", "bernoulli_max"):
            assert sg3[k] == so3[k], (k, cap)
    ctx.close()
    orc.close()


@pytest.mark.parametrize("biased", [1, 0])
def test_ss_estimate_dome(gpu, oracle_mod, biased):
    sc = scenes.sphere_dome()
    cfg = abi.SmsConfig.default(biased=biased, max_trials=64 if biased else 100000)
    ctx = gpu.Context(sc, cfg)
    orc = oracle_mod.Oracle(sc, cfg)
    q = workloads.plane_queries(sc, 3000, 2, extent=0.9)
    rg, sg = ctx.ss_estimate(q, 7)
    ro, so = orc.ss_estimate(q, 7)
    print("gpu", sg, "oracle", so, flush=True)
    assert (ro[:, 0] > 0).mean() > 0.2
    np.testing.assert_allclose(rg, ro, rtol=1e-5, atol=0)
    assert np.array_equal(rg, ro, equal_nan=True)
    for k in ("walks", "walks_converged", "newton_iterations"):
        assert sg[k] == so[k], k
    if not biased:
        for first_pass in (0, 1):
            for lane_trials in (1, 8, 64):
                ctx.set_schedule(lane_trials=lane_trials, ss_first_pass=first_pass)
                rg2, sg2 = ctx.ss_estimate(q, 7)
                if lane_trials == 1:
                    assert sg2["overflow_queries"] > 0
                assert np.array_equal(rg2, ro, equal_nan=True), (first_pass, lane_trials)
                for k in ("walks", "walks_converged", "newton_iterations", "bernoulli_calls", "bernoulli_trials",
                          "bernoulli_max"):
                    assert sg2[k] == so[k], (k, first_pass, lane_trials)
    ctx.close()
    orc.close()


def test_ss_estimate_sphere_above_plane_is_zero(gpu, oracle_mod):
    """Sphere wholly above the plane: every refraction solution is occluded from the light by
    the sphere itself (manifold_ss.cpp:88-102), on both sides."""
    sc = scenes.sphere_caustic()
    cfg = abi.SmsConfig.default(biased=1, max_trials=64)
    ctx = gpu.Context(sc, cfg)
    orc = oracle_mod.Oracle(sc, cfg)
    q = workloads.plane_queries(sc, 1000, 4, extent=1.2)
    rg, sg = ctx.ss_estimate(q, 3)
    ro, so = orc.ss_estimate(q, 3)
    assert so["walks_converged"] > 0
    assert np.array_equal(rg, ro, equal_nan=True)
    assert not ro.any()
    ctx.close()
    orc.close()


def _pool_walks(sc, n, seed):
    rng = np.random.default_rng(seed)
    w = np.zeros(n, abi.SS_WALK_IN)
    w["si_p"] = np.stack([rng.uniform(-0.9, 0.9, n), np.full(n, -1.0), rng.uniform(-0.9, 0.9, n)], -1)
    w["shape"] = 0
    ls = sc.shapes[2]
    w["ei"]["p"] = np.asarray(ls["center"]) + ls["radius"] * workloads.uniform_sphere(rng.random(n), rng.random(n))
    w["ei"]["kind"] = abi.EI_AREA
    w["n_offset"] = (0, 0, 1)
    w["sample"] = rng.random((n, 2))
    return w


def test_pool_walks_through_bvh(gpu, oracle_mod):
    """524,288-triangle Gerstner pool: every reprojection traverses the GPU BVH; the oracle
    traverses its own BVH, so agreement also checks closest-hit equivalence."""
    sc = scenes.pool(n=512)
    cfg = abi.SmsConfig.default(max_trials=100000)
    ctx = gpu.Context(sc, cfg)
    orc = oracle_mod.Oracle(sc, cfg)
    w = _pool_walks(sc, 20000, 5)
    _check_ss_walks(ctx.ss_walks(w), orc.ss_walks(w))
    q = workloads.plane_queries(sc, 4000, 2, extent=0.9, shape=1)
    q["p"][:, 1] = -1.0
    ro, so = orc.ss_estimate(q, 7)
    for first_pass in (0, 1):
        ctx.set_schedule(ss_first_pass=first_pass)
        rg, sg = ctx.ss_estimate(q, 7)
        print("gpu", first_pass, sg, "oracle", so, flush=True)
        assert (ro[:, 0] > 0).mean() > 0.05
        assert np.array_equal(rg, ro, equal_nan=True)
        for k in ("walks", "walks_converged", "newton_iterations", "bernoulli_calls", "bernoulli_trials"):
            assert sg[k] == so[k], (k, first_pass)
    ctx.close()
    orc.close()


def test_nonfinite_walk_matches_oracle(gpu, oracle_mod):
    """The pool walk whose Newton step turns NaN (tests/golden/pool_nan_walk.npy) matches the
    oracle.  NaN reprojection rays skip the BVH (a full traversal by one wavefront took about
    0.4 s per ray, 8 s for the walk); the time is printed, and the test runner's per-test
    timeout, not a wall-clock assertion, stands guard over a regression."""
    import os
    import time
    w = oracle_lib.upgrade_ss_walk_in(
        np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pool_nan_walk.npy")))
    sc = scenes.pool(512, 512)
    cfg = abi.SmsConfig.default(max_trials=1000000, uniqueness_threshold=1e-5)
    ctx = gpu.Context(sc, cfg)
    orc = oracle_mod.Oracle(sc, cfg)
    ctx.ss_walks(w)                      # first launch of the kernel
    t = time.perf_counter()
    go = ctx.ss_walks(w)
    dt = time.perf_counter() - t
    print(f"nan walk: {dt * 1e3:.2f} ms", flush=True)
    _check_ss_walks(go, orc.ss_walks(w))
    ctx.close()
    orc.close()


def _check_ms_walks(go, oo):
    print(f"ms walks={len(go)} seed={oo['seed_valid'].mean():.3f} success={oo['success'].mean():.3f} "
          f"iters={oo['iterations'].mean():.2f}", flush=True)
    for k in ("success", "seed_valid", "iterations", "n_vertices", "shapes"):
        assert np.array_equal(go[k], oo[k]), k
    ok = oo["success"] == 1
    assert _rel_dist(go["p"][ok].reshape(-1, 3), oo["p"][ok].reshape(-1, 3)).max(initial=0) <= REL_TOL
    assert np.array_equal(go["p"], oo["p"]), "chain vertices not bit-equal"


@pytest.fixture(scope="module")
def slab_scene():
    return scenes.slab(width=64, height=64, n=256)


@pytest.mark.parametrize("bounces,halfvector", [(2, 0), (2, 1), (1, 0), (3, 0)])
def test_ms_walks_slab(gpu, oracle_mod, slab_scene, bounces, halfvector):
    """Two-vertex chains through the displaced slab (bottom face -> top face); bounces=1 and 3
    exercise chains that stop early or run out of specular surfaces, and the 4-vertex kernel."""
    cfg = abi.SmsConfig.default(bounces=bounces, halfvector_constraints=halfvector)
    ctx = gpu.Context(slab_scene, cfg)
    orc = oracle_mod.Oracle(slab_scene, cfg)
    w = workloads.ms_walk_inputs(slab_scene, 20000, 21, extent=0.35)
    oo = orc.ms_walks(w)
    _check_ms_walks(ctx.ms_walks(w), oo)
    if bounces == 2 and not halfvector:
        assert oo["success"].mean() > 0.02
        w2 = workloads.ms_walk_inputs(slab_scene, 20000, 22, extent=0.35, first_path=False, seed_shapes=(1, 0, -1, -1))
        _check_ms_walks(ctx.ms_walks(w2), orc.ms_walks(w2))
    ctx.close()
    orc.close()


@pytest.mark.parametrize("biased", [1, 0])
def test_ms_estimate_slab(gpu, oracle_mod, slab_scene, biased):
    cfg = abi.SmsConfig.default(bounces=2, biased=biased, max_trials=32 if biased else 100000)
    ctx = gpu.Context(slab_scene, cfg)
    orc = oracle_mod.Oracle(slab_scene, cfg)
    q = workloads.plane_queries(slab_scene, 1500, 5, extent=0.35, shape=2)
    rg, sg = ctx.ms_estimate(q, 9)
    ro, so = orc.ms_estimate(q, 9)
    print("gpu", sg, "oracle", so, flush=True)
    assert (ro[:, 0] > 0).mean() > 0.02
    assert np.array_equal(rg, ro, equal_nan=True)
    for k in ("walks", "walks_converged", "newton_iterations"):
        assert sg[k] == so[k], k
    if not biased:
        for k in ("bernoulli_calls", "bernoulli_trials", "bernoulli_max"):
            assert sg[k] == so[k], k
        # persistent first pass (0) and lane per query (1), long and short lane trial runs
        for first_pass in (0, 1):
            for lane_trials in (256, 8, 1):
                ctx.set_schedule(lane_trials=lane_trials, ss_first_pass=first_pass)
                rg2, sg2 = ctx.ms_estimate(q, 9)
                print("first pass", first_pass, "lane_trials", lane_trials, sg2, flush=True)
                if lane_trials == 1:
                    assert sg2["overflow_queries"] > 0
                assert np.array_equal(rg2, ro, equal_nan=True), (first_pass, lane_trials)
                for k in ("walks", "walks_converged", "newton_iterations", "bernoulli_calls", "bernoulli_trials",
                          "bernoulli_max"):
                    assert sg2[k] == so[k], (k, first_pass, lane_trials)
    ctx.close()
    orc.close()


def test_ms_rejects_unbuilt_chain_lengths(gpu, slab_scene):
    ctx = gpu.Context(slab_scene, abi.SmsConfig.default(bounces=5))
    with pytest.raises(gpu.SmsError):
        ctx.ms_walks(workloads.ms_walk_inputs(slab_scene, 4, 1))
    with pytest.raises(gpu.SmsError):
        ctx.ms_estimate(workloads.plane_queries(slab_scene, 4, 1, extent=0.3, shape=2), 1)
    ctx.close()


def test_empty_batches(gpu):
    sc = scenes.sphere_dome()
    ctx = gpu.Context(sc, abi.SmsConfig.default())
    assert len(ctx.ss_walks(np.zeros(0, abi.SS_WALK_IN))) == 0
    rg, st = ctx.ss_estimate(np.zeros(0, abi.SS_QUERY), 1)
    assert rg.shape == (0, 3)
    ctx.close()
