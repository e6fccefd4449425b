"""Path integrators and film on the GPU, through the C ABI: per-sample radiance, film positions
and SMS statistics bit-equal to the oracle's (path_sms_ss and path_filtered_ss, caustic term and
whole image, unbiased and MNEE configurations, passes of more than one estimator batch); the
film's weighted sums bit-equal; refused scenes and arguments."""
import numpy as np
import pytest

from smsamd import abi, scenes

pytestmark = pytest.mark.gpu


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


def _compare(ctx, orc, pc, spp, seed):
    rg, pg, sg = ctx.path_samples(pc, spp, seed)
    ro, po, so = orc.path_samples(pc, spp, seed)
    print(f"integrator {pc.integrator} caustics_only {pc.caustics_only} spp {spp}: mean {ro.mean():.6g}, "
          f"nonzero {(ro[:, 0] != 0).mean():.3f}, stats {so}", flush=True)
    assert np.array_equal(pg, po)
    assert np.array_equal(rg, ro, equal_nan=True)
    # the estimator's counters (overflow and tail counts describe the GPU's schedule)
    for k in ("walks", "walks_converged", "newton_iterations", "bernoulli_calls", "bernoulli_trials", "bernoulli_max"):
        assert sg[k] == so[k], k
    return ro


@pytest.mark.parametrize("integrator", [abi.PATH_SMS_SS, abi.PATH_FILTERED_SS])
@pytest.mark.parametrize("caustics_only", [0, 1])
def test_path_samples_dome(gpu, oracle_mod, integrator, caustics_only):
    sc = scenes.sphere_dome(32, 24, light_radius=0.5, light_radiance=4.0)
    cfg = abi.SmsConfig.default(max_trials=100000)
    ctx = gpu.Context(sc, cfg)
    orc = oracle_mod.Oracle(sc, cfg)
    for max_depth in (6, -1):
        pc = abi.SmsPathConfig.default(integrator=integrator, max_depth=max_depth, caustics_only=caustics_only)
        r = _compare(ctx, orc, pc, 16, 3)
        assert (r[:, 0] > 0).mean() > (0.0 if caustics_only else 0.5)
    ctx.close()
    orc.close()


def test_path_samples_pool_mnee_and_batches(gpu, oracle_mod):
    """MNEE with its completions (mnee_init, biased, one trial, path_sms_ss.cpp:196-241), biased
    MNEE, plain SMS; and a pass of 1.5 estimator batches."""
    sc = scenes.pool(48, 40, n=64, light_radius=0.3, light_radiance=60.0)
    cfg_mnee = abi.SmsConfig.default(mnee_init=1, biased=1, max_trials=1, halfvector_constraints=1)
    ctx = gpu.Context(sc, cfg_mnee)
    orc = oracle_mod.Oracle(sc, cfg_mnee)
    for biased_mnee in (0, 1):
        _compare(ctx, orc, abi.SmsPathConfig.default(max_depth=7, biased_mnee=biased_mnee), 8, 5)
    cfg = abi.SmsConfig.default(max_trials=100000, uniqueness_threshold=1e-5, solver_threshold=1e-5)
    ctx.set_config(cfg)
    orc.set_config(cfg)
    _compare(ctx, orc, abi.SmsPathConfig.default(max_depth=7), 8, 6)
    # 48 x 40 x 820 samples = 1.5 batches of SMS_PATH_CHUNK: the second batch's estimator seed
    spp = (3 * abi.PATH_CHUNK // 2) // (48 * 40) + 1
    _compare(ctx, orc, abi.SmsPathConfig.default(max_depth=7, caustics_only=1), spp, 7)
    ctx.close()
    orc.close()


def test_path_samples_normal_mapped(gpu, oracle_mod):
    """Normal-mapped casters in both single-scatter path tracers: a glass dome with a wavy map
    (nested dielectric; frames from slopes and from decoded RGB) and a mirror with a scratch map
    (nested conductor)."""
    u = np.linspace(0, 2 * np.pi, 64, endpoint=False)
    sx, sy = 0.12 * np.sin(2 * u)[None, :] * np.ones((64, 1)), 0.12 * np.cos(3 * u)[:, None] * np.ones((1, 64))
    n = np.stack([-sx, -sy, np.ones_like(sx)], -1)
    n /= np.linalg.norm(n, axis=2, keepdims=True)
    dome = scenes.sphere_dome(24, 20, light_radius=0.5, light_radiance=4.0)
    nm = dome.add_normalmap((0.5 * n + 0.5).astype(np.float32))
    dome.shapes[0].update(bsdf=abi.BSDF_NORMALMAP, normalmap=nm, tiles=2.0, k=0.0, eta=1.5)
    wall = scenes.mirror_wall(nm_res=64, n_scratches=16)
    wall.shapes[2]["radius"] = 0.3
    wall.emitters[0]["radiance"] = (5.0, 5.0, 5.0)
    encoded = scenes.sphere_dome(24, 20, light_radius=0.5, light_radiance=4.0)
    encoded.shapes[0].update(bsdf=abi.BSDF_NORMALMAP, normalmap=encoded.add_normalmap((0.5 * n + 0.5).astype(np.float32), encoded=True),
                             tiles=2.0, k=0.0, eta=1.5)
    for sc in (dome, wall, encoded):
        cfg = abi.SmsConfig.default(max_trials=100000)
        ctx = gpu.Context(sc, cfg)
        orc = oracle_mod.Oracle(sc, cfg)
        for integrator in (abi.PATH_SMS_SS, abi.PATH_FILTERED_SS):
            for only in (1, 0):
                _compare(ctx, orc, abi.SmsPathConfig.default(integrator=integrator, max_depth=6, caustics_only=only), 8, 3)
        ctx.close()
        orc.close()


def test_path_samples_slab_multi_scatter(gpu, oracle_mod):
    """path_sms_ms (unbiased SMS; MNEE with and without its completions) and path_filtered_ms on
    the slab, caustic term and whole image."""
    sc = scenes.slab(24, 20, n=32, amplitude=0.01, caster="bottom")
    sc.shapes[3]["radius"] = 0.3
    sc.emitters[0]["radiance"] = (50.0, 50.0, 50.0)
    cfg = abi.SmsConfig.default(bounces=2, max_trials=100000)
    cfg_mnee = abi.SmsConfig.default(bounces=2, mnee_init=1, biased=1, max_trials=1, halfvector_constraints=1)
    ctx = gpu.Context(sc, cfg)
    orc = oracle_mod.Oracle(sc, cfg)
    for only in (1, 0):
        for integrator in (abi.PATH_SMS_MS, abi.PATH_FILTERED_MS):
            r = _compare(ctx, orc, abi.SmsPathConfig.default(integrator=integrator, max_depth=8, caustics_only=only), 16, 9)
            assert (r[:, 0] > 0).mean() > 0.001
    ctx.set_config(cfg_mnee)
    orc.set_config(cfg_mnee)
    for biased_mnee in (0, 1):
        _compare(ctx, orc, abi.SmsPathConfig.default(integrator=abi.PATH_SMS_MS, max_depth=8,
                                                     biased_mnee=biased_mnee), 16, 10)
    ctx.close()
    orc.close()


@pytest.mark.parametrize("stddev", [0.5, 0.3, 0.2, 0.0])
def test_film_splat(gpu, oracle_mod, stddev):
    sc = scenes.sphere_dome(40, 30, light_radius=0.5, light_radiance=4.0)
    cfg = abi.SmsConfig.default(max_trials=100000)
    ctx = gpu.Context(sc, cfg)
    orc = oracle_mod.Oracle(sc, cfg)
    pc = abi.SmsPathConfig.default(max_depth=6)
    rgb, pos, _ = orc.path_samples(pc, 8, 9)
    rgb[5] = (np.nan, 0.0, 0.0)                    # dropped whole
    rgb[77] = (-1.0, 0.0, 0.0)
    ig, wg = ctx.film_splat(rgb, pos, 8, stddev)
    io, wo = orc.film_splat(rgb, pos, 8, stddev)
    assert np.array_equal(wg, wo)
    assert np.array_equal(ig, io)
    assert (wo > 0).all() and np.isfinite(io).all()
    ctx.close()
    orc.close()


def test_refused(gpu):
    rough = scenes.sphere_dome(16, 16)
    rough.shapes[0].update(bsdf=abi.BSDF_CONDUCTOR, eta=0.0, k=1.0, roughness=0.1)
    ctx = gpu.Context(rough, abi.SmsConfig.default())
    with pytest.raises(gpu.SmsError, match="rough"):
        ctx.path_samples(abi.SmsPathConfig.default(), 1, 1)
    ctx.close()
    sc = scenes.sphere_dome(16, 16)
    ctx = gpu.Context(sc, abi.SmsConfig.default())
    for bad in (dict(max_depth=17), dict(max_depth=-2), dict(integrator=7)):
        with pytest.raises(gpu.SmsError):
            ctx.path_samples(abi.SmsPathConfig.default(**bad), 1, 1)
    rgb, pos, _ = ctx.path_samples(abi.SmsPathConfig.default(), 1, 1)
    with pytest.raises(gpu.SmsError, match="radius"):
        ctx.film_splat(rgb, pos, 1, 0.6)
    ctx.close()
