"""Camera pass of the single-scatter integrator and the film, GPU against the oracle.

The camera pass uses the oracle's arithmetic contract, so queries and path weights must be
bit-equal; the estimator run over camera queries (including samples that reach no receiver)
must match the oracle's estimate bit for bit; the film's per-pixel sums must equal numpy's
in sample order.
"""
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


@pytest.mark.parametrize("scene_fn", ["pool", "sphere"])
def test_camera_queries_and_estimates_match_oracle(gpu, oracle_mod, scene_fn):
    if scene_fn == "pool":
        sc = scenes.pool(width=64, height=48, n=512)
        cfg = abi.SmsConfig.default(max_trials=100000, uniqueness_threshold=1e-5)
    else:
        sc = scenes.sphere_caustic(64, 48)
        cfg = abi.SmsConfig.default(max_trials=100000)
    ctx = gpu.Context(sc, cfg)
    orc = oracle_mod.Oracle(sc, cfg)
    qg, wg = ctx.ss_queries(4, 21)
    qo, wo = orc.ss_queries(4, 21)
    print(f"{scene_fn}: receivers {(qo['shape'] >= 0).mean():.3f} weights {np.unique(wo)[:6]}", flush=True)
    assert np.array_equal(qg.view(np.uint8), qo.view(np.uint8)), "camera queries differ"
    assert np.array_equal(wg, wo)
    ro, so = orc.ss_estimate(qo, 5)
    for first_pass in (0, 1):
        ctx.set_schedule(ss_first_pass=first_pass)
        rg, sg = ctx.ss_estimate(qg, 5)
        print("gpu", first_pass, sg, "oracle", so, flush=True)
        assert np.array_equal(rg, ro, equal_nan=True)
        for k in ("walks", "walks_converged", "newton_iterations", "bernoulli_calls", "bernoulli_trials"):
            assert sg[k] == so[k], k
        assert not rg[qg["shape"] < 0].any()
    ctx.close()
    orc.close()


def test_film_accumulate(gpu):
    """Per-pixel sums in sample order; samples with a non-finite or negative weighted channel are
    dropped and the count of kept samples is accumulated (src/librender/imageblock.cpp:85-107)."""
    sc = scenes.sphere_caustic(32, 16)
    ctx = gpu.Context(sc, abi.SmsConfig.default())
    pixels, spp = 32 * 16, 5
    rng = np.random.default_rng(1)
    rgb = rng.random((pixels * spp, 3)).astype(np.float32)
    w = rng.random(pixels * spp).astype(np.float32)
    rgb[7, 1] = np.nan
    rgb[30, 0] = np.inf
    rgb[44, 2] = -1.0
    rgb[45, 2] = -1e-6                   # within the reference's tolerance: kept
    w[61] = 0.0
    rgb[61, 0] = np.inf                  # 0 * inf = NaN: dropped
    img0 = rng.random((pixels, 3)).astype(np.float32)
    n0 = np.full(pixels, 3.0, np.float32)
    bufs = [gpu.DeviceBuffer(ctx, a.nbytes) for a in (rgb, w, img0, img0, n0)]
    for b, a in zip(bufs, (rgb, w, img0, np.zeros_like(img0), n0)):
        b.upload(a)
    ctx.film_accumulate_device(bufs[0], bufs[1], pixels, spp, bufs[2], bufs[3], bufs[4])
    img, sq, cnt = np.zeros_like(img0), np.zeros_like(img0), np.zeros_like(n0)
    bufs[2].download(img)
    bufs[3].download(sq)
    bufs[4].download(cnt)
    with np.errstate(invalid="ignore"):
        v = (w[:, None] * rgb).reshape(pixels, spp, 3)
    keep = (np.isfinite(v) & (v >= np.float32(-1e-5))).all(2)
    assert keep.sum() == pixels * spp - 4
    s = np.zeros((pixels, 3), np.float32)
    s2 = np.zeros((pixels, 3), np.float32)
    for k in range(spp):
        vk = np.where(keep[:, k, None], v[:, k], np.float32(0))
        s = s + vk
        s2 = s2 + vk * vk
    assert np.array_equal(img, img0 + s)
    assert np.array_equal(sq, s2)
    assert np.array_equal(cnt, n0 + keep.sum(1).astype(np.float32))
    # without the count buffer
    ctx.film_accumulate_device(bufs[0], bufs[1], pixels, spp, bufs[2])
    bufs[2].download(img)
    assert np.array_equal(img, (img0 + s) + s)
    for b in bufs:
        b.free()
    ctx.close()
