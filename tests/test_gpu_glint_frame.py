"""The glint estimate of a whole camera frame without a query buffer (sms_glint_estimate_frame,
sms_glint_frame_graph_create): each kernel generates the camera query it reads; the results and
statistics must equal sms_glint_make_queries + sms_glint_estimate_batch bit for bit, and the
oracle's -- unbiased and biased, with an area light (BSDF strategy), with a learned footprint
proposal, eager and as graph replays."""
import numpy as np
import pytest

from smsamd import abi, scenes

pytestmark = pytest.mark.gpu
KEYS = ("walks", "walks_converged", "newton_iterations", "bernoulli_calls", "bernoulli_trials", "bernoulli_max")


@pytest.fixture(scope="module")
def gpu():
    from smsamd import gpu as g
    g.lib()
    return g


def _same(a, b):
    (ra, sa), (rb, sb) = a, b
    assert np.array_equal(ra, rb, equal_nan=True)
    for k in KEYS:
        assert sa[k] == sb[k], k


@pytest.mark.parametrize("biased", [0, 1])
@pytest.mark.parametrize("area", [False, True])
def test_frame_equals_queries_and_estimate(gpu, biased, area):
    import oracle_lib
    sc = scenes.glint_plane(64, 48, nm_res=256, n_scratches=64, seed=3,
                            area_light=((0.3, 1.0, -0.2), 0.1, (40.0, 40.0, 40.0)) if area else None)
    cfg = abi.SmsConfig.default(max_trials=32 if biased else 100000, biased=biased)
    ctx = gpu.Context(sc, cfg)
    for spp, seed in ((2, 5), (3, 11)):
        fr = ctx.glint_estimate_frame(spp, seed)
        q = ctx.glint_queries(spp, seed)
        _same(fr, ctx.glint_estimate(q, seed))
        if spp == 2:
            orc = oracle_lib.Oracle(sc, cfg)
            _same(fr, orc.glint_estimate(orc.glint_queries(spp, seed), seed))
            orc.close()
        print(biased, area, spp, fr[1], flush=True)
    ctx.close()


def test_frame_with_proposal_and_graph(gpu):
    sc = scenes.glint_plane(64, 48, nm_res=256, n_scratches=64, seed=3)
    cfg = abi.SmsConfig.default(max_trials=100000)
    ctx = gpu.Context(sc, cfg)
    mlp = gpu.Mlp(ctx, seed=5)
    ctx.set_proposal(mlp, 0, 0.2)
    fr = ctx.glint_estimate_frame(2, 9)
    _same(fr, ctx.glint_estimate(ctx.glint_queries(2, 9), 9))
    ctx.set_proposal(None, -1)
    n = 64 * 48 * 4
    out = gpu.DeviceBuffer(ctx, n * 12)
    g = ctx.glint_frame_graph(4, out)
    for seed in (21, 22, 21):
        g.launch(seed)
        st = g.stats(reset=True)
        rgb = np.zeros((n, 3), np.float32)
        out.download(rgb)
        ref = ctx.glint_estimate_frame(4, seed)                 # eager work between replays
        _same((rgb, st), ref)
    g.close()
    out.free()
    mlp.close()
    ctx.close()
