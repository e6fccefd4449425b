"""Frame graphs (sms_*_graph_create / sms_graph_launch): every replay must give the results of
the eager estimator call with the same seed, bit for bit, and the folded statistics their sums.
Replays run both back to back and with the program's own eager work and synchronisations
between them; a graph whose context changed since its capture must refuse to launch.
"""
import numpy as np
import pytest

from smsamd import abi, scenes, workloads

pytestmark = pytest.mark.gpu

SUMMED = ("walks", "walks_converged", "newton_iterations", "bernoulli_calls", "bernoulli_trials",
          "overflow_queries", "tail_iterations")


@pytest.fixture(scope="module")
def gpu():
    from smsamd import gpu as g
    g.lib()
    return g


def _fold(stats):
    tot = {k: sum(s[k] for s in stats) for k in SUMMED}
    tot["bernoulli_max"] = max(s["bernoulli_max"] for s in stats)
    return tot


def _check_graph(ctx, graph, outbuf, n, seeds, eager, eager_stats, between=None):
    """Replays one by one (sync and eager work between them), then back to back."""
    out = np.zeros((n, 3), np.float32)
    graph.stats(reset=True)
    for s in seeds:
        graph.launch(s)
        ctx.sync()
        outbuf.download(out)
        assert np.array_equal(out, eager[s], equal_nan=True), f"replay with seed {s}"
        if between is not None:
            between()
    tot = graph.stats(reset=True)
    ref = _fold([eager_stats[s] for s in seeds])
    for k in ref:
        assert tot[k] == ref[k], k
    for s in seeds:
        graph.launch(s)
    tot = graph.stats(reset=True)
    outbuf.download(out)
    assert np.array_equal(out, eager[seeds[-1]], equal_nan=True)
    for k in ref:
        assert tot[k] == ref[k], k


def test_glint_graph(gpu):
    sc = scenes.glint_plane(64, 48, nm_res=256, n_scratches=64, seed=3)
    ctx = gpu.Context(sc, abi.SmsConfig.default(max_trials=100000))
    qbuf, n = ctx.glint_queries_device(2, 5)
    out, other = gpu.DeviceBuffer(ctx, n * 12), gpu.DeviceBuffer(ctx, n * 12)
    seeds = [9, 10, 11]
    eager, est = {}, {}
    for s in seeds:
        est[s] = ctx.glint_estimate_device(qbuf, n, s, out)
        eager[s] = np.zeros((n, 3), np.float32)
        out.download(eager[s])
    assert est[9]["bernoulli_calls"] > 20
    g = ctx.glint_graph(qbuf, n, out)
    _check_graph(ctx, g, out, n, seeds, eager, est, between=lambda: ctx.glint_estimate_device(qbuf, n, 77, other))
    # a changed schedule invalidates the capture
    ctx.set_schedule(**ctx.get_schedule())
    with pytest.raises(gpu.SmsError):
        g.launch(9)
    g.close()
    g = ctx.glint_graph(qbuf, n, out)
    _check_graph(ctx, g, out, n, seeds[:1], eager, est)
    g.close()
    for b in (qbuf, out, other):
        b.free()
    ctx.close()


@pytest.mark.parametrize("first_pass,proposal", [(0, False), (1, False), (0, True)])
def test_ss_graph(gpu, first_pass, proposal):
    sc = scenes.pool(width=64, height=48, n=512)
    ctx = gpu.Context(sc, abi.SmsConfig.default(max_trials=100000, uniqueness_threshold=1e-5))
    ctx.set_schedule(ss_first_pass=first_pass, lane_trials=8)
    mlp = None
    if proposal:
        mlp = gpu.Mlp(ctx, seed=5)
        ctx.set_proposal(mlp, 0, 0.2)
    pass_spp = 4
    n = 64 * 48 * pass_spp
    qbuf, wbuf = gpu.DeviceBuffer(ctx, n * abi.SS_QUERY.itemsize), gpu.DeviceBuffer(ctx, n * 4)
    ctx.ss_queries_device(pass_spp, 3, qbuf, wbuf)
    out, other = gpu.DeviceBuffer(ctx, n * 12), gpu.DeviceBuffer(ctx, n * 12)
    seeds = [21, 22, 23]
    eager, est = {}, {}
    for s in seeds:
        est[s] = ctx.ss_estimate_device(qbuf, n, s, out)
        eager[s] = np.zeros((n, 3), np.float32)
        out.download(eager[s])
    assert est[21]["walks_converged"] > 100
    g = ctx.ss_graph(qbuf, n, out)
    _check_graph(ctx, g, out, n, seeds, eager, est, between=lambda: ctx.ss_estimate_device(qbuf, n, 99, other))
    g.close()
    # a graph of a smaller batch stays valid across eager calls of that size
    g = ctx.ss_graph(qbuf, n // 2, other)
    e2 = np.zeros((n // 2, 3), np.float32)
    st2 = ctx.ss_estimate_device(qbuf, n // 2, 31, other)
    other.download(e2)
    _check_graph(ctx, g, other, n // 2, [31], {31: e2}, {31: st2})
    g.close()
    for b in (qbuf, wbuf, out, other):
        b.free()
    if mlp is not None:
        mlp.close()
    ctx.close()


def test_ss_pass_graph(gpu):
    """Camera pass + estimate + film in one graph against the same three calls made eagerly."""
    from smsamd import render
    sc = scenes.pool(width=64, height=48, n=512)
    ctx = gpu.Context(sc, abi.SmsConfig.default(max_trials=100000, uniqueness_threshold=1e-5))
    runner = render.PassRunner(ctx, 4)
    film_e, film_g = render.Film(ctx, runner.pixels), render.Film(ctx, runner.pixels)
    q2, w2, rgb2 = (gpu.DeviceBuffer(ctx, b.nbytes) for b in (runner.q, runner.w, runner.rgb))
    g = ctx.ss_pass_graph(4, q2, w2, rgb2, film_g.sum, film_g.sq, film_g.n)
    for s in (41, 42, 43):
        runner.run(s, film_e)
        g.launch(s)
        ctx.sync()
        for a, b in ((film_e.sum, film_g.sum), (film_e.sq, film_g.sq), (film_e.n, film_g.n)):
            x, y = np.zeros(a.nbytes // 4, np.float32), np.zeros(a.nbytes // 4, np.float32)
            a.download(x)
            b.download(y)
            assert np.array_equal(x, y, equal_nan=True), s
    assert g.stats()["walks"] > 0
    g.close()
    for b in (q2, w2, rgb2):
        b.free()
    runner.free()
    film_e.free()
    film_g.free()
    ctx.close()


def test_ms_graph(gpu):
    sc = scenes.slab(width=64, height=64, n=256)
    ctx = gpu.Context(sc, abi.SmsConfig.default(bounces=2, max_trials=100000))
    q = workloads.plane_queries(sc, 1500, 5, extent=0.35, shape=2)
    n = len(q)
    qbuf, out = gpu.DeviceBuffer(ctx, q.nbytes), gpu.DeviceBuffer(ctx, n * 12)
    qbuf.upload(q)
    seeds = [9, 10]
    eager, est = {}, {}
    for s in seeds:
        est[s] = ctx.ms_estimate_device(qbuf, n, s, out)
        eager[s] = np.zeros((n, 3), np.float32)
        out.download(eager[s])
    g = ctx.ms_graph(qbuf, n, out)
    _check_graph(ctx, g, out, n, seeds, eager, est)
    # a larger eager batch grows the work buffers: the capture is stale
    big = np.concatenate([q] * 400)
    bbuf, bout = gpu.DeviceBuffer(ctx, big.nbytes), gpu.DeviceBuffer(ctx, len(big) * 12)
    bbuf.upload(big)
    ctx.ms_estimate_device(bbuf, len(big), 3, bout)
    with pytest.raises(gpu.SmsError):
        g.launch(9)
    g.close()
    for b in (qbuf, out, bbuf, bout):
        b.free()
    ctx.close()
