"""Learned seed proposal on glint footprints, on the GPU through the C ABI.

The glint estimator's footprint sample xi in [0,1)^2 (manifold_glints.cpp:252-259) is drawn
from a defensive mixture of the network's Gaussian mixture and the uniform footprint sample,
for the first walk and every Bernoulli trial (manifold_glints.cpp:70-111):

* network inputs of a glint query: bit-equal to the oracle's;
* estimates with the mixture seeds: bit-equal to the oracle fed the same raw network outputs,
  over the trial-chain schedules, and for the biased estimator;
* capture of converged first walks as footprint samples, and a training run whose network
  must put more density than the uniform seed at held-out solutions.
"""
import math
import os
import sys

import numpy as np
import pytest

from smsamd import abi, scenes

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as M  # noqa: E402

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


@pytest.fixture(scope="module")
def glint_scene():
    return scenes.glint_plane(64, 48, nm_res=256, n_scratches=64, seed=3)


def _perturbed_params(mlp, seed):
    p = mlp.get_params()
    rng = np.random.default_rng(seed)
    for (n, k), (w0, b0) in zip(M.LAYERS, M.OFFSETS):
        p[b0:b0 + n] = rng.normal(0, 0.1, n).astype(np.float32)
    p[M.OFFSETS[4][0]:M.OFFSETS[4][1]] *= 10.0
    return p


def _queries(orc, spp, seed):
    q = orc.glint_queries(spp, seed)
    return np.ascontiguousarray(q[q["shape"] >= 0])


def test_glint_proposal_inputs_bit_exact(gpu, oracle_mod, glint_scene):
    cfg = abi.SmsConfig.default(max_trials=100000)
    ctx = gpu.Context(glint_scene, cfg)
    orc = oracle_mod.Oracle(glint_scene, cfg)
    q = _queries(orc, 2, 5)
    xg, xo = ctx.glint_proposal_inputs(q, 7), orc.glint_prop_inputs(q, 7)
    assert np.array_equal(xg, xo)
    assert np.abs(xg[:, :6]).max() <= 1.0 + 1e-6       # shading point in the scene box, unit normal
    # the inputs do not depend on whether a proposal is set
    mlp = gpu.Mlp(ctx, seed=1)
    ctx.set_proposal(mlp, 0, 0.2)
    assert np.array_equal(ctx.glint_proposal_inputs(q, 7), xo)
    mlp.close()
    ctx.close()
    orc.close()


def test_glint_proposal_rejects_non_glint_shapes(gpu):
    sc = scenes.glint_plane(32, 32, nm_res=128, n_scratches=8, seed=1, area_light=((0.3, 1.0, 0.0), 0.05, (50.0, 50.0, 50.0)))
    ctx = gpu.Context(sc, abi.SmsConfig.default())
    mlp = gpu.Mlp(ctx, seed=1)
    with pytest.raises(gpu.SmsError):
        ctx.set_proposal(mlp, 1, 0.2)                 # the diffuse light sphere
    with pytest.raises(gpu.SmsError):
        ctx.set_proposal(mlp, 0, 0.0)                 # defensive weight outside (0, 1]
    ctx.set_proposal(mlp, 0, 1.0)
    ctx.set_proposal(None, -1)
    mlp.close()
    ctx.close()


@pytest.mark.parametrize("kind", ["random", "perturbed"])
def test_glint_estimate_with_proposal_matches_oracle(gpu, oracle_mod, glint_scene, kind):
    cfg = abi.SmsConfig.default(max_trials=100000)
    ctx = gpu.Context(glint_scene, cfg)
    orc = oracle_mod.Oracle(glint_scene, cfg)
    mlp = gpu.Mlp(ctx, seed=11)
    if kind == "perturbed":
        mlp.set_params(_perturbed_params(mlp, 2))
    ctx.set_proposal(mlp, 0, 0.2)
    q = _queries(orc, 2, 5)
    raw = ctx.glint_proposal_outputs(q, 9)
    assert np.isfinite(raw).all()
    ro, so = orc.glint_estimate_prop(q, 9, raw, 0.2)
    ru, su = orc.glint_estimate(q, 9)
    print(kind, "oracle", so, "uniform", su, flush=True)
    assert so["bernoulli_calls"] > 20
    assert so["walks"] != su["walks"]
    rg, sg = ctx.glint_estimate(q, 9)
    assert np.array_equal(rg, ro, equal_nan=True)
    for k in ("walks", "walks_converged", "newton_iterations", "bernoulli_calls", "bernoulli_trials", "bernoulli_max"):
        assert sg[k] == so[k], k
    schedules = [dict(trial_group=1, lane_trials=1, tail_rounds=0),
                 dict(trial_group=1, lane_trials=1, tail_rounds=0, long_chains=8, long_trials=64),
                 dict(trial_group=4, lane_trials=8, tail_rounds=2),
                 dict(trial_group=16, lane_trials=16, tail_rounds=3)]
    for sch in schedules:
        ctx.set_schedule(**sch)
        rg2, sg2 = ctx.glint_estimate(q, 9)
        assert np.array_equal(rg2, ro, equal_nan=True), sch
        for k in ("walks", "walks_converged", "newton_iterations", "bernoulli_trials", "bernoulli_max"):
            assert sg2[k] == so[k], (k, sch)
    # biased trial sets: every member seeded by the mixture
    cfg_b = abi.SmsConfig.default(biased=1, max_trials=16)
    ctx.set_config(cfg_b)
    orc.set_config(cfg_b)
    rg, sg = ctx.glint_estimate(q, 9)
    ro, so = orc.glint_estimate_prop(q, 9, raw, 0.2)
    assert sg["walks"] == so["walks"] and sg["walks_converged"] == so["walks_converged"]
    assert np.array_equal(rg, ro, equal_nan=True)
    # proposal cleared: back to the reference's uniform footprint seeds
    ctx.set_config(cfg)
    ctx.set_proposal(None, -1)
    rg, sg = ctx.glint_estimate(q, 9)
    assert np.array_equal(rg, ru, equal_nan=True) and sg["walks"] == su["walks"]
    mlp.close()
    ctx.close()
    orc.close()


def test_glint_capture(gpu, oracle_mod, glint_scene):
    cfg = abi.SmsConfig.default(max_trials=100000)
    ctx = gpu.Context(glint_scene, cfg)
    orc = oracle_mod.Oracle(glint_scene, cfg)
    ctx.set_proposal(None, 0)
    ctx.set_training_capture(1 << 16)
    q = _queries(orc, 4, 5)
    rg, sg = ctx.glint_estimate(q, 3)
    x, xi, w = ctx.training_samples()
    assert (w == 1).all()
    xin = ctx.glint_proposal_inputs(q, 3)
    rows = {}
    for i, r in enumerate(xin):
        rows.setdefault(r.tobytes(), []).append(i)
    idx = [rows.get(r.tobytes(), [-1]) for r in x]
    print(f"captured {len(x)} first-walk solutions of {len(q)} queries; nonzero estimates {(rg[:, 0] > 0).sum()}",
          flush=True)
    assert len(x) > 20 and all(i[0] >= 0 for i in idx)
    # the footprint sample of a solution inside the footprint lies in [0,1]^2 (up to rounding)
    assert (xi > -1e-3).all() and (xi < 1 + 1e-3).all()
    # one record per converged first walk: a non-zero estimate's first walk converged
    assert len(x) <= len(q)
    assert len(x) >= (rg[:, 0] > 0).sum()
    # a walk started at the captured footprint sample converges at once: the record is the solution
    k = min(len(x), 300)
    wq = q[[i[0] for i in idx[:k]]]
    win = np.zeros(k, abi.GLINT_WALK_IN)
    for a, b in (("sensor_p", "sensor_p"), ("si_p", "p"), ("frame_s", "sh_s"), ("frame_t", "sh_t"),
                 ("frame_n", "sh_n"), ("uv", "uv"), ("duv_dx", "duv_dx"), ("duv_dy", "duv_dy"), ("shape", "shape")):
        win[a] = wq[b]
    win["ei_p"] = glint_scene.emitters[0]["position"]
    win["n_offset"] = (0.0, 0.0, 1.0)
    win["xi_start"] = xi[:k]
    out = ctx.glint_walks(win)
    print(f"re-walks from captured xi: success {out['success'].mean():.3f} iterations {out['iterations'].mean():.2f}",
          flush=True)
    assert out["success"].mean() > 0.95 and out["iterations"].mean() < 2.0
    # capture does not change the estimate
    ctx.set_training_capture(0)
    rg2, _ = ctx.glint_estimate(q, 3)
    assert np.array_equal(rg, rg2, equal_nan=True)
    # records do not depend on the schedule
    outs = []
    for sch in (dict(trial_group=1, lane_trials=1, tail_rounds=0), dict(trial_group=8, lane_trials=8, tail_rounds=2)):
        ctx.set_schedule(**sch)
        ctx.set_training_capture(1 << 16)
        ctx.clear_training()
        ctx.glint_estimate(q, 3)
        outs.append(ctx.training_samples())
    for a, b in zip(*outs):
        assert np.array_equal(a, b)
    ctx.set_training_capture(0)
    ctx.close()
    orc.close()


def _finite(r):
    """Estimates the film keeps: a sample with a non-finite channel is dropped, as the reference's
    ImageBlock::put drops it (the LEAN MIS weight is NaN where a footprint's slope covariance is
    singular, lean.h:127-140)."""
    return r[np.isfinite(r).all(1), 0].astype(np.float64)


def test_glint_training_learns_footprint_density(gpu):
    """Capture first-walk solutions under uniform footprint seeds on a glint frame whose
    footprints cover about 4 texels (the 512^2 frame / 2048^2 map ratio at 128^2 / 512^2) and
    train on the GPU.  The trained mixture puts more density than the uniform seed at held-out
    solutions, and the estimate's mean is unchanged within its standard error.  It does not
    lower the Bernoulli trial count: given a first-walk success the expected trial count is
    N / P (N solutions reachable in the footprint, P the seed's total convergence probability),
    uniform seeds already reach P of about 0.75 in these footprints, and mixture samples that
    fall outside the footprint are failed walks -- the counts are printed, not asserted."""
    sc = scenes.glint_plane(128, 128, nm_res=512, n_scratches=256, seed=5)
    cfg = abi.SmsConfig.default(max_trials=100000)
    ctx = gpu.Context(sc, cfg)
    ctx.set_proposal(None, 0)
    ctx.set_training_capture(1 << 20)           # no capture list fills: the records are reproducible
    seed = 100
    while ctx.training_count() < 200000 and seed < 600:
        ctx.glint_estimate(ctx.glint_queries(4, seed), seed)
        seed += 1
    xb, ub, wb, n = ctx.training_samples_device()
    mlp = gpu.Mlp(ctx, seed=3)
    rng = np.random.default_rng(0)
    bs = 8192
    hist = mlp.train_device(xb, ub, n, bs, rng.integers(0, n - bs, 3000), 1e-3, want_losses=True)
    xb.free()
    ub.free()
    wb.free()
    # held-out solutions: another capture, other pixel jitter
    ctx.clear_training()
    ctx.set_training_capture(1 << 16)
    for s in range(5000, 5010):
        ctx.glint_estimate(ctx.glint_queries(4, s), s)
    xh, uh, _ = ctx.training_samples()
    held, _ = mlp.grad(xh, uh)
    ctx.set_training_capture(0)
    print(f"samples {n} in {seed - 100} passes; loss {hist[0]:.3f} -> {hist[-1]:.3f}; held-out {held:.3f} "
          f"on {len(xh)} solutions", flush=True)
    assert hist[-1] < hist[0] - 1.0
    assert held < -0.25                # held-out solutions: > 1.28x the uniform footprint density
    q = ctx.glint_queries(64, 999)
    ctx.set_proposal(None, 0)
    ru, su = ctx.glint_estimate(q, 5)
    ctx.set_proposal(mlp, 0, 0.2)
    rl, sl = ctx.glint_estimate(q, 5)
    cu, cl = su["walks_converged"] / su["walks"], sl["walks_converged"] / sl["walks"]
    tu, tl = su["bernoulli_trials"] / su["bernoulli_calls"], sl["bernoulli_trials"] / sl["bernoulli_calls"]
    lu, ll = _finite(ru), _finite(rl)
    se = math.sqrt(lu.var() / len(lu) + ll.var() / len(ll))
    print(f"uniform: conv {cu:.4f} trials/estimate {tu:.3f} mean {lu.mean():.5g} var {lu.var():.4g}; "
          f"learned: conv {cl:.4f} trials/estimate {tl:.3f} mean {ll.mean():.5g} var {ll.var():.4g}; se {se:.3g}; "
          f"dropped {len(q) - len(lu)} {len(q) - len(ll)}", flush=True)
    assert abs(lu.mean() - ll.mean()) < 4 * se
    mlp.close()
    ctx.close()
