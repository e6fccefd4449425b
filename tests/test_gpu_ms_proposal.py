"""Learned seed proposal on the multi-scatter slab (the chain's seed point on the top face),
GPU against the oracle.

Given the network's raw outputs (read back with sms_ms_proposal_outputs) the oracle's
multi-scatter estimator with the defensive mixture must reproduce the GPU's estimates and
statistics bit for bit, through both the lane pass and the cooperative tail; the training
capture records the uv of the chain's vertex on the caster; a network trained on captured
solutions must converge more often with the estimate's mean unchanged.
"""
import math
import os
import sys

import numpy as np
import pytest

from smsamd import abi, scenes, workloads

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
def slab_scene():
    return scenes.slab(width=64, height=64, n=256)


def _floor_queries(sc, n, seed, extent=0.35):
    return workloads.plane_queries(sc, n, seed, extent=extent, shape=2)


def _concentrated(mlp, sc, sigma=0.1):
    """Network whose mixture is one fixed spread of components around the top-face point above
    the floor's centre on the way to the light (last layer: zero weights)."""
    p = mlp.get_params()
    w0, b0 = M.OFFSETS[4]
    p[w0:b0] = 0.0
    light = np.asarray(sc.shapes[3]["center"])
    hit = 0.3 * light                      # floor centre to the light, at the top face's height
    u, v = hit[0] + 0.5, hit[2] + 0.5
    b = np.zeros(48, np.float32)
    for c in range(8):
        b[6 * c + 1] = math.log(u / (1 - u)) + 0.3 * (c % 4 - 1.5)
        b[6 * c + 2] = math.log(v / (1 - v)) + 0.3 * (c // 4 - 0.5)
        b[6 * c + 3] = b[6 * c + 4] = math.log(sigma)
        b[6 * c + 5] = 0.1 * (c - 4)
    p[b0:b0 + 48] = b
    mlp.set_params(p)


@pytest.mark.parametrize("kind", ["random", "concentrated"])
def test_ms_estimator_with_proposal_matches_oracle(gpu, oracle_mod, slab_scene, kind):
    cfg = abi.SmsConfig.default(bounces=2, max_trials=100000)
    ctx = gpu.Context(slab_scene, cfg)
    orc = oracle_mod.Oracle(slab_scene, cfg)
    mlp = gpu.Mlp(ctx, seed=11)
    if kind == "concentrated":
        _concentrated(mlp, slab_scene)
    ctx.set_proposal(mlp, 0, 0.2)
    q = _floor_queries(slab_scene, 1500, 5)
    assert np.array_equal(ctx.ms_proposal_inputs(q, 9), orc.ms_prop_inputs(q, 9))
    raw = ctx.ms_proposal_outputs(q, 9)
    ro, so = orc.ms_estimate_prop(q, 9, raw, 0, 0.2)
    ru, su = orc.ms_estimate(q, 9)
    for lane_trials in (8, 1):
        ctx.set_schedule(lane_trials=lane_trials)
        rg, sg = ctx.ms_estimate(q, 9)
        print(kind, lane_trials, "gpu", sg, "oracle", so, "uniform", su, flush=True)
        for k in ("walks", "walks_converged", "newton_iterations", "bernoulli_calls", "bernoulli_trials",
                  "bernoulli_max"):
            assert sg[k] == so[k], k
        assert np.array_equal(rg, ro, equal_nan=True)
    assert (ro[:, 0] > 0).sum() >= 8
    if kind == "concentrated":
        assert so["walks_converged"] / so["walks"] > su["walks_converged"] / su["walks"]
    # the biased estimator keeps uniform seeds
    cfg_b = abi.SmsConfig.default(bounces=2, biased=1, max_trials=16)
    ctx.set_config(cfg_b)
    orc.set_config(cfg_b)
    rg, _ = ctx.ms_estimate(q[:500], 9)
    rb, _ = orc.ms_estimate(q[:500], 9)
    assert np.array_equal(rg, rb, equal_nan=True)
    mlp.close()
    ctx.close()
    orc.close()


def test_ms_capture(gpu, oracle_mod, slab_scene):
    cfg = abi.SmsConfig.default(bounces=2, max_trials=100000)
    ctx = gpu.Context(slab_scene, cfg)
    ctx.set_proposal(None, 0)
    q = _floor_queries(slab_scene, 4000, 5)
    xin = ctx.ms_proposal_inputs(q, 3)
    rows = {r.tobytes(): i for i, r in enumerate(xin)}
    top = slab_scene.shapes[0]
    pos = np.asarray(top["positions"]).reshape(-1, 3)
    for lane_trials in (8, 1):
        ctx.set_schedule(lane_trials=lane_trials)
        ctx.set_training_capture(1 << 16, abi.CAPTURE_SOLUTIONS)
        rg, st = ctx.ms_estimate(q, 3)
        x, uv, w = ctx.training_samples()
        idx = np.array([rows.get(r.tobytes(), -1) for r in x])
        print(f"lane_trials {lane_trials}: captured {len(x)} solutions of {len(np.unique(idx))} queries; "
              f"converged walks {st['walks_converged']}", flush=True)
        assert (w == 1).all() and (idx >= 0).all()
        assert ((uv >= 0) & (uv <= 1)).all()
        # every converged walk the estimator counts is recorded (first walks and trials up to
        # the match), less the rare chain whose caster vertex a ray cast does not find again
        assert 0.999 * st["walks_converged"] <= len(x) <= st["walks_converged"]
        nz = np.nonzero(rg[:, 0] > 0)[0]
        counts = np.bincount(idx, minlength=len(q))
        assert (counts[nz] >= 2).all()
        # the recorded uv is on the top face where the chain leaves the slab: its grid point
        # lies above the query, displaced toward the light
        i0 = int(nz[0])
        j = np.nonzero(idx == i0)[0][0]
        gx = int(round(float(uv[j, 0]) * 256)), int(round(float(uv[j, 1]) * 256))
        p = pos[gx[1] * 257 + gx[0]]
        assert abs(p[1] - 0.6) < 0.02
        assert np.linalg.norm(p[[0, 2]] - q["p"][i0][[0, 2]]) < 0.5
        # estimate-weighted records: one per non-zero estimate, weight = its RGB mean
        ctx.set_training_capture(1 << 16, abi.CAPTURE_ESTIMATES)
        rg2, _ = ctx.ms_estimate(q, 3)
        assert np.array_equal(rg, rg2, equal_nan=True)
        x, uv, w = ctx.training_samples()
        idx = np.array([rows.get(r.tobytes(), -1) for r in x])
        assert (idx >= 0).all() and set(idx.tolist()) <= set(nz.tolist()) and len(set(idx.tolist())) == len(idx)
        assert len(idx) >= len(nz) - max(1, len(nz) // 100)
        np.testing.assert_allclose(w, rg[idx].mean(1), rtol=1e-5)
    ctx.set_training_capture(0)
    ctx.close()


def test_ms_training_beats_uniform(gpu, slab_scene):
    """Capture estimate-weighted chain solutions under uniform seeding, train on the GPU, and
    check that the trained proposal converges more often and needs fewer Bernoulli trials, with
    the estimate's mean unchanged within its standard error."""
    cfg = abi.SmsConfig.default(bounces=2, max_trials=100000)
    ctx = gpu.Context(slab_scene, cfg)
    ctx.set_proposal(None, 0)
    ctx.set_training_capture(1 << 18, abi.CAPTURE_ESTIMATES)
    seed = 100
    while ctx.training_count() < 30000 and seed < 200:
        ctx.ms_estimate(_floor_queries(slab_scene, 65536, seed), seed)
        seed += 1
    xb, ub, wb, n = ctx.training_samples_device()
    mlp = gpu.Mlp(ctx, seed=3)
    rng = np.random.default_rng(0)
    hist = mlp.train_device(xb, ub, n, 8192, rng.integers(0, n - 8192, 400), 1e-3, wbuf=wb, want_losses=True)
    for b in (xb, ub, wb):
        b.free()
    print(f"samples {n} in {seed - 100} passes; loss {hist[0]:.3f} -> {hist[-1]:.3f}", flush=True)
    assert hist[-1] < hist[0] - 1.0
    q = _floor_queries(slab_scene, 65536, 999)
    ctx.set_training_capture(0)
    ctx.set_proposal(None, 0)
    ru, su = ctx.ms_estimate(q, 5)
    ctx.set_proposal(mlp, 0, 0.2)
    rl, sl = ctx.ms_estimate(q, 5)
    cu, cl = su["walks_converged"] / su["walks"], sl["walks_converged"] / sl["walks"]
    tu, tl = su["bernoulli_trials"] / max(su["bernoulli_calls"], 1), sl["bernoulli_trials"] / max(sl["bernoulli_calls"], 1)
    lu, ll = ru[:, 0].astype(np.float64), rl[:, 0].astype(np.float64)
    se = math.sqrt(lu.var() / len(lu) + ll.var() / len(ll))
    print(f"uniform: conv {cu:.4f} trials/estimate {tu:.2f} mean {lu.mean():.5f} var {lu.var():.4f}; "
          f"learned: conv {cl:.4f} trials/estimate {tl:.2f} mean {ll.mean():.5f} var {ll.var():.4f}; se {se:.5f}",
          flush=True)
    assert cl > 1.3 * cu
    assert tl < tu
    assert abs(lu.mean() - ll.mean()) < 4 * se
    mlp.close()
    ctx.close()


def test_ms_autotune_keeps_results(gpu, slab_scene):
    """sms_ms_autotune picks a first pass and lane_trials from its candidates; the estimate does not change."""
    from smsamd import gpu as g
    cfg = abi.SmsConfig.default(bounces=2, max_trials=100000)
    ctx = g.Context(slab_scene, cfg)
    q = _floor_queries(slab_scene, 20000, 7)
    r0, s0 = ctx.ms_estimate(q, 4)
    qb, out = g.DeviceBuffer(ctx, q.nbytes), g.DeviceBuffer(ctx, len(q) * 12)
    qb.upload(q)
    sched = ctx.ms_autotune_device(qb, len(q), 11, out)
    assert sched["lane_trials"] in ((4, 16, 64, 256, 1024) if sched["ss_first_pass"] == 0 else (1, 2, 4, 8, 16, 32))
    assert ctx.get_schedule() == sched
    r1, s1 = ctx.ms_estimate(q, 4)
    assert np.array_equal(r0, r1, equal_nan=True)
    for k in ("walks", "walks_converged", "newton_iterations", "bernoulli_trials"):
        assert s0[k] == s1[k], k
    qb.free()
    out.free()
    ctx.close()
