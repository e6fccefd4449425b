"""Learned seed proposal on the GPU, called through the C ABI.

* network inputs of a query: bit-equal to the oracle's;
* MLP forward / gradient (BF16 matrix cores, FP32 accumulation): against the oracle's
  BF16-rounding restatement (tight) and its exact float64 form (loose), and the gradient
  against a torch float32 autograd reference;
* Adam: against the oracle's step;
* estimator with the defensive-mixture seed: bit-equal to the oracle fed the same raw network
  outputs (unbiased and biased estimators);
* capture of training samples, and a training run whose network must beat uniform seeding.
"""
import math
import os
import sys
import time

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
def pool_scene():
    return scenes.pool(width=64, height=64, n=512)


def _floor_queries(sc, n, seed, extent=0.9):
    q = workloads.plane_queries(sc, n, seed, extent=extent, shape=1)
    q["p"][:, 1] = -1.0
    return q


def _batch(n, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1, 1, (n, 9)).astype(np.float32), rng.uniform(0.05, 0.95, (n, 2)).astype(np.float32)


def _perturbed_params(mlp, seed):
    p = mlp.get_params()
    rng = np.random.default_rng(seed)
    for (n, k), (w0, b0) in zip(M.LAYERS, M.OFFSETS):
        p[b0:b0 + n] = rng.normal(0, 0.1, n).astype(np.float32)
    p[M.OFFSETS[4][0]:M.OFFSETS[4][1]] *= 10.0      # a non-trivial mixture
    return p


def test_proposal_inputs_bit_exact(gpu, oracle_mod, pool_scene):
    cfg = abi.SmsConfig.default()
    ctx = gpu.Context(pool_scene, cfg)
    orc = oracle_mod.Oracle(pool_scene, cfg)
    q = _floor_queries(pool_scene, 20000, 3)
    xg, xo = ctx.proposal_inputs(q, 7), orc.prop_inputs(q, 7)
    assert np.array_equal(xg, xo)
    assert np.abs(xg).max() <= 1.0 + 1e-6
    ctx.close()
    orc.close()


def test_mlp_forward(gpu, pool_scene):
    ctx = gpu.Context(pool_scene, abi.SmsConfig.default())
    mlp = gpu.Mlp(ctx, seed=1)
    p = _perturbed_params(mlp, 2)
    mlp.set_params(p)
    assert np.array_equal(mlp.get_params(), p)
    x, _ = _batch(1000, 3)                           # not a multiple of the 64-sample tile
    out = mlp.infer(x)
    ref_bf, _ = M.forward(p, x, "bf16")
    ref64, _ = M.forward(p, x, "f64")
    scale = np.abs(ref64).max()
    e_bf = np.abs(out - ref_bf).max() / scale
    e_64 = np.abs(out - ref64).max() / scale
    print(f"forward: max|o|={scale:.3f} err vs bf16 oracle {e_bf:.2e} vs f64 {e_64:.2e}", flush=True)
    assert e_bf < 5e-3
    assert e_64 < 5e-2
    mlp.close()
    ctx.close()


def _torch_fp32_grad(p, x, uv):
    import torch
    pt = torch.tensor(p, dtype=torch.float32, requires_grad=True)
    xt = torch.tensor(x)
    feats = [xt]
    for k in range(8):
        a = float(np.float32(np.pi) * np.float32(1 << k)) * xt
        feats += [torch.sin(a), torch.cos(a)]
    h = torch.cat(feats + [torch.zeros(len(x), 7)], 1)
    o = 0
    for i, (n, k) in enumerate(M.LAYERS):
        W = pt[o:o + n * k].reshape(n, k)
        b = pt[o + n * k:o + n * k + n]
        o += n * k + n
        h = h @ W.T + b
        if i < 4:
            h = torch.relu(h)
    r = h.reshape(-1, 8, 6)
    mu = torch.sigmoid(r[..., 1:3])
    sig = torch.exp(torch.clamp(r[..., 3:5], -7.0, 0.0))
    rho = 0.95 * torch.tanh(r[..., 5])
    d = torch.tensor(uv)[:, None, :] - mu
    z1 = d[..., 0] / sig[..., 0]
    z2 = (d[..., 1] / sig[..., 1] - rho * z1) / torch.sqrt(1 - rho * rho)
    logn = (-0.5 * (z1 * z1 + z2 * z2) - math.log(2 * math.pi) - torch.log(sig[..., 0]) - torch.log(sig[..., 1])
            - 0.5 * torch.log(1 - rho * rho))
    loss = -torch.logsumexp(torch.log_softmax(r[..., 0], -1) + logn, -1).mean()
    loss.backward()
    return float(loss.detach()), pt.grad.numpy().astype(np.float64)


def test_mlp_gradient(gpu, pool_scene):
    ctx = gpu.Context(pool_scene, abi.SmsConfig.default())
    mlp = gpu.Mlp(ctx, seed=4)
    p = _perturbed_params(mlp, 5)
    mlp.set_params(p)
    x, uv = _batch(1000, 6)
    loss, g = mlp.grad(x, uv)
    l_bf, g_bf = M.loss_and_grad(p, x, uv, "bf16")
    l_64, g_64 = M.loss_and_grad(p, x, uv, "f64")
    l_t, g_t = _torch_fp32_grad(p, x, uv)
    g = g.astype(np.float64)

    def rel(a, b):
        return np.linalg.norm(a - b) / np.linalg.norm(b)

    def cos(a, b):
        return a @ b / np.linalg.norm(a) / np.linalg.norm(b)
    print(f"loss gpu {loss:.6f} bf16 {l_bf:.6f} f64 {l_64:.6f} torch {l_t:.6f}; grad rel err vs bf16 oracle "
          f"{rel(g, g_bf):.2e} vs f64 {rel(g, g_64):.2e} vs torch fp32 {rel(g, g_t):.2e}", flush=True)
    assert abs(loss - l_bf) < 1e-3 * abs(l_bf)
    assert abs(loss - l_64) < 2e-2 * abs(l_64)
    assert rel(g, g_bf) < 2e-2 and cos(g, g_bf) > 0.999
    assert cos(g, g_64) > 0.99 and cos(g, g_t) > 0.99
    # per layer as well: no layer's gradient is off
    for i, ((n, k), (w0, b0)) in enumerate(zip(M.LAYERS, M.OFFSETS)):
        sl = slice(w0, b0 + n)
        assert cos(g[sl], g_bf[sl]) > 0.995, i
    # weighted negative log-likelihood (heavy-tailed weights, as estimates are)
    wts = np.random.default_rng(9).lognormal(0.0, 1.5, len(x)).astype(np.float32)
    loss_w, gw = mlp.grad(x, uv, wts)
    lw_bf, gw_bf = M.loss_and_grad(p, x, uv, "bf16", wts)
    lw_64, gw_64 = M.loss_and_grad(p, x, uv, "f64", wts)
    print(f"weighted: loss gpu {loss_w:.6f} bf16 {lw_bf:.6f} f64 {lw_64:.6f}; grad rel err vs bf16 "
          f"{rel(gw.astype(np.float64), gw_bf):.2e} vs f64 {rel(gw.astype(np.float64), gw_64):.2e}", flush=True)
    assert abs(loss_w - lw_bf) < 1e-3 * abs(lw_bf)
    assert rel(gw.astype(np.float64), gw_bf) < 2e-2 and cos(gw.astype(np.float64), gw_64) > 0.99
    mlp.close()
    ctx.close()


def test_adam_step(gpu, pool_scene):
    ctx = gpu.Context(pool_scene, abi.SmsConfig.default())
    mlp = gpu.Mlp(ctx, seed=7)
    p0 = mlp.get_params()
    x, uv = _batch(512, 8)
    _, g = mlp.grad(x, uv)
    mlp.adam(1e-3)
    p1 = mlp.get_params()
    z = np.zeros_like(p0)
    ref, m, v = M.adam(p0, g, z, z, 1, 1e-3)
    np.testing.assert_allclose(p1, ref, rtol=1e-6, atol=1e-9)
    # the BF16 copy follows the master parameters
    out = mlp.infer(x[:128])
    ref_bf, _ = M.forward(p1, x[:128], "bf16")
    assert np.abs(out - ref_bf).max() / np.abs(ref_bf).max() < 5e-3
    # second step uses the kept moments
    _, g2 = mlp.grad(x, uv)
    mlp.adam(1e-3)
    ref2, _, _ = M.adam(p1, g2, m, v, 2, 1e-3)
    np.testing.assert_allclose(mlp.get_params(), ref2, rtol=1e-6, atol=1e-9)
    mlp.close()
    ctx.close()


def test_train_loop_matches_steps_and_oracle(gpu, pool_scene):
    """sms_mlp_train: bit-identical to the same steps through sms_mlp_grad + sms_mlp_adam,
    bit-reproducible from run to run, and following the oracle's BF16 training trajectory."""
    ctx = gpu.Context(pool_scene, abi.SmsConfig.default())
    x, uv = _batch(3000, 12)
    wts = np.random.default_rng(13).lognormal(0.0, 1.0, len(x)).astype(np.float32)
    bufs = [gpu.DeviceBuffer(ctx, a.nbytes) for a in (x, uv, wts)]
    for b, a in zip(bufs, (x, uv, wts)):
        b.upload(a)
    batch, offs = 1024, np.array([0, 1500, 977, 1976, 3], np.int64)
    mlps = [gpu.Mlp(ctx, seed=21) for _ in range(3)]
    p0 = mlps[0].get_params()
    losses = mlps[0].train_device(bufs[0], bufs[1], len(x), batch, offs, 1e-3, wbuf=bufs[2], want_losses=True)
    step_losses = []
    for o in offs:
        loss, _ = mlps[1].grad_device(bufs[0], bufs[1], batch, want_loss=True, offset=int(o), wbuf=bufs[2])
        step_losses.append(loss)
        mlps[1].adam(1e-3)
    mlps[2].train_device(bufs[0], bufs[1], len(x), batch, offs, 1e-3, wbuf=bufs[2])
    pa, pb, pc = (m.get_params() for m in mlps)
    assert np.array_equal(pa, pb), "train loop differs from grad + adam steps"
    assert np.array_equal(pa, pc), "training is not reproducible"
    assert np.array_equal(losses, np.array(step_losses, np.float32))
    # oracle trajectory (BF16 restatement, float32 Adam)
    p, m, v = p0.copy(), np.zeros_like(p0), np.zeros_like(p0)
    ref_losses = []
    for t, o in enumerate(offs, 1):
        sl = slice(int(o), int(o) + batch)
        lo, g = M.loss_and_grad(p, x[sl], uv[sl], "bf16", wts[sl])
        ref_losses.append(lo)
        p, m, v = M.adam(p, g, m, v, t, 1e-3)
    print(f"losses gpu {losses} oracle {np.array(ref_losses)}; params max diff {np.abs(pa - p).max():.2e}", flush=True)
    np.testing.assert_allclose(losses, ref_losses, rtol=2e-3)
    # Adam moves every parameter by about lr per step whatever the gradient's size, so a
    # gradient entry near zero may take the other sign: most parameters must agree closely
    close = np.abs(pa - p) < 2e-4
    assert close.mean() > 0.97, close.mean()
    assert np.abs(pa - p).max() < 1.2e-2
    with pytest.raises(ValueError):
        mlps[0].train_device(bufs[0], bufs[1], len(x), batch, [len(x) - batch + 1], 1e-3)
    for b in bufs:
        b.free()
    for mm in mlps:
        mm.close()
    ctx.close()


def _constant_mixture(mlp, centre=(0.5, 0.5), sigma=0.08):
    """Network whose output is one fixed mixture for every query (last layer: zero weights)."""
    p = mlp.get_params()
    w0, b0 = M.OFFSETS[4]
    p[w0:b0] = 0.0
    b = np.zeros(48, np.float32)
    for c in range(8):
        b[6 * c] = 0.0
        b[6 * c + 1] = math.log(centre[0] / (1 - centre[0])) + 0.3 * (c % 4 - 1.5)
        b[6 * c + 2] = math.log(centre[1] / (1 - centre[1])) + 0.3 * (c // 4 - 0.5)
        b[6 * c + 3] = b[6 * c + 4] = math.log(sigma)
        b[6 * c + 5] = 0.1 * (c - 4)
    p[b0:b0 + 48] = b
    mlp.set_params(p)


@pytest.mark.parametrize("kind,radius", [("random", 0.0), ("constant", 0.0), ("random", 0.125), ("constant", 0.125)])
def test_estimator_with_proposal_matches_oracle(gpu, oracle_mod, pool_scene, kind, radius):
    cfg = abi.SmsConfig.default(max_trials=100000)
    ctx = gpu.Context(pool_scene, cfg)
    orc = oracle_mod.Oracle(pool_scene, cfg)
    mlp = gpu.Mlp(ctx, seed=11)
    if kind == "constant":
        _constant_mixture(mlp)
    ctx.set_proposal(mlp, 0, 0.2)
    ctx.set_proposal_anchor(radius)
    q = _floor_queries(pool_scene, 3000, 4, extent=0.7)
    raw = ctx.proposal_outputs(q, 9)
    ro, so = orc.ss_estimate_prop(q, 9, raw, 0, 0.2, radius=radius)
    for first_pass in (0, 1):
        ctx.set_schedule(ss_first_pass=first_pass)
        rg, sg = ctx.ss_estimate(q, 9)
        print(kind, first_pass, "gpu", sg, "oracle", so, flush=True)
        for k in ("walks", "walks_converged", "newton_iterations", "bernoulli_calls", "bernoulli_trials",
                  "bernoulli_max"):
            assert sg[k] == so[k], k
        assert (ro[:, 0] > 0).mean() > 0.05
        assert np.array_equal(rg, ro, equal_nan=True)
    # biased estimator with the same seeds
    cfg_b = abi.SmsConfig.default(biased=1, max_trials=32)
    ctx.set_config(cfg_b)
    orc.set_config(cfg_b)
    rg, sg = ctx.ss_estimate(q[:1000], 9)
    ro, so = orc.ss_estimate_prop(q[:1000], 9, raw[:1000], 0, 0.2, radius=radius)
    assert sg["walks"] == so["walks"] and sg["walks_converged"] == so["walks_converged"]
    assert np.array_equal(rg, ro, equal_nan=True)
    mlp.close()
    ctx.close()
    orc.close()


def test_anchors_match_oracle(gpu, oracle_mod, pool_scene):
    """Anchors of camera-pass queries (every pixel, receivers or not) bit-equal to the oracle's."""
    cfg = abi.SmsConfig.default(max_trials=100000)
    ctx = gpu.Context(pool_scene, cfg)
    orc = oracle_mod.Oracle(pool_scene, cfg)
    ctx.set_proposal(None, 0)
    q, _ = orc.ss_queries(1, 21)
    q = np.concatenate([q, _floor_queries(pool_scene, 4000, 6, extent=1.0)])
    ag = ctx.ss_proposal_anchors(q, 13)
    ao = orc.prop_anchors(q, 13, 0)
    have = np.isfinite(ao[:, 0])
    print(f"{have.sum()} of {len(q)} queries have anchors", flush=True)
    assert have.sum() > len(q) // 4 and (~have).sum() > 0
    assert np.array_equal(ag, ao, equal_nan=True)
    ctx.close()
    orc.close()


def test_anchored_capture(gpu, oracle_mod, pool_scene):
    """Anchored training records: the same estimates as absolute records, targets mapped to the
    anchor's box xi = (uv - a) / (2 r) + 1/2; queries without anchors record nothing."""
    cfg = abi.SmsConfig.default(max_trials=100000)
    ctx = gpu.Context(pool_scene, cfg)
    ctx.set_proposal(None, 0)
    q = _floor_queries(pool_scene, 4000, 5, extent=1.0)
    recs = {}
    for radius in (0.0, 0.125):
        ctx.set_proposal_anchor(radius)
        ctx.set_training_capture(1 << 16, abi.CAPTURE_ESTIMATES)
        rg, _ = ctx.ss_estimate(q, 3)
        recs[radius] = ctx.training_samples()
        ctx.set_training_capture(0)
    x0, uv0, w0 = recs[0.0]
    x1, xi1, w1 = recs[0.125]
    xin = ctx.proposal_inputs(q, 3)
    rows = {r.tobytes(): i for i, r in enumerate(xin)}
    i0 = np.array([rows[r.tobytes()] for r in x0])
    i1 = np.array([rows[r.tobytes()] for r in x1])
    a = ctx.ss_proposal_anchors(q, 3)
    have = np.isfinite(a[:, 0])
    assert sorted(i1.tolist()) == sorted(i for i in i0.tolist() if have[i])
    assert len(i1) < len(i0) or have[i0].all()
    m0 = dict(zip(i0.tolist(), zip(uv0, w0)))
    for j, i in enumerate(i1.tolist()):
        uv, w = m0[i]
        assert w1[j] == w
        np.testing.assert_allclose(a[i] + 0.25 * (xi1[j] - 0.5), uv, atol=2e-6)
    ctx.close()


def test_capture(gpu, oracle_mod, pool_scene):
    cfg = abi.SmsConfig.default(max_trials=100000)
    ctx = gpu.Context(pool_scene, cfg)
    orc = oracle_mod.Oracle(pool_scene, cfg)
    ctx.set_proposal(None, 0)
    ctx.set_training_capture(1 << 16)
    q = _floor_queries(pool_scene, 4000, 5)
    rg, _ = ctx.ss_estimate(q, 3)
    x, uv, w = ctx.training_samples()
    assert (w == 1).all()
    xin = ctx.proposal_inputs(q, 3)
    rows = {r.tobytes(): i for i, r in enumerate(xin)}
    idx = np.array([rows.get(r.tobytes(), -1) for r in x])
    print(f"captured {len(x)} solutions of {len(np.unique(idx))} of {len(q)} queries; nonzero estimates "
          f"{(rg[:, 0] > 0).sum()}", flush=True)
    assert len(x) > 0 and (idx >= 0).all()
    assert ((uv >= 0) & (uv <= 1)).all()
    # every converged walk is captured: each non-zero estimate's first walk and its matching trial
    nz = np.nonzero(rg[:, 0] > 0)[0]
    assert set(nz.tolist()) <= set(idx.tolist())
    counts = np.bincount(idx, minlength=len(q))
    assert (counts[nz] >= 2).all()
    # the estimate is unchanged by capturing
    ctx.set_training_capture(0)
    rg2, _ = ctx.ss_estimate(q, 3)
    assert np.array_equal(rg, rg2, equal_nan=True)
    # capture accumulates over calls until cleared
    ctx.set_training_capture(1 << 16)
    ctx.ss_estimate(q, 3)
    n0 = ctx.training_count()
    ctx.ss_estimate(q, 4)
    assert ctx.training_count() > n0
    ctx.clear_training()
    assert ctx.training_count() == 0
    # estimate-weighted records: one per non-zero estimate, weight = the estimate's RGB mean
    ctx.set_training_capture(1 << 16, abi.CAPTURE_ESTIMATES)
    rg, _ = ctx.ss_estimate(q, 3)
    x, uv, w = ctx.training_samples()
    idx = np.array([rows.get(r.tobytes(), -1) for r in x])
    nz = np.nonzero(rg[:, 0] > 0)[0]
    assert (idx >= 0).all() and sorted(idx.tolist()) == sorted(nz.tolist())
    np.testing.assert_allclose(w, rg[idx].mean(1), rtol=1e-5)
    # records come out in an order that does not depend on scheduling: the walker and the lane
    # pass append them in different orders and return them identically
    outs = []
    for first_pass in (0, 1, 0):
        ctx.set_schedule(ss_first_pass=first_pass)
        ctx.clear_training()
        ctx.ss_estimate(q, 3)
        outs.append(ctx.training_samples())
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert np.array_equal(a, b)
    ctx.set_training_capture(0)
    ctx.close()
    orc.close()


def test_training_beats_uniform(gpu, pool_scene):
    """Capture estimate-weighted solutions under uniform seeding, train on the GPU, and check
    that the trained proposal converges more often, needs fewer Bernoulli trials and has a
    lower variance, with the estimate's mean unchanged within its standard error."""
    cfg = abi.SmsConfig.default(max_trials=100000)
    ctx = gpu.Context(pool_scene, cfg)
    ctx.set_proposal(None, 0)
    ctx.set_training_capture(1 << 18, abi.CAPTURE_ESTIMATES)
    t0 = time.time()
    seed = 100
    while ctx.training_count() < 150000 and seed < 200:
        ctx.ss_estimate(_floor_queries(pool_scene, 65536, seed), seed)
        seed += 1
    xb, ub, wb, n = ctx.training_samples_device()
    mlp = gpu.Mlp(ctx, seed=3)
    rng = np.random.default_rng(0)
    bs = 8192
    offs = rng.integers(0, n - bs, 400)
    hist = mlp.train_device(xb, ub, n, bs, offs, 1e-3, wbuf=wb, want_losses=True)
    losses = [float(hist[i]) for i in list(range(0, 400, 50)) + [399]]
    xb.free()
    ub.free()
    wb.free()
    print(f"samples {n} in {seed - 100} passes; losses {['%.3f' % l for l in losses]}; {time.time() - t0:.1f}s",
          flush=True)
    assert losses[-1] < losses[0] - 1.0
    q = _floor_queries(pool_scene, 65536, 999)
    ctx.set_training_capture(0)
    ctx.set_proposal(None, 0)
    ru, su = ctx.ss_estimate(q, 5)
    ctx.set_proposal(mlp, 0, 0.2)
    rl, sl = ctx.ss_estimate(q, 5)
    cu, cl = su["walks_converged"] / su["walks"], sl["walks_converged"] / sl["walks"]
    tu, tl = su["bernoulli_trials"] / max(su["bernoulli_calls"], 1), sl["bernoulli_trials"] / max(sl["bernoulli_calls"], 1)
    lu, ll = ru[:, 0].astype(np.float64), rl[:, 0].astype(np.float64)
    se = math.sqrt(lu.var() / len(lu) + ll.var() / len(ll))
    print(f"uniform: conv {cu:.4f} trials/estimate {tu:.1f} mean {lu.mean():.5f} var {lu.var():.4f}; "
          f"learned: conv {cl:.4f} trials/estimate {tl:.1f} mean {ll.mean():.5f} var {ll.var():.4f}; se {se:.5f}",
          flush=True)
    assert cl > 1.8 * cu
    assert tl < 0.6 * tu
    assert ll.var() < 0.7 * lu.var()
    assert abs(lu.mean() - ll.mean()) < 4 * se
    mlp.close()
    ctx.close()
