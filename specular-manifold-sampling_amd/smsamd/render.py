"""Frame rendering around the single-scatter estimator, all on the GPU through the C ABI:
camera pass (sms_ss_make_queries) -> estimator (sms_ss_estimate_batch) -> box film
(sms_film_accumulate), in passes of a few samples per pixel; and the online training of the
learned seed proposal from the solutions those passes capture.

The caustic term only (what path_sms_ss adds at the first receiver,
src/integrators/path_sms_ss.cpp:118-124): pixel value = mean over samples of path weight x
SMS estimate.
"""
import time

import numpy as np

from . import abi, gpu


class Film:
    """Per-pixel sums of weight * rgb, of its square and the count of samples kept (device);
    invalid samples are dropped by the film kernel, as the reference's image block drops them."""

    def __init__(self, ctx, pixels):
        self.ctx = ctx
        self.pixels = pixels
        self.sum = gpu.DeviceBuffer(ctx, pixels * 12)
        self.sq = gpu.DeviceBuffer(ctx, pixels * 12)
        self.n = gpu.DeviceBuffer(ctx, pixels * 4)
        self.spp = 0
        self.clear()

    def clear(self):
        z = np.zeros((self.pixels, 3), np.float32)
        self.sum.upload(z)
        self.sq.upload(z)
        self.n.upload(np.zeros(self.pixels, np.float32))
        self.spp = 0

    def download(self):
        """(mean, per-sample variance) images, float64, shape (pixels, 3)."""
        s = np.zeros((self.pixels, 3), np.float32)
        s2 = np.zeros((self.pixels, 3), np.float32)
        cnt = np.zeros(self.pixels, np.float32)
        self.sum.download(s)
        self.sq.download(s2)
        self.n.download(cnt)
        n = np.maximum(cnt.astype(np.float64), 1.0)[:, None]
        mean = s.astype(np.float64) / n
        var = np.maximum(s2.astype(np.float64) / n - mean * mean, 0.0) * (n / np.maximum(n - 1, 1))
        return mean, var

    def dropped(self):
        """samples the film dropped as invalid"""
        cnt = np.zeros(self.pixels, np.float32)
        self.n.download(cnt)
        return int(self.spp * self.pixels - cnt.astype(np.int64).sum())

    def free(self):
        self.sum.free()
        self.sq.free()
        self.n.free()


class PassRunner:
    """Device buffers for passes of ``pass_spp`` samples per pixel over the camera's frame;
    ``estimator``: "ss" (single scatter) or "ms" (multi scatter) at the first receiver, or
    "glint" (the glint estimator at the camera ray's glinty hit, sample weight 1)."""

    def __init__(self, ctx, pass_spp, estimator="ss"):
        cam = ctx.scene.camera
        self.ctx = ctx
        self.pixels = cam["width"] * cam["height"]
        self.pass_spp = pass_spp
        self.n = self.pixels * pass_spp
        if estimator not in ("ss", "ms", "glint"):
            raise ValueError(estimator)
        self.estimator = estimator
        qsize = abi.GLINT_QUERY.itemsize if estimator == "glint" else abi.SS_QUERY.itemsize
        self.q = gpu.DeviceBuffer(ctx, self.n * qsize)
        self.w = gpu.DeviceBuffer(ctx, self.n * 4)
        self.rgb = gpu.DeviceBuffer(ctx, self.n * 12)
        if estimator == "glint":
            self.w.upload(np.ones(self.n, np.float32))
        self.estimate = {"ss": ctx.ss_estimate_device, "ms": ctx.ms_estimate_device,
                         "glint": ctx.glint_estimate_device}[estimator]

    def run(self, seed, film, stats=None):
        """One pass: camera samples of seed `seed`, estimates with the same seed, into film."""
        if self.estimator == "glint":
            self.ctx.glint_queries_device(self.pass_spp, seed, self.q)
        else:
            self.ctx.ss_queries_device(self.pass_spp, seed, self.q, self.w)
        st = self.estimate(self.q, self.n, seed, self.rgb)
        self.ctx.film_accumulate_device(self.rgb, self.w, self.pixels, self.pass_spp, film.sum, film.sq, film.n)
        film.spp += self.pass_spp
        if stats is not None:
            for k, v in st.items():
                stats[k] = max(stats.get(k, 0), v) if k == "bernoulli_max" else stats.get(k, 0) + v
        return st

    def free(self):
        for b in (self.q, self.w, self.rgb):
            b.free()


def render_online(ctx, runner, film, mlp, caster, budget_s, samples=2_000_000, batch=8192, lr=1e-3, epochs=16,
                  uniform_passes=1, defensive=0.2, seed0=1 << 24, rng_seed=0, stats=None, max_steps_per_pass=None,
                  stop_when_trained=False, uniform_runner=None):
    """One frame of `budget_s` seconds whose seed proposal is trained while it renders.  The
    first `uniform_passes` passes use uniform seeds, every later pass the defensive mixture with
    the network as trained so far (the uniform passes with `uniform_runner`, a runner of fewer
    samples per pixel, when given).  Each pass captures estimate-weighted records on mesh
    `caster` (a glint runner: its converged first walks, weight 1); a record's weight is its estimate, so the expected total weight at a solution is
    the solution's contribution whatever proposal found it, and records of learned passes are
    as valid as uniform ones.  After a pass, epochs x (records // batch) Adam steps (at most
    max_steps_per_pass) on windows at random offsets of its records run before the next pass,
    until fewer than `batch` of `samples` records are left (stop_when_trained: the frame ends
    then).  Every pass
    counts in `film`; full passes stop when the next one would end after the budget, then passes
    of `uniform_runner` (with the proposal as trained) fill the time left.  Returns a dict
    of passes, steps, losses (every 32nd step), used (records trained on) and train_s (seconds
    spent training)."""
    rng = np.random.default_rng(rng_seed)
    ctx.set_proposal(None, caster)
    # glint estimates capture their converged first walks (weight 1) as footprint samples
    mode = abi.CAPTURE_SOLUTIONS if runner.estimator == "glint" else abi.CAPTURE_ESTIMATES
    ctx.set_training_capture(max(runner.n, uniform_runner.n if uniform_runner else 0), mode)
    t0 = time.perf_counter()
    out = {"passes": 0, "steps": 0, "losses": [], "used": 0, "train_s": 0.0}
    capturing, last, last_small = True, 0.0, None
    try:
        while True:
            t = time.perf_counter()
            first = uniform_runner is not None and out["passes"] < uniform_passes
            if out["passes"] and (t - t0 + last > budget_s or (stop_when_trained and not capturing)):
                # close the frame with passes of the smaller runner while they fit
                if stop_when_trained or uniform_runner is None or last_small is None:
                    break
                est = min(last_small, last * uniform_runner.pass_spp / runner.pass_spp)
                if t - t0 + est > budget_s:
                    break
                if not first:
                    ctx.set_proposal(mlp, caster, defensive)
                uniform_runner.run(seed0 + out["passes"], film, stats)
                out["passes"] += 1
                ctx.sync()
                last_small = time.perf_counter() - t
                last = float("inf")
                continue
            (uniform_runner if first else runner).run(seed0 + out["passes"], film, stats)
            out["passes"] += 1
            if capturing:
                xb, ub, wb, have = ctx.training_samples_device()
                ctx.clear_training()
                try:
                    n_new = min(have, samples - out["used"]) // batch * batch
                    n_steps = epochs * (n_new // batch)
                    if max_steps_per_pass is not None:
                        n_steps = min(n_steps, max_steps_per_pass)
                    if n_steps:
                        offs = rng.integers(0, have - batch + 1, n_steps)
                        t1 = time.perf_counter()
                        hist = mlp.train_device(xb, ub, have, batch, offs, lr, wbuf=wb, want_losses=True)
                        out["train_s"] += time.perf_counter() - t1
                        out["losses"].extend(float(v) for v in hist[(-out["steps"]) % 32::32])
                        out["steps"] += n_steps
                    out["used"] += n_new
                finally:
                    xb.free()
                    ub.free()
                    wb.free()
                if samples - out["used"] < batch:          # no whole batch left to train on
                    capturing = False
                    ctx.set_training_capture(0)
                if out["passes"] >= uniform_passes:
                    ctx.set_proposal(mlp, caster, defensive)
            ctx.sync()
            if first:
                last_small = time.perf_counter() - t
            else:
                last = time.perf_counter() - t
    finally:
        ctx.set_training_capture(0)
    return out


def rmse_trimmed(img, ref, ref_var_of_mean, drop=1e-3):
    """rmse() over the pixels left after dropping the fraction `drop` with the largest squared
    error (a firefly-robust companion of the plain RMSE)."""
    d2 = (((img - ref) ** 2 - ref_var_of_mean).mean(1))
    keep = max(1, int(len(d2) * (1.0 - drop)))
    return float(np.sqrt(max(np.partition(d2, keep - 1)[:keep].mean(), 0.0)))


def rmse(img, ref, ref_var_of_mean, mask=None):
    """RMSE of img against ref with the reference's own noise removed: mean over pixels (and
    channels) of (img - ref)^2 - Var[ref mean], floored at 0."""
    d2 = (img - ref) ** 2 - ref_var_of_mean
    if mask is not None:
        d2 = d2[mask]
    return float(np.sqrt(max(d2.mean(), 0.0)))


def timed(fn):
    t = time.perf_counter()
    r = fn()
    return r, time.perf_counter() - t
