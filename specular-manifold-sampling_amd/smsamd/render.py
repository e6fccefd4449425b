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
    """Per-pixel sums of weight * rgb and of its square over the samples taken (device)."""

    def __init__(self, ctx, pixels):
        self.ctx = ctx
        self.pixels = pixels
        self.sum = gpu.DeviceBuffer(ctx, pixels * 12)
        self.sq = gpu.DeviceBuffer(ctx, pixels * 12)
        self.spp = 0
        self.clear()

    def clear(self):
        z = np.zeros((self.pixels, 3), np.float32)
        self.sum.upload(z)
        self.sq.upload(z)
        self.spp = 0

    def download(self):
        """(mean, per-sample variance) images, float64, shape (pixels, 3)."""
        s = np.zeros((self.pixels, 3), np.float32)
        s2 = np.zeros((self.pixels, 3), np.float32)
        self.sum.download(s)
        self.sq.download(s2)
        n = max(self.spp, 1)
        mean = s.astype(np.float64) / n
        var = np.maximum(s2.astype(np.float64) / n - mean * mean, 0.0) * (n / max(n - 1, 1))
        return mean, var

    def free(self):
        self.sum.free()
        self.sq.free()


class PassRunner:
    """Device buffers for passes of ``pass_spp`` samples per pixel over the camera's frame."""

    def __init__(self, ctx, pass_spp):
        cam = ctx.scene.camera
        self.ctx = ctx
        self.pixels = cam["width"] * cam["height"]
        self.pass_spp = pass_spp
        self.n = self.pixels * pass_spp
        self.q = gpu.DeviceBuffer(ctx, self.n * abi.SS_QUERY.itemsize)
        self.w = gpu.DeviceBuffer(ctx, self.n * 4)
        self.rgb = gpu.DeviceBuffer(ctx, self.n * 12)

    def run(self, seed, film, stats=None):
        """One pass: camera samples of seed `seed`, estimates with the same seed, into film."""
        self.ctx.ss_queries_device(self.pass_spp, seed, self.q, self.w)
        st = self.ctx.ss_estimate_device(self.q, self.n, seed, self.rgb)
        self.ctx.film_accumulate_device(self.rgb, self.w, self.pixels, self.pass_spp, film.sum, film.sq)
        film.spp += self.pass_spp
        if stats is not None:
            for k, v in st.items():
                stats[k] = stats.get(k, 0) + v
        return st

    def free(self):
        for b in (self.q, self.w, self.rgb):
            b.free()


def train_online(ctx, runner, film, mlp, caster, samples=2_000_000, batch=8192, lr=1e-3, seed0=1 << 20,
                 stats=None, rng_seed=0, epochs=1, mode=abi.CAPTURE_SOLUTIONS, weight_power=1.0, deadline=None):
    """Render uniform-seed passes into `film` while capturing training records on mesh
    `caster` (`mode`: every converged solution with weight 1, or the solutions behind non-zero
    estimates weighted by the estimate, raised to `weight_power`); after each pass,
    `epochs` x (records // batch) Adam steps on windows of `batch` consecutive records at random
    offsets of that pass's capture, until `samples` records were trained on or, if given, the
    time.perf_counter() `deadline` has passed.  Returns (passes, steps, losses)."""
    rng = np.random.default_rng(rng_seed)
    ctx.set_proposal(None, caster)
    ctx.set_training_capture(runner.n, mode)
    used, steps, passes, losses = 0, 0, 0, []
    weighted = mode == abi.CAPTURE_ESTIMATES
    try:
        while used < samples and (deadline is None or time.perf_counter() < deadline):
            runner.run(seed0 + passes, film, stats)
            passes += 1
            xb, ub, wb, have = ctx.training_samples_device()
            ctx.clear_training()
            try:
                if weighted and weight_power != 1.0 and have:
                    w = np.zeros(have, np.float32)
                    wb.download(w)
                    wb.upload(np.power(np.maximum(w, 0.0), weight_power).astype(np.float32))
                n_new = min(have, samples - used) // batch * batch
                for _ in range(epochs * (n_new // batch)):
                    off = int(rng.integers(0, have - batch + 1))
                    loss, _ = mlp.grad_device(xb, ub, batch, want_loss=(steps % 32 == 0), offset=off,
                                              wbuf=wb if weighted else None)
                    if loss is not None:
                        losses.append(loss)
                    mlp.adam(lr)
                    steps += 1
                used += max(n_new, batch if have >= batch else 0)
            finally:
                xb.free()
                ub.free()
                wb.free()
            if have < batch and passes > 1000:
                raise RuntimeError("the caster yields too few converged solutions to train on")
    finally:
        ctx.set_training_capture(0)
    return passes, steps, losses


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
