"""sms_ss_estimate_batch_noalloc, the estimator entry point that allocates nothing (its scratch is
a fixed device array of the library), called through the C ABI: the same results and counters as
sms_ss_estimate_batch, bit for bit, on two calls in a row (the second smaller), no growth of the
device's used memory, and sizes beyond the scratch rejected with an error."""
import ctypes as C

import numpy as np
import pytest

from smsamd import abi, scenes, workloads

pytestmark = pytest.mark.gpu

MAX_ITEMS = 1 << 18


@pytest.fixture(scope="module")
def gpu():
    from smsamd import gpu as g
    g.lib()
    return g


def _free_bytes():
    # the HIP runtime the library itself is linked against (dlsym through its handle), not
    # another copy that an earlier import may have loaded into the process
    hip = C.CDLL(abi.LIB_PATH)
    free, total = C.c_size_t(0), C.c_size_t(0)
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def _noalloc(gpu, ctx, qbuf, n, seed, outbuf):
    st = abi.SmsStats()
    rc = gpu.lib().sms_ss_estimate_batch_noalloc(ctx.h, qbuf.ptr, n, seed, outbuf.ptr, C.byref(st))
    return rc, st.as_dict()


def _run(gpu, sc, cfg, queries, first_pass=0):
    ctx = gpu.Context(sc, cfg)
    ctx.set_schedule(ss_first_pass=first_pass)
    nmax = max(len(q) for q in queries)
    qbuf = gpu.DeviceBuffer(ctx, nmax * abi.SS_QUERY.itemsize)
    out_a, out_b = gpu.DeviceBuffer(ctx, nmax * 12), gpu.DeviceBuffer(ctx, nmax * 12)
    small = queries[-1][:16].copy()
    qbuf.upload(small)
    ctx.ss_estimate_device(qbuf, len(small), 1, out_a)          # first launches; buffers for 16 queries
    ctx.sync()
    for k, q in enumerate(queries):
        n = len(q)
        qbuf.upload(q)
        seed = 7 + k
        free0 = _free_bytes()
        rc, st_b = _noalloc(gpu, ctx, qbuf, n, seed, out_b)
        assert rc == 0, gpu.lib().sms_ctx_last_error(ctx.h)
        free1 = _free_bytes()
        print(f"call {k}: n={n} noalloc used {free0 - free1} bytes", flush=True)
        assert free0 - free1 < (1 << 20)
        st_a = ctx.ss_estimate_device(qbuf, n, seed, out_a)
        ra = out_a.download(np.zeros((n, 3), np.float32))
        rb = out_b.download(np.zeros((n, 3), np.float32))
        print("allocating", st_a, "noalloc", st_b, flush=True)
        assert (ra[:, 0] > 0).sum() > 0
        assert np.array_equal(ra, rb, equal_nan=True)
        for f in ("walks", "walks_converged", "newton_iterations", "bernoulli_calls", "bernoulli_trials",
                  "bernoulli_max", "overflow_queries"):
            assert st_a[f] == st_b[f], f
    return ctx, qbuf, out_b


@pytest.mark.parametrize("biased,first_pass", [(0, 0), (0, 1), (1, 0)])
def test_noalloc_matches_estimator_on_the_dome(gpu, biased, first_pass):
    sc = scenes.sphere_dome()
    cfg = abi.SmsConfig.default(biased=biased, max_trials=16 if biased else 100000)
    q1 = workloads.plane_queries(sc, 60000, 2, extent=0.9)
    q2 = workloads.plane_queries(sc, 5000, 3, extent=0.9)
    ctx, qbuf, out = _run(gpu, sc, cfg, [q1, q2], first_pass)
    # sizes beyond the scratch are refused, with the reason
    big = gpu.DeviceBuffer(ctx, (MAX_ITEMS + 1) * abi.SS_QUERY.itemsize)
    obig = gpu.DeviceBuffer(ctx, (MAX_ITEMS + 1) * 12)
    rc, _ = _noalloc(gpu, ctx, big, MAX_ITEMS + 1, 1, obig)
    assert rc == -1                                  # SMS_E_INVALID
    assert b"262144" in gpu.lib().sms_ctx_last_error(ctx.h)
    big.free(); obig.free(); qbuf.free(); out.free()
    ctx.close()


def test_noalloc_matches_estimator_on_the_pool(gpu):
    sc = scenes.pool(64, 64, n=256)
    cfg = abi.SmsConfig.default(max_trials=100000)
    q1 = workloads.plane_queries(sc, 40000, 2, extent=0.9, shape=1)
    q1["p"][:, 1] = -1.0
    q2 = q1[:3000].copy()
    ctx, qbuf, out = _run(gpu, sc, cfg, [q1, q2])
    qbuf.free(); out.free()
    ctx.close()
