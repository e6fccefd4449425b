"""The op-counting build of the oracle (oracle/opcount.h, oracle/opcount.cpp) computes what the
oracle computes, bit for bit, and the frozen per-iteration FLOP counts the benchmark's rooflines
use (tests/golden/flop_counts.json, written by tests/golden/make_flop_counts.py from larger
samples of the benchmark workloads) agree with a recount on a small sample."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import oracle_lib
from smsamd import abi, scenes, workloads

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = json.load(open(os.path.join(HERE, "golden", "flop_counts.json")))


def _counting(sc, cfg):
    o = oracle_lib.Oracle(sc, cfg, threads=1, lib_path=oracle_lib.OPCOUNT_SO)
    o.lib.opcount_read.argtypes = [C.c_void_p]
    return o


def _flop32(o):
    v = np.zeros(12, np.uint64)
    o.lib.opcount_read(v.ctypes.data)
    add, mul, div, fma, sqrt = (int(x) for x in v[:5])
    return add + mul + div + 2 * fma + sqrt


def test_counting_build_is_the_oracle_and_counts_add_up():
    sc = scenes.sphere_dome(64, 64)
    cfg = abi.SmsConfig.default(biased=1, max_trials=64)
    plain = oracle_lib.Oracle(sc, cfg, threads=4)
    q, _ = plain.ss_queries(1, 1)
    q = np.ascontiguousarray(q[q["shape"] >= 0][:60])
    cnt = _counting(sc, cfg)
    cnt.lib.opcount_reset()
    a, sa = plain.ss_estimate(q, 1)
    b, sb = cnt.ss_estimate(q, 1)
    assert np.array_equal(a, b, equal_nan=True) and sa == sb
    per_iter = _flop32(cnt) / sb["newton_iterations"]
    ref = FIXTURE["dome_biased"]["flop32_per_iteration"]
    print(f"dome biased: {per_iter:.1f} FP32 FLOP per iteration (fixture {ref:.1f})")
    assert per_iter == pytest.approx(ref, rel=0.1)
    # walks: a second call adds to the per-thread counters
    w = workloads.ss_walk_inputs(sc, 200, 3, extent=0.9)
    cnt.lib.opcount_reset()
    o1 = cnt.ss_walks(w)
    f1 = _flop32(cnt)
    cnt.ss_walks(w)
    assert _flop32(cnt) == 2 * f1
    assert np.array_equal(o1["p"], plain.ss_walks(w)["p"])
    plain.close()
    cnt.close()


def test_fixture_is_complete():
    for k in ("dome_biased", "glint", "pool", "pool_walks", "slab"):
        r = FIXTURE[k]
        assert r["flop32"] == (r["fp32"]["add"] + r["fp32"]["mul"] + r["fp32"]["div"] + 2 * r["fp32"]["fma"]
                               + r["fp32"]["sqrt"])
        assert r["flop32_per_iteration"] == pytest.approx(r["flop32"] / r["newton_iterations"])
        assert 50 < r["flop32_per_iteration"] < 10000
