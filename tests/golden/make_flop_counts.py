"""Freeze the floating-point work of the SMS walks per Newton iteration with the op-counting
build of the CPU oracle (oracle/liboracle_opcount.so: the oracle's own source compiled with
counting float/double types, oracle/opcount.h), on samples of the benchmark workloads:

  dome_biased   sphere-caustic dome, camera queries of a 256^2 x 1 spp pass, biased M = 64
  glint         512^2 x 4 spp footprint queries on the 2048^2 scratch/bump map, unbiased
                (cap 1e5, thresholds 4e-5 / 2e-4)
  pool          1024^2 x 8 spp-pass camera queries on the 524,288-triangle Gerstner pool,
                unbiased (cap 1e7, thresholds 1e-5)
  slab          1024^2 camera queries on the value-noise slab, two-vertex chains, unbiased
  pool_walks    single pool walks (seed, BVH seed ray, Newton solve) without the estimator

For each: FP32 and FP64 adds, multiplies, divides, FMAs (2 FLOP), square roots and the
non-FLOP compare/abs/floor count, per Newton iteration of the sample (the estimator's whole work
-- emitter sample, seed rays, walks, contributions, shadow rays, trial loops -- divided by its
Newton iterations).  Before counting, every sample's outputs are checked equal, bit for bit, to
the plain oracle's: the counts are of the same computation.

Writes tests/golden/flop_counts.json.  Run: python tests/golden/make_flop_counts.py
"""
import ctypes as C
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "specular-manifold-sampling_amd"))
sys.path.insert(0, os.path.dirname(HERE))

import oracle_lib  # noqa: E402
from smsamd import abi, scenes, workloads  # noqa: E402

NAMES = ("add", "mul", "div", "fma", "sqrt", "misc")


def counting(sc, cfg):
    o = oracle_lib.Oracle(sc, cfg, threads=1, lib_path=oracle_lib.OPCOUNT_SO)
    o.lib.opcount_read.argtypes = [C.c_void_p]
    o.lib.opcount_reset.argtypes = []
    return o


def read(o):
    v = np.zeros(12, np.uint64)
    o.lib.opcount_read(v.ctypes.data)
    f32 = dict(zip(NAMES, (int(x) for x in v[:6])))
    f64 = dict(zip(NAMES, (int(x) for x in v[6:])))
    return f32, f64


def flop(d):
    return d["add"] + d["mul"] + d["div"] + 2 * d["fma"] + d["sqrt"]


def record(name, f32, f64, iters, walks, extra):
    r = {"fp32": f32, "fp64": f64, "flop32": flop(f32), "flop64": flop(f64), "newton_iterations": int(iters),
         "walks": int(walks), "flop32_per_iteration": flop(f32) / max(iters, 1),
         "flop64_per_iteration": flop(f64) / max(iters, 1), "flop32_per_walk": flop(f32) / max(walks, 1)}
    r.update(extra)
    print(f"{name}: {r['flop32_per_iteration']:.1f} FP32 FLOP + {r['flop64_per_iteration']:.2f} FP64 FLOP "
          f"per Newton iteration ({iters} iterations, {walks} walks)", flush=True)
    return r


def estimate_counts(name, sc, cfg, q, seed, kind, extra):
    plain = oracle_lib.Oracle(sc, cfg, threads=8)
    cnt = counting(sc, cfg)
    if kind == "ss":
        ref, st_ref = plain.ss_estimate(q, seed)
        cnt.lib.opcount_reset()
        got, st = cnt.ss_estimate(q, seed)
    elif kind == "ms":
        ref, st_ref = plain.ms_estimate(q, seed)
        cnt.lib.opcount_reset()
        got, st = cnt.ms_estimate(q, seed)
    else:
        ref, st_ref = plain.glint_estimate(q, seed)
        cnt.lib.opcount_reset()
        got, st = cnt.glint_estimate(q, seed)
    f32, f64 = read(cnt)
    assert np.array_equal(ref, got, equal_nan=True), f"{name}: counting build differs from the oracle"
    assert st == st_ref, name
    plain.close()
    cnt.close()
    return record(name, f32, f64, st["newton_iterations"], st["walks"], dict(extra, queries=len(q)))


def sample(q, n, shape_ok=True, seed=0):
    if shape_ok:
        q = q[q["shape"] >= 0]
    idx = np.random.default_rng(seed).choice(len(q), size=min(n, len(q)), replace=False)
    return np.ascontiguousarray(q[np.sort(idx)])


def main():
    subprocess_make()
    out = {"generator": "tests/golden/make_flop_counts.py",
           "counting": "oracle/opcount.h: add, sub = 1; mul = 1; div = 1; fma = 2; sqrt = 1 FLOP; "
                       "compare/abs/floor/min/max = misc, not counted as FLOP"}
    t0 = time.time()
    # C1: dome, biased M = 64
    sc = scenes.sphere_dome(256, 256)
    cfg = abi.SmsConfig.default(biased=1, max_trials=64)
    q, _ = oracle_lib.Oracle(sc, cfg, threads=8).ss_queries(1, 1)
    out["dome_biased"] = estimate_counts("dome_biased", sc, cfg, sample(q, 600), 1, "ss",
                                         {"workload": "sphere_dome 256^2 camera queries, biased M=64"})
    # C2: glint frame, full map
    sc = scenes.glint_plane(512, 512, nm_res=2048, n_scratches=4096, seed=1)
    cfg = abi.SmsConfig.default(max_trials=100000, solver_threshold=4e-5, uniqueness_threshold=2e-4)
    q = oracle_lib.Oracle(sc, cfg, threads=8).glint_queries(4, 1)
    out["glint"] = estimate_counts("glint", sc, cfg, sample(q, 60000), 1, "glint",
                                   {"workload": "glint 512^2x4spp footprint queries, 2048^2 map, unbiased"})
    # C3: pool
    sc = scenes.pool(1024, 1024)
    cfg = abi.SmsConfig.default(max_trials=10_000_000, uniqueness_threshold=1e-5, solver_threshold=1e-5)
    q, _ = oracle_lib.Oracle(sc, cfg, threads=8).ss_queries(8, 1)
    out["pool"] = estimate_counts("pool", sc, cfg, sample(q, 6000), 1, "ss",
                                  {"workload": "pool 1024^2x8spp camera queries, unbiased, uniform seeds"})
    # pool walks alone
    rng = np.random.default_rng(5)
    n = 4000
    w = np.zeros(n, abi.SS_WALK_IN)
    w["si_p"] = np.stack([rng.uniform(-0.9, 0.9, n), np.full(n, -1.0), rng.uniform(-0.9, 0.9, n)], -1)
    ls = sc.shapes[2]
    w["ei"]["p"] = np.asarray(ls["center"]) + ls["radius"] * workloads.uniform_sphere(rng.random(n), rng.random(n))
    w["ei"]["kind"] = abi.EI_AREA
    w["n_offset"] = (0, 0, 1)
    w["sample"] = rng.random((n, 2))
    plain = oracle_lib.Oracle(sc, cfg, threads=8)
    cnt = counting(sc, cfg)
    ref = plain.ss_walks(w)
    cnt.lib.opcount_reset()
    got = cnt.ss_walks(w)
    for k in ("success", "iterations", "p"):
        assert np.array_equal(ref[k], got[k]), k
    f32, f64 = read(cnt)
    out["pool_walks"] = record("pool_walks", f32, f64, int(ref["iterations"].sum()), n,
                               {"workload": "pool walks: uniform seed, seed ray, Newton solve"})
    plain.close()
    cnt.close()
    # C5: slab, two-vertex chains
    sc = scenes.slab(1024, 1024, n=256)
    cfg = abi.SmsConfig.default(bounces=2, max_trials=10_000_000, uniqueness_threshold=1e-5, solver_threshold=1e-5)
    q, _ = oracle_lib.Oracle(sc, cfg, threads=8).ss_queries(1, 1)
    out["slab"] = estimate_counts("slab", sc, cfg, sample(q, 1500), 1, "ms",
                                  {"workload": "slab 1024^2 camera queries, 2-vertex chains, unbiased"})
    out["seconds"] = round(time.time() - t0, 1)
    with open(os.path.join(HERE, "flop_counts.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


def subprocess_make():
    import subprocess
    subprocess.run(["make", "-s", "-C", oracle_lib.ORACLE_DIR, "liboracle.so", "liboracle_opcount.so"], check=True)


if __name__ == "__main__":
    main()
