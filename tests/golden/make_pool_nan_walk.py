"""Writes pool_nan_walk.npy: one single-scatter walk on the 512^2 Gerstner pool whose Newton
step turns non-finite (its reprojection rays are NaN), found in the camera-query workload
(512x512x8 camera samples, seed 1007, query 1567059, caster 0, walk 10).  The walk inputs are
rebuilt from the oracle's camera pass, emitter sample and PCG32 walk stream.

    python tests/golden/make_pool_nan_walk.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "specular-manifold-sampling_amd"))
import oracle_lib  # noqa: E402
from smsamd import abi, scenes  # noqa: E402

SEED, QID, WALK = 1007, 1567059, 10
M64 = (1 << 64) - 1


def mix64(z):
    z = ((z ^ (z >> 30)) * 0xbf58476d1ce4e5b9) & M64
    z = ((z ^ (z >> 27)) * 0x94d049bb133111eb) & M64
    return z ^ (z >> 31)


def main():
    sc = scenes.pool(512, 512)
    cfg = abi.SmsConfig.default(max_trials=1000000, uniqueness_threshold=1e-5)
    orc = oracle_lib.Oracle(sc, cfg)
    q, _ = orc.ss_queries(8, SEED)
    g = q[QID]
    ei = np.zeros(1, abi.EI_DTYPE)
    p = np.ascontiguousarray(g["p"], np.float32)
    orc.lib.oracle_sample_emitter_interaction(orc.h, p.ctypes.data, SEED, QID, ei.ctypes.data)
    key = mix64(SEED ^ mix64((QID + 0x9E3779B97F4A7C15) & M64))
    u = np.zeros(2, np.uint32)
    orc.lib.oracle_pcg32_first(key, WALK + 1, u.ctypes.data, 2)
    w = np.zeros(1, abi.SS_WALK_IN)
    w["si_p"] = g["p"]
    w["shape"] = 0
    w["ei"] = ei[0]
    w["n_offset"] = (0, 0, 1)
    w["sample"] = ((u >> 9) | 0x3f800000).view(np.float32) - 1.0
    out = orc.ss_walks(w)
    print(w, out)
    np.save(os.path.join(HERE, "pool_nan_walk.npy"), w)
    orc.close()


if __name__ == "__main__":
    main()
