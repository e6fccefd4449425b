"""Writes the golden fixtures of the solver path (inputs and the oracle's outputs):

  ss_sphere_walks.npz   5000 single-scatter walks, glass sphere over the floor (half with a
                        p_start seed override), angle-difference constraints
  glint_walks.npz       5000 glint walks on a 64x64 plane with a seeded 256^2 scratch map
  ms_slab_walks.npz     1000 two-vertex multi-scatter walks through a 32^2-cell slab
  tridiag.npz           200 block-tridiagonal systems (2..4 blocks): Newton step of the oracle and
                        a float64 dense solve
  mlp_64.npz            learned-proposal network on 64 samples: parameters, inputs, uv, raw outputs
                        and the mean-NLL gradient (float64 restatement and its BF16-rounding form)

Scenes come from smsamd.scenes with fixed seeds, so only walk records are stored.  The GPU
tests compare the HIP kernels with these records; the CPU tests check that the oracle still
reproduces them.

    python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "specular-manifold-sampling_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_oracle  # noqa: E402
import oracle_lib  # noqa: E402
from smsamd import abi, scenes, workloads  # noqa: E402


def ss_scene():
    return scenes.sphere_caustic(64, 64)


def glint_scene():
    return scenes.glint_plane(64, 64, nm_res=256, n_scratches=64, seed=3)


def ms_scene():
    return scenes.slab(width=32, height=32, n=32)


def ss_inputs():
    w = workloads.ss_walk_inputs(ss_scene(), 5000, 101, extent=1.2)
    rng = np.random.default_rng(102)
    half = np.arange(5000) % 2 == 1
    w["p_start"][half] = (np.array([0.0, 1.5, 0.0])
                          + workloads.uniform_sphere(rng.random(half.sum()), rng.random(half.sum()))).astype(np.float32)
    return w


def glint_inputs():
    return workloads.glint_walk_inputs(glint_scene(), 5000, 103, footprint_texels=3.0)


def ms_inputs():
    return workloads.ms_walk_inputs(ms_scene(), 1000, 104, extent=0.35)


def tridiag_inputs():
    rng = np.random.default_rng(105)
    n = rng.integers(2, 5, 200)
    dprev = rng.normal(0, 0.3, (200, 4, 4)).astype(np.float32)
    dnext = rng.normal(0, 0.3, (200, 4, 4)).astype(np.float32)
    dcur = rng.normal(0, 0.3, (200, 4, 4)).astype(np.float32)
    dcur[:, :, 0] += 2.0
    dcur[:, :, 3] += 2.0
    C = rng.normal(0, 1.0, (200, 4, 2)).astype(np.float32)
    return n, dprev, dcur, dnext, C


def dense_step(n, dprev, dcur, dnext, C):
    A = np.zeros((2 * n, 2 * n))
    for i in range(n):
        A[2 * i:2 * i + 2, 2 * i:2 * i + 2] = dcur[i].reshape(2, 2)
        if i > 0:
            A[2 * i:2 * i + 2, 2 * i - 2:2 * i] = dprev[i].reshape(2, 2)
        if i + 1 < n:
            A[2 * i:2 * i + 2, 2 * i + 2:2 * i + 4] = dnext[i].reshape(2, 2)
    return np.linalg.solve(A, C[:n].reshape(-1).astype(np.float64)).reshape(n, 2)


def mlp_inputs():
    p = mlp_oracle.init_like_library(106)
    rng = np.random.default_rng(107)
    for (nn, k), (w0, b0) in zip(mlp_oracle.LAYERS, mlp_oracle.OFFSETS):
        p[b0:b0 + nn] = rng.normal(0, 0.1, nn)
    p[mlp_oracle.OFFSETS[4][0]:mlp_oracle.OFFSETS[4][1]] *= 10.0
    x = rng.uniform(-1, 1, (64, 9)).astype(np.float32)
    uv = rng.uniform(0.05, 0.95, (64, 2)).astype(np.float32)
    return p.astype(np.float32), x, uv


def main():
    cfg = abi.SmsConfig.default()
    orc = oracle_lib.Oracle(ss_scene(), cfg)
    w = ss_inputs()
    np.savez(os.path.join(HERE, "ss_sphere_walks.npz"), inputs=w, outputs=orc.ss_walks(w))
    orc.close()
    cfg_g = abi.SmsConfig.default(max_trials=100000)
    orc = oracle_lib.Oracle(glint_scene(), cfg_g)
    w = glint_inputs()
    np.savez(os.path.join(HERE, "glint_walks.npz"), inputs=w, outputs=orc.glint_walks(w))
    orc.close()
    cfg_m = abi.SmsConfig.default(bounces=2)
    orc = oracle_lib.Oracle(ms_scene(), cfg_m)
    w = ms_inputs()
    np.savez(os.path.join(HERE, "ms_slab_walks.npz"), inputs=w, outputs=orc.ms_walks(w))
    orc.close()
    lib = oracle_lib.load()
    n, dprev, dcur, dnext, C = tridiag_inputs()
    dx = np.zeros((200, 4, 2), np.float32)
    ok = np.zeros(200, np.int32)
    dense = np.zeros((200, 4, 2))
    for i in range(200):
        out = np.zeros((4, 2), np.float32)
        ok[i] = lib.oracle_ms_tridiag_step(int(n[i]), np.ascontiguousarray(dprev[i]).ctypes.data,
                                           np.ascontiguousarray(dcur[i]).ctypes.data,
                                           np.ascontiguousarray(dnext[i]).ctypes.data,
                                           np.ascontiguousarray(C[i]).ctypes.data, out.ctypes.data)
        dx[i] = out
        dense[i, :n[i]] = dense_step(int(n[i]), dprev[i], dcur[i], dnext[i], C[i])
    np.savez(os.path.join(HERE, "tridiag.npz"), n=n, dprev=dprev, dcur=dcur, dnext=dnext, C=C, dx=dx, ok=ok,
             dense=dense)
    p, x, uv = mlp_inputs()
    o64, _ = mlp_oracle.forward(p, x, "f64")
    obf, _ = mlp_oracle.forward(p, x, "bf16")
    l64, g64 = mlp_oracle.loss_and_grad(p, x, uv, "f64")
    lbf, gbf = mlp_oracle.loss_and_grad(p, x, uv, "bf16")
    np.savez(os.path.join(HERE, "mlp_64.npz"), params=p, x=x, uv=uv, out_f64=o64, out_bf16=obf,
             loss_f64=np.float64(l64), grad_f64=g64, loss_bf16=np.float64(lbf), grad_bf16=gbf.astype(np.float32))
    for f in sorted(os.listdir(HERE)):
        print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
