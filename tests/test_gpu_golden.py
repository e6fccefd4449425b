"""The HIP kernels against the committed golden fixtures (tests/golden/): walks bit for bit,
the network's forward pass and gradient within the BF16 tolerances of its own tests."""
import os
import sys

import numpy as np
import pytest

import oracle_lib
from smsamd import abi

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLD)
import make_fixtures as F  # noqa: E402


@pytest.fixture(scope="module")
def gpu():
    from smsamd import gpu as g
    g.lib()
    return g


def _load(name):
    d = np.load(os.path.join(GOLD, name), allow_pickle=False)
    return {k: oracle_lib.upgrade_ss_walk_in(d[k]) for k in d.files}


@pytest.mark.parametrize("name,scene,cfg,fn", [
    ("ss_sphere_walks.npz", F.ss_scene, abi.SmsConfig.default, "ss_walks"),
    ("glint_walks.npz", F.glint_scene, lambda: abi.SmsConfig.default(max_trials=100000), "glint_walks"),
    ("ms_slab_walks.npz", F.ms_scene, lambda: abi.SmsConfig.default(bounces=2), "ms_walks"),
])
def test_gpu_walks_match_fixture(gpu, name, scene, cfg, fn):
    d = _load(name)
    ctx = gpu.Context(scene(), cfg())
    out = getattr(ctx, fn)(d["inputs"])
    print(name, "success", d["outputs"]["success"].mean(), flush=True)
    assert out.tobytes() == d["outputs"].tobytes()
    ctx.close()


def test_gpu_mlp_matches_fixture(gpu):
    d = _load("mlp_64.npz")
    ctx = gpu.Context(F.ss_scene(), abi.SmsConfig.default())
    mlp = gpu.Mlp(ctx, seed=0)
    mlp.set_params(d["params"])
    out = mlp.infer(d["x"])
    scale = np.abs(d["out_f64"]).max()
    assert np.abs(out - d["out_bf16"]).max() / scale < 5e-3
    loss, g = mlp.grad(d["x"], d["uv"])
    gb = d["grad_bf16"].astype(np.float64)
    rel = np.linalg.norm(g - gb) / np.linalg.norm(gb)
    print(f"mlp fixture: loss {loss:.6f} vs {float(d['loss_bf16']):.6f}, grad rel err {rel:.2e}", flush=True)
    assert abs(loss - float(d["loss_bf16"])) < 1e-3 * abs(float(d["loss_bf16"]))
    assert rel < 2e-2
    mlp.close()
    ctx.close()
