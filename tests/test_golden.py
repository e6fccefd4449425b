"""The committed golden fixtures (tests/golden/, written by make_fixtures.py) against the oracle
(CPU): walks are reproduced bit for bit, the tridiagonal Newton steps agree with a float64
dense solve, and the network restatement reproduces its forward pass and gradient."""
import os
import sys

import numpy as np
import pytest

import oracle_lib
from smsamd import abi

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLD)
import make_fixtures as F  # noqa: E402
import mlp_oracle as M  # noqa: E402


def _load(name):
    d = np.load(os.path.join(GOLD, name), allow_pickle=False)
    return {k: oracle_lib.upgrade_ss_walk_in(d[k]) for k in d.files}


def test_fixture_inputs_are_regenerated_identically():
    assert np.array_equal(_load("ss_sphere_walks.npz")["inputs"], F.ss_inputs())
    assert np.array_equal(_load("glint_walks.npz")["inputs"], F.glint_inputs())
    assert np.array_equal(_load("ms_slab_walks.npz")["inputs"], F.ms_inputs())


@pytest.mark.parametrize("name,scene,cfg,fn", [
    ("ss_sphere_walks.npz", F.ss_scene, abi.SmsConfig.default, "ss_walks"),
    ("glint_walks.npz", F.glint_scene, lambda: abi.SmsConfig.default(max_trials=100000), "glint_walks"),
    ("ms_slab_walks.npz", F.ms_scene, lambda: abi.SmsConfig.default(bounces=2), "ms_walks"),
])
def test_oracle_reproduces_walk_fixture(name, scene, cfg, fn):
    d = _load(name)
    orc = oracle_lib.Oracle(scene(), cfg())
    out = getattr(orc, fn)(d["inputs"])
    assert out.tobytes() == d["outputs"].tobytes()
    ok = d["outputs"]["success"] == 1
    assert 0.005 < ok.mean() < 0.98
    orc.close()


def test_tridiagonal_fixture_matches_dense_solve():
    d = _load("tridiag.npz")
    assert (d["ok"] == 1).all()
    for i in range(len(d["n"])):
        n = int(d["n"][i])
        np.testing.assert_allclose(d["dx"][i, :n], d["dense"][i, :n], rtol=2e-4, atol=2e-5)


def test_mlp_fixture_reproduced():
    d = _load("mlp_64.npz")
    o64, _ = M.forward(d["params"], d["x"], "f64")
    np.testing.assert_array_equal(o64, d["out_f64"])
    obf, _ = M.forward(d["params"], d["x"], "bf16")
    np.testing.assert_array_equal(obf, d["out_bf16"])
    l64, g64 = M.loss_and_grad(d["params"], d["x"], d["uv"], "f64")
    assert l64 == float(d["loss_f64"])
    np.testing.assert_array_equal(g64, d["grad_f64"])
