"""The C++ class interface include/sms.hpp, through the program tests/cpp/sms_hpp_check (built by
__graft_entry__.build()): the integrators' property mapping against the keys and defaults of
path_sms_ss.cpp:47-61, path_sms_ms.cpp:149-164, path_sms_glints.cpp:42-54 and
path_filtered_ss.cpp:33-35 (CPU), and renders through SingleScatterSMSPathIntegrator,
SingleScatterFilteredPathIntegrator and SpecularManifoldSingleScatter bit-equal to the oracle
(GPU)."""
import os
import subprocess

import numpy as np
import pytest

from smsamd import abi, scenes

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "cpp", "sms_hpp_check")


def _exe():
    if not os.path.exists(EXE):
        pytest.skip("sms_hpp_check not built (__graft_entry__.build())")
    return EXE


def test_property_mapping():
    out = subprocess.run([_exe(), "config"], check=True, capture_output=True, text=True).stdout
    v = dict(line.split() for line in out.strip().splitlines())
    f = {k: float(x) for k, x in v.items()}
    # results/Figure_14_15_MainComparison/sphere_sms.xml:26-41
    assert f["sphere.max_depth"] == 6 and f["sphere.rr_depth"] == 5
    assert f["sphere.max_trials"] == 10_000_000 and f["sphere.max_iterations"] == 20
    assert f["sphere.solver_threshold"] == np.float32(1e-5) and f["sphere.uniqueness_threshold"] == np.float32(1e-5)
    assert f["sphere.biased"] == 0 and f["sphere.bounces"] == 1 and f["sphere.integrator"] == abi.PATH_SMS_SS
    assert f["sphere.filter_stddev"] == 0.5 and f["sphere.unqueried"] == 0
    # the MNEE setup of path_sms_ss.cpp:30-36; defaults where the keys are absent
    assert (f["mnee.mnee_init"], f["mnee.biased"], f["mnee.max_trials"], f["mnee.halfvector_constraints"],
            f["mnee.biased_mnee"]) == (1, 1, 1, 1, 1)
    assert f["mnee.uniqueness_threshold"] == np.float32(1e-4) and f["mnee.max_depth"] == 7
    assert f["filtered.integrator"] == abi.PATH_FILTERED_SS and f["filtered.disable_reflected_emitters"] == 1
    assert f["filtered.biased_mnee"] == 0
    assert f["ms.bounces"] == 2 and f["ms.step_scale"] == 0.5 and f["ms.max_trials"] == 32 and f["ms.biased"] == 1
    assert f["glints.sms_strategy_only"] == 1 and f["glints.bsdf_strategy_only"] == 0
    assert f["glints.solver_threshold"] == np.float32(4e-5) and f["glints.uniqueness_threshold"] == np.float32(2e-4)
    assert f["glints.max_trials"] == 100000 and f["glints.unqueried"] == 0
    assert f["type_error"] == -1                     # SMS_E_INVALID
    assert f["int_as_float"] == 2.0


@pytest.mark.gpu
def test_class_interface_renders_match_oracle(tmp_path):
    import oracle_lib
    oracle_lib.build()
    out_file = tmp_path / "render.bin"
    r = subprocess.run([_exe(), "render", str(out_file)], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    got = np.fromfile(out_file, np.float32)
    w, h = 24, 16
    px = w * h
    sc = scenes.sphere_dome(w, h, light_radius=0.5, light_radiance=4.0)
    cfg = abi.SmsConfig.default(max_trials=100000)
    orc = oracle_lib.Oracle(sc, cfg)

    def render(pc, passes):
        img, wt = np.zeros((px, 3), np.float32), np.zeros(px, np.float32)
        walks = 0
        for spp, seed in passes:
            rgb, pos, st = orc.path_samples(pc, spp, seed)
            orc.film_splat(rgb, pos, spp, 0.5, img, wt)
            walks += st["walks"]
        dev = np.where(wt[:, None] > 0, img / np.where(wt > 0, wt, 1)[:, None], 0).astype(np.float32)
        return dev.ravel(), walks

    a, wa = render(abi.SmsPathConfig.default(max_depth=6), [(16, 11), (16, 12), (8, 13)])
    b, wb = render(abi.SmsPathConfig.default(max_depth=6, caustics_only=1), [(16, 21)])
    c, _ = render(abi.SmsPathConfig.default(integrator=abi.PATH_FILTERED_SS, max_depth=6), [(8, 31), (8, 32)])
    q, _ = orc.ss_queries(2, 41)
    e, se = orc.ss_estimate(q, 41)
    n = 3 * px
    assert got.size == 3 * n + e.size + 3
    assert np.array_equal(got[:n], a)
    assert np.array_equal(got[n:2 * n], b)
    assert np.array_equal(got[2 * n:3 * n], c)
    assert np.array_equal(got[3 * n:3 * n + e.size], e.ravel())
    assert list(got[-3:]) == [float(wa), float(wb), float(se["walks"])]
    orc.close()
