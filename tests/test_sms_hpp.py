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


# ---------------------------------------------------------------- scene files through SceneFile
def _scene_xml(tmp_path, integrator, extra=""):
    from test_io import DOME_XML
    xml = DOME_XML.replace('<integrator type="path_sms_ss">', f'<integrator type="{integrator}">')
    xml = xml.replace('<integer name="max_depth" value="3"/>', f'<integer name="max_depth" value="6"/>{extra}')
    # the dome: the glass sphere centred on the receiver, a larger light
    xml = xml.replace('<point name="center" x="0" y="1.5" z="0"/>', '<point name="center" x="0" y="0" z="0"/>')
    xml = xml.replace('<scale value="0.1"/>', '<scale value="0.5"/>').replace('value="100"', 'value="4"')
    p = tmp_path / f"{integrator}.xml"
    p.write_text(xml)
    return p


@pytest.mark.parametrize("integrator,code", [("path_sms_ss", abi.PATH_SMS_SS), ("path_filtered_ss", abi.PATH_FILTERED_SS),
                                             ("path_sms_ms", abi.PATH_SMS_MS), ("path_filtered_ms", abi.PATH_FILTERED_MS)])
def test_scene_file_integrator(tmp_path, integrator, code):
    """SceneFile reads the file's integrator and creates it with the file's properties."""
    extra = '<integer name="bounces" value="3"/><boolean name="biased_mnee" value="true"/>' if "ms" in integrator else ""
    p = _scene_xml(tmp_path, integrator, extra)
    out = subprocess.run([_exe(), "scene", str(p), "res=20"], check=True, capture_output=True, text=True).stdout
    lines = out.strip().splitlines()
    assert lines[0] == f"integrator {integrator}"
    v = dict(line.split() for line in lines[1:])
    f = {k: float(x) for k, x in v.items()}
    assert f["spp"] == 16 and f["shapes"] == 3 and f["emitters"] == 1 and f["notes"] == 0
    assert f["file.integrator"] == code and f["file.max_depth"] == 6
    if integrator == "path_sms_ms":
        assert f["file.bounces"] == 3 and f["file.biased_mnee"] == 1 and f["file.max_trials"] == 64
    if integrator == "path_filtered_ms":
        assert f["file.bounces"] == 3 and f["file.biased_mnee"] == 0
    if integrator.endswith("_ss"):
        assert f["file.bounces"] == 1
    # glints are rendered by their camera-pass integrator, not a path integrator
    q = _scene_xml(tmp_path, "path_sms_glints")
    r = subprocess.run([_exe(), "scene", str(q)], capture_output=True, text=True)
    assert r.returncode == 1 and "not a built path integrator" in r.stderr


@pytest.mark.gpu
def test_scene_file_render_matches_oracle(tmp_path):
    """SceneFile -> make_path_integrator -> render, bit-equal to the oracle's render of the same
    file read through the Python reader."""
    import oracle_lib
    from smsamd import files
    oracle_lib.build()
    p = _scene_xml(tmp_path, "path_sms_ss")
    out_file = tmp_path / "img.bin"
    r = subprocess.run([_exe(), "scene-render", str(p), str(out_file), "24", "res=20"], capture_output=True,
                       text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    got = np.fromfile(out_file, np.float32)
    sc, info = files.load_scene(p, defines={"res": 20})
    props = info["properties"]
    cfg = abi.SmsConfig.default(biased=int(props["biased"]), max_trials=props["max_trials"],
                                uniqueness_threshold=props["uniqueness_threshold"])
    orc = oracle_lib.Oracle(sc, cfg)
    px = 20 * 20
    img, wt = np.zeros((px, 3), np.float32), np.zeros(px, np.float32)
    for spp, seed in ((8, 7), (8, 8), (8, 9)):
        rgb, pos, _ = orc.path_samples(abi.SmsPathConfig.default(max_depth=6), spp, seed)
        orc.film_splat(rgb, pos, spp, 0.5, img, wt)
    dev = np.where(wt[:, None] > 0, img / np.where(wt > 0, wt, 1)[:, None], 0).astype(np.float32)
    assert got.size == 3 * px and (dev > 0).mean() > 0.3
    assert np.array_equal(got, dev.ravel())
    orc.close()
