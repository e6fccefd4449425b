"""Scenes read from files, rendered on the GPU through the C ABI and checked bit for bit
against the oracle: area lights on rectangles and meshes in both path integrators; a scene
file in the form of the reference's sphere caustic scenes (results/Figure_8_Constraints/
sphere.xml: a rectangle area light, a diffuse receiver, a dielectric sphere) written here with
an OBJ caster (a glass dome: the sphere centred on the receiver, so light reaches the floor
through one refraction); and single-scatter SMS estimates on a caster whose normal map comes from an
OpenEXR file."""
import numpy as np
import pytest

from smsamd import abi, files, scenes
from test_gpu_path import _compare
from test_io import OBJ, write_exr
from test_oracle_path import _area_light_scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    from smsamd import gpu as g
    g.lib()
    return g


@pytest.fixture(scope="module")
def oracle_mod():
    import oracle_lib
    oracle_lib.build()
    return oracle_lib


@pytest.mark.parametrize("light", ["rectangle", "mesh"])
def test_area_lights_on_rectangles_and_meshes(gpu, oracle_mod, light):
    sc, _ = _area_light_scene(light, -0.9, 0.6)
    sc.set_camera((0.0, 4.0, 3.0), (0.0, 0.0, 0.0), (0, 1, 0), 60.0, 24, 16)
    cfg = abi.SmsConfig.default(max_trials=1000)
    ctx = gpu.Context(sc, cfg)
    orc = oracle_mod.Oracle(sc, cfg)
    for integrator in (abi.PATH_SMS_SS, abi.PATH_FILTERED_SS):
        r = _compare(ctx, orc, abi.SmsPathConfig.default(integrator=integrator, max_depth=4), 8, 11)
        assert (r[:, 0] > 0).mean() > 0.5
    ctx.close()
    orc.close()


SPHERE_XML = """<scene version="2.0.0">
    <default name="spp" value="8"/>
    <integrator type="path_sms_ss">
        <integer name="max_depth" value="6"/>
        <integer name="max_trials" value="10000"/>
        <float name="solver_threshold" value="0.0001"/>
        <float name="uniqueness_threshold" value="0.0001"/>
    </integrator>
    <sensor type="perspective">
        <string name="fov_axis" value="smaller"/>
        <float name="fov" value="40"/>
        <transform name="to_world"><lookat origin="0, 3, 6" target="0, 0.5, 0" up="0, 1, 0"/></transform>
        <sampler type="independent"><integer name="sample_count" value="$spp"/></sampler>
        <film type="hdrfilm"><integer name="width" value="40"/><integer name="height" value="32"/>
            <rfilter type="gaussian"/></film>
    </sensor>
    <shape type="rectangle">
        <bsdf type="diffuse"><spectrum name="reflectance" value="0.0"/></bsdf>
        <emitter type="area">
            <boolean name="caustic_emitter_single" value="true"/>
            <spectrum name="radiance" value="300"/>
        </emitter>
        <transform name="to_world">
            <rotate x="1" y="0" z="0" angle="90"/>
            <scale x="0.2" y="1" z="0.2"/>
            <translate x="0" y="5" z="0"/>
            <rotate x="1" y="0" z="0" angle="20"/>
        </transform>
    </shape>
    <shape type="rectangle">
        <boolean name="caustic_receiver" value="true"/>
        <bsdf type="diffuse"><rgb name="reflectance" value="0.6, 0.5, 0.4"/></bsdf>
        <transform name="to_world"><scale value="4"/><rotate x="1" angle="-90"/></transform>
    </shape>
    <shape type="obj">
        <string name="filename" value="caster.obj"/>
        <boolean name="caustic_caster_single" value="true"/>
        <bsdf type="dielectric"><float name="int_ior" value="1.5"/><float name="ext_ior" value="1.0"/></bsdf>
        <transform name="to_world"><scale value="1.2"/></transform>
    </shape>
</scene>
"""


def _uv_sphere(n=24):
    th = np.linspace(0, np.pi, n + 1)
    ph = np.linspace(0, 2 * np.pi, 2 * n + 1)
    lines = []
    for t in th:
        for p in ph:
            lines.append(f"v {np.sin(t) * np.cos(p):.7f} {np.cos(t):.7f} {np.sin(t) * np.sin(p):.7f}")
            lines.append(f"vt {p / (2 * np.pi):.7f} {t / np.pi:.7f}")
    w = 2 * n + 1
    for i in range(n):
        for j in range(2 * n):
            a, b, c, d = i * w + j + 1, i * w + j + 2, (i + 1) * w + j + 2, (i + 1) * w + j + 1
            lines.append(f"f {a}/{a} {b}/{b} {c}/{c} {d}/{d}")
    return "\n".join(lines) + "\n"


def test_scene_file_renders_match_oracle(gpu, oracle_mod, tmp_path):
    (tmp_path / "caster.obj").write_text(_uv_sphere())
    (tmp_path / "s.xml").write_text(SPHERE_XML)
    sc, info = files.load_scene(tmp_path / "s.xml")
    p = info["properties"]
    cfg = abi.SmsConfig.default(max_trials=p["max_trials"], solver_threshold=p["solver_threshold"],
                                uniqueness_threshold=p["uniqueness_threshold"])
    ctx = gpu.Context(sc, cfg)
    orc = oracle_mod.Oracle(sc, cfg)
    for integrator in (abi.PATH_SMS_SS, abi.PATH_FILTERED_SS):
        for caustics_only in (1, 0):
            pc = abi.SmsPathConfig.default(integrator=integrator, max_depth=p["max_depth"], caustics_only=caustics_only)
            r = _compare(ctx, orc, pc, info["spp"], 5)
            if integrator == abi.PATH_SMS_SS and caustics_only:
                assert (r[:, 0] > 0).mean() > 0.02
    ctx.close()
    orc.close()


@pytest.mark.parametrize("encoded", [False, True])
def test_exr_normal_mapped_caster_estimates_match_oracle(gpu, oracle_mod, tmp_path, encoded):
    """A normal-mapped mirror whose map is read from an OpenEXR file (values as stored, no 8-bit
    quantisation), its frames from interpolated slopes or from decoded interpolated RGB:
    single-scatter estimates and two-stage walks equal to the oracle's."""
    rng = np.random.default_rng(4)
    res = 64
    u = np.linspace(0, 4 * np.pi, res)
    sx = 0.15 * np.sin(u)[None, :] + 0.02 * rng.standard_normal((res, res))
    sy = 0.15 * np.cos(u)[:, None] + 0.02 * rng.standard_normal((res, res))
    n = np.stack([-sx, -sy, np.ones_like(sx)], -1)
    n /= np.linalg.norm(n, axis=2, keepdims=True)
    write_exr(tmp_path / "nm.exr", (0.5 * n + 0.5).astype(np.float32), 3, 2)
    nm = files.load_image(tmp_path / "nm.exr")
    sc = scenes.mirror_wall(normalmap=np.zeros((res, res, 3), np.uint8))
    sc.normalmaps[0] = (res, nm)
    if encoded:
        sc.encoded_normalmaps.add(0)
    cfg = abi.SmsConfig.default(max_trials=10000)
    q = None
    for twostage in (0, 1):
        cfg.twostage = twostage
        ctx = gpu.Context(sc, cfg)
        orc = oracle_mod.Oracle(sc, cfg)
        if q is None:
            q, _ = orc.ss_queries(1, 3)
        rg, sg = ctx.ss_estimate(q, 17)
        ro, so = orc.ss_estimate(q, 17)
        print(f"twostage {twostage}: {so}", flush=True)
        assert so["walks_converged"] > 0
        assert np.array_equal(rg, ro, equal_nan=True)
        for k in ("walks", "walks_converged", "newton_iterations", "bernoulli_trials"):
            assert sg[k] == so[k], k
        ctx.close()
        orc.close()
