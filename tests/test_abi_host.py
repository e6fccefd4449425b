"""Host-side behaviour of the C ABI that needs no GPU: scene validation before any device call,
the scene footprint, and containment of C++ exceptions (a host allocation failure comes back as
SMS_E_OOM, the process lives on)."""
import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest

from smsamd import abi, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
E_INVALID, E_OOM = -1, -3


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(abi.LIB_PATH):
        pytest.skip("libsms_amd.so not built")
    return abi.load_library()


def _footprint(lib, sc):
    b = C.c_uint64(0)
    err = C.create_string_buffer(256)
    st = lib.sms_scene_footprint(C.byref(sc.desc()), C.byref(b), err, 256)
    return st, b.value, err.value.decode()


def test_footprint_counts_the_scene_arrays(lib):
    n = 64
    sc = scenes.pool(32, 32, n=n, n_waves=4)
    st, b, err = _footprint(lib, sc)
    assert st == 0, err
    nv, nf = (n + 1) ** 2, 2 * n * n
    mesh = nv * (16 + 16 + 8) + nf * 16 + nf * 48 + nf * 8          # vertices, faces, triangle records, area CDF/PMF
    nodes_max = 2 * nf                                               # a binary tree over nf leaves' items
    assert mesh < b < mesh + 32 * nodes_max + 4096


def test_footprint_of_a_normal_map(lib):
    sc = scenes.glint_plane(16, 16, nm_res=64, n_scratches=4)
    st, b, err = _footprint(lib, sc)
    assert st == 0, err
    slopes = 65 * 65 * 8
    lean = sum((64 >> l) ** 2 for l in range(7)) * 5 * 4
    assert slopes + lean < b < slopes + lean + 4096


def test_footprint_of_an_encoded_normal_map(lib):
    """A map whose frames decode interpolated RGB (use_slopes false) keeps its RGB texels in the
    padded layout beside the slopes; glinty shapes need slopes and refuse it."""
    sc = scenes.mirror_wall(nm_res=64, n_scratches=4)
    st, b0, err = _footprint(lib, sc)
    assert st == 0, err
    sc.encoded_normalmaps.add(0)
    st, b1, err = _footprint(lib, sc)
    assert st == 0, err
    assert b1 - b0 == 65 * 65 * 16
    g = scenes.glint_plane(16, 16, nm_res=64, n_scratches=4)
    g.encoded_normalmaps.add(0)
    st, _, err = _footprint(lib, g)
    assert st == -5 and "slope normal map" in err


def test_bad_scene_is_rejected_before_any_device_call(lib):
    faces = np.array([[0, 1, 7]], np.uint32)                       # vertex 7 does not exist
    sc = scenes.Scene()
    sc.add_mesh(np.zeros((3, 3), np.float32), faces)
    st, b, err = _footprint(lib, sc)
    assert st == E_INVALID and b == 0 and "out of range" in err
    h = C.c_void_p()
    cfg = abi.SmsConfig.default()
    assert lib.sms_ctx_create(C.byref(sc.desc()), C.byref(cfg), C.byref(h)) == E_INVALID
    assert not h.value


CHILD = r"""
import ctypes as C, resource, sys
import numpy as np
sys.path.insert(0, {pkg!r})
from smsamd import abi, scenes
lib = abi.load_library()
n = 1024
t = np.linspace(-1, 1, n + 1, dtype=np.float32)
X, Z = np.meshgrid(t, t)
P = np.stack([X, np.zeros_like(X), Z], -1).reshape(-1, 3)
idx = np.arange((n + 1) ** 2, dtype=np.uint32).reshape(n + 1, n + 1)
a, b, c, d = idx[:-1, :-1].ravel(), idx[:-1, 1:].ravel(), idx[1:, :-1].ravel(), idx[1:, 1:].ravel()
F = np.concatenate([np.stack([a, c, b], -1), np.stack([b, c, d], -1)])
sc = scenes.Scene()
sc.add_mesh(P, F, bsdf=abi.BSDF_DIELECTRIC, flags=abi.FLAG_CAUSTIC_CASTER_SS, eta=1.33)
desc = sc.desc()
cfg = abi.SmsConfig.default()
out = C.c_uint64(0)
err = C.create_string_buffer(128)
h = C.c_void_p()
vm = [int(l.split()[1]) for l in open("/proc/self/status") if l.startswith("VmSize")][0] * 1024
resource.setrlimit(resource.RLIMIT_AS, (vm + (48 << 20), resource.RLIM_INFINITY))
s1 = lib.sms_scene_footprint(C.byref(desc), C.byref(out), err, 128)
s2 = lib.sms_ctx_create(C.byref(desc), C.byref(cfg), C.byref(h))
resource.setrlimit(resource.RLIMIT_AS, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))
s3 = lib.sms_scene_footprint(C.byref(desc), C.byref(out), err, 128)
print("STATUS", s1, s2, s3, err.value.decode(), out.value, flush=True)
"""


def test_host_allocation_failure_is_a_status_code(lib):
    """A 2M-triangle mesh under an address-space limit 48 MiB above the process's size: the host
    BVH build runs out of memory inside the library; both entry points return SMS_E_OOM instead
    of letting std::bad_alloc cross the C ABI, and the process goes on (the same call without the
    limit succeeds)."""
    code = CHILD.format(pkg=os.path.join(ROOT, "specular-manifold-sampling_amd"))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    line = [l for l in r.stdout.splitlines() if l.startswith("STATUS")][0].split()
    assert int(line[1]) == E_OOM and int(line[2]) == E_OOM, line
    assert int(line[3]) == 0 and int(line[-1]) > 2 * 1024 * 1024 * 48, line
