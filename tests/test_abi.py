"""The C-ABI library loads without a GPU, exports every entry point include/sms.h declares, and
the numpy/ctypes record layouts equal the C struct layouts (checked with a gcc-compiled probe)."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

from smsamd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sms.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:sms_status|void|const char \*|const sms_scene_desc \*)\s*(sms_\w+)\s*\(", text, re.M)))


def test_header_and_python_export_lists_agree():
    assert declared_functions() == sorted(abi.EXPORTS)


@pytest.fixture(scope="module")
def library():
    if not os.path.exists(abi.LIB_PATH):
        pytest.skip("libsms_amd.so not built")
    return C.CDLL(abi.LIB_PATH)


def test_library_exports_every_declared_symbol(library):
    for name in declared_functions():
        assert hasattr(library, name), name


def test_config_default_matches_reference_defaults(library):
    lib = abi.load_library()
    cfg = abi.SmsConfig()
    lib.sms_config_default(C.byref(cfg))
    ref = abi.SmsConfig.default()
    for f, _ in abi.SmsConfig._fields_:
        assert getattr(cfg, f) == pytest.approx(getattr(ref, f)), f
    # SMSConfig() / path_sms_ss.cpp:49-57 defaults
    assert cfg.max_iterations == 20 and cfg.max_trials == -1 and cfg.biased == 0
    assert cfg.solver_threshold == pytest.approx(1e-5) and cfg.uniqueness_threshold == pytest.approx(1e-4)


def test_null_arguments_are_errors_not_crashes(library):
    lib = abi.load_library()
    h = C.c_void_p()
    assert lib.sms_ctx_create(None, None, C.byref(h)) == -1
    assert lib.sms_ctx_destroy(None) == -1
    assert lib.sms_ss_walk_batch(None, None, 0, None) == -1


def test_struct_layouts_match_c():
    names = list(abi.C_SIZES)
    fields = {n: list(abi.DTYPES[n].names) if n in abi.DTYPES else [] for n in names}
    src = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    for n in names:
        src.append(f'printf("{n} %zu\\n", sizeof({n}));')
        for f in fields[n]:
            src.append(f'printf("{n}.{f} %zu\\n", offsetof({n}, {f}));')
    src.append("return 0;}")
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "probe.c")
        exe = os.path.join(d, "probe")
        open(c, "w").write("\n".join(src))
        subprocess.run(["gcc", "-std=c99", "-o", exe, c], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    got = dict(line.split() for line in out.strip().splitlines())
    for n in names:
        assert int(got[n]) == abi.C_SIZES[n], n
        if n in abi.DTYPES:
            dt = abi.DTYPES[n]
            assert dt.itemsize == abi.C_SIZES[n], n
            for f in dt.names:
                assert int(got[f"{n}.{f}"]) == dt.fields[f][1], (n, f)


CTYPES_STRUCTS = {"sms_config": abi.SmsConfig, "sms_stats": abi.SmsStats, "sms_schedule": abi.SmsSchedule,
                  "sms_shape_desc": abi.SmsShapeDesc, "sms_emitter_desc": abi.SmsEmitterDesc,
                  "sms_normalmap_desc": abi.SmsNormalmapDesc, "sms_camera_desc": abi.SmsCameraDesc,
                  "sms_scene_desc": abi.SmsSceneDesc, "sms_path_config": abi.SmsPathConfig}


def test_ctypes_structs_match_c():
    """Every ctypes mirror has the C struct's size and field offsets."""
    src = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    for n, cls in CTYPES_STRUCTS.items():
        src.append(f'printf("{n} %zu\\n", sizeof({n}));')
        for f, _ in cls._fields_:
            src.append(f'printf("{n}.{f} %zu\\n", offsetof({n}, {f}));')
    src.append("return 0;}")
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "probe.c")
        exe = os.path.join(d, "probe")
        open(c, "w").write("\n".join(src))
        subprocess.run(["gcc", "-std=c99", "-o", exe, c], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    got = dict(line.split() for line in out.strip().splitlines())
    for n, cls in CTYPES_STRUCTS.items():
        assert int(got[n]) == C.sizeof(cls), n
        for f, _ in cls._fields_:
            assert int(got[f"{n}.{f}"]) == getattr(cls, f).offset, (n, f)


def test_library_defines_its_kernels():
    """Every kernel and device-side symbol the host code launches is defined in the library
    (a kernel declared in one namespace and defined in another links as an undefined symbol)."""
    if not os.path.exists(abi.LIB_PATH):
        pytest.skip("library not built")
    out = subprocess.run(["nm", "-D", "--undefined-only", abi.LIB_PATH], check=True, capture_output=True,
                         text=True).stdout
    missing = [line.split()[-1] for line in out.splitlines() if "smsdev" in line]
    assert not missing, missing
