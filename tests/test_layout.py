"""Upload contract: record sizes and offsets (rustracer-renderer/tests/gpu_layout.rs:12-41,
rustracer-core/src/material.rs:186-221) and the exported C ABI. CPU only."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from rustracer_amd import _ffi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_record_sizes_match_wgsl_spans():
    # gpu_layout.rs:12-41: Light 64, Material 64, Triangle 128, BVHNode 48, Camera 96, Photon 64
    assert _ffi.LIGHT.itemsize == 64
    assert _ffi.MATERIAL.itemsize == 64
    assert _ffi.TRIANGLE.itemsize == 128
    assert _ffi.BVH_NODE.itemsize == 48
    assert C.sizeof(_ffi.CameraUniform) == 96
    assert _ffi.PHOTON.itemsize == 64
    assert C.sizeof(_ffi.ImageParams) == 80
    assert C.sizeof(_ffi.GridParams) == 48


def test_material_offsets():
    # material.rs:193-216 writes 3..11 into the fields and reads them back at these offsets
    m = np.zeros(1, _ffi.MATERIAL)
    m["albedo"] = 1.0
    m["emissive"] = 2.0
    m["roughness"], m["metallic"], m["ior"] = 3.0, 4.0, 5.0
    m["kind"], m["flags"], m["tex_albedo"], m["tex_mr"], m["tex_emissive"], m["tex_normal"] = 6, 7, 8, 9, 10, 11
    b = m.tobytes()
    f = lambda o: np.frombuffer(b[o:o + 4], np.float32)[0]
    u = lambda o: np.frombuffer(b[o:o + 4], np.uint32)[0]
    assert f(28) == 3.0 and f(32) == 4.0
    assert u(40) == 6 and u(44) == 7 and u(48) == 8 and u(56) == 10 and u(60) == 11
    assert _ffi.TEX_NONE == 2**32 - 1


def test_light_layout_tail_at_48():
    # scene.rs:128-140: data 0, color 16, intensity 28, kind 32, pad to 48, _pad vec3 at 48
    assert _ffi.LIGHT.fields["kind"][1] == 32
    assert _ffi.LIGHT.fields["_pad_b"][1] == 48
    assert _ffi.TRIANGLE.fields["material_id"][1] == 120


@pytest.mark.parametrize("lang", ["c", "c++"])
def test_header_static_asserts_compile(tmp_path, lang):
    src = tmp_path / ("t.c" if lang == "c" else "t.cpp")
    src.write_text('#include "rt_types.h"\n#include "rustracer_hip.h"\nint main(void){return 0;}\n')
    cc = "gcc" if lang == "c" else "g++"
    std = "-std=c11" if lang == "c" else "-std=c++17"
    subprocess.run([cc, std, "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(tmp_path / "a.out")], check=True)


def _declared_functions():
    names = set()
    for hdr in ("rustracer_hip.h",):
        text = open(os.path.join(ROOT, "include", hdr)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(rt_\w+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_ffi.LIB_PATH)
    declared = _declared_functions()
    assert len(declared) > 50
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding covers exactly the header
    assert declared == set(_ffi.SIGNATURES), declared ^ set(_ffi.SIGNATURES)


def test_library_loads_without_device_and_errors_cleanly(rt):
    cfg = rt.RenderConfig(width=4, height=4)
    with pytest.raises(rt.RtError) as e:
        rt.Renderer(cfg, rt.Scene.builtin("smoke"), 0)
    assert e.value.code in (_ffi.RT_ERR_NO_DEVICE, _ffi.RT_ERR_HIP) or os.environ.get("HIP_VISIBLE_DEVICES")
