"""The present pass (fullscreen.wgsl through the ClampToEdge linear quad
sampler, lib.rs:555-562 and 834-850) on the GPU against the oracle, bit for
bit: surfaces of the render size, enlarged, reduced, of other aspect ratios
and wider than one wavefront's 64 columns; the pitched device-memory form
leaves the row padding untouched; argument checks."""
import ctypes as C

import numpy as np
import pytest

import rustracer_amd as rt
from rustracer_amd import _ffi

from . import oracle_ffi as orc

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def _rendered(w=64, h=48):
    scene = rt.Scene.builtin("cornell_box", 3)
    cam = rt.Scene.builtin_camera("cornell_box", 3)
    r = rt.Renderer(rt.RenderConfig(width=w, height=h, spp=2, enable_denoise=True, accumulate_frames=4), scene, 0)
    r.update_camera(cam)
    r.render_frames(3)
    return r


@pytest.mark.parametrize("sw,sh", [(64, 48), (128, 96), (37, 29), (100, 31), (1, 1), (1000, 3), (5, 700)])
def test_present_bit_exact(sw, sh):
    print(f"present 64x48 -> {sw}x{sh}", flush=True)
    with _rendered() as r:
        out = r.read_output_rgba8()
        got = r.read_present(sw, sh)
        again = r.read_present(sw, sh)
    want = orc.present(out, sw, sh)
    assert np.array_equal(got, want), (sw, sh, np.argwhere(got != want)[:4])
    assert np.array_equal(again, got)
    if (sw, sh) == (64, 48):
        assert np.array_equal(got, out[..., [2, 1, 0, 3]])


def test_present_pitched_device_surface():
    """rt_present into device memory with padded rows: every pixel as the
    oracle has it, the padding bytes as they were."""
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    sw, sh, pitch = 150, 77, 4 * 150 + 36
    print(f"pitched present {sw}x{sh} pitch {pitch}", flush=True)
    with _rendered(40, 30) as r:
        out = r.read_output_rgba8()
        d = C.c_void_p()
        assert hip.hipMalloc(C.byref(d), pitch * sh) == 0
        try:
            fill = np.full(pitch * sh, 0xAB, np.uint8)
            assert hip.hipMemcpy(d, fill.ctypes.data, fill.nbytes, H2D) == 0
            r.present(d.value, sw, sh, pitch)
            r.synchronize()
            back = np.zeros(pitch * sh, np.uint8)
            assert hip.hipMemcpy(back.ctypes.data, d, back.nbytes, D2H) == 0
        finally:
            hip.hipFree(d)
    rows = back.reshape(sh, pitch)
    assert np.array_equal(rows[:, : 4 * sw].reshape(sh, sw, 4), orc.present(out, sw, sh))
    assert (rows[:, 4 * sw:] == 0xAB).all()


def test_present_argument_checks():
    with _rendered(16, 16) as r:
        lib = _ffi.lib()
        buf = np.zeros(16 * 16 * 4, np.uint8)
        for sw, sh, n in ((0, 4, buf.size), (4, 0, buf.size), (40000, 1, 10 ** 6), (16, 16, buf.size - 1)):
            assert lib.rt_read_present(r._h, sw, sh, buf.ctypes.data, n) != 0, (sw, sh, n)
        for pitch in (63, 60):  # not a multiple of 4; shorter than 16 pixels
            assert lib.rt_present(r._h, C.c_void_p(1 << 20), 16, 16, pitch) != 0, pitch
        assert lib.rt_present(r._h, None, 16, 16, 64) != 0
        # the renderer still works after the rejected calls
        assert r.read_present(8, 8).shape == (8, 8, 4)
