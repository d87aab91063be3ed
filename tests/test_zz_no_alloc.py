"""The allocation-free photon-grid build (rt_photon_grid_build_noalloc) against
the allocating one (rt_build_photon_grid_from) on the same photon records:
grid_meta, the stored photons in sorted order, their count and the dense gather
records, bit for bit, on two calls in a row (the second with fewer slots, over
the workspace the first one left behind). The test allocates its own device
buffers through the HIP runtime; the entry point under test allocates none."""
import ctypes as C

import numpy as np
import pytest

import rustracer_amd as rt
from rustracer_amd import _ffi

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


class _Hip:
    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipSetDevice.argtypes = [C.c_int]
        self.held = []

    def ok(self, rc, what):
        assert rc == 0, f"{what}: hip error {rc}"

    def malloc(self, nbytes):
        p = C.c_void_p()
        self.ok(self.lib.hipMalloc(C.byref(p), max(int(nbytes), 4)), "hipMalloc")
        self.held.append(p)
        return p

    def upload(self, arr):
        a = np.ascontiguousarray(arr)
        p = self.malloc(a.nbytes)
        if a.nbytes:
            self.ok(self.lib.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")
        return p

    def download(self, p, dtype, count):
        out = np.zeros(count, dtype)
        if out.nbytes:
            self.ok(self.lib.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    def sync(self):
        self.ok(self.lib.hipDeviceSynchronize(), "hipDeviceSynchronize")

    def free_all(self):
        for p in self.held:
            self.lib.hipFree(p)
        self.held = []


def _noalloc_build(hip, d_photons, n_slots, gp, meta_cells, outs):
    d_meta, d_sorted, d_rec, d_n = outs
    rc = _ffi.lib().rt_photon_grid_build_noalloc(d_photons, n_slots, C.byref(gp), d_meta, meta_cells, d_sorted,
                                                 d_rec, d_n, None)
    assert rc == 0, _ffi.lib().rt_last_error()
    hip.sync()
    n = int(hip.download(d_n, np.uint32, 1)[0])
    meta = hip.download(d_meta, np.uint32, 2 * meta_cells).reshape(-1, 2)
    srt = hip.download(d_sorted, _ffi.PHOTON, n)
    rec = hip.download(d_rec, np.float32, 12 * n).reshape(-1, 12)
    return n, meta, srt, rec


def test_noalloc_grid_build_equals_allocating_build():
    scene = rt.Scene.builtin("photon_plain", 1)
    cfg = rt.RenderConfig(width=16, height=16, enable_photon_map=True, photon_count=150_000, max_depth=6)
    hip = _Hip()
    hip.ok(hip.lib.hipSetDevice(0), "hipSetDevice")
    print("no-alloc grid build: photon_plain, 150000 photons", flush=True)
    try:
        with rt.Renderer(cfg, scene, 0) as r:
            r.rebuild_photons()
            photons = r.read_photons()
            gp = r.grid_params()
            meta_cells = len(r.read_grid_meta())
            assert len(photons) <= 1 << 20  # RT_GRID_NOALLOC_MAX_SLOTS
            outs = (hip.malloc(8 * meta_cells), hip.malloc(photons.nbytes), hip.malloc(48 * len(photons)),
                    hip.malloc(4))
            for n_slots in (len(photons), 200_003):
                print(f"  n_slots={n_slots}", flush=True)
                part = photons[:n_slots]
                r.build_photon_grid_from(part)
                ref_meta, ref_sorted = r.read_grid_meta(), r.read_sorted_photons()
                d_ph = hip.upload(part)
                n, meta, srt, rec = _noalloc_build(hip, d_ph, n_slots, gp, meta_cells, outs)
                assert n == len(ref_sorted) and n > 1000, (n, len(ref_sorted))
                assert np.array_equal(meta, ref_meta)
                assert srt.tobytes() == ref_sorted.tobytes()
                # gather records: (position, power.r), (power.gb, incident.xy), (incident.z, 0, 0, 0)
                f = srt.view(np.float32).reshape(n, 16)
                want = np.zeros((n, 12), np.float32)
                want[:, 0:3], want[:, 3] = f[:, 0:3], f[:, 4]
                want[:, 4:6], want[:, 6:8] = f[:, 5:7], f[:, 8:10]
                want[:, 8] = f[:, 10]
                assert rec.tobytes() == want.tobytes()
            # no gather records requested: the rest is unchanged
            d_ph = hip.upload(photons[:5000])
            rc = _ffi.lib().rt_photon_grid_build_noalloc(d_ph, 5000, C.byref(gp), outs[0], meta_cells, outs[1],
                                                         None, outs[3], None)
            assert rc == 0, _ffi.lib().rt_last_error()
            hip.sync()
            r.build_photon_grid_from(photons[:5000])
            n = int(hip.download(outs[3], np.uint32, 1)[0])
            assert n == len(r.read_sorted_photons())
            assert hip.download(outs[1], _ffi.PHOTON, n).tobytes() == r.read_sorted_photons().tobytes()
    finally:
        hip.free_all()


def test_noalloc_grid_build_rejects_sizes_beyond_its_workspace():
    gp = _ffi.GridParams()
    gp.num_cells = 8
    dummy = C.c_void_p(16)  # never dereferenced: the size check comes first
    rc = _ffi.lib().rt_photon_grid_build_noalloc(dummy, (1 << 20) + 1, C.byref(gp), dummy, 8, dummy, None, dummy,
                                                 None)
    assert rc == -5  # RT_ERR_CAPACITY
    assert b"exceeds the allocation-free workspace" in _ffi.lib().rt_last_error()
