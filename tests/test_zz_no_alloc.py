"""The allocation-free entry point nft_laplc_noalloc (no reference counterpart), through the
C ABI: laplc (lippsLFroude_implgwopks3D_unstr3.f:1288-1353) on device arrays with its
intermediate in the library's fixed scratch array, against nft_laplc bit for bit.

The device arrays of the test come from the HIP runtime that libnft_hip.so itself links
(hipMalloc / hipMemcpy / hipFree through ctypes), so one runtime serves the process."""
import ctypes

import numpy as np
import pytest

from common import BAND_KW, mesh_for, periodic_cfg, seeded_fields, small_cfg, tiny_cfg
from src_amd import mesh as M
from src_amd import nft

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2   # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


class _Hip:
    """hipMalloc/hipMemcpy/hipFree of the runtime libnft_hip.so is linked against."""

    def __init__(self):
        nft.lib()   # loads libnft_hip.so and with it its HIP runtime
        self.rt = ctypes.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), nbytes) == 0
        return p

    def put(self, dst, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert self.rt.hipMemcpy(dst, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, H2D) == 0

    def get(self, src, n):
        out = np.empty(n)
        assert self.rt.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), src, out.nbytes, D2H) == 0
        return out

    def free(self, p):
        assert self.rt.hipFree(p) == 0


def test_laplc_noalloc_equals_laplc():
    """Calls in a row on the 40,000-node mesh, then on the 15,840-node periodic box (its
    set sums in the context), then on the 120-node mesh (each a smaller input while the
    shared scratch still holds the larger intermediate), both impfl: every result equals
    nft_laplc's."""
    hip = _Hip()
    for key, cfg in (("small", small_cfg(**BAND_KW)), ("periodic2", periodic_cfg()), ("tiny", tiny_cfg())):
        print(f"case {key}", flush=True)
        m = mesh_for(key)
        s = nft.Solver(m, nft.make_params(cfg))
        n = m.npoin
        p = seeded_fields(m)["p"]
        dp, dout = hip.alloc(8 * n), hip.alloc(8 * n)
        try:
            hip.put(dp, p)
            for impfl in (1, 0):
                want = s.laplc(p, impfl)
                hip.put(dout, np.full(n, np.nan))
                s.laplc_noalloc(dp.value, dout.value, impfl)
                got = hip.get(dout, n)
                assert np.array_equal(got, want), (key, impfl)
        finally:
            hip.free(dp)
            hip.free(dout)
            s.close()


def test_laplc_noalloc_rejects_meshes_beyond_its_scratch():
    """1,058,200 nodes: more than the 1,048,576 the fixed array holds; rejected before any
    kernel runs. A null array is rejected too."""
    hip = _Hip()
    m = M.generate(M.tiny_spec(nx=130, ny=110, nlev=74, seed=4))
    assert m.npoin > 1 << 20
    s = nft.Solver(m, nft.make_params(tiny_cfg()))
    d = hip.alloc(64)
    try:
        with pytest.raises(nft.NftError, match="fixed scratch"):
            s.laplc_noalloc(d.value, d.value, 1)
        with pytest.raises(nft.NftError, match="bad argument"):
            s.laplc_noalloc(0, d.value, 1)
    finally:
        hip.free(d)
        s.close()
