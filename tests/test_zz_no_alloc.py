"""owt_f2p_noalloc: Fourier2Physical_dealiased (src/FFT.cpp:49-128) on
caller-owned device buffers, allocating nothing (its intermediate lives in the
library's scratch array). Checked through the C ABI against owt_f2p, the
allocating entry point with the same kernels: bit for bit, on two calls in a
row, the second with a smaller input on a smaller channel (the scratch keeps
the first call's contents). An input too large for the scratch is rejected.

Device buffers come from the HIP runtime that libowt itself is linked against,
so their addresses are valid for it."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RE180 = dict(nu=1 / 2800, dt=1e-3, dkx=0.5, dkz=1.0, alpha_mesh=1.5)


def hip_runtime(owt_mod):
    """hipMalloc and friends as libowt resolves them: dlsym on libowt's handle
    searches libowt's own dependencies, so these are the functions of the HIP
    runtime it is linked against (another runtime in the process, such as the
    one torch bundles, would not share its devices' address spaces)."""
    owt_mod.lib()
    H = ctypes.CDLL(owt_mod.LIB_PATH)
    H.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    H.hipFree.argtypes = [ctypes.c_void_p]
    H.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return H


class DevBuf:
    def __init__(self, H, nbytes):
        self.H, self.n, self.p = H, nbytes, ctypes.c_void_p()
        assert H.hipMalloc(ctypes.byref(self.p), nbytes) == 0

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.n
        assert self.H.hipMemcpy(self.p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1) == 0  # host to device

    def get(self, shape, dtype):
        out = np.empty(shape, dtype=dtype)
        assert self.H.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        self.H.hipFree(self.p)


def check(owt_mod, H, shape, nf, seed):
    p = dict(RE180, NX=shape[0], NY=shape[1], NZ=shape[2], precision=0, method=1)
    print(f"[gpu] f2p_noalloc {shape} nf={nf}", flush=True)
    with owt_mod.Channel(params=p) as ch:
        rng = np.random.default_rng(seed)
        F = rng.standard_normal((nf,) + ch.wave_shape) + 1j * rng.standard_normal((nf,) + ch.wave_shape)
        ref = ch.f2p(F).reshape((nf,) + ch.phys_shape)
        w = DevBuf(H, F.nbytes)
        ph = DevBuf(H, ref.nbytes)
        try:
            w.put(F)
            ch.f2p_noalloc(w.p.value, nf, ph.p.value)
            got = ph.get((nf,) + ch.phys_shape, np.float64)
        finally:
            w.free()
            ph.free()
        assert np.array_equal(got, ref)


def test_f2p_noalloc_equals_f2p_twice_then_smaller(owt_mod):
    H = hip_runtime(owt_mod)
    check(owt_mod, H, (64, 65, 64), 6, 1)   # 6 x 3.2 MB of scratch
    check(owt_mod, H, (16, 17, 16), 1, 2)   # smaller, over the first call's contents


def test_f2p_noalloc_rejects_inputs_beyond_the_scratch(owt_mod):
    H = hip_runtime(owt_mod)
    p = dict(RE180, NX=256, NY=33, NZ=256, precision=0, method=1)
    with owt_mod.Channel(params=p) as ch:
        # 3 fields x 384 x 128 x 33 x 16 B = 78 MB > 64 MiB: refused before any launch
        w = DevBuf(H, 16)
        ph = DevBuf(H, 16)
        try:
            with pytest.raises(owt_mod.OwtError) as e:
                ch.f2p_noalloc(w.p.value, 3, ph.p.value)
            assert e.value.code == 2 and "scratch" in str(e.value)
        finally:
            w.free()
            ph.free()
