"""pf_timestep_noalloc (no allocation, scratch in a file-scope device array) gives
the same result as pf_timestep on the same particle velocities, bit for bit, on two
calls in a row, the second with a smaller input.

The test's own device buffers come from the HIP runtime the library links
(libamdhip64, loaded through ctypes), so the process holds one HIP runtime."""
import ctypes as C
import os

import numpy as np
import pytest

import pfflip_amd as G

pytestmark = pytest.mark.gpu


def _hip():
    """the libamdhip64 the library itself loaded (same runtime, same device context)"""
    G.lib()
    with open("/proc/self/maps") as f:
        paths = [line.split()[-1] for line in f if "libamdhip64.so" in line]
    h = C.CDLL(paths[0] if paths else "libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipDeviceSynchronize.argtypes = []
    return h


def _noalloc_fn():
    f = G.lib().pf_timestep_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p] * 3 + [C.c_int64, C.c_double, C.c_double, C.c_int32, C.c_double, C.c_double,
                                     C.c_void_p] + [C.c_void_p] * 4
    return f


def _noalloc(dev, n, dx, dt, cfl, min_dt):
    e, vm, cd, lim = C.c_double(), C.c_double(), C.c_double(), C.c_int32()
    st = _noalloc_fn()(dev[0], dev[1], dev[2], n, dx, dt, 1, cfl, min_dt, None, C.byref(e), C.byref(vm),
                       C.byref(cd), C.byref(lim))
    assert st == 0
    return e.value, vm.value, cd.value, lim.value


def test_timestep_noalloc_matches_pf_timestep():
    hip = _hip()
    n0 = 48
    dx = 1.0 / n0
    rng = np.random.default_rng(3)
    for n in (5_000_003, 777_777):  # the second call smaller: the scratch keeps the first call's partials
        x = rng.uniform(0.1, 0.9, n).astype(np.float32)
        y = rng.uniform(0.1, 0.9, n).astype(np.float32)
        z = rng.uniform(0.1, 0.9, n).astype(np.float32)
        u, v, w = (rng.normal(scale=2.0, size=n).astype(np.float32) for _ in range(3))
        u[rng.integers(0, n, 50)] = np.inf  # non-finite velocities are skipped
        w[rng.integers(0, n, 50)] = np.nan
        if n < 1_000_000:
            u *= 0.25; v *= 0.25; w *= 0.25  # a smaller maximum than the first call's
        ctx = G.Context(n0, n0, n0, dx)
        try:
            ctx.upload(x, y, z, u, v, w)
            want = ctx.timestep(1e-2, True, 0.9, 1e-5)
        finally:
            ctx.close()
        dev = [C.c_void_p() for _ in range(3)]
        try:
            for d, a in zip(dev, (u, v, w)):
                assert hip.hipMalloc(C.byref(d), a.nbytes) == 0
                assert hip.hipMemcpy(d, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0  # host to device
            assert hip.hipDeviceSynchronize() == 0
            print(f"[noalloc] n {n}: calling pf_timestep_noalloc", flush=True)
            got = _noalloc(dev, n, dx, 1e-2, 0.9, 1e-5)
        finally:
            for d in dev:
                if d.value:
                    hip.hipFree(d)
        fin = np.isfinite(u) & np.isfinite(v) & np.isfinite(w)
        sp = np.sqrt(u[fin] * u[fin] + v[fin] * v[fin] + w[fin] * w[fin]).max()
        print(f"[noalloc] n {n}: {got} vs {want}", flush=True)
        assert got == want
        assert abs(got[1] - float(sp)) <= 1e-6 * float(sp)  # the FP32 speed the host computes
        assert got[3] == 1


def test_timestep_noalloc_rejects_oversize():
    f = _noalloc_fn()
    assert f(None, None, None, (1 << 28) + 1, 0.1, 0.1, 1, 1.0, 0.0, None, None, None, None, None) == 1
    assert f(None, None, None, -1, 0.1, 0.1, 1, 1.0, 0.0, None, None, None, None, None) == 1
