"""nb_compute_forces_noalloc: the allocation-free force entry point.

Called through the C ABI on device arrays the test allocates itself (through
the HIP runtime libnbody.so links), it must give bitwise the accelerations of
nb_compute_forces with the same parameters -- on two calls in a row, the
second with fewer particles, so the static scratch is reused with the
previous call's contents still in it.
"""
import ctypes

import numpy as np
import pytest

from nbodyssey_amd import _lib, engine, ic
from nbodyssey_amd._lib import NbParams

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip():
    """The HIP runtime libnbody.so was linked against (already loaded)."""
    _lib.load()
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


class DeviceArray:
    def __init__(self, h, nbytes):
        self.h = h
        self.p = ctypes.c_void_p()
        assert h.hipMalloc(ctypes.byref(self.p), nbytes) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert self.h.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, shape, dtype=np.float32):
        out = np.zeros(shape, dtype)
        assert self.h.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        if self.p.value:
            self.h.hipFree(self.p)
            self.p = ctypes.c_void_p()


def params(**kw):
    p = NbParams()
    _lib.load().nb_default_params(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def noalloc(L, d_pos, n, p, d_acc):
    fn = L.nb_compute_forces_noalloc
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(NbParams), ctypes.c_void_p,
                   ctypes.c_void_p]
    return fn(d_pos.p, n, ctypes.byref(p), d_acc.p, None)


CASES = {
    # name: (nb_params fields, Simulation keywords)
    "bh-wave64-30": (dict(force=1, traverse=1, group_width=64, key_bits=30, theta=0.5,
                          epsilon=0.02),
                     dict(force="bh", traverse="warp", group_width=64, key_bits=30, theta=0.5,
                          eps=0.02)),
    "bh-thread-63": (dict(force=1, traverse=0, group_width=32, key_bits=63, theta=0.5,
                          epsilon=0.05),
                     dict(force="bh", traverse="thread", group_width=32, key_bits=63, theta=0.5,
                          eps=0.05)),
    "bh-kahan-32": (dict(force=1, traverse=1, group_width=32, key_bits=30, theta=0.7,
                         epsilon=0.05, precision=1),
                    dict(force="bh", traverse="warp", group_width=32, key_bits=30, theta=0.7,
                         eps=0.05, compensated=True)),
    "naive": (dict(force=0, epsilon=0.05), dict(force="naive", eps=0.05)),
}


@pytest.mark.parametrize("case", list(CASES))
def test_noalloc_matches_compute_forces(case):
    L = _lib.load()
    h = hip()
    cfields, skw = CASES[case]
    p = params(**cfields)
    sizes = [65536, 20011]  # the maximum, then a smaller one on the same scratch
    d_pos = DeviceArray(h, 16 * sizes[0])
    d_acc = DeviceArray(h, 16 * sizes[0])
    try:
        for k, n in enumerate(sizes):
            pos, _ = ic.two_disk_encounter(n, seed=70 + k, eps=0.05)
            print(f"noalloc {case}: call {k} n={n}", flush=True)
            d_pos.put(pos)
            assert noalloc(L, d_pos, n, p, d_acc) == 0, L.nb_last_error().decode()
            assert h.hipDeviceSynchronize() == 0
            got = d_acc.get((n, 4))
            with engine.Simulation(pos, **skw) as sim:
                ref = sim.compute_forces()
            assert np.all(np.isfinite(got))
            assert np.array_equal(got, ref), (case, n, np.abs(got - ref).max())
    finally:
        d_pos.free()
        d_acc.free()


def test_noalloc_rejects_bad_arguments():
    L = _lib.load()
    h = hip()
    p = params(force=1, traverse=1, group_width=64, key_bits=30)
    d_pos = DeviceArray(h, 16 * 1024)
    d_acc = DeviceArray(h, 16 * 1024)
    try:
        over = _lib.NB_ERR_INVALID
        # larger than the static scratch
        assert noalloc(L, d_pos, 65537, p, d_acc) == over
        assert "65536" in L.nb_last_error().decode()
        # host memory instead of device memory
        host = np.zeros((1024, 4), np.float32)

        class Host:
            pass
        hp = Host()
        hp.p = ctypes.c_void_p(host.ctypes.data)
        assert noalloc(L, hp, 1024, p, d_acc) == over
        assert "device memory" in L.nb_last_error().decode()
        bad = params(force=1, traverse=1, group_width=48)
        assert noalloc(L, d_pos, 1024, bad, d_acc) == over
        # n = 1 under Barnes-Hut: zero force, as the reference
        d_acc.put(np.ones((1, 4), np.float32))
        d_pos.put(np.array([[1, 2, 3, 1]], np.float32))
        assert noalloc(L, d_pos, 1, p, d_acc) == 0
        assert h.hipDeviceSynchronize() == 0
        assert np.all(d_acc.get((1, 4)) == 0)
    finally:
        d_pos.free()
        d_acc.free()
