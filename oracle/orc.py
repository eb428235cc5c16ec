"""ctypes wrappers over liboracle.so (the C restatement in nbody_oracle.c).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  Arrays are numpy; pos4 and
vel4 are (n, 4) float32 in the reference's float4 layout (docs/FORMATS.md).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_vp = ctypes.c_void_p
_i = ctypes.c_int
_f = ctypes.c_float
_d = ctypes.c_double


def build():
    subprocess.run(["make", "-C", _HERE, "-B", "-s"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_spread21.restype = ctypes.c_uint64
        L.orc_spread21.argtypes = [ctypes.c_uint64]
        L.orc_quant_params.argtypes = [_f32p, _i, _f64p, ctypes.POINTER(ctypes.c_double)]
        L.orc_morton.argtypes = [_f32p, _i, _i, _u64p]
        L.orc_sort_pairs.argtypes = [_u64p, _i, _u64p, _u32p]
        L.orc_karras.argtypes = [_u64p, _i, _i, _i32p, _i32p]
        L.orc_radix_topdown.argtypes = [_u64p, _i, _i, _i32p, _i32p]
        L.orc_com.argtypes = [_f32p, _i, _i32p, _f32p, _f32p, _f32p, _f32p]
        L.orc_allpairs_f32.argtypes = [_f32p, _i, _f, _f, _f32p]
        L.orc_allpairs_f64.argtypes = [_f32p, _i, _d, _d, _vp, _i, _f64p]
        L.orc_traverse_thread.argtypes = [_f32p, _i, _i32p, _f32p, _f32p, _f, _f, _f, _i, _i,
                                          _vp, _vp, _vp]
        L.orc_traverse_group.argtypes = [_f32p, _i, _i32p, _f32p, _f32p, _f, _f, _f, _i,
                                         _vp, _vp, _vp]
        L.orc_traverse_group_f64sum.argtypes = [_f32p, _i, _i32p, _f32p, _f32p, _f, _f, _d, _i,
                                                _vp, _vp]
        L.orc_traverse_thread_f64.argtypes = [_f64p, _i, _i32p, _f64p, _f64p, _d, _d, _i, _i,
                                              _f64p, _vp]
        L.orc_energy_f64.argtypes = [_f32p, _f32p, _i, _d, _d, _f64p]
        L.orc_kick.argtypes = [_f32p, _f32p, _i, _f]
        L.orc_drift.argtypes = [_f32p, _f32p, _i, _f]
        L.orc_kdk_step_naive.argtypes = [_f32p, _f32p, _f32p, _i, _f, _f, _f]
        L.orc_num_threads.restype = _i
        _lib = L
    return _lib


def _c4(a):
    return np.ascontiguousarray(a, dtype=np.float32).reshape(-1, 4)


def _ptr(a):
    return None if a is None else a.ctypes.data


def spread21(v):
    return int(lib().orc_spread21(int(v)))


def quant_params(pos4):
    pos4 = _c4(pos4)
    lo = np.zeros(3)
    inv = ctypes.c_double()
    lib().orc_quant_params(pos4, len(pos4), lo, ctypes.byref(inv))
    return lo, inv.value


def morton(pos4, key_bits=63):
    pos4 = _c4(pos4)
    keys = np.zeros(len(pos4), np.uint64)
    lib().orc_morton(pos4, len(pos4), key_bits, keys)
    return keys


def sort_pairs(keys):
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    ks = np.zeros_like(keys)
    perm = np.zeros(len(keys), np.uint32)
    if lib().orc_sort_pairs(keys, len(keys), ks, perm) != 0:
        raise MemoryError("orc_sort_pairs")
    return ks, perm


def karras(keys_sorted, key_bits=63):
    keys_sorted = np.ascontiguousarray(keys_sorted, dtype=np.uint64)
    n = len(keys_sorted)
    ch = np.zeros(max(2 * (n - 1), 1), np.int32)
    par = np.zeros(max(2 * n - 1, 1), np.int32)
    lib().orc_karras(keys_sorted, n, key_bits, ch, par)
    return ch[: 2 * (n - 1)].reshape(-1, 2), par[: 2 * n - 1]


def radix_topdown(keys_sorted, key_bits=63):
    keys_sorted = np.ascontiguousarray(keys_sorted, dtype=np.uint64)
    n = len(keys_sorted)
    ch = np.zeros(max(2 * (n - 1), 1), np.int32)
    par = np.zeros(max(2 * n - 1, 1), np.int32)
    if lib().orc_radix_topdown(keys_sorted, n, key_bits, ch, par) != 0:
        raise MemoryError("orc_radix_topdown")
    return ch[: 2 * (n - 1)].reshape(-1, 2), par[: 2 * n - 1]


def com(pos_sorted4, children):
    pos_sorted4 = _c4(pos_sorted4)
    n = len(pos_sorted4)
    m = max(n - 1, 1)
    ch = np.ascontiguousarray(children, dtype=np.int32).reshape(-1)
    if ch.size == 0:
        ch = np.zeros(2, np.int32)
    com4 = np.zeros((m, 4), np.float32)
    lo = np.zeros((m, 4), np.float32)
    hi = np.zeros((m, 4), np.float32)
    size2 = np.zeros(m, np.float32)
    if lib().orc_com(pos_sorted4, n, ch, com4, lo, hi, size2) != 0:
        raise MemoryError("orc_com")
    k = n - 1
    return com4[:k], lo[:k], hi[:k], size2[:k]


def allpairs_f32(pos4, eps, G=1.0):
    pos4 = _c4(pos4)
    acc = np.zeros_like(pos4)
    lib().orc_allpairs_f32(pos4, len(pos4), eps, G, acc)
    return acc


def allpairs_f64(pos4, eps, G=1.0, targets=None):
    pos4 = _c4(pos4)
    if targets is None:
        nt = len(pos4)
        tp = None
    else:
        targets = np.ascontiguousarray(targets, dtype=np.int32)
        nt = len(targets)
        tp = targets
    acc = np.zeros((nt, 3))
    lib().orc_allpairs_f64(pos4, len(pos4), eps, G, _ptr(tp), nt, acc)
    return acc


class Tree:
    """Everything the CPU LBVH pipeline produces for one position set."""

    def __init__(self, pos4, key_bits=63):
        pos4 = _c4(pos4)
        self.n = n = len(pos4)
        self.key_bits = key_bits
        self.keys = morton(pos4, key_bits)
        self.keys_sorted, self.perm = sort_pairs(self.keys)
        self.pos_sorted = np.ascontiguousarray(pos4[self.perm])
        self.children, self.parent = karras(self.keys_sorted, key_bits)
        if n >= 2:
            self.com4, self.box_lo, self.box_hi, self.size2 = com(self.pos_sorted, self.children)
        else:
            self.com4 = np.zeros((0, 4), np.float32)
            self.size2 = np.zeros(0, np.float32)

    def _ch(self):
        c = np.ascontiguousarray(self.children, dtype=np.int32).reshape(-1)
        return c if c.size else np.zeros(2, np.int32)

    def _nodes(self):
        c4 = self.com4 if len(self.com4) else np.zeros((1, 4), np.float32)
        s2 = self.size2 if len(self.size2) else np.zeros(1, np.float32)
        return np.ascontiguousarray(c4), np.ascontiguousarray(s2)

    def traverse_thread(self, theta, eps, G=1.0, t0=0, t1=None):
        """Returns (acc4 in original order for [t0,t1) sorted targets scattered,
        sorted-order acc4, interactions, opens, max depth)."""
        t1 = self.n if t1 is None else t1
        acc_s = np.zeros((self.n, 4), np.float32)
        inter = np.zeros(self.n, np.int32)
        opens = np.zeros(self.n, np.int32)
        c4, s2 = self._nodes()
        depth = lib().orc_traverse_thread(self.pos_sorted, self.n, self._ch(), c4, s2, theta, eps,
                                          G, t0, t1, acc_s.ctypes.data, inter.ctypes.data,
                                          opens.ctypes.data)
        if depth < 0:
            raise RuntimeError("oracle traversal stack overflow")
        acc = np.zeros_like(acc_s)
        acc[self.perm] = acc_s
        return acc, acc_s, inter, opens, depth

    def traverse_group(self, theta, eps, width, G=1.0):
        ng = (self.n + width - 1) // width
        acc_s = np.zeros((self.n, 4), np.float32)
        gi = np.zeros(ng, np.int32)
        go = np.zeros(ng, np.int32)
        c4, s2 = self._nodes()
        depth = lib().orc_traverse_group(self.pos_sorted, self.n, self._ch(), c4, s2, theta, eps,
                                         G, width, acc_s.ctypes.data, gi.ctypes.data,
                                         go.ctypes.data)
        if depth < 0:
            raise RuntimeError("oracle traversal stack overflow")
        acc = np.zeros_like(acc_s)
        acc[self.perm] = acc_s
        return acc, acc_s, gi, go, depth

    def traverse_group_f64sum(self, theta, eps, width, G=1.0):
        """The group walk's interaction lists (its FP32 open/accept decisions)
        summed in FP64: (acc (n, 3) in original order, per-group counts)."""
        acc_s = np.zeros((self.n, 3))
        gi = np.zeros((self.n + width - 1) // width, np.int32)
        c4, s2 = self._nodes()
        depth = lib().orc_traverse_group_f64sum(self.pos_sorted, self.n, self._ch(), c4, s2,
                                                theta, eps, G, width, acc_s.ctypes.data,
                                                gi.ctypes.data)
        if depth < 0:
            raise RuntimeError("oracle traversal stack overflow")
        acc = np.zeros_like(acc_s)
        acc[self.perm] = acc_s
        return acc, gi


def traverse_thread_f64(pos_sorted4_f64, children, com4d, size2d, theta, eps, t0, t1):
    p = np.ascontiguousarray(pos_sorted4_f64, dtype=np.float64)
    n = len(p)
    ch = np.ascontiguousarray(children, dtype=np.int32).reshape(-1)
    if ch.size == 0:
        ch = np.zeros(2, np.int32)
    c4 = np.ascontiguousarray(com4d, dtype=np.float64)
    s2 = np.ascontiguousarray(size2d, dtype=np.float64)
    if c4.size == 0:
        c4 = np.zeros((1, 4))
        s2 = np.zeros(1)
    acc = np.zeros((t1 - t0, 3))
    inter = np.zeros(t1 - t0, np.int32)
    depth = lib().orc_traverse_thread_f64(p, n, ch, c4, s2, theta, eps, t0, t1, acc,
                                          inter.ctypes.data)
    if depth < 0:
        raise RuntimeError("oracle traversal stack overflow")
    return acc, inter, depth


def energy_f64(pos4, vel4, eps, G=1.0):
    pos4 = _c4(pos4)
    vel4 = _c4(vel4)
    out = np.zeros(3)
    lib().orc_energy_f64(pos4, vel4, len(pos4), eps, G, out)
    return out


def kdk_step_naive(pos4, vel4, acc4, dt, eps, G=1.0):
    """In-place naive KDK step on float32 (n,4) arrays."""
    lib().orc_kdk_step_naive(pos4, vel4, acc4, len(pos4), dt, eps, G)


def num_threads():
    return int(lib().orc_num_threads())
