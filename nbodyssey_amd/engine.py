"""Python host interface over libnbody.so.

``Simulation`` wraps one engine handle (device state + stream).  The module
functions mirror the reference's CPU-script interfaces on the GPU path, so a
caller of those scripts can switch over:

* ``accelerations(pos, mass, eps2, G)``      -- reference_nbody.py:45-61
* ``total_energy(pos, vel, mass, eps2, G)``  -- reference_nbody.py:64-81
* ``morton_keys(pos)``                       -- lbvh_check.py:45-54
* ``build_tree(keys)``                       -- lbvh_check.py:57-100
* ``barnes_hut_accelerations(...)``          -- forces_bh.cu:463-527
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import NbParams, as_f4, check, ptr

FORCES = {"naive": _lib.FORCE_NAIVE, "bh": _lib.FORCE_BH}
TRAVERSALS = {"thread": _lib.TRAVERSE_THREAD, "warp": _lib.TRAVERSE_WARP}


def device_count():
    return int(_lib.load().nb_device_count())


class Simulation:
    """One particle system on the GPU (ParticleSystem + the force modules +
    the integrator of the reference, simulation.hpp / forces.cuh /
    integrator.cuh)."""

    def __init__(self, pos4, vel4=None, dt=0.01, eps=0.05, G=1.0, theta=0.5, force="naive",
                 traverse="thread", key_bits=63, group_width=32, compensated=False,
                 far16=False):
        L = _lib.load()
        pos4 = as_f4(pos4)
        self.n = len(pos4)
        h = ctypes.c_void_p()
        check(L.nb_create(self.n, ctypes.byref(h)))
        self._h = h
        self._L = L
        self.set_params(dt=dt, eps=eps, G=G, theta=theta, force=force, traverse=traverse,
                        key_bits=key_bits, group_width=group_width, compensated=compensated,
                        far16=far16)
        self.upload(pos4, vel4)

    # -- lifetime
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.nb_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- parameters
    def set_params(self, dt=0.01, eps=0.05, G=1.0, theta=0.5, force="naive", traverse="thread",
                   key_bits=63, group_width=32, compensated=False, far16=False):
        p = NbParams()
        p.dt, p.epsilon, p.G, p.theta = dt, eps, G, theta
        p.force = FORCES[force] if isinstance(force, str) else int(force)
        p.traverse = TRAVERSALS[traverse] if isinstance(traverse, str) else int(traverse)
        p.key_bits = key_bits
        p.group_width = group_width
        p.precision = (_lib.PREC_FP16_FAR if far16 else
                       _lib.PREC_COMPENSATED if compensated else _lib.PREC_FP32)
        check(self._L.nb_set_params(self._h, ctypes.byref(p)))
        self.params = dict(dt=dt, eps=eps, G=G, theta=theta, force=force, traverse=traverse,
                           key_bits=key_bits, group_width=group_width, compensated=compensated,
                           far16=far16)

    def update(self, **kw):
        p = dict(self.params)
        p.update(kw)
        self.set_params(**p)

    # -- state
    def upload(self, pos4, vel4=None):
        pos4 = as_f4(pos4, self.n)
        vel4 = None if vel4 is None else as_f4(vel4, self.n)
        check(self._L.nb_upload(self._h, ptr(pos4), ptr(vel4)))

    def download(self):
        pos = np.zeros((self.n, 4), np.float32)
        vel = np.zeros((self.n, 4), np.float32)
        check(self._L.nb_download(self._h, ptr(pos), ptr(vel)))
        return pos, vel

    def acc(self):
        a = np.zeros((self.n, 4), np.float32)
        check(self._L.nb_download_acc(self._h, ptr(a)))
        return a

    # -- compute
    def compute_forces(self):
        a = np.zeros((self.n, 4), np.float32)
        check(self._L.nb_compute_forces(self._h, ptr(a)))
        return a

    def step(self, nsteps=1):
        check(self._L.nb_step(self._h, int(nsteps)))

    def synchronize(self):
        check(self._L.nb_synchronize(self._h))

    def time_steps(self, nsteps):
        ms = ctypes.c_double()
        check(self._L.nb_time_steps(self._h, int(nsteps), ctypes.byref(ms)))
        return ms.value

    def total_energy(self):
        out = np.zeros(3)
        check(self._L.nb_total_energy(self._h, ptr(out)))
        return out

    def launch(self):
        """The launch configuration (nb_get_launch) as a dict; 0 = automatic."""
        l = _lib.NbLaunch()
        check(self._L.nb_get_launch(self._h, ctypes.byref(l)))
        return {k: getattr(l, k) for k, _ in _lib.NbLaunch._fields_}

    def set_launch(self, **kw):
        """Change fields of the launch configuration (nb_set_launch):
        allpairs_split, allpairs_tile, walk_stack, walk_grid."""
        cur = self.launch()
        unknown = set(kw) - set(cur)
        if unknown:
            raise ValueError(f"unknown launch fields {sorted(unknown)}")
        cur.update(kw)
        check(self._L.nb_set_launch(self._h, ctypes.byref(_lib.NbLaunch(**cur))))

    def autotune(self, steps=5, max_median_err=0.01):
        """Pick the fastest launch configuration within the error bound
        (nb_autotune); returns (choice dict, per-candidate log lines)."""
        r = _lib.NbTuneResult()
        buf = ctypes.create_string_buffer(8192)
        check(self._L.nb_autotune(self._h, int(steps), float(max_median_err), ctypes.byref(r),
                                  buf, len(buf)))
        choice = {k: getattr(r, k) for k, _ in _lib.NbTuneResult._fields_}
        if self.params.get("force") in ("bh", _lib.FORCE_BH) and not hasattr(self, "nsys"):
            self.params.update(group_width=r.group_width, key_bits=r.key_bits)
        return choice, buf.value.decode().splitlines()

    def total_energy_tree(self, theta=0.3):
        """(total, kinetic, potential) with the potential from a tree walk at
        theta (O(n log n); Barnes-Hut engines only)."""
        out = np.zeros(3)
        check(self._L.nb_total_energy_tree(self._h, float(theta), ptr(out)))
        return out

    def direct_f64(self, targets):
        """FP64 direct-sum accelerations (nt, 3) of the given original indices."""
        t = np.ascontiguousarray(targets, dtype=np.int32)
        out = np.zeros((len(t), 3))
        check(self._L.nb_direct_f64(self._h, ptr(t), len(t), ptr(out)))
        return out

    def force_error(self, targets, acc=None):
        """Relative error |a - a_fp64| / |a_fp64| of this engine's accelerations
        (the last force evaluation, or `acc`) on the given targets."""
        t = np.ascontiguousarray(targets, dtype=np.int32)
        a = self.acc() if acc is None else acc
        ref = self.direct_f64(t)
        return np.linalg.norm(a[t, :3].astype(np.float64) - ref, axis=1) / \
            np.linalg.norm(ref, axis=1)

    # -- timing
    def force_ms_total(self):
        return float(self._L.nb_force_ms_total(self._h))

    def force_calls(self):
        return int(self._L.nb_force_calls(self._h))

    def reset_force_timing(self):
        self._L.nb_reset_force_timing(self._h)

    def set_phase_timing(self, on):
        check(self._L.nb_set_phase_timing(self._h, 1 if on else 0))

    def bh_phase_ms(self):
        out = np.zeros(6)
        calls = ctypes.c_long()
        check(self._L.nb_bh_phase_ms(self._h, ptr(out), ctypes.byref(calls)))
        return out, calls.value

    def set_graph(self, on):
        """Replay each leapfrog step as a captured HIP graph."""
        check(self._L.nb_set_graph(self._h, 1 if on else 0))

    def bh_release(self):
        check(self._L.nb_bh_release(self._h))

    # -- tree introspection
    def tree(self):
        n = self.n
        keys = np.zeros(n, np.uint64)
        perm = np.zeros(n, np.uint32)
        children = np.zeros((max(n - 1, 1), 2), np.int32)
        parent = np.zeros(max(2 * n - 1, 1), np.int32)
        com4 = np.zeros((max(n - 1, 1), 4), np.float32)
        size2 = np.zeros(max(n - 1, 1), np.float32)
        pos_sorted = np.zeros((n, 4), np.float32)
        check(self._L.nb_get_tree(self._h, ptr(keys), ptr(perm), ptr(children), ptr(parent),
                                  ptr(com4), ptr(size2), ptr(pos_sorted)))
        return dict(keys=keys, perm=perm, children=children[: n - 1], parent=parent[: 2 * n - 1],
                    com4=com4[: n - 1], size2=size2[: n - 1], pos_sorted=pos_sorted)

    def traverse_counts(self, traverse="thread", group_width=32):
        t = TRAVERSALS[traverse] if isinstance(traverse, str) else int(traverse)
        slots = self.n if t == _lib.TRAVERSE_THREAD else -(-self.n // group_width)
        inter = np.zeros(slots, np.int32)
        opens = np.zeros(slots, np.int32)
        check(self._L.nb_traverse_counts(self._h, t, group_width, ptr(inter), ptr(opens)))
        return inter, opens


class Ensemble(Simulation):
    """Independent systems batched in one engine (nb_create_ensemble).

    ``systems`` is a list of (pos4, vel4) arrays; they are stored back to
    back and stepped together.  Barnes-Hut with the wave64 walk on
    (system, 30-bit Morton) keys; dt, eps, G, theta apply to every system.
    """

    def __init__(self, systems, dt=0.01, eps=0.05, G=1.0, theta=0.5, compensated=False,
                 far16=False):
        L = _lib.load()
        pos = [as_f4(p) for p, _ in systems]
        vel = [np.zeros_like(p) if v is None else as_f4(v, len(p))
               for p, (_, v) in zip(pos, systems)]
        self.counts = np.array([len(p) for p in pos], np.int32)
        self.n = int(self.counts.sum())
        self.nsys = len(pos)
        h = ctypes.c_void_p()
        check(L.nb_create_ensemble(self.nsys, ptr(self.counts), ctypes.byref(h)))
        self._h = h
        self._L = L
        self.offsets = np.zeros(self.nsys + 1, np.int32)
        check(L.nb_ensemble_offsets(self._h, ptr(self.offsets)))
        self.set_params(dt=dt, eps=eps, G=G, theta=theta, force="bh", traverse="warp",
                        key_bits=30, group_width=64, compensated=compensated, far16=far16)
        self.upload(np.concatenate(pos), np.concatenate(vel))

    def split(self, a):
        """Per-system views of an (n, ...) array in ensemble order."""
        return [a[self.offsets[s]:self.offsets[s + 1]] for s in range(self.nsys)]

    def items(self):
        """Walk work items (system, first sorted target, count, 0) in queue order."""
        k = ctypes.c_int()
        check(self._L.nb_ensemble_items(self._h, None, ctypes.byref(k)))
        out = np.zeros((k.value, 4), np.int32)
        check(self._L.nb_ensemble_items(self._h, ptr(out), ctypes.byref(k)))
        return out

    def roots(self):
        out = np.zeros(self.nsys, np.int32)
        check(self._L.nb_ensemble_roots(self._h, ptr(out)))
        return out

    def system_energy(self):
        """(nsys, 3): total, kinetic, potential of every system (FP64)."""
        out = np.zeros((self.nsys, 3))
        check(self._L.nb_system_energy(self._h, ptr(out)))
        return out

    def traverse_counts(self, traverse="warp", group_width=64):
        slots = len(self.items())
        inter = np.zeros(slots, np.int32)
        opens = np.zeros(slots, np.int32)
        t = TRAVERSALS[traverse] if isinstance(traverse, str) else int(traverse)
        check(self._L.nb_traverse_counts(self._h, t, group_width, ptr(inter), ptr(opens)))
        return inter, opens


def morton_keys(pos, key_bits=63):
    """GPU Morton keys of (n,3) or (n,4) positions (lbvh_check.py:45-54)."""
    pos = np.asarray(pos)
    if pos.ndim == 2 and pos.shape[1] == 3:
        pos = np.column_stack([pos, np.ones(len(pos))])
    pos4 = as_f4(pos)
    keys = np.zeros(len(pos4), np.uint64)
    check(_lib.load().nb_morton(ptr(pos4), len(pos4), key_bits, ptr(keys)))
    return keys


def sort_pairs(keys, key_bits=63):
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    out = np.zeros_like(keys)
    perm = np.zeros(len(keys), np.uint32)
    check(_lib.load().nb_sort_pairs(ptr(keys), len(keys), key_bits, ptr(out), ptr(perm)))
    return out, perm


def build_tree(keys_sorted, key_bits=63):
    """GPU Karras build (lbvh_check.py:57-100): (children (n-1,2), parent (2n-1,))."""
    keys_sorted = np.ascontiguousarray(keys_sorted, dtype=np.uint64)
    n = len(keys_sorted)
    children = np.zeros((n - 1, 2), np.int32)
    parent = np.zeros(2 * n - 1, np.int32)
    check(_lib.load().nb_karras_build(ptr(keys_sorted), n, key_bits, ptr(children), ptr(parent)))
    return children, parent


def accelerations(pos, mass, eps2, G=1.0, chunk=None, dtype=np.float32):
    """GPU all-pairs with the signature of reference_nbody.accelerations
    (chunk is accepted for compatibility and ignored)."""
    pos = np.asarray(pos, dtype=np.float32)
    pos4 = np.column_stack([pos, np.asarray(mass, dtype=np.float32)])
    with Simulation(pos4, eps=float(np.sqrt(eps2)), G=float(G), force="naive") as s:
        a = s.compute_forces()
    return a[:, :3].astype(dtype)


def total_energy(pos, vel, mass, eps2, G=1.0, chunk=None):
    """GPU FP64 energy with the signature of reference_nbody.total_energy."""
    pos4 = np.column_stack([np.asarray(pos, np.float32), np.asarray(mass, np.float32)])
    vel4 = np.column_stack([np.asarray(vel, np.float32), np.zeros(len(pos4), np.float32)])
    with Simulation(pos4, vel4, eps=float(np.sqrt(eps2)), G=float(G)) as s:
        e = s.total_energy()
    return float(e[0]), float(e[1]), float(e[2])


def barnes_hut_accelerations(pos4, eps, theta=0.5, G=1.0, traverse="warp", group_width=64,
                             key_bits=63):
    with Simulation(pos4, eps=eps, G=G, theta=theta, force="bh", traverse=traverse,
                    group_width=group_width, key_bits=key_bits) as s:
        return s.compute_forces()
