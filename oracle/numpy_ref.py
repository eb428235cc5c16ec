"""NumPy restatements of the reference's CPU scripts -- TEST INFRASTRUCTURE ONLY.

* ``accelerations`` / ``total_energy`` / ``kdk_run`` restate
  scripts/reference_nbody.py:45-81,130-136 (chunked (b, n, 3) all-pairs in the
  integration dtype, FP64 energy, kick-drift-kick).  This is the CPU baseline
  that bench.py times (optionally over a multiprocessing pool, one block of
  targets per worker).  Pinned by: two-body closed forms, Newton's third law,
  agreement with the C FP64 oracle, and the committed reference energy trace
  (E0 of benchmarks/energy_gpu_naive.csv).
* ``com_f64`` restates lbvh_check.py:121-148 in FP64 (post-order COM/AABB).
  Pinned by: root mass == total mass and root COM == mass-weighted mean.
"""
import numpy as np


def _target_block(pos, mass, eps2, chunk, dtype, t0, t1):
    """Accelerations (without G) of targets [t0, t1), chunk targets at a time."""
    acc = np.zeros((t1 - t0, 3), dtype=dtype)
    for lo in range(t0, t1, chunk):
        hi = min(lo + chunk, t1)
        d = pos[np.newaxis, :, :] - pos[lo:hi, np.newaxis, :]
        r = np.sqrt(np.sum(d * d, axis=2) + eps2)
        inv3 = dtype(1.0) / (r * r * r)
        acc[lo - t0:hi - t0] = np.einsum("bn,bnk->bk", mass[np.newaxis, :] * inv3, d,
                                         dtype=dtype)
    return acc


def _pool_block(args):
    return _target_block(*args)


def accelerations(pos, mass, eps2, G, chunk=2048, dtype=np.float32, pool=None):
    """reference_nbody.py:45-61: softened acceleration, chunked over targets.
    With a multiprocessing pool, the targets are split into one contiguous
    block per worker and each block runs the same chunked expression."""
    n = pos.shape[0]
    if pool is None:
        acc = _target_block(pos, mass, eps2, chunk, dtype, 0, n)
    else:
        k = pool._processes
        edges = np.linspace(0, n, k + 1).astype(int)
        parts = pool.map(_pool_block, [(pos, mass, eps2, chunk, dtype, int(edges[i]),
                                        int(edges[i + 1])) for i in range(k)])
        acc = np.concatenate(parts)
    return (G * acc).astype(dtype)


def total_energy(pos, vel, mass, eps2, G, chunk=2048):
    """reference_nbody.py:64-81: FP64 KE + PE with the self terms removed."""
    p = pos.astype(np.float64)
    v = vel.astype(np.float64)
    m = mass.astype(np.float64)
    ke = 0.5 * np.sum(m * np.sum(v * v, axis=1))
    n = p.shape[0]
    pe = 0.0
    for lo in range(0, n, chunk):
        hi = min(lo + chunk, n)
        d = p[np.newaxis, :, :] - p[lo:hi, np.newaxis, :]
        inv = 1.0 / np.sqrt(np.sum(d * d, axis=2) + float(eps2))
        pe += np.sum((m[lo:hi, np.newaxis] * m[np.newaxis, :]) * inv)
    self_term = np.sum(m * m) / np.sqrt(float(eps2))
    pe = -float(G) * 0.5 * (pe - self_term)
    return ke + pe, ke, pe


def kdk_run(pos, vel, mass, eps, dt, steps, G=1.0, chunk=2048, dtype=np.float32, pool=None,
            budget_s=None):
    """reference_nbody.py:129-136 without I/O: returns final (pos, vel, acc).
    budget_s: stop after the first step that ends past this many seconds
    (the function attribute `steps_done` then says how many ran)."""
    import time
    eps2 = dtype(eps * eps)
    dt = dtype(dt)
    G = dtype(G)
    pos = pos.astype(dtype).copy()
    vel = vel.astype(dtype).copy()
    mass = mass.astype(dtype)
    t0 = time.perf_counter()
    acc = accelerations(pos, mass, eps2, G, chunk, dtype, pool)
    kdk_run.steps_done = 0
    for _ in range(steps):
        vel += dtype(0.5) * dt * acc
        pos += dt * vel
        acc = accelerations(pos, mass, eps2, G, chunk, dtype, pool)
        vel += dtype(0.5) * dt * acc
        kdk_run.steps_done += 1
        if budget_s is not None and time.perf_counter() - t0 > budget_s:
            break
    return pos, vel, acc


def com_f64(children, pos_sorted, mass_sorted):
    """lbvh_check.py:121-148 in FP64: COM (xyz, m) and longest-edge size."""
    n = len(pos_sorted)
    children = np.asarray(children).reshape(-1, 2)
    com = np.zeros((n - 1, 4))
    lo = np.zeros((n - 1, 3))
    hi = np.zeros((n - 1, 3))
    order, stack = [], [0]
    while stack:
        i = stack.pop()
        if i < n - 1:
            order.append(i)
            stack.extend(int(c) for c in children[i])

    def data(i):
        if i >= n - 1:
            k = i - (n - 1)
            p = pos_sorted[k].astype(np.float64)
            return np.append(p, float(mass_sorted[k])), p, p
        return com[i], lo[i], hi[i]

    for i in reversed(order):
        (cl, ll, hl), (cr, lr, hr) = (data(int(c)) for c in children[i])
        m = cl[3] + cr[3]
        com[i, :3] = (cl[:3] * cl[3] + cr[:3] * cr[3]) / m
        com[i, 3] = m
        lo[i] = np.minimum(ll, lr)
        hi[i] = np.maximum(hl, hr)
    size = (hi - lo).max(axis=1) if n > 1 else np.zeros(0)
    return com, size
