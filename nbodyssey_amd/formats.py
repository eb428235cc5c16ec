"""On-disk formats of the reference (docs/FORMATS.md:1-31, main.cu:114-179).

* ``ic.bin``: int32 n, float4 pos[n] (x, y, z, m), float4 vel[n] (vx, vy, vz, 0),
  little-endian (FORMATS.md:7-18, main.cu:115-128).
* ``frame_%05d.bin``: int32 n, float4 pos[n] (FORMATS.md:20-28, main.cu:130-138).
* energy CSV: header ``step,time,energy,kinetic,potential,rel_error`` and rows
  ``%d,%.5f,%.8e,%.8e,%.8e,%.3e`` (main.cu:159-160,172-179).
"""
import os

import numpy as np

ENERGY_HEADER = "step,time,energy,kinetic,potential,rel_error\n"


class FormatError(ValueError):
    pass


def write_ic(path, pos4, vel4):
    pos4 = np.ascontiguousarray(pos4, dtype="<f4").reshape(-1, 4)
    vel4 = np.ascontiguousarray(vel4, dtype="<f4").reshape(-1, 4)
    if len(pos4) != len(vel4):
        raise FormatError("pos and vel lengths differ")
    with open(path, "wb") as f:
        f.write(np.int32(len(pos4)).astype("<i4").tobytes())
        f.write(pos4.tobytes())
        f.write(vel4.tobytes())


def read_ic(path):
    """Same checks and messages as main.cu:115-128."""
    try:
        f = open(path, "rb")
    except OSError:
        raise FormatError(f"cannot open IC file: {path}")
    with f:
        head = f.read(4)
        if len(head) < 4:
            raise FormatError(f"bad particle count in {path}")
        n = int(np.frombuffer(head, "<i4")[0])
        if n <= 0:
            raise FormatError(f"bad particle count in {path}")
        body = f.read(32 * n)
        if len(body) < 32 * n:
            raise FormatError(f"truncated IC file: {path}")
    arr = np.frombuffer(body, "<f4")
    pos4 = arr[: 4 * n].reshape(n, 4).copy()
    vel4 = arr[4 * n:].reshape(n, 4).copy()
    return pos4, vel4


def write_frame(path, pos4):
    pos4 = np.ascontiguousarray(pos4, dtype="<f4").reshape(-1, 4)
    with open(path, "wb") as f:
        f.write(np.int32(len(pos4)).astype("<i4").tobytes())
        f.write(pos4.tobytes())


def frame_path(out_dir, frame):
    return os.path.join(out_dir, "frame_%05d.bin" % frame)


def read_frame(path):
    with open(path, "rb") as f:
        n = int(np.frombuffer(f.read(4), "<i4")[0])
        pos4 = np.frombuffer(f.read(16 * n), "<f4").reshape(n, 4).copy()
    return pos4


def energy_row(step, dt, e, e0):
    """One CSV row, main.cu:174-177: e and e0 are (total, kinetic, potential)."""
    rel = (e[0] - e0[0]) / abs(e0[0])
    return "%d,%.5f,%.8e,%.8e,%.8e,%.3e\n" % (step, step * dt, e[0], e[1], e[2], rel)


def read_energy_csv(path):
    """Parse an energy log into a dict of numpy columns."""
    with open(path) as f:
        header = f.readline()
        if header != ENERGY_HEADER:
            raise FormatError(f"unexpected energy header in {path}: {header!r}")
        rows = [line.strip().split(",") for line in f if line.strip()]
    cols = list(zip(*rows)) if rows else [[]] * 6
    return {
        "step": np.array(cols[0], dtype=np.int64),
        "time": np.array(cols[1], dtype=np.float64),
        "energy": np.array(cols[2], dtype=np.float64),
        "kinetic": np.array(cols[3], dtype=np.float64),
        "potential": np.array(cols[4], dtype=np.float64),
        "rel_error": np.array(cols[5], dtype=np.float64),
    }
