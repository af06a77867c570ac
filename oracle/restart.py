"""The reference's restart file format, restated with numpy (test infrastructure: the checker of the
library's atjup_hip_save_state / atjup_hip_load_state).

cJupiterModel::save_state (planet/FileIO_Jup.cpp:584-624): a header of five int32 {magic, im, jm,
km, iter} with magic "JUP2" (0x4A555032) when the carried arrays follow and "JUP1" (0x4A555031)
otherwise, then every array of restart_arrays() (:281-293) as im*jm*km FP64 values in [i][j][k]
order, in JUP2 files the arrays of restart_arrays_carried() (:297-305) and vel_star ([j][k]).
load_state (:626-684) reads either magic and rejects a grid mismatch or a short file.
"""
from __future__ import annotations

import numpy as np

from atjup_amd import abi

MAGIC = {True: 0x4A555032, False: 0x4A555031}


def write(path, fields: dict, vel_star: np.ndarray, iteration: int, carried: bool = True):
    """fields: name -> [im][jm][km] array, by the reference's names (n-level copies as 'tn' etc.)."""
    shape = fields["t"].shape
    with open(path, "wb") as f:
        np.array([MAGIC[carried], *shape, iteration], dtype="<i4").tofile(f)
        names = abi.RESTART_FIXED + (abi.RESTART_CARRIED if carried else ())
        for n in names:
            np.ascontiguousarray(fields[n], dtype="<f8").tofile(f)
        if carried:
            np.ascontiguousarray(vel_star, dtype="<f8").tofile(f)


def read(path, shape=None):
    """(iteration, carried, {name: array}, vel_star or None)."""
    raw = np.fromfile(path, dtype="<i4", count=5)
    if len(raw) != 5 or raw[0] not in MAGIC.values():
        raise ValueError("bad header")
    carried = bool(raw[0] == MAGIC[True])
    im, jm, km = (int(x) for x in raw[1:4])
    if shape is not None and tuple(shape) != (im, jm, km):
        raise ValueError("grid mismatch")
    n = im * jm * km
    names = abi.RESTART_FIXED + (abi.RESTART_CARRIED if carried else ())
    data = np.fromfile(path, dtype="<f8", offset=20)
    need = len(names) * n + (jm * km if carried else 0)
    if data.size < need:
        raise ValueError("truncated file")
    out = {name: data[a * n:(a + 1) * n].reshape(im, jm, km) for a, name in enumerate(names)}
    vs = data[len(names) * n:len(names) * n + jm * km].reshape(jm, km) if carried else None
    return int(raw[4]), carried, out, vs
