"""Reporting reductions of planet/Reporting.h restated in numpy (test infrastructure: the oracle for
the device reductions, never the product path).

search_minmax_3d: searchMinMax_3D (:261-297), loops j, k, i with `>` for the max and `else if <`
  for the min, both starting from the cell (0, 0, 0): the first cell of that order on ties, NaNs
  never taken unless the starting cell is NaN.
search_minmax_2d: searchMinMax_2D (:299-330), starting from (0, 0), the interior j in [1, jm-2],
  k in [1, km-2].
continuity_residual: steadyQuery's residual (:372-398), loops i, j, k over the interior with
  `fabs(r) >= minimum` from 0: the last cell on ties; NaNs never taken.
steady_changes: steadyQuery's per-field changes (:399-410), |f - f_n| over every cell, loops i, j, k
  with `d >= val` from 0: the last cell on ties; NaNs never taken; (0.0, (0, 0, 0)) when none is.
"""
import math

import numpy as np


def _first_extreme(vals, order_first, fn):
    ok = ~np.isnan(vals)
    if not ok.any():
        return None
    best = fn(vals[ok])
    idx = np.flatnonzero(ok & (vals == best))
    return int(idx[0] if order_first else idx[-1])


def search_minmax_3d(f):
    """(max, (i, j, k), min, (i, j, k)) of an (im, jm, km) field."""
    im, jm, km = f.shape
    flat = np.ascontiguousarray(np.transpose(f, (1, 2, 0))).ravel()      # the j, k, i loop order
    if np.isnan(flat[0]):
        return flat[0], (0, 0, 0), flat[0], (0, 0, 0)
    out = []
    for fn in (np.max, np.min):
        p = _first_extreme(flat, True, fn)
        j, r = divmod(p, km * im)
        k, i = divmod(r, im)
        out.append((float(flat[p]), (i, j, k)))
    return out[0][0], out[0][1], out[1][0], out[1][1]


def search_minmax_2d(m):
    """(max, (j, k), min, (j, k)) of a (jm, km) map."""
    jm, km = m.shape
    if np.isnan(m[0, 0]):
        return m[0, 0], (0, 0), m[0, 0], (0, 0)
    inner = m[1:jm - 1, 1:km - 1]
    vals = np.concatenate([[m[0, 0]], inner.ravel()])
    out = []
    for fn in (np.max, np.min):
        p = _first_extreme(vals, True, fn)
        if p == 0:
            out.append((float(vals[0]), (0, 0)))
        else:
            j, k = divmod(p - 1, km - 2)
            out.append((float(vals[p]), (j + 1, k + 1)))
    return out[0][0], out[0][1], out[1][0], out[1][1]


def grid_axes(G):
    """rad.z and the.z as the model accumulates them (r0 + dr + dr ..., 0 + dthe + dthe ...)."""
    rad, the = [G.r0], [0.0]
    for _ in range(1, G.im):
        rad.append(rad[-1] + G.dr)
    for _ in range(1, G.jm):
        the.append(the[-1] + G.dthe)
    return np.array(rad), np.array(the)


def continuity_residual(u, v, w, rad, the, dr, dthe, dphi):
    """(max |residual|, (i, j, k)); (0.0, (-1, -1, -1)) when no residual is finite."""
    im, jm, km = u.shape
    sinthe = np.array([math.sin(t) for t in the])[None, 1:jm - 1, None]
    costhe = np.array([math.cos(t) for t in the])[None, 1:jm - 1, None]
    rm = np.asarray(rad, dtype=np.float64)[1:im - 1, None, None]
    rmsinthe = rm * sinthe
    c = (slice(1, im - 1), slice(1, jm - 1), slice(1, km - 1))
    dudr = (u[2:, 1:jm - 1, 1:km - 1] - u[:im - 2, 1:jm - 1, 1:km - 1]) / (2. * dr)
    dvdthe = (v[1:im - 1, 2:, 1:km - 1] - v[1:im - 1, :jm - 2, 1:km - 1]) / (2. * dthe)
    dwdphi = (w[1:im - 1, 1:jm - 1, 2:] - w[1:im - 1, 1:jm - 1, :km - 2]) / (2. * dphi)
    res = dudr + 2. * u[c] / rm + dvdthe / rm + costhe / rmsinthe * v[c] + dwdphi / rmsinthe
    a = np.abs(res).ravel()                                  # the i, j, k loop order
    p = _first_extreme(a, False, np.max)
    if p is None:
        return 0.0, (-1, -1, -1)
    i, r = divmod(p, (jm - 2) * (km - 2))
    j, k = divmod(r, km - 2)
    return float(a[p]), (i + 1, j + 1, k + 1)


STEADY_FIELDS = ("p_dyn", "u", "v", "w", "t", "h2o", "h2o_cloud", "h2o_ice", "nh3", "nh3_cloud", "nh3_ice", "nh4sh")


def steady_changes(fields, fields_n):
    """[(name, max |f - f_n|, (i, j, k))] in the report's order (Reporting.h:356-369, :399-410)."""
    out = []
    for name in STEADY_FIELDS:
        f, fn = fields[name], fields_n[name]
        d = np.abs(f - fn).ravel()                      # the i, j, k loop order of a C array
        p = _first_extreme(d, False, np.max)
        if p is None:
            out.append((name, 0.0, (0, 0, 0)))
            continue
        i, r = divmod(p, f.shape[1] * f.shape[2])
        j, k = divmod(r, f.shape[2])
        out.append((name, float(d[p]), (int(i), int(j), int(k))))
    return out
