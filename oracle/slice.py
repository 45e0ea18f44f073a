"""ORACLE - TEST INFRASTRUCTURE ONLY.

numpy restatement of the device slice sampler (pcfd_sample_slice): an nu x nv image of a field
on the plane x[axis] = coord, each pixel the cell-centred value of the finest leaf cell that
contains the pixel centre, NaN where that cell is solid or the point lies outside the grid.
Fields as in the reference viewer's list (gui/slice_field.h:20-31): 0 |v|, 1 u, 2 v, 3 w,
4 pressure, 5 Cp.  Only tests/ import this module.
"""
from __future__ import annotations

import numpy as np


def brick_maps(topo, instance=0):
    """Per level: array [z, y, x] of brick ids (-1 where no stored brick) of one instance."""
    info = topo.info
    levels = int(info.levels)
    base = [int(info.base_bricks[d]) for d in range(3)]
    lev = topo.array("brick_level")
    coord = topo.array("brick_coord").reshape(-1, 3)
    inst = topo.instance(instance) if topo.instances > 1 else {"brick_begin": 0, "brick_end": int(info.bricks)}
    maps = []
    for lv in range(levels):
        dims = [b << lv for b in base]
        m = np.full((dims[2], dims[1], dims[0]), -1, dtype=np.int64)
        sel = np.arange(inst["brick_begin"], inst["brick_end"])
        sel = sel[lev[sel] == lv]
        c = coord[sel]
        m[c[:, 2], c[:, 1], c[:, 0]] = sel
        maps.append(m)
    return maps


def sample_slice(topo, v, p, field, axis, coord, lo, hi, nu, nv, qinf=1.0, instance=0):
    """v: (vx, vy, vz) arrays over the dofs, p: pressure; returns float32 [nv, nu]."""
    info = topo.info
    levels = int(info.levels)
    glo = np.array([info.lo[d] for d in range(3)])
    hs = topo.array("level_h")
    mask = topo.array("mask")
    maps = brick_maps(topo, instance)
    a1, a2 = (1, 2) if axis == 0 else ((0, 2) if axis == 1 else (0, 1))
    i = np.arange(nu, dtype=np.float64)[None, :].repeat(nv, 0)
    j = np.arange(nv, dtype=np.float64)[:, None].repeat(nu, 1)
    x = [None, None, None]
    x[axis] = np.full((nv, nu), float(coord))
    x[a1] = lo[0] + (i + 0.5) * (hi[0] - lo[0]) / nu
    x[a2] = lo[1] + (j + 0.5) * (hi[1] - lo[1]) / nv
    out = np.full((nv, nu), np.nan, dtype=np.float32)
    found = np.zeros((nv, nu), dtype=bool)
    for lv in range(levels - 1, -1, -1):
        t = [np.floor((x[d] - glo[d]) / hs[lv]) for d in range(3)]
        dims = maps[lv].shape[::-1]
        inside = np.ones((nv, nu), dtype=bool)
        for d in range(3):
            inside &= (t[d] >= 0) & (t[d] < 8.0 * dims[d])
        c = [np.where(inside, t[d], 0).astype(np.int64) for d in range(3)]
        b = maps[lv][c[2] >> 3, c[1] >> 3, c[0] >> 3]
        hit = inside & (b >= 0) & ~found
        a = b * 512 + (c[0] & 7) + 8 * (c[1] & 7) + 64 * (c[2] & 7)
        a = np.where(hit, a, 0)
        active = hit & ((mask[a] & 64) != 0)
        vx, vy, vz = (comp[a] for comp in v)
        vals = {0: np.sqrt(vx * vx + vy * vy + vz * vz), 1: vx, 2: vy, 3: vz, 4: p[a], 5: p[a] / qinf}[field]
        out[active] = vals[active].astype(np.float32)
        found |= hit
    return out
