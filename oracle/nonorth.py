"""ORACLE - TEST INFRASTRUCTURE ONLY.

numpy/scipy restatement of the composite pressure operator with the non-orthogonal correction
(CompositeAmrPressureSystem::apply_cpu, amr_pressure.cpp:498-521, with
reconstruct_pressure_gradients / add_nonorthogonal_pressure_flux /
regular_pressure_derivative_delta, amr_pressure.cpp:193-212) and of the face correction that
goes with it (correct_fluxes, amr_pressure.cu:2040-2069), as sparse matrices over the dofs of a
topology on which build_nonorthogonal() has run.  Only tests/ import this module.

  orthogonal_matrix: structured faces h (p_a - p_b), compact edges coef (p_a - p_b), the outlet
                     2h p at the x-max face (the operator the PCG path solves)
  gradient_matrix:   g[n] = sum_q w_q (p[nbr_q] - p[dof_n]), rows 3n + c
  face_extra_matrix: per correction face, delta (p_u - p_l) + k . ((1-w) g_l + w g_u)
  full_matrix:       orthogonal + scatter of F_f = -area * extra_f (out[l] += F, out[u] -= F)
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def _structured_pairs(topo):
    """(lower, upper, h) of every structured face (+axis bit of the lower cell)."""
    mask = topo.array("mask")
    nbr = topo.array("brick_nbr").reshape(-1, 6)
    h = topo.array("level_h")[topo.array("brick_level")]
    n = mask.size
    a = np.arange(n)
    b_of, q = a >> 9, a & 511
    loc = np.stack([q & 7, (q >> 3) & 7, q >> 6], axis=1)
    lows, ups, hs = [], [], []
    for axis in range(3):
        sel = np.flatnonzero((mask & (1 << axis)) != 0)
        stride = (1, 8, 64)[axis]
        inner = loc[sel, axis] < 7
        up = np.where(inner, sel + stride, nbr[b_of[sel], 2 * axis + 1] * 512 + (q[sel] - 7 * stride))
        lows.append(sel)
        ups.append(up)
        hs.append(h[b_of[sel]])
    return np.concatenate(lows), np.concatenate(ups), np.concatenate(hs)


def orthogonal_matrix(topo):
    n = int(topo.ndof)
    lo, up, h = _structured_pairs(topo)
    ea, eb, ec = topo.array("edge_a"), topo.array("edge_b"), topo.array("edge_coef")
    mask = topo.array("mask")
    act = (mask & 64) != 0
    keep = act[ea] & act[eb]
    ea, eb, ec = ea[keep], eb[keep], ec[keep]
    i = np.concatenate([lo, up, ea, eb])
    j = np.concatenate([up, lo, eb, ea])
    c = np.concatenate([h, h, ec, ec])
    off = sp.coo_matrix((-c, (i, j)), shape=(n, n))
    diag = np.bincount(i, weights=c, minlength=n)
    if topo.info.outlet:
        flags = topo.array("brick_flags")
        hb = topo.array("level_h")[topo.array("brick_level")]
        a = np.arange(n)
        out = act & ((flags[a >> 9] & 2) != 0) & ((a & 7) == 7)
        diag[out] += 2.0 * hb[a[out] >> 9]
    return (off + sp.diags(diag)).tocsr()


def gradient_matrix(topo):
    n = int(topo.ndof)
    off, nbr, w, dof = topo.array("ng_off"), topo.array("ng_nbr"), topo.array("ng_w").reshape(-1, 3), topo.array("ng_dof")
    nn = dof.size
    seg = np.repeat(np.arange(nn), np.diff(off))
    rows = (3 * seg[:, None] + np.arange(3)[None, :]).ravel()
    G = sp.coo_matrix((w.ravel(), (rows, np.repeat(nbr, 3))), shape=(3 * nn, n))
    G = G - sp.coo_matrix((w.ravel(), (rows, np.repeat(dof[seg], 3))), shape=(3 * nn, n))
    return G.tocsr()


def face_extra_matrix(topo):
    """Rows: delta (p_u - p_l) + k . g_f of each correction face (the part of dp/dn the
    orthogonal operator does not carry)."""
    n = int(topo.ndof)
    lo, up = topo.array("nf_lower"), topo.array("nf_upper")
    delta, k, w = topo.array("nf_delta"), topo.array("nf_k").reshape(-1, 3), topo.array("nf_w")
    ln, un = topo.array("nf_lnode"), topo.array("nf_unode")
    nf = lo.size
    f = np.arange(nf)
    E = sp.coo_matrix((np.concatenate([delta, -delta]), (np.concatenate([f, f]), np.concatenate([up, lo]))), shape=(nf, n)).tocsr()
    G = gradient_matrix(topo)
    has = (ln >= 0) & (un >= 0)
    fh = f[has]
    if fh.size:
        nn3 = G.shape[0]
        rows, cols, vals = [], [], []
        for c in range(3):
            rows += [fh, fh]
            cols += [3 * ln[has] + c, 3 * un[has] + c]
            vals += [k[has, c] * (1 - w[has]), k[has, c] * w[has]]
        K = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(nf, nn3)).tocsr()
        E = E + K @ G
    return E.tocsr()


def scatter_matrix(topo):
    lo, up = topo.array("nf_lower"), topo.array("nf_upper")
    nf = lo.size
    f = np.arange(nf)
    return sp.coo_matrix((np.concatenate([np.ones(nf), -np.ones(nf)]), (np.concatenate([lo, up]), np.concatenate([f, f]))),
                         shape=(int(topo.ndof), nf)).tocsr()


def full_matrix(topo):
    area = topo.array("nf_area")
    return (orthogonal_matrix(topo) - scatter_matrix(topo) @ sp.diags(area) @ face_extra_matrix(topo)).tocsr()


def corrected_faces(topo, faces, p, rho, dt):
    """Face velocities after the full correction u_f -= (dt/rho) dp/dn, from the faces after the
    orthogonal correction: the extra derivative of each correction face on its structured face
    (f[axis][lower]) or compact edge."""
    fx, fy, fz, fe = (np.array(c, dtype=np.float64) for c in faces)
    extra = face_extra_matrix(topo) @ p
    s = dt / rho
    edge, lo = topo.array("nf_edge"), topo.array("nf_lower")
    comp = edge >= 0
    np.subtract.at(fe, edge[comp], s * extra[comp])
    for axis, arr in enumerate((fx, fy, fz)):
        sel = edge == -1 - axis
        np.subtract.at(arr, lo[sel], s * extra[sel])
    return fx, fy, fz, fe
