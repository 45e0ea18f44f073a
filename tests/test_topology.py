"""Invariants of the shared brick / cut-cell topology and the multigrid hierarchy
(host preprocessing, no GPU).  Closed forms: sphere volume and area, discrete closure of
every control volume, 2:1 nesting, symmetric coarse operators."""
import math

import numpy as np
import pytest

from paracfd_amd import _lib as L, cases


@pytest.fixture(scope="module")
def sphere():
    g, d = cases.sphere_box(cells_x=64)
    return L.Topology(g, d)


@pytest.fixture(scope="module")
def sphere_amr():
    g, d = cases.sphere_box(cells_x=32, levels=2, band=0.3)
    return L.Topology(g, d)


@pytest.fixture(scope="module")
def wing():
    g, d = cases.naca_wing(digits=2412, alpha_deg=4.0, cells_x=48, levels=2, bands=(0.3,))
    return L.Topology(g, d)


def cell_closure(t):
    """sum over each cell's open faces of (outward normal * open area) plus its wall area vector"""
    n = t.ndof
    mask = t.array("mask")
    h = t.array("level_h")[t.array("brick_level")]
    hd = np.repeat(h, 512)
    flags = np.repeat(t.array("brick_flags"), 512)
    q = np.arange(n) % 512
    ijk = [q & 7, (q >> 3) & 7, q >> 6]
    clo = np.zeros((n, 3))
    A = hd * hd
    for ax in range(3):
        clo[:, ax] += np.where(mask & (1 << ax), A, 0.0)
        clo[:, ax] -= np.where(mask & (8 << ax), A, 0.0)
        lo_flag, hi_flag = 1 << (2 * ax), 2 << (2 * ax)
        clo[:, ax] -= np.where((ijk[ax] == 0) & (flags & lo_flag != 0), A, 0.0)
        clo[:, ax] += np.where((ijk[ax] == 7) & (flags & hi_flag != 0), A, 0.0)
    ea, eb, ax_, area = t.array("edge_a"), t.array("edge_b"), t.array("edge_axis"), t.array("edge_area")
    np.add.at(clo, (ea, ax_), area)
    np.add.at(clo, (eb, ax_), -area)
    S = t.array("patch_S").reshape(-1, 3)
    np.add.at(clo, t.array("patch_dof"), S)
    act = (mask & 64) != 0
    return clo[act], A[act]


@pytest.mark.parametrize("name", ["sphere", "sphere_amr", "wing", "glider"])
def test_every_control_volume_closes(name, request):
    t = request.getfixturevalue(name)
    clo, A = cell_closure(t)
    assert np.max(np.abs(clo) / A[:, None]) < 1e-12


def test_sphere_volume_and_area(sphere):
    r = 0.25  # 4 cells per radius here; cut cells below 10 % fluid become solid
    assert abs(sphere.info.body_volume - 4.0 / 3.0 * math.pi * r ** 3) < 0.03 * 4.0 / 3.0 * math.pi * r ** 3
    S = sphere.array("patch_S").reshape(-1, 3)
    assert np.allclose(S.sum(axis=0), 0.0, atol=1e-12)
    area = np.linalg.norm(S, axis=1).sum()
    assert abs(area - 4 * math.pi * r ** 2) < 0.05 * 4 * math.pi * r ** 2
    # wall normals point from the fluid into the body (towards the centre)
    X = sphere.array("patch_x").reshape(-1, 3)
    outward = (X - np.array([1.0, 1.0, 1.0]))
    assert np.mean(np.sum(S * outward, axis=1) < 0) > 0.99


def test_wing_volume(wing):
    # area of the NACA 2412 section (closed trailing edge) times span
    info = wing.info
    assert abs(info.body_volume - 0.0822 * 2.03) < 0.03 * 0.0822 * 2.03
    assert info.reference_area == pytest.approx(2.03)


@pytest.mark.parametrize("name", ["sphere_amr", "wing"])
def test_coarse_fine_edges(name, request):
    t = request.getfixturevalue(name)
    lev = np.repeat(t.array("brick_level"), 512)
    h = t.array("level_h")
    ea, eb = t.array("edge_a"), t.array("edge_b")
    la, lb = lev[ea], lev[eb]
    assert np.all(np.abs(la - lb) <= 1)
    ha, hb = h[la], h[lb]
    coef, area, w = t.array("edge_coef"), t.array("edge_area"), t.array("edge_w")
    assert np.allclose(coef, area / (0.5 * (ha + hb)), rtol=1e-14)
    # between two full cells the centroid weight is the cell-size weight h_a / (h_a + h_b); cut
    # cells use the face's position between their centroids (test_oracle_pins.py)
    vol = h[lev] ** 3
    vol[t.array("special_dof")] = t.array("special_volume")
    full = np.abs(vol - h[lev] ** 3) <= 1e-15 * h[lev] ** 3
    both = full[ea] & full[eb]
    assert both.sum() > 0
    assert np.allclose(w[both], (ha / (ha + hb))[both], rtol=1e-14)
    cf = la != lb
    assert cf.sum() > 0
    # a fine face of a coarse/fine edge is never larger than the fine cell face
    assert np.all(area[cf] <= np.minimum(ha, hb)[cf] ** 2 * (1 + 1e-12))


def test_structured_faces_are_symmetric(sphere_amr):
    t = sphere_amr
    mask = t.array("mask").astype(np.int64)
    nbr = t.array("brick_nbr").reshape(-1, 6)
    n = t.ndof
    idx = np.arange(n)
    b, q = idx // 512, idx % 512
    for ax in range(3):
        has = (mask & (1 << ax)) != 0
        stride = 8 ** ax
        c = (q >> (3 * ax)) & 7
        inside = c < 7
        other = np.where(inside, idx + stride, nbr[b, 2 * ax + 1] * 512 + q - 7 * stride)
        sel = has
        assert np.all(nbr[b[sel & ~inside], 2 * ax + 1] >= 0)
        assert np.all((mask[other[sel]] & (8 << ax)) != 0)
        assert np.all((mask[other[sel]] & 64) != 0)


@pytest.mark.parametrize("name", ["sphere", "wing", "glider"])
def test_multigrid_levels(name, request):
    t = request.getfixturevalue(name)
    prev = t.n_active
    for m in range(1, t.mg_levels):
        ro, col, coef, diag = (t.mg_array(m, k) for k in ("row_off", "col", "coef", "diag"))
        n = diag.size
        assert n < prev
        prev = n
        A = np.zeros((n, n)) if n <= 3000 else None
        rows = np.repeat(np.arange(n), np.diff(ro))
        assert np.all(coef > 0)
        assert np.all(diag >= np.bincount(rows, weights=coef, minlength=n) * (1 - 1e-12))
        if A is not None:
            A[rows, col] = -coef
            A[np.arange(n), np.arange(n)] = diag
            assert np.allclose(A, A.T, rtol=0, atol=1e-14 * diag.max())
        parent = t.mg_array(m - 1, "parent")
        if m == 1:
            act = (t.array("mask") & 64) != 0
            assert np.all(parent[act] >= 0) and np.all(parent[~act] == -1)
            counts = np.bincount(parent[act], minlength=n)
        else:
            counts = np.bincount(parent, minlength=n)
        assert np.all(counts >= 1) and np.all(counts <= 8)
    inv = t.mg_array(t.mg_levels - 1, "coarse_inv")
    n = t.mg_n(t.mg_levels - 1)
    ro, col, coef, diag = (t.mg_array(t.mg_levels - 1, k) for k in ("row_off", "col", "coef", "diag"))
    A = np.zeros((n, n))
    rows = np.repeat(np.arange(n), np.diff(ro))
    A[rows, col] = -coef
    A[np.arange(n), np.arange(n)] = diag
    assert np.allclose(A @ inv.reshape(n, n), np.eye(n), atol=1e-8)


def test_refinement_band(wing):
    """level-1 bricks exist only near the wing, the far field stays at level 0"""
    lev = wing.array("brick_level")
    coord = wing.array("brick_coord").reshape(-1, 3)
    h = wing.array("level_h")
    lo = np.array(wing.info.lo[:])
    centre = lo + (coord + 0.5) * 8 * h[lev][:, None]
    fine = lev == 1
    assert fine.sum() > 0
    # every fine brick centre lies within chord + band + brick diagonal of the wing box
    box_lo, box_hi = np.array([-0.1, -1.1, -0.2]), np.array([1.1, 1.1, 0.2])
    d = np.maximum(0, np.maximum(box_lo - centre[fine], centre[fine] - box_hi))
    assert np.all(np.linalg.norm(d, axis=1) < 0.3 + 2 * 8 * h[0] * math.sqrt(3))


@pytest.fixture(scope="module")
def glider():
    g, d = cases.paraglider(cells_x=96, levels=3)
    return L.Topology(g, d)


def naca_section_area(digits, n=4000):
    """Area of the closed-trailing-edge NACA 4-digit section in chord^2 (thickness integral;
    camber does not change it to first order)."""
    t = (digits % 100) / 100.0
    x = 0.5 * (1 - np.cos(np.linspace(0, math.pi, n)))
    yt = 5 * t * (0.2969 * np.sqrt(x) - 0.1260 * x - 0.3516 * x ** 2 + 0.2843 * x ** 3 - 0.1036 * x ** 4)
    return float(np.trapezoid(2 * yt, x))


def test_paraglider_volume_and_area(glider):
    # elliptic chord c(s) = c0 sqrt(1 - (2s/b)^2) up to |2s/b| = 0.98:
    # volume = A_section * int c^2 ds = A_section * c0^2 b (0.98 - (8/3) 0.49^3)
    b = 8.0
    c0 = 4 * b / (5 * math.pi)
    vol = naca_section_area(2412) * c0 * c0 * b * (0.98 - 8.0 / 3.0 * 0.49 ** 3)
    assert glider.info.body_volume == pytest.approx(vol, rel=0.04)
    assert glider.info.reference_area == pytest.approx(math.pi / 4 * c0 * b)  # = b^2 / 5


def test_paraglider_arc_sweep_and_symmetry(glider):
    x = glider.array("patch_x").reshape(-1, 3)
    S = glider.array("patch_S").reshape(-1, 3)
    b = 8.0
    c0 = 4 * b / (5 * math.pi)
    half = math.radians(25.0)
    R = 0.5 * b / half
    # the wing tips droop to z ~ -R (1 - cos 25 deg) at |y| ~ R sin(0.98 * 25 deg) (pitch 6 deg
    # moves them by less than the tip chord)
    tip_y = R * math.sin(0.98 * half)
    assert np.max(np.abs(x[:, 1])) == pytest.approx(tip_y, abs=0.1)
    tips = np.abs(x[:, 1]) > tip_y - 0.2
    assert np.mean(x[tips, 2]) == pytest.approx(-R * (1 - math.cos(0.98 * half)), abs=0.3)
    # swept: the tips sit downstream of the root leading edge
    assert np.mean(x[tips, 0]) > 0.5 * tip_y * math.tan(math.radians(8.0))
    # mirror symmetry in y: closed surface (sum of wall area vectors ~ 0), equal halves
    assert np.linalg.norm(S.sum(axis=0)) < 1e-6 * np.abs(S).sum()
    left, right = x[:, 1] < 0, x[:, 1] > 0
    assert np.abs(S[left]).sum() == pytest.approx(np.abs(S[right]).sum(), rel=1e-3)
    assert S[left, 1].sum() == pytest.approx(-S[right, 1].sum(), rel=1e-3)


def test_redistribution_groups_follow_the_rule():
    """The groups recomputed from the exported arrays: every active cut cell with a fluid
    fraction below merge_volume_fraction, or an axis whose open area lies in (0, fraction h^2),
    joins the neighbour across its largest compact edge (first in edge order); the groups are
    the connected components of those links, ordered by their lowest cell, members ascending,
    volumes the members' summed volumes."""
    from paracfd_amd import _lib as L, cases
    g, d = cases.naca_wing(digits=2412, alpha_deg=4.0, cells_x=40, levels=3, bands=(1.0, 0.25))
    t = L.Topology(g, d)
    fm = d.merge_volume_fraction
    sd, sv = t.array("special_dof"), t.array("special_volume")
    h = t.array("level_h")[t.array("brick_level")[sd // 512]]
    ea, eb, ax, ar = t.array("edge_a"), t.array("edge_b"), t.array("edge_axis"), t.array("edge_area")
    rank = {int(a): s for s, a in enumerate(sd)}
    inc = [[] for _ in sd]
    for e in range(len(ea)):
        inc[rank[int(ea[e])]].append(e)
        inc[rank[int(eb[e])]].append(e)
    parent = list(range(len(sd)))

    def root(s):
        while parent[s] != s:
            s = parent[s]
        return s
    seeds = []
    for s in range(len(sd)):
        if not sv[s] < h[s] ** 3:
            continue
        area = np.zeros(3)
        best, donor = 0.0, -1
        for e in sorted(inc[s]):
            area[ax[e]] += ar[e]
            if ar[e] > best:
                best, donor = ar[e], rank[int(eb[e] if ea[e] == sd[s] else ea[e])]
        if sv[s] < fm * h[s] ** 3 or np.any((area > 0) & (area < fm * h[s] ** 2)):
            ra, rb = root(s), root(donor)
            parent[max(ra, rb)] = min(ra, rb)
            seeds.append(s)
    grouped = {root(s) for s in seeds}
    comps = {}
    for s in range(len(sd)):
        if root(s) in grouped:
            comps.setdefault(root(s), []).append(s)
    groups = sorted(comps.values(), key=lambda m: m[0])
    assert len(groups) > 0
    off, memb, vol, gid = t.array("redist_off"), t.array("redist_member"), t.array("redist_volume"), t.array("redist_group")
    assert len(off) == len(groups) + 1
    for k, m in enumerate(groups):
        assert list(memb[off[k]:off[k + 1]]) == m
        assert abs(vol[k] - sv[m].sum()) <= 1e-14 * vol[k]
        assert all(gid[s] == k for s in m)
    assert int((gid >= 0).sum()) == sum(len(m) for m in groups)
