"""Slice sampling for a viewer (pcfd_sample_slice).

CPU: the numpy oracle (oracle/slice.py) is pinned by closed forms -- with the pressure set to
the x coordinate of each cell centre, every pixel lies within half a cell of the value it
samples, and within half a *finest* cell inside the refined band around the body (the finest
leaf wins); pixels inside the solid sphere and outside the grid are NaN; |v| of a uniform field
is its speed.  A planted mistake (sampling the coarsest covering level first) fails the band check.
GPU: the device sampler equals the oracle on the same state (pressure and components exactly,
|v| to one float ulp, the device may fuse the sum of squares) on a 2-level sphere, a wing and
one instance of a batch."""
import numpy as np
import pytest

from oracle.slice import sample_slice
from paracfd_amd import _lib as L, cases


@pytest.fixture(scope="module")
def sphere():
    g, d = cases.sphere_box(cells_x=32, levels=2, band=0.3)
    return L.Topology(g, d)


def cell_centres(topo):
    coord = topo.array("brick_coord").reshape(-1, 3)
    lev = topo.array("brick_level")
    h = topo.array("level_h")
    q = np.arange(512)
    loc = np.stack([q & 7, (q >> 3) & 7, q >> 6], axis=1)
    lo = np.array([topo.info.lo[d] for d in range(3)])
    c = (8 * coord[:, None, :] + loc[None, :, :] + 0.5) * h[lev][:, None, None] + lo
    return c.reshape(-1, 3), np.repeat(h[lev], 512)


def test_pixels_sample_the_finest_leaf(sphere):
    topo = sphere
    cen, hc = cell_centres(topo)
    n = topo.ndof
    z = np.zeros(n)
    p = cen[:, 0].copy()
    lo, hi, nu, nv = (0.0, 0.0), (4.0, 2.0), 400, 200
    img = sample_slice(topo, (z, z, z), p, 4, 2, 1.0 + 1e-3, lo, hi, nu, nv)
    px = lo[0] + (np.arange(nu) + 0.5) * (hi[0] - lo[0]) / nu
    py = lo[1] + (np.arange(nv) + 0.5) * (hi[1] - lo[1]) / nv
    X, Y = np.meshgrid(px, py)
    ok = np.isfinite(img)
    hmin = float(topo.array("level_h").min())
    hmax = float(topo.array("level_h").max())
    err = np.abs(img[ok] - X[ok])
    assert err.max() <= 0.5 * hmax + 1e-6
    # inside the refined band (within 0.2 of the sphere surface) the finest cells answer
    r = np.sqrt((X - 1.0) ** 2 + (Y - 1.0) ** 2 + 1e-6)
    band = ok & (r > 0.27) & (r < 0.4)
    assert band.sum() > 100
    assert np.abs(img[band] - X[band]).max() <= 0.5 * hmin + 1e-6
    # the solid sphere is NaN, the far field is not
    assert np.isnan(img[(r < 0.2)]).all()
    assert ok[(r > 0.5)].mean() > 0.99


def test_planted_coarse_first_lookup_fails_the_band_check(sphere):
    topo = sphere
    cen, _ = cell_centres(topo)
    n = topo.ndof
    z = np.zeros(n)
    p = cen[:, 0].copy()
    # a level-0-only lookup (as if the finest leaf were not searched first) cannot resolve the band
    import oracle.slice as S
    maps = S.brick_maps(topo)
    assert len(maps) == 2 and (maps[1] >= 0).sum() > 0
    orig = S.brick_maps
    try:
        S.brick_maps = lambda t, i=0: [orig(t, i)[0], np.full_like(orig(t, i)[1], -1)]
        img = S.sample_slice(topo, (z, z, z), p, 4, 2, 1.0 + 1e-3, (0.0, 0.0), (4.0, 2.0), 400, 200)
    finally:
        S.brick_maps = orig
    px = (np.arange(400) + 0.5) * 4.0 / 400
    X = np.broadcast_to(px, (200, 400))
    hmin = float(topo.array("level_h").min())
    ok = np.isfinite(img)
    assert np.abs(img[ok] - X[ok]).max() > 0.5 * hmin + 1e-6


def test_outside_the_grid_is_nan_and_speed_of_uniform_flow(sphere):
    topo = sphere
    n = topo.ndof
    act = (topo.array("mask") & 64) != 0
    vx, vy, vz = np.where(act, 3.0, 0.0), np.where(act, 4.0, 0.0), np.zeros(n)
    img = sample_slice(topo, (vx, vy, vz), np.zeros(n), 0, 1, 1.0, (-1.0, -1.0), (5.0, 3.0), 60, 40)
    ok = np.isfinite(img)
    assert np.all(img[ok] == np.float32(5.0))
    assert np.isnan(img[:, 0]).all() and np.isnan(img[0, :]).all()   # x < 0 or z < 0: outside


def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["sphere", "wing"])
def test_device_slice_matches_oracle(case):
    _gpu()
    from paracfd_amd import ExternalAeroCore, default_options
    if case == "sphere":
        g, d = cases.sphere_box(cells_x=32, levels=2, band=0.3)
        planes = [(2, 1.0), (1, 1.03), (0, 1.2)]
    else:
        g, d = cases.naca_wing(digits=2412, alpha_deg=4.0, cells_x=24, levels=2, bands=(0.3,))
        planes = [(1, 1.0), (2, 0.05), (0, 0.5)]
    topo = L.Topology(g, d)
    o = default_options(rho=1.2, nu=1e-3, speed=2.0, cfl=0.5, rtol=1e-8, max_iterations=200)
    core = ExternalAeroCore(topo, o)
    core.initialize()
    for _ in range(2):
        core.step()
    v = core.download_cell_momentum_state()
    p = core.download_pressure()
    info = topo.info
    q = 0.5 * 1.2 * 2.0 ** 2
    for axis, c in planes:
        a1, a2 = [d for d in range(3) if d != axis]
        lo = (info.lo[a1], info.lo[a2])
        hi = (info.lo[a1] + 8 * info.base_bricks[a1] * info.h0, info.lo[a2] + 8 * info.base_bricks[a2] * info.h0)
        for field in range(6):
            print(f"[slice] {case} axis={axis} field={field}", flush=True)
            got = core.sample_slice(field, axis, c, lo, hi, 173, 91)
            ref = sample_slice(topo, v, p, field, axis, c, lo, hi, 173, 91, qinf=q)
            assert np.array_equal(np.isnan(got), np.isnan(ref))
            ok = np.isfinite(ref)
            assert ok.sum() > 1000
            if field == 0:
                assert np.allclose(got[ok], ref[ok], rtol=2.5e-7, atol=0)
            else:
                assert np.array_equal(got[ok], ref[ok])


@pytest.mark.gpu
def test_device_slice_of_a_batch_instance():
    _gpu()
    from paracfd_amd import ExternalAeroCore, default_options
    parts = []
    for alpha in (-2.0, 8.0):
        g, d = cases.naca_wing(digits=4412, alpha_deg=alpha, cells_x=24, levels=2, bands=(0.3,))
        parts.append(L.Topology(g, d))
    merged = L.Topology.merge(parts)
    o = default_options(rho=1.0, nu=1e-3, speed=1.0, cfl=0.5, rtol=1e-8, max_iterations=200)
    core = ExternalAeroCore(merged, o)
    core.initialize()
    core.step()
    v = core.download_cell_momentum_state()
    p = core.download_pressure()
    info = merged.info
    lo = (info.lo[0], info.lo[2])
    hi = (info.lo[0] + 8 * info.base_bricks[0] * info.h0, info.lo[2] + 8 * info.base_bricks[2] * info.h0)
    for inst in range(2):
        got = core.sample_slice("u", 1, 1.0, lo, hi, 120, 80, instance=inst)
        ref = sample_slice(merged, v, p, 1, 1, 1.0, lo, hi, 120, 80, instance=inst)
        assert np.array_equal(np.isnan(got), np.isnan(ref))
        assert np.array_equal(got[np.isfinite(ref)], ref[np.isfinite(ref)])
    a = core.sample_slice("u", 1, 1.0, lo, hi, 120, 80, instance=0)
    b = core.sample_slice("u", 1, 1.0, lo, hi, 120, 80, instance=1)
    assert not np.array_equal(a, b)   # two angles of attack: each instance answers with its own field


@pytest.mark.gpu
def test_device_slice_rejects_bad_requests():
    _gpu()
    from paracfd_amd import ExternalAeroCore, default_options
    g, d = cases.uniform_box(n=8, h=0.125, outlet=True)
    topo = L.Topology(g, d)
    core = ExternalAeroCore(topo, default_options())
    core.initialize()
    good = core.sample_slice("p", 2, 0.5, (0.0, 0.0), (1.0, 1.0), 8, 8)
    assert good.shape == (8, 8) and np.isfinite(good).all()
    for kw in (dict(field=6), dict(axis=3), dict(lo=(1.0, 0.0)), dict(nu=0), dict(instance=1), dict(coord=np.nan)):
        args = dict(field=4, axis=2, coord=0.5, lo=(0.0, 0.0), hi=(1.0, 1.0), nu=8, nv=8, instance=0)
        args.update(kw)
        with pytest.raises(L.PcfdError):
            core.sample_slice(**args)
    # a window reaching outside the grid gives NaN there
    img = core.sample_slice("u", 2, 0.5, (-1.0, 0.0), (1.0, 1.0), 16, 4)
    assert np.isnan(img[:, :8]).all() and np.isfinite(img[:, 8:]).all()
