"""Case descriptions: the sphere box and the NACA wings at the sizes the benchmarks use.

Cell-count interpretation (cubic cells, 8^3 bricks): the wing cases keep the stated domain
[-3,6]x[-3,3]x[-2,3.03], chord 1, span 2.03 and refinement bands; the base resolution along x
is chosen so the active-cell counts land near the stated ~4 M (2-level) and ~20 M (3-level).
"""
from __future__ import annotations

import math

from . import _lib as L


def sphere_box(cells_x=128, levels=1, subsamples=4, min_volume_fraction=0.1, outlet=True, band=0.25,
               merge_volume_fraction=0.25):
    """Sphere r=0.25 at (1,1,1) in [0,4]x[0,2]x[0,2]; h = 4/cells_x (128 -> h=1/32, 128x64x64)."""
    g = L.GeometryDesc()
    g.kind = L.GEOM_SPHERE
    g.center[:] = (1.0, 1.0, 1.0)
    g.radius = 0.25
    d = L.GridDesc()
    d.lo[:] = (0.0, 0.0, 0.0)
    d.hi[:] = (4.0, 2.0, 2.0)
    d.base_cells_x = cells_x
    d.levels = levels
    d.refine_band[:] = (band, band * 0.5, band * 0.25, 0.0)
    d.min_volume_fraction = min_volume_fraction
    d.subsamples = subsamples
    d.merge_volume_fraction = merge_volume_fraction
    d.outlet_xmax = 1 if outlet else 0
    return g, d


def naca_wing(digits=2412, alpha_deg=4.0, cells_x=128, levels=2, bands=(0.3, 0.0, 0.0), subsamples=4,
              min_volume_fraction=0.1, chord=1.0, span=2.03, merge_volume_fraction=0.25):
    """Straight NACA 4-digit wing, leading edge at the origin, span along y, pitched about +y."""
    g = L.GeometryDesc()
    g.kind = L.GEOM_NACA4_WING
    g.center[:] = (0.0, 0.0, 0.0)
    g.naca_digits = digits
    g.chord = chord
    g.span = span
    g.alpha_deg = alpha_deg
    d = L.GridDesc()
    d.lo[:] = (-3.0, -3.0, -2.0)
    d.hi[:] = (6.0, 3.0, 3.03)
    d.base_cells_x = cells_x
    d.levels = levels
    d.refine_band[:] = tuple(bands) + (0.0,) * (4 - len(bands))
    d.min_volume_fraction = min_volume_fraction
    d.subsamples = subsamples
    d.merge_volume_fraction = merge_volume_fraction
    d.outlet_xmax = 1
    return g, d


def paraglider(span=8.0, alpha_deg=6.0, anhedral_deg=25.0, sweep_deg=8.0, digits=2412, reflex=0.02, chord=0.0,
               cells_x=160, levels=3, bands=(1.0, 0.25), subsamples=4, min_volume_fraction=0.1,
               merge_volume_fraction=0.25):
    """Synthetic paraglider-like wing: elliptic planform of aspect ratio 5 (chord 0) on a circular
    span arc drooping anhedral_deg at each tip, leading-edge sweep, NACA section with a reflexed
    (upturned) trailing edge, pitched alpha about +y through the root leading edge at the origin.
    Domain: 3 root chords upstream, 6 downstream, 1.5 spans across, from 2.5 chords below the
    tips to 2.5 chords above the root."""
    g = L.GeometryDesc()
    g.kind = L.GEOM_PARAGLIDER
    g.center[:] = (0.0, 0.0, 0.0)
    g.naca_digits = digits
    g.chord = chord
    g.span = span
    g.alpha_deg = alpha_deg
    g.sweep_deg = sweep_deg
    g.anhedral_deg = anhedral_deg
    g.reflex = reflex
    c0 = chord if chord > 0 else 4.0 * span / (5.0 * math.pi)
    half = math.radians(anhedral_deg)
    drop = (0.5 * span / half) * (1 - math.cos(half)) if half > 0 else 0.0
    d = L.GridDesc()
    d.lo[:] = (-3.0 * c0, -0.75 * span, -drop - 2.5 * c0)
    d.hi[:] = (6.0 * c0, 0.75 * span, 2.5 * c0)
    d.base_cells_x = cells_x
    d.levels = levels
    d.refine_band[:] = tuple(b * c0 for b in bands) + (0.0,) * (4 - len(bands))
    d.min_volume_fraction = min_volume_fraction
    d.subsamples = subsamples
    d.merge_volume_fraction = merge_volume_fraction
    d.outlet_xmax = 1
    return g, d


def uniform_box(n=16, h=0.05, outlet=False):
    """Empty n^3 box (legacy MAC golden-vector grid when n=16, h=0.05, no outlet)."""
    g = L.GeometryDesc()
    g.kind = L.GEOM_NONE
    d = L.GridDesc()
    d.lo[:] = (0.0, 0.0, 0.0)
    d.hi[:] = (n * h, n * h, n * h)
    d.base_cells_x = n
    d.levels = 1
    d.min_volume_fraction = 0.1
    d.subsamples = 2
    d.outlet_xmax = 1 if outlet else 0
    return g, d


# Configurations of the benchmark table (cells_x chosen as described in the module docstring).
CASES = {
    "sphere64": dict(kind="sphere", args=dict(cells_x=128), rho=1.0, nu=1e-3, speed=1.0, cfl=0.5, steps=100),
    "naca2412_2lvl": dict(kind="naca", args=dict(digits=2412, alpha_deg=4.0, cells_x=208, levels=2, bands=(0.3,)),
                          rho=1.0, nu=1e-5, speed=1.0, cfl=0.5, steps=500),
    "naca2412_3lvl": dict(kind="naca", args=dict(digits=2412, alpha_deg=4.0, cells_x=288, levels=3, bands=(1.0, 0.25)),
                          rho=1.0, nu=1e-5, speed=1.0, cfl=0.5, steps=200),
    "paraglider_3lvl": dict(kind="paraglider", args=dict(alpha_deg=6.0, cells_x=400, levels=3, bands=(1.0, 0.25)),
                            rho=1.0, nu=1.5e-5, speed=10.0, cfl=0.5, steps=200),
}


# Angle-of-attack sweeps run as one batch of instances (one merged topology).  The 4412 sweep:
# 16 wings at alpha = -4 + 1.2 k deg, 2-level AMR with the 0.3-chord band, cells_x = 200
# (~3.6 M active cells per wing, ~58 M in total).
SWEEPS = {
    "naca4412_sweep16": dict(args=dict(digits=4412, cells_x=200, levels=2, bands=(0.3,)),
                             alphas=tuple(-4.0 + 1.2 * k for k in range(16)),
                             rho=1.0, nu=1e-5, speed=1.0, cfl=0.5, steps=300),
}


def build_sweep(name, device=None, **overrides):
    """Topologies of every instance of a sweep (built one after another, on the host threads or
    with the per-cell stages on HIP device `device`) and the sweep's parameters."""
    spec = dict(SWEEPS[name])
    args = dict(spec["args"])
    args.update(overrides)
    parts = []
    for a in spec["alphas"]:
        g, d = naca_wing(alpha_deg=a, **args)
        parts.append(L.Topology(g, d, device=device))
    return parts, spec


def build_case(name, **overrides):
    spec = dict(CASES[name])
    args = dict(spec["args"])
    args.update(overrides)
    if spec["kind"] == "sphere":
        g, d = sphere_box(**args)
    elif spec["kind"] == "paraglider":
        g, d = paraglider(**args)
    else:
        g, d = naca_wing(**args)
    return g, d, spec
