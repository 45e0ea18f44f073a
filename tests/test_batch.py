"""Batched instances: a merged topology is the concatenation of its parts with every index
offset into the merged numbering (CPU), and a solver on it advances each instance exactly as
that instance's own solver does (GPU, FP64: per-instance iterations, timestep, fields and
loads agree to round-off of the reduction order, which is the same by construction)."""
import numpy as np
import pytest

from paracfd_amd import _lib as L, cases


def wing(alpha, cells_x=24):
    g, d = cases.naca_wing(digits=4412, alpha_deg=alpha, cells_x=cells_x, levels=2, bands=(0.3,))
    return L.Topology(g, d)


@pytest.fixture(scope="module")
def parts():
    return [wing(-4.0), wing(6.8), wing(14.0)]


@pytest.fixture(scope="module")
def merged(parts):
    return L.Topology.merge(parts)


def test_merged_sizes_and_ranges(parts, merged):
    assert merged.instances == 3
    assert merged.info.dofs == sum(p.info.dofs for p in parts)
    assert merged.info.active_cells == sum(p.info.active_cells for p in parts)
    assert merged.info.compact_edges == sum(p.info.compact_edges for p in parts)
    assert merged.info.patches == sum(p.info.patches for p in parts)
    dof0 = brick0 = patch0 = 0
    for i, p in enumerate(parts):
        inst = merged.instance(i)
        assert (inst["dof_begin"], inst["dof_end"]) == (dof0, dof0 + p.info.dofs)
        assert (inst["brick_begin"], inst["brick_end"]) == (brick0, brick0 + p.info.bricks)
        assert (inst["patch_begin"], inst["patch_end"]) == (patch0, patch0 + p.info.patches)
        assert inst["active_cells"] == p.info.active_cells
        assert inst["reference_area"] == pytest.approx(p.info.reference_area)
        dof0 += p.info.dofs
        brick0 += p.info.bricks
        patch0 += p.info.patches
    with pytest.raises(RuntimeError):
        merged.instance(3)


def test_merged_arrays_are_offset_copies(parts, merged):
    offs = np.cumsum([0] + [p.info.dofs for p in parts])
    boffs = np.cumsum([0] + [p.info.bricks for p in parts])
    eoffs = np.cumsum([0] + [p.info.compact_edges for p in parts])
    mask = merged.array("mask")
    nbr = merged.array("brick_nbr").reshape(-1, 6)
    ea, eb = merged.array("edge_a"), merged.array("edge_b")
    inc = merged.array("special_inc")
    sdof = merged.array("special_dof")
    for i, p in enumerate(parts):
        assert np.array_equal(mask[offs[i]:offs[i + 1]], p.array("mask"))
        pn = p.array("brick_nbr").reshape(-1, 6)
        mn = nbr[boffs[i]:boffs[i + 1]]
        assert np.array_equal(mn, np.where(pn >= 0, pn + boffs[i], pn))
        assert np.array_equal(ea[eoffs[i]:eoffs[i + 1]], p.array("edge_a") + offs[i])
        assert np.array_equal(eb[eoffs[i]:eoffs[i + 1]], p.array("edge_b") + offs[i])
    # every compact-edge incidence of a special cell names an edge that touches that cell
    off = merged.array("special_inc_off")
    for s in range(0, len(sdof), max(1, len(sdof) // 500)):
        for k in range(off[s], off[s + 1]):
            raw = inc[k]
            e = raw if raw >= 0 else ~raw
            assert (ea[e] if raw >= 0 else eb[e]) == sdof[s]
    # no structured neighbour or edge crosses instances
    inst_of_brick = np.searchsorted(boffs, np.arange(merged.info.bricks), side="right") - 1
    ok = nbr >= 0
    assert np.all(inst_of_brick[np.where(ok, nbr, 0)][ok] == np.repeat(inst_of_brick, 6).reshape(-1, 6)[ok])
    assert np.all(np.searchsorted(offs, ea, side="right") == np.searchsorted(offs, eb, side="right"))


def test_merged_multigrid_levels(parts, merged):
    depth = max(p.mg_levels for p in parts)
    assert merged.mg_levels == depth
    for m in range(1, depth):
        sizes = [p.mg_n(m) if m < p.mg_levels else 0 for p in parts]
        assert merged.mg_n(m) == sum(sizes)
        par = merged.mg_array(m - 1, "parent")
        assert par.max() < merged.mg_n(m)


def test_merge_rejects_mismatched_levels():
    g1, d1 = cases.naca_wing(digits=4412, alpha_deg=0.0, cells_x=24, levels=2, bands=(0.3,))
    g2, d2 = cases.naca_wing(digits=4412, alpha_deg=0.0, cells_x=32, levels=2, bands=(0.3,))
    a, b = L.Topology(g1, d1), L.Topology(g2, d2)
    with pytest.raises(RuntimeError):
        L.Topology.merge([a, b])


@pytest.mark.gpu
@pytest.mark.parametrize("wall", [0, 1])
def test_batch_matches_individual_runs(parts, merged, wall):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from paracfd_amd import ExternalAeroCore, default_options
    o = default_options(rho=1.0, nu=1e-4, speed=1.0, cfl=0.5, rtol=1e-8, max_iterations=300,
                        field_precision=L.FP64, vcycle_precision=L.FP64, smooth_wall=wall)
    print(f"[batch] {merged.instances} instances, {merged.info.dofs} dofs, smooth wall {wall}", flush=True)
    batch = ExternalAeroCore(merged, o)
    singles = [ExternalAeroCore(p, o) for p in parts]
    for i, p in enumerate(parts):
        batch.set_instance_reference(i, p.info.reference_area)
        singles[i].set_reference(p.info.reference_area)
    batch.initialize()
    for c in singles:
        c.initialize()
    for step in range(4):
        bs = batch.step()
        ss = [c.step() for c in singles]
        assert bs["converged"] == 1
        for i, s in enumerate(ss):
            st = batch.instance_stats(i)
            assert st["iterations"] == s["iterations"], (step, i)
            assert st["dt"] == pytest.approx(s["dt"], rel=1e-13)
            assert st["converged"] == s["converged"] == 1
        assert bs["iterations"] == max(s["iterations"] for s in ss)
    bx, by, bz = batch.download_cell_momentum_state()
    bp = batch.download_pressure()
    for i, (p, c) in enumerate(zip(parts, singles)):
        inst = merged.instance(i)
        sl = slice(inst["dof_begin"], inst["dof_end"])
        x, y, z = c.download_cell_momentum_state()
        for a, b in ((bx[sl], x), (by[sl], y), (bz[sl], z), (bp[sl], c.download_pressure())):
            assert np.max(np.abs(a - b)) <= 1e-12 * max(1.0, np.max(np.abs(b)))
        bl, sl_ = batch.instance_loads(i), c.pressure_loads()
        for k in ("cl", "cd", "cd_viscous", "cl_viscous", "total_cd"):
            assert bl[k] == pytest.approx(sl_[k], rel=1e-11, abs=1e-14)
        assert bl["viscous_valid"] == sl_["viscous_valid"] == wall
        assert (sl_["cd_viscous"] > 0) == bool(wall)
    # all instances' loads in one call are the per-instance integrals
    every = batch.batch_loads()
    assert len(every) == merged.instances
    for i in range(merged.instances):
        one = batch.instance_loads(i)
        for k in ("cl", "cd", "drag", "lift"):
            assert every[i][k] == one[k]


@pytest.fixture(scope="module")
def mixed():
    """A sphere box and a wing on the same cell size (h = 1/16) whose multigrid hierarchies end
    at different levels, so the batch has a dense coarse block above its coarsest level."""
    g, d = cases.sphere_box(cells_x=64, levels=2)
    s = L.Topology(g, d)
    g, d = cases.naca_wing(digits=4412, alpha_deg=6.0, cells_x=144, levels=2, bands=(0.3,))
    w = L.Topology(g, d)
    return [s, w], L.Topology.merge([s, w])


def test_mixed_depth_merge(mixed):
    parts, merged = mixed
    assert parts[0].mg_levels != parts[1].mg_levels
    assert merged.mg_levels == max(p.mg_levels for p in parts)


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(-1, 256, 1), (4, 128, 0)])
def test_mixed_depth_batch_matches_individual_runs(mixed, shape):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from paracfd_amd import ExternalAeroCore, default_options
    parts, merged = mixed
    o = default_options(rho=1.0, nu=1e-4, speed=1.0, cfl=0.5, rtol=1e-8, max_iterations=300,
                        field_precision=L.FP64, vcycle_precision=L.FP64)
    print(f"[batch] mixed depths, launch shape {shape}", flush=True)
    batch = ExternalAeroCore(merged, o)
    batch.set_tuning(xcd_chunk_log2=shape[0], coarse_block=shape[1], fused_vcycle=shape[2])
    singles = [ExternalAeroCore(p, o) for p in parts]
    batch.initialize()
    for c in singles:
        c.initialize()
    for step in range(3):
        batch.step()
        ss = [c.step() for c in singles]
        for i, s in enumerate(ss):
            assert batch.instance_stats(i)["iterations"] == s["iterations"], (step, i)
    bx, _, _ = batch.download_cell_momentum_state()
    bp = batch.download_pressure()
    for i, c in enumerate(singles):
        inst = merged.instance(i)
        sl = slice(inst["dof_begin"], inst["dof_end"])
        x, _, _ = c.download_cell_momentum_state()
        for a, b in ((bx[sl], x), (bp[sl], c.download_pressure())):
            assert np.max(np.abs(a - b)) <= 1e-12 * max(1.0, np.max(np.abs(b)))


@pytest.mark.gpu
def test_batch_with_an_instance_without_body():
    """An empty box (no body: zero right-hand side, no surface patches) batched with a wing:
    the empty instance's projection ends at once with zero pressure and zero loads while the
    wing advances exactly as it does alone."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from paracfd_amd import ExternalAeroCore, default_options
    gw, dw = cases.naca_wing(digits=4412, alpha_deg=4.0, cells_x=24, levels=2, bands=(0.3,))
    w = L.Topology(gw, dw)
    ge = L.GeometryDesc()
    ge.kind = L.GEOM_NONE
    e = L.Topology(ge, dw)
    merged = L.Topology.merge([e, w])
    assert e.info.patches == 0
    o = default_options(rho=1.0, nu=1e-4, speed=1.0, cfl=0.5, rtol=1e-8, max_iterations=300,
                        field_precision=L.FP64, vcycle_precision=L.FP64)
    print(f"[batch] empty instance + wing, {merged.info.dofs} dofs", flush=True)
    batch, single = ExternalAeroCore(merged, o), ExternalAeroCore(w, o)
    batch.set_instance_reference(1, w.info.reference_area)
    single.set_reference(w.info.reference_area)
    batch.initialize()
    single.initialize()
    for _ in range(3):
        bs, ss = batch.step(), single.step()
        assert bs["converged"] == 1
        assert batch.instance_stats(0)["iterations"] == 0
        assert batch.instance_stats(1)["iterations"] == ss["iterations"]
    e0, w0 = merged.instance(0), merged.instance(1)
    bp = batch.download_pressure()
    assert np.all(bp[e0["dof_begin"]:e0["dof_end"]] == 0.0)
    sl = slice(w0["dof_begin"], w0["dof_end"])
    assert np.max(np.abs(bp[sl] - single.download_pressure())) <= 1e-12 * np.max(np.abs(bp[sl]))
    loads = batch.batch_loads()
    assert loads[0]["drag"] == 0.0 and loads[0]["lift"] == 0.0 and loads[0]["coefficients_valid"] == 0
    assert loads[1]["cd"] == pytest.approx(single.pressure_loads()["cd"], rel=1e-11)
