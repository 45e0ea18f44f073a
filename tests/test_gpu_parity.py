"""HIP path against the FP64 oracle and the reference golden vectors (needs an MI355X).

Every call goes through the C ABI (libparacfd.so).  Tolerances:
  * multigrid primitives in FP64: 1e-12 relative max-norm
  * FP64 step (fields, pressure, loads): 1e-10 relative L2 (the sphere case's stated bar)
  * FP32 fields with the FP64 outer solve: 1e-4 relative L2 on velocity, 1e-3 on loads
"""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def relmax(a, ref):
    return float(np.max(np.abs(a - ref)) / max(np.max(np.abs(ref)), 1e-300))


def rel_l2(a, ref):
    return float(np.linalg.norm(a - ref) / max(np.linalg.norm(ref), 1e-300))


@pytest.fixture(scope="module")
def lib():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from paracfd_amd import _lib as L
    return L.load()


def make(case, **kw):
    from paracfd_amd import _lib as L, cases
    if case == "sphere":
        g, d = cases.sphere_box(**kw)
    elif case == "naca":
        g, d = cases.naca_wing(**kw)
    elif case == "paraglider":
        g, d = cases.paraglider(**kw)
    else:
        g, d = cases.uniform_box(**kw)
    return L.Topology(g, d)


def opts(**kw):
    from paracfd_amd import default_options
    base = dict(rho=1.0, nu=1e-3, speed=1.0, cfl=0.5, rtol=1e-8, max_iterations=200, pre_smooth=2, post_smooth=2)
    base.update(kw)
    return default_options(**base)


def pair(topo, o):
    from oracle.oracle import Oracle
    from paracfd_amd import ExternalAeroCore
    print(f"[parity] dofs={topo.ndof} active={topo.n_active} edges={topo.info.compact_edges} "
          f"field_precision={o.field_precision} vcycle={o.vcycle_precision}", flush=True)
    gpu = ExternalAeroCore(topo, o)
    orc = Oracle(topo, o)
    gpu.set_reference(topo.info.reference_area)
    orc.set_reference(topo.info.reference_area)
    return gpu, orc


def golden_index_map(topo, n=16):
    """legacy index i + n j + n^2 k -> brick-major dof of the 16^3 box"""
    coord = topo.array("brick_coord").reshape(-1, 3)
    brick_of = {tuple(c): b for b, c in enumerate(coord)}
    idx = np.empty(n ** 3, dtype=np.int64)
    for k in range(n):
        for j in range(n):
            for i in range(n):
                b = brick_of[(i // 8, j // 8, k // 8)]
                idx[i + n * j + n * n * k] = b * 512 + (i % 8) + 8 * (j % 8) + 64 * (k % 8)
    return idx


def test_golden_primitives_on_gpu(lib, oracle_built):
    """The composite-level operator, residual and damped Jacobi kernels reproduce the reference's
    poisson_apply / poisson_residual / jacobi_smooth snapshots (volume-integrated: x h^3)."""
    from oracle.oracle import Legacy
    leg = Legacy()
    fill = leg.fn("legacy_fill")
    fill.argtypes = [C.c_int64, C.c_uint32, C.c_double, C.c_double, C.POINTER(C.c_double)]
    p = np.empty(4096); rhs = np.empty(4096)
    fill(4096, 104, -1.0, 1.0, p.ctypes.data_as(C.POINTER(C.c_double)))
    fill(4096, 105, -1.0, 1.0, rhs.ctypes.data_as(C.POINTER(C.c_double)))
    topo = make("box", n=16, h=0.05, outlet=False)
    idx = golden_index_map(topo)
    from paracfd_amd import ExternalAeroCore
    gpu = ExternalAeroCore(topo, opts())
    h3 = 0.05 ** 3
    x = np.zeros(topo.ndof); x[idx] = p
    b = np.zeros(topo.ndof); b[idx] = rhs * h3
    ap = gpu.mg_primitive(0, 0, x=x)[idx] / h3
    assert relmax(ap, np.fromfile(os.path.join(GOLDEN, "poisson_apply.f64"))) <= 1e-12
    res = gpu.mg_primitive(1, 0, x=x, b=b)[idx] / h3
    assert relmax(res, np.fromfile(os.path.join(GOLDEN, "poisson_residual.f64"))) <= 1e-12
    jac = gpu.mg_primitive(2, 0, x=x, b=b, omega=2.0 / 3.0, sweeps=2)[idx]
    assert relmax(jac, np.fromfile(os.path.join(GOLDEN, "jacobi_smooth.f64"))) <= 1e-12


@pytest.mark.parametrize("case,kw", [("sphere", dict(cells_x=64)), ("sphere", dict(cells_x=32, levels=2, band=0.3))])
def test_multigrid_primitives_match_oracle(lib, oracle_built, case, kw):
    topo = make(case, **kw)
    gpu, orc = pair(topo, opts())
    rng = np.random.default_rng(7)
    mask = topo.array("mask")
    act = (mask & 64) != 0
    for level in range(topo.mg_levels - 1):
        n = topo.mg_n(level)
        x = rng.uniform(-1, 1, n)
        b = rng.uniform(-1, 1, n)
        if level == 0:
            x[~act] = 0.0; b[~act] = 0.0
        for op in (0, 1, 2):
            g = gpu.mg_primitive(op, level, x=x, b=b, omega=2.0 / 3.0, sweeps=3)
            o = orc.mg_primitive(op, level, x=x, b=b, omega=2.0 / 3.0, sweeps=3)
            if level == 0:
                g = g[act]; o = o[act]
            assert relmax(g, o) <= 1e-12, (level, op)
    b = rng.uniform(-1, 1, topo.ndof); b[~act] = 0.0
    g = gpu.mg_primitive(3, 0, b=b)
    o = orc.mg_primitive(3, 0, b=b)
    assert relmax(g[act], o[act]) <= 1e-12


def run_both(gpu, orc, steps):
    gi = gpu.initialize()
    oi = orc.initialize()
    assert gi["converged"] and oi.converged
    assert gi["iterations"] == oi.iterations
    hist = []
    for s in range(steps):
        gs = gpu.step()
        os_ = orc.step()
        hist.append((gs, os_))
        assert gs["converged"] and os_.converged
    return gi, oi, hist


@pytest.mark.parametrize("case,kw,steps", [
    ("sphere", dict(cells_x=64), 20),
    ("sphere", dict(cells_x=32, levels=2, band=0.3), 10),
    ("naca", dict(digits=2412, alpha_deg=4.0, cells_x=24, levels=2, bands=(0.3,)), 5),
    ("paraglider", dict(cells_x=48, levels=2, bands=(0.5,)), 4),
])
def test_fp64_step_matches_oracle(lib, oracle_built, case, kw, steps):
    topo = make(case, **kw)
    gpu, orc = pair(topo, opts())
    gi, oi, hist = run_both(gpu, orc, steps)
    for gs, os_ in hist:
        assert gs["iterations"] == os_.iterations
        assert abs(gs["dt"] - os_.dt) <= 1e-12 * os_.dt
    act = (topo.array("mask") & 64) != 0
    gx, gy, gz = gpu.download_cell_momentum_state()
    ox, oy, oz, op = orc.state()
    for g, o in ((gx, ox), (gy, oy), (gz, oz)):
        assert rel_l2(g[act], o[act]) <= 1e-10
    assert rel_l2(gpu.download_pressure()[act], op[act]) <= 1e-10
    gl, ol = gpu.pressure_loads(), orc.loads()
    for c in range(3):
        assert abs(gl["force"][c] - ol["force"][c]) <= 1e-10 * max(1e-12, np.linalg.norm(ol["force"]))
    gf = gpu.download_faces()
    of = orc.faces()
    for g, o in zip(gf, of):
        if len(o):
            assert rel_l2(g, o) <= 1e-10


def test_single_brick_box_matches_oracle(lib, oracle_built):
    """The smallest topology (one 8^3 brick, no body, outlet at x-max) from a perturbed
    velocity field: every neighbour of the brick is a domain boundary and the hierarchy is one
    coarse level solved exactly."""
    topo = make("box", n=8, h=0.125, outlet=True)
    assert topo.info.bricks == 1
    gpu, orc = pair(topo, opts())
    gpu.initialize()
    orc.initialize()
    rng = np.random.default_rng(7)
    act = ((topo.array("mask") & 64) != 0).astype(float)
    v = [1.0 * act + 0.2 * rng.standard_normal(topo.ndof) * act for _ in range(3)]
    v[1] -= 1.0 * act
    v[2] -= 1.0 * act
    gpu.upload_state(*v)
    orc.set_state(*v)
    gpu.mark_initialized(0.0)
    orc.mark_initialized(0.0)
    for _ in range(3):
        gs, os_ = gpu.step(), orc.step()
        assert gs["converged"] and os_.converged
        assert gs["iterations"] == os_.iterations and gs["iterations"] > 0
        assert abs(gs["dt"] - os_.dt) <= 1e-12 * os_.dt
    a = act > 0
    gx, gy, gz = gpu.download_cell_momentum_state()
    ox, oy, oz, op = orc.state()
    for g, o in ((gx, ox), (gy, oy), (gz, oz)):
        assert rel_l2(g[a], o[a]) <= 1e-10
    assert rel_l2(gpu.download_pressure()[a], op[a]) <= 1e-10


@pytest.mark.parametrize("case,kw", [
    ("sphere", dict(cells_x=32, levels=2, band=0.3)),
    ("naca", dict(digits=2412, alpha_deg=4.0, cells_x=24, levels=2, bands=(0.3,))),
])
def test_fp64_smagorinsky_step_matches_oracle(lib, oracle_built, case, kw):
    """Eddy viscosity (central / one-sided differences on regular cells, least squares on
    special cells), the diffusion substep count from the largest rate and the variable-
    viscosity diffusion, through whole steps, at the FP64 bar."""
    topo = make(case, **kw)
    gpu, orc = pair(topo, opts(smagorinsky_cs=0.17, nu=1e-4))
    gi, oi, hist = run_both(gpu, orc, 6)
    for gs, os_ in hist:
        assert gs["diffusion_substeps"] == os_.diffusion_substeps
        assert gs["iterations"] == os_.iterations
    act = (topo.array("mask") & 64) != 0
    gx, gy, gz = gpu.download_cell_momentum_state()
    ox, oy, oz, op = orc.state()
    for g, o in ((gx, ox), (gy, oy), (gz, oz)):
        assert rel_l2(g[act], o[act]) <= 1e-10
    assert rel_l2(gpu.download_pressure()[act], op[act]) <= 1e-10


def test_fp32_fields_track_oracle(lib, oracle_built):
    from paracfd_amd import _lib as L
    topo = make("sphere", cells_x=64)
    gpu, orc = pair(topo, opts(field_precision=L.FP32))
    gi, oi, hist = run_both(gpu, orc, 10)
    act = (topo.array("mask") & 64) != 0
    gx, _, _ = gpu.download_cell_momentum_state()
    ox, _, _, op = orc.state()
    assert rel_l2(gx[act], ox[act]) <= 1e-4
    gl, ol = gpu.pressure_loads(), orc.loads()
    assert abs(gl["drag"] - ol["drag"]) <= 1e-3 * abs(ol["drag"])


@pytest.mark.parametrize("prec", [0, 1])
def test_fused_graph_path_matches_plain_launches(lib, prec):
    """The fused composite-level passes replayed from a HIP graph give the same result as
    the unfused kernels launched one by one."""
    from paracfd_amd import ExternalAeroCore
    topo = make("sphere", cells_x=32, levels=2, band=0.3)
    res = []
    for fused in (1, 0):
        core = ExternalAeroCore(topo, opts(vcycle_precision=prec, check_interval=4))
        core.set_tuning(fused_vcycle=fused, graph_replay=fused)
        core.initialize()
        its = [core.step()["iterations"] for _ in range(3)]
        res.append((its, core.download_pressure(), core.download_cell_momentum_state()[0]))
        core.close()
    assert res[0][0] == res[1][0]
    assert rel_l2(res[0][1], res[1][1]) <= 1e-13
    assert rel_l2(res[0][2], res[1][2]) <= 1e-13


@pytest.mark.parametrize("prec", [1, 2])
def test_mixed_precision_vcycle_converges(lib, oracle_built, prec):
    """FP32 / BF16 V-cycle under the FP64 outer (flexible) CG reaches rtol 1e-8; the
    projected field agrees with the FP64-everywhere oracle to the solver tolerance."""
    topo = make("sphere", cells_x=64)
    gpu, orc = pair(topo, opts(vcycle_precision=prec))
    gi, oi, hist = run_both(gpu, orc, 5)
    act = (topo.array("mask") & 64) != 0
    gx, _, _ = gpu.download_cell_momentum_state()
    ox, _, _, op = orc.state()
    assert rel_l2(gx[act], ox[act]) <= 1e-6
    assert rel_l2(gpu.download_pressure()[act], op[act]) <= 1e-5


@pytest.mark.parametrize("case,kw", [
    ("sphere", dict(cells_x=32, levels=2, band=0.3)),
    ("naca", dict(digits=2412, alpha_deg=4.0, cells_x=24, levels=2, bands=(0.3,))),
])
def test_fp64_smooth_wall_step_matches_oracle(lib, oracle_built, case, kw):
    """Smooth-wall shear (Spalding friction velocity, implicit 3x3 update per wall cell) in
    every step, and the viscous loads it reports, at the FP64 bar."""
    topo = make(case, **kw)
    gpu, orc = pair(topo, opts(smooth_wall=1, nu=2e-4, speed=2.0))
    before = gpu.pressure_loads()
    assert before["viscous_valid"] == 0 and before["viscous_force"] == [0.0, 0.0, 0.0]
    gi, oi, hist = run_both(gpu, orc, 5)
    for gs, os_ in hist:
        assert gs["iterations"] == os_.iterations
        assert abs(gs["dt"] - os_.dt) <= 1e-12 * os_.dt
    act = (topo.array("mask") & 64) != 0
    gx, gy, gz = gpu.download_cell_momentum_state()
    ox, oy, oz, op = orc.state()
    for g, o in ((gx, ox), (gy, oy), (gz, oz)):
        assert rel_l2(g[act], o[act]) <= 1e-10
    assert rel_l2(gpu.download_pressure()[act], op[act]) <= 1e-10
    gl, ol = gpu.pressure_loads(), orc.loads()
    assert gl["viscous_valid"] == 1 and ol["viscous_valid"] == 1
    assert ol["viscous_drag"] > 0
    for key in ("viscous_force", "viscous_moment", "total_force"):
        ref = np.array(ol[key])
        assert np.linalg.norm(np.array(gl[key]) - ref) <= 1e-10 * np.linalg.norm(ref), key
    assert gl["total_cd"] == pytest.approx(ol["total_cd"], rel=1e-10)


def test_fp32_smooth_wall_tracks_oracle(lib, oracle_built):
    from paracfd_amd import _lib as L
    topo = make("sphere", cells_x=32, levels=2, band=0.3)
    gpu, orc = pair(topo, opts(smooth_wall=1, nu=2e-4, speed=2.0, field_precision=L.FP32, vcycle_precision=L.FP32))
    run_both(gpu, orc, 5)
    act = (topo.array("mask") & 64) != 0
    gx, _, _ = gpu.download_cell_momentum_state()
    ox, _, _, _ = orc.state()
    assert rel_l2(gx[act], ox[act]) <= 1e-4
    gl, ol = gpu.pressure_loads(), orc.loads()
    assert abs(gl["viscous_drag"] - ol["viscous_drag"]) <= 1e-3 * abs(ol["viscous_drag"])


def test_smooth_wall_needs_a_molecular_viscosity(lib):
    from paracfd_amd import ExternalAeroCore, _lib as L
    topo = make("sphere", cells_x=32)
    with pytest.raises(L.PcfdError, match="viscosity"):
        ExternalAeroCore(topo, opts(smooth_wall=1, nu=0.0))
    core = ExternalAeroCore(topo, opts(smooth_wall=0, nu=0.0))   # without the model nu = 0 is allowed
    assert core.initialize()["converged"]
