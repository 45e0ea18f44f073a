"""The reference's physics contracts on the HIP path itself, through the C ABI
(tools/physics_contract_probe.cpp:287-326, 366-392; the reference's gates are 2e-3 for the
traction, 1e-6 N for the body force and 2e-5 for the reaction):

  * linear pressure: uniform traction -grad p on every interior control volume, cut cells
    included, and the body force -V_body grad p (FP64 fields: 1e-10)
  * a pressure bump: the fluid impulse equals minus the surface reaction
  * affine velocity: exact face values at every compact face centroid
and the same quantities equal to the FP64 oracle's.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from paracfd_amd import _lib as L
    return L.load()


@pytest.fixture(scope="module")
def wing(lib):
    from paracfd_amd import _lib as L, cases
    g, d = cases.naca_wing(cells_x=48, levels=2, bands=(0.3,))
    return L.Topology(g, d)


def opts(**kw):
    from paracfd_amd import default_options, _lib as L
    base = dict(rho=1.0, nu=1e-3, speed=0.0, cfl=0.5, rtol=1e-8, max_iterations=200, pre_smooth=2, post_smooth=2,
                field_precision=L.FP64, vcycle_precision=L.FP64)
    base.update(kw)
    return default_options(**base)


def cells(t):
    cen = t.array("centroid").reshape(-1, 3)
    h = np.repeat(t.array("level_h")[t.array("brick_level")], 512)
    vol = h ** 3
    vol[t.array("special_dof")] = t.array("special_volume")
    act = (t.array("mask") & 64) != 0
    vol[~act] = 0.0
    return cen, h, vol, act


def test_gpu_linear_pressure_traction_and_body_force(wing, oracle_built):
    from oracle.oracle import Oracle
    from paracfd_amd import ExternalAeroCore
    t = wing
    cen, h, vol, act = cells(t)
    print(f"[contracts] wing {t.n_active} cells, {t.info.compact_edges} compact edges", flush=True)
    lo = np.array(t.info.lo[:])
    hi = lo + np.array(t.info.base_bricks[:]) * 8 * t.info.h0
    inner = act & np.all((cen > lo + 2 * h[:, None]) & (cen < hi - 2 * h[:, None]), axis=1)
    cut = inner & (vol < h ** 3 * (1 - 1e-12))
    G = np.array([0.6, -0.35, 0.8])
    p = (3.0 + cen @ G) * act
    zero = np.zeros(t.ndof)
    core = ExternalAeroCore(t, opts())
    core.upload_state(zero, zero, zero)
    core.upload_pressure(p)
    core.apply_pressure_impulse(1.0)
    v = np.stack(core.download_cell_momentum_state(), axis=1)
    assert cut.sum() > 100
    assert np.max(np.abs(v + G)[inner]) < 1e-10
    # the same impulse as the oracle's
    orc = Oracle(t, opts())
    orc.set_state(zero, zero, zero, p)
    orc.impulse(1.0)
    vo = np.stack(orc.state()[:3], axis=1)
    assert np.max(np.abs(v - vo)[act]) < 1e-12
    # body force
    V = t.info.body_volume
    core.set_reference(t.info.reference_area)
    for Gk in (np.array([1.0, 0, 0]), np.array([0, 0, 1.0])):
        core.upload_pressure((3.0 + cen @ Gk) * act)
        F = np.array(core.pressure_loads(0.0)["force"])
        assert np.max(np.abs(F + V * Gk)) <= 1e-10 * V
    core.close()


def test_gpu_fluid_impulse_equals_surface_reaction(wing, oracle_built):
    from paracfd_amd import ExternalAeroCore
    t = wing
    cen, h, vol, act = cells(t)
    centre, radius = np.array([0.5, 0.0, 0.0]), np.array([1.6, 1.9, 1.3])
    dd = cen - centre
    bump = np.prod(np.maximum(0.0, 1.0 - dd * dd / radius ** 2), axis=1)
    p = 100.0 * bump ** 2 * (1.0 + 0.5 * dd[:, 2] / radius[2]) * act
    zero = np.zeros(t.ndof)
    dt, rho = 1e-3, 1.2
    core = ExternalAeroCore(t, opts(rho=rho))
    core.set_reference(t.info.reference_area)
    core.upload_state(zero, zero, zero)
    core.upload_pressure(p)
    F = np.array(core.pressure_loads(0.0)["force"])
    core.apply_pressure_impulse(dt)
    v = np.stack(core.download_cell_momentum_state(), axis=1)
    reaction = -rho * (vol[:, None] * v).sum(axis=0) / dt
    assert np.linalg.norm(reaction - F) / max(1.0, np.linalg.norm(F)) < 1e-12
    core.close()


def test_gpu_affine_velocity_reconstructs_on_compact_faces(wing, oracle_built):
    from paracfd_amd import ExternalAeroCore
    t = wing
    cen, h, vol, act = cells(t)
    A = np.array([[0.0, 0.3, -0.2], [-0.4, 0.0, 0.1], [0.2, 0.3, 0.0]])
    b0 = np.array([2.0, 0.1, -0.3])
    v = (cen @ A.T + b0) * act[:, None]
    ex = t.array("edge_x").reshape(-1, 3)
    ax = t.array("edge_axis")
    exact = (ex @ A.T + b0)[np.arange(len(ax)), ax]
    zero = np.zeros(t.ndof)
    core = ExternalAeroCore(t, opts(nu=0.0))
    core.upload_state(v[:, 0], v[:, 1], v[:, 2])
    core.upload_faces(zero, zero, zero, np.zeros(t.info.compact_edges))
    core.advect(1e-9)   # no flux moves; the faces are reconstructed at the end
    fe = core.download_faces()[3]
    assert np.max(np.abs(fe - exact)) < 1e-12
    core.close()
