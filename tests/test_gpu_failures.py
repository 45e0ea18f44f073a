"""Failure paths of the C ABI on the device, each forced from the host (no kernel faults):

  * a non-finite velocity before a step: PCFD_ENONFINITE, nothing advances
    (external_aero_core.cpp:308)
  * step before a converged initialization: PCFD_ESTATE (:306)
  * a projection that does not reach the tolerance: a status, not an error; physical time does
    not advance (:327)
  * a device-memory budget below what the solver needs: PCFD_ENOMEM before any allocation
    beyond it (:91-95), and a budget just above it succeeds
  * bad options: PCFD_EINVAL
The oracle implements the same statuses for the same conditions (compared where it has them).
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
def sphere(lib):
    from paracfd_amd import _lib as L, cases
    g, d = cases.sphere_box(cells_x=32)
    return L.Topology(g, d)


def opts(**kw):
    from paracfd_amd import default_options
    base = dict(rho=1.0, nu=1e-3, speed=1.0, cfl=0.5, rtol=1e-8, max_iterations=200)
    base.update(kw)
    return default_options(**base)


def status_of(fn):
    from paracfd_amd import _lib as L
    try:
        fn()
    except L.PcfdError as e:
        return e.status, str(e)
    return L.PCFD_OK, ""


def test_nonfinite_velocity_is_refused_before_the_step(sphere, oracle_built):
    from oracle.oracle import Oracle
    from paracfd_amd import ExternalAeroCore, _lib as L
    t = sphere
    core = ExternalAeroCore(t, opts())
    st = core.initialize()
    assert st["converged"]
    x, y, z = core.download_cell_momentum_state()
    p0 = core.download_pressure()
    act = np.flatnonzero((t.array("mask") & 64) != 0)
    x = x.copy()
    x[act[len(act) // 2]] = np.nan
    core.upload_state(x, y, z)
    status, msg = status_of(core.step)
    assert status == L.PCFD_ENONFINITE and "non-finite" in msg
    # nothing advanced: the pressure and the other components are as they were
    assert np.array_equal(core.download_pressure(), p0)
    assert np.array_equal(core.download_cell_momentum_state()[1], y)
    core.close()
    # the oracle refuses the same state with the same status
    orc = Oracle(t, opts())
    assert orc.initialize().converged
    orc.set_state(x, y, z, None)
    with pytest.raises(RuntimeError, match="status 4"):
        orc.step()


def test_step_before_initialize_is_a_state_error(sphere):
    from paracfd_amd import ExternalAeroCore, _lib as L
    core = ExternalAeroCore(sphere, opts())
    status, msg = status_of(core.step)
    assert status == L.PCFD_ESTATE and "initialization" in msg
    core.close()


def test_unconverged_projection_is_a_status_and_time_stays(sphere):
    from paracfd_amd import ExternalAeroCore, _lib as L
    core = ExternalAeroCore(sphere, opts(max_iterations=1))
    st = core.initialize()
    assert not st["converged"] and st["iterations"] <= 1
    # an unconverged initialization leaves the solver uninitialized
    assert status_of(core.step)[0] == L.PCFD_ESTATE
    core.mark_initialized(0.25)   # resume from a checkpoint at t = 0.25
    st = core.step()              # returns normally
    assert not st["converged"]
    assert st["relative_residual"] > 1e-8
    assert st["physical_time"] == 0.25
    st = core.step()
    assert not st["converged"] and st["physical_time"] == 0.25
    core.close()


def test_memory_budget(sphere):
    from paracfd_amd import ExternalAeroCore, _lib as L
    full = ExternalAeroCore(sphere, opts())
    need = full.device_bytes()
    full.close()
    status, msg = status_of(lambda: ExternalAeroCore(sphere, opts(memory_budget_bytes=float(need // 2))))
    assert status == L.PCFD_ENOMEM and "budget" in msg
    # the failed creation released what it had allocated: a solver within budget still fits
    ok = ExternalAeroCore(sphere, opts(memory_budget_bytes=float(need + (1 << 20))))
    assert ok.device_bytes() == need
    st = ok.initialize()
    assert st["converged"]
    ok.close()


def test_bad_options_are_invalid_arguments(sphere):
    from paracfd_amd import ExternalAeroCore, _lib as L
    for kw in (dict(rho=-1.0), dict(field_precision=L.BF16), dict(preconditioner=7), dict(pre_smooth=0),
               dict(smoother_weight=2.0, smoother_weight2=2.0), dict(smagorinsky_cs=-0.1)):
        status, msg = status_of(lambda: ExternalAeroCore(sphere, opts(**kw)))
        assert status == L.PCFD_EINVAL, (kw, status, msg)
