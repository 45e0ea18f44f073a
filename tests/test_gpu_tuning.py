"""Launch-shape tuning: every configuration computes the same arithmetic, so results are
bit-identical across configurations, and the autotuner leaves the solution where it was."""
import numpy as np
import pytest

from paracfd_amd import _lib as L, cases

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def topo():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    g, d = cases.naca_wing(digits=2412, alpha_deg=4.0, cells_x=48, levels=2, bands=(0.3,))
    return L.Topology(g, d)


def options(field, vcycle):
    from paracfd_amd import default_options
    return default_options(rho=1.0, nu=1e-5, speed=1.0, cfl=0.5, rtol=1e-8, max_iterations=300,
                           field_precision=field, vcycle_precision=vcycle)


def state(core):
    return np.concatenate(core.download_cell_momentum_state() + (core.download_pressure(),))


@pytest.mark.parametrize("field,vcycle", [(L.FP64, L.FP64), (L.FP32, L.FP32), (L.FP32, L.BF16)])
def test_tunings_give_identical_results(topo, field, vcycle):
    from paracfd_amd import ExternalAeroCore
    print(f"[tuning] identical results field={field} vcycle={vcycle}", flush=True)
    cores = [ExternalAeroCore(topo, options(field, vcycle)) for _ in range(4)]
    cores[1].set_tuning(xcd_chunk_log2=2, coarse_block=64, fused_vcycle=0)
    cores[2].set_tuning(xcd_chunk_log2=0, coarse_block=512, fused_vcycle=1)
    cores[3].set_tuning(xcd_chunk_log2=4, coarse_block=128, fused_vcycle=0)
    assert cores[1].tuning()["coarse_block"] == 64
    assert cores[3].tuning()["fused_vcycle"] == 0
    its = [[], [], [], []]
    for i, c in enumerate(cores):
        c.initialize()
        for _ in range(3):
            its[i].append(c.step()["iterations"])
    assert its[0] == its[1] == its[2] == its[3]
    ref = state(cores[0])
    for c in cores[1:]:
        assert np.array_equal(state(c), ref)


def test_autotune_keeps_the_solution(topo):
    from paracfd_amd import ExternalAeroCore
    print("[tuning] autotune", flush=True)
    tuned, plain = (ExternalAeroCore(topo, options(L.FP32, L.FP32)) for _ in range(2))
    for c in (tuned, plain):
        c.initialize()
        c.step()
    before = state(tuned)
    res = tuned.autotune(repetitions=1)
    assert res["candidates"] >= 8
    assert res["xcd_chunk_log2"] in (-1, 0, 2, 4, 6)
    assert res["coarse_block"] in (64, 128, 256, 512)
    assert 0 < res["step_ms"] <= res["default_step_ms"]
    assert tuned.tuning() == res
    assert np.array_equal(state(tuned), before)
    for _ in range(2):
        a, b = tuned.step(), plain.step()
        assert a["iterations"] == b["iterations"]
        assert a["physical_time"] == b["physical_time"]
    assert np.array_equal(state(tuned), state(plain))
    with pytest.raises(RuntimeError):
        tuned.set_tuning(coarse_block=100)


@pytest.mark.parametrize("field,vcycle", [(L.FP64, L.FP64), (L.FP32, L.FP32)])
def test_kernel_timing_hooks_leave_the_state(topo, field, vcycle):
    """The measurement launches (bench roofline) touch only scratch vectors: the state is unchanged
    and the next steps match a solver that was never measured, bit for bit."""
    from paracfd_amd import ExternalAeroCore
    print(f"[tuning] timing hooks field={field} vcycle={vcycle}", flush=True)
    timed, plain = (ExternalAeroCore(topo, options(field, vcycle)) for _ in range(2))
    for c in (timed, plain):
        c.initialize()
        c.step()
    before = state(timed)
    for k in (0, 1, 2, 3, 5, 6):
        ms, by = timed.time_kernel(k, reps=2)
        assert ms > 0 and (by > 0 or k == 3)
    assert np.array_equal(state(timed), before)
    for _ in range(2):
        a, b = timed.step(), plain.step()
        assert a["iterations"] == b["iterations"]
    assert np.array_equal(state(timed), state(plain))
