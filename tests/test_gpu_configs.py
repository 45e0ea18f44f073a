"""The benchmark configurations on the HIP path against the FP64 oracle, at sizes the oracle
steps in seconds (the full-size cases run in bench.py):

  * sphere box 128 x 64 x 64 (h = 1/32), 100 FP64 steps from the freestream start: velocity,
    pressure, faces and drag within 1e-10 relative L2, iteration counts and dt identical
  * NACA 2412 at alpha 4 deg on a 3-level AMR (bands 1.0 and 0.25 chord): the FP64 step at 1e-10;
    the FP32 and BF16 V-cycles under the FP64 outer solve, and FP32 fields with the FP32
    V-cycle, within the mixed-precision bounds (the loads within 0.1 %)
  * a batch of NACA 4412 wings at uneven angles of attack: every instance of the batch against
    the oracle stepping that instance alone, at 1e-10
  * a 3-level paraglider with the learned preconditioner (trained on the device) in flexible
    PCG: the converged fields and loads against the FP64 oracle at the rtol-consistent bound
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel_l2(a, ref):
    return float(np.linalg.norm(a - ref) / max(np.linalg.norm(ref), 1e-300))


@pytest.fixture(scope="module")
def lib():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from paracfd_amd import _lib as L
    return L.load()


def opts(**kw):
    from paracfd_amd import default_options
    base = dict(rho=1.0, nu=1e-3, speed=1.0, cfl=0.5, rtol=1e-8, max_iterations=300, pre_smooth=2, post_smooth=2)
    base.update(kw)
    return default_options(**base)


def pair(topo, o):
    from oracle.oracle import Oracle
    from paracfd_amd import ExternalAeroCore
    gpu = ExternalAeroCore(topo, o)
    orc = Oracle(topo, o)
    gpu.set_reference(topo.info.reference_area)
    orc.set_reference(topo.info.reference_area)
    return gpu, orc


def compare(topo, gpu, orc, bound_v, bound_p, bound_loads):
    act = (topo.array("mask") & 64) != 0
    g = gpu.download_cell_momentum_state()
    o = orc.state()
    for c in range(3):
        assert rel_l2(g[c][act], o[c][act]) <= bound_v, c
    assert rel_l2(gpu.download_pressure()[act], o[3][act]) <= bound_p
    gl, ol = gpu.pressure_loads(), orc.loads()
    fo = np.linalg.norm(ol["force"])
    assert np.linalg.norm(np.array(gl["force"]) - np.array(ol["force"])) <= bound_loads * fo
    for key in ("cl", "cd"):
        assert abs(gl[key] - ol[key]) <= bound_loads * max(abs(ol[key]), 1e-3 * fo / max(1e-30, topo.info.reference_area)), key


def test_sphere_box_128x64x64_100_steps(lib, oracle_built):
    from paracfd_amd import _lib as L, cases
    g, d = cases.sphere_box(cells_x=128)
    topo = L.Topology(g, d)
    print(f"[configs] sphere box {topo.n_active} cells, 100 FP64 steps", flush=True)
    gpu, orc = pair(topo, opts(field_precision=L.FP64, vcycle_precision=L.FP64))
    gi, oi = gpu.initialize(), orc.initialize()
    assert gi["converged"] and oi.converged and gi["iterations"] == oi.iterations
    for s in range(100):
        gs, os_ = gpu.step(), orc.step()
        assert gs["converged"] and os_.converged
        assert gs["iterations"] == os_.iterations, s
        assert abs(gs["dt"] - os_.dt) <= 1e-12 * os_.dt, s
        if s % 25 == 24:
            print(f"[configs] step {s + 1}: {gs['iterations']} iterations, dt {gs['dt']:.4e}", flush=True)
    compare(topo, gpu, orc, 1e-10, 1e-10, 1e-10)
    for a, b in zip(gpu.download_faces(), orc.faces()):
        if len(b):
            assert rel_l2(a, b) <= 1e-10


@pytest.fixture(scope="module")
def wing3(lib):
    from paracfd_amd import _lib as L, cases
    g, d = cases.naca_wing(digits=2412, alpha_deg=4.0, cells_x=40, levels=3, bands=(1.0, 0.25))
    return L.Topology(g, d)


def test_naca2412_3level_fp64_step(wing3, oracle_built):
    from paracfd_amd import _lib as L
    topo = wing3
    assert topo.info.levels == 3
    print(f"[configs] NACA 2412 3-level {topo.n_active} cells, FP64", flush=True)
    gpu, orc = pair(topo, opts(nu=1e-5, field_precision=L.FP64, vcycle_precision=L.FP64))
    gi, oi = gpu.initialize(), orc.initialize()
    assert gi["iterations"] == oi.iterations
    for _ in range(5):
        gs, os_ = gpu.step(), orc.step()
        assert gs["converged"] and os_.converged and gs["iterations"] == os_.iterations
    compare(topo, gpu, orc, 1e-10, 1e-10, 1e-10)


@pytest.mark.parametrize("field,vcycle,bv,bp,bl", [
    (0, 1, 1e-6, 1e-5, 1e-5),     # FP64 fields, FP32 V-cycle, FP64 outer CG
    (0, 2, 1e-6, 1e-5, 1e-5),     # FP64 fields, BF16 V-cycle storage
    (1, 1, 1e-4, 1e-3, 1e-3),     # FP32 fields and V-cycle (the bench's mixed precision)
])
def test_naca2412_3level_mixed_precision(wing3, oracle_built, field, vcycle, bv, bp, bl):
    from paracfd_amd import _lib as L
    topo = wing3
    from oracle.oracle import Oracle
    from paracfd_amd import ExternalAeroCore
    print(f"[configs] NACA 2412 3-level field={field} vcycle={vcycle}", flush=True)
    gpu = ExternalAeroCore(topo, opts(nu=1e-5, field_precision=field, vcycle_precision=vcycle))
    gpu.set_reference(topo.info.reference_area)
    orc = Oracle(topo, opts(nu=1e-5, field_precision=L.FP64, vcycle_precision=L.FP64))
    orc.set_reference(topo.info.reference_area)
    assert gpu.initialize()["converged"] and orc.initialize().converged
    for _ in range(5):
        gs, os_ = gpu.step(), orc.step()
        assert gs["converged"] and os_.converged
        assert gs["relative_residual"] <= 1e-8
    compare(topo, gpu, orc, bv, bp, bl)


def test_naca4412_batch_each_instance_against_the_oracle(lib, oracle_built):
    from oracle.oracle import Oracle
    from paracfd_amd import ExternalAeroCore, _lib as L, cases
    alphas = (-4.0, 3.2, 9.2, 14.0)
    parts = []
    for a in alphas:
        g, d = cases.naca_wing(digits=4412, alpha_deg=a, cells_x=32, levels=2, bands=(0.3,))
        parts.append(L.Topology(g, d))
    merged = L.Topology.merge(parts)
    o = opts(nu=1e-5, field_precision=L.FP64, vcycle_precision=L.FP64)
    print(f"[configs] NACA 4412 batch of {len(parts)}: {[p.n_active for p in parts]} cells", flush=True)
    batch = ExternalAeroCore(merged, o)
    for i, p in enumerate(parts):
        batch.set_instance_reference(i, p.info.reference_area)
    oracles = []
    for p in parts:
        orc = Oracle(p, o)
        orc.set_reference(p.info.reference_area)
        assert orc.initialize().converged
        oracles.append(orc)
    assert batch.initialize()["converged"]
    for _ in range(3):
        batch.step()
        for i, orc in enumerate(oracles):
            os_ = orc.step()
            bs = batch.instance_stats(i)
            assert bs["converged"] and os_.converged
            assert bs["iterations"] == os_.iterations, i
            assert abs(bs["dt"] - os_.dt) <= 1e-12 * os_.dt, i
    gv = batch.download_cell_momentum_state()
    gp = batch.download_pressure()
    loads = batch.batch_loads()
    off = 0
    for i, (p, orc) in enumerate(zip(parts, oracles)):
        n = p.ndof
        act = (p.array("mask") & 64) != 0
        ov = orc.state()
        for c in range(3):
            assert rel_l2(gv[c][off:off + n][act], ov[c][act]) <= 1e-10, (i, c)
        assert rel_l2(gp[off:off + n][act], ov[3][act]) <= 1e-10, i
        ol = orc.loads()
        fo = np.linalg.norm(ol["force"])
        assert np.linalg.norm(np.array(loads[i]["force"]) - np.array(ol["force"])) <= 1e-10 * fo, i
        off += n


def test_paraglider_3level_learned_against_the_oracle(lib, oracle_built):
    from oracle.oracle import Oracle
    from paracfd_amd import ExternalAeroCore, _lib as L, cases
    g, d = cases.paraglider(cells_x=40, levels=3, bands=(1.0, 0.25))
    topo = L.Topology(g, d)
    assert topo.info.levels == 3
    print(f"[configs] paraglider 3-level {topo.n_active} cells, learned preconditioner", flush=True)
    o = opts(nu=1.5e-5, speed=10.0, field_precision=L.FP64, vcycle_precision=L.FP32, preconditioner=L.PC_LEARNED,
             learned_bricks=0, learned_capacity=64)
    gpu = ExternalAeroCore(topo, o)
    gpu.set_reference(topo.info.reference_area)
    o64 = opts(nu=1.5e-5, speed=10.0, field_precision=L.FP64, vcycle_precision=L.FP64)
    orc = Oracle(topo, o64)
    orc.set_reference(topo.info.reference_area)
    assert gpu.initialize()["converged"] and orc.initialize().converged
    gs, os_ = gpu.step(), orc.step()
    for _ in range(2):
        gpu.learned_collect(max_pairs=16)
    tr = gpu.learned_train(steps=200, batch=4, lr=1e-3)
    assert tr["loss_last"] < tr["loss_first"]
    used = 0
    for _ in range(4):
        gs, os_ = gpu.step(), orc.step()
        assert gs["converged"] and os_.converged and gs["relative_residual"] <= 1e-8
        used += gs["learned_used"]
    assert used >= 1   # the network ran inside flexible PCG (before any validation verdict)
    compare(topo, gpu, orc, 1e-6, 1e-5, 1e-5)
