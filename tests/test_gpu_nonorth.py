"""Projection with the non-orthogonal operator on the device (options.nonorthogonal = 1):
flexible GMRES right-preconditioned by the V-cycle, and the face correction with the same
derivative.  Parity against the oracle's full operator (oracle/nonorth.py, itself checked
against the reference's apply_cpu in test_nonorth_topology.py), independent of the Krylov path:
  * the device pressure solves A_full p = rhs (relative residual <= rtol, with A_full and the
    right-hand side -(rho/dt) integrated divergence assembled on the host from the faces the
    device projected)
  * it equals scipy's sparse direct solution of that system
  * the corrected faces satisfy  div(after) = div(before) + (dt/rho) A_full p  exactly (the face
    correction carries the operator's derivative), so the divergence after projection is at the
    solver tolerance; the same identity with the orthogonal operator holds for the PCG path
  * the correction changes the pressure (the orthogonal solution differs)
  * an FP32 V-cycle (a variable preconditioner) converges to the same pressure
  * batches and the learned preconditioner are refused with this option"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import nonorth as NO
from oracle.oracle import Oracle
from paracfd_amd import _lib as L, cases

pytestmark = pytest.mark.gpu


def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _topo(name):
    if name == "sphere":
        g, d = cases.sphere_box(cells_x=32, levels=2, band=0.3)
    else:
        g, d = cases.naca_wing(digits=2412, alpha_deg=4.0, cells_x=24, levels=2, bands=(0.3,))
    return L.Topology(g, d)


def _run(topo, nonorth, vcycle=L.FP64, rtol=1e-10, faces=None, dt=None):
    from paracfd_amd import ExternalAeroCore, default_options
    o = default_options(rho=1.2, nu=1e-3, speed=2.0, cfl=0.5, rtol=rtol, max_iterations=400,
                        field_precision=L.FP64, vcycle_precision=vcycle, nonorthogonal=nonorth)
    core = ExternalAeroCore(topo, o)
    core.initialize()
    core.step()
    used = core.advect()
    if faces is not None:
        core.upload_faces(*faces)
    else:
        dt = used
    f0 = core.download_faces()
    st = core.project(dt, warm=False)
    return core, dt, f0, st, core.download_pressure(), core.download_faces()


def _div(topo, faces):
    o = L.Options()
    o.rho, o.nu, o.speed, o.cfl, o.rtol, o.max_iterations = 1.2, 1e-3, 2.0, 0.5, 1e-10, 10
    orc = Oracle(topo, o)
    orc.set_faces(*faces)
    return orc.divergence()


@pytest.mark.parametrize("case", ["sphere", "wing"])
def test_full_operator_projection_matches_the_oracle(case):
    _gpu()
    topo = _topo(case)
    print(f"[nonorth] {case} fp64", flush=True)
    core, dt, f0, st, p, f1 = _run(topo, 1)
    assert st["converged"] and st["iterations"] > 0
    topo.build_nonorthogonal()
    A = NO.full_matrix(topo)
    act = (topo.array("mask") & 64) != 0
    s = dt / 1.2
    d0, d1 = _div(topo, f0), _div(topo, f1)
    rhs = np.where(act, -d0 / s, 0.0)
    r = (rhs - A @ p)[act]
    rel = np.linalg.norm(r) / np.linalg.norm(rhs[act])
    assert rel <= 2e-10, rel
    assert st["relative_residual"] <= 1e-10
    # direct solution of the same system
    idx = np.flatnonzero(act)
    Aa = A[idx][:, idx].tocsc()
    pd = spla.spsolve(Aa, rhs[idx])
    assert np.linalg.norm(p[idx] - pd) <= 1e-7 * np.linalg.norm(pd)
    # the face correction carries the full operator's derivative
    lhs = d1 - d0
    assert np.abs((lhs - s * (A @ p))[act]).max() <= 1e-10 * np.abs(d0[act]).max()
    assert np.abs(d1[act]).max() <= 1e-7 * np.abs(d0[act]).max()
    # the correction matters: the orthogonal projection of the same state differs
    _, _, f0o, sto, po, f1o = _run(topo, 0, faces=f0, dt=dt)
    assert all(np.array_equal(a, b) for a, b in zip(f0o, f0))
    assert np.linalg.norm(po[idx] - p[idx]) > 1e-4 * np.linalg.norm(p[idx])
    A0 = NO.orthogonal_matrix(topo)
    assert np.abs((_div(topo, f1o) - d0 - s * (A0 @ po))[act]).max() <= 1e-10 * np.abs(d0[act]).max()


def test_fp32_vcycle_converges_to_the_same_pressure():
    _gpu()
    topo = _topo("wing")
    print("[nonorth] wing fp64 vs fp32 V-cycle", flush=True)
    *_, st64, p64, _ = _run(topo, 1)
    *_, st32, p32, _ = _run(topo, 1, vcycle=L.FP32)
    assert st64["converged"] and st32["converged"]
    act = (topo.array("mask") & 64) != 0
    assert np.linalg.norm((p32 - p64)[act]) <= 1e-8 * np.linalg.norm(p64[act])


def test_refused_combinations():
    _gpu()
    from paracfd_amd import ExternalAeroCore, default_options
    parts = [_topo("wing"), _topo("wing")]
    merged = L.Topology.merge(parts)
    with pytest.raises(L.PcfdError):
        ExternalAeroCore(merged, default_options(nonorthogonal=1))
    with pytest.raises(L.PcfdError):
        ExternalAeroCore(parts[0], default_options(nonorthogonal=1, preconditioner=L.PC_LEARNED))


def test_captured_and_direct_fgmres_agree_bitwise():
    """The per-basis-vector graphs replay the same kernels with the same arguments."""
    _gpu()
    from paracfd_amd import ExternalAeroCore, default_options
    topo = _topo("wing")
    out = []
    for graph in (1, 0):
        print(f"[nonorth] wing graph_replay={graph}", flush=True)
        o = default_options(rho=1.2, nu=1e-3, speed=2.0, cfl=0.5, rtol=1e-8, max_iterations=400, nonorthogonal=1)
        core = ExternalAeroCore(topo, o)
        core.set_tuning(graph_replay=graph)
        core.initialize()
        sts = [core.step() for _ in range(3)]
        assert all(s["converged"] for s in sts)
        out.append((core.download_pressure(), [s["iterations"] for s in sts]))
    assert out[0][1] == out[1][1]
    assert np.array_equal(out[0][0], out[1][0])


def test_restarted_fgmres_cycles():
    """gmres_restart = 3: the projection needs several cycles (the restart rebuilds the first basis
    vector from the true residual and reuses the per-basis-index graphs); it still solves the
    oracle's full system to the tolerance, equals the direct solution, and the captured and
    directly launched solves agree bitwise."""
    _gpu()
    from paracfd_amd import ExternalAeroCore, default_options
    topo = _topo("wing")
    out = []
    for graph in (1, 0):
        print(f"[nonorth] wing restart=3 graph_replay={graph}", flush=True)
        o = default_options(rho=1.2, nu=1e-3, speed=2.0, cfl=0.5, rtol=1e-10, max_iterations=400, nonorthogonal=1,
                            gmres_restart=3, field_precision=L.FP64, vcycle_precision=L.FP64)
        core = ExternalAeroCore(topo, o)
        core.set_tuning(graph_replay=graph)
        core.initialize()
        core.step()
        dt = core.advect()
        f0 = core.download_faces()
        st = core.project(dt, warm=False)
        assert st["converged"] and st["iterations"] > 3, st["iterations"]   # more than one cycle
        out.append((core.download_pressure(), st["iterations"], f0, dt))
    assert out[0][1] == out[1][1]
    assert np.array_equal(out[0][0], out[1][0])
    p, _, f0, dt = out[0]
    topo.build_nonorthogonal()
    A = NO.full_matrix(topo)
    act = (topo.array("mask") & 64) != 0
    rhs = np.where(act, -_div(topo, f0) / (dt / 1.2), 0.0)
    idx = np.flatnonzero(act)
    assert np.linalg.norm((rhs - A @ p)[act]) <= 2e-10 * np.linalg.norm(rhs[act])
    pd = spla.spsolve(A[idx][:, idx].tocsc(), rhs[idx])
    assert np.linalg.norm(p[idx] - pd) <= 1e-7 * np.linalg.norm(pd)


def test_gcro_recycling_across_steps():
    """GCRO recycling (gmres_recycle = 3, the reference's capacity) against plain flexible GMRES (0, the
    default) over six
    steps of the wing with short cycles (gmres_restart = 8): every projection converges, the
    recycled solves need no more iterations in total, the last projection's pressure solves the
    oracle's full system to the tolerance and equals the direct solution, and the two runs'
    velocities agree to the solver tolerance."""
    _gpu()
    from paracfd_amd import ExternalAeroCore, default_options
    topo = _topo("wing")
    runs = {}
    for rec in (3, 0):
        print(f"[nonorth] wing gmres_recycle={rec}", flush=True)
        o = default_options(rho=1.2, nu=1e-3, speed=2.0, cfl=0.5, rtol=1e-10, max_iterations=400, nonorthogonal=1,
                            gmres_restart=8, gmres_recycle=rec, field_precision=L.FP64, vcycle_precision=L.FP64)
        core = ExternalAeroCore(topo, o)
        assert core.initialize()["converged"]
        its = []
        for _ in range(6):
            st = core.step()
            assert st["converged"]
            its.append(st["iterations"])
        dt = core.advect()
        f0 = core.download_faces()
        st = core.project(dt, warm=True)
        assert st["converged"]
        its.append(st["iterations"])
        runs[rec] = (its, core.download_pressure(), f0, dt, np.concatenate(core.download_cell_momentum_state()))
        core.close()
    print(f"[nonorth] iterations recycled {runs[3][0]} plain {runs[0][0]}", flush=True)
    assert sum(runs[3][0]) <= sum(runs[0][0])
    _, p, f0, dt, v3 = runs[3]
    topo.build_nonorthogonal()
    A = NO.full_matrix(topo)
    act = (topo.array("mask") & 64) != 0
    rhs = np.where(act, -_div(topo, f0) / (dt / 1.2), 0.0)
    idx = np.flatnonzero(act)
    assert np.linalg.norm((rhs - A @ p)[act]) <= 2e-10 * np.linalg.norm(rhs[act])
    pd = spla.spsolve(A[idx][:, idx].tocsc(), rhs[idx])
    assert np.linalg.norm(p[idx] - pd) <= 1e-7 * np.linalg.norm(pd)
    v0 = runs[0][4]
    assert np.linalg.norm(v3 - v0) <= 1e-7 * np.linalg.norm(v0)
