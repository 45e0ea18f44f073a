"""The oracle's composite pressure operator against the reference's own code.

oracle/_ref/libref.so is built (oracle/Makefile, target ref) from the reference's
amr_pressure.cpp / amr_grid.cpp / triangle_bvh.cpp with a driver of this project that maps the
shared topology onto the reference's structures (oracle/ref_apply.cpp).  Its
CompositeAmrPressureSystem::apply_cpu and diagonal_cpu (amr_pressure.cpp:498-543) then
evaluate, on cut-cell AMR topologies, the operator that the oracle and the GPU kernels apply:
structured faces h (p_a - p_b), compact edges (A/d)(p_a - p_b), the outlet 2h p.  The library
exists only where the reference sources are present; without it the test is skipped."""
import ctypes as C
import os

import numpy as np
import pytest

from paracfd_amd import _lib as L, cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "_ref", "libref.so")


@pytest.fixture(scope="module")
def ref():
    if not os.path.exists(LIB):
        pytest.skip("reference operator library not built (needs the reference sources)")
    lib = C.CDLL(LIB)
    lib.pcfdref_operator.restype = C.c_int
    return lib


def reference_operator(lib, topo, p):
    lev = topo.array("brick_level")
    nlev = int(lev.max()) + 1
    first = np.searchsorted(lev, np.arange(nlev + 1)).astype(np.int32)
    assert np.all(np.diff(lev) >= 0), "bricks are stored level by level"
    h = np.ascontiguousarray(topo.array("level_h")[:nlev], dtype=np.float64)
    arrs = dict(nbr=np.ascontiguousarray(topo.array("brick_nbr"), dtype=np.int32),
                flags=np.ascontiguousarray(topo.array("brick_flags"), dtype=np.uint8),
                mask=np.ascontiguousarray(topo.array("mask"), dtype=np.uint8),
                ea=np.ascontiguousarray(topo.array("edge_a"), dtype=np.int32),
                eb=np.ascontiguousarray(topo.array("edge_b"), dtype=np.int32),
                area=np.ascontiguousarray(topo.array("edge_area"), dtype=np.float64),
                coef=np.ascontiguousarray(topo.array("edge_coef"), dtype=np.float64))
    p = np.ascontiguousarray(p, dtype=np.float64)
    ap, dg = np.empty(topo.ndof), np.empty(topo.ndof)
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)
    rc = lib.pcfdref_operator(C.c_int(nlev), ptr(h), ptr(first), ptr(arrs["nbr"]), ptr(arrs["flags"]),
                              C.c_int(1 if topo.info.outlet else 0), ptr(arrs["mask"]),
                              C.c_int(len(arrs["ea"])), ptr(arrs["ea"]), ptr(arrs["eb"]), ptr(arrs["area"]),
                              ptr(arrs["coef"]), ptr(p), ptr(ap), ptr(dg))
    assert rc == 0
    return ap, dg


CASES = {
    "sphere_amr": lambda: cases.sphere_box(cells_x=32, levels=2, band=0.3),
    "naca_2lvl": lambda: cases.naca_wing(digits=2412, alpha_deg=4.0, cells_x=48, levels=2, bands=(0.3,)),
    "naca_3lvl": lambda: cases.naca_wing(digits=4412, alpha_deg=8.0, cells_x=32, levels=3, bands=(0.5, 0.2)),
    "paraglider": lambda: cases.paraglider(cells_x=48, levels=2, bands=(0.5,)),
    "box_outlet": lambda: cases.uniform_box(n=16, h=0.0625, outlet=True),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_operator_matches_reference_code(ref, oracle_built, name):
    from oracle.oracle import Oracle
    from paracfd_amd import default_options
    topo = L.Topology(*CASES[name]())
    orc = Oracle(topo, default_options())
    act = (topo.array("mask") & 64) != 0
    rng = np.random.default_rng(11)
    p = rng.standard_normal(topo.ndof) * act
    ap_ref, dg_ref = reference_operator(ref, topo, p)
    ap = orc.mg_primitive(0, 0, x=p)
    scale = np.max(np.abs(ap_ref[act]))
    assert np.max(np.abs(ap[act] - ap_ref[act])) <= 1e-13 * scale
    # diagonal: one Jacobi sweep from zero with b = 1, weight 1 gives 1 / D
    inv_d = orc.mg_primitive(2, 0, b=act.astype(float), omega=1.0, sweeps=1)
    assert np.allclose(1.0 / inv_d[act], dg_ref[act], rtol=1e-13, atol=0)
    # a non-trivial share of the rows comes from compact edges (cut cells, coarse/fine faces)
    if name != "box_outlet":
        assert topo.info.compact_edges > 0
