"""The oracle's restatement of the reference's uniform MAC / V-cycle primitives against the
reference's frozen FP64 snapshots (tests/golden/*.f64, written by tools/parity_probe.cu:102-178).

Inputs are regenerated exactly as the probe makes them: std::mt19937 seeds 101-109 with
std::uniform_real_distribution<double> on a 16^3 grid, h = 0.05.  The probe's own gate is a
relative max-norm of 1e-5 (parity_probe.cu:59-89); the restatement is held to 1e-12.
"""
import ctypes as C
import os

import numpy as np
import pytest

from oracle.oracle import Legacy

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NX = NY = NZ = 16
H = 0.05
RHO, NU, DT, CS, OMEGA = 1.225, 1e-3, 1e-3, 0.16, 2.0 / 3.0
NP, NU_C, NV_C, NW_C = NX * NY * NZ, (NX + 1) * NY * NZ, NX * (NY + 1) * NZ, NX * NY * (NZ + 1)
TOL = 1e-12


def golden(name):
    return np.fromfile(os.path.join(GOLDEN, name + ".f64"), dtype="<f8")


def relmax(a, ref):
    return np.max(np.abs(a - ref)) / max(np.max(np.abs(ref)), 1e-30)


@pytest.fixture(scope="module")
def leg(oracle_built):
    return Legacy()


def dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


@pytest.fixture(scope="module")
def inputs(leg):
    f = leg.fn("legacy_fill")
    f.argtypes = [C.c_int64, C.c_uint32, C.c_double, C.c_double, C.POINTER(C.c_double)]

    def filled(n, seed, lo, hi):
        out = np.empty(n)
        f(n, seed, lo, hi, dp(out))
        return out

    return dict(u=filled(NU_C, 101, -1, 1), v=filled(NV_C, 102, -1, 1), w=filled(NW_C, 103, -1, 1),
                p=filled(NP, 104, -1, 1), rhs=filled(NP, 105, -1, 1), nut=filled(NP, 106, 0, 0.05),
                fine=filled(NP, 107, -1, 1), coarse=filled(NP // 8, 108, -1, 1), fine2=filled(NP, 109, -1, 1))


def call(leg, name, *args):
    fn = leg.fn(name)
    conv = []
    types = []
    for a in args:
        if isinstance(a, np.ndarray):
            conv.append(dp(a)); types.append(C.POINTER(C.c_double))
        elif isinstance(a, float):
            conv.append(a); types.append(C.c_double)
        else:
            conv.append(a); types.append(C.c_int)
    fn.argtypes = types
    return fn(*conv)


def test_inputs_regenerate(inputs):
    # first values of std::mt19937(101) through uniform_real_distribution<double>(-1, 1)
    assert inputs["u"].shape == (NU_C,)
    assert np.all(inputs["nut"] >= 0) and np.all(inputs["nut"] <= 0.05)


def test_poisson_rhs(leg, inputs):
    out = np.empty(NP)
    call(leg, "legacy_poisson_rhs", NX, NY, NZ, H, inputs["u"], inputs["v"], inputs["w"], RHO, DT, out)
    assert relmax(out, golden("poisson_rhs")) <= TOL


def test_poisson_apply(leg, inputs):
    out = np.empty(NP)
    call(leg, "legacy_poisson_apply", NX, NY, NZ, H, inputs["p"], out)
    assert relmax(out, golden("poisson_apply")) <= TOL


def test_poisson_residual(leg, inputs):
    out = np.empty(NP)
    call(leg, "legacy_poisson_residual", NX, NY, NZ, H, inputs["p"], inputs["rhs"], out)
    assert relmax(out, golden("poisson_residual")) <= TOL


def test_jacobi_smooth(leg, inputs):
    p = inputs["p"].copy()
    call(leg, "legacy_jacobi", NX, NY, NZ, H, p, inputs["rhs"], OMEGA, 2)
    assert relmax(p, golden("jacobi_smooth")) <= TOL


def test_gs_band(leg, inputs):
    p = inputs["p"].copy()
    call(leg, "legacy_gs_band", NX, NY, NZ, H, p, inputs["rhs"], 2, 2, 1)
    assert relmax(p, golden("gs_band")) <= TOL


def test_restrict(leg, inputs):
    out = np.empty(NP // 8)
    call(leg, "legacy_restrict", NX, NY, NZ, H, inputs["fine"], out)
    assert relmax(out, golden("restrict")) <= TOL


def test_prolong_add(leg, inputs):
    fine = inputs["fine2"].copy()
    call(leg, "legacy_prolong_add", NX, NY, NZ, H, inputs["coarse"], fine)
    assert relmax(fine, golden("prolong_add")) <= TOL


def test_subtract_gradient(leg, inputs):
    u, v, w = inputs["u"].copy(), inputs["v"].copy(), inputs["w"].copy()
    call(leg, "legacy_subtract_gradient", NX, NY, NZ, H, u, v, w, inputs["p"], RHO, DT)
    for name, a in (("u", u), ("v", v), ("w", w)):
        assert relmax(a, golden("subtract_gradient_" + name)) <= TOL


def test_max_abs_divergence(leg, inputs):
    fn = leg.fn("legacy_max_abs_divergence")
    fn.restype = C.c_double
    fn.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double] + [C.POINTER(C.c_double)] * 3
    m = fn(NX, NY, NZ, H, dp(inputs["u"]), dp(inputs["v"]), dp(inputs["w"]))
    assert abs(m - golden("max_abs_divergence")[0]) <= TOL * abs(m)


def test_mac_advect(leg, inputs):
    uo, vo, wo = np.empty(NU_C), np.empty(NV_C), np.empty(NW_C)
    call(leg, "legacy_mac_advect", NX, NY, NZ, H, inputs["u"], inputs["v"], inputs["w"], uo, vo, wo, DT, 1)
    for name, a in (("u", uo), ("v", vo), ("w", wo)):
        assert relmax(a, golden("mac_advect_" + name)) <= TOL


def test_smagorinsky_nut(leg, inputs):
    out = np.empty(NP)
    call(leg, "legacy_smagorinsky_nut", NX, NY, NZ, H, inputs["u"], inputs["v"], inputs["w"], CS, out)
    assert relmax(out, golden("smagorinsky_nut")) <= TOL


def test_mac_diffuse(leg, inputs):
    uo, vo, wo = np.empty(NU_C), np.empty(NV_C), np.empty(NW_C)
    call(leg, "legacy_mac_diffuse", NX, NY, NZ, H, inputs["u"], inputs["v"], inputs["w"], inputs["nut"], DT, NU,
         uo, vo, wo)
    for name, a in (("u", uo), ("v", vo), ("w", wo)):
        assert relmax(a, golden("mac_diffuse_" + name)) <= TOL


def test_planted_mistake_is_caught(leg, inputs):
    """A dropped face term in the restated operator must fail the golden comparison."""
    out = np.empty(NP)
    call(leg, "legacy_poisson_apply", NX, NY, NZ, H, inputs["p"], out)
    broken = out - inputs["p"] / (H * H)  # one neighbour's weight missing on the diagonal
    assert relmax(broken, golden("poisson_apply")) > 1e-3
