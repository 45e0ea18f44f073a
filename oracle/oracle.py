"""ORACLE - TEST INFRASTRUCTURE ONLY.

ctypes binding of the serial FP64 oracle (oracle/oracle.cpp) and of the legacy uniform-MAC
restatement (oracle/mac_legacy.cpp).  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module; the product path never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
LEGACY_SO = os.path.join(HERE, "libmac_legacy.so")


def build(force=False):
    """Compile the oracle libraries with g++ (host only)."""
    subprocess.run(["make", "-C", HERE, "-s"] + (["-B"] if force else []), check=True)


REFERENCE_SRC = "/root/reference/src"


def build_reference():
    """Compile the reference's own composite-operator code into _ref/libref.so (test
    infrastructure).  Only where the reference sources are present; elsewhere the tests that
    use it skip.  Returns whether the library was built."""
    if not os.path.isdir(os.path.join(REFERENCE_SRC, "core", "fluid")):
        return False
    subprocess.run(["make", "-C", HERE, "-s", "-B", "ref", f"REF={REFERENCE_SRC}"], check=True)
    return True


def _need(path):
    if not os.path.exists(path):
        build()
    return C.CDLL(path)


class OracleTopology(C.Structure):
    _fields_ = [("nbricks", C.c_int64), ("ndof", C.c_int64), ("nspecial", C.c_int64), ("nedges", C.c_int64),
                ("npatches", C.c_int64), ("levels", C.c_int), ("outlet", C.c_int),
                ("level_h", C.c_void_p), ("brick_level", C.c_void_p), ("brick_flags", C.c_void_p),
                ("brick_nbr", C.c_void_p), ("mask", C.c_void_p), ("special_dof", C.c_void_p),
                ("special_volume", C.c_void_p), ("special_patch", C.c_void_p), ("edge_a", C.c_void_p),
                ("edge_b", C.c_void_p), ("edge_axis", C.c_void_p), ("edge_area", C.c_void_p),
                ("edge_coef", C.c_void_p), ("edge_w", C.c_void_p), ("patch_dof", C.c_void_p),
                ("patch_S", C.c_void_p), ("patch_x", C.c_void_p), ("patch_D", C.c_void_p),
                ("edge_x", C.c_void_p), ("edge_dcoef", C.c_void_p), ("finest_h", C.c_double),
                ("max_diffusion_coefficient", C.c_double), ("mg_levels", C.c_int), ("mg_n", C.c_void_p),
                ("mg_parent", C.c_void_p), ("mg_row_off", C.c_void_p), ("mg_col", C.c_void_p),
                ("mg_coef", C.c_void_p), ("mg_diag", C.c_void_p), ("coarse_inv", C.c_void_p),
                ("centroid", C.c_void_p), ("ngroups", C.c_int64), ("redist_group", C.c_void_p),
                ("redist_off", C.c_void_p), ("redist_member", C.c_void_p), ("redist_volume", C.c_void_p)]


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None and a.size else None


class Oracle:
    """Serial FP64 step on the topology of a paracfd_amd._lib.Topology."""

    def __init__(self, topo, options, finest_h=None, max_diffusion_coefficient=None):
        from paracfd_amd import _lib as L  # options / stats structs are shared with the ABI header
        self.L = L
        lib = _need(ORACLE_SO)
        self.lib = lib
        P = C.c_void_p
        dp = C.POINTER(C.c_double)
        lib.oracle_create.restype = P
        lib.oracle_create.argtypes = [C.POINTER(OracleTopology), C.POINTER(L.Options)]
        lib.oracle_destroy.argtypes = [P]
        lib.oracle_last_error.restype = C.c_char_p
        for name, args in {
            "oracle_initialize": [P, C.POINTER(L.StepStats)], "oracle_step": [P, C.POINTER(L.StepStats)],
            "oracle_advect": [P, C.c_double, dp, C.POINTER(C.c_int)],
            "oracle_project": [P, C.c_double, C.c_int, C.POINTER(L.StepStats)],
            "oracle_impulse": [P, C.c_double], "oracle_loads": [P, C.c_double, C.POINTER(L.Loads)],
            "oracle_get_state": [P, dp, dp, dp, dp], "oracle_set_state": [P, dp, dp, dp, dp],
            "oracle_get_faces": [P, dp, dp, dp, dp], "oracle_set_faces": [P, dp, dp, dp, dp],
            "oracle_mark_initialized": [P, C.c_double],
            "oracle_mg_primitive": [P, C.c_int, C.c_int, dp, dp, dp, C.c_double, C.c_int],
            "oracle_dense_operator": [P, dp], "oracle_divergence": [P, dp], "oracle_viscosity": [P, dp],
            "oracle_smooth_wall": [P, C.c_double, dp], "oracle_gradients": [P, dp, dp],
            "oracle_plant": [P, C.c_int], "oracle_rhs": [P, C.c_double, dp], "oracle_correct": [P, C.c_double],
        }.items():
            f = getattr(lib, name)
            f.restype = C.c_int
            f.argtypes = args
        lib.oracle_set_reference.argtypes = [P, C.c_double, dp]
        lib.oracle_outflow_rate.restype = C.c_double
        lib.oracle_outflow_rate.argtypes = [P]
        lib.oracle_time_steps.restype = C.c_double
        lib.oracle_time_steps.argtypes = [P, C.c_int]
        lib.oracle_wall_friction.restype = C.c_double
        lib.oracle_wall_friction.argtypes = [C.c_double, C.c_double, C.c_double]

        t = topo
        self.topo = topo
        keep = {}
        for name in ["level_h", "brick_level", "brick_flags", "brick_nbr", "mask", "special_dof", "special_volume",
                     "special_patch", "edge_a", "edge_b", "edge_axis", "edge_area", "edge_coef", "edge_w",
                     "patch_dof", "patch_S", "patch_x", "patch_D", "edge_x", "edge_dcoef", "centroid",
                     "redist_group", "redist_off", "redist_member", "redist_volume"]:
            keep[name] = np.ascontiguousarray(t.array(name))
        nl = t.mg_levels
        mg_n = np.array([t.mg_n(m) for m in range(nl)], dtype=np.int64)
        parents = [np.ascontiguousarray(t.mg_array(m, "parent")) for m in range(nl)]
        rows = [None] + [np.ascontiguousarray(t.mg_array(m, "row_off")) for m in range(1, nl)]
        cols = [None] + [np.ascontiguousarray(t.mg_array(m, "col")) for m in range(1, nl)]
        coefs = [None] + [np.ascontiguousarray(t.mg_array(m, "coef")) for m in range(1, nl)]
        diags = [None] + [np.ascontiguousarray(t.mg_array(m, "diag")) for m in range(1, nl)]
        cinv = np.ascontiguousarray(t.mg_array(nl - 1, "coarse_inv")) if nl > 1 else None

        def ptr_array(lst):
            arr = (C.c_void_p * len(lst))()
            for i, a in enumerate(lst):
                arr[i] = _ptr(a) if a is not None else None
            return arr

        self._keep = (keep, mg_n, parents, rows, cols, coefs, diags, cinv)
        ot = OracleTopology()
        info = t.info
        ot.nbricks, ot.ndof, ot.nspecial, ot.nedges, ot.npatches = (info.bricks, info.dofs, info.special_cells,
                                                                    info.compact_edges, info.patches)
        ot.levels = info.levels
        ot.outlet = 1 if t.outlet else 0
        for name, a in keep.items():
            setattr(ot, name, _ptr(a))
        ot.ngroups = int(keep["redist_volume"].size)
        ot.finest_h = float(keep["level_h"][-1])
        ot.max_diffusion_coefficient = t.max_diffusion_coefficient
        ot.mg_levels = nl
        self._arrs = [ptr_array(parents), ptr_array(rows), ptr_array(cols), ptr_array(coefs), ptr_array(diags)]
        ot.mg_n = _ptr(mg_n)
        ot.mg_parent = C.cast(self._arrs[0], C.c_void_p)
        ot.mg_row_off = C.cast(self._arrs[1], C.c_void_p)
        ot.mg_col = C.cast(self._arrs[2], C.c_void_p)
        ot.mg_coef = C.cast(self._arrs[3], C.c_void_p)
        ot.mg_diag = C.cast(self._arrs[4], C.c_void_p)
        ot.coarse_inv = _ptr(cinv)
        self._ot = ot
        h = lib.oracle_create(C.byref(ot), C.byref(options))
        if not h:
            raise RuntimeError(lib.oracle_last_error().decode())
        self.h = h
        self.n = int(info.dofs)
        self.nedges = int(info.compact_edges)

    def close(self):
        if self.h:
            self.lib.oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _stats(self, fn, *args):
        st = self.L.StepStats()
        rc = fn(self.h, *args, C.byref(st))
        if rc != 0:
            raise RuntimeError(f"oracle status {rc}: {self.lib.oracle_last_error().decode()}")
        return st

    def initialize(self):
        return self._stats(self.lib.oracle_initialize)

    def step(self):
        return self._stats(self.lib.oracle_step)

    def advect(self, dt=0.0):
        used = C.c_double()
        subs = C.c_int()
        self.lib.oracle_advect(self.h, dt, C.byref(used), C.byref(subs))
        self.last_substeps = subs.value
        return used.value

    def project(self, dt, warm=True):
        return self._stats(self.lib.oracle_project, C.c_double(dt), C.c_int(1 if warm else 0))

    def impulse(self, dt):
        self.lib.oracle_impulse(self.h, dt)

    def set_reference(self, area, origin=(0.0, 0.0, 0.0)):
        o = (C.c_double * 3)(*origin)
        self.lib.oracle_set_reference(self.h, area, o)

    def loads(self, pref=0.0):
        out = self.L.Loads()
        self.lib.oracle_loads(self.h, pref, C.byref(out))
        return out.as_dict()

    def wall_friction(self, ut, y, nu):
        """Friction velocity of the smooth-wall law for tangential speed ut at wall distance y."""
        return self.lib.oracle_wall_friction(ut, y, nu)

    def smooth_wall(self, dt):
        """Apply only the smooth-wall model to the current velocity; returns the patch forces [npatch, 3]."""
        npatch = int(self.topo.info.patches)
        f = np.zeros(3 * max(npatch, 1))
        rc = self.lib.oracle_smooth_wall(self.h, dt, f.ctypes.data_as(C.POINTER(C.c_double)))
        if rc != 0:
            raise RuntimeError(self.lib.oracle_last_error().decode())
        return f[:3 * npatch].reshape(npatch, 3)

    def state(self):
        x, y, z, p = (np.empty(self.n) for _ in range(4))
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        self.lib.oracle_get_state(self.h, dp(x), dp(y), dp(z), dp(p))
        return x, y, z, p

    def set_state(self, x=None, y=None, z=None, p=None):
        dp = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(C.POINTER(C.c_double))
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z, p)]
        self.lib.oracle_set_state(self.h, *[dp(a) for a in arrs])

    def faces(self):
        fx, fy, fz = (np.empty(self.n) for _ in range(3))
        fe = np.empty(max(self.nedges, 1))
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        self.lib.oracle_get_faces(self.h, dp(fx), dp(fy), dp(fz), dp(fe))
        return fx, fy, fz, fe[: self.nedges]

    def set_faces(self, fx=None, fy=None, fz=None, fe=None):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (fx, fy, fz, fe)]
        dp = lambda a: None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))
        self.lib.oracle_set_faces(self.h, *[dp(a) for a in arrs])

    def mark_initialized(self, time=0.0):
        self.lib.oracle_mark_initialized(self.h, time)

    def mg_primitive(self, op, level, x=None, b=None, omega=2.0 / 3.0, sweeps=1):
        n = self.topo.mg_n(level)
        out = np.empty(n)
        xa = None if x is None else np.ascontiguousarray(x, dtype=np.float64)
        ba = None if b is None else np.ascontiguousarray(b, dtype=np.float64)
        dp = lambda a: None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))
        if xa is None:
            xa = np.zeros(n)
        rc = self.lib.oracle_mg_primitive(self.h, op, level, dp(xa), dp(ba), dp(out), omega, sweeps)
        if rc != 0:
            raise RuntimeError(self.lib.oracle_last_error().decode())
        return out

    def dense_operator(self):
        A = np.empty((self.n, self.n))
        self.lib.oracle_dense_operator(self.h, A.ctypes.data_as(C.POINTER(C.c_double)))
        return A

    def divergence(self):
        d = np.empty(self.n)
        self.lib.oracle_divergence(self.h, d.ctypes.data_as(C.POINTER(C.c_double)))
        return d

    def viscosity(self):
        """Cell viscosity nu_mol + (Cs V^(1/3))^2 |S| of the current velocity."""
        nu = np.empty(self.n)
        rc = self.lib.oracle_viscosity(self.h, nu.ctypes.data_as(C.POINTER(C.c_double)))
        if rc != 0:
            raise RuntimeError(self.lib.oracle_last_error().decode())
        return nu

    def gradients(self):
        """(velocity gradients [nspecial, 3, 3] with [s, c, r] = d v_c / d x_r, pressure
        gradients [nspecial, 3]) at the special cells, least squares on the current state."""
        ns = int(self.topo.info.special_cells)
        gv = np.zeros(9 * max(ns, 1))
        gp = np.zeros(3 * max(ns, 1))
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        self.lib.oracle_gradients(self.h, dp(gv), dp(gp))
        return gv[:9 * ns].reshape(ns, 3, 3), gp[:3 * ns].reshape(ns, 3)

    def plant(self, defect):
        """Plant mistake `defect` in this oracle (contract tests show that it fails them); 0 removes it."""
        self.lib.oracle_plant(self.h, int(defect))

    def correct(self, dt):
        """Face correction u -= (dt/rho) dp/dn with the current pressure (no solve)."""
        self.lib.oracle_correct(self.h, dt)

    def rhs(self, dt):
        """-(rho/dt) times the integrated divergence of the current faces."""
        b = np.empty(self.n)
        self.lib.oracle_rhs(self.h, dt, b.ctypes.data_as(C.POINTER(C.c_double)))
        return b

    def outflow_rate(self):
        return self.lib.oracle_outflow_rate(self.h)

    def time_steps(self, steps):
        return self.lib.oracle_time_steps(self.h, steps)


class Legacy:
    """Restatement of the reference's uniform FP64 MAC primitives (project.cu, advect.cu, turbulence.cu)."""

    def __init__(self):
        self.lib = _need(LEGACY_SO)

    def fn(self, name):
        return getattr(self.lib, name)
