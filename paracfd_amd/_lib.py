"""ctypes binding of the C ABI declared in include/paracfd.h (libparacfd.so, built in-tree).

The library is the product: topology preprocessing on the host plus the HIP solver for gfx950.
Loading fails loudly when it has not been built; there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libparacfd.so")

PCFD_OK, PCFD_EINVAL, PCFD_EDEVICE, PCFD_ENOTCONVERGED, PCFD_ENONFINITE, PCFD_ESTATE, PCFD_ENOMEM = range(7)
GEOM_NONE, GEOM_SPHERE, GEOM_NACA4_WING, GEOM_PARAGLIDER = range(4)
FP64, FP32, BF16 = range(3)
PC_MG, PC_JACOBI, PC_LEARNED = range(3)

ARR = dict(
    brick_level=(0, np.int32), brick_coord=(1, np.int32), brick_flags=(2, np.uint8), brick_nbr=(3, np.int32),
    mask=(4, np.uint8), special_dof=(5, np.int32), special_volume=(6, np.float64), special_inc_off=(7, np.int32),
    special_inc=(8, np.int32), special_patch=(9, np.int32), brick_special_off=(10, np.int32), edge_a=(11, np.int32),
    edge_b=(12, np.int32), edge_axis=(13, np.int8), edge_area=(14, np.float64), edge_coef=(15, np.float64),
    edge_w=(16, np.float64), patch_dof=(17, np.int32), patch_S=(18, np.float64), patch_x=(19, np.float64),
    patch_D=(20, np.float64), edge_x=(21, np.float64), edge_dcoef=(22, np.float64),
    centroid=(23, np.float64), level_h=(24, np.float64),
    nf_lower=(25, np.int32), nf_upper=(26, np.int32), nf_edge=(27, np.int32), nf_area=(28, np.float64),
    nf_delta=(29, np.float64), nf_k=(30, np.float64), nf_w=(31, np.float64), nf_lnode=(32, np.int32),
    nf_unode=(33, np.int32), ng_dof=(34, np.int32), ng_off=(35, np.int32), ng_nbr=(36, np.int32),
    ng_w=(37, np.float64), ng_rank=(38, np.uint8), ng_ring=(39, np.uint8),
    redist_group=(40, np.int32), redist_off=(41, np.int32), redist_member=(42, np.int32),
    redist_volume=(43, np.float64),
)
MG_ARR = dict(row_off=(0, np.int32), col=(1, np.int32), coef=(2, np.float64), diag=(3, np.float64),
              parent=(4, np.int32), coarse_inv=(5, np.float64))


class GeometryDesc(C.Structure):
    _fields_ = [("kind", C.c_int), ("center", C.c_double * 3), ("radius", C.c_double), ("naca_digits", C.c_int),
                ("chord", C.c_double), ("span", C.c_double), ("alpha_deg", C.c_double), ("sweep_deg", C.c_double),
                ("anhedral_deg", C.c_double), ("reflex", C.c_double)]


class GridDesc(C.Structure):
    _fields_ = [("lo", C.c_double * 3), ("hi", C.c_double * 3), ("base_cells_x", C.c_int), ("levels", C.c_int),
                ("refine_band", C.c_double * 4), ("min_volume_fraction", C.c_double), ("subsamples", C.c_int),
                ("outlet_xmax", C.c_int), ("merge_volume_fraction", C.c_double)]


class TopologyInfo(C.Structure):
    _fields_ = [("bricks", C.c_int64), ("dofs", C.c_int64), ("active_cells", C.c_int64),
                ("special_cells", C.c_int64), ("compact_edges", C.c_int64), ("patches", C.c_int64),
                ("mg_levels", C.c_int64), ("mg_cells", C.c_int64 * 12), ("levels", C.c_int), ("h0", C.c_double),
                ("lo", C.c_double * 3), ("base_bricks", C.c_int * 3), ("body_volume", C.c_double),
                ("reference_area", C.c_double), ("finest_h", C.c_double),
                ("max_diffusion_coefficient", C.c_double), ("outlet", C.c_int), ("instances", C.c_int)]


class InstanceInfo(C.Structure):
    _fields_ = [("brick_begin", C.c_int64), ("brick_end", C.c_int64), ("dof_begin", C.c_int64),
                ("dof_end", C.c_int64), ("patch_begin", C.c_int64), ("patch_end", C.c_int64),
                ("active_cells", C.c_int64), ("reference_area", C.c_double), ("body_volume", C.c_double)]


class Options(C.Structure):
    _fields_ = [("rho", C.c_double), ("nu", C.c_double), ("speed", C.c_double), ("cfl", C.c_double),
                ("smagorinsky_cs", C.c_double), ("rtol", C.c_double), ("max_iterations", C.c_int),
                ("field_precision", C.c_int), ("vcycle_precision", C.c_int), ("preconditioner", C.c_int),
                ("pre_smooth", C.c_int), ("post_smooth", C.c_int), ("check_interval", C.c_int),
                ("smoother_weight", C.c_double), ("learned_bricks", C.c_int), ("learned_capacity", C.c_int),
                ("smooth_wall", C.c_int), ("smoother_weight2", C.c_double),
                ("nonorthogonal", C.c_int), ("gmres_restart", C.c_int), ("memory_budget_bytes", C.c_double),
                ("gmres_recycle", C.c_int)]


NN_C = 16
NN_SETS = 4      # one weight set per AMR level (levels >= 3 share the last)
NN_LAYERS = 4
NN_WEIGHTS = NN_SETS * NN_LAYERS * 27 * NN_C * NN_C  # [set][layer][tap][cout][cin]
NN_BIASES = NN_SETS * NN_LAYERS * NN_C


class LearnedStats(C.Structure):
    _fields_ = [("pairs", C.c_int), ("train_steps", C.c_int), ("loss_first", C.c_double),
                ("loss_last", C.c_double), ("collect_ms", C.c_double), ("train_ms", C.c_double),
                ("bricks", C.c_int), ("cells", C.c_int64)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Tuning(C.Structure):
    _fields_ = [("xcd_chunk_log2", C.c_int), ("coarse_block", C.c_int), ("fused_vcycle", C.c_int),
                ("graph_replay", C.c_int), ("candidates", C.c_int), ("step_ms", C.c_double),
                ("default_step_ms", C.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class StepStats(C.Structure):
    _fields_ = [("dt", C.c_double), ("physical_time", C.c_double), ("iterations", C.c_int), ("converged", C.c_int),
                ("relative_residual", C.c_double), ("max_abs_velocity", C.c_double), ("cfl_rate", C.c_double),
                ("max_abs_divergence", C.c_double), ("diffusion_substeps", C.c_int), ("fallback_used", C.c_int),
                ("learned_used", C.c_int), ("learned_rejected", C.c_int),
                ("advect_ms", C.c_double), ("project_ms", C.c_double), ("step_ms", C.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Loads(C.Structure):
    _fields_ = [("force", C.c_double * 3), ("moment", C.c_double * 3), ("drag", C.c_double), ("side", C.c_double),
                ("lift", C.c_double), ("cd", C.c_double), ("cs", C.c_double), ("cl", C.c_double),
                ("coefficients_valid", C.c_int), ("viscous_valid", C.c_int),
                ("viscous_force", C.c_double * 3), ("viscous_moment", C.c_double * 3),
                ("viscous_drag", C.c_double), ("viscous_side", C.c_double), ("viscous_lift", C.c_double),
                ("cd_viscous", C.c_double), ("cs_viscous", C.c_double), ("cl_viscous", C.c_double),
                ("total_force", C.c_double * 3), ("total_moment", C.c_double * 3),
                ("total_drag", C.c_double), ("total_side", C.c_double), ("total_lift", C.c_double),
                ("total_cd", C.c_double), ("total_cs", C.c_double), ("total_cl", C.c_double)]

    def as_dict(self):
        d = {name: getattr(self, name) for name, _ in self._fields_}
        for name in ("force", "moment", "viscous_force", "viscous_moment", "total_force", "total_moment"):
            d[name] = list(getattr(self, name))
        return d


# every symbol the public header declares
EXPORTED = [
    "pcfd_last_error", "pcfd_abi_version", "pcfd_topology_build", "pcfd_topology_free", "pcfd_topology_get_info",
    "pcfd_topology_array_size", "pcfd_topology_export", "pcfd_topology_build_nonorthogonal", "pcfd_create", "pcfd_destroy", "pcfd_default_options",
    "pcfd_initialize", "pcfd_step", "pcfd_advect", "pcfd_project", "pcfd_apply_pressure_impulse",
    "pcfd_integrate_loads", "pcfd_set_reference", "pcfd_get_velocity", "pcfd_set_velocity", "pcfd_get_pressure",
    "pcfd_set_pressure", "pcfd_get_faces", "pcfd_set_faces", "pcfd_mark_initialized", "pcfd_device_bytes",
    "pcfd_mg_primitive", "pcfd_time_kernel", "pcfd_learned_collect", "pcfd_learned_train",
    "pcfd_learned_get_weights", "pcfd_learned_set_weights", "pcfd_learned_enable", "pcfd_learned_forward",
    "pcfd_learned_gradient", "pcfd_topology_merge", "pcfd_topology_instance", "pcfd_instance_stats",
    "pcfd_integrate_loads_instance", "pcfd_set_reference_instance", "pcfd_integrate_loads_batch",
    "pcfd_autotune", "pcfd_get_tuning", "pcfd_set_tuning", "pcfd_sample_slice", "pcfd_probe_start",
    "pcfd_probe_stop", "pcfd_sample_slice_static", "pcfd_topology_build_on_device", "pcfd_topology_build_times",
]

_lib = None


def load():
    """Load libparacfd.so; raise if it has not been built (no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"paracfd_amd: {LIB_PATH} is not built; run __graft_entry__.build()")
    # PyTorch ships its own HIP runtime.  When both are used in one process, torch's runtime
    # must be loaded first so that libparacfd binds to the same one: two HIP runtimes in one
    # process cannot both open the device.  Without torch, the ROCm runtime is used alone.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    dp = C.POINTER(C.c_double)
    fp = C.POINTER(C.c_float)
    sig = {
        "pcfd_last_error": (C.c_char_p, []),
        "pcfd_abi_version": (C.c_int, []),
        "pcfd_topology_build": (C.c_int, [C.POINTER(GeometryDesc), C.POINTER(GridDesc), C.POINTER(P)]),
        "pcfd_topology_build_on_device": (C.c_int, [C.POINTER(GeometryDesc), C.POINTER(GridDesc), C.c_int,
                                                    C.POINTER(P)]),
        "pcfd_topology_build_times": (C.c_int, [P, C.POINTER(C.c_double), C.POINTER(C.c_int)]),
        "pcfd_topology_free": (None, [P]),
        "pcfd_topology_get_info": (C.c_int, [P, C.POINTER(TopologyInfo)]),
        "pcfd_topology_array_size": (C.c_int64, [P, C.c_int]),
        "pcfd_topology_export": (C.c_int, [P, C.c_int, P, C.c_size_t]),
        "pcfd_topology_build_nonorthogonal": (C.c_int, [P]),
        "pcfd_create": (C.c_int, [P, C.POINTER(Options), C.POINTER(P)]),
        "pcfd_destroy": (None, [P]),
        "pcfd_default_options": (None, [C.POINTER(Options)]),
        "pcfd_initialize": (C.c_int, [P, C.POINTER(StepStats)]),
        "pcfd_step": (C.c_int, [P, C.POINTER(StepStats)]),
        "pcfd_advect": (C.c_int, [P, C.c_double, dp, C.POINTER(StepStats)]),
        "pcfd_project": (C.c_int, [P, C.c_double, C.c_int, C.POINTER(StepStats)]),
        "pcfd_apply_pressure_impulse": (C.c_int, [P, C.c_double]),
        "pcfd_integrate_loads": (C.c_int, [P, C.c_double, C.POINTER(Loads)]),
        "pcfd_set_reference": (C.c_int, [P, C.c_double, dp]),
        "pcfd_get_velocity": (C.c_int, [P, dp, dp, dp]),
        "pcfd_set_velocity": (C.c_int, [P, dp, dp, dp]),
        "pcfd_get_pressure": (C.c_int, [P, dp]),
        "pcfd_set_pressure": (C.c_int, [P, dp]),
        "pcfd_get_faces": (C.c_int, [P, dp, dp, dp, dp]),
        "pcfd_set_faces": (C.c_int, [P, dp, dp, dp, dp]),
        "pcfd_mark_initialized": (C.c_int, [P, C.c_double]),
        "pcfd_device_bytes": (C.c_int64, [P]),
        "pcfd_mg_primitive": (C.c_int, [P, C.c_int, C.c_int, dp, dp, dp, C.c_double, C.c_int]),
        "pcfd_time_kernel": (C.c_int, [P, C.c_int, C.c_int, dp, dp]),
        "pcfd_probe_start": (C.c_int, [P, C.c_uint]),
        "pcfd_probe_stop": (C.c_int, [P, C.POINTER(C.c_int64), dp]),
        "pcfd_learned_collect": (C.c_int, [P, C.c_int, C.POINTER(LearnedStats)]),
        "pcfd_learned_train": (C.c_int, [P, C.c_int, C.c_int, C.c_double, C.POINTER(LearnedStats)]),
        "pcfd_learned_get_weights": (C.c_int, [P, fp, fp]),
        "pcfd_learned_set_weights": (C.c_int, [P, fp, fp]),
        "pcfd_learned_enable": (C.c_int, [P, C.c_int]),
        "pcfd_learned_forward": (C.c_int, [P, dp, fp, fp]),
        "pcfd_learned_gradient": (C.c_int, [P, dp, dp, fp, fp]),
        "pcfd_topology_merge": (C.c_int, [C.POINTER(P), C.c_int, C.POINTER(P)]),
        "pcfd_topology_instance": (C.c_int, [P, C.c_int, C.POINTER(InstanceInfo)]),
        "pcfd_instance_stats": (C.c_int, [P, C.c_int, C.POINTER(StepStats)]),
        "pcfd_integrate_loads_instance": (C.c_int, [P, C.c_int, C.c_double, C.POINTER(Loads)]),
        "pcfd_set_reference_instance": (C.c_int, [P, C.c_int, C.c_double, dp]),
        "pcfd_integrate_loads_batch": (C.c_int, [P, C.c_double, C.POINTER(Loads), C.c_int]),
        "pcfd_autotune": (C.c_int, [P, C.c_int, C.POINTER(Tuning)]),
        "pcfd_get_tuning": (C.c_int, [P, C.POINTER(Tuning)]),
        "pcfd_set_tuning": (C.c_int, [P, C.POINTER(Tuning)]),
        "pcfd_sample_slice": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_double, dp, dp, C.c_int, C.c_int, fp]),
        "pcfd_sample_slice_static": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_double, dp, dp, C.c_int, C.c_int, fp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class PcfdError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"pcfd status {status}: {message}")
        self.status = status


def check(status):
    if status != PCFD_OK:
        raise PcfdError(status, load().pcfd_last_error().decode())
    return status


def dptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def fptr(a):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Topology:
    """Host topology (bricks, cut cells, compact edges, patches, multigrid hierarchy)."""

    BUILD_STAGES = ("geometry", "refine", "cut_list", "sampling", "components", "finalize", "multigrid", "total",
                    "device_kernel")

    def __init__(self, geometry: GeometryDesc = None, grid: GridDesc = None, _handle=None, device=None):
        """Build on the host (pcfd_topology_build), or with the cut cells sub-sampled on HIP
        device `device` (pcfd_topology_build_on_device: bit-identical arrays)."""
        lib = load()
        self._lib = lib
        if _handle is None:
            h = C.c_void_p()
            if device is None:
                check(lib.pcfd_topology_build(C.byref(geometry), C.byref(grid), C.byref(h)))
            else:
                check(lib.pcfd_topology_build_on_device(C.byref(geometry), C.byref(grid), int(device), C.byref(h)))
        else:
            h = _handle
        self.handle = h
        self.info = TopologyInfo()
        check(lib.pcfd_topology_get_info(h, C.byref(self.info)))
        self._cache = {}

    @classmethod
    def merge(cls, parts):
        """One batch topology of independent single cases (pcfd_topology_merge)."""
        lib = load()
        arr = (C.c_void_p * len(parts))(*[p.handle.value for p in parts])
        h = C.c_void_p()
        check(lib.pcfd_topology_merge(arr, len(parts), C.byref(h)))
        return cls(_handle=h)

    @property
    def instances(self):
        return int(self.info.instances)

    def build_times(self):
        """Wall milliseconds per build stage (pcfd_topology_build_times) and the sampling device
        (-1: host)."""
        ms = (C.c_double * 9)()
        dev = C.c_int()
        check(self._lib.pcfd_topology_build_times(self.handle, ms, C.byref(dev)))
        return dict(zip(self.BUILD_STAGES, list(ms))), int(dev.value)

    def instance(self, i):
        info = InstanceInfo()
        check(self._lib.pcfd_topology_instance(self.handle, i, C.byref(info)))
        return {name: getattr(info, name) for name, _ in InstanceInfo._fields_}

    def close(self):
        if self.handle:
            self._lib.pcfd_topology_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _export(self, aid, dtype):
        n = self._lib.pcfd_topology_array_size(self.handle, aid)
        if n < 0:
            raise KeyError(aid)
        out = np.empty(n, dtype=dtype)
        check(self._lib.pcfd_topology_export(self.handle, aid, out.ctypes.data_as(C.c_void_p), out.nbytes))
        return out

    def build_nonorthogonal(self):
        """Build the non-orthogonal correction (pcfd_topology_build_nonorthogonal); its arrays
        nf_* (faces) and ng_* (gradient nodes) are exported afterwards."""
        check(self._lib.pcfd_topology_build_nonorthogonal(self.handle))
        for k in [k for k in self._cache if isinstance(k, str) and k[:3] in ("nf_", "ng_")]:
            del self._cache[k]
        return self

    def array(self, name):
        if name not in self._cache:
            aid, dt = ARR[name]
            self._cache[name] = self._export(aid, dt)
        return self._cache[name]

    def mg_array(self, level, name):
        key = (level, name)
        if key not in self._cache:
            kind, dt = MG_ARR[name]
            self._cache[key] = self._export(100 + 8 * level + kind, dt)
        return self._cache[key]

    @property
    def ndof(self):
        return int(self.info.dofs)

    @property
    def n_active(self):
        return int(self.info.active_cells)

    @property
    def mg_levels(self):
        return int(self.info.mg_levels)

    @property
    def outlet(self):
        return bool(self.info.outlet)

    @property
    def max_diffusion_coefficient(self):
        return float(self.info.max_diffusion_coefficient)

    def mg_n(self, level):
        return self.ndof if level == 0 else int(self.mg_array(level, "diag").size)
