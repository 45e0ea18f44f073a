"""Python mirror of the reference's step interface over the C ABI.

`ExternalAeroCore` follows src/core/fluid/external_aero_core.h:124-182: initialize(), step(),
pressure_loads(pref), and the download/upload helpers.  Every call goes through
libparacfd.so (HIP on gfx950); there is no CPU implementation behind it.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


def default_options(**kw) -> L.Options:
    o = L.Options()
    L.load().pcfd_default_options(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise AttributeError(f"unknown option {k}")
        setattr(o, k, v)
    return o


class ExternalAeroCore:
    """One solver instance on the current HIP device (reference: ExternalAeroCore)."""

    def __init__(self, topology: L.Topology, options: L.Options):
        self._lib = L.load()
        self.topology = topology
        self.options = options
        h = C.c_void_p()
        L.check(self._lib.pcfd_create(topology.handle, C.byref(options), C.byref(h)))
        self.handle = h
        self.n = topology.ndof
        self.nedges = int(topology.info.compact_edges)

    def close(self):
        if getattr(self, "handle", None):
            self._lib.pcfd_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _stats(self, fn, *args) -> dict:
        st = L.StepStats()
        L.check(fn(self.handle, *args, C.byref(st)))
        return st.as_dict()

    # reference: ExternalAeroCore::initialize (external_aero_core.cpp:245-302)
    def initialize(self) -> dict:
        return self._stats(self._lib.pcfd_initialize)

    # reference: ExternalAeroCore::step (external_aero_core.cpp:304-328)
    def step(self) -> dict:
        return self._stats(self._lib.pcfd_step)

    def advect(self, dt=0.0) -> float:
        used = C.c_double()
        st = L.StepStats()
        L.check(self._lib.pcfd_advect(self.handle, dt, C.byref(used), C.byref(st)))
        return used.value

    def project(self, dt, warm=True) -> dict:
        return self._stats(self._lib.pcfd_project, C.c_double(dt), C.c_int(1 if warm else 0))

    def apply_pressure_impulse(self, dt):
        L.check(self._lib.pcfd_apply_pressure_impulse(self.handle, dt))

    def set_reference(self, area, origin=(0.0, 0.0, 0.0)):
        o = (C.c_double * 3)(*origin)
        L.check(self._lib.pcfd_set_reference(self.handle, area, o))

    # reference: ExternalAeroCore::pressure_loads (external_aero_core.cpp:330-337)
    def pressure_loads(self, pref=0.0) -> dict:
        out = L.Loads()
        L.check(self._lib.pcfd_integrate_loads(self.handle, pref, C.byref(out)))
        return out.as_dict()

    # reference: download_cell_momentum_state / upload_state (amr_advection.cu:1301-1309)
    def download_cell_momentum_state(self):
        x, y, z = (np.empty(self.n) for _ in range(3))
        L.check(self._lib.pcfd_get_velocity(self.handle, L.dptr(x), L.dptr(y), L.dptr(z)))
        return x, y, z

    def upload_state(self, x, y, z):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z)]
        L.check(self._lib.pcfd_set_velocity(self.handle, *[L.dptr(a) for a in arrs]))

    def download_pressure(self):
        p = np.empty(self.n)
        L.check(self._lib.pcfd_get_pressure(self.handle, L.dptr(p)))
        return p

    def upload_pressure(self, p):
        p = np.ascontiguousarray(p, dtype=np.float64)
        L.check(self._lib.pcfd_set_pressure(self.handle, L.dptr(p)))

    def download_faces(self):
        fx, fy, fz = (np.empty(self.n) for _ in range(3))
        fe = np.empty(max(self.nedges, 1))
        L.check(self._lib.pcfd_get_faces(self.handle, L.dptr(fx), L.dptr(fy), L.dptr(fz), L.dptr(fe)))
        return fx, fy, fz, fe[: self.nedges]

    def upload_faces(self, fx, fy, fz, fe):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (fx, fy, fz)]
        fe = np.ascontiguousarray(fe if len(fe) else np.zeros(1), dtype=np.float64)
        L.check(self._lib.pcfd_set_faces(self.handle, *[L.dptr(a) for a in arrs], L.dptr(fe)))

    def mark_initialized(self, time=0.0):
        L.check(self._lib.pcfd_mark_initialized(self.handle, time))

    def device_bytes(self) -> int:
        return int(self._lib.pcfd_device_bytes(self.handle))

    def time_kernel(self, kernel, reps=20):
        """(average ms per launch, algorithmic bytes per launch) of a hot kernel."""
        ms, by = C.c_double(), C.c_double()
        L.check(self._lib.pcfd_time_kernel(self.handle, kernel, reps, C.byref(ms), C.byref(by)))
        return ms.value, by.value

    def probe_start(self, kernels):
        """Time every launch of the given hot kernels (ids of time_kernel: 1, 2, 5, 6) in the
        calls that follow, with HIP events on the solver's stream."""
        mask = 0
        for k in kernels:
            mask |= 1 << int(k)
        L.check(self._lib.pcfd_probe_start(self.handle, mask))

    def probe_stop(self):
        """{kernel id: (launches, total ms)} of the launches timed since probe_start."""
        n = (C.c_int64 * 8)()
        ms = (C.c_double * 8)()
        L.check(self._lib.pcfd_probe_stop(self.handle, n, ms))
        return {k: (int(n[k]), float(ms[k])) for k in range(8) if n[k] > 0}

    def mg_primitive(self, op, level, x=None, b=None, omega=2.0 / 3.0, sweeps=1):
        n = self.topology.mg_n(level)
        out = np.empty(n)
        xa = None if x is None else np.ascontiguousarray(x, dtype=np.float64)
        ba = None if b is None else np.ascontiguousarray(b, dtype=np.float64)
        L.check(self._lib.pcfd_mg_primitive(self.handle, op, level, L.dptr(xa), L.dptr(ba), L.dptr(out), omega,
                                            sweeps))
        return out

    # batches (merged topologies); no reference counterpart
    def instance_stats(self, i) -> dict:
        st = L.StepStats()
        L.check(self._lib.pcfd_instance_stats(self.handle, i, C.byref(st)))
        return st.as_dict()

    def instance_loads(self, i, pref=0.0) -> dict:
        out = L.Loads()
        L.check(self._lib.pcfd_integrate_loads_instance(self.handle, i, pref, C.byref(out)))
        return out.as_dict()

    # launch-shape tuning; no reference counterpart
    def autotune(self, repetitions=2) -> dict:
        out = L.Tuning()
        L.check(self._lib.pcfd_autotune(self.handle, repetitions, C.byref(out)))
        return out.as_dict()

    def tuning(self) -> dict:
        out = L.Tuning()
        L.check(self._lib.pcfd_get_tuning(self.handle, C.byref(out)))
        return out.as_dict()

    def set_tuning(self, xcd_chunk_log2=-1, coarse_block=256, fused_vcycle=1, graph_replay=1):
        t = L.Tuning(xcd_chunk_log2, coarse_block, fused_vcycle, graph_replay, 0, 0.0, 0.0)
        L.check(self._lib.pcfd_set_tuning(self.handle, C.byref(t)))

    def batch_loads(self, pref=0.0) -> list:
        n = self.topology.instances
        out = (L.Loads * n)()
        L.check(self._lib.pcfd_integrate_loads_batch(self.handle, pref, out, n))
        return [out[i].as_dict() for i in range(n)]

    def set_instance_reference(self, i, area, origin=(0.0, 0.0, 0.0)):
        o = (C.c_double * 3)(*origin)
        L.check(self._lib.pcfd_set_reference_instance(self.handle, i, area, o))

    SLICE_FIELDS = {"speed": 0, "u": 1, "v": 2, "w": 3, "p": 4, "cp": 5}

    def sample_slice(self, field, axis, coord, lo, hi, nu, nv, instance=0) -> np.ndarray:
        """Image [nv, nu] of a field on the plane x[axis] = coord over the window lo..hi of the
        two other axes: the finest leaf cell's value at each pixel centre, NaN where solid."""
        f = self.SLICE_FIELDS[field] if isinstance(field, str) else int(field)
        out = np.empty(nu * nv, dtype=np.float32)
        lo2 = (C.c_double * 2)(*lo)
        hi2 = (C.c_double * 2)(*hi)
        L.check(self._lib.pcfd_sample_slice(self.handle, instance, f, axis, coord, lo2, hi2, nu, nv, L.fptr(out)))
        return out.reshape(nv, nu)

    # learned preconditioner (options.preconditioner == PC_LEARNED); no reference counterpart
    def learned_collect(self, max_pairs=8) -> dict:
        """Store post-V-cycle residuals from a plain-MGPCG solve of the next step's projection."""
        st = L.LearnedStats()
        L.check(self._lib.pcfd_learned_collect(self.handle, max_pairs, C.byref(st)))
        return st.as_dict()

    def learned_train(self, steps, batch=1, lr=1e-3) -> dict:
        st = L.LearnedStats()
        L.check(self._lib.pcfd_learned_train(self.handle, steps, batch, lr, C.byref(st)))
        return st.as_dict()

    def learned_weights(self):
        """FP32 master weights [sets, layers, 27, 16, 16] ([set][layer][tap][cout][cin]) and biases
        [sets, layers, 16]; set = AMR level of the brick (levels beyond the last set share it)."""
        w = np.empty(L.NN_WEIGHTS, dtype=np.float32)
        b = np.empty(L.NN_BIASES, dtype=np.float32)
        L.check(self._lib.pcfd_learned_get_weights(self.handle, L.fptr(w), L.fptr(b)))
        return (w.reshape(L.NN_SETS, L.NN_LAYERS, 27, L.NN_C, L.NN_C), b.reshape(L.NN_SETS, L.NN_LAYERS, L.NN_C))

    def set_learned_weights(self, w, b):
        w = np.ascontiguousarray(w, dtype=np.float32).reshape(-1)
        b = np.ascontiguousarray(b, dtype=np.float32).reshape(-1)
        if w.size != L.NN_WEIGHTS or b.size != L.NN_BIASES:
            raise ValueError(f"weights must hold {L.NN_WEIGHTS} values and biases {L.NN_BIASES}")
        L.check(self._lib.pcfd_learned_set_weights(self.handle, L.fptr(w), L.fptr(b)))

    def learned_enable(self, on=True):
        L.check(self._lib.pcfd_learned_enable(self.handle, 1 if on else 0))

    def learned_forward(self, r):
        """Raw network output y (per dof, float32) and per-brick residual scales s."""
        r = np.ascontiguousarray(r, dtype=np.float64)
        y = np.empty(self.n, dtype=np.float32)
        s = np.empty(self.n // 512, dtype=np.float32)
        L.check(self._lib.pcfd_learned_forward(self.handle, L.dptr(r), L.fptr(y), L.fptr(s)))
        return y, s

    def learned_gradient(self, residual):
        """(loss, dW [sets,4,27,16,16], db [sets,4,16]) of one post-V-cycle residual (read as FP32)
        on the network's bricks with the training kernels."""
        r = np.ascontiguousarray(residual, dtype=np.float64)
        loss = C.c_double()
        gw = np.empty(L.NN_WEIGHTS, dtype=np.float32)
        gb = np.empty(L.NN_BIASES, dtype=np.float32)
        L.check(self._lib.pcfd_learned_gradient(self.handle, L.dptr(r), C.byref(loss), L.fptr(gw), L.fptr(gb)))
        return (loss.value, gw.reshape(L.NN_SETS, L.NN_LAYERS, 27, L.NN_C, L.NN_C),
                gb.reshape(L.NN_SETS, L.NN_LAYERS, L.NN_C))
